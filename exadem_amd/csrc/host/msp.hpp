// msp.hpp — the YAML subset of exaNBody decks (.msp): block and flow mappings and sequences,
// plain / quoted scalars, literal blocks ("|"), comments, "includes" merged recursively.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace xdem_host {

struct Node {
  enum Kind { Null, Scalar, Seq, Map } kind = Null;
  std::string s;                                   // Scalar
  std::vector<Node> seq;                           // Seq
  std::vector<std::pair<std::string, Node>> map;   // Map (insertion order kept)
  int line = 0;

  bool is_null() const { return kind == Null; }
  bool is_scalar() const { return kind == Scalar; }
  bool is_seq() const { return kind == Seq; }
  bool is_map() const { return kind == Map; }
  const Node* get(const std::string& key) const;   // map lookup, nullptr if absent
  Node* get(const std::string& key);
  void set(const std::string& key, const Node& v);  // replace or append
  static Node scalar(const std::string& v, int line = 0) {
    Node n;
    n.kind = Scalar;
    n.s = v;
    n.line = line;
    return n;
  }
};

struct ParseError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Parses one document.
Node parse_msp(const std::string& text, const std::string& source = "<string>");
// Reads a file, resolves "includes" (relative to the including file, then to each search dir) and
// merges: included files first, in order, then the file's own keys; maps merge key by key,
// everything else is replaced.
Node load_msp(const std::string& path, const std::vector<std::string>& search_dirs);
void merge_into(Node& dst, const Node& src);

// Scalar conversions. Numbers may carry a unit suffix ("1.0e-4 s", "2 m"): SI units are taken as
// they are; anything else after the number is an error.
double to_double(const Node& n);
long long to_int(const Node& n);
bool to_bool(const Node& n);
std::vector<double> to_doubles(const Node& n);
std::string dump(const Node& n, int indent = 0);

}  // namespace xdem_host
