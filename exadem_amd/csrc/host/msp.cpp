// msp.cpp — parser of the deck dialect (see msp.hpp).
#include "msp.hpp"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace xdem_host {

const Node* Node::get(const std::string& key) const {
  if (kind != Map) return nullptr;
  for (auto& kv : map)
    if (kv.first == key) return &kv.second;
  return nullptr;
}
Node* Node::get(const std::string& key) {
  if (kind != Map) return nullptr;
  for (auto& kv : map)
    if (kv.first == key) return &kv.second;
  return nullptr;
}
void Node::set(const std::string& key, const Node& v) {
  if (kind == Null) kind = Map;
  for (auto& kv : map)
    if (kv.first == key) {
      kv.second = v;
      return;
    }
  map.emplace_back(key, v);
}

namespace {

struct Line {
  int indent;
  std::string text;  // without indentation and comment, right-trimmed
  int no;
};

std::string rtrim(const std::string& s) {
  size_t e = s.size();
  while (e > 0 && std::isspace((unsigned char)s[e - 1])) e--;
  return s.substr(0, e);
}
std::string trim(const std::string& s) {
  size_t b = 0;
  while (b < s.size() && std::isspace((unsigned char)s[b])) b++;
  return rtrim(s.substr(b));
}

// removes a comment: '#' at the start or after whitespace, outside quotes
std::string strip_comment(const std::string& s) {
  char q = 0;
  for (size_t i = 0; i < s.size(); i++) {
    const char c = s[i];
    if (q) {
      if (c == q) q = 0;
      continue;
    }
    if (c == '"' || c == '\'') {
      q = c;
      continue;
    }
    if (c == '#' && (i == 0 || std::isspace((unsigned char)s[i - 1]))) return s.substr(0, i);
  }
  return s;
}

class Parser {
 public:
  Parser(const std::string& text, const std::string& src) : src_(src) {
    std::istringstream in(text);
    std::string raw;
    int no = 0;
    while (std::getline(in, raw)) {
      no++;
      raw_.push_back(raw);
      for (char& c : raw)
        if (c == '\t') c = ' ';
      const std::string t = rtrim(strip_comment(raw));
      size_t ind = 0;
      while (ind < t.size() && t[ind] == ' ') ind++;
      if (ind == t.size()) continue;
      lines_.push_back({(int)ind, t.substr(ind), no});
    }
  }

  Node parse() {
    if (lines_.empty()) return Node{};
    Node n = block(lines_[0].indent);
    if (pos_ < lines_.size()) fail(lines_[pos_], "unexpected indentation");
    return n;
  }

 private:
  [[noreturn]] void fail(const Line& l, const std::string& what) {
    throw ParseError(src_ + ":" + std::to_string(l.no) + ": " + what + ": '" + l.text + "'");
  }

  static bool is_item(const std::string& t) { return t == "-" || (t.size() >= 2 && t[0] == '-' && t[1] == ' '); }

  // position of the ':' that ends a mapping key (followed by space or end), outside quotes/brackets
  static size_t key_colon(const std::string& t) {
    char q = 0;
    int depth = 0;
    for (size_t i = 0; i < t.size(); i++) {
      const char c = t[i];
      if (q) {
        if (c == q) q = 0;
        continue;
      }
      if (c == '"' || c == '\'') {
        if (i == 0) {
          q = c;
          continue;
        }
      }
      if (c == '[' || c == '{') depth++;
      if (c == ']' || c == '}') depth--;
      if (c == ':' && depth == 0 && (i + 1 == t.size() || t[i + 1] == ' ')) return i;
    }
    return std::string::npos;
  }

  Node block(int indent) {
    const Line& l = lines_[pos_];
    if (l.indent != indent) fail(l, "bad indentation");
    if (is_item(l.text)) return sequence(indent);
    return mapping(indent, std::string(), -1);
  }

  Node sequence(int indent) {
    Node n;
    n.kind = Node::Seq;
    n.line = lines_[pos_].no;
    while (pos_ < lines_.size() && lines_[pos_].indent == indent && is_item(lines_[pos_].text)) {
      const Line l = lines_[pos_];
      const std::string rest = l.text.size() > 1 ? trim(l.text.substr(1)) : std::string();
      const int col = indent + (int)(l.text.size() - rest.size());
      if (rest.empty()) {
        pos_++;
        if (pos_ < lines_.size() && lines_[pos_].indent > indent) n.seq.push_back(block(lines_[pos_].indent));
        else n.seq.push_back(Node{});
      } else if (rest[0] != '[' && rest[0] != '{' && key_colon(rest) != std::string::npos) {
        // mapping that starts on the item line; its other keys sit at column `col`
        n.seq.push_back(mapping(col, rest, l.no));
      } else {
        pos_++;
        n.seq.push_back(value(rest, l));
      }
    }
    return n;
  }

  // mapping at `indent`; when `first` is set it is the text of the first key line (already inside
  // an item, not in lines_[pos_] as a line of its own)
  Node mapping(int indent, const std::string& first, int first_no) {
    Node n;
    n.kind = Node::Map;
    bool use_first = !first.empty();
    n.line = use_first ? first_no : lines_[pos_].no;
    while (true) {
      Line l;
      if (use_first) {
        l = {indent, first, first_no};
        pos_++;  // the item line
        use_first = false;
      } else {
        if (pos_ >= lines_.size() || lines_[pos_].indent != indent || is_item(lines_[pos_].text)) break;
        l = lines_[pos_++];
      }
      const size_t c = key_colon(l.text);
      if (c == std::string::npos) fail(l, "expected 'key: value'");
      std::string key = trim(l.text.substr(0, c));
      if (key.size() >= 2 && (key[0] == '"' || key[0] == '\'') && key.back() == key[0]) key = key.substr(1, key.size() - 2);
      const std::string rest = trim(l.text.substr(c + 1));
      Node v;
      if (rest.empty()) {
        if (pos_ < lines_.size() && lines_[pos_].indent > indent) v = block(lines_[pos_].indent);
        else if (pos_ < lines_.size() && lines_[pos_].indent == indent && is_item(lines_[pos_].text)) v = sequence(indent);
      } else if (rest == "|" || rest == "|-" || rest == ">" || rest == ">-") {
        // literal block: raw lines deeper than the key
        std::string text;
        while (pos_ < lines_.size() && lines_[pos_].indent > indent) {
          text += lines_[pos_].text + "\n";
          pos_++;
        }
        v = Node::scalar(text, l.no);
      } else {
        v = value(rest, l);
      }
      v.line = l.no;
      n.set(key, v);
    }
    return n;
  }

  // inline value: flow collection (possibly continued on following deeper lines) or scalar
  Node value(std::string rest, const Line& l) {
    if (rest[0] == '[' || rest[0] == '{') {
      while (!balanced(rest) && pos_ < lines_.size()) rest += " " + lines_[pos_++].text;
      size_t p = 0;
      Node n = flow(rest, p, l);
      skip_ws(rest, p);
      if (p != rest.size()) fail(l, "trailing text after a flow collection");
      return n;
    }
    return Node::scalar(unquote(rest), l.no);
  }

  static bool balanced(const std::string& s) {
    int d = 0;
    char q = 0;
    for (char c : s) {
      if (q) {
        if (c == q) q = 0;
        continue;
      }
      if (c == '"' || c == '\'') q = c;
      else if (c == '[' || c == '{') d++;
      else if (c == ']' || c == '}') d--;
    }
    return d == 0;
  }
  static void skip_ws(const std::string& s, size_t& p) {
    while (p < s.size() && std::isspace((unsigned char)s[p])) p++;
  }
  static std::string unquote(const std::string& v) {
    if (v.size() >= 2 && (v[0] == '"' || v[0] == '\'') && v.back() == v[0]) return v.substr(1, v.size() - 2);
    return v;
  }

  Node flow(const std::string& s, size_t& p, const Line& l) {
    skip_ws(s, p);
    if (p >= s.size()) fail(l, "unterminated flow collection");
    if (s[p] == '[') {
      Node n;
      n.kind = Node::Seq;
      p++;
      skip_ws(s, p);
      if (p < s.size() && s[p] == ']') {
        p++;
        return n;
      }
      while (true) {
        n.seq.push_back(flow(s, p, l));
        skip_ws(s, p);
        if (p < s.size() && s[p] == ',') {
          p++;
          continue;
        }
        if (p < s.size() && s[p] == ']') {
          p++;
          return n;
        }
        fail(l, "expected ',' or ']'");
      }
    }
    if (s[p] == '{') {
      Node n;
      n.kind = Node::Map;
      p++;
      skip_ws(s, p);
      if (p < s.size() && s[p] == '}') {
        p++;
        return n;
      }
      while (true) {
        skip_ws(s, p);
        size_t k0 = p;
        while (p < s.size() && s[p] != ':' && s[p] != ',' && s[p] != '}') p++;
        if (p >= s.size() || s[p] != ':') fail(l, "expected 'key: value' in a flow mapping");
        const std::string key = unquote(trim(s.substr(k0, p - k0)));
        p++;
        Node v = flow(s, p, l);
        n.set(key, v);
        skip_ws(s, p);
        if (p < s.size() && s[p] == ',') {
          p++;
          continue;
        }
        if (p < s.size() && s[p] == '}') {
          p++;
          return n;
        }
        fail(l, "expected ',' or '}'");
      }
    }
    // scalar inside a flow collection
    size_t b = p;
    if (s[p] == '"' || s[p] == '\'') {
      const char q = s[p++];
      while (p < s.size() && s[p] != q) p++;
      if (p >= s.size()) fail(l, "unterminated quote");
      p++;
      return Node::scalar(s.substr(b + 1, p - b - 2), l.no);
    }
    while (p < s.size() && s[p] != ',' && s[p] != ']' && s[p] != '}') p++;
    return Node::scalar(trim(s.substr(b, p - b)), l.no);
  }

  std::string src_;
  std::vector<std::string> raw_;
  std::vector<Line> lines_;
  size_t pos_ = 0;
};

bool file_exists(const std::string& p) {
  std::ifstream f(p);
  return f.good();
}
std::string dir_of(const std::string& p) {
  const size_t k = p.find_last_of('/');
  return k == std::string::npos ? std::string(".") : p.substr(0, k);
}

Node load_rec(const std::string& path, const std::vector<std::string>& dirs, int depth) {
  if (depth > 32) throw ParseError(path + ": includes nested too deeply");
  std::ifstream f(path);
  if (!f) throw ParseError("cannot open deck file " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  Node doc = parse_msp(ss.str(), path);
  if (doc.is_null()) doc.kind = Node::Map;
  if (!doc.is_map()) throw ParseError(path + ": a deck is a mapping of operator names");
  Node out;
  out.kind = Node::Map;
  if (const Node* inc = doc.get("includes")) {
    if (!inc->is_seq()) throw ParseError(path + ": includes must be a list");
    for (const Node& it : inc->seq) {
      if (!it.is_scalar()) throw ParseError(path + ": include entries are file names");
      std::string found;
      std::vector<std::string> cand{dir_of(path) + "/" + it.s};
      for (auto& d : dirs) cand.push_back(d + "/" + it.s);
      for (auto& c : cand)
        if (file_exists(c)) {
          found = c;
          break;
        }
      if (found.empty()) throw ParseError(path + ": included file not found: " + it.s);
      merge_into(out, load_rec(found, dirs, depth + 1));
    }
  }
  for (auto& kv : doc.map)
    if (kv.first != "includes") {
      Node one;
      one.kind = Node::Map;
      one.map.push_back(kv);
      merge_into(out, one);
    }
  return out;
}

}  // namespace

Node parse_msp(const std::string& text, const std::string& source) { return Parser(text, source).parse(); }

void merge_into(Node& dst, const Node& src) {
  if (!(dst.is_map() && src.is_map())) {
    dst = src;
    return;
  }
  for (auto& kv : src.map) {
    Node* d = dst.get(kv.first);
    if (d && d->is_map() && kv.second.is_map()) merge_into(*d, kv.second);
    else dst.set(kv.first, kv.second);
  }
}

Node load_msp(const std::string& path, const std::vector<std::string>& search_dirs) {
  return load_rec(path, search_dirs, 0);
}

double to_double(const Node& n) {
  if (!n.is_scalar()) throw ParseError("line " + std::to_string(n.line) + ": expected a number");
  const char* b = n.s.c_str();
  char* e = nullptr;
  const double v = std::strtod(b, &e);
  if (e == b) throw ParseError("line " + std::to_string(n.line) + ": not a number: '" + n.s + "'");
  std::string unit = trim(std::string(e));
  if (unit.empty()) return v;
  // SI units are the internal units; a few length prefixes are converted
  static const std::pair<const char*, double> scale[] = {{"mm", 1e-3}, {"cm", 1e-2}, {"um", 1e-6}, {"nm", 1e-9},
                                                         {"ms", 1e-3}, {"us", 1e-6}};
  for (auto& p : scale)
    if (unit == p.first) return v * p.second;
  for (char c : unit)
    if (!(std::isalpha((unsigned char)c) || c == '/' || c == '^' || c == '.' || c == '-' || std::isdigit((unsigned char)c)))
      throw ParseError("line " + std::to_string(n.line) + ": bad unit in '" + n.s + "'");
  return v;
}
long long to_int(const Node& n) { return (long long)std::llround(to_double(n)); }
bool to_bool(const Node& n) {
  if (!n.is_scalar()) throw ParseError("line " + std::to_string(n.line) + ": expected a boolean");
  if (n.s == "true" || n.s == "True" || n.s == "yes" || n.s == "1") return true;
  if (n.s == "false" || n.s == "False" || n.s == "no" || n.s == "0") return false;
  throw ParseError("line " + std::to_string(n.line) + ": not a boolean: '" + n.s + "'");
}
std::vector<double> to_doubles(const Node& n) {
  std::vector<double> v;
  if (n.is_seq())
    for (auto& x : n.seq) v.push_back(to_double(x));
  else
    v.push_back(to_double(n));
  return v;
}

std::string dump(const Node& n, int indent) {
  const std::string pad(indent, ' ');
  std::string out;
  switch (n.kind) {
    case Node::Null: return "~";
    case Node::Scalar: return n.s;
    case Node::Seq:
      out = "[";
      for (size_t k = 0; k < n.seq.size(); k++) out += (k ? ", " : "") + dump(n.seq[k], indent);
      return out + "]";
    case Node::Map:
      out = "{";
      for (size_t k = 0; k < n.map.size(); k++) out += (k ? ", " : "") + n.map[k].first + ": " + dump(n.map[k].second, indent);
      return out + "}";
  }
  return out;
}

}  // namespace xdem_host
