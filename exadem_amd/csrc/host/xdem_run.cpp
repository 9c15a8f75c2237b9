// xdem_run.cpp — command-line deck runner: xdem-run deck.msp [--config DIR] [--set name=value]...
// [--device N] [--dump FILE]. Prints the run statistics as one JSON line.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/xdem_host.h"

int main(int argc, char** argv) {
  std::string deck_path, dump;
  std::vector<std::string> dirs, sets;
  int device = 0;
  for (int k = 1; k < argc; k++) {
    const std::string a = argv[k];
    if (a == "--config" && k + 1 < argc) dirs.push_back(argv[++k]);
    else if (a == "--set" && k + 1 < argc) sets.push_back(argv[++k]);
    else if (a == "--device" && k + 1 < argc) device = std::atoi(argv[++k]);
    else if (a == "--dump" && k + 1 < argc) dump = argv[++k];
    else if (deck_path.empty()) deck_path = a;
    else {
      std::fprintf(stderr, "usage: xdem-run deck.msp [--config DIR] [--set name=value] [--device N] [--dump FILE]\n");
      return 2;
    }
  }
  if (deck_path.empty()) {
    std::fprintf(stderr, "usage: xdem-run deck.msp [--config DIR] [--set name=value] [--device N] [--dump FILE]\n");
    return 2;
  }
  std::vector<const char*> cdirs;
  for (auto& d : dirs) cdirs.push_back(d.c_str());
  xdem_deck* deck = nullptr;
  char msg[1024];
  if (xdem_deck_load(deck_path.c_str(), cdirs.data(), (int32_t)cdirs.size(), &deck) != XDEM_OK) {
    xdem_deck_last_error(deck, msg, sizeof(msg));
    std::fprintf(stderr, "xdem-run: %s\n", msg);
    xdem_deck_destroy(deck);
    return 1;
  }
  for (auto& s : sets) {
    const size_t eq = s.find('=');
    if (eq == std::string::npos) {
      std::fprintf(stderr, "xdem-run: --set expects name=value\n");
      return 2;
    }
    xdem_deck_set_global(deck, s.substr(0, eq).c_str(), s.substr(eq + 1).c_str());
  }
  xdem_ctx* ctx = nullptr;
  if (xdem_ctx_create(device, &ctx) != XDEM_OK) {
    std::fprintf(stderr, "xdem-run: no device %d\n", device);
    return 1;
  }
  xdem_deck_stats st{};
  const int rc = xdem_deck_run(deck, ctx, &st);
  if (rc != XDEM_OK) {
    xdem_deck_last_error(deck, msg, sizeof(msg));
    std::fprintf(stderr, "xdem-run: %s\n", msg);
  } else if (!dump.empty()) {
    const int64_t n = xdem_num_particles(ctx);
    std::vector<uint64_t> id(n);
    std::vector<double> x(n), y(n), z(n), vx(n), vy(n), vz(n);
    xdem_particles p{};
    p.n = n;
    p.id = id.data();
    p.rx = x.data(); p.ry = y.data(); p.rz = z.data();
    p.vx = vx.data(); p.vy = vy.data(); p.vz = vz.data();
    xdem_particles_download(ctx, &p);
    FILE* f = std::fopen(dump.c_str(), "w");
    if (f) {
      std::fprintf(f, "id,rx,ry,rz,vx,vy,vz\n");
      for (int64_t i = 0; i < n; i++)
        std::fprintf(f, "%llu,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g\n", (unsigned long long)id[i], x[i], y[i], z[i], vx[i],
                     vy[i], vz[i]);
      std::fclose(f);
    }
  }
  std::printf("{\"particles\": %lld, \"iterations\": %lld, \"rebuilds\": %lld, \"operator_calls\": %lld, "
              "\"seconds\": %.6f, \"status\": %d}\n",
              (long long)st.particles, (long long)st.iterations, (long long)st.rebuilds, (long long)st.operator_calls,
              st.seconds, rc);
  xdem_ctx_destroy(ctx);
  xdem_deck_destroy(deck);
  return rc == XDEM_OK ? 0 : 1;
}
