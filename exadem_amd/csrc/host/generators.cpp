// generators.cpp — seeded synthetic inputs for the sphere workloads (host code, libxdem_host.so).
//
// Random numbers come from a counter-based SplitMix64 stream: value k of stream (seed, tag) is
// splitmix64(splitmix64(seed ^ tag<<32) + k), so every draw is reproducible and independent of the
// order in which it is requested. Uniform doubles take the top 53 bits; normals use Box-Muller.
//
// Packing: random sequential addition (RSA) of the spheres at a reduced radius scale, largest
// first, inside a periodic box; then the radii are restored and overlaps are relaxed by Jacobi
// push-apart sweeps over a cell grid. Pure RSA jams far below the target solid fractions of the
// configurations (0.55-0.58), so this "RSA + growth + relaxation" is the stated resolution.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Stream {
  uint64_t key;
  Stream(uint64_t seed, uint64_t tag) : key(splitmix64(seed ^ (tag << 32))) {}
  uint64_t bits(uint64_t k) const { return splitmix64(key + k * 0x9E3779B97F4A7C15ull); }
  double uniform(uint64_t k) const { return (double)(bits(k) >> 11) * (1.0 / 9007199254740992.0); }
  double normal(uint64_t k) const {
    const double u1 = 1.0 - uniform(2 * k), u2 = uniform(2 * k + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

struct CellGrid {
  double lo[3], L[3], h[3];
  int dims[3];
  std::vector<int64_t> start, items;
  void init(const double* lo_, const double* hi_, double cell) {
    for (int a = 0; a < 3; a++) {
      lo[a] = lo_[a];
      L[a] = hi_[a] - lo_[a];
      dims[a] = std::max(1, (int)std::floor(L[a] / cell));
      h[a] = L[a] / dims[a];
    }
  }
  int64_t ncells() const { return (int64_t)dims[0] * dims[1] * dims[2]; }
  int coord(double x, int a) const {
    double t = x - lo[a];
    t -= L[a] * std::floor(t / L[a]);
    int k = (int)(t / h[a]);
    return std::min(std::max(k, 0), dims[a] - 1);
  }
  int64_t cell_of(double x, double y, double z) const {
    return ((int64_t)coord(z, 2) * dims[1] + coord(y, 1)) * dims[0] + coord(x, 0);
  }
  void build(const double* x, const double* y, const double* z, int64_t n) {
    const int64_t nc = ncells();
    start.assign(nc + 1, 0);
    items.resize(n);
    std::vector<int64_t> c(n);
    for (int64_t i = 0; i < n; i++) {
      c[i] = cell_of(x[i], y[i], z[i]);
      start[c[i] + 1]++;
    }
    for (int64_t k = 0; k < nc; k++) start[k + 1] += start[k];
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    for (int64_t i = 0; i < n; i++) items[fill[c[i]]++] = i;
  }
};

}  // namespace

extern "C" {

// Uniform radii in [rmin, rmax] (stream tag 1) and Gaussian velocities (stream tag 2, per component).
void xdem_gen_uniform(uint64_t seed, uint64_t tag, int64_t n, double a, double b, double* out) {
  Stream s(seed, tag);
  for (int64_t i = 0; i < n; i++) out[i] = a + (b - a) * s.uniform((uint64_t)i);
}

void xdem_gen_normal(uint64_t seed, uint64_t tag, int64_t n, double mean, double sigma, double* out) {
  Stream s(seed, tag);
  for (int64_t i = 0; i < n; i++) out[i] = mean + sigma * s.normal((uint64_t)i);
}

// RSA + growth + relaxation in a box periodic on every axis. Radii are given; positions are
// written to x, y, z. Returns the number of relaxation sweeps; *max_overlap receives the largest
// remaining overlap (r_i + r_j - d, 0 when none). scale is the RSA radius scale (e.g. 0.8).
int64_t xdem_gen_pack_periodic(uint64_t seed, int64_t n, const double* rad, const double* lo, const double* hi,
                               double scale, int64_t max_sweeps, double tol, double* x, double* y, double* z,
                               double* max_overlap) {
  double rmax = 0;
  for (int64_t i = 0; i < n; i++) rmax = std::max(rmax, rad[i]);
  double L[3];
  for (int a = 0; a < 3; a++) L[a] = hi[a] - lo[a];
  // --- RSA, largest radius first
  std::vector<int64_t> order(n);
  for (int64_t i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return rad[a] > rad[b]; });
  CellGrid g;
  g.init(lo, hi, 2.0 * rmax * scale);
  const int64_t nc = g.ncells();
  std::vector<int64_t> head(nc, -1), next(n, -1);  // per-cell linked lists of placed particles
  Stream s(seed, 3);
  uint64_t draw = 0;
  int64_t attempts = 0;
  for (int64_t k = 0; k < n; k++) {
    const int64_t i = order[k];
    const double ri = rad[i] * scale;
    for (int attempt = 0;; attempt++) {
      attempts++;
      const double px = lo[0] + L[0] * s.uniform(draw++);
      const double py = lo[1] + L[1] * s.uniform(draw++);
      const double pz = lo[2] + L[2] * s.uniform(draw++);
      const int cx = g.coord(px, 0), cy = g.coord(py, 1), cz = g.coord(pz, 2);
      bool ok = true;
      for (int dz = -1; dz <= 1 && ok; dz++) {
        int nz = cz + dz;
        double sz = 0;
        if (nz < 0) { nz += g.dims[2]; sz = -L[2]; } else if (nz >= g.dims[2]) { nz -= g.dims[2]; sz = L[2]; }
        for (int dy = -1; dy <= 1 && ok; dy++) {
          int ny = cy + dy;
          double sy = 0;
          if (ny < 0) { ny += g.dims[1]; sy = -L[1]; } else if (ny >= g.dims[1]) { ny -= g.dims[1]; sy = L[1]; }
          for (int dx = -1; dx <= 1 && ok; dx++) {
            int nx = cx + dx;
            double sx = 0;
            if (nx < 0) { nx += g.dims[0]; sx = -L[0]; } else if (nx >= g.dims[0]) { nx -= g.dims[0]; sx = L[0]; }
            for (int64_t j = head[((int64_t)nz * g.dims[1] + ny) * g.dims[0] + nx]; j >= 0; j = next[j]) {
              const double ex = px - (x[j] + sx), ey = py - (y[j] + sy), ez = pz - (z[j] + sz);
              const double rr = ri + rad[j] * scale;
              if (ex * ex + ey * ey + ez * ez < rr * rr) {
                ok = false;
                break;
              }
            }
          }
        }
      }
      if (ok || attempt > 100000) {
        x[i] = px;
        y[i] = py;
        z[i] = pz;
        const int64_t c = g.cell_of(px, py, pz);
        next[i] = head[c];
        head[c] = i;
        break;
      }
    }
  }
  // --- full radii: Jacobi push-apart sweeps
  CellGrid g2;
  g2.init(lo, hi, 2.0 * rmax);
  std::vector<double> dx(n), dy(n), dz(n), X(n), Y(n), Z(n), R(n);
  int64_t sweep = 0;
  double worst = 0;
  for (; sweep <= max_sweeps; sweep++) {
    g2.build(x, y, z, n);
    // gather in cell order so that the stencil reads contiguous memory
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; k++) {
      const int64_t i = g2.items[k];
      X[k] = x[i];
      Y[k] = y[i];
      Z[k] = z[i];
      R[k] = rad[i];
    }
    worst = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(max : worst)
    for (int64_t k = 0; k < n; k++) {
      double ax = 0, ay = 0, az = 0;
      const double xi = X[k], yi = Y[k], zi = Z[k], ri = R[k];
      const int cx = g2.coord(xi, 0), cy = g2.coord(yi, 1), cz = g2.coord(zi, 2);
      for (int oz = -1; oz <= 1; oz++) {
        int nz = cz + oz;
        double sz = 0;
        if (nz < 0) { nz += g2.dims[2]; sz = -L[2]; } else if (nz >= g2.dims[2]) { nz -= g2.dims[2]; sz = L[2]; }
        for (int oy = -1; oy <= 1; oy++) {
          int ny = cy + oy;
          double sy = 0;
          if (ny < 0) { ny += g2.dims[1]; sy = -L[1]; } else if (ny >= g2.dims[1]) { ny -= g2.dims[1]; sy = L[1]; }
          for (int ox = -1; ox <= 1; ox++) {
            int nx = cx + ox;
            double sx = 0;
            if (nx < 0) { nx += g2.dims[0]; sx = -L[0]; } else if (nx >= g2.dims[0]) { nx -= g2.dims[0]; sx = L[0]; }
            const int64_t cb = ((int64_t)nz * g2.dims[1] + ny) * g2.dims[0] + nx;
            for (int64_t t = g2.start[cb]; t < g2.start[cb + 1]; t++) {
              if (t == k) continue;
              const double ex = xi - (X[t] + sx), ey = yi - (Y[t] + sy), ez = zi - (Z[t] + sz);
              const double d2 = ex * ex + ey * ey + ez * ez;
              const double rr = ri + R[t];
              if (d2 < rr * rr && d2 > 0) {
                const double d = std::sqrt(d2);
                const double ov = rr - d;
                worst = std::max(worst, ov);
                // move by this particle's share (inverse to its size) of the overlap, slightly over-relaxed
                const double share = 0.55 * ov * R[t] / (rr * d);
                ax += share * ex;
                ay += share * ey;
                az += share * ez;
              }
            }
          }
        }
      }
      dx[k] = ax;
      dy[k] = ay;
      dz[k] = az;
    }
    if (worst <= tol || sweep == max_sweeps) break;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; k++) {
      const int64_t i = g2.items[k];
      double px = x[i] + dx[k], py = y[i] + dy[k], pz = z[i] + dz[k];
      px -= L[0] * std::floor((px - lo[0]) / L[0]);
      py -= L[1] * std::floor((py - lo[1]) / L[1]);
      pz -= L[2] * std::floor((pz - lo[2]) / L[2]);
      x[i] = px;
      y[i] = py;
      z[i] = pz;
    }
  }
  if (max_overlap) *max_overlap = worst;
  return sweep;
}

}  // extern "C"
