// api.hip — C ABI of libxdem_hip.so: context, buffers, upload/download, device scans,
// operator entry points and the fused multi-step engine.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unordered_set>
#include <vector>

#include "xdem_internal.h"

using namespace xdem;

namespace xdem {

int set_err(xdem_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_check(xdem_ctx* c, hipError_t e, const char* what) {
  return set_err(c, XDEM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
static int dalloc(xdem_ctx* c, T** p, int64_t count) {
  if (*p) {
    hipError_t e = hipFree(*p);
    *p = nullptr;
    if (e != hipSuccess) return hip_check(c, e, "hipFree");
  }
  if (count <= 0) return XDEM_OK;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)count);
  if (e != hipSuccess) return hip_check(c, e, "hipMalloc");
  return XDEM_OK;
}

template <typename T>
static void dfree(T** p) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
}

static int alloc_pstate(xdem_ctx* c, PState& s, int64_t n) {
  int rc = XDEM_OK;
#define A(ptr)                                 \
  if ((rc = dalloc(c, &s.ptr, n)) != XDEM_OK) \
    return rc;
  A(core)
  A(ax) A(ay) A(az) A(mx) A(my) A(mz) A(alx) A(aly) A(alz) A(ix) A(iy) A(iz) A(bx) A(by) A(bz) A(bq)
  A(id) A(type) A(group)
#undef A
  return XDEM_OK;
}

static void free_pstate(PState& s) {
  dfree(&s.core);
  dfree(&s.ax); dfree(&s.ay); dfree(&s.az); dfree(&s.mx); dfree(&s.my); dfree(&s.mz);
  dfree(&s.alx); dfree(&s.aly); dfree(&s.alz); dfree(&s.ix); dfree(&s.iy); dfree(&s.iz);
  dfree(&s.bx); dfree(&s.by); dfree(&s.bz); dfree(&s.bq);
  dfree(&s.id); dfree(&s.type); dfree(&s.group);
}

static void free_nl(NbrList& l) {
  dfree(&l.word); dfree(&l.sx); dfree(&l.hx); dfree(&l.hy); dfree(&l.hz); dfree(&l.mhx); dfree(&l.mhy); dfree(&l.mhz);
  dfree(&l.fx32); dfree(&l.fy32); dfree(&l.fz32); dfree(&l.mfx32); dfree(&l.mfy32); dfree(&l.mfz32);
  dfree(&l.cnt); dfree(&l.kmax); dfree(&l.base);
  l.slots = l.cap_slots = l.pairs = 0;
}

static void free_wl(WallList& w) {
  dfree(&w.mask); dfree(&w.hx); dfree(&w.hy); dfree(&w.hz); dfree(&w.mhx); dfree(&w.mhy); dfree(&w.mhz);
}

int ensure_capacity(xdem_ctx* c, int64_t n) {
  if (n <= c->cap) return XDEM_OK;
  for (int k = 0; k < 2; k++) {
    free_pstate(c->ps[k]);
    int rc = alloc_pstate(c, c->ps[k], n);
    if (rc) return rc;
  }
  int rc;
  dfree(&c->core_alt);
  if ((rc = dalloc(c, &c->core_alt, n))) return rc;
  if ((rc = dalloc(c, &c->cell_of, n))) return rc;
  if ((rc = dalloc(c, &c->perm, n))) return rc;
  if ((rc = dalloc(c, &c->perm_tmp, n))) return rc;
  if ((rc = dalloc(c, &c->sort_key, n))) return rc;
  const int64_t nw = (n + kWave - 1) / kWave;
  for (int k = 0; k < 2; k++) {
    NbrList& l = c->nl[k];
    if ((rc = dalloc(c, &l.cnt, n))) return rc;
    if ((rc = dalloc(c, &l.kmax, nw + 1))) return rc;
    if ((rc = dalloc(c, &l.base, nw + 1))) return rc;
  }
  c->cap = n;
  c->wl_nw = -1;  // force wall-list reallocation
  return XDEM_OK;
}

int ensure_cells(xdem_ctx* c, int64_t ncells) {
  if (ncells > c->cap_cells) {
    int rc;
    if ((rc = dalloc(c, &c->cell_count, ncells + 1))) return rc;
    if ((rc = dalloc(c, &c->cell_start, ncells + 1))) return rc;
    if ((rc = dalloc(c, &c->cell_rank, ncells))) return rc;
    c->cap_cells = ncells;
  }
  // rank of every cell along a Morton (Z-order) curve over the grid: consecutive particles then
  // cover compact 3-D regions, which keeps the 27-cell stencils of a block inside one L2
  const Grid& g = c->grid;
  auto spread = [](uint64_t v) {
    uint64_t r = 0;
    for (int b = 0; b < 21; b++) r |= ((v >> b) & 1ull) << (3 * b);
    return r;
  };
  std::vector<std::pair<uint64_t, uint32_t>> key(ncells);
  for (int z = 0; z < g.dims[2]; z++)
    for (int y = 0; y < g.dims[1]; y++)
      for (int x = 0; x < g.dims[0]; x++) {
        const uint32_t cid = (uint32_t)(((int64_t)z * g.dims[1] + y) * g.dims[0] + x);
        key[cid] = {spread(x) | (spread(y) << 1) | (spread(z) << 2), cid};
      }
  std::sort(key.begin(), key.end());
  std::vector<uint32_t> rank(ncells);
  for (int64_t r = 0; r < ncells; r++) rank[key[r].second] = (uint32_t)r;
  XDEM_HIP(c, hipMemcpyAsync(c->cell_rank, rank.data(), sizeof(uint32_t) * ncells, hipMemcpyHostToDevice, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  return XDEM_OK;
}

int alloc_walls(xdem_ctx* c, int nw) {
  if (nw == c->wl_nw) return XDEM_OK;
  int rc;
  for (int k = 0; k < 2; k++) {
    WallList& w = c->wl[k];
    free_wl(w);
    if (nw > 0 && c->cap > 0) {
      const int64_t m = (int64_t)nw * c->cap;
      if ((rc = dalloc(c, &w.mask, c->cap))) return rc;
      if ((rc = dalloc(c, &w.hx, m))) return rc;
      if ((rc = dalloc(c, &w.hy, m))) return rc;
      if ((rc = dalloc(c, &w.hz, m))) return rc;
      if ((rc = dalloc(c, &w.mhx, m))) return rc;
      if ((rc = dalloc(c, &w.mhy, m))) return rc;
      if ((rc = dalloc(c, &w.mhz, m))) return rc;
      XDEM_HIP(c, hipMemsetAsync(w.mask, 0, c->cap, c->stream));
    }
  }
  c->wl_nw = nw;
  return XDEM_OK;
}

// ---------------------------------------------------------------- exclusive scans
constexpr int kScanBS = 256;
constexpr int kScanIPT = 4;
constexpr int kScanTile = kScanBS * kScanIPT;

template <typename TI, typename TO>
__global__ __launch_bounds__(kScanBS) void k_scan_reduce(const TI* __restrict__ in, int64_t n, TO mul,
                                                         TO* __restrict__ block_sums) {
  __shared__ TO sm[kScanBS];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  TO s = 0;
  for (int k = 0; k < kScanIPT; k++) {
    const int64_t i = base + (int64_t)k * kScanBS + threadIdx.x;
    if (i < n) s += (TO)in[i] * mul;
  }
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int o = kScanBS / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) block_sums[blockIdx.x] = sm[0];
}

// single block: exclusive scan of block sums in place, writes the total to *total
template <typename TO>
__global__ __launch_bounds__(kScanBS) void k_scan_blocks(TO* __restrict__ bs, int64_t nb, TO* __restrict__ total) {
  __shared__ TO sm[kScanBS];
  TO carry = 0;
  for (int64_t b0 = 0; b0 < nb; b0 += kScanBS) {
    const int64_t i = b0 + threadIdx.x;
    TO v = (i < nb) ? bs[i] : (TO)0;
    sm[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < kScanBS; o <<= 1) {
      TO t = (threadIdx.x >= (unsigned)o) ? sm[threadIdx.x - o] : (TO)0;
      __syncthreads();
      sm[threadIdx.x] += t;
      __syncthreads();
    }
    const TO incl = sm[threadIdx.x];
    if (i < nb) bs[i] = carry + incl - v;
    __syncthreads();
    carry += sm[kScanBS - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

template <typename TI, typename TO>
__global__ __launch_bounds__(kScanBS) void k_scan_apply(const TI* __restrict__ in, int64_t n, TO mul,
                                                        const TO* __restrict__ block_off, TO* __restrict__ out) {
  __shared__ TO sm[kScanBS];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanIPT;
  TO v[kScanIPT];
  TO s = 0;
  for (int k = 0; k < kScanIPT; k++) {
    const int64_t i = base + k;
    v[k] = (i < n) ? (TO)in[i] * mul : (TO)0;
    s += v[k];
  }
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < kScanBS; o <<= 1) {
    TO t = (threadIdx.x >= (unsigned)o) ? sm[threadIdx.x - o] : (TO)0;
    __syncthreads();
    sm[threadIdx.x] += t;
    __syncthreads();
  }
  TO run = block_off[blockIdx.x] + sm[threadIdx.x] - s;
  for (int k = 0; k < kScanIPT; k++) {
    const int64_t i = base + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
}

template <typename TI, typename TO>
static int scan_impl(xdem_ctx* c, const TI* in, TO* out, int64_t n, TO mul) {
  const int64_t nb = std::max<int64_t>(1, (n + kScanTile - 1) / kScanTile);
  const int64_t need = (int64_t)sizeof(TO) * (nb + 1);
  if (need > c->scan_tmp_bytes) {
    if (c->scan_tmp) (void)hipFree(c->scan_tmp);
    c->scan_tmp = nullptr;
    XDEM_HIP(c, hipMalloc(&c->scan_tmp, need));
    c->scan_tmp_bytes = need;
  }
  TO* bs = (TO*)c->scan_tmp;
  hipLaunchKernelGGL((k_scan_reduce<TI, TO>), dim3((unsigned)nb), dim3(kScanBS), 0, c->stream, in, n, mul, bs);
  hipLaunchKernelGGL((k_scan_blocks<TO>), dim3(1), dim3(kScanBS), 0, c->stream, bs, nb, out + n);
  hipLaunchKernelGGL((k_scan_apply<TI, TO>), dim3((unsigned)nb), dim3(kScanBS), 0, c->stream, in, n, mul, bs, out);
  XDEM_HIP(c, hipGetLastError());
  return XDEM_OK;
}

int scan_u32(xdem_ctx* c, const uint32_t* in, uint32_t* out, int64_t n) {
  return scan_impl<uint32_t, uint32_t>(c, in, out, n, 1u);
}
int scan_u64_from_u32(xdem_ctx* c, const uint32_t* in, uint64_t* out, int64_t n, uint32_t mul) {
  return scan_impl<uint32_t, uint64_t>(c, in, out, n, (uint64_t)mul);
}

int read_u64(xdem_ctx* c, const uint64_t* d, uint64_t* h) {
  XDEM_HIP(c, hipMemcpyAsync(c->h_u32tmp, d, sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  std::memcpy(h, c->h_u32tmp, sizeof(uint64_t));
  return XDEM_OK;
}

// ---------------------------------------------------------------- upload / download kernels
__global__ void k_pack_in(int64_t n, const double* rx, const double* ry, const double* rz, const double* rad,
                          const double* vx, const double* vy, const double* vz, const double* m, const double* wx,
                          const double* wy, const double* wz, const double* qw, const double* qx, const double* qy,
                          const double* qz, const double* hom, PState s) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Core c;
  c.xr = {rx[i], ry[i], rz[i], rad[i]};
  c.vm = {vx[i], vy[i], vz[i], m[i]};
  c.w = {wx[i], wy[i], wz[i], hom[i]};
  c.q = {qw[i], qx[i], qy[i], qz[i]};
  s.core[i] = c;
  s.bq[i] = c.q;
}

__global__ void k_unpack_out(int64_t n, PState s, double* rx, double* ry, double* rz, double* rad, double* vx,
                             double* vy, double* vz, double* m, double* wx, double* wy, double* wz, double* qw,
                             double* qx, double* qy, double* qz, double* hom) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Core c = s.core[i];
  const d4 a = c.xr, b = c.vm, w = c.w, q = c.q;
  rx[i] = a.x; ry[i] = a.y; rz[i] = a.z; rad[i] = a.w;
  vx[i] = b.x; vy[i] = b.y; vz[i] = b.z; m[i] = b.w;
  wx[i] = w.x; wy[i] = w.y; wz[i] = w.z; hom[i] = w.w;
  qw[i] = q.x; qx[i] = q.y; qy[i] = q.z; qz[i] = q.w;
}

}  // namespace xdem

// ================================================================== C ABI
namespace xdem {
static inline bool motion_on(const Motion& m, double time) { return time >= m.t_start && time <= m.t_end; }
static inline double shaker_signal(const Motion& m, double time) {
  time -= m.t_start;
  return m.amplitude * std::sin(m.omega * time);
}
static inline V3 shaker_velocity(const Motion& m, double time) {
  time -= m.t_start;
  return m.amplitude * m.omega * std::cos(m.omega * time) * m.shaker_dir;
}
// push_f_v_r of a moving driver: Surface::push_f_v_r (surface.hpp:260-299, no acceleration for
// kinematic motions), Cylinder (cylinder.hpp:160-173), Ball (ball.hpp:203-213)
static void push_f_v_r_one(Surface& s, const Motion& m, double time, double dt) {
  if (m.type == XDEM_MOTION_STATIONARY) return;
  if (s.kind == XDEM_DRIVER_SURFACE) {
    const double acc = 0.0;
    double displ = dt * dot(s.vel, s.normal) + 0.5 * dt * dt * acc;
    if (m.type == XDEM_MOTION_SHAKER) {
      const double next = shaker_signal(m, time + dt), cur = shaker_signal(m, time);
      displ = (next - cur) * dot(m.shaker_dir, s.normal);
      s.vel = shaker_velocity(m, time + dt);
    }
    s.center = s.center + displ * s.normal;
    s.offset += displ;
    s.center_proj = s.center_proj + displ * s.normal;
  } else if (s.kind == XDEM_DRIVER_CYLINDER) {
    if (m.type == XDEM_MOTION_SHAKER) {
      const double displ = shaker_signal(m, time + dt) - shaker_signal(m, time);
      s.center = s.center + displ * m.shaker_dir;
      s.vel = shaker_velocity(m, time + dt);
    }
  } else {
    s.center = s.center + dt * s.vel;
  }
}
// push_f_v (surface.hpp:234-252, ball.hpp:219-234; nothing for a cylinder, cylinder.hpp:149-151)
static void push_f_v_one(Surface& s, const Motion& m) {
  if (s.kind == XDEM_DRIVER_SURFACE) {
    if (m.type == XDEM_MOTION_STATIONARY) s.vel = {0, 0, 0};
    else if (m.type == XDEM_MOTION_LINEAR) s.vel = m.const_vel * m.mvec;
  } else if (s.kind == XDEM_DRIVER_BALL) {
    if (m.type == XDEM_MOTION_LINEAR) s.vel = m.mvec * m.const_vel;
  }
}
void drivers_push_f_v(Surface* w, const Motion* m, int nw) {
  for (int k = 0; k < nw; k++) push_f_v_one(w[k], m[k]);
}
// push_f_v_r_driver (gated by is_motion_triggered at time + dt, push_f_v_r_driver.cpp:28-37) then
// push_f_v_driver (config_numerical_schemes.msp:13-14)
void drivers_advance(Surface* w, const Motion* m, int nw, double time, double dt) {
  for (int k = 0; k < nw; k++)
    if (motion_on(m[k], time + dt)) push_f_v_r_one(w[k], m[k], time, dt);
  drivers_push_f_v(w, m, nw);
}
// driver_displ_over (driver_displ_over.cu:59-130): projected centre of a surface, centre of a ball; a
// cylinder never triggers
bool drivers_displaced(const Surface* now, const Surface* bk, int nw, double thr) {
  const double r2 = thr * thr;
  for (int k = 0; k < nw; k++) {
    V3 d{0, 0, 0};
    if (now[k].kind == XDEM_DRIVER_SURFACE) d = now[k].center_proj - bk[k].center_proj;
    else if (now[k].kind == XDEM_DRIVER_BALL) d = now[k].center - bk[k].center;
    else continue;
    if (dot(d, d) >= r2) return true;
  }
  return false;
}
static int upload_walls(xdem_ctx* c) {
  XDEM_HIP(c, hipMemcpyAsync(c->d_walls, c->walls, sizeof(Surface) * XDEM_MAX_SURFACES, hipMemcpyHostToDevice,
                             c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  return XDEM_OK;
}
}  // namespace xdem

extern "C" {

int xdem_ctx_create(int32_t device, xdem_ctx** out) {
  if (!out) return XDEM_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return XDEM_ERR_HIP;
  if (device < 0 || device >= ndev) return XDEM_ERR_ARG;
  xdem_ctx* c = new xdem_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc((void**)&c->d_flags, 4 * sizeof(int32_t)) != hipSuccess ||
      hipHostMalloc((void**)&c->h_flags, 4 * sizeof(int32_t), hipHostMallocDefault) != hipSuccess ||
      hipMalloc((void**)&c->d_u32tmp, 16) != hipSuccess || hipMalloc((void**)&c->d_rmax, 8) != hipSuccess ||
      hipMalloc((void**)&c->d_pairs, 16) != hipSuccess ||
      hipHostMalloc((void**)&c->h_u32tmp, 16, hipHostMallocDefault) != hipSuccess ||
      hipMalloc((void**)&c->d_walls, sizeof(Surface) * XDEM_MAX_SURFACES) != hipSuccess) {
    delete c;
    return XDEM_ERR_HIP;
  }
  (void)hipMemsetAsync(c->d_flags, 0, 4 * sizeof(int32_t), c->stream);
  (void)hipStreamSynchronize(c->stream);
  // this GPU's no-allocation scratch address (xdem_simulation_state_noalloc), once per device
  if (noalloc_scratch_init(device) != XDEM_OK) {
    xdem_ctx_destroy(c);
    return XDEM_ERR_HIP;
  }
  *out = c;
  return XDEM_OK;
}

void xdem_ctx_destroy(xdem_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (int k = 0; k < 2; k++) {
    free_pstate(c->ps[k]);
    free_nl(c->nl[k]);
    free_wl(c->wl[k]);
    for (auto& e : c->timers[k].ev) (void)hipEventDestroy(e);
  }
  mesh_free(c);
  dfree(&c->core_alt);
  dfree(&c->cell_of); dfree(&c->cell_count); dfree(&c->cell_start); dfree(&c->cell_rank); dfree(&c->perm);
  dfree(&c->d_rmax); dfree(&c->nbh_tmp); dfree(&c->mxr); dfree(&c->mid); dfree(&c->blk_u32);
  dfree(&c->perm_tmp);
  dfree(&c->sort_key);
  dfree(&c->d_flags); dfree(&c->d_u32tmp); dfree(&c->d_pairs); dfree(&c->d_walls);
  dfree(&c->cstate); dfree(&c->wstate); dfree(&c->simpart); dfree(&c->d_shapes); dfree(&c->vtx); dfree(&c->frange); dfree(&c->fcontrib); dfree(&c->fcontact); dfree(&c->fdn);
  dfree(&c->feat_tmp); dfree(&c->oslot); dfree(&c->feat_ovf); dfree(&c->feat_fpool); dfree(&c->feat_pool); dfree(&c->d_nown);
  dfree(&c->forder); dfree(&c->fclass_blk); dfree(&c->fclass_info); dfree(&c->d_walls_ring);
  if (c->h_walls_ring) (void)hipHostFree(c->h_walls_ring);
  for (int k = 0; k < 2; k++) {
    PolyList& P = c->pl[k];
    dfree(&P.start); dfree(&P.feat); dfree(&P.hx); dfree(&P.hy); dfree(&P.hz);
    dfree(&P.mhx); dfree(&P.mhy); dfree(&P.mhz); dfree(&P.fcount); dfree(&P.fslot); dfree(&P.fown);
    PolyWall& W = c->pw[k];
    dfree(&W.mask); dfree(&W.hx); dfree(&W.hy); dfree(&W.hz); dfree(&W.mhx); dfree(&W.mhy); dfree(&W.mhz);
  }
  if (c->scan_tmp) (void)hipFree(c->scan_tmp);
  if (c->h_flags) (void)hipHostFree(c->h_flags);
  if (c->h_u32tmp) (void)hipHostFree(c->h_u32tmp);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

int xdem_last_error(const xdem_ctx* c, char* buf, int64_t n) {
  if (!c || !buf || n <= 0) return XDEM_ERR_ARG;
  const size_t k = std::min<size_t>((size_t)(n - 1), c->err.size());
  std::memcpy(buf, c->err.data(), k);
  buf[k] = 0;
  return XDEM_OK;
}

int xdem_synchronize(xdem_ctx* c) {
  if (!c) return XDEM_ERR_ARG;
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  return XDEM_OK;
}

int xdem_set_domain(xdem_ctx* c, const xdem_domain* d) {
  if (!c || !d) return XDEM_ERR_ARG;
  if (!(d->cell_size > 0)) return set_err(c, XDEM_ERR_ARG, "domain: cell_size must be > 0");
  Grid g{};
  int64_t nc = 1;
  for (int a = 0; a < 3; a++) {
    const double L = d->hi[a] - d->lo[a];
    if (!(L > 0)) return set_err(c, XDEM_ERR_ARG, "domain: hi must exceed lo on every axis");
    int nd = (int)std::floor(L / d->cell_size);
    if (nd < 1) nd = 1;
    if (d->periodic[a] && nd < 3)
      return set_err(c, XDEM_ERR_ARG, "domain: a periodic axis needs at least 3 cells of size >= cell_size");
    g.lo[a] = d->lo[a];
    g.L[a] = L;
    g.dims[a] = nd;
    g.h[a] = L / nd;
    g.inv_h[a] = 1.0 / g.h[a];
    g.per[a] = d->periodic[a] ? 1 : 0;
    nc *= nd;
  }
  if (nc > (int64_t)UINT32_MAX - 1) return set_err(c, XDEM_ERR_CAPACITY, "domain: too many cells");
  g.ncells = nc;
  c->grid = g;
  c->have_domain = true;
  c->nl_valid = c->nl_prev_valid = c->perm_valid = false;
  return ensure_cells(c, nc);
}

int xdem_set_surfaces(xdem_ctx* c, const xdem_surface* w, int32_t n) {
  if (!c || (n > 0 && !w)) return XDEM_ERR_ARG;
  if (n < 0 || n > XDEM_MAX_SURFACES) return set_err(c, XDEM_ERR_CAPACITY, "at most 8 surfaces");
  for (int k = 0; k < n; k++) {
    Surface s{};
    s.kind = w[k].kind;
    if (s.kind != XDEM_DRIVER_SURFACE && s.kind != XDEM_DRIVER_CYLINDER && s.kind != XDEM_DRIVER_BALL)
      return set_err(c, XDEM_ERR_ARG, "driver: kind is XDEM_DRIVER_SURFACE, _CYLINDER or _BALL");
    s.offset = w[k].offset;
    s.normal = {w[k].normal[0], w[k].normal[1], w[k].normal[2]};
    s.radius = w[k].radius;
    s.axis = {w[k].axis[0], w[k].axis[1], w[k].axis[2]};
    if (s.kind == XDEM_DRIVER_SURFACE) {
      if (!(norm(s.normal) > 0)) return set_err(c, XDEM_ERR_ARG, "surface: normal must be non-zero");
    } else if (!(s.radius > 0)) {
      return set_err(c, XDEM_ERR_ARG, "cylinder / ball: radius must be positive");
    }
    s.center = {w[k].center[0], w[k].center[1], w[k].center[2]};
    if (s.kind == XDEM_DRIVER_SURFACE && s.center.x == 0 && s.center.y == 0 && s.center.z == 0)
      s.center = s.offset * s.normal;
    if (s.kind == XDEM_DRIVER_CYLINDER)  // Cylinder::initialize (cylinder.hpp:104)
      s.center = {s.axis.x * s.center.x, s.axis.y * s.center.y, s.axis.z * s.center.z};
    s.vel = {w[k].vel[0], w[k].vel[1], w[k].vel[2]};
    s.vrot = {w[k].vrot[0], w[k].vrot[1], w[k].vrot[2]};
    s.center_proj = s.offset * s.normal;
    c->walls[k] = s;
  }
  c->nwalls = n;
  for (int k = 0; k < XDEM_MAX_SURFACES; k++) {
    c->motion[k] = Motion{};
    c->walls_bk[k] = c->walls[k];
  }
  c->drivers_move = false;
  XDEM_HIP(c, hipMemcpyAsync(c->d_walls, c->walls, sizeof(Surface) * XDEM_MAX_SURFACES, hipMemcpyHostToDevice,
                             c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  c->nl_valid = false;  // wall candidates must be rebuilt
  return XDEM_OK;
}

// ------------------------------------------------------------------ driver motion
// Motion of each driver as register_surface / register_cylinder / register_ball set it up:
// check_motion_coherence (driver_params.hpp:108-214) normalises a LINEAR motion vector or a shaker
// direction when |v|^2 - 1 >= 1e-14, and a linear surface moves along its normal unless its motion vector
// is +-normal (Surface::initialize, surface.hpp:183-195). Valid motion types per driver:
// surface STATIONARY / LINEAR_MOTION / SHAKER, cylinder STATIONARY / SHAKER, ball STATIONARY / LINEAR_MOTION.
int xdem_set_driver_motion(xdem_ctx* c, const xdem_driver_motion* m, int32_t n) {
  if (!c || (n > 0 && !m)) return XDEM_ERR_ARG;
  if (n != 0 && n != c->nwalls)
    return set_err(c, XDEM_ERR_ARG, "set_driver_motion: one motion per registered driver (or none)");
  Motion mo[XDEM_MAX_SURFACES];
  for (int k = 0; k < n; k++) {
    Motion& a = mo[k];
    a.type = m[k].motion_type;
    a.mvec = {m[k].motion_vector[0], m[k].motion_vector[1], m[k].motion_vector[2]};
    a.const_vel = m[k].const_vel;
    a.t_start = m[k].motion_start_threshold;
    a.t_end = m[k].motion_end_threshold;
    a.amplitude = m[k].amplitude;
    a.omega = m[k].omega;
    a.shaker_dir = {m[k].shaker_dir[0], m[k].shaker_dir[1], m[k].shaker_dir[2]};
    const int kind = c->walls[k].kind;
    const bool ok = a.type == XDEM_MOTION_STATIONARY ||
                    (a.type == XDEM_MOTION_LINEAR && kind != XDEM_DRIVER_CYLINDER) ||
                    (a.type == XDEM_MOTION_SHAKER && kind != XDEM_DRIVER_BALL);
    if (!ok) {
      if (a.type == XDEM_MOTION_LINEAR || a.type == XDEM_MOTION_SHAKER)
        return set_err(c, XDEM_ERR_ARG, "set_driver_motion: motion type not valid for this driver kind");
      return set_err(c, XDEM_ERR_UNSUPPORTED, "set_driver_motion: only STATIONARY, LINEAR_MOTION and SHAKER are provided");
    }
    if (a.type == XDEM_MOTION_SHAKER) {
      if (!(a.amplitude > 0.0) || !(a.omega > 0.0))
        return set_err(c, XDEM_ERR_ARG, "set_driver_motion: SHAKER needs amplitude > 0 and omega > 0");
      if (dot(a.shaker_dir, a.shaker_dir) - 1 >= 1e-14) a.shaker_dir = a.shaker_dir / norm(a.shaker_dir);
    }
    if (a.type == XDEM_MOTION_LINEAR) {
      if (a.mvec.x == 0 && a.mvec.y == 0 && a.mvec.z == 0)
        return set_err(c, XDEM_ERR_ARG, "set_driver_motion: LINEAR_MOTION needs a motion vector");
      if (dot(a.mvec, a.mvec) - 1 >= 1e-14) a.mvec = a.mvec / norm(a.mvec);
      if (kind == XDEM_DRIVER_SURFACE) {
        const V3 nn = c->walls[k].normal, mn = -1.0 * nn;
        const bool same = a.mvec.x == nn.x && a.mvec.y == nn.y && a.mvec.z == nn.z;
        const bool opp = a.mvec.x == mn.x && a.mvec.y == mn.y && a.mvec.z == mn.z;
        if (!same && !opp) a.mvec = nn;
      }
    }
  }
  bool moving = false;
  for (int k = 0; k < XDEM_MAX_SURFACES; k++) {
    c->motion[k] = k < n ? mo[k] : Motion{};
    moving = moving || c->motion[k].type != XDEM_MOTION_STATIONARY;
  }
  c->drivers_move = moving;
  return XDEM_OK;
}


int xdem_push_f_v_r_driver(xdem_ctx* c, double time, double dt) {
  if (!c) return XDEM_ERR_ARG;
  if (c->nwalls == 0) return XDEM_OK;
  for (int k = 0; k < c->nwalls; k++)
    if (motion_on(c->motion[k], time + dt)) push_f_v_r_one(c->walls[k], c->motion[k], time, dt);
  return upload_walls(c);
}

int xdem_push_f_v_driver(xdem_ctx* c, double dt) {
  if (!c) return XDEM_ERR_ARG;
  (void)dt;  // kinematic motions set the velocity outright
  if (c->nwalls == 0) return XDEM_OK;
  drivers_push_f_v(c->walls, c->motion, c->nwalls);
  return upload_walls(c);
}

int xdem_driver_displ_over(xdem_ctx* c, double threshold, int32_t* result) {
  if (!c || !result) return XDEM_ERR_ARG;
  *result = drivers_displaced(c->walls, c->walls_bk, c->nwalls, threshold) ? 1 : 0;
  return XDEM_OK;
}

int xdem_drivers_download(xdem_ctx* c, xdem_surface* out, int32_t n) {
  if (!c || (n > 0 && !out)) return XDEM_ERR_ARG;
  if (n > c->nwalls) return set_err(c, XDEM_ERR_ARG, "drivers_download: n exceeds the registered drivers");
  for (int k = 0; k < n; k++) {
    const Surface& s = c->walls[k];
    xdem_surface r{};
    r.offset = s.offset;
    const V3* src[4] = {&s.normal, &s.center, &s.vel, &s.vrot};
    double* dst[4] = {r.normal, r.center, r.vel, r.vrot};
    for (int q = 0; q < 4; q++) {
      dst[q][0] = src[q]->x;
      dst[q][1] = src[q]->y;
      dst[q][2] = src[q]->z;
    }
    r.kind = s.kind;
    r.radius = s.radius;
    r.axis[0] = s.axis.x;
    r.axis[1] = s.axis.y;
    r.axis[2] = s.axis.z;
    out[k] = r;
  }
  return XDEM_OK;
}

int xdem_set_physical_time(xdem_ctx* c, double time) {
  if (!c) return XDEM_ERR_ARG;
  c->physical_time = time;
  return XDEM_OK;
}

int xdem_physical_time(xdem_ctx* c, double* time) {
  if (!c || !time) return XDEM_ERR_ARG;
  *time = c->physical_time;
  return XDEM_OK;
}

int xdem_next_time_step(xdem_ctx* c, double dt) {
  if (!c) return XDEM_ERR_ARG;
  c->physical_time += dt;
  return XDEM_OK;
}

int xdem_particles_upload(xdem_ctx* c, const xdem_particles* p) {
  if (!c || !p) return XDEM_ERR_ARG;
  const int64_t n = p->n;
  if (n < 0) return XDEM_ERR_ARG;
  if (n > XDEM_MAX_PARTICLES) return set_err(c, XDEM_ERR_CAPACITY, "too many particles");
  if (n > 0 && (!p->id || !p->rx || !p->ry || !p->rz || !p->mass || !p->radius))
    return set_err(c, XDEM_ERR_ARG, "upload: id, rx, ry, rz, mass and radius are required");
  double rmax = 0.0;
  {
    std::unordered_set<uint64_t> seen;
    seen.reserve((size_t)n * 2);
    for (int64_t i = 0; i < n; i++) {
      if (!seen.insert(p->id[i]).second) return set_err(c, XDEM_ERR_ARG, "upload: particle ids must be unique");
      if (!(p->mass[i] > 0) || !(p->radius[i] > 0))
        return set_err(c, XDEM_ERR_ARG, "upload: mass and radius must be > 0");
      rmax = std::max(rmax, p->radius[i]);
    }
  }
  XDEM_HIP(c, hipSetDevice(c->device));
  int rc = ensure_capacity(c, std::max<int64_t>(n, 1));
  if (rc) return rc;
  c->n = n;
  c->rmax_host = rmax;  // radii change only here
  c->cur = 0;
  PState& s = c->ps[0];
  // staging: host vectors with defaults for absent fields
  std::vector<double> zero(n, 0.0), one(n, 1.0);
  auto or0 = [&](double* a) { return a ? a : zero.data(); };
  auto or1 = [&](double* a) { return a ? a : one.data(); };
  std::vector<double*> dtmp;
  auto up = [&](const double* h) -> double* {
    double* d = nullptr;
    if (hipMalloc((void**)&d, sizeof(double) * std::max<int64_t>(n, 1)) != hipSuccess) return nullptr;
    (void)hipMemcpyAsync(d, h, sizeof(double) * n, hipMemcpyHostToDevice, c->stream);
    dtmp.push_back(d);
    return d;
  };
  double* t[16] = {up(p->rx), up(p->ry), up(p->rz), up(p->radius), up(or0(p->vx)), up(or0(p->vy)), up(or0(p->vz)),
                   up(p->mass), up(or0(p->vrotx)), up(or0(p->vroty)), up(or0(p->vrotz)), up(or1(p->qw)),
                   up(or0(p->qx)), up(or0(p->qy)), up(or0(p->qz)), up(or1(p->homothety))};
  for (auto* d : t)
    if (!d) {
      for (auto* x : dtmp) (void)hipFree(x);
      return set_err(c, XDEM_ERR_HIP, "upload: staging allocation failed");
    }
  if (n > 0)
    hipLaunchKernelGGL(k_pack_in, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, n, t[0], t[1], t[2], t[3], t[4],
                       t[5], t[6], t[7], t[8], t[9], t[10], t[11], t[12], t[13], t[14], t[15], s);
  auto cp = [&](double* d, const double* h) {
    return hipMemcpyAsync(d, h, sizeof(double) * n, hipMemcpyHostToDevice, c->stream);
  };
  XDEM_HIP(c, cp(s.ax, or0(p->fx)));
  XDEM_HIP(c, cp(s.ay, or0(p->fy)));
  XDEM_HIP(c, cp(s.az, or0(p->fz)));
  XDEM_HIP(c, cp(s.mx, or0(p->momx)));
  XDEM_HIP(c, cp(s.my, or0(p->momy)));
  XDEM_HIP(c, cp(s.mz, or0(p->momz)));
  XDEM_HIP(c, cp(s.alx, or0(p->arotx)));
  XDEM_HIP(c, cp(s.aly, or0(p->aroty)));
  XDEM_HIP(c, cp(s.alz, or0(p->arotz)));
  XDEM_HIP(c, cp(s.ix, or1(p->inx)));
  XDEM_HIP(c, cp(s.iy, or1(p->iny)));
  XDEM_HIP(c, cp(s.iz, or1(p->inz)));
  XDEM_HIP(c, cp(s.bx, p->rx));
  XDEM_HIP(c, cp(s.by, p->ry));
  XDEM_HIP(c, cp(s.bz, p->rz));
  XDEM_HIP(c, hipMemcpyAsync(s.id, p->id, sizeof(uint64_t) * n, hipMemcpyHostToDevice, c->stream));
  std::vector<uint32_t> z32(n, 0);
  c->ids_dense = true;
  {
    std::vector<uint8_t> seen(n, 0);
    for (int64_t i = 0; i < n && c->ids_dense; i++) {
      if (p->id[i] >= (uint64_t)n || seen[p->id[i]]) c->ids_dense = false;
      else seen[p->id[i]] = 1;
    }
  }
  c->max_type = 0;
  if (p->type)
    for (int64_t i = 0; i < n; i++) c->max_type = std::max(c->max_type, p->type[i]);
  XDEM_HIP(c, hipMemcpyAsync(s.type, p->type ? p->type : z32.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice,
                             c->stream));
  XDEM_HIP(c, hipMemcpyAsync(s.group, p->group ? p->group : z32.data(), sizeof(uint32_t) * n,
                             hipMemcpyHostToDevice, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  for (auto* x : dtmp) (void)hipFree(x);
  XDEM_HIP(c, hipGetLastError());
  c->nl_valid = c->nl_prev_valid = c->perm_valid = false;
  c->ml_valid = false;  // mesh candidates (and their sphere index) refer to the previous particles
  c->pending_kick = false;
  c->wl_nw = -1;
  return XDEM_OK;
}

int xdem_particles_download(xdem_ctx* c, xdem_particles* p) {
  if (!c || !p) return XDEM_ERR_ARG;
  if (p->n < c->n) return set_err(c, XDEM_ERR_CAPACITY, "download: arrays shorter than the particle count");
  XDEM_HIP(c, hipSetDevice(c->device));
  if (c->pending_kick) {
    int rc = launch_kick(c, c->pending_dt);
    if (rc) return rc;
    c->pending_kick = false;
  }
  const int64_t n = c->n;
  if (n == 0) return XDEM_OK;
  PState& s = live(c);
  std::vector<double*> dtmp;
  auto dt = [&]() -> double* {
    double* d = nullptr;
    if (hipMalloc((void**)&d, sizeof(double) * n) != hipSuccess) return nullptr;
    dtmp.push_back(d);
    return d;
  };
  double* t[16];
  for (int k = 0; k < 16; k++) {
    t[k] = dt();
    if (!t[k]) {
      for (auto* x : dtmp) (void)hipFree(x);
      return set_err(c, XDEM_ERR_HIP, "download: staging allocation failed");
    }
  }
  hipLaunchKernelGGL(k_unpack_out, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, n, s, t[0], t[1], t[2], t[3],
                     t[4], t[5], t[6], t[7], t[8], t[9], t[10], t[11], t[12], t[13], t[14], t[15]);
  double* hosts[16] = {p->rx, p->ry, p->rz, p->radius, p->vx, p->vy, p->vz, p->mass,
                       p->vrotx, p->vroty, p->vrotz, p->qw, p->qx, p->qy, p->qz, p->homothety};
  for (int k = 0; k < 16; k++)
    if (hosts[k]) XDEM_HIP(c, hipMemcpyAsync(hosts[k], t[k], sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  auto cp = [&](double* h, const double* d) -> hipError_t {
    if (!h) return hipSuccess;
    return hipMemcpyAsync(h, d, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream);
  };
  XDEM_HIP(c, cp(p->fx, s.ax));
  XDEM_HIP(c, cp(p->fy, s.ay));
  XDEM_HIP(c, cp(p->fz, s.az));
  XDEM_HIP(c, cp(p->momx, s.mx));
  XDEM_HIP(c, cp(p->momy, s.my));
  XDEM_HIP(c, cp(p->momz, s.mz));
  XDEM_HIP(c, cp(p->arotx, s.alx));
  XDEM_HIP(c, cp(p->aroty, s.aly));
  XDEM_HIP(c, cp(p->arotz, s.alz));
  XDEM_HIP(c, cp(p->inx, s.ix));
  XDEM_HIP(c, cp(p->iny, s.iy));
  XDEM_HIP(c, cp(p->inz, s.iz));
  if (p->id) XDEM_HIP(c, hipMemcpyAsync(p->id, s.id, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, c->stream));
  if (p->type) XDEM_HIP(c, hipMemcpyAsync(p->type, s.type, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
  if (p->group)
    XDEM_HIP(c, hipMemcpyAsync(p->group, s.group, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  for (auto* x : dtmp) (void)hipFree(x);
  return XDEM_OK;
}

int64_t xdem_num_particles(const xdem_ctx* c) { return c ? c->n : -1; }

// ------------------------------------------------------------------ operators
static int flush_kick(xdem_ctx* c) {
  if (!c->pending_kick) return XDEM_OK;
  c->pending_kick = false;
  return launch_kick(c, c->pending_dt);
}

int xdem_combined_compute_prolog(xdem_ctx* c, double dt) {
  if (!c) return XDEM_ERR_ARG;
  if (c->n == 0) return XDEM_OK;
  const bool kick = c->pending_kick;
  const double kdt = c->pending_dt;
  c->pending_kick = false;
  return launch_prolog(c, dt, kick, kdt, false, 0.0, false);
}

int xdem_combined_compute_epilog(xdem_ctx* c, double dt) {
  if (!c) return XDEM_ERR_ARG;
  int rc = flush_kick(c);
  if (rc || c->n == 0) return rc;
  return launch_epilog(c, dt);
}

int xdem_reset_force_moment(xdem_ctx* c) {
  if (!c) return XDEM_ERR_ARG;
  int rc = flush_kick(c);
  if (rc || c->n == 0) return rc;
  return launch_reset(c);
}

int xdem_gravity_force(xdem_ctx* c, const double g[3]) {
  if (!c || !g) return XDEM_ERR_ARG;
  int rc = flush_kick(c);
  if (rc || c->n == 0) return rc;
  return launch_gravity(c, {g[0], g[1], g[2]});
}

int xdem_force_to_accel(xdem_ctx* c) {
  if (!c) return XDEM_ERR_ARG;
  int rc = flush_kick(c);
  if (rc || c->n == 0) return rc;
  return launch_to_accel(c);
}

int xdem_particle_displ_over(xdem_ctx* c, double threshold, int32_t* result) {
  if (!c || !result) return XDEM_ERR_ARG;
  *result = 0;
  int rc = flush_kick(c);
  if (rc || c->n == 0) return rc;
  XDEM_HIP(c, hipMemsetAsync(c->d_flags, 0, sizeof(int32_t), c->stream));
  rc = launch_displ(c, threshold * threshold);
  if (rc) return rc;
  XDEM_HIP(c, hipMemcpyAsync(c->h_flags, c->d_flags, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  *result = c->h_flags[0] ? 1 : 0;
  XDEM_HIP(c, hipMemsetAsync(c->d_flags, 0, sizeof(int32_t), c->stream));
  return XDEM_OK;
}

int xdem_move_particles(xdem_ctx* c) {
  if (!c) return XDEM_ERR_ARG;
  if (!c->have_domain) return set_err(c, XDEM_ERR_STATE, "move_particles: no domain set");
  int rc = flush_kick(c);
  if (rc) return rc;
  return do_move_particles(c);
}

int xdem_chunk_neighbors_contact(xdem_ctx* c, double rcut_inc, double nbh_dist) {
  if (!c) return XDEM_ERR_ARG;
  if (!c->have_domain) return set_err(c, XDEM_ERR_STATE, "chunk_neighbors_contact: no domain set");
  if (!(rcut_inc >= 0)) return set_err(c, XDEM_ERR_ARG, "chunk_neighbors_contact: rcut_inc must be >= 0");
  int rc = flush_kick(c);
  if (rc) return rc;
  return do_chunk_neighbors(c, rcut_inc, nbh_dist);
}

int xdem_nbh_sphere(xdem_ctx* c, double rcut_inc, int32_t symetric) {
  if (!c) return XDEM_ERR_ARG;
  if (!symetric) return set_err(c, XDEM_ERR_UNSUPPORTED, "nbh_sphere: only symetric: true is provided");
  int rc = flush_kick(c);
  if (rc) return rc;
  return do_nbh_sphere(c, rcut_inc);
}

int xdem_classify_interactions(xdem_ctx* c) {
  if (!c) return XDEM_ERR_ARG;
  if (!c->nl_valid && c->n > 0) return set_err(c, XDEM_ERR_STATE, "classify_interactions: no interaction list");
  // polyhedra: the feature interactions split into one wave per type (the contact pass runs a type per
  // wavefront). Spheres hold one interaction type per structure already: particle pairs in the
  // neighbour list, each driver's candidates in its own arrays, mesh candidates in the mesh lists.
  if (c->nshapes > 0) return poly_classify(c);
  return XDEM_OK;
}

static ContactParams to_cp(const xdem_contact_params* p) {
  ContactParams c{};
  if (p) c = {p->kn, p->kt, p->kr, p->mu, p->damp_rate, p->fc, p->dncut, p->gamma};
  return c;
}

static int check_law(xdem_ctx* c, const xdem_contact_params* p, const char* what) {
  if (!p) return XDEM_OK;
  if (p->dncut > 0 || p->fc != 0)
    return set_err(c, XDEM_ERR_UNSUPPORTED, std::string(what) + ": cohesive laws are not provided (dncut/fc)");
  if (p->gamma > 0) return set_err(c, XDEM_ERR_UNSUPPORTED, std::string(what) + ": DMT adhesion is not provided");
  if (!(p->kn >= 0) || !(p->kt >= 0) || !(p->mu >= 0) || !(p->damp_rate >= 0))
    return set_err(c, XDEM_ERR_ARG, std::string(what) + ": kn, kt, mu and damp_rate must be >= 0");
  return XDEM_OK;
}

int xdem_contact_sphere(xdem_ctx* c, const xdem_contact_params* cp, const xdem_contact_params* cpd, double dt,
                        int32_t form) {
  if (!c || !cp) return XDEM_ERR_ARG;
  int rc;
  if ((rc = check_law(c, cp, "contact_sphere config"))) return rc;
  if ((rc = check_law(c, cpd, "contact_sphere config_driver"))) return rc;
  if (c->nwalls > 0 && !cpd)
    return set_err(c, XDEM_ERR_ARG, "contact_sphere: config_driver is required when surfaces are registered");
  if (form != XDEM_FORM_FP64 && form != XDEM_FORM_FP32) return set_err(c, XDEM_ERR_ARG, "contact_sphere: bad form");
  if ((rc = flush_kick(c))) return rc;
  if (c->n == 0) return XDEM_OK;
  if (!c->nl_valid) return set_err(c, XDEM_ERR_STATE, "contact_sphere: no interaction list (run nbh_sphere)");
  return launch_force(c, to_cp(cp), to_cp(cpd), dt, form, false, false, {0, 0, 0}, false, 0);
}

static int rebuild(xdem_ctx* c, double rcut_inc) {
  int rc;
  if ((rc = do_move_particles(c))) return rc;
  if ((rc = do_chunk_neighbors(c, rcut_inc, -1.0))) return rc;
  if (c->nshapes > 0) {
    if ((rc = poly_nbh(c, rcut_inc))) return rc;
    return poly_classify(c);
  }
  return do_nbh_sphere(c, rcut_inc);
}

// one force pass of the engine: spheres (one kernel) or polyhedra (pair + gather kernels)
static int engine_force(xdem_ctx* c, const ContactParams& cp, const ContactParams& cpd, const xdem_scheme* s, V3 g,
                        bool guarded, int32_t tag) {
  if (c->nshapes > 0) return poly_force(c, cp, cpd, s->dt, true, s->gravity_on != 0, g, guarded, tag);
  return launch_force(c, cp, cpd, s->dt, s->form, true, s->gravity_on != 0, g, guarded, tag);
}

static int record_timer(xdem_ctx* c, int which, bool start) {
  KernelTimer& t = c->timers[which];
  hipEvent_t e;
  XDEM_HIP(c, hipEventCreate(&e));
  XDEM_HIP(c, hipEventRecord(e, c->stream));
  t.ev.push_back(e);
  (void)start;
  return XDEM_OK;
}

static int harvest_timers(xdem_ctx* c) {
  for (int k = 0; k < 2; k++) {
    KernelTimer& t = c->timers[k];
    for (size_t i = 0; i + 1 < t.ev.size(); i += 2) {
      float ms = 0;
      XDEM_HIP(c, hipEventSynchronize(t.ev[i + 1]));
      XDEM_HIP(c, hipEventElapsedTime(&ms, t.ev[i], t.ev[i + 1]));
      t.total_ms += ms;
      t.count++;
    }
    for (auto& e : t.ev) (void)hipEventDestroy(e);
    t.ev.clear();
  }
  return XDEM_OK;
}

int xdem_run_steps(xdem_ctx* c, const xdem_scheme* s, int64_t nsteps, xdem_run_stats* stats) {
  if (!c || !s || nsteps < 0) return XDEM_ERR_ARG;
  int rc;
  if ((rc = check_law(c, &s->cp, "scheme cp"))) return rc;
  if (c->nwalls > 0 && (rc = check_law(c, &s->cpd, "scheme cpd"))) return rc;
  if (s->form != XDEM_FORM_FP64 && s->form != XDEM_FORM_FP32) return set_err(c, XDEM_ERR_ARG, "run_steps: bad form");
  if (!(s->dt > 0)) return set_err(c, XDEM_ERR_ARG, "run_steps: dt must be > 0");
  if (c->nshapes > 0 && s->form != XDEM_FORM_FP64)
    return set_err(c, XDEM_ERR_UNSUPPORTED, "run_steps: polyhedra run in the FP64 form only");
  if (c->nshapes > 0 && c->n > 0 && c->max_type >= (uint32_t)c->nshapes)
    return set_err(c, XDEM_ERR_ARG, "run_steps: a particle type has no registered shape");
  if (stats) *stats = {0, 0, 0};
  if (c->n == 0 || nsteps == 0) return XDEM_OK;
  if (!c->nl_valid) return set_err(c, XDEM_ERR_STATE, "run_steps: run first_iteration (neighbour build) first");
  XDEM_HIP(c, hipSetDevice(c->device));
  const ContactParams cp = to_cp(&s->cp), cpd = to_cp(&s->cpd);
  const V3 g{s->gravity[0], s->gravity[1], s->gravity[2]};
  const double thr = 0.5 * s->rcut_inc;
  const int poll = s->poll > 0 ? s->poll : 16;
  const int nw = c->nwalls;
  // Drivers. A step applies push_f_v_r_driver and push_f_v_driver after the particle prolog; with
  // kinematic motions the drivers' states do not depend on the particles, so the host plans them for
  // each step of a chunk (the same arithmetic as the operators), copies the chunk's states to the device
  // once, and each force pass reads its step's state. A driver that moved by rcut_inc/2 since the last
  // neighbour build raises the rebuild trigger of its step, as a particle does. Without motion, every
  // step's push_f_v leaves the same state (a stationary surface has zero velocity), set here once.
  const bool moving = nw > 0 && c->drivers_move;
  if (nw > 0 && !moving) {
    drivers_push_f_v(c->walls, c->motion, nw);
    XDEM_HIP(c, hipMemcpyAsync(c->d_walls, c->walls, sizeof(Surface) * XDEM_MAX_SURFACES, hipMemcpyHostToDevice,
                               c->stream));
  }
  if (moving && c->cap_ring < poll) {
    if (c->h_walls_ring) (void)hipHostFree(c->h_walls_ring);
    c->h_walls_ring = nullptr;
    XDEM_HIP(c, hipHostMalloc((void**)&c->h_walls_ring, sizeof(Surface) * XDEM_MAX_SURFACES * poll,
                              hipHostMallocDefault));
    if ((rc = dev_realloc(c, &c->d_walls_ring, (int64_t)XDEM_MAX_SURFACES * poll))) return rc;
    c->cap_ring = poll;
  }
  XDEM_HIP(c, hipMemsetAsync(c->d_flags, 0, 4 * sizeof(int32_t), c->stream));
  c->trig_slot = 0;
  // spheres run a step as one kernel (force pass + kick + next prolog into the other record buffer,
  // launch_step) between a chain-start prolog and the last step of the run, which keeps the separate
  // kernels so that the run ends in the same state (pending kick flushed) as the operator sequence.
  // (Polyhedra keep separate kernels: an integrating gather that also recomputes the vertices measured
  // no faster than the separate vertex pass.)
  const bool fuse = c->nshapes == 0;
  std::vector<Core*> input(poll);   // record buffer read by the integrating pass of each step of a chunk
  std::vector<double> times(poll + 1);  // physical time at the start of each step of a chunk
  bool chain = false;               // the live records hold the prolog of the next step
  int64_t done = 0, rebuilds = 0, polls = 0;
  double t_done = c->physical_time;
  while (done < nsteps) {
    // issue up to `poll` steps. A step is either K1 (kicks of the previous pass + prolog +
    // displacement test) and K2 (force pass), or, inside a chain, one integrating pass. A force pass
    // turns a raised trigger into a stall that makes every later launch a no-op.
    const int64_t upto = std::min<int64_t>(nsteps, done + poll);
    int64_t drv_trig = -1;  // first step of the chunk whose drivers moved by rcut_inc/2
    times[0] = t_done;
    for (int64_t step = done; step < upto; step++) times[step - done + 1] = times[step - done] + s->dt;
    if (moving) {
      Surface w[XDEM_MAX_SURFACES];
      std::memcpy(w, c->walls, sizeof(w));
      for (int64_t step = done; step < upto; step++) {
        drivers_advance(w, c->motion, nw, times[step - done], s->dt);
        std::memcpy(c->h_walls_ring + (step - done) * XDEM_MAX_SURFACES, w, sizeof(w));
        if (drv_trig < 0 && drivers_displaced(w, c->walls_bk, nw, thr)) drv_trig = step;
      }
      XDEM_HIP(c, hipMemcpyAsync(c->d_walls_ring, c->h_walls_ring, sizeof(Surface) * XDEM_MAX_SURFACES * (upto - done),
                                 hipMemcpyHostToDevice, c->stream));
    }
    for (int64_t step = done; step < upto; step++) {
      const bool kick = c->pending_kick;
      const double kdt = c->pending_dt;
      const bool timed = c->timing && step % c->timing_every == 0;
      const bool poly = c->nshapes > 0;
      input[step - done] = nullptr;
      if (!chain) {
        if (timed && (rc = record_timer(c, 0, true))) return rc;
        // spheres test the centre displacement in the prolog; polyhedra test vertices (vertex_displ_over)
        if (poly) {
          if ((rc = launch_prolog_verts(c, s->dt, kick, kdt, thr * thr, true))) return rc;
        } else if ((rc = launch_prolog(c, s->dt, kick, kdt, true, thr * thr, true))) {
          return rc;
        }
        if (timed && (rc = record_timer(c, 0, false))) return rc;
        c->trig_slot = 0;
      }
      if (moving) {
        c->d_walls_pass = c->d_walls_ring + (step - done) * XDEM_MAX_SURFACES;
        if (step == drv_trig) XDEM_HIP(c, hipMemsetAsync(c->d_flags + c->trig_slot, 1, 1, c->stream));
      }
      if (timed && (rc = record_timer(c, 1, true))) return rc;
      if (fuse && step + 1 < nsteps) {
        if ((rc = launch_step(c, cp, cpd, s->dt, s->form, s->gravity_on != 0, g, thr * thr, (int32_t)step,
                              &input[step - done])))
          return rc;
        chain = true;
        c->pending_kick = false;
      } else {
        if ((rc = engine_force(c, cp, cpd, s, g, true, (int32_t)step))) return rc;
        chain = false;
        c->pending_kick = true;
        c->pending_dt = s->dt;
      }
      if (timed && (rc = record_timer(c, 1, false))) return rc;
    }
    c->d_walls_pass = nullptr;
    XDEM_HIP(c, hipMemcpyAsync(c->h_flags, c->d_flags, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipStreamSynchronize(c->stream));
    polls++;
    if (c->h_flags[1]) {
      // step `stalled` has its prolog done and found the trigger raised: put back the record buffer
      // that holds that state (integrating passes after it were no-ops), rebuild, redo its force pass
      const int64_t stalled = (int64_t)c->h_flags[2];
      Core* state = stalled >= done && stalled < upto ? input[stalled - done] : nullptr;
      if (state && live(c).core != state) std::swap(live(c).core, c->core_alt);
      if (moving && stalled >= done && stalled < upto) {
        std::memcpy(c->walls, c->h_walls_ring + (stalled - done) * XDEM_MAX_SURFACES, sizeof(c->walls));
        XDEM_HIP(c, hipMemcpyAsync(c->d_walls, c->walls, sizeof(Surface) * XDEM_MAX_SURFACES, hipMemcpyHostToDevice,
                                   c->stream));
      }
      XDEM_HIP(c, hipMemsetAsync(c->d_flags, 0, 4 * sizeof(int32_t), c->stream));
      c->trig_slot = 0;
      if ((rc = rebuild(c, s->rcut_inc))) return rc;
      rebuilds++;
      if ((rc = engine_force(c, cp, cpd, s, g, false, 0))) return rc;
      chain = false;
      c->pending_kick = true;
      c->pending_dt = s->dt;
      t_done = times[stalled - done + 1];
      done = stalled + 1;
    } else {
      if (moving) {
        std::memcpy(c->walls, c->h_walls_ring + (upto - 1 - done) * XDEM_MAX_SURFACES, sizeof(c->walls));
        XDEM_HIP(c, hipMemcpyAsync(c->d_walls, c->walls, sizeof(Surface) * XDEM_MAX_SURFACES, hipMemcpyHostToDevice,
                                   c->stream));
      }
      t_done = times[upto - done];
      done = upto;
    }
  }
  c->physical_time = t_done;
  c->trig_slot = 0;
  if ((rc = flush_kick(c))) return rc;
  if (c->timing && (rc = harvest_timers(c))) return rc;
  if (stats) *stats = {nsteps, rebuilds, polls};
  return XDEM_OK;
}

int xdem_interactions_count(xdem_ctx* c, int64_t* n_pp, int64_t* n_pw);
int xdem_interactions_dump(xdem_ctx* c, int64_t cap, uint64_t* id_i, uint64_t* id_j, uint16_t* type,
                           double* friction, int64_t* written);

int xdem_nbh_stats(xdem_ctx* c, int64_t* pairs, int64_t* slots, int32_t dims[3]) {
  if (!c) return XDEM_ERR_ARG;
  const NbrList& l = c->nl[c->nl_cur];
  if (pairs) *pairs = 0;
  if (pairs && c->nl_valid) {
    const int rc = nl_pairs(c, pairs);
    if (rc) return rc;
  }
  if (slots) *slots = c->nl_valid ? l.slots : 0;
  if (dims)
    for (int a = 0; a < 3; a++) dims[a] = c->grid.dims[a];
  return XDEM_OK;
}

int xdem_kernel_timing(xdem_ctx* c, int32_t which, int32_t reset, double* avg_ms, int64_t* count) {
  if (!c || which < 0 || which > 1) return XDEM_ERR_ARG;
  int rc = harvest_timers(c);
  if (rc) return rc;
  KernelTimer& t = c->timers[which];
  if (avg_ms) *avg_ms = t.count ? t.total_ms / (double)t.count : 0.0;
  if (count) *count = t.count;
  if (reset) {
    t.count = 0;
    t.total_ms = 0;
  }
  return XDEM_OK;
}

int xdem_set_timing(xdem_ctx* c, int32_t on) {
  if (!c) return XDEM_ERR_ARG;
  if (on < 0) return set_err(c, XDEM_ERR_ARG, "set_timing: on must be >= 0");
  c->timing = on != 0;
  c->timing_every = on > 0 ? on : 1;
  return XDEM_OK;
}

}  // extern "C"

// engine pieces shared with the surrogate driver
namespace xdem {
int engine_rebuild(xdem_ctx* c, double rcut_inc) { return rebuild(c, rcut_inc); }
int engine_force_pass(xdem_ctx* c, const xdem_scheme* s) {
  int rc;
  if ((rc = check_law(c, &s->cp, "scheme cp"))) return rc;
  if (c->nwalls > 0 && (rc = check_law(c, &s->cpd, "scheme cpd"))) return rc;
  const V3 g{s->gravity[0], s->gravity[1], s->gravity[2]};
  return engine_force(c, to_cp(&s->cp), to_cp(&s->cpd), s, g, false, 0);
}
int engine_flush_kick(xdem_ctx* c) { return flush_kick(c); }
}  // namespace xdem
