// integrate.hip — per-particle velocity-Verlet pieces (HBM-streaming kernels).
//
//   prolog  combined_compute_prolog (src/numericalscheme/combined_compute_prolog.cu:35-47,94-116):
//           r += v dt + a dt^2/2 ; v += a dt/2 ; Q <- normalize(Q + dot(Q,w) dt) ; w += alpha dt/2
//   epilog  combined_compute_epilog (src/numericalscheme/combined_compute_epilog.cu:35-46,94-113):
//           alpha = Q (I^-1 (Q* mom - w_b x (I w_b))) ; w += alpha dt/2 ; v += a dt/2
//   kick    the two epilog updates alone, when alpha was already produced by the fused force pass
// The fused engine runs "kick(n-1) + prolog(n) + displacement test" as one kernel: the
// per-particle operation sequence is the same as running the operators one after the other.
#include "step.h"

namespace xdem {

// quat * vec3 (src/shape/include/exaDEM/type/quat.hpp:84-91)
__device__ inline V3 qrot(Q4 q, V3 v) {
  const V3 qv{q.x, q.y, q.z};
  V3 uv = cross(qv, v);
  V3 uuv = cross(qv, uv);
  uv = uv * (2.0 * q.w);
  uuv = uuv * 2.0;
  return v + uv + uuv;
}

// PushToAngularAccelerationFunctor (src/numericalscheme/include/exaDEM/angular_acceleration.hpp:28-42)
__device__ inline V3 angular_accel(Q4 Q, V3 mom, V3 vrot, V3 I) {
  const Q4 Qi{Q.w, -Q.x, -Q.y, -Q.z};
  const V3 om = qrot(Qi, vrot);
  const V3 M = qrot(Qi, mom);
  const V3 dom{(M.x - (I.z - I.y) * om.y * om.z) / I.x, (M.y - (I.x - I.z) * om.z * om.x) / I.y,
               (M.z - (I.y - I.x) * om.x * om.y) / I.z};
  return qrot(Q, dom);
}

// The 128-byte records of a block's 256 particles are moved between HBM and LDS with fully
// coalesced 16-byte accesses (a record-per-lane access pattern would touch 64 lines per
// instruction). LDS rows are padded to 144 bytes to spread the per-lane reads over the banks.
constexpr int kStageBS = 256;
constexpr int kStageRow = 9;  // uint4 per record row in LDS (8 used + 1 pad)

__device__ inline void stage_in(const Core* __restrict__ core, int64_t first, int cnt, uint4* lds) {
  const uint4* src = reinterpret_cast<const uint4*>(core + first);
  for (int j = threadIdx.x; j < cnt * 8; j += kStageBS) lds[(j >> 3) * kStageRow + (j & 7)] = src[j];
}
__device__ inline void stage_out(Core* __restrict__ core, int64_t first, int cnt, const uint4* lds) {
  uint4* dst = reinterpret_cast<uint4*>(core + first);
  for (int j = threadIdx.x; j < cnt * 8; j += kStageBS) dst[j] = lds[(j >> 3) * kStageRow + (j & 7)];
}

// LDS rows are only 16-byte aligned: move the 32-byte fields with memcpy
__device__ inline d4 lds_get(const uint4* p) {
  d4 v;
  __builtin_memcpy(&v, p, sizeof(d4));
  return v;
}
__device__ inline void lds_put(uint4* p, const d4& v) { __builtin_memcpy(p, &v, sizeof(d4)); }

template <bool KICK, bool DISPL, bool GUARD>
__global__ __launch_bounds__(kStageBS) void k_prolog(int64_t n, PState s, double dt, double kdt_2, double thr2,
                                                     int32_t* __restrict__ flags) {
  if (GUARD && flags[1]) return;
  __shared__ uint4 lds[kStageBS * kStageRow];
  const int64_t first = (int64_t)blockIdx.x * kStageBS;
  const int cnt = (int)((n - first) < kStageBS ? (n - first) : kStageBS);
  stage_in(s.core, first, cnt, lds);
  __syncthreads();
  const int t = threadIdx.x;
  const int64_t i = first + t;
  if (t < cnt) {
    uint4* row = lds + t * kStageRow;
    d4 xr = lds_get(row);
    d4 vm = lds_get(row + 2);
    d4 w = lds_get(row + 4);
    d4 qq = lds_get(row + 6);
    kick_prolog<KICK>(xr, vm, w, qq, s.ax[i], s.ay[i], s.az[i], s.alx[i], s.aly[i], s.alz[i], dt, kdt_2);
    lds_put(row, xr);
    lds_put(row + 2, vm);
    lds_put(row + 4, w);
    lds_put(row + 6, qq);
    if (DISPL) {
      const double dx = xr.x - s.bx[i], dy = xr.y - s.by[i], dz = xr.z - s.bz[i];
      if (dx * dx + dy * dy + dz * dz >= thr2) flags[0] = 1;
    }
  }
  __syncthreads();
  stage_out(s.core, first, cnt, lds);
}

// The prolog of polyhedra (k_prolog without the centre test) fused with compute_vertices and
// vertex_displ_over: the world vertices of the new pose go to vtx, and the trigger (flags[0]) is raised
// when a vertex moved by sqrt(thr2) or more from its place at the last rebuild. The same values as
// k_prolog followed by k_poly_vertices, which reads the pose back from the record.
template <bool KICK, bool GUARD>
__global__ __launch_bounds__(kStageBS) void k_prolog_verts(int64_t n, PState s, double dt, double kdt_2, double thr2,
                                                           int32_t* __restrict__ flags,
                                                           const xdem_shape* __restrict__ shapes, d4* __restrict__ vtx) {
  if (GUARD && flags[1]) return;
  __shared__ uint4 lds[kStageBS * kStageRow];
  const int64_t first = (int64_t)blockIdx.x * kStageBS;
  const int cnt = (int)((n - first) < kStageBS ? (n - first) : kStageBS);
  stage_in(s.core, first, cnt, lds);
  __syncthreads();
  const int t = threadIdx.x;
  const int64_t i = first + t;
  if (t < cnt) {
    uint4* row = lds + t * kStageRow;
    d4 xr = lds_get(row);
    d4 vm = lds_get(row + 2);
    d4 w = lds_get(row + 4);
    d4 qq = lds_get(row + 6);
    kick_prolog<KICK>(xr, vm, w, qq, s.ax[i], s.ay[i], s.az[i], s.alx[i], s.aly[i], s.alz[i], dt, kdt_2);
    lds_put(row, xr);
    lds_put(row + 2, vm);
    lds_put(row + 4, w);
    lds_put(row + 6, qq);
    const xdem_shape& sh = shapes[s.type[i]];
    const V3 p{xr.x, xr.y, xr.z};
    const double h = w.w;
    const V3 pb{s.bx[i], s.by[i], s.bz[i]};
    const d4 qb = s.bq[i];
    bool moved = false;
    for (int k = 0; k < sh.nv; k++) {
      const V3 v = get_vertex(sh, k, p, h, qq);
      vtx[i * XDEM_MAX_SHAPE_VERTICES + k] = {v.x, v.y, v.z, 0.0};
      const V3 a = get_vertex(sh, k, pb, h, qb);
      const V3 d = v - a;
      if (dot(d, d) >= thr2) moved = true;
    }
    if (moved) flags[0] = 1;
  }
  __syncthreads();
  stage_out(s.core, first, cnt, lds);
}

__global__ __launch_bounds__(256) void k_epilog(int64_t n, PState s, double dt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dt_2 = dt * 0.5;
  const d4 qq = s.core[i].q;
  d4 w = s.core[i].w;
  d4 vm = s.core[i].vm;
  const V3 al = angular_accel(Q4{qq.x, qq.y, qq.z, qq.w}, V3{s.mx[i], s.my[i], s.mz[i]}, V3{w.x, w.y, w.z},
                              V3{s.ix[i], s.iy[i], s.iz[i]});
  s.alx[i] = al.x;
  s.aly[i] = al.y;
  s.alz[i] = al.z;
  w.x = w.x + al.x * dt_2;
  w.y = w.y + al.y * dt_2;
  w.z = w.z + al.z * dt_2;
  vm.x = vm.x + s.ax[i] * dt_2;
  vm.y = vm.y + s.ay[i] * dt_2;
  vm.z = vm.z + s.az[i] * dt_2;
  s.core[i].w = w;
  s.core[i].vm = vm;
}

__global__ __launch_bounds__(256) void k_kick(int64_t n, PState s, double dt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dt_2 = dt * 0.5;
  d4 w = s.core[i].w;
  d4 vm = s.core[i].vm;
  w.x = w.x + s.alx[i] * dt_2;
  w.y = w.y + s.aly[i] * dt_2;
  w.z = w.z + s.alz[i] * dt_2;
  vm.x = vm.x + s.ax[i] * dt_2;
  vm.y = vm.y + s.ay[i] * dt_2;
  vm.z = vm.z + s.az[i] * dt_2;
  s.core[i].w = w;
  s.core[i].vm = vm;
}

__global__ __launch_bounds__(256) void k_reset(int64_t n, PState s) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  s.ax[i] = 0.0; s.ay[i] = 0.0; s.az[i] = 0.0;
  s.mx[i] = 0.0; s.my[i] = 0.0; s.mz[i] = 0.0;
}

// GravityForceFunctor (src/forcefield/include/exaDEM/forcefield/gravity_force.hpp:38-46)
__global__ __launch_bounds__(256) void k_gravity(int64_t n, PState s, double gx, double gy, double gz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double m = s.core[i].vm.w;
  s.ax[i] = s.ax[i] + gx * m;
  s.ay[i] = s.ay[i] + gy * m;
  s.az[i] = s.az[i] + gz * m;
}

// ForceToAccelFunctor (src/numericalscheme/include/exaDEM/force_to_accel.hpp:26-33)
__global__ __launch_bounds__(256) void k_to_accel(int64_t n, PState s) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double m = s.core[i].vm.w;
  s.ax[i] = s.ax[i] / m;
  s.ay[i] = s.ay[i] / m;
  s.az[i] = s.az[i] / m;
}

__global__ __launch_bounds__(256) void k_displ(int64_t n, PState s, double thr2, int32_t* __restrict__ flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const d4 xr = s.core[i].xr;
  const double dx = xr.x - s.bx[i], dy = xr.y - s.by[i], dz = xr.z - s.bz[i];
  if (dx * dx + dy * dy + dz * dz >= thr2) flags[0] = 1;
}

#define LAUNCH1(kern, ...)                                                                          \
  do {                                                                                              \
    hipLaunchKernelGGL(kern, dim3(grid_blocks(c->n, 256)), dim3(256), 0, c->stream, __VA_ARGS__); \
    XDEM_HIP(c, hipGetLastError());                                                                 \
  } while (0)
static_assert(kStageBS == 256, "prolog launches use 256-thread blocks");

int launch_prolog(xdem_ctx* c, double dt, bool with_kick, double kick_dt, bool displ, double thr2, bool guarded) {
  PState s = live(c);
  const double kdt_2 = kick_dt * 0.5;
  const int64_t n = c->n;
  // guarded launches (fused engine) become no-ops once a stall is raised, with or without the
  // centre-displacement test (polyhedra test vertices in a kernel of their own)
  if (with_kick) {
    if (displ && guarded) LAUNCH1((k_prolog<true, true, true>), n, s, dt, kdt_2, thr2, c->d_flags);
    else if (displ) LAUNCH1((k_prolog<true, true, false>), n, s, dt, kdt_2, thr2, c->d_flags);
    else if (guarded) LAUNCH1((k_prolog<true, false, true>), n, s, dt, kdt_2, thr2, c->d_flags);
    else LAUNCH1((k_prolog<true, false, false>), n, s, dt, kdt_2, thr2, c->d_flags);
  } else {
    if (displ && guarded) LAUNCH1((k_prolog<false, true, true>), n, s, dt, kdt_2, thr2, c->d_flags);
    else if (displ) LAUNCH1((k_prolog<false, true, false>), n, s, dt, kdt_2, thr2, c->d_flags);
    else if (guarded) LAUNCH1((k_prolog<false, false, true>), n, s, dt, kdt_2, thr2, c->d_flags);
    else LAUNCH1((k_prolog<false, false, false>), n, s, dt, kdt_2, thr2, c->d_flags);
  }
  c->cells_valid = false;  // positions moved: the binning no longer matches them
  return XDEM_OK;
}

int launch_prolog_verts(xdem_ctx* c, double dt, bool with_kick, double kick_dt, double thr2, bool guarded) {
  if (c->n == 0) return XDEM_OK;
  int rc = poly_vtx_alloc(c);
  if (rc) return rc;
  PState s = live(c);
  const double kdt_2 = kick_dt * 0.5;
  const int64_t n = c->n;
  if (with_kick) {
    if (guarded) LAUNCH1((k_prolog_verts<true, true>), n, s, dt, kdt_2, thr2, c->d_flags, c->d_shapes, c->vtx);
    else LAUNCH1((k_prolog_verts<true, false>), n, s, dt, kdt_2, thr2, c->d_flags, c->d_shapes, c->vtx);
  } else {
    if (guarded) LAUNCH1((k_prolog_verts<false, true>), n, s, dt, kdt_2, thr2, c->d_flags, c->d_shapes, c->vtx);
    else LAUNCH1((k_prolog_verts<false, false>), n, s, dt, kdt_2, thr2, c->d_flags, c->d_shapes, c->vtx);
  }
  c->cells_valid = false;
  return XDEM_OK;
}

int launch_epilog(xdem_ctx* c, double dt) {
  LAUNCH1(k_epilog, c->n, live(c), dt);
  return XDEM_OK;
}
int launch_kick(xdem_ctx* c, double dt) {
  if (c->n == 0) return XDEM_OK;
  LAUNCH1(k_kick, c->n, live(c), dt);
  return XDEM_OK;
}
int launch_reset(xdem_ctx* c) {
  LAUNCH1(k_reset, c->n, live(c));
  return XDEM_OK;
}
int launch_gravity(xdem_ctx* c, V3 g) {
  LAUNCH1(k_gravity, c->n, live(c), g.x, g.y, g.z);
  return XDEM_OK;
}
int launch_to_accel(xdem_ctx* c) {
  LAUNCH1(k_to_accel, c->n, live(c));
  return XDEM_OK;
}
int launch_displ(xdem_ctx* c, double thr2) {
  LAUNCH1(k_displ, c->n, live(c), thr2, c->d_flags);
  return XDEM_OK;
}

}  // namespace xdem
