// analysis.hip — the thermodynamic simulation state, reduced on the device (the interaction counts of
// polyhedra come from poly_sim_counts in poly.hip)
// (simulation_state, src/io/simulation_state.cu:116-200, with ReduceSimulationStateFunctor of
// src/io/include/exaDEM/simulation_state.hpp:48-57 and the interaction counter of
// src/interaction/include/exaDEM/itools/counter.hpp:140-153).
//
// Per particle: kinetic energy 1/2 m v*v and rotation energy 1/2 m w*w per component (the reference
// weights the spin by the mass), mass, count. Per interaction (owner entries of the neighbour list,
// counted twice as the reference counts particle pairs under symetric: true; walls and mesh features
// once): the total,
// the active ones (overlap dn < 0 or non-zero friction history) and the smallest dn over the active
// ones (0 when none overlaps). Blocks reduce in a fixed tree order and the host adds the block
// partials in block order, so the result does not depend on scheduling.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "mesh.h"

namespace xdem {

// Block partials and results of xdem_simulation_state_noalloc, one copy per GPU, held until the process
// ends. Host code never passes `scratch` itself to a kernel or a copy: it is a host-side stand-in; the
// device address of each GPU's copy comes from hipGetSymbolAddress at context set-up
// (noalloc_scratch_init) and is kept per device in g_scratch.
constexpr int64_t kScratchDoubles = 196608;  // 1.5 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
static double* g_scratch[kMaxDevices] = {};
static std::mutex g_scratch_mu;

int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return XDEM_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return XDEM_OK;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != device) return XDEM_ERR_STATE;
  double* ptr = nullptr;
  if (hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return XDEM_ERR_HIP;
  g_scratch[device] = ptr;
  return XDEM_OK;
}

static double* scratch_of(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch[device];
}

// out[q] for q < nvals: the block partials part[k * nvals + q] folded in block order k = 0, 1, ..., as
// the host does in xdem_simulation_state (sums from 0; the last value a minimum from 0 as std::min folds
// it), so both entry points give the same bits.
__global__ __launch_bounds__(64) void k_fold_partials(const double* __restrict__ part, int64_t nblocks, int nvals,
                                                      double* __restrict__ out) {
  const int q = (int)threadIdx.x;
  if (q >= nvals) return;
  const bool is_min = q == nvals - 1;
  double acc = 0.0;
  constexpr int kAhead = 16;
  int64_t k = 0;
  for (; k + kAhead <= nblocks; k += kAhead) {
    double v[kAhead];
#pragma unroll
    for (int u = 0; u < kAhead; u++) v[u] = part[(k + u) * nvals + q];
#pragma unroll
    for (int u = 0; u < kAhead; u++) acc = is_min ? (v[u] < acc ? v[u] : acc) : acc + v[u];
  }
  for (; k < nblocks; k++) {
    const double v = part[k * nvals + q];
    acc = is_min ? (v < acc ? v : acc) : acc + v;
  }
  out[q] = acc;
}

int fold_partials(xdem_ctx* c, const double* part, int64_t nblocks, int nvals, double* out_dev) {
  hipLaunchKernelGGL(k_fold_partials, dim3(1), dim3(64), 0, c->stream, part, nblocks, nvals, out_dev);
  XDEM_HIP(c, hipGetLastError());
  return XDEM_OK;
}

constexpr int kSimVals = 10;  // ke (3), re (3), mass, interactions, active interactions, min dn

__global__ __launch_bounds__(256) void k_sim_state(int64_t n, PState s, NbrList nl, Grid g, WallList wl,
                                                   const Surface* __restrict__ walls, int nw, int64_t wstride,
                                                   MeshView mesh, MeshList ml, double* __restrict__ part) {
  __shared__ double red[kSimVals][256];
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[kSimVals];
#pragma unroll
  for (int q = 0; q < kSimVals; q++) v[q] = 0.0;
  if (x < n) {
    const Core me = s.core[x];
    const double m = me.vm.w;
    v[0] = 0.5 * me.vm.x * me.vm.x * m;
    v[1] = 0.5 * me.vm.y * me.vm.y * m;
    v[2] = 0.5 * me.vm.z * me.vm.z * m;
    v[3] = 0.5 * me.w.x * me.w.x * m;
    v[4] = 0.5 * me.w.y * me.w.y * m;
    v[5] = 0.5 * me.w.z * me.w.z * m;
    v[6] = m;
    const V3 rx{me.xr.x, me.xr.y, me.xr.z};
    const double Rx = me.xr.w;
    const uint32_t cnt = nl.cnt ? nl.cnt[x] : 0u;
    const uint64_t b = cnt ? nl.base[x >> 6] + (uint64_t)(x & 63) : 0;
    for (uint32_t k = 0; k < cnt; k++) {
      const uint32_t wd = nl.word[b + (uint64_t)k * kWave];
      if (!(wd & kOwnerBit)) continue;
      const Core py = s.core[wd & kIdxMask];
      const V3 sh = shift_vec(g, (int)((wd >> kShiftShift) & 31u));
      const V3 d{rx.x - (py.xr.x + sh.x), rx.y - (py.xr.y + sh.y), rx.z - (py.xr.z + sh.z)};
      const double dn = sqrt(dot(d, d)) - (Rx + py.xr.w);
      v[7] += 2.0;
      if (dn < 0.0 || (wd & kActiveBit)) {
        v[8] += 2.0;
        v[9] = fmin(v[9], dn < 0.0 ? dn : 0.0);
      }
    }
    if (nw > 0) {
      const uint32_t mask = wl.mask[x];
      for (int w = 0; w < nw; w++) {
        if (!((mask >> w) & 1u)) continue;
        double dn;
        V3 nrm, cpos;
        driver_detect(walls[w], rx, Rx, dn, nrm, cpos);
        const int64_t j = (int64_t)w * wstride + x;
        v[7] += 1.0;
        if (dn < 0.0 || wl.hx[j] != 0.0 || wl.hy[j] != 0.0 || wl.hz[j] != 0.0) {
          v[8] += 1.0;
          v[9] = fmin(v[9], dn < 0.0 ? dn : 0.0);
        }
      }
    }
    if (mesh.cap) {  // mesh candidates, once each
      const uint32_t mc = min(ml.cnt[x], (uint32_t)mesh.cap);
      for (uint32_t k = 0; k < mc; k++) {
        const int64_t j = (int64_t)k * mesh.stride + x;
        double dn = 0.0;
        V3 nrm, cpos;
        const bool over = mesh_detect(mesh, ml.word[j], rx, Rx, dn, nrm, cpos) && dn < 0.0;
        v[7] += 1.0;
        if (over || ml.hx[j] != 0.0 || ml.hy[j] != 0.0 || ml.hz[j] != 0.0) {
          v[8] += 1.0;
          v[9] = fmin(v[9], over ? dn : 0.0);
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < kSimVals; q++) red[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
#pragma unroll
      for (int q = 0; q < kSimVals - 1; q++) red[q][threadIdx.x] += red[q][threadIdx.x + st];
      red[kSimVals - 1][threadIdx.x] = fmin(red[kSimVals - 1][threadIdx.x], red[kSimVals - 1][threadIdx.x + st]);
    }
    __syncthreads();
  }
  if (threadIdx.x < kSimVals) part[(int64_t)blockIdx.x * kSimVals + threadIdx.x] = red[threadIdx.x][0];
}

}  // namespace xdem

using namespace xdem;

// The state of simulation_state with the block partials in `part` (the library's buffer, grown on demand,
// or the caller's fixed buffer of part_cap doubles when fixed is set). Fixed buffers fold the partials on
// the device after them and copy the results only; the library buffer copies the partials and folds them
// on the host. Both fold in block order.
static int sim_state_impl(xdem_ctx* c, xdem_sim_state* out, double* fixed, int64_t part_cap) {
  *out = {};
  const bool poly = c->nshapes > 0;
  const int64_t n = c->n;
  out->n_particles = n;
  out->volume = c->grid.L[0] * c->grid.L[1] * c->grid.L[2];
  if (n == 0) return XDEM_OK;
  const int blocks = grid_blocks(n, 256);
  double* part = fixed;
  if (fixed) {
    if ((int64_t)blocks * kSimVals + kSimVals > part_cap)
      return set_err(c, XDEM_ERR_CAPACITY, "simulation_state_noalloc: more than " +
                                               std::to_string((part_cap - kSimVals) / kSimVals * 256) +
                                               " particles");
  } else {
    if ((int64_t)blocks * kSimVals > c->cap_simpart) {
      int rc = dev_realloc(c, &c->simpart, (int64_t)blocks * kSimVals);
      if (rc) return rc;
      c->cap_simpart = (int64_t)blocks * kSimVals;
    }
    part = c->simpart;
  }
  NbrList nl{};
  if (!poly && c->nl_valid && c->hist_ready) nl = c->nl[c->nl_cur];  // polyhedra: energies only here
  const bool walls = c->nwalls > 0 && c->wl_valid && nl.cnt;
  WallList wl{};
  if (walls) wl = c->wl[c->wl_cur];
  MeshView mesh{};
  MeshList ml{};
  if (!poly && nl.cnt && c->has_mesh && c->ml_valid) {
    int rc = mesh_sync(c);
    if (rc) return rc;
    mesh = c->mesh;
    ml = c->ml[c->ml_cur];
  }
  hipLaunchKernelGGL(k_sim_state, dim3(blocks), dim3(256), 0, c->stream, n, live(c), nl, c->grid, wl, c->d_walls,
                     walls ? c->nwalls : 0, c->cap, mesh, ml, part);
  XDEM_HIP(c, hipGetLastError());
  double acc[kSimVals] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (fixed) {
    double* res = part + (int64_t)blocks * kSimVals;
    int rc = fold_partials(c, part, blocks, kSimVals, res);
    if (rc) return rc;
    XDEM_HIP(c, hipMemcpyAsync(acc, res, sizeof(acc), hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipStreamSynchronize(c->stream));
  } else {
    std::vector<double> h((size_t)blocks * kSimVals);
    XDEM_HIP(c, hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < blocks; k++) {
      for (int q = 0; q < kSimVals - 1; q++) acc[q] += h[(size_t)k * kSimVals + q];
      acc[kSimVals - 1] = std::min(acc[kSimVals - 1], h[(size_t)k * kSimVals + kSimVals - 1]);
    }
  }
  for (int a = 0; a < 3; a++) {
    out->kinetic_energy[a] = acc[a];
    out->rotation_energy[a] = acc[3 + a];
  }
  out->mass = acc[6];
  out->n_interactions = (int64_t)acc[7];
  out->n_active = (int64_t)acc[8];
  out->min_dn = acc[9];
  if (poly) {
    double tot = 0, act = 0, mn = 0;
    const int rc = poly_sim_counts(c, &tot, &act, &mn, fixed, part_cap);
    if (rc) return rc;
    out->n_interactions = (int64_t)tot;
    out->n_active = (int64_t)act;
    out->min_dn = mn;
  }
  return XDEM_OK;
}

extern "C" int xdem_simulation_state(xdem_ctx* c, xdem_sim_state* out) {
  if (!c || !out) return XDEM_ERR_ARG;
  return sim_state_impl(c, out, nullptr, 0);
}

extern "C" int xdem_simulation_state_noalloc(xdem_ctx* c, xdem_sim_state* out) {
  if (!c || !out) return XDEM_ERR_ARG;
  *out = {};
  double* ptr = scratch_of(c->device);
  if (!ptr) return set_err(c, XDEM_ERR_STATE, "simulation_state_noalloc: no scratch address for this device");
  XDEM_HIP(c, hipSetDevice(c->device));
  return sim_state_impl(c, out, ptr, kScratchDoubles);
}
