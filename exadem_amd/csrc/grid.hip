// grid.hip — rebuild path: periodic wrap, cell binning and reorder, neighbour lists,
// interaction update with friction-history carry-over, wall candidates, inspection dumps.
//
//   move_particles          migrate_cell_particles_interaction + backup_r (config_spheres.msp:14-17)
//   chunk_neighbors_contact src/grid/chunk_neighbors_contact.cpp:45-57 (DEM filter)
//   nbh_sphere              src/sphere/nbh_sphere.cu:64-309 (sym mode owner choice, history by key)
//
// Binning is a counting sort over the cells' Morton ranks (histogram, exclusive scan, scatter),
// followed by a per-cell insertion sort on particle id so that the order is deterministic. Neighbour lists are built in
// two passes (count, fill) over the 27-cell stencil into a per-wavefront ELL block; the history
// carry-over looks each new entry up among the owner's previous entries (owner id, partner id key).
#include <algorithm>
#include <cstring>
#include <unordered_map>

#include "xdem_internal.h"

namespace xdem {

__device__ inline int cell_coord(double x, double lo, double h, int dims) {
  int k = (int)floor((x - lo) / h);
  k = k < 0 ? 0 : k;
  k = k >= dims ? dims - 1 : k;
  return k;
}

__global__ __launch_bounds__(256) void k_wrap_bin(int64_t n, PState s, Grid g, const uint32_t* __restrict__ rank,
                                                  uint32_t* __restrict__ cell_of, uint32_t* __restrict__ count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  d4 xr = s.core[i].xr;
  double c[3] = {xr.x, xr.y, xr.z};
  int ci[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (g.per[a]) {
      if (c[a] < g.lo[a]) c[a] += g.L[a];
      else if (c[a] >= g.lo[a] + g.L[a]) c[a] -= g.L[a];
    }
    ci[a] = cell_coord(c[a], g.lo[a], g.h[a], g.dims[a]);
  }
  xr.x = c[0];
  xr.y = c[1];
  xr.z = c[2];
  s.core[i].xr = xr;
  const uint32_t cell = rank[((int64_t)ci[2] * g.dims[1] + ci[1]) * g.dims[0] + ci[0]];
  cell_of[i] = cell;
  atomicAdd(&count[cell], 1u);
}

__global__ __launch_bounds__(256) void k_scatter(int64_t n, const uint32_t* __restrict__ cell_of,
                                                 const uint32_t* __restrict__ start, uint32_t* __restrict__ fill,
                                                 uint32_t* __restrict__ perm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = cell_of[i];
  const uint32_t pos = start[c] + atomicAdd(&fill[c], 1u);
  perm[pos] = (uint32_t)i;
}

// order the particles of each cell by ascending id (cells hold a handful of particles)
__global__ __launch_bounds__(256) void k_cell_sort(int64_t ncells, const uint32_t* __restrict__ start,
                                                   const uint64_t* __restrict__ id, uint32_t* __restrict__ perm) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const uint32_t b = start[c], e = start[c + 1];
  for (uint32_t i = b + 1; i < e; i++) {
    const uint32_t v = perm[i];
    const uint64_t key = id[v];
    uint32_t j = i;
    while (j > b && id[perm[j - 1]] > key) {
      perm[j] = perm[j - 1];
      j--;
    }
    perm[j] = v;
  }
}

__global__ __launch_bounds__(256) void k_gather(int64_t n, const uint32_t* __restrict__ perm, PState src, PState dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t o = perm[i];
  const Core c = src.core[o];
  const d4 xr = c.xr;
  dst.core[i] = c;
  dst.ax[i] = src.ax[o]; dst.ay[i] = src.ay[o]; dst.az[i] = src.az[o];
  dst.mx[i] = src.mx[o]; dst.my[i] = src.my[o]; dst.mz[i] = src.mz[o];
  dst.alx[i] = src.alx[o]; dst.aly[i] = src.aly[o]; dst.alz[i] = src.alz[o];
  dst.ix[i] = src.ix[o]; dst.iy[i] = src.iy[o]; dst.iz[i] = src.iz[o];
  dst.id[i] = src.id[o];
  dst.type[i] = src.type[o];
  dst.group[i] = src.group[o];
  dst.bx[i] = xr.x; dst.by[i] = xr.y; dst.bz[i] = xr.z;  // backup_r
  dst.bq[i] = c.q;
}

__global__ __launch_bounds__(256) void k_max_radius(int64_t n, PState s, uint64_t* __restrict__ out_bits) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double r = 0.0;
  if (i < n) r = s.core[i].xr.w;
  // radii are positive: the IEEE bit pattern orders like the value
  unsigned long long bits = (unsigned long long)__double_as_longlong(r);
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long other = __shfl_down(bits, o, 64);
    bits = other > bits ? other : bits;
  }
  if ((threadIdx.x & 63) == 0) atomicMax((unsigned long long*)out_bits, bits);
}

int do_move_particles(xdem_ctx* c) {
  const int64_t n = c->n;
  const Grid g = c->grid;
  if (n == 0) {
    c->cells_valid = true;
    return XDEM_OK;
  }
  XDEM_HIP(c, hipMemsetAsync(c->cell_count, 0, sizeof(uint32_t) * (g.ncells + 1), c->stream));
  PState src = live(c);
  hipLaunchKernelGGL(k_wrap_bin, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, n, src, g, c->cell_rank,
                     c->cell_of, c->cell_count);
  int rc = scan_u32(c, c->cell_count, c->cell_start, g.ncells);
  if (rc) return rc;
  XDEM_HIP(c, hipMemsetAsync(c->cell_count, 0, sizeof(uint32_t) * (g.ncells + 1), c->stream));
  hipLaunchKernelGGL(k_scatter, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, n, c->cell_of, c->cell_start,
                     c->cell_count, c->perm);
  hipLaunchKernelGGL(k_cell_sort, dim3(grid_blocks(g.ncells, 256)), dim3(256), 0, c->stream, g.ncells,
                     c->cell_start, src.id, c->perm);
  PState dst = c->ps[1 - c->cur];
  hipLaunchKernelGGL(k_gather, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, n, c->perm, src, dst);
  XDEM_HIP(c, hipGetLastError());
  c->cur = 1 - c->cur;
  // the list built for the previous order can seed the history carry-over through perm
  c->nl_prev_valid = c->nl_valid;
  c->nl_valid = false;
  c->perm_valid = true;
  c->cells_valid = true;
  return XDEM_OK;
}

// ---------------------------------------------------------------- neighbour lists
struct NbhArgs {
  Grid g;
  double rcut_inc;
  double rcut2;
  const uint32_t* cell_start;  // indexed by Morton rank
  const uint32_t* rank;
  const Core* core;
  const uint64_t* id;
  int64_t n;
  int poly;  // polyhedra: owner = lower id always (nbh_polyhedron.cpp:234-236)
};

// Visits the stencil of particle x in stream order: dz, dy, dx in -1..1 (dx fastest), then the
// particles of each neighbour cell in cell order. fn(y, code) is called for each accepted partner.
template <typename F>
__device__ inline void for_each_neighbour(const NbhArgs& a, int64_t x, F&& fn) {
  const d4 px = a.core[x].xr;
  const int cx = cell_coord(px.x, a.g.lo[0], a.g.h[0], a.g.dims[0]);
  const int cy = cell_coord(px.y, a.g.lo[1], a.g.h[1], a.g.dims[1]);
  const int cz = cell_coord(px.z, a.g.lo[2], a.g.h[2], a.g.dims[2]);
  for (int dz = -1; dz <= 1; dz++) {
    int nz = cz + dz, sz = 0;
    if (nz < 0) {
      if (!a.g.per[2]) continue;
      nz += a.g.dims[2];
      sz = -1;
    } else if (nz >= a.g.dims[2]) {
      if (!a.g.per[2]) continue;
      nz -= a.g.dims[2];
      sz = 1;
    }
    for (int dy = -1; dy <= 1; dy++) {
      int ny = cy + dy, sy = 0;
      if (ny < 0) {
        if (!a.g.per[1]) continue;
        ny += a.g.dims[1];
        sy = -1;
      } else if (ny >= a.g.dims[1]) {
        if (!a.g.per[1]) continue;
        ny -= a.g.dims[1];
        sy = 1;
      }
      for (int dx = -1; dx <= 1; dx++) {
        int nx = cx + dx, sx = 0;
        if (nx < 0) {
          if (!a.g.per[0]) continue;
          nx += a.g.dims[0];
          sx = -1;
        } else if (nx >= a.g.dims[0]) {
          if (!a.g.per[0]) continue;
          nx -= a.g.dims[0];
          sx = 1;
        }
        const int code = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
        const V3 sh = shift_vec(a.g, code);
        const uint32_t cb = a.rank[((int64_t)nz * a.g.dims[1] + ny) * a.g.dims[0] + nx];
        const uint32_t b = a.cell_start[cb], e = a.cell_start[cb + 1];
        for (uint32_t y = b; y < e; y++) {
          if ((int64_t)y == x && code == kNoShift) continue;
          const d4 py = a.core[y].xr;
          const V3 img{py.x + sh.x, py.y + sh.y, py.z + sh.z};
          const V3 d{px.x - img.x, px.y - img.y, px.z - img.z};
          const double d2 = dot(d, d);
          const double rmax = px.w + py.w + a.rcut_inc;
          if (d2 > 0.0 && d2 < rmax * rmax && d2 < a.rcut2) fn(y, code);
        }
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_nbh_count(NbhArgs a, uint32_t* __restrict__ cnt) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= a.n) return;
  uint32_t k = 0;
  for_each_neighbour(a, x, [&](uint32_t, int) { k++; });
  cnt[x] = k;
}

// Work balancing: inside each group of kBalance consecutive particles (a compact Morton region),
// a stable counting sort by decreasing list length, so that the 64 lanes of a wavefront walk lists of
// nearly equal length. perm2[new] = morton index, inv2[morton] = new index. A group is one block of
// the force pass (256 particles): the block's LDS copy of its records then holds a compact region,
// where most partners are. 1M spheres, FP64 step: 189.5 / 182 / 196.5 / 217 / 212.5 us with groups of
// 128 / 256 / 512 / 1024 / 2048.
constexpr int kBalance = 256;
constexpr int kBalanceBins = 64;
__global__ __launch_bounds__(64) void k_balance(int64_t n, const uint32_t* __restrict__ cnt,
                                                uint32_t* __restrict__ perm2, uint32_t* __restrict__ inv2) {
  __shared__ uint32_t start[kBalanceBins];
  __shared__ uint8_t bin[kBalance];
  const int64_t g0 = (int64_t)blockIdx.x * kBalance;
  const int m = (int)(((g0 + kBalance) < n ? (g0 + kBalance) : n) - g0);
  for (int i = threadIdx.x; i < m; i += 64) {
    const uint32_t c = cnt[g0 + i];
    bin[i] = (uint8_t)(kBalanceBins - 1 - (c < kBalanceBins ? c : kBalanceBins - 1));  // longest lists first
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int b = 0; b < kBalanceBins; b++) start[b] = 0;
  for (int i = 0; i < m; i++) start[bin[i]]++;
  uint32_t run = (uint32_t)g0;
  for (int b = 0; b < kBalanceBins; b++) {
    const uint32_t t = start[b];
    start[b] = run;
    run += t;
  }
  for (int i = 0; i < m; i++) {
    const uint32_t pos = start[bin[i]]++;
    perm2[pos] = (uint32_t)(g0 + i);
    inv2[g0 + i] = pos;
  }
}

__global__ __launch_bounds__(256) void k_balance_apply(int64_t n, const uint32_t* __restrict__ perm2,
                                                       const uint32_t* __restrict__ cnt_m, uint32_t* __restrict__ cnt,
                                                       uint32_t* __restrict__ kmax, uint32_t* __restrict__ max_rows,
                                                       const uint32_t* __restrict__ perm_old,
                                                       uint32_t* __restrict__ perm_new) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t k = 0;
  if (x < n) {
    const uint32_t m = perm2[x];
    k = cnt_m[m];
    cnt[x] = k;
    if (perm_old) perm_new[x] = perm_old[m];
  }
  uint32_t mx = k;
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t t = __shfl_down(mx, o, 64);
    mx = t > mx ? t : mx;
  }
  const int64_t w = x >> 6;
  if ((threadIdx.x & 63) == 0 && (w << 6) < n) {
    kmax[w] = mx;
    atomicMax(max_rows, mx);
  }
}

__global__ __launch_bounds__(256) void k_nbh_fill(NbhArgs a, const uint32_t* __restrict__ inv2,
                                                  const uint64_t* __restrict__ base, uint32_t* __restrict__ word,
                                                  unsigned long long* __restrict__ pairs) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t owned = 0;
  if (x < a.n) {
    const uint32_t xn = inv2[x];
    const uint64_t b = base[xn >> 6] + (uint64_t)(xn & 63);
    const uint64_t idx = a.id[x];
    uint32_t k = 0;
    for_each_neighbour(a, x, [&](uint32_t y, int code) {
      const bool owner = a.poly ? (idx < a.id[y]) : !(idx >= a.id[y] && code == kNoShift);
      word[b + (uint64_t)k * kWave] = inv2[y] | ((uint32_t)code << kShiftShift) | (owner ? kOwnerBit : 0u);
      owned += owner ? 1u : 0u;
      k++;
    });
  }
  for (int o = 32; o > 0; o >>= 1) owned += __shfl_down(owned, o, 64);
  if ((threadIdx.x & 63) == 0 && owned) atomicAdd(pairs, (unsigned long long)owned);
}

template <typename T>
static int grow(xdem_ctx* c, T** p, int64_t count) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  XDEM_HIP(c, hipMalloc((void**)p, sizeof(T) * (size_t)std::max<int64_t>(count, 1)));
  return XDEM_OK;
}

int do_chunk_neighbors(xdem_ctx* c, double rcut_inc, double nbh_dist) {
  if (!c->cells_valid) return set_err(c, XDEM_ERR_STATE, "chunk_neighbors_contact: run move_particles first");
  const int64_t n = c->n;
  const Grid g = c->grid;
  if (n > 0 && nbh_dist <= 0) {
    // dem_rcut_max (src/field/dem_rcut_max.cpp:55-65) + nbh_dist = rcut_max + rcut_inc
    uint64_t* bits = (uint64_t*)c->d_u32tmp;
    XDEM_HIP(c, hipMemsetAsync(bits, 0, sizeof(uint64_t), c->stream));
    hipLaunchKernelGGL(k_max_radius, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, n, live(c), bits);
    uint64_t hb = 0;
    int rc = read_u64(c, bits, &hb);
    if (rc) return rc;
    double rmax;
    std::memcpy(&rmax, &hb, sizeof(double));
    nbh_dist = 2.0 * rmax + rcut_inc;
  }
  // check_rcut (src/field/check_rcut.cpp:67-118): the one-layer stencil needs cells >= nbh_dist
  for (int a = 0; a < 3; a++)
    if (g.h[a] < nbh_dist * (1.0 - 1e-12))
      return set_err(c, XDEM_ERR_ARG, "chunk_neighbors_contact: cell edge smaller than the neighbour distance");
  c->rcut_inc = rcut_inc;
  c->nbh_dist = nbh_dist;
  const int nxt = 1 - c->nl_cur;
  NbrList& L = c->nl[nxt];
  if (n == 0) {
    L.slots = 0;
    L.pairs = 0;
    c->nl_cur = nxt;
    c->nl_valid = true;
    return XDEM_OK;
  }
  NbhArgs a{g, rcut_inc, nbh_dist * nbh_dist, c->cell_start, c->cell_rank, live(c).core, live(c).id, n,
            c->nshapes > 0 ? 1 : 0};
  const int64_t nw = (n + kWave - 1) / kWave;
  uint32_t* max_rows = c->d_u32tmp + 2;
  uint32_t* cnt_m = c->cell_of;  // list lengths in Morton order (cell_of is free after the binning)
  uint32_t* perm2 = reinterpret_cast<uint32_t*>(c->sort_key);
  uint32_t* inv2 = perm2 + c->cap;
  uint32_t* perm_tmp = c->perm_tmp;
  XDEM_HIP(c, hipMemsetAsync(max_rows, 0, sizeof(uint32_t), c->stream));
  hipLaunchKernelGGL(k_nbh_count, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, a, cnt_m);
  hipLaunchKernelGGL(k_balance, dim3(grid_blocks(n, kBalance)), dim3(64), 0, c->stream, n, cnt_m, perm2, inv2);
  hipLaunchKernelGGL(k_balance_apply, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, n, perm2, cnt_m, L.cnt,
                     L.kmax, max_rows, c->perm_valid ? c->perm : nullptr, perm_tmp);
  int rc = scan_u64_from_u32(c, L.kmax, L.base, nw, (uint32_t)kWave);
  if (rc) return rc;
  XDEM_HIP(c, hipMemcpyAsync(c->h_u32tmp + 2, max_rows, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  uint64_t slots = 0;
  if ((rc = read_u64(c, L.base + nw, &slots))) return rc;
  c->nl_max_rows = (int)c->h_u32tmp[2];
  if ((int64_t)slots > L.cap_slots) {
    const int64_t cap = (int64_t)(slots * 1.25) + 4096;
    if ((rc = grow(c, &L.word, cap)) || (rc = grow(c, &L.hx, cap)) || (rc = grow(c, &L.hy, cap)) ||
        (rc = grow(c, &L.hz, cap)) || (rc = grow(c, &L.mhx, cap)) || (rc = grow(c, &L.mhy, cap)) ||
        (rc = grow(c, &L.mhz, cap)) || (rc = grow(c, &L.fx32, cap)) || (rc = grow(c, &L.fy32, cap)) ||
        (rc = grow(c, &L.fz32, cap)) || (rc = grow(c, &L.mfx32, cap)) || (rc = grow(c, &L.mfy32, cap)) ||
        (rc = grow(c, &L.mfz32, cap)))
      return rc;
    L.cap_slots = cap;
  }
  L.slots = (int64_t)slots;
  unsigned long long* pairs = (unsigned long long*)c->d_u32tmp;
  XDEM_HIP(c, hipMemsetAsync(pairs, 0, sizeof(uint64_t), c->stream));
  hipLaunchKernelGGL(k_nbh_fill, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, a, inv2, L.base, L.word,
                     pairs);
  // move every field into the balanced order; compose the history permutation (new -> before binning)
  PState dst = c->ps[1 - c->cur];
  hipLaunchKernelGGL(k_gather, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, n, perm2, live(c), dst);
  if (c->perm_valid)
    XDEM_HIP(c, hipMemcpyAsync(c->perm, perm_tmp, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, c->stream));
  XDEM_HIP(c, hipGetLastError());
  c->cur = 1 - c->cur;
  c->cells_valid = false;  // cell ranges refer to the Morton order
  uint64_t hp = 0;
  if ((rc = read_u64(c, (const uint64_t*)pairs, &hp))) return rc;
  L.pairs = (int64_t)hp;
  c->nl_cur = nxt;
  c->nl_valid = true;
  c->hist_ready = false;
  c->nl_epoch++;
  return XDEM_OK;
}

// ---------------------------------------------------------------- interaction update (nbh_sphere)
struct CarryArgs {
  int64_t n;
  NbrList nl;        // new list
  NbrList ol;        // previous list (valid when have_old)
  const uint32_t* perm;
  int have_old;
};

template <typename T>
struct HistView {
  T *x, *y, *z, *mx, *my, *mz;
};
__host__ __device__ inline HistView<double> hist_view(const NbrList& l, double*) {
  return {l.hx, l.hy, l.hz, l.mhx, l.mhy, l.mhz};
}
__host__ __device__ inline HistView<float> hist_view(const NbrList& l, float*) {
  return {l.fx32, l.fy32, l.fz32, l.mfx32, l.mfy32, l.mfz32};
}

// One thread per (particle, entry). New entries start from zero history; an entry whose key
// (owner id, partner id) matches an active owner entry of the previous list inherits its friction
// and moment (interaction_manager.hpp:88-112 with placeholder_interaction.hpp:398-408).
template <typename T>
__global__ __launch_bounds__(256) void k_hist_carry(CarryArgs a) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= a.n) return;
  const HistView<T> nh = hist_view(a.nl, (T*)nullptr), oh = hist_view(a.ol, (T*)nullptr);
  const uint32_t cnt = a.nl.cnt[x];
  const uint64_t b = a.nl.base[x >> 6] + (uint64_t)(x & 63);
  for (uint32_t k = 0; k < cnt; k++) {
    const uint64_t s = b + (uint64_t)k * kWave;
    uint32_t wd = a.nl.word[s] & ~kActiveBit;
    T hx = 0, hy = 0, hz = 0, mx = 0, my = 0, mz = 0;
    bool act = false;
    if (a.have_old) {
      const uint32_t y = wd & kIdxMask;
      const bool owner = (wd & kOwnerBit) != 0;
      const uint32_t o = owner ? (uint32_t)x : y;
      const uint32_t p = owner ? y : (uint32_t)x;
      const uint32_t oo = a.perm[o], op = a.perm[p];
      const uint32_t ocnt = a.ol.cnt[oo];
      const uint64_t ob = a.ol.base[oo >> 6] + (uint64_t)(oo & 63);
      for (uint32_t k2 = 0; k2 < ocnt; k2++) {
        const uint64_t s2 = ob + (uint64_t)k2 * kWave;
        const uint32_t w2 = a.ol.word[s2];
        if ((w2 & kOwnerBit) && (w2 & kActiveBit) && (w2 & kIdxMask) == op) {
          hx = oh.x[s2]; hy = oh.y[s2]; hz = oh.z[s2];
          mx = oh.mx[s2]; my = oh.my[s2]; mz = oh.mz[s2];
          act = true;
          break;
        }
      }
    }
    nh.x[s] = hx; nh.y[s] = hy; nh.z[s] = hz;
    nh.mx[s] = mx; nh.my[s] = my; nh.mz[s] = mz;
    a.nl.word[s] = wd | (act ? kActiveBit : 0u);
  }
}

// Driver candidates (the driver's filter with rad + rcut_inc, e.g. Surface::filter, surface.hpp:320-324)
// and carry-over.
__global__ __launch_bounds__(256) void k_wall_carry(int64_t n, int nw, const Surface* __restrict__ walls, PState s,
                                                    double rcut_inc, WallList nwl, WallList owl,
                                                    const uint32_t* __restrict__ perm, int have_old, int64_t stride) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const d4 xr = s.core[x].xr;
  const V3 p{xr.x, xr.y, xr.z};
  const double rc = xr.w + rcut_inc;
  uint8_t mask = 0;
  const uint32_t o = have_old ? perm[x] : 0u;
  const uint8_t omask = have_old ? owl.mask[o] : (uint8_t)0;
  for (int w = 0; w < nw; w++) {
    const Surface& sf = walls[w];
    const int64_t j = (int64_t)w * stride + x;
    double hx = 0, hy = 0, hz = 0, mx = 0, my = 0, mz = 0;
    if (driver_filter(sf, p, rc)) {
      mask |= (uint8_t)(1u << w);
      if ((omask >> w) & 1u) {
        const int64_t jo = (int64_t)w * stride + o;
        hx = owl.hx[jo]; hy = owl.hy[jo]; hz = owl.hz[jo];
        mx = owl.mhx[jo]; my = owl.mhy[jo]; mz = owl.mhz[jo];
      }
    }
    nwl.hx[j] = hx; nwl.hy[j] = hy; nwl.hz[j] = hz;
    nwl.mhx[j] = mx; nwl.mhy[j] = my; nwl.mhz[j] = mz;
  }
  nwl.mask[x] = mask;
}

int do_nbh_sphere(xdem_ctx* c, double rcut_inc) {
  if (!c->nl_valid) return set_err(c, XDEM_ERR_STATE, "nbh_sphere: run chunk_neighbors_contact first");
  const int64_t n = c->n;
  if (n == 0) {
    c->hist_ready = true;
    return XDEM_OK;
  }
  const bool have_old = c->nl_prev_valid && c->perm_valid;
  CarryArgs a{n, c->nl[c->nl_cur], c->nl[1 - c->nl_cur], c->perm, have_old ? 1 : 0};
  if (c->hist_fp32)
    hipLaunchKernelGGL(k_hist_carry<float>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, a);
  else
    hipLaunchKernelGGL(k_hist_carry<double>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, a);
  XDEM_HIP(c, hipGetLastError());
  // walls
  int rc = alloc_walls(c, c->nwalls);
  if (rc) return rc;
  if (c->nwalls > 0) {
    const bool wold = have_old && c->wl_valid;
    const int nxt = 1 - c->wl_cur;
    hipLaunchKernelGGL(k_wall_carry, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, n, c->nwalls, c->d_walls,
                       live(c), rcut_inc, c->wl[nxt], c->wl[c->wl_cur], c->perm, wold ? 1 : 0, c->cap);
    XDEM_HIP(c, hipGetLastError());
    c->wl_cur = nxt;
    c->wl_valid = true;
  }
  if ((rc = mesh_carry(c, rcut_inc, have_old))) return rc;
  c->nl_prev_valid = false;
  c->perm_valid = false;
  c->hist_ready = true;
  return XDEM_OK;
}

// ---------------------------------------------------------------- inspection
__global__ __launch_bounds__(256) void k_count_pp(int64_t n, NbrList l, unsigned long long* __restrict__ out) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t own = 0;
  if (x < n) {
    const uint64_t b = l.base[x >> 6] + (uint64_t)(x & 63);
    for (uint32_t k = 0; k < l.cnt[x]; k++) own += (l.word[b + (uint64_t)k * kWave] & kOwnerBit) ? 1u : 0u;
  }
  for (int o = 32; o > 0; o >>= 1) own += __shfl_down(own, o, 64);
  if ((threadIdx.x & 63) == 0 && own) atomicAdd(out, (unsigned long long)own);
}

}  // namespace xdem

using namespace xdem;

extern "C" int xdem_interactions_count(xdem_ctx* c, int64_t* n_pp, int64_t* n_pw) {
  if (!c) return XDEM_ERR_ARG;
  if (n_pp) *n_pp = 0;
  if (n_pw) *n_pw = 0;
  if (!c->nl_valid || c->n == 0) return XDEM_OK;
  if (n_pp) *n_pp = c->nl[c->nl_cur].pairs;
  if (n_pw && c->nwalls > 0 && c->wl_valid) {
    std::vector<uint8_t> m(c->n);
    XDEM_HIP(c, hipMemcpyAsync(m.data(), c->wl[c->wl_cur].mask, c->n, hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipStreamSynchronize(c->stream));
    int64_t k = 0;
    for (auto v : m) k += __builtin_popcount(v);
    *n_pw = k;
  }
  return XDEM_OK;
}

// The interaction rows of xdem_interactions_dump with, per row, the friction history or (moments) the
// rolling-moment history.
static int interactions_rows(xdem_ctx* c, int64_t cap, uint64_t* id_i, uint64_t* id_j, uint16_t* type,
                             double* friction, bool moments, int64_t* written) {
  if (written) *written = 0;
  if (!c->nl_valid || c->n == 0) return XDEM_OK;
  if (c->pending_kick) {
    // history does not depend on the pending kicks; nothing to flush for the dump
  }
  const int64_t n = c->n;
  const NbrList& L = c->nl[c->nl_cur];
  const int64_t nw = (n + kWave - 1) / kWave;
  std::vector<uint32_t> cnt(n), word(L.slots);
  std::vector<uint64_t> base(nw + 1), ids(n);
  std::vector<double> hx(L.slots), hy(L.slots), hz(L.slots);
  XDEM_HIP(c, hipMemcpyAsync(cnt.data(), L.cnt, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(base.data(), L.base, sizeof(uint64_t) * (nw + 1), hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(ids.data(), live(c).id, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, c->stream));
  std::vector<float> fx(c->hist_fp32 ? L.slots : 0), fy(fx.size()), fz(fx.size());
  if (L.slots > 0) {
    XDEM_HIP(c, hipMemcpyAsync(word.data(), L.word, sizeof(uint32_t) * L.slots, hipMemcpyDeviceToHost, c->stream));
    if (c->hist_fp32) {
      XDEM_HIP(c, hipMemcpyAsync(fx.data(), moments ? L.mfx32 : L.fx32, sizeof(float) * L.slots, hipMemcpyDeviceToHost,
                                 c->stream));
      XDEM_HIP(c, hipMemcpyAsync(fy.data(), moments ? L.mfy32 : L.fy32, sizeof(float) * L.slots, hipMemcpyDeviceToHost,
                                 c->stream));
      XDEM_HIP(c, hipMemcpyAsync(fz.data(), moments ? L.mfz32 : L.fz32, sizeof(float) * L.slots, hipMemcpyDeviceToHost,
                                 c->stream));
    } else {
      XDEM_HIP(c, hipMemcpyAsync(hx.data(), moments ? L.mhx : L.hx, sizeof(double) * L.slots, hipMemcpyDeviceToHost,
                                 c->stream));
      XDEM_HIP(c, hipMemcpyAsync(hy.data(), moments ? L.mhy : L.hy, sizeof(double) * L.slots, hipMemcpyDeviceToHost,
                                 c->stream));
      XDEM_HIP(c, hipMemcpyAsync(hz.data(), moments ? L.mhz : L.hz, sizeof(double) * L.slots, hipMemcpyDeviceToHost,
                                 c->stream));
    }
  }
  std::vector<uint8_t> mask;
  std::vector<double> whx, why, whz;
  if (c->nwalls > 0 && c->wl_valid) {
    const WallList& W = c->wl[c->wl_cur];
    const int64_t m = (int64_t)c->nwalls * c->cap;
    mask.resize(n);
    whx.resize(m); why.resize(m); whz.resize(m);
    XDEM_HIP(c, hipMemcpyAsync(mask.data(), W.mask, n, hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipMemcpyAsync(whx.data(), moments ? W.mhx : W.hx, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipMemcpyAsync(why.data(), moments ? W.mhy : W.hy, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipMemcpyAsync(whz.data(), moments ? W.mhz : W.hz, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
  }
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  if (c->hist_fp32)
    for (int64_t i = 0; i < L.slots; i++) {
      hx[i] = fx[i];
      hy[i] = fy[i];
      hz[i] = fz[i];
    }
  int64_t k = 0;
  for (int64_t x = 0; x < n; x++) {
    if (!mask.empty())
      for (int w = 0; w < c->nwalls; w++)
        if ((mask[x] >> w) & 1) {
          if (k >= cap) return set_err(c, XDEM_ERR_CAPACITY, "interactions_dump: cap too small");
          const int64_t j = (int64_t)w * c->cap + x;
          id_i[k] = ids[x];
          id_j[k] = (uint64_t)w;
          type[k] = 5;
          friction[3 * k] = whx[j];
          friction[3 * k + 1] = why[j];
          friction[3 * k + 2] = whz[j];
          k++;
        }
    const uint64_t b = base[x >> 6] + (uint64_t)(x & 63);
    for (uint32_t e = 0; e < cnt[x]; e++) {
      const uint64_t s = b + (uint64_t)e * kWave;
      const uint32_t wd = word[s];
      if (!(wd & kOwnerBit)) continue;
      if (k >= cap) return set_err(c, XDEM_ERR_CAPACITY, "interactions_dump: cap too small");
      id_i[k] = ids[x];
      id_j[k] = ids[wd & kIdxMask];
      type[k] = 0;
      friction[3 * k] = hx[s];
      friction[3 * k + 1] = hy[s];
      friction[3 * k + 2] = hz[s];
      k++;
    }
  }
  if (written) *written = k;
  return XDEM_OK;
}

extern "C" int xdem_interactions_dump(xdem_ctx* c, int64_t cap, uint64_t* id_i, uint64_t* id_j, uint16_t* type,
                                      double* friction, int64_t* written) {
  if (!c || !id_i || !id_j || !type || !friction) return XDEM_ERR_ARG;
  return interactions_rows(c, cap, id_i, id_j, type, friction, false, written);
}

extern "C" int xdem_interactions_moments(xdem_ctx* c, int64_t cap, double* moment, int64_t* written) {
  if (!c || !moment || cap < 0) return XDEM_ERR_ARG;
  std::vector<uint64_t> a((size_t)std::max<int64_t>(cap, 1)), b(a.size());
  std::vector<uint16_t> t(a.size());
  return interactions_rows(c, cap, a.data(), b.data(), t.data(), moment, true, written);
}

// Full (both-direction) neighbour streams: particle id, neighbour id, image code, owner flag.
extern "C" int xdem_streams_dump(xdem_ctx* c, int64_t cap, uint64_t* a, uint64_t* b, int32_t* code, uint8_t* owner,
                                 int64_t* written) {
  if (!c || !a || !b || !code) return XDEM_ERR_ARG;
  if (written) *written = 0;
  if (!c->nl_valid || c->n == 0) return XDEM_OK;
  const int64_t n = c->n;
  const NbrList& L = c->nl[c->nl_cur];
  const int64_t nw = (n + kWave - 1) / kWave;
  std::vector<uint32_t> cnt(n), word(L.slots);
  std::vector<uint64_t> base(nw + 1), ids(n);
  XDEM_HIP(c, hipMemcpyAsync(cnt.data(), L.cnt, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(base.data(), L.base, sizeof(uint64_t) * (nw + 1), hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(ids.data(), live(c).id, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, c->stream));
  if (L.slots > 0)
    XDEM_HIP(c, hipMemcpyAsync(word.data(), L.word, sizeof(uint32_t) * L.slots, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  int64_t k = 0;
  for (int64_t x = 0; x < n; x++) {
    const uint64_t bb = base[x >> 6] + (uint64_t)(x & 63);
    for (uint32_t e = 0; e < cnt[x]; e++) {
      if (k >= cap) return set_err(c, XDEM_ERR_CAPACITY, "streams_dump: cap too small");
      const uint32_t wd = word[bb + (uint64_t)e * kWave];
      a[k] = ids[x];
      b[k] = ids[wd & kIdxMask];
      code[k] = (int32_t)((wd >> kShiftShift) & 31u);
      if (owner) owner[k] = (wd & kOwnerBit) ? 1 : 0;
      k++;
    }
  }
  if (written) *written = k;
  return XDEM_OK;
}

// Restart with interaction history: the history of the current lists' interactions is set from a
// saved interaction list, matched by key, as the reference's restart hands dumped histories to the
// interaction update (extract_history / update, interaction_manager.hpp:88-112 and 205-217).
namespace {
struct IKey {
  uint64_t i, j;
  uint16_t t;
  bool operator==(const IKey& o) const { return i == o.i && j == o.j && t == o.t; }
};
struct IKeyHash {
  size_t operator()(const IKey& k) const {
    uint64_t h = k.i * 0x9e3779b97f4a7c15ull;
    h ^= k.j + 0x7f4a7c159e3779b9ull + (h << 6) + (h >> 2);
    h ^= (uint64_t)k.t + (h << 6) + (h >> 2);
    return (size_t)h;
  }
};
}  // namespace

extern "C" int xdem_interactions_load(xdem_ctx* c, int64_t m, const uint64_t* id_i, const uint64_t* id_j,
                                      const uint16_t* type, const double* friction, const double* moment,
                                      int64_t* matched) {
  if (!c || m < 0 || (m > 0 && (!id_i || !id_j || !type || !friction))) return XDEM_ERR_ARG;
  if (matched) *matched = 0;
  if (c->nshapes > 0) return set_err(c, XDEM_ERR_UNSUPPORTED, "interactions_load: sphere interactions only");
  if (!c->nl_valid || !c->hist_ready) return set_err(c, XDEM_ERR_STATE, "interactions_load: run nbh_sphere first");
  int rc = set_history_form(c, false);
  if (rc) return rc;
  const int64_t n = c->n;
  if (n == 0) return XDEM_OK;
  std::unordered_map<IKey, int64_t, IKeyHash> saved;
  saved.reserve((size_t)m * 2);
  for (int64_t k = 0; k < m; k++)
    if (!saved.emplace(IKey{id_i[k], id_j[k], type[k]}, k).second)
      return set_err(c, XDEM_ERR_ARG, "interactions_load: duplicate interaction key");
  NbrList& L = c->nl[c->nl_cur];
  const int64_t nw = (n + kWave - 1) / kWave;
  std::vector<uint32_t> cnt(n), word(L.slots);
  std::vector<uint64_t> base(nw + 1), ids(n);
  XDEM_HIP(c, hipMemcpyAsync(cnt.data(), L.cnt, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(base.data(), L.base, sizeof(uint64_t) * (nw + 1), hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(ids.data(), live(c).id, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, c->stream));
  if (L.slots > 0)
    XDEM_HIP(c, hipMemcpyAsync(word.data(), L.word, sizeof(uint32_t) * L.slots, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  // both copies of a pair (the owner entry and the partner's entry) take the pair's history
  std::vector<double> h[6];
  for (auto& a : h) a.assign((size_t)L.slots, 0.0);
  int64_t hit = 0;
  for (int64_t x = 0; x < n; x++) {
    const uint64_t b = base[x >> 6] + (uint64_t)(x & 63);
    for (uint32_t e = 0; e < cnt[x]; e++) {
      const uint64_t s = b + (uint64_t)e * kWave;
      const uint32_t wd = word[s];
      const uint64_t py = ids[wd & kIdxMask];
      const bool own = (wd & kOwnerBit) != 0;
      const auto it = saved.find(own ? IKey{ids[x], py, 0} : IKey{py, ids[x], 0});
      bool act = false;
      if (it != saved.end()) {
        const int64_t k = it->second;
        for (int q = 0; q < 3; q++) {
          h[q][s] = friction[3 * k + q];
          h[3 + q][s] = moment ? moment[3 * k + q] : 0.0;
          act = act || h[q][s] != 0.0 || h[3 + q][s] != 0.0;
        }
        if (own) hit++;
      }
      word[s] = act ? (wd | kActiveBit) : (wd & ~kActiveBit);
    }
  }
  if (L.slots > 0) {
    double* dst[6] = {L.hx, L.hy, L.hz, L.mhx, L.mhy, L.mhz};
    for (int q = 0; q < 6; q++)
      XDEM_HIP(c, hipMemcpyAsync(dst[q], h[q].data(), sizeof(double) * L.slots, hipMemcpyHostToDevice, c->stream));
    XDEM_HIP(c, hipMemcpyAsync(L.word, word.data(), sizeof(uint32_t) * L.slots, hipMemcpyHostToDevice, c->stream));
  }
  if (c->nwalls > 0 && c->wl_valid) {
    WallList& W = c->wl[c->wl_cur];
    const int64_t mw = (int64_t)c->nwalls * c->cap;
    std::vector<uint8_t> mask(n);
    XDEM_HIP(c, hipMemcpyAsync(mask.data(), W.mask, n, hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<double> wh[6];
    for (auto& a : wh) a.assign((size_t)mw, 0.0);
    for (int64_t x = 0; x < n; x++)
      for (int w = 0; w < c->nwalls; w++) {
        if (!((mask[x] >> w) & 1)) continue;
        const auto it = saved.find(IKey{ids[x], (uint64_t)w, 5});
        if (it == saved.end()) continue;
        const int64_t k = it->second, j = (int64_t)w * c->cap + x;
        for (int q = 0; q < 3; q++) {
          wh[q][j] = friction[3 * k + q];
          wh[3 + q][j] = moment ? moment[3 * k + q] : 0.0;
        }
        hit++;
      }
    double* dst[6] = {W.hx, W.hy, W.hz, W.mhx, W.mhy, W.mhz};
    for (int q = 0; q < 6; q++)
      XDEM_HIP(c, hipMemcpyAsync(dst[q], wh[q].data(), sizeof(double) * mw, hipMemcpyHostToDevice, c->stream));
  }
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  if (moment)
    for (int64_t k = 0; k < 3 * m && !c->kr_used; k++) c->kr_used = moment[k] != 0.0;
  if (matched) *matched = hit;
  return XDEM_OK;
}
