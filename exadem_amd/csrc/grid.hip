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

// Periodic wrap and cell histogram; also the largest radius (dem_rcut_max, src/field/dem_rcut_max.cpp:55-65)
// as IEEE bits (radii are positive, so the bit pattern orders like the value): one atomic per block.
__global__ __launch_bounds__(256) void k_wrap_bin(int64_t n, PState s, Grid g, const uint32_t* __restrict__ rank,
                                                  uint32_t* __restrict__ cell_of, uint32_t* __restrict__ count,
                                                  unsigned long long* __restrict__ rmax_bits) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long bits = 0;
  if (i < n) {
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
    bits = (unsigned long long)__double_as_longlong(xr.w);
    xr.x = c[0];
    xr.y = c[1];
    xr.z = c[2];
    s.core[i].xr = xr;
    const uint32_t cell = rank[((int64_t)ci[2] * g.dims[1] + ci[1]) * g.dims[0] + ci[0]];
    cell_of[i] = cell;
    atomicAdd(&count[cell], 1u);
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long other = __shfl_down(bits, o, 64);
    bits = other > bits ? other : bits;
  }
  __shared__ unsigned long long wmax[4];
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = bits;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = wmax[0];
    for (int w = 1; w < 4; w++) m = wmax[w] > m ? wmax[w] : m;
    if (m) atomicMax(rmax_bits, m);
  }
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

// Positions (with radius) and ids in Morton order, for the neighbour build; the full particle state is
// reordered once, by the neighbour build, into its final (balanced) order.
__global__ __launch_bounds__(256) void k_morton_pack(int64_t n, const uint32_t* __restrict__ perm, const Core* core,
                                                     const uint64_t* __restrict__ id, d4* __restrict__ mxr,
                                                     uint64_t* __restrict__ mid) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t o = perm[i];
  mxr[i] = core[o].xr;
  mid[i] = id[o];
}

int do_move_particles(xdem_ctx* c) {
  // backup_drivers (config_move_particles.msp:149-153): the reference of driver_displ_over
  for (int k = 0; k < XDEM_MAX_SURFACES; k++) c->walls_bk[k] = c->walls[k];
  const int64_t n = c->n;
  const Grid g = c->grid;
  if (n == 0) {
    c->cells_valid = true;
    return XDEM_OK;
  }
  XDEM_HIP(c, hipMemsetAsync(c->cell_count, 0, sizeof(uint32_t) * (g.ncells + 1), c->stream));
  XDEM_HIP(c, hipMemsetAsync(c->d_rmax, 0, sizeof(uint64_t), c->stream));
  PState src = live(c);
  hipLaunchKernelGGL(k_wrap_bin, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, n, src, g, c->cell_rank,
                     c->cell_of, c->cell_count, (unsigned long long*)c->d_rmax);
  int rc = scan_u32(c, c->cell_count, c->cell_start, g.ncells);
  if (rc) return rc;
  XDEM_HIP(c, hipMemsetAsync(c->cell_count, 0, sizeof(uint32_t) * (g.ncells + 1), c->stream));
  hipLaunchKernelGGL(k_scatter, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, n, c->cell_of, c->cell_start,
                     c->cell_count, c->perm);
  hipLaunchKernelGGL(k_cell_sort, dim3(grid_blocks(g.ncells, 256)), dim3(256), 0, c->stream, g.ncells,
                     c->cell_start, src.id, c->perm);
  if (c->cap_morton < c->cap) {
    if ((rc = dev_realloc(c, &c->mxr, c->cap)) || (rc = dev_realloc(c, &c->mid, c->cap))) return rc;
    c->cap_morton = c->cap;
  }
  hipLaunchKernelGGL(k_morton_pack, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, n, c->perm, src.core, src.id,
                     c->mxr, c->mid);
  XDEM_HIP(c, hipGetLastError());
  // perm: Morton position -> current index. The particles keep their order until the neighbour build
  // moves them into the final one; the list built for the current order seeds the history carry-over.
  c->nl_prev_valid = c->nl_valid;
  c->nl_valid = false;
  c->perm_valid = true;
  c->cells_valid = true;
  return XDEM_OK;
}

// ---------------------------------------------------------------- neighbour lists
// Candidate lists in one pass, one wavefront per row segment of kRowCells cells along x (same y and z).
// Each particle's candidates come from its 27-cell stencil, in stream order: stencil cell (dz, dy, dx in
// -1..1, dx fastest; periodic images wrap with their image code), then cell order. The segment's cells
// share one stencil of 9 (kRowCells + 2) cells: lane l looks up cell (dz, dy, dxx) with l = ((dz + 1) 3
// + dy + 1) (kRowCells + 2) + dxx + 1, and a wavefront prefix sum numbers the candidates t = 0 .. T-1 in
// lane order. For a particle in column q of the segment, the candidates of lanes with dxx in q-1..q+1 are
// exactly its stencil, and in its stream order. Its particles are tested 64 candidates at a time with
// the DEM filter of chunk_neighbors_contact.cpp:45-57 (0 < d^2 < (r_x + r_y + rcut_inc)^2 and
// d^2 < nbh_dist^2, strict), and a ballot + prefix count writes the accepted ones to x's row of `tmp` in
// candidate order. Sharing the stencil lookups and the dependent loads over a segment cuts the
// per-wavefront latency chains per cell by kRowCells. Rows hold at most `cap` entries; the count is exact,
// so a longer row is detected and the pass repeated with a larger cap.
// Entry: partner index (Morton order) | image code << 25 | owner << 31 (owner: lower id, or a periodic
// image, as nbh_sphere's sym mode keeps it; polyhedra: lower id always, nbh_polyhedron.cpp:234-236).
constexpr int kRowCells = 3;
struct CellNbhArgs {
  Grid g;
  double rcut_inc;
  double rcut2;
  const uint32_t* cell_start;  // indexed by Morton rank
  const uint32_t* rank;        // linear cell -> Morton rank
  const d4* xr;                // positions and radii, Morton order
  const uint64_t* id;          // ids, Morton order
  int64_t ncells;
  int poly;
  uint32_t* tmp;  // [n * cap]
  uint32_t cap;
  uint32_t* cnt;  // [n] list lengths
};

__device__ inline bool stencil_cell(const Grid& g, int c, int d, int a, int& out, int& sh) {
  int v = c + d;
  sh = 0;
  if (v < 0) {
    if (!g.per[a]) return false;
    v += g.dims[a];
    sh = -1;
  } else if (v >= g.dims[a]) {
    if (!g.per[a]) return false;
    v -= g.dims[a];
    sh = 1;
  }
  out = v;
  return true;
}

// value of lane j (wavefront-uniform j) in every lane
__device__ inline uint32_t bcast_u32(uint32_t v, int j) { return (uint32_t)__builtin_amdgcn_readlane((int)v, j); }
__device__ inline double bcast_f64(double v, int j) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = bcast_u32((uint32_t)b, j), hi = bcast_u32((uint32_t)(b >> 32), j);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ inline uint64_t bcast_u64(uint64_t v, int j) {
  return ((uint64_t)bcast_u32((uint32_t)(v >> 32), j) << 32) | bcast_u32((uint32_t)v, j);
}

// Each chunk of 64 candidates is loaded once and tested against up to 64 particles of the segment (lane j
// holding particle j's position and running count), so a candidate record is read once per segment.
__global__ __launch_bounds__(256) void k_nbh_rows(CellNbhArgs a) {
  constexpr int NC = kRowCells + 2;  // stencil columns dxx = -1 .. kRowCells
  constexpr int NS = 9 * NC;         // stencil lanes
  static_assert(NS <= 64, "one stencil cell per lane");
  const int lane = (int)threadIdx.x & 63;
  const Grid& g = a.g;
  const int nxb = (g.dims[0] + kRowCells - 1) / kRowCells;
  const int64_t w = (int64_t)blockIdx.x * 4 + ((int)threadIdx.x >> 6);
  if (w >= (int64_t)nxb * g.dims[1] * g.dims[2]) return;
  const int cy = (int)((w / nxb) % g.dims[1]);
  const int cz = (int)(w / ((int64_t)nxb * g.dims[1]));
  const int x0 = (int)(w % nxb) * kRowCells;
  const int nq = min(kRowCells, g.dims[0] - x0);  // columns of the segment
  // the segment's own cells: lane q < nq holds column q's particle range
  uint32_t ob = 0, on = 0;
  if (lane < nq) {
    const uint32_t r = a.rank[((int64_t)cz * g.dims[1] + cy) * g.dims[0] + x0 + lane];
    ob = a.cell_start[r];
    on = a.cell_start[r + 1] - ob;
  }
  uint32_t cbeg[kRowCells], cend[kRowCells];  // wavefront-uniform: first particle, end of the prefix
  uint32_t own = 0;
#pragma unroll
  for (int q = 0; q < kRowCells; q++) {
    cbeg[q] = bcast_u32(ob, q);
    own += bcast_u32(on, q);
    cend[q] = own;
  }
  if (own == 0) return;
  uint32_t sstart = 0, scnt = 0;
  int scode = kNoShift, scol = -2;
  if (lane < NS) {
    const int dz = lane / (3 * NC) - 1, dy = (lane / NC) % 3 - 1, dxx = lane % NC - 1;
    int nx, ny, nz, sx, sy, sz;
    // columns past the segment's last neighbour column stay empty (x0 + dxx <= dims: one wrap at most)
    if (dxx <= nq && stencil_cell(g, cz, dz, 2, nz, sz) && stencil_cell(g, cy, dy, 1, ny, sy) &&
        stencil_cell(g, x0, dxx, 0, nx, sx)) {
      const uint32_t cb = a.rank[((int64_t)nz * g.dims[1] + ny) * g.dims[0] + nx];
      sstart = a.cell_start[cb];
      scnt = a.cell_start[cb + 1] - sstart;
      scode = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
      scol = dxx;
    }
  }
  uint32_t inc = scnt;  // inclusive prefix over the stencil lanes (lanes NS..63 then hold the total)
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  const uint32_t total = __shfl(inc, NS - 1, 64);
  const uint64_t below = (1ull << lane) - 1ull;
  for (uint32_t g0 = 0; g0 < own; g0 += 64) {
    const int ng = (int)min(64u, own - g0);
    // lane j: particle g0 + j of the segment, its column and index
    const uint32_t tj = g0 + (uint32_t)lane;
    int myq = 0;
    uint32_t myx = cbeg[0] + tj;
#pragma unroll
    for (int q = 1; q < kRowCells; q++)
      if (tj >= cend[q - 1]) {
        myq = q;
        myx = cbeg[q] + (tj - cend[q - 1]);
      }
    d4 mine{0, 0, 0, 0};
    uint64_t myid = 0;
    if (lane < ng) {
      mine = a.xr[myx];
      myid = a.id[myx];
    }
    uint32_t run = 0;  // lane j: entries of particle g0 + j so far
    for (uint32_t t0 = 0; t0 < total; t0 += 64) {
      const uint32_t t = t0 + (uint32_t)lane;
      // stencil lane holding candidate t: the first lane whose inclusive count exceeds t (branch-free
      // binary search over lanes 0..63, every lane taking part)
      int pos = 0;
#pragma unroll
      for (int step = 32; step >= 1; step >>= 1) {
        const uint32_t v = __shfl(inc, pos + step - 1, 64);
        pos += v <= t ? step : 0;
      }
      const int sc = pos < NS ? pos : NS - 1;
      const uint32_t first = __shfl(inc, sc, 64) - __shfl(scnt, sc, 64);
      const uint32_t ystart = __shfl(sstart, sc, 64);
      const int code = __shfl(scode, sc, 64);
      const int col = __shfl(scol, sc, 64);
      const bool valid = t < total;
      const uint32_t y = ystart + (t - first);
      d4 py{0, 0, 0, 0};
      uint64_t idy = 0;
      if (valid) {
        py = a.xr[y];
        idy = a.id[y];
      }
      const V3 sh = shift_vec(g, code);
      const V3 img{py.x + sh.x, py.y + sh.y, py.z + sh.z};
      for (int j = 0; j < ng; j++) {
        const double pxx = bcast_f64(mine.x, j), pxy = bcast_f64(mine.y, j), pxz = bcast_f64(mine.z, j);
        const double pxw = bcast_f64(mine.w, j);
        const uint32_t x = bcast_u32(myx, j);
        const int qj = (int)bcast_u32((uint32_t)myq, j);
        bool acc = false;
        if (valid && col >= qj - 1 && col <= qj + 1 && !(y == x && code == kNoShift)) {
          const V3 d{pxx - img.x, pxy - img.y, pxz - img.z};
          const double d2 = dot(d, d);
          const double rmax = pxw + py.w + a.rcut_inc;
          acc = d2 > 0.0 && d2 < rmax * rmax && d2 < a.rcut2;
        }
        const uint64_t m = __ballot(acc);
        if (m) {
          const uint32_t base = bcast_u32(run, j);
          const uint64_t idx = bcast_u64(myid, j);
          if (acc) {
            const bool owner = a.poly ? (idx < idy) : !(idx >= idy && code == kNoShift);
            const uint32_t p = base + (uint32_t)__popcll(m & below);
            if (p < a.cap)
              a.tmp[(uint64_t)x * a.cap + p] = y | ((uint32_t)code << kShiftShift) | (owner ? kOwnerBit : 0u);
          }
          run += lane == j ? (uint32_t)__popcll(m) : 0u;
        }
      }
    }
    if (lane < ng) a.cnt[myx] = run;
  }
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

// List lengths in the balanced order, the per-wave row count of the ELL blocks, and the permutation from
// the balanced order to the particles' current order (perm_new = perm_old o perm2).
__global__ __launch_bounds__(256) void k_balance_apply(int64_t n, const uint32_t* __restrict__ perm2,
                                                       const uint32_t* __restrict__ cnt_m, uint32_t* __restrict__ cnt,
                                                       uint32_t* __restrict__ kmax,
                                                       const uint32_t* __restrict__ perm_old,
                                                       uint32_t* __restrict__ perm_new) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t k = 0;
  if (x < n) {
    const uint32_t m = perm2[x];
    k = cnt_m[m];
    cnt[x] = k;
    perm_new[x] = perm_old[m];
  }
  uint32_t mx = k;
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t t = __shfl_down(mx, o, 64);
    mx = t > mx ? t : mx;
  }
  const int64_t w = x >> 6;
  if ((threadIdx.x & 63) == 0 && (w << 6) < n) kmax[w] = mx;
}

// out[0] = max of v[0..m) (one block)
__global__ __launch_bounds__(1024) void k_max_u32(int64_t m, const uint32_t* __restrict__ v, uint32_t* __restrict__ out) {
  __shared__ uint32_t red[16];
  uint32_t mx = 0;
  for (int64_t i = threadIdx.x; i < m; i += blockDim.x) mx = v[i] > mx ? v[i] : mx;
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t t = __shfl_down(mx, o, 64);
    mx = t > mx ? t : mx;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) mx = red[w] > mx ? red[w] : mx;
    out[0] = mx;
  }
}

// out[0] = sum of v[0..m) (one block)
__global__ __launch_bounds__(1024) void k_sum_u32(int64_t m, const uint32_t* __restrict__ v,
                                                  unsigned long long* __restrict__ out) {
  __shared__ unsigned long long red[16];
  unsigned long long sm = 0;
  for (int64_t i = threadIdx.x; i < m; i += blockDim.x) sm += v[i];
  for (int o = 32; o > 0; o >>= 1) sm += __shfl_down(sm, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sm;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) sm += red[w];
    out[0] = sm;
  }
}

// Rows of tmp (Morton order) into the per-wavefront ELL blocks of the balanced order: particle xn takes
// the row of perm2[xn]; partner indices are mapped to the balanced order through inv2.
__global__ __launch_bounds__(256) void k_nbh_emit(int64_t n, const uint32_t* __restrict__ perm2,
                                                  const uint32_t* __restrict__ inv2, const uint32_t* __restrict__ tmp,
                                                  uint32_t cap, const uint32_t* __restrict__ cnt,
                                                  const uint64_t* __restrict__ base,
                                                  const uint32_t* __restrict__ kmax, uint32_t* __restrict__ word,
                                                  uint32_t* __restrict__ sx, uint32_t* __restrict__ blk_owned) {
  const int64_t xn = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t owned = 0;
  if (xn < n) {
    const uint32_t* row = tmp + (uint64_t)perm2[xn] * cap;
    const uint32_t k = cnt[xn];
    const uint64_t b = base[xn >> 6] + (uint64_t)(xn & 63);
    const uint32_t rows = kmax[xn >> 6];
    for (uint32_t e = 0; e < rows; e++) sx[b + (uint64_t)e * kWave] = e < k ? (uint32_t)xn : 0xffffffffu;
    uint32_t e = 0;
    for (; e + 4 <= k; e += 4) {  // four rows' words and partner indices in flight
      uint32_t w[4], q[4];
#pragma unroll
      for (int u = 0; u < 4; u++) w[u] = row[e + u];
#pragma unroll
      for (int u = 0; u < 4; u++) q[u] = inv2[w[u] & kIdxMask];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        word[b + (uint64_t)(e + u) * kWave] = q[u] | (w[u] & ~kIdxMask);
        owned += w[u] >> 31;
      }
    }
    for (; e < k; e++) {
      const uint32_t w = row[e];
      word[b + (uint64_t)e * kWave] = inv2[w & kIdxMask] | (w & ~kIdxMask);
      owned += w >> 31;
    }
  } else if ((xn >> 6) == ((n - 1) >> 6)) {
    // lanes past the last particle in the last wave: their slots of the ELL block are padding
    const uint64_t b = base[xn >> 6] + (uint64_t)(xn & 63);
    const uint32_t rows = kmax[xn >> 6];
    for (uint32_t e = 0; e < rows; e++) sx[b + (uint64_t)e * kWave] = 0xffffffffu;
  }
  // owner entries of the block (pairs of the half list), summed by k_sum_u32
  __shared__ uint32_t red[4];
  for (int o = 32; o > 0; o >>= 1) owned += __shfl_down(owned, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = owned;
  __syncthreads();
  if (threadIdx.x == 0) blk_owned[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
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
  int rc;
  if (n > 0 && nbh_dist <= 0) {
    // dem_rcut_max (src/field/dem_rcut_max.cpp:55-65): the largest radius, reduced by move_particles'
    // binning pass; nbh_dist = rcut_max + rcut_inc
    // binning pass; nbh_dist = rcut_max + rcut_inc. Radii change only at upload, which keeps their
    // maximum on the host, so the reduction is read back only without it.
    double rmax = c->rmax_host;
    if (rmax < 0) {
      uint64_t hb = 0;
      if ((rc = read_u64(c, c->d_rmax, &hb))) return rc;
      std::memcpy(&rmax, &hb, sizeof(double));
    }
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
  const int64_t nw = (n + kWave - 1) / kWave;
  uint32_t* max_rows = c->d_u32tmp + 2;
  uint32_t* cnt_m = c->cell_of;  // list lengths in Morton order (cell_of is free after the binning)
  uint32_t* perm2 = reinterpret_cast<uint32_t*>(c->sort_key);
  uint32_t* inv2 = perm2 + c->cap;
  uint32_t* perm_tmp = c->perm_tmp;
  if (c->nbh_tmp_cap == 0 || c->nbh_tmp_n < c->cap) {
    const uint32_t rows = c->nbh_tmp_cap ? c->nbh_tmp_cap : 64u;
    if ((rc = grow(c, &c->nbh_tmp, (int64_t)rows * c->cap))) return rc;
    c->nbh_tmp_cap = rows;
    c->nbh_tmp_n = c->cap;
  }
  CellNbhArgs a{g, rcut_inc, nbh_dist * nbh_dist, c->cell_start, c->cell_rank, c->mxr, c->mid,
                g.ncells, c->nshapes > 0 ? 1 : 0, c->nbh_tmp, c->nbh_tmp_cap, cnt_m};
  const int64_t nseg = (int64_t)((g.dims[0] + kRowCells - 1) / kRowCells) * g.dims[1] * g.dims[2];
  const dim3 cgrid((uint32_t)((nseg + 3) / 4));
  hipLaunchKernelGGL(k_nbh_rows, cgrid, dim3(256), 0, c->stream, a);
  hipLaunchKernelGGL(k_balance, dim3(grid_blocks(n, kBalance)), dim3(64), 0, c->stream, n, cnt_m, perm2, inv2);
  hipLaunchKernelGGL(k_balance_apply, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, n, perm2, cnt_m, L.cnt,
                     L.kmax, c->perm, perm_tmp);
  hipLaunchKernelGGL(k_max_u32, dim3(1), dim3(1024), 0, c->stream, nw, L.kmax, max_rows);
  if ((rc = scan_u64_from_u32(c, L.kmax, L.base, nw, (uint32_t)kWave))) return rc;
  XDEM_HIP(c, hipMemcpyAsync(c->h_u32tmp + 2, max_rows, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  uint64_t slots = 0;
  if ((rc = read_u64(c, L.base + nw, &slots))) return rc;
  c->nl_max_rows = (int)c->h_u32tmp[2];
  if ((uint32_t)c->nl_max_rows > c->nbh_tmp_cap) {
    // a row outgrew the candidate buffer: redo the pass with rows of the next power of two
    uint32_t rows = 64;
    while (rows < (uint32_t)c->nl_max_rows) rows <<= 1;
    if ((rc = grow(c, &c->nbh_tmp, (int64_t)rows * c->cap))) return rc;
    c->nbh_tmp_cap = rows;
    c->nbh_tmp_n = c->cap;
    a.tmp = c->nbh_tmp;
    a.cap = rows;
    hipLaunchKernelGGL(k_nbh_rows, cgrid, dim3(256), 0, c->stream, a);
  }
  if ((int64_t)slots > L.cap_slots) {
    const int64_t cap = (int64_t)(slots * 1.25) + 4096;
    if ((rc = grow(c, &L.word, cap)) || (rc = grow(c, &L.sx, cap)) || (rc = grow(c, &L.hx, cap)) ||
        (rc = grow(c, &L.hy, cap)) ||
        (rc = grow(c, &L.hz, cap)) || (rc = grow(c, &L.mhx, cap)) || (rc = grow(c, &L.mhy, cap)) ||
        (rc = grow(c, &L.mhz, cap)) || (rc = grow(c, &L.fx32, cap)) || (rc = grow(c, &L.fy32, cap)) ||
        (rc = grow(c, &L.fz32, cap)) || (rc = grow(c, &L.mfx32, cap)) || (rc = grow(c, &L.mfy32, cap)) ||
        (rc = grow(c, &L.mfz32, cap)))
      return rc;
    L.cap_slots = cap;
  }
  L.slots = (int64_t)slots;
  unsigned long long* pairs = (unsigned long long*)(c->d_pairs + nxt);
  const int64_t nblk = grid_blocks(n, 256);
  if (nblk > c->cap_blk) {
    if ((rc = grow(c, &c->blk_u32, nblk))) return rc;
    c->cap_blk = nblk;
  }
  hipLaunchKernelGGL(k_nbh_emit, dim3((uint32_t)nblk), dim3(256), 0, c->stream, n, perm2, inv2, c->nbh_tmp,
                     c->nbh_tmp_cap, L.cnt, L.base, L.kmax, L.word, L.sx, c->blk_u32);
  hipLaunchKernelGGL(k_sum_u32, dim3(1), dim3(1024), 0, c->stream, nblk, c->blk_u32, pairs);
  // every field moves once, from the current order into the balanced one (perm_tmp: balanced ->
  // current); perm keeps that map for the history carry-over
  PState dst = c->ps[1 - c->cur];
  hipLaunchKernelGGL(k_gather, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, n, perm_tmp, live(c), dst);
  XDEM_HIP(c, hipMemcpyAsync(c->perm, perm_tmp, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, c->stream));
  XDEM_HIP(c, hipGetLastError());
  c->cur = 1 - c->cur;
  c->cells_valid = false;  // cell ranges refer to the Morton order
  L.pairs = -1;  // on the device until asked for (nl_pairs)
  c->nl_cur = nxt;
  c->nl_valid = true;
  c->hist_ready = false;
  c->nl_epoch++;
  return XDEM_OK;
}

// ---------------------------------------------------------------- interaction update (nbh_sphere)

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

// History carry-over by scatter from the previous list: thread per previous particle, for each of its
// owner entries with history (interaction_manager.hpp:88-112 extracts the active ones; update() copies
// the payload to the new interaction with the same key, placeholder_interaction.hpp:398-408). The key
// is (owner id, partner id): the new owner entry of the same direction takes the history, and so does
// the partner's entry of the pair when it is not an owner entry (its copy of the pair). A new entry has
// at most one such source, so every target is written by one thread. Entries without a source keep
// the active bit clear, which is what every reader of the history checks; their history is not
// written.
__global__ __launch_bounds__(256) void k_invert(int64_t n, const uint32_t* __restrict__ perm,
                                                uint32_t* __restrict__ inv) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < n) inv[perm[x]] = (uint32_t)x;
}

// slot of the entry of particle x with partner y in list l, or ~0ull. kh: a rank to try first (the
// entry's rank in the previous list, right when nothing moved in front of it); only when a partner
// appears at most once per list (hint_ok), since the first match is the one wanted.
__device__ inline uint64_t find_entry(const NbrList& l, uint32_t x, uint32_t y, uint32_t kh = ~0u) {
  const uint32_t cnt = l.cnt[x];
  const uint64_t b = l.base[x >> 6] + (uint64_t)(x & 63);
  if (kh < cnt && (l.word[b + (uint64_t)kh * kWave] & kIdxMask) == y) return b + (uint64_t)kh * kWave;
  for (uint32_t k = 0; k < cnt; k += 4) {  // four words in flight per step
    uint32_t w[4];
#pragma unroll
    for (int u = 0; u < 4; u++) w[u] = k + u < cnt ? l.word[b + (uint64_t)(k + u) * kWave] : 0u;
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (k + u < cnt && (w[u] & kIdxMask) == y) return b + (uint64_t)(k + u) * kWave;
  }
  return ~0ull;
}

// One thread per slot of the previous list; only owner entries with history do work.
template <typename T, bool KR>
__global__ __launch_bounds__(256) void k_hist_scatter(int64_t oslots, NbrList nl, NbrList ol,
                                                      const uint32_t* __restrict__ inv, int hint_ok) {
  const int64_t s2 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s2 >= oslots) return;
  const uint32_t ox = ol.sx[s2];
  if (ox == 0xffffffffu) return;
  const uint32_t w2 = ol.word[s2];
  if (!(w2 & kOwnerBit) || !(w2 & kActiveBit)) return;
  const HistView<T> nh = hist_view(nl, (T*)nullptr), oh = hist_view(ol, (T*)nullptr);
  const uint32_t nx = inv[ox], ny = inv[w2 & kIdxMask];
  const T hx = oh.x[s2], hy = oh.y[s2], hz = oh.z[s2];
  T mx = 0, my = 0, mz = 0;
  if (KR) {
    mx = oh.mx[s2];
    my = oh.my[s2];
    mz = oh.mz[s2];
  }
  // moment histories are meaningful only once a rolling stiffness is used (use_kr clears them then)
  // rank of this entry in ox's previous list (slot = base + lane + rank * 64)
  const uint32_t kh = hint_ok ? (uint32_t)((uint64_t)(s2 - (int64_t)ol.base[ox >> 6] - (int64_t)(ox & 63)) / kWave) : ~0u;
  const uint64_t t1 = find_entry(nl, nx, ny, kh);
  if (t1 != ~0ull && (nl.word[t1] & kOwnerBit)) {
    nh.x[t1] = hx; nh.y[t1] = hy; nh.z[t1] = hz;
    if (KR) {
      nh.mx[t1] = mx; nh.my[t1] = my; nh.mz[t1] = mz;
    }
    nl.word[t1] |= kActiveBit;
  }
  const uint64_t t2 = find_entry(nl, ny, nx);
  if (t2 != ~0ull && !(nl.word[t2] & kOwnerBit)) {
    nh.x[t2] = hx; nh.y[t2] = hy; nh.z[t2] = hz;
    if (KR) {
      nh.mx[t2] = mx; nh.my[t2] = my; nh.mz[t2] = mz;
    }
    nl.word[t2] |= kActiveBit;
  }
}

// The same carry-over as a gather, one thread per slot of the new list (its writes coalesce), when a
// partner appears at most once per list (hint_ok). For new entry s = (nx -> ny) with nx, ny at old
// indices ox, oy: an owner entry takes the history of the old entry (ox -> oy) when that was an owner
// entry with history (the scatter's first target); a non-owner entry takes that of the old owner entry
// (oy -> ox) (the scatter's second target). When the old (ox -> oy) was itself a non-owner entry, it is
// the copy of (oy -> ox), written by every pass with the same values and active bit, and is read in
// its place; only a pair whose image changed at this rebuild (old (ox -> oy) an owner) looks up
// (oy -> ox) in oy's old list.
template <typename T, bool KR>
__global__ __launch_bounds__(256) void k_hist_gather(int64_t nslots, NbrList nl, NbrList ol,
                                                     const uint32_t* __restrict__ perm) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslots) return;
  const uint32_t nx = nl.sx[s];
  if (nx == 0xffffffffu) return;
  const uint32_t w = nl.word[s];
  const uint32_t ox = perm[nx], oy = perm[w & kIdxMask];
  // rank of this entry in nx's new list: the first place to look in ox's old list
  const uint32_t kh = (uint32_t)((uint64_t)(s - (int64_t)nl.base[nx >> 6] - (int64_t)(nx & 63)) / kWave);
  const uint64_t e = find_entry(ol, ox, oy, kh);
  if (e == ~0ull) return;
  const uint32_t we = ol.word[e];
  uint64_t src = ~0ull;
  if (w & kOwnerBit) {
    if ((we & kOwnerBit) && (we & kActiveBit)) src = e;
  } else if (!(we & kOwnerBit)) {
    if (we & kActiveBit) src = e;
  } else {
    const uint64_t e2 = find_entry(ol, oy, ox);
    if (e2 != ~0ull) {
      const uint32_t w2 = ol.word[e2];
      if ((w2 & kOwnerBit) && (w2 & kActiveBit)) src = e2;
    }
  }
  if (src == ~0ull) return;
  const HistView<T> nh = hist_view(nl, (T*)nullptr), oh = hist_view(ol, (T*)nullptr);
  nh.x[s] = oh.x[src];
  nh.y[s] = oh.y[src];
  nh.z[s] = oh.z[src];
  if (KR) {
    nh.mx[s] = oh.mx[src];
    nh.my[s] = oh.my[src];
    nh.mz[s] = oh.mz[src];
  }
  nl.word[s] = w | kActiveBit;
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

int nl_pairs(xdem_ctx* c, int64_t* out) {
  NbrList& L = c->nl[c->nl_cur];
  if (L.pairs < 0) {
    uint64_t hp = 0;
    const int rc = read_u64(c, c->d_pairs + c->nl_cur, &hp);
    if (rc) return rc;
    L.pairs = (int64_t)hp;
  }
  *out = L.pairs;
  return XDEM_OK;
}

int do_nbh_sphere(xdem_ctx* c, double rcut_inc) {
  if (!c->nl_valid) return set_err(c, XDEM_ERR_STATE, "nbh_sphere: run chunk_neighbors_contact first");
  const int64_t n = c->n;
  if (n == 0) {
    c->hist_ready = true;
    return XDEM_OK;
  }
  const bool have_old = c->nl_prev_valid && c->perm_valid;
  if (have_old) {
    const NbrList& nl = c->nl[c->nl_cur];
    const NbrList& ol = c->nl[1 - c->nl_cur];
    const int64_t os = ol.slots, ns = nl.slots;
    // a partner appears at most once per list unless a periodic direction has fewer than 3 cells (which
    // set_domain refuses); the scatter form stays for that case
    int hint_ok = 1;
    for (int d = 0; d < 3; d++)
      if (c->grid.per[d] && c->grid.dims[d] < 3) hint_ok = 0;
    // moment histories are zero (and not carried) until a rolling stiffness is used
    if (hint_ok && os > 0 && ns > 0) {
      const dim3 gs(grid_blocks(ns, 256)), bs(256);
      if (c->hist_fp32) {
        if (c->kr_used) hipLaunchKernelGGL((k_hist_gather<float, true>), gs, bs, 0, c->stream, ns, nl, ol, c->perm);
        else hipLaunchKernelGGL((k_hist_gather<float, false>), gs, bs, 0, c->stream, ns, nl, ol, c->perm);
      } else {
        if (c->kr_used) hipLaunchKernelGGL((k_hist_gather<double, true>), gs, bs, 0, c->stream, ns, nl, ol, c->perm);
        else hipLaunchKernelGGL((k_hist_gather<double, false>), gs, bs, 0, c->stream, ns, nl, ol, c->perm);
      }
      XDEM_HIP(c, hipGetLastError());
    } else if (os > 0) {
      uint32_t* inv = c->perm_tmp;  // free after the neighbour build
      hipLaunchKernelGGL(k_invert, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, n, c->perm, inv);
      const dim3 gs(grid_blocks(os, 256)), bs(256);
      if (c->hist_fp32) {
        if (c->kr_used) hipLaunchKernelGGL((k_hist_scatter<float, true>), gs, bs, 0, c->stream, os, nl, ol, inv, hint_ok);
        else hipLaunchKernelGGL((k_hist_scatter<float, false>), gs, bs, 0, c->stream, os, nl, ol, inv, hint_ok);
      } else {
        if (c->kr_used) hipLaunchKernelGGL((k_hist_scatter<double, true>), gs, bs, 0, c->stream, os, nl, ol, inv, hint_ok);
        else hipLaunchKernelGGL((k_hist_scatter<double, false>), gs, bs, 0, c->stream, os, nl, ol, inv, hint_ok);
      }
      XDEM_HIP(c, hipGetLastError());
    }
  }
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

// A rolling stiffness comes into use: the moment histories of the current sphere list become live. Until
// then no pass writes them (their value is zero by definition), so they are cleared here once.
int use_kr(xdem_ctx* c) {
  if (c->kr_used) return XDEM_OK;
  const NbrList& L = c->nl[c->nl_cur];
  if (c->nl_valid && L.slots > 0) {
    const size_t b64 = sizeof(double) * (size_t)L.slots, b32 = sizeof(float) * (size_t)L.slots;
    for (double* p : {L.mhx, L.mhy, L.mhz}) XDEM_HIP(c, hipMemsetAsync(p, 0, b64, c->stream));
    for (float* p : {L.mfx32, L.mfy32, L.mfz32}) XDEM_HIP(c, hipMemsetAsync(p, 0, b32, c->stream));
  }
  c->kr_used = true;
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
  if (n_pp) {
    const int rc = nl_pairs(c, n_pp);
    if (rc) return rc;
  }
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
      // an entry without the active bit has no history (its memory is not kept at zero)
      const bool act = (wd & kActiveBit) != 0 && (!moments || c->kr_used);
      friction[3 * k] = act ? hx[s] : 0.0;
      friction[3 * k + 1] = act ? hy[s] : 0.0;
      friction[3 * k + 2] = act ? hz[s] : 0.0;
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
