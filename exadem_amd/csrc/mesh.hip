// mesh.hip — stationary triangle-mesh driver for spheres: registration, host-built bins, per-particle
// candidate lists with history carry-over at the neighbour build, and the candidate/history dump.
//
//   registration  register_stl_mesh (src/driver/register_stl_mesh.cpp:30-140): world vertex k is
//                 center + Q (v_k) (shape::get_vertex with homothety 1, shape.hpp:259-262)
//   candidates    nbh_sphere's RShape branch (src/sphere/nbh_sphere.cu:221-229) calling
//                 detection_sphere_driver (nbh_sphere.hpp:25-86) with rVerlet = rcut_inc: vertices, then
//                 edges, then faces, each by index
//   history       update() of the interaction manager: an interaction keeps its friction and moment when
//                 the same (particle, driver, type, feature) was in the previous list
//
// Bins replace the reference's per-cell projection of the driver (rshape_grid.hpp): a feature is listed in
// every bin its bounding box, enlarged by reach = max R + rcut_inc + rm, overlaps, so the bin of a sphere's
// centre holds every feature that can pass its filter; the filter itself decides, so the candidate set and
// its order (ascending word = kind, then index) do not depend on the binning.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>

#include "mesh.h"

namespace xdem {

static V3 host_qrot(const double q[4], V3 v) {  // q = (w, x, y, z)
  const V3 qv{q[1], q[2], q[3]};
  V3 uv = cross(qv, v);
  V3 uuv = cross(qv, uv);
  uv = uv * (2.0 * q[0]);
  uuv = uuv * 2.0;
  return v + uv + uuv;
}

static void free_ml(MeshList& l) {
  for (void* p : {(void*)l.cnt, (void*)l.word, (void*)l.hx, (void*)l.hy, (void*)l.hz, (void*)l.mhx, (void*)l.mhy,
                  (void*)l.mhz})
    if (p) (void)hipFree(p);
  l = MeshList{};
}

void mesh_free(xdem_ctx* c) {
  for (void* p : {(void*)c->d_mesh_v, (void*)c->d_mesh_e, (void*)c->d_mesh_f, (void*)c->d_mesh_start,
                  (void*)c->d_mesh_feat, (void*)c->d_mesh_over})
    if (p) (void)hipFree(p);
  c->d_mesh_v = nullptr;
  c->d_mesh_e = nullptr;
  c->d_mesh_f = nullptr;
  c->d_mesh_start = nullptr;
  c->d_mesh_feat = nullptr;
  c->d_mesh_over = nullptr;
  if (c->mstate) (void)hipFree(c->mstate);
  c->mstate = nullptr;
  c->cap_mstate = 0;
  for (void* p : {(void*)c->mesh_cand, (void*)c->mesh_idx, (void*)c->mesh_terms, (void*)c->mesh_geo,
                  (void*)c->mesh_wcnt, (void*)c->mesh_cc, (void*)c->mesh_off})
    if (p) (void)hipFree(p);
  c->mesh_cc = nullptr;
  c->mesh_off = nullptr;
  c->mesh_wcnt = nullptr;
  c->mesh_geo = nullptr;
  c->mesh_cand = nullptr;
  c->mesh_idx = nullptr;
  c->mesh_terms = nullptr;
  c->cap_mesh_cand = 0;
  c->mesh_ncand = 0;
  free_ml(c->ml[0]);
  free_ml(c->ml[1]);
  free_ml(c->mc);
  c->mc_valid = c->mc_dirty = false;
  c->ml_alloc = 0;
  c->ml_valid = false;
  c->has_mesh = false;
  c->mesh = MeshView{};
  c->mesh_reach = -1;
  c->mesh_rv = -1;
  c->mesh_hv.clear();
  c->mesh_he.clear();
  c->mesh_hf.clear();
}

static int alloc_ml(xdem_ctx* c) {
  if (c->ml_alloc == c->cap && c->ml[0].cnt) return XDEM_OK;
  const int64_t m = (int64_t)c->mesh_cap * c->cap;
  int rc;
  for (int k = 0; k < 2; k++) {
    MeshList& l = c->ml[k];
    free_ml(l);
    if ((rc = dev_realloc(c, &l.cnt, c->cap))) return rc;
    if ((rc = dev_realloc(c, &l.word, m))) return rc;
    if ((rc = dev_realloc(c, &l.hx, m))) return rc;
    if ((rc = dev_realloc(c, &l.hy, m))) return rc;
    if ((rc = dev_realloc(c, &l.hz, m))) return rc;
    if ((rc = dev_realloc(c, &l.mhx, m))) return rc;
    if ((rc = dev_realloc(c, &l.mhy, m))) return rc;
    if ((rc = dev_realloc(c, &l.mhz, m))) return rc;
  }
  if ((rc = dev_realloc(c, &c->mesh_idx, c->cap))) return rc;
  if ((rc = dev_realloc(c, &c->mesh_cc, c->cap))) return rc;
  if ((rc = dev_realloc(c, &c->mesh_off, c->cap + 1))) return rc;
  if ((rc = dev_realloc(c, &c->mesh_wcnt, (c->cap + 63) / 64))) return rc;
  c->ml_alloc = c->cap;
  c->mesh.stride = c->cap;
  c->ml_valid = false;  // the lists no longer match the particle capacity
  return XDEM_OK;
}

// Bins for the current reach (host side; the mesh is stationary, so they are rebuilt only when the
// reach grows).
// rv: the edge filter admits projections r in [-rv, 1 + rv] of the edge, so edges are extended by rv of
// their length at both ends before the enlargement.
static int build_bins(xdem_ctx* c, double reach, double rv) {
  const auto& hv = c->mesh_hv;
  V3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
  for (const V3& v : hv) {
    lo = {std::min(lo.x, v.x), std::min(lo.y, v.y), std::min(lo.z, v.z)};
    hi = {std::max(hi.x, v.x), std::max(hi.y, v.y), std::max(hi.z, v.z)};
  }
  const double ext = std::max({hi.x - lo.x, hi.y - lo.y, hi.z - lo.z});
  const double h = std::max(reach, ext / 96.0);
  const double pad = reach * (1.0 + 1e-9) + 1e-12 * std::max(1.0, ext);
  const V3 blo{lo.x - pad, lo.y - pad, lo.z - pad};
  const double inv_h = 1.0 / h;
  int dim[3];
  const double span[3] = {hi.x - lo.x, hi.y - lo.y, hi.z - lo.z};
  for (int a = 0; a < 3; a++) dim[a] = (int)std::floor((span[a] + 2 * pad) * inv_h) + 1;
  const int64_t nb = (int64_t)dim[0] * dim[1] * dim[2];
  auto bin_range = [&](V3 a, V3 b, int r0[3], int r1[3]) {
    const double mn[3] = {a.x - pad, a.y - pad, a.z - pad}, mx[3] = {b.x + pad, b.y + pad, b.z + pad};
    const double l[3] = {blo.x, blo.y, blo.z};
    for (int q = 0; q < 3; q++) {
      r0[q] = std::max(0, std::min(dim[q] - 1, (int)std::floor((mn[q] - l[q]) * inv_h)));
      r1[q] = std::max(0, std::min(dim[q] - 1, (int)std::floor((mx[q] - l[q]) * inv_h)));
    }
  };
  std::vector<uint32_t> count((size_t)nb + 1, 0);
  std::vector<std::pair<uint32_t, std::array<int, 6>>> feats;
  auto add = [&](uint32_t wd, V3 a, V3 b) {
    std::array<int, 6> r;
    bin_range(a, b, &r[0], &r[3]);
    feats.push_back({wd, r});
  };
  auto mn3 = [](V3 a, V3 b) { return V3{std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; };
  auto mx3 = [](V3 a, V3 b) { return V3{std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; };
  for (size_t k = 0; k < hv.size(); k++) add((uint32_t)k, hv[k], hv[k]);
  for (size_t k = 0; k < c->mesh_he.size(); k++) {
    const uint2 e = c->mesh_he[k];
    const V3 fs = hv[e.y] - hv[e.x];
    const V3 a = hv[e.x] - fs * rv, b = hv[e.y] + fs * rv;
    add((1u << kMeshKindShift) | (uint32_t)k, mn3(a, b), mx3(a, b));
  }
  for (size_t k = 0; k < c->mesh_hf.size(); k++) {
    const uint4 f = c->mesh_hf[k];
    add((2u << kMeshKindShift) | (uint32_t)k, mn3(mn3(hv[f.x], hv[f.y]), hv[f.z]),
        mx3(mx3(hv[f.x], hv[f.y]), hv[f.z]));
  }
  int64_t total = 0;
  for (auto& fr : feats) {
    const auto& r = fr.second;
    total += (int64_t)(r[3] - r[0] + 1) * (r[4] - r[1] + 1) * (r[5] - r[2] + 1);
  }
  if (total > (int64_t)1 << 28)
    return set_err(c, XDEM_ERR_CAPACITY, "mesh: bin lists exceed 2^28 entries (features too large for the bins)");
  for (auto& fr : feats) {
    const auto& r = fr.second;
    for (int z = r[2]; z <= r[5]; z++)
      for (int y = r[1]; y <= r[4]; y++)
        for (int x = r[0]; x <= r[3]; x++) count[((size_t)z * dim[1] + y) * dim[0] + x]++;
  }
  std::vector<uint32_t> start((size_t)nb + 1, 0);
  for (int64_t b = 0; b < nb; b++) start[b + 1] = start[b] + count[b];
  std::vector<uint32_t> fill(start.begin(), start.end() - 1), feat((size_t)std::max<int64_t>(total, 1));
  for (auto& fr : feats) {  // words in ascending order, so each bin's list is sorted
    const auto& r = fr.second;
    for (int z = r[2]; z <= r[5]; z++)
      for (int y = r[1]; y <= r[4]; y++)
        for (int x = r[0]; x <= r[3]; x++) feat[fill[((size_t)z * dim[1] + y) * dim[0] + x]++] = fr.first;
  }
  int rc;
  if ((rc = dev_realloc(c, &c->d_mesh_start, nb + 1))) return rc;
  if ((rc = dev_realloc(c, &c->d_mesh_feat, (int64_t)feat.size()))) return rc;
  XDEM_HIP(c, hipMemcpyAsync(c->d_mesh_start, start.data(), sizeof(uint32_t) * start.size(), hipMemcpyHostToDevice,
                             c->stream));
  XDEM_HIP(c, hipMemcpyAsync(c->d_mesh_feat, feat.data(), sizeof(uint32_t) * feat.size(), hipMemcpyHostToDevice,
                             c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));  // host vectors go out of scope
  c->mesh.lo = blo;
  c->mesh.inv_h = inv_h;
  for (int a = 0; a < 3; a++) c->mesh.dim[a] = dim[a];
  c->mesh.start = c->d_mesh_start;
  c->mesh.feat = c->d_mesh_feat;
  c->mesh_reach = reach;
  c->mesh_rv = rv;
  return XDEM_OK;
}

// Largest radius (bit pattern of a non-negative double orders like the value).
__global__ __launch_bounds__(256) void k_max_radius(int64_t n, const Core* __restrict__ core,
                                                    unsigned long long* __restrict__ out) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long b = 0;
  if (x < n) {
    const double r = core[x].xr.w;
    b = r > 0 ? (unsigned long long)__double_as_longlong(r) : 0ull;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long t = __shfl_down(b, o, 64);
    b = t > b ? t : b;
  }
  // one atomic per block: per-wavefront atomics on the one address serialise (190 us at 1M)
  __shared__ unsigned long long wmax[4];
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = b;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = wmax[0];
    for (int w = 1; w < 4; w++) m = wmax[w] > m ? wmax[w] : m;
    if (m) atomicMax(out, m);
  }
}

__global__ __launch_bounds__(256) void k_mesh_carry(int64_t n, PState s, MeshView m, double rcut_inc, MeshList nl,
                                                    MeshList ol, const uint32_t* __restrict__ perm, int have_old,
                                                    int32_t* __restrict__ over) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const d4 xr = s.core[x].xr;
  const V3 p{xr.x, xr.y, xr.z};
  const double R = xr.w;
  const double fx = (p.x - m.lo.x) * m.inv_h, fy = (p.y - m.lo.y) * m.inv_h, fz = (p.z - m.lo.z) * m.inv_h;
  uint32_t cnt = 0;
  if (fx >= 0.0 && fy >= 0.0 && fz >= 0.0 && fx < (double)m.dim[0] && fy < (double)m.dim[1] &&
      fz < (double)m.dim[2]) {
    const int64_t b = ((int64_t)(int)fz * m.dim[1] + (int)fy) * m.dim[0] + (int)fx;
    const uint32_t q0 = m.start[b], q1 = m.start[b + 1];
    const uint32_t o = have_old ? perm[x] : 0u;
    const uint32_t oc = have_old ? min(ol.cnt[o], (uint32_t)m.cap) : 0u;
    for (uint32_t q = q0; q < q1; q++) {
      const uint32_t wd = m.feat[q];
      if (!mesh_filter(m, wd, p, R, rcut_inc)) continue;
      if (cnt < (uint32_t)m.cap) {
        const int64_t j = (int64_t)cnt * m.stride + x;
        double hx = 0, hy = 0, hz = 0, mx = 0, my = 0, mz = 0;
        for (uint32_t k = 0; k < oc; k++) {
          const int64_t jo = (int64_t)k * m.stride + o;
          if (ol.word[jo] == wd) {
            hx = ol.hx[jo]; hy = ol.hy[jo]; hz = ol.hz[jo];
            mx = ol.mhx[jo]; my = ol.mhy[jo]; mz = ol.mhz[jo];
            break;
          }
        }
        nl.word[j] = wd;
        nl.hx[j] = hx; nl.hy[j] = hy; nl.hz[j] = hz;
        nl.mhx[j] = mx; nl.mhy[j] = my; nl.mhz[j] = mz;
      }
      cnt++;
    }
  }
  nl.cnt[x] = cnt;
  if (cnt > (uint32_t)m.cap) atomicMax(over, (int32_t)cnt);
}

// Flat candidate index in particle order: the counts (clamped to the capacity) are scanned into off[0..n]
// (k_mesh_count, then a scan); idx[x] = (off[x], count), (0, 0) without candidates; wcnt[x >> 6]: the
// wave's candidates.
__global__ __launch_bounds__(256) void k_mesh_count(int64_t n, const uint32_t* __restrict__ cnt, uint32_t cap,
                                                    uint32_t* __restrict__ cc) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < n) cc[x] = min(cnt[x], cap);
}
__global__ __launch_bounds__(256) void k_mesh_index(int64_t n, const uint32_t* __restrict__ off,
                                                    uint2* __restrict__ idx, uint32_t* __restrict__ wcnt) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const uint32_t o = off[x], c = off[x + 1] - o;
  idx[x] = c ? uint2{o, c} : uint2{0u, 0u};
  if ((x & 63) == 0) wcnt[x >> 6] = off[x + 64 < n ? x + 64 : n] - o;
}
// The flat candidates: cand[t] = (particle, rank), one thread per particle over its run.
__global__ __launch_bounds__(256) void k_mesh_cand(int64_t n, const uint2* __restrict__ idx,
                                                   uint2* __restrict__ cand) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const uint2 r = idx[x];
  for (uint32_t k = 0; k < r.y; k++) cand[r.x + k] = uint2{(uint32_t)x, k};
}

// One thread per flat candidate. (Running each XCD's candidates on that XCD, for the L2, measured slower:
// the candidates of a plane cluster in a few XCDs' particle ranges, which then do all the work.)
template <bool KR, bool GUARD>
__global__ __launch_bounds__(256) void k_mesh_pass(MeshView m, MeshList l, MeshTerms t, const Core* __restrict__ core,
                                                   double dt, ContactParams cpd, const int32_t* __restrict__ flags,
                                                   int32_t trig_in, double* cs) {
  const int64_t jc = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (jc >= t.ncand) return;
  // the candidate's own loads (entry, cached feature, history) go out before the guard is tested, which
  // then waits on them (the particle's record needs the entry)
  const uint2 ck = t.cand[jc];
  const double2* gp = reinterpret_cast<const double2*>(t.geo + jc * 10);
  const double2 g0 = gp[0], g1 = gp[1], g2 = gp[2], g3 = gp[3], g4 = gp[4];
  const V3 ft0{l.hx[jc], l.hy[jc], l.hz[jc]};
  V3 mint0{0.0, 0.0, 0.0};
  if (KR) mint0 = {l.mhx[jc], l.mhy[jc], l.mhz[jc]};
  if (GUARD && (flags[1] || flags[trig_in]) && ck.x != 0xffffffffu) return;
  const V3 tg[3] = {V3{g0.x, g0.y, g1.x}, V3{g1.y, g2.x, g2.y}, V3{g3.x, g3.y, g4.x}};
  const int64_t x = ck.x;
  const Core me = core[x];
  const V3 rx{me.xr.x, me.xr.y, me.xr.z}, vx{me.vm.x, me.vm.y, me.vm.z}, wx{me.w.x, me.w.y, me.w.z};
  mesh_candidate<KR>(m, l, jc, (int64_t)ck.y * m.stride + x, (uint32_t)g4.y, tg, rx, me.xr.w, me.vm.w, vx, wx, dt,
                     cpd, t.terms + jc * 6, cs, ft0, mint0);
}

// Compact copy <-> lists, one thread per flat candidate: the gather copies the words and histories (and,
// for spheres, geo non-null, caches the feature's geometry and kind), the write-back the histories.
__global__ __launch_bounds__(256) void k_mesh_gather(MeshTerms t, MeshList l, MeshList lc, MeshView m, double* geo) {
  const int64_t jc = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (jc >= t.ncand) return;
  const uint2 ck = t.cand[jc];
  const int64_t j = (int64_t)ck.y * m.stride + ck.x;
  const uint32_t wd = l.word[j];
  lc.word[jc] = wd;
  if (geo) {
    V3 tg[3];
    const uint32_t kind = mesh_feature(m, wd, tg);
    double2* gp = reinterpret_cast<double2*>(geo + jc * 10);
    gp[0] = double2{tg[0].x, tg[0].y};
    gp[1] = double2{tg[0].z, tg[1].x};
    gp[2] = double2{tg[1].y, tg[1].z};
    gp[3] = double2{tg[2].x, tg[2].y};
    gp[4] = double2{tg[2].z, (double)kind};
  }
  lc.hx[jc] = l.hx[j]; lc.hy[jc] = l.hy[j]; lc.hz[jc] = l.hz[j];
  lc.mhx[jc] = l.mhx[j]; lc.mhy[jc] = l.mhy[j]; lc.mhz[jc] = l.mhz[j];
}
__global__ __launch_bounds__(256) void k_mesh_writeback(MeshTerms t, MeshList l, MeshList lc, int64_t stride) {
  const int64_t jc = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (jc >= t.ncand) return;
  const uint2 ck = t.cand[jc];
  const int64_t j = (int64_t)ck.y * stride + ck.x;
  l.hx[j] = lc.hx[jc]; l.hy[j] = lc.hy[jc]; l.hz[j] = lc.hz[jc];
  l.mhx[j] = lc.mhx[jc]; l.mhy[j] = lc.mhy[jc]; l.mhz[j] = lc.mhz[jc];
}

static MeshTerms terms_view(xdem_ctx* c);

int mesh_compact(xdem_ctx* c) {
  if (c->mc_valid || !c->has_mesh || !c->ml_valid || c->mesh_ncand == 0) return XDEM_OK;
  const dim3 grid(grid_blocks(c->mesh_ncand, 256)), block(256);
  // the feature geometry cache serves the sphere pass (polyhedra words have another layout)
  hipLaunchKernelGGL(k_mesh_gather, grid, block, 0, c->stream, terms_view(c), c->ml[c->ml_cur], c->mc, c->mesh,
                     c->mesh_poly ? nullptr : c->mesh_geo);
  XDEM_HIP(c, hipGetLastError());
  c->mc_valid = true;
  c->mc_dirty = false;
  return XDEM_OK;
}

int mesh_sync(xdem_ctx* c) {
  if (!c->mc_dirty) return XDEM_OK;
  c->mc_dirty = false;
  if (!c->has_mesh || !c->ml_valid || c->mesh_ncand == 0) return XDEM_OK;
  const dim3 grid(grid_blocks(c->mesh_ncand, 256)), block(256);
  hipLaunchKernelGGL(k_mesh_writeback, grid, block, 0, c->stream, terms_view(c), c->ml[c->ml_cur], c->mc,
                     c->mesh.stride);
  XDEM_HIP(c, hipGetLastError());
  return XDEM_OK;
}

static MeshTerms terms_view(xdem_ctx* c) {
  MeshTerms t;
  t.cand = c->mesh_cand;
  t.idx = c->mesh_idx;
  t.wcnt = c->mesh_wcnt;
  t.terms = c->mesh_terms;
  t.geo = c->mesh_geo;
  t.ncand = c->mesh_ncand;
  return t;
}

int launch_mesh_pass(xdem_ctx* c, const ContactParams& cpd, double dt, bool guarded, int32_t trig_in, double* cs,
                     MeshTerms* t) {
  *t = MeshTerms{};
  if (!c->has_mesh || !c->ml_valid || c->mesh_ncand == 0) return XDEM_OK;
  if (c->mesh_poly)
    return set_err(c, XDEM_ERR_STATE, "mesh contacts of spheres: the mesh lists were built by nbh_polyhedron");
  int rc = mesh_compact(c);
  if (rc) return rc;
  *t = terms_view(c);
  const dim3 grid(grid_blocks(c->mesh_ncand, 256)), block(256);
  const MeshList& l = c->mc;
  c->mc_dirty = true;
  const Core* core = live(c).core;
  if (c->kr_used) {
    if (guarded)
      hipLaunchKernelGGL((k_mesh_pass<true, true>), grid, block, 0, c->stream, c->mesh, l, *t, core, dt, cpd,
                         c->d_flags, trig_in, cs);
    else
      hipLaunchKernelGGL((k_mesh_pass<true, false>), grid, block, 0, c->stream, c->mesh, l, *t, core, dt, cpd,
                         c->d_flags, trig_in, cs);
  } else {
    if (guarded)
      hipLaunchKernelGGL((k_mesh_pass<false, true>), grid, block, 0, c->stream, c->mesh, l, *t, core, dt, cpd,
                         c->d_flags, trig_in, cs);
    else
      hipLaunchKernelGGL((k_mesh_pass<false, false>), grid, block, 0, c->stream, c->mesh, l, *t, core, dt, cpd,
                         c->d_flags, trig_in, cs);
  }
  XDEM_HIP(c, hipGetLastError());
  return XDEM_OK;
}

int mesh_carry(xdem_ctx* c, double rcut_inc, bool have_old, bool poly) {
  if (!c->has_mesh || c->n == 0) return XDEM_OK;
  int rc = mesh_sync(c);  // the previous lists' histories, as the last pass left them
  if (rc) return rc;
  c->mc_valid = false;
  if ((rc = alloc_ml(c))) return rc;
  const int64_t n = c->n;
  // reach of the bins
  unsigned long long* d_rmax = (unsigned long long*)c->d_mesh_over;
  XDEM_HIP(c, hipMemsetAsync(c->d_mesh_over, 0, 8 * sizeof(int32_t), c->stream));
  hipLaunchKernelGGL(k_max_radius, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, n, live(c).core, d_rmax);
  XDEM_HIP(c, hipGetLastError());
  unsigned long long rb = 0;
  XDEM_HIP(c, hipMemcpyAsync(&rb, d_rmax, sizeof(rb), hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  double rmax;
  memcpy(&rmax, &rb, sizeof(rmax));
  // a sphere's candidates lie within R + rcut_inc + r_m of its centre; a polyhedron's (radius: its
  // bounding sphere) also through its own edges, which the vertex-edge filter extends by rcut_inc of their
  // length (at most 2 R) at both ends
  const double reach = (poly ? rmax * (1.0 + 2.0 * rcut_inc) : rmax) + rcut_inc + c->mesh.rm;
  if (!(reach <= c->mesh_reach) || rcut_inc != c->mesh_rv) {
    if ((rc = build_bins(c, reach * 1.05, rcut_inc))) return rc;
  }
  const bool old = have_old && c->ml_valid;
  const int nxt = 1 - c->ml_cur;
  int32_t* d_over = c->d_mesh_over + 2;
  if (poly) {
    if ((rc = launch_poly_mesh_carry(c, rcut_inc, c->ml[nxt], c->ml[c->ml_cur], old, d_over))) return rc;
  } else {
    hipLaunchKernelGGL(k_mesh_carry, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, n, live(c), c->mesh,
                       rcut_inc, c->ml[nxt], c->ml[c->ml_cur], c->perm, old ? 1 : 0, d_over);
    XDEM_HIP(c, hipGetLastError());
  }
  hipLaunchKernelGGL(k_mesh_count, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, n, c->ml[nxt].cnt,
                     (uint32_t)c->mesh_cap, c->mesh_cc);
  if ((rc = scan_u32(c, c->mesh_cc, c->mesh_off, n))) return rc;
  hipLaunchKernelGGL(k_mesh_index, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, n, c->mesh_off, c->mesh_idx,
                     c->mesh_wcnt);
  XDEM_HIP(c, hipGetLastError());
  int32_t over[2] = {0, 0};  // largest overflowing count, candidates
  XDEM_HIP(c, hipMemcpyAsync(over, d_over, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(over + 1, c->mesh_off + n, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  c->ml_cur = nxt;
  c->mesh_poly = poly;
  c->mesh_ncand = over[1];
  if (c->mesh_ncand > c->cap_mesh_cand) {
    const int64_t cap = c->mesh_ncand + c->mesh_ncand / 4 + 1024;
    MeshList& l = c->mc;
    free_ml(l);
    if ((rc = dev_realloc(c, &c->mesh_cand, cap)) || (rc = dev_realloc(c, &c->mesh_terms, cap * 6)) ||
        (rc = dev_realloc(c, &c->mesh_geo, cap * 10)) ||
        (rc = dev_realloc(c, &l.word, cap)) || (rc = dev_realloc(c, &l.hx, cap)) || (rc = dev_realloc(c, &l.hy, cap)) ||
        (rc = dev_realloc(c, &l.hz, cap)) || (rc = dev_realloc(c, &l.mhx, cap)) || (rc = dev_realloc(c, &l.mhy, cap)) ||
        (rc = dev_realloc(c, &l.mhz, cap)))
      return rc;
    c->cap_mesh_cand = cap;
  }
  if (c->mesh_ncand > 0) {
    hipLaunchKernelGGL(k_mesh_cand, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, n, c->mesh_idx,
                       c->mesh_cand);
    XDEM_HIP(c, hipGetLastError());
  }
  if (over[0] > 0) {
    c->ml_valid = false;
    return set_err(c, XDEM_ERR_CAPACITY,
                   "mesh: a particle has " + std::to_string(over[0]) + " candidate mesh features, more than the " +
                       std::to_string(c->mesh_cap) + " slots given to xdem_set_mesh");
  }
  c->ml_valid = true;
  return XDEM_OK;
}

}  // namespace xdem

using namespace xdem;

extern "C" int xdem_set_mesh(xdem_ctx* c, int32_t id, const double* vertices, int64_t nv, const uint32_t* edges,
                             int64_t ne, const uint32_t* faces, int64_t nf, double minkowski, const double* center,
                             const double* quat, int32_t max_candidates) {
  if (!c) return XDEM_ERR_ARG;
  XDEM_HIP(c, hipSetDevice(c->device));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  mesh_free(c);
  if (nv == 0) return XDEM_OK;  // no mesh
  if (nv < 0 || ne < 0 || nf < 0 || !vertices || (ne && !edges) || (nf && !faces) || !center || !quat)
    return set_err(c, XDEM_ERR_ARG, "set_mesh: null arrays or negative counts");
  if (nv > (int64_t)kMeshIdxMask || ne > (int64_t)kMeshIdxMask || nf > (int64_t)kMeshIdxMask)
    return set_err(c, XDEM_ERR_ARG, "set_mesh: more than 2^30 vertices, edges or faces");
  if (!(minkowski >= 0.0) || id < 0 || max_candidates < 1 || max_candidates > 4096)
    return set_err(c, XDEM_ERR_ARG, "set_mesh: minkowski >= 0, id >= 0 and 1 <= max_candidates <= 4096");
  if (c->nshapes > 0) return set_err(c, XDEM_ERR_ARG, "set_mesh: mesh drivers act on spheres only");
  for (int64_t k = 0; k < 2 * ne; k++)
    if (edges[k] >= (uint64_t)nv) return set_err(c, XDEM_ERR_ARG, "set_mesh: edge vertex index out of range");
  for (int64_t k = 0; k < 3 * nf; k++)
    if (faces[k] >= (uint64_t)nv) return set_err(c, XDEM_ERR_ARG, "set_mesh: face vertex index out of range");
  const V3 ctr{center[0], center[1], center[2]};
  c->mesh_hv.resize(nv);
  for (int64_t k = 0; k < nv; k++)
    c->mesh_hv[k] = ctr + host_qrot(quat, V3{vertices[3 * k], vertices[3 * k + 1], vertices[3 * k + 2]});
  c->mesh_he.resize(ne);
  for (int64_t k = 0; k < ne; k++) c->mesh_he[k] = make_uint2(edges[2 * k], edges[2 * k + 1]);
  c->mesh_hf.resize(nf);
  for (int64_t k = 0; k < nf; k++) c->mesh_hf[k] = make_uint4(faces[3 * k], faces[3 * k + 1], faces[3 * k + 2], 0u);
  int rc;
  if ((rc = dev_realloc(c, &c->d_mesh_v, nv))) return rc;
  if ((rc = dev_realloc(c, &c->d_mesh_e, std::max<int64_t>(ne, 1)))) return rc;
  if ((rc = dev_realloc(c, &c->d_mesh_f, std::max<int64_t>(nf, 1)))) return rc;
  if ((rc = dev_realloc(c, &c->d_mesh_over, 8))) return rc;
  XDEM_HIP(c, hipMemcpy(c->d_mesh_v, c->mesh_hv.data(), sizeof(V3) * nv, hipMemcpyHostToDevice));
  if (ne) XDEM_HIP(c, hipMemcpy(c->d_mesh_e, c->mesh_he.data(), sizeof(uint2) * ne, hipMemcpyHostToDevice));
  if (nf) XDEM_HIP(c, hipMemcpy(c->d_mesh_f, c->mesh_hf.data(), sizeof(uint4) * nf, hipMemcpyHostToDevice));
  c->has_mesh = true;
  c->mesh_id = id;
  c->mesh_cap = max_candidates;
  c->mesh.v = c->d_mesh_v;
  c->mesh.e = c->d_mesh_e;
  c->mesh.f = c->d_mesh_f;
  c->mesh.rm = minkowski;
  c->mesh.center = ctr;
  c->mesh.cap = max_candidates;
  c->hist_ready = false;  // the next neighbour build makes the candidate lists
  return XDEM_OK;
}

// Rows (owner id, feature index, type 7/8/9, friction, moment) of the current candidate lists, sorted by
// (owner id, type, feature). written = row count; more rows than cap fail with XDEM_ERR_CAPACITY.
extern "C" int xdem_mesh_interactions(xdem_ctx* c, int64_t cap, uint64_t* id_i, uint32_t* sub_j, uint16_t* type,
                                      double* friction, double* moment, int64_t* written) {
  if (!c || !written) return XDEM_ERR_ARG;
  *written = 0;
  if (!c->has_mesh || !c->ml_valid || c->n == 0) return XDEM_OK;
  XDEM_HIP(c, hipSetDevice(c->device));
  int rc = mesh_sync(c);
  if (rc) return rc;
  const int64_t n = c->n, st = c->cap;
  const MeshList& L = c->ml[c->ml_cur];
  std::vector<uint32_t> cnt(n);
  std::vector<uint64_t> ids(n);
  XDEM_HIP(c, hipMemcpyAsync(cnt.data(), L.cnt, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(ids.data(), live(c).id, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  uint32_t kmax = 0;
  for (auto v : cnt) kmax = std::max(kmax, std::min(v, (uint32_t)c->mesh_cap));
  const int64_t m = (int64_t)kmax * st;
  std::vector<uint32_t> wd(m);
  std::vector<double> h[6];
  double* src[6] = {L.hx, L.hy, L.hz, L.mhx, L.mhy, L.mhz};
  if (m) {
    XDEM_HIP(c, hipMemcpyAsync(wd.data(), L.word, sizeof(uint32_t) * m, hipMemcpyDeviceToHost, c->stream));
    for (int q = 0; q < 6; q++) {
      h[q].resize(m);
      XDEM_HIP(c, hipMemcpyAsync(h[q].data(), src[q], sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
    }
    XDEM_HIP(c, hipStreamSynchronize(c->stream));
  }
  struct Row {
    uint64_t id;
    uint32_t wd;
    int64_t j;
  };
  std::vector<Row> rows;
  for (int64_t x = 0; x < n; x++) {
    const uint32_t k1 = std::min(cnt[x], (uint32_t)c->mesh_cap);
    for (uint32_t k = 0; k < k1; k++) rows.push_back({ids[x], wd[(int64_t)k * st + x], (int64_t)k * st + x});
  }
  std::sort(rows.begin(), rows.end(), [](const Row& a, const Row& b) { return a.id != b.id ? a.id < b.id : a.wd < b.wd; });
  *written = (int64_t)rows.size();
  if ((int64_t)rows.size() > cap) return set_err(c, XDEM_ERR_CAPACITY, "mesh_interactions: output arrays too small");
  for (size_t r = 0; r < rows.size(); r++) {
    const int64_t j = rows[r].j;
    if (id_i) id_i[r] = rows[r].id;
    if (sub_j) sub_j[r] = rows[r].wd & kMeshIdxMask;
    if (type) type[r] = (uint16_t)(7 + (rows[r].wd >> kMeshKindShift));
    if (friction) {
      friction[3 * r] = h[0][j];
      friction[3 * r + 1] = h[1][j];
      friction[3 * r + 2] = h[2][j];
    }
    if (moment) {
      moment[3 * r] = h[3][j];
      moment[3 * r + 1] = h[4][j];
      moment[3 * r + 2] = h[5][j];
    }
  }
  return XDEM_OK;
}

// Restart: saved mesh interaction histories (rows as xdem_mesh_interactions writes them) handed to the
// current candidate lists by key (owner id, type, feature); unmatched candidates keep a zero history.
// *matched = rows that found their candidate.
extern "C" int xdem_mesh_interactions_load(xdem_ctx* c, int64_t m, const uint64_t* id_i, const uint32_t* sub_j,
                                           const uint16_t* type, const double* friction, const double* moment,
                                           int64_t* matched) {
  if (!c || (m > 0 && (!id_i || !sub_j || !type || !friction))) return XDEM_ERR_ARG;
  if (matched) *matched = 0;
  if (m == 0) return XDEM_OK;
  if (!c->has_mesh || !c->ml_valid) return set_err(c, XDEM_ERR_STATE, "mesh_interactions_load: no mesh candidate lists (run nbh_sphere)");
  XDEM_HIP(c, hipSetDevice(c->device));
  {
    const int rs = mesh_sync(c);  // unmatched candidates keep the history the last pass left
    if (rs) return rs;
  }
  struct Key {
    uint64_t id;
    uint32_t wd;
    bool operator==(const Key& o) const { return id == o.id && wd == o.wd; }
  };
  struct KeyHash {
    size_t operator()(const Key& k) const { return std::hash<uint64_t>()(k.id * 0x9E3779B97F4A7C15ull ^ k.wd); }
  };
  std::unordered_map<Key, int64_t, KeyHash> rows;
  rows.reserve((size_t)m * 2);
  for (int64_t r = 0; r < m; r++) {
    if (type[r] < 7 || type[r] > 9 || sub_j[r] > kMeshIdxMask)
      return set_err(c, XDEM_ERR_ARG, "mesh_interactions_load: type must be 7, 8 or 9");
    rows[{id_i[r], ((uint32_t)(type[r] - 7) << kMeshKindShift) | sub_j[r]}] = r;
  }
  const int64_t n = c->n, st = c->cap;
  const MeshList& L = c->ml[c->ml_cur];
  std::vector<uint32_t> cnt(n);
  std::vector<uint64_t> ids(n);
  XDEM_HIP(c, hipMemcpyAsync(cnt.data(), L.cnt, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(ids.data(), live(c).id, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  uint32_t kmax = 0;
  for (auto v : cnt) kmax = std::max(kmax, std::min(v, (uint32_t)c->mesh_cap));
  const int64_t tot = (int64_t)kmax * st;
  if (tot == 0) return XDEM_OK;
  std::vector<uint32_t> wd(tot);
  std::vector<double> h[6];
  double* dst[6] = {L.hx, L.hy, L.hz, L.mhx, L.mhy, L.mhz};
  XDEM_HIP(c, hipMemcpyAsync(wd.data(), L.word, sizeof(uint32_t) * tot, hipMemcpyDeviceToHost, c->stream));
  for (int q = 0; q < 6; q++) {
    h[q].resize(tot);
    XDEM_HIP(c, hipMemcpyAsync(h[q].data(), dst[q], sizeof(double) * tot, hipMemcpyDeviceToHost, c->stream));
  }
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  int64_t hits = 0;
  for (int64_t x = 0; x < n; x++) {
    const uint32_t k1 = std::min(cnt[x], (uint32_t)c->mesh_cap);
    for (uint32_t k = 0; k < k1; k++) {
      const int64_t j = (int64_t)k * st + x;
      auto it = rows.find({ids[x], wd[j]});
      if (it == rows.end()) continue;
      const int64_t r = it->second;
      for (int a = 0; a < 3; a++) {
        h[a][j] = friction[3 * r + a];
        h[3 + a][j] = moment ? moment[3 * r + a] : 0.0;
      }
      hits++;
    }
  }
  for (int q = 0; q < 6; q++)
    XDEM_HIP(c, hipMemcpyAsync(dst[q], h[q].data(), sizeof(double) * tot, hipMemcpyHostToDevice, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  c->mc_valid = false;  // the passes gather the loaded histories again
  if (moment && std::any_of(moment, moment + 3 * m, [](double v) { return v != 0.0; })) {
    const int rk = use_kr(c);
    if (rk) return rk;
  }
  if (matched) *matched = hits;
  return XDEM_OK;
}
