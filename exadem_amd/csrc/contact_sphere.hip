// contact_sphere.hip — sphere contact force pass (contact_sphere / contact_singlemat_sphere_hooke_none,
// src/sphere/contact_force_sphere.cu.in:44-333 and src/sphere/include/exaDEM/sphere/contact_sphere.hpp:121-297).
//
// One thread per particle walks its full (both-direction) neighbour list. An entry the particle owns
// is evaluated with the particle as i; an entry owned by the partner is evaluated in the partner's
// orientation (partner as i) and the particle takes -f and the partner-side moment, exactly as the
// reference's symmetric functor applies them to j. Both copies of the friction history are computed
// from identical operands and stay identical, so no thread ever writes another particle's force:
// there are no atomics and the summation order is fixed (stream order, then walls).
//
// The fused variant also performs reset_force_moment + gravity_force before the contacts and
// force_to_accel + the angular-acceleration part of combined_compute_epilog after them.
#include "contact_law.h"

namespace xdem {

struct Q4f {
  double w, x, y, z;
};
__device__ inline V3 qrot_d(Q4f q, V3 v) {
  const V3 qv{q.x, q.y, q.z};
  V3 uv = cross(qv, v);
  V3 uuv = cross(qv, uv);
  uv = uv * (2.0 * q.w);
  uuv = uuv * 2.0;
  return v + uv + uuv;
}
__device__ inline V3 angular_accel_d(Q4f Q, V3 mom, V3 vrot, V3 I) {
  const Q4f Qi{Q.w, -Q.x, -Q.y, -Q.z};
  const V3 om = qrot_d(Qi, vrot);
  const V3 M = qrot_d(Qi, mom);
  const V3 dom{(M.x - (I.z - I.y) * om.y * om.z) / I.x, (M.y - (I.x - I.z) * om.z * om.x) / I.y,
               (M.z - (I.y - I.x) * om.x * om.y) / I.z};
  return qrot_d(Q, dom);
}

struct ForceArgs {
  int64_t n;
  PState s;
  NbrList nl;
  WallList wl;
  const Surface* walls;
  int nw;
  int64_t wstride;
  Grid g;
  ContactParams cp, cpd;
  double dt;
  V3 grav;
  int gravity_on;
  int32_t* flags;
  int32_t step_tag;
};


// One entry of a particle's list with the partner data it needs, loaded ahead of use.
struct Entry {
  uint32_t wd;
  d4 p;     // partner xr
  d4 v;     // partner vm
  d4 w;     // partner w
  V3 ft;    // friction history (zero when the active bit is clear)
  V3 mint;  // rolling-moment history
};

template <bool KR>
__device__ inline void fetch_entry(const ForceArgs& a, uint64_t sl, uint32_t wd, Entry& e, const Core* tile,
                                   int64_t b0) {
  const Core* __restrict__ core = a.s.core;
  const uint32_t y = wd & kIdxMask;
  e.wd = wd;
  partner_global(core, b0, y, e.p, e.v, e.w);
  if (wd & kActiveBit) {
    e.ft = {a.nl.hx[sl], a.nl.hy[sl], a.nl.hz[sl]};
    if (KR) e.mint = {a.nl.mhx[sl], a.nl.mhy[sl], a.nl.mhz[sl]};
    else e.mint = {0.0, 0.0, 0.0};
  } else {
    e.ft = {0.0, 0.0, 0.0};
    e.mint = {0.0, 0.0, 0.0};
  }
}

template <bool FUSED, bool KR, bool WALLS, bool GUARD>
__global__ __launch_bounds__(256) void k_force_fp64(ForceArgs a) {
  if (GUARD) {
    if (a.flags[1]) return;
    if (a.flags[0]) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.flags[1] = 1;
        a.flags[2] = a.step_tag;
      }
      return;
    }
  }
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t b0 = blk * blockDim.x;
  // the block's own 256 records (32 KB, coalesced 16-byte copies): partners inside the block are
  // read from LDS, the others from global memory
  __shared__ Core tile[256];
  stage_tile(tile, a.s.core, b0, a.n);
  __syncthreads();
  const int64_t x = b0 + threadIdx.x;
  if (x >= a.n) return;
  const PState& s = a.s;
  uint32_t* __restrict__ word = a.nl.word;
  const uint32_t cnt = a.nl.cnt[x];
  const uint64_t b = a.nl.base[x >> 6] + (uint64_t)(x & 63);
  // software pipeline: entry k+1 (record + history) and word k+2 are in flight while k is computed
  Entry nxt;
  uint32_t wd2 = 0;
  if (cnt > 0) fetch_entry<KR>(a, b, word[b], nxt, tile, b0);
  if (cnt > 1) wd2 = word[b + kWave];
  const Core me = tile[threadIdx.x];
  const V3 rx{me.xr.x, me.xr.y, me.xr.z};
  const V3 vx{me.vm.x, me.vm.y, me.vm.z};
  const V3 wx{me.w.x, me.w.y, me.w.z};
  const double mx = me.vm.w, Rx = me.xr.w;
  V3 F, M;
  if (FUSED) {
    F = {0.0, 0.0, 0.0};
    M = {0.0, 0.0, 0.0};
    if (a.gravity_on) F = {F.x + a.grav.x * mx, F.y + a.grav.y * mx, F.z + a.grav.z * mx};
  } else {
    F = {s.ax[x], s.ay[x], s.az[x]};
    M = {s.mx[x], s.my[x], s.mz[x]};
  }
  for (uint32_t k = 0; k < cnt; k++) {
    const uint64_t sl = b + (uint64_t)k * kWave;
    Entry e = nxt;
    if (k + 1 < cnt) {
      fetch_entry<KR>(a, sl + kWave, wd2, nxt, tile, b0);
      if (k + 2 < cnt) wd2 = word[sl + 2 * kWave];
    }
    const uint32_t wd = e.wd;
    partner_tile(tile, b0, wd & kIdxMask, e.p, e.v, e.w);
    const bool owner = (wd & kOwnerBit) != 0;
    V3 ri, rj;
    double Ri, Rj;
    if (owner) {
      const V3 sh = shift_vec(a.g, (int)((wd >> kShiftShift) & 31u));
      ri = rx;
      Ri = Rx;
      rj = {e.p.x + sh.x, e.p.y + sh.y, e.p.z + sh.z};
      Rj = e.p.w;
    } else {
      ri = {e.p.x, e.p.y, e.p.z};
      Ri = e.p.w;
      rj = rx;
      Rj = Rx;
    }
    // detection_vertex_vertex_core
    const V3 dist = ri - rj;
    const double dnorm = norm(dist);
    const double dn = dnorm - (Ri + Rj);
    if (dn < 0) {
      const double inv = 1.0 / dnorm;
      const V3 nrm = dist * inv;
      const V3 cpos = ri - nrm * (Ri + 0.5 * dn);
      const V3 vy{e.v.x, e.v.y, e.v.z}, wy{e.w.x, e.w.y, e.w.z};
      const double my = e.v.w;
      const double mi = owner ? mx : my, mj = owner ? my : mx;
      const double meff = (mi * mj) / (mi + mj);
      V3 ft = e.ft;
      V3 mint = e.mint;
      const V3 f = owner ? hooke_force<KR>(dn, nrm, a.dt, a.cp, meff, ft, mint, cpos, ri, vx, wx, rj, vy, wy)
                         : hooke_force<KR>(dn, nrm, a.dt, a.cp, meff, ft, mint, cpos, ri, vy, wy, rj, vx, wx);
      const bool act = !(ft.x == 0 && ft.y == 0 && ft.z == 0 && mint.x == 0 && mint.y == 0 && mint.z == 0);
      if (act || (wd & kActiveBit)) {
        a.nl.hx[sl] = ft.x;
        a.nl.hy[sl] = ft.y;
        a.nl.hz[sl] = ft.z;
        if (KR) {
          a.nl.mhx[sl] = mint.x;
          a.nl.mhy[sl] = mint.y;
          a.nl.mhz[sl] = mint.z;
        }
      }
      const uint32_t nwd = act ? (wd | kActiveBit) : (wd & ~kActiveBit);
      if (nwd != wd) word[sl] = nwd;
      if (owner) {
        F = F + f;
        M = M + (cross(cpos - rx, f) + mint);
      } else {
        const V3 nf = -f;
        M = M + (cross(cpos - rx, nf) + (-mint));
        F = F + nf;
      }
    } else if (wd & kActiveBit) {
      a.nl.hx[sl] = 0.0;
      a.nl.hy[sl] = 0.0;
      a.nl.hz[sl] = 0.0;
      if (KR) {
        a.nl.mhx[sl] = 0.0;
        a.nl.mhy[sl] = 0.0;
        a.nl.mhz[sl] = 0.0;
      }
      word[sl] = wd & ~kActiveBit;
    }
  }
  if (WALLS) {
    // ContactLawDriverFunc<Surface> (contact_sphere.hpp:216-297) with Surface::detector (surface.hpp:337-350)
    const uint32_t mask = a.wl.mask[x];
    for (int w = 0; w < a.nw; w++) {
      if (!((mask >> w) & 1u)) continue;
      const Surface& sf = a.walls[w];
      const int64_t j = (int64_t)w * a.wstride + x;
      const V3 proj = dot(rx, sf.normal) * sf.normal;
      const V3 s2p = -(sf.center_proj - proj);
      const double d = norm(s2p);
      const double dn = d - Rx;
      if (dn > 0) {
        a.wl.hx[j] = 0.0; a.wl.hy[j] = 0.0; a.wl.hz[j] = 0.0;
        a.wl.mhx[j] = 0.0; a.wl.mhy[j] = 0.0; a.wl.mhz[j] = 0.0;
        continue;
      }
      const V3 nrm = s2p / d;
      const V3 cpos = rx - nrm * (Rx + 0.5 * dn);
      V3 ft{a.wl.hx[j], a.wl.hy[j], a.wl.hz[j]};
      V3 mint{0.0, 0.0, 0.0};
      if (KR) mint = {a.wl.mhx[j], a.wl.mhy[j], a.wl.mhz[j]};
      const V3 f = hooke_force<KR>(dn, nrm, a.dt, a.cpd, mx, ft, mint, cpos, rx, vx, wx, sf.center, sf.vel, sf.vrot);
      a.wl.hx[j] = ft.x; a.wl.hy[j] = ft.y; a.wl.hz[j] = ft.z;
      if (KR) {
        a.wl.mhx[j] = mint.x; a.wl.mhy[j] = mint.y; a.wl.mhz[j] = mint.z;
      }
      F = F + f;
      M = M + (cross(cpos - rx, f) + mint);
    }
  }
  if (FUSED) {
    s.ax[x] = F.x / mx;
    s.ay[x] = F.y / mx;
    s.az[x] = F.z / mx;
    s.mx[x] = M.x;
    s.my[x] = M.y;
    s.mz[x] = M.z;
    const d4 q = s.core[x].q;
    const V3 al = angular_accel_d(Q4f{q.x, q.y, q.z, q.w}, M, wx, V3{s.ix[x], s.iy[x], s.iz[x]});
    s.alx[x] = al.x;
    s.aly[x] = al.y;
    s.alz[x] = al.z;
  } else {
    s.ax[x] = F.x;
    s.ay[x] = F.y;
    s.az[x] = F.z;
    s.mx[x] = M.x;
    s.my[x] = M.y;
    s.mz[x] = M.z;
  }
}

// ---------------------------------------------------------------------------------------------
// Compacted variant: one wavefront per ELL tile (64 particles).
//   pass 1  every lane tests its entries k < cnt for overlap (loads only, independent iterations);
//           the contact pattern of row k is the wavefront ballot rows[k]; prefix[k] counts the
//           contacts of rows < k (tile contacts are numbered row-major: k, then lane);
//   pass 2  the 64 lanes evaluate the tile's contacts 64 at a time in that numbering, write each
//           contribution to LDS, and every lane then adds the contributions that belong to its own
//           particle in increasing k (its stream order) -> same sums, same order as the per-lane loop.
constexpr int kTileMaxK = 64;

__device__ inline int nth_set_bit(uint64_t m, int r) {
  // position of the r-th (0-based) set bit of m
  int pos = 0;
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) {
    const uint64_t low = m & ((w == 64 ? ~0ull : ((1ull << w) - 1ull)));
    const int c = __popcll(low);
    if (r >= c) {
      r -= c;
      m >>= w;
      pos += w;
    } else {
      m = low;
    }
  }
  return pos;
}

struct alignas(16) Contrib {
  double fx, fy, fz, mx, my, mz;
};

template <bool FUSED, bool KR, bool WALLS, bool GUARD>
__global__ __launch_bounds__(64) void k_force_fp64_tile(ForceArgs a) {
  if (GUARD) {
    if (a.flags[1]) return;
    if (a.flags[0]) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.flags[1] = 1;
        a.flags[2] = a.step_tag;
      }
      return;
    }
  }
  __shared__ uint64_t rows[kTileMaxK];
  __shared__ uint32_t prefix[kTileMaxK + 1];
  __shared__ Contrib contrib[64];
  const int64_t tile = xcd_block(blockIdx.x, gridDim.x);
  const int lane = threadIdx.x;
  const int64_t x = tile * 64 + lane;
  const bool live = x < a.n;
  const PState& s = a.s;
  uint32_t* __restrict__ word = a.nl.word;
  const Core* __restrict__ core = s.core;
  const int K = (int)a.nl.kmax[tile];
  const uint64_t tb = a.nl.base[tile];
  const uint32_t cnt = live ? a.nl.cnt[x] : 0u;
  d4 px = {0, 0, 0, 0};
  if (live) px = core[x].xr;
  const V3 rx{px.x, px.y, px.z};
  const double Rx = px.w;
  // ---- pass 1: overlap test of every entry
  uint64_t mycontacts = 0;
  uint32_t run = 0;
  for (int k = 0; k < K; k++) {
    bool contact = false;
    if ((uint32_t)k < cnt) {
      const uint64_t sl = tb + (uint64_t)k * kWave + lane;
      const uint32_t wd = word[sl];
      const uint32_t y = wd & kIdxMask;
      const bool owner = (wd & kOwnerBit) != 0;
      const d4 py = core[y].xr;
      V3 ri, rj;
      double Rs;
      if (owner) {
        const V3 sh = shift_vec(a.g, (int)((wd >> kShiftShift) & 31u));
        ri = rx;
        rj = {py.x + sh.x, py.y + sh.y, py.z + sh.z};
        Rs = Rx + py.w;
      } else {
        ri = {py.x, py.y, py.z};
        rj = rx;
        Rs = py.w + Rx;
      }
      const V3 dist = ri - rj;
      const double dn = norm(dist) - Rs;
      contact = dn < 0;
      if (!contact && (wd & kActiveBit)) {
        a.nl.hx[sl] = 0.0;
        a.nl.hy[sl] = 0.0;
        a.nl.hz[sl] = 0.0;
        if (KR) {
          a.nl.mhx[sl] = 0.0;
          a.nl.mhy[sl] = 0.0;
          a.nl.mhz[sl] = 0.0;
        }
        word[sl] = wd & ~kActiveBit;
      }
    }
    const uint64_t b = __ballot(contact);
    if (contact) mycontacts |= 1ull << k;
    if (lane == 0) {
      rows[k] = b;
      prefix[k] = run;
    }
    run += (uint32_t)__popcll(b);
  }
  if (lane == 0) prefix[K] = run;
  __syncthreads();
  const uint32_t total = run;
  const uint64_t below = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  // own accumulators
  d4 pv = {0, 0, 0, 1};
  d4 pw = {0, 0, 0, 0};
  if (live) {
    pv = s.core[x].vm;
    pw = s.core[x].w;
  }
  const V3 vx{pv.x, pv.y, pv.z};
  const V3 wx{pw.x, pw.y, pw.z};
  const double mx = pv.w;
  V3 F, M;
  if (FUSED) {
    F = {0.0, 0.0, 0.0};
    M = {0.0, 0.0, 0.0};
    if (a.gravity_on) F = {F.x + a.grav.x * mx, F.y + a.grav.y * mx, F.z + a.grav.z * mx};
  } else if (live) {
    F = {s.ax[x], s.ay[x], s.az[x]};
    M = {s.mx[x], s.my[x], s.mz[x]};
  } else {
    F = {0, 0, 0};
    M = {0, 0, 0};
  }
  // my next contact (k) and its tile-wide number
  uint64_t pending = mycontacts;
  for (uint32_t c0 = 0; c0 < total; c0 += 64) {
    // ---- pass 2a: lane evaluates tile contact number g
    const uint32_t g = c0 + (uint32_t)lane;
    if (g < total) {
      int k = 0;
      while (prefix[k + 1] <= g) k++;
      const int plane = nth_set_bit(rows[k], (int)(g - prefix[k]));
      const int64_t xo = tile * 64 + plane;
      const uint64_t sl = tb + (uint64_t)k * kWave + plane;
      const uint32_t wd = word[sl];
      const uint32_t y = wd & kIdxMask;
      const bool owner = (wd & kOwnerBit) != 0;
      const d4 qa = core[xo].xr;
      const d4 qb = core[y].xr;
      const d4 va = s.core[xo].vm;
      const d4 vb = s.core[y].vm;
      const d4 wa = s.core[xo].w;
      const d4 wb = s.core[y].w;
      V3 ri, rj, vi, vj, wi, wj;
      double Ri, Rj, mi, mj;
      if (owner) {
        const V3 sh = shift_vec(a.g, (int)((wd >> kShiftShift) & 31u));
        ri = {qa.x, qa.y, qa.z};
        Ri = qa.w;
        rj = {qb.x + sh.x, qb.y + sh.y, qb.z + sh.z};
        Rj = qb.w;
        vi = {va.x, va.y, va.z};
        vj = {vb.x, vb.y, vb.z};
        wi = {wa.x, wa.y, wa.z};
        wj = {wb.x, wb.y, wb.z};
        mi = va.w;
        mj = vb.w;
      } else {
        ri = {qb.x, qb.y, qb.z};
        Ri = qb.w;
        rj = {qa.x, qa.y, qa.z};
        Rj = qa.w;
        vi = {vb.x, vb.y, vb.z};
        vj = {va.x, va.y, va.z};
        wi = {wb.x, wb.y, wb.z};
        wj = {wa.x, wa.y, wa.z};
        mi = vb.w;
        mj = va.w;
      }
      const V3 dist = ri - rj;
      const double dnorm = norm(dist);
      const double dn = dnorm - (Ri + Rj);
      const double inv = 1.0 / dnorm;
      const V3 nrm = dist * inv;
      const V3 cpos = ri - nrm * (Ri + 0.5 * dn);
      const double meff = (mi * mj) / (mi + mj);
      V3 ft{a.nl.hx[sl], a.nl.hy[sl], a.nl.hz[sl]};
      V3 mint{0.0, 0.0, 0.0};
      if (KR) mint = {a.nl.mhx[sl], a.nl.mhy[sl], a.nl.mhz[sl]};
      const V3 f = hooke_force<KR>(dn, nrm, a.dt, a.cp, meff, ft, mint, cpos, ri, vi, wi, rj, vj, wj);
      a.nl.hx[sl] = ft.x;
      a.nl.hy[sl] = ft.y;
      a.nl.hz[sl] = ft.z;
      if (KR) {
        a.nl.mhx[sl] = mint.x;
        a.nl.mhy[sl] = mint.y;
        a.nl.mhz[sl] = mint.z;
      }
      const bool act = !(ft.x == 0 && ft.y == 0 && ft.z == 0 && mint.x == 0 && mint.y == 0 && mint.z == 0);
      const uint32_t nwd = act ? (wd | kActiveBit) : (wd & ~kActiveBit);
      if (nwd != wd) word[sl] = nwd;
      const V3 ra{qa.x, qa.y, qa.z};
      Contrib cb;
      if (owner) {
        const V3 t = cross(cpos - ra, f) + mint;
        cb = {f.x, f.y, f.z, t.x, t.y, t.z};
      } else {
        const V3 nf = -f;
        const V3 t = cross(cpos - ra, nf) + (-mint);
        cb = {nf.x, nf.y, nf.z, t.x, t.y, t.z};
      }
      contrib[lane] = cb;
    }
    __syncthreads();
    // ---- pass 2b: add my contributions of this chunk, in stream order
    while (pending) {
      const int k = __ffsll((unsigned long long)pending) - 1;
      const uint32_t num = prefix[k] + (uint32_t)__popcll(rows[k] & below);
      if (num >= c0 + 64) break;
      const Contrib& cb = contrib[num - c0];
      F = F + V3{cb.fx, cb.fy, cb.fz};
      M = M + V3{cb.mx, cb.my, cb.mz};
      pending &= pending - 1;
    }
    __syncthreads();
  }
  if (!live) return;
  if (WALLS) {
    const uint32_t mask = a.wl.mask[x];
    for (int w = 0; w < a.nw; w++) {
      if (!((mask >> w) & 1u)) continue;
      const Surface& sf = a.walls[w];
      const int64_t j = (int64_t)w * a.wstride + x;
      const V3 proj = dot(rx, sf.normal) * sf.normal;
      const V3 s2p = -(sf.center_proj - proj);
      const double d = norm(s2p);
      const double dn = d - Rx;
      if (dn > 0) {
        a.wl.hx[j] = 0.0; a.wl.hy[j] = 0.0; a.wl.hz[j] = 0.0;
        a.wl.mhx[j] = 0.0; a.wl.mhy[j] = 0.0; a.wl.mhz[j] = 0.0;
        continue;
      }
      const V3 nrm = s2p / d;
      const V3 cpos = rx - nrm * (Rx + 0.5 * dn);
      V3 ft{a.wl.hx[j], a.wl.hy[j], a.wl.hz[j]};
      V3 mint{0.0, 0.0, 0.0};
      if (KR) mint = {a.wl.mhx[j], a.wl.mhy[j], a.wl.mhz[j]};
      const V3 f = hooke_force<KR>(dn, nrm, a.dt, a.cpd, mx, ft, mint, cpos, rx, vx, wx, sf.center, sf.vel, sf.vrot);
      a.wl.hx[j] = ft.x; a.wl.hy[j] = ft.y; a.wl.hz[j] = ft.z;
      if (KR) {
        a.wl.mhx[j] = mint.x; a.wl.mhy[j] = mint.y; a.wl.mhz[j] = mint.z;
      }
      F = F + f;
      M = M + (cross(cpos - rx, f) + mint);
    }
  }
  if (FUSED) {
    s.ax[x] = F.x / mx;
    s.ay[x] = F.y / mx;
    s.az[x] = F.z / mx;
    s.mx[x] = M.x;
    s.my[x] = M.y;
    s.mz[x] = M.z;
    const d4 q = s.core[x].q;
    const V3 al = angular_accel_d(Q4f{q.x, q.y, q.z, q.w}, M, wx, V3{s.ix[x], s.iy[x], s.iz[x]});
    s.alx[x] = al.x;
    s.aly[x] = al.y;
    s.alz[x] = al.z;
  } else {
    s.ax[x] = F.x;
    s.ay[x] = F.y;
    s.az[x] = F.z;
    s.mx[x] = M.x;
    s.my[x] = M.y;
    s.mz[x] = M.z;
  }
}

template <bool FUSED, bool KR, bool WALLS>
static void launch_g(xdem_ctx* c, const ForceArgs& a, bool guarded) {
  if (c->nl_max_rows <= kTileMaxK && !c->force_per_lane) {
    const dim3 grid(grid_blocks(a.n, 64)), block(64);
    if (guarded) hipLaunchKernelGGL((k_force_fp64_tile<FUSED, KR, WALLS, true>), grid, block, 0, c->stream, a);
    else hipLaunchKernelGGL((k_force_fp64_tile<FUSED, KR, WALLS, false>), grid, block, 0, c->stream, a);
    return;
  }
  const dim3 grid(grid_blocks(a.n, 256)), block(256);
  if (guarded) hipLaunchKernelGGL((k_force_fp64<FUSED, KR, WALLS, true>), grid, block, 0, c->stream, a);
  else hipLaunchKernelGGL((k_force_fp64<FUSED, KR, WALLS, false>), grid, block, 0, c->stream, a);
}

template <bool FUSED>
static void launch_f(xdem_ctx* c, const ForceArgs& a, bool kr, bool walls, bool guarded) {
  if (kr) {
    if (walls) launch_g<FUSED, true, true>(c, a, guarded);
    else launch_g<FUSED, true, false>(c, a, guarded);
  } else {
    if (walls) launch_g<FUSED, false, true>(c, a, guarded);
    else launch_g<FUSED, false, false>(c, a, guarded);
  }
}

int launch_force_fp32(xdem_ctx* c, const ContactParams& cp, const ContactParams& cpd, double dt, bool fused,
                      bool gravity_on, V3 g, bool guarded, int32_t step_tag);

int launch_force(xdem_ctx* c, const ContactParams& cp, const ContactParams& cpd, double dt, int form, bool fused,
                 bool gravity_on, V3 g, bool guarded, int32_t step_tag) {
  if (!c->hist_ready) return set_err(c, XDEM_ERR_STATE, "contact: run nbh_sphere after chunk_neighbors_contact");
  if (form == XDEM_FORM_FP32) return launch_force_fp32(c, cp, cpd, dt, fused, gravity_on, g, guarded, step_tag);
  int rc = set_history_form(c, false);
  if (rc) return rc;
  if (cp.kr != 0.0 || cpd.kr != 0.0) c->kr_used = true;
  const bool walls = c->nwalls > 0 && c->wl_valid;
  ForceArgs a{};
  a.n = c->n;
  a.s = live(c);
  a.nl = c->nl[c->nl_cur];
  if (walls) a.wl = c->wl[c->wl_cur];
  a.walls = c->d_walls;
  a.nw = c->nwalls;
  a.wstride = c->cap;
  a.g = c->grid;
  a.cp = cp;
  a.cpd = cpd;
  a.dt = dt;
  a.grav = g;
  a.gravity_on = gravity_on ? 1 : 0;
  a.flags = c->d_flags;
  a.step_tag = step_tag;
  if (fused) launch_f<true>(c, a, c->kr_used, walls, guarded);
  else launch_f<false>(c, a, c->kr_used, walls, guarded);
  XDEM_HIP(c, hipGetLastError());
  return XDEM_OK;
}

}  // namespace xdem
