// contact_sphere_fp32.hip — FP32-force form of the sphere contact pass.
//
// What stays FP64: absolute positions, the overlap sign test dn = |ri - rj| - (Ri + Rj) < 0 (so the
// contact set is exactly the FP64 one), walls, gravity, force_to_accel, the angular acceleration and
// the integrator. What is FP32: the contact geometry on relative vectors (the FP64 separation rounded
// once; contact arms cp - ri = -n (Ri + dn/2) and cp - rj = (cp - ri) + (ri - rj), never absolute
// coordinates), velocities, the Hooke + Coulomb law (same expression order as
// src/forcefield/include/exaDEM/forcefield/contact_force.hpp:146-235) and the friction history (12 bytes
// per slot instead of 24). Per-particle sums of the contact forces and moments are accumulated in FP32
// with TwoSum compensation and added to the FP64 gravity term at the end.
#include "contact_law.h"
#include "mesh.h"
#include "step.h"

namespace xdem {

struct F3 {
  float x, y, z;
};
__device__ inline F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline F3 operator-(F3 a) { return {-a.x, -a.y, -a.z}; }
__device__ inline F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ inline F3 operator*(float s, F3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ inline float dotf(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline F3 crossf(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ inline F3 to_f3(double x, double y, double z) { return {(float)x, (float)y, (float)z}; }

// TwoSum-compensated accumulator (Knuth): s + c carries the sum with about twice FP32 precision.
struct CSum {
  float s, c;
};
__device__ inline void cadd(CSum& a, float x) {
  const float t = a.s + x;
  const float bp = t - a.s;
  a.c += (a.s - (t - bp)) + (x - bp);
  a.s = t;
}
__device__ inline double cval(const CSum& a) { return (double)a.s + (double)a.c; }

struct ParamsF {
  float kn, kt, kr, mu, damp_rate;
};

// contact_force_core<Hooke,None> in FP32 on relative vectors: bi = cp - pos_i, bj = cp - pos_j.
template <bool KR>
__device__ inline F3 hooke_force_f32(float dn, F3 n, float dt, const ParamsF& p, float meff, F3& ft, F3& mom,
                                     F3 bi, F3 bj, F3 vel_i, F3 vrot_i, F3 vel_j, F3 vrot_j) {
  const float damp = p.damp_rate * 2.0f * sqrtf(p.kn * meff);
  const F3 ci = vel_i - crossf(bi, vrot_i);
  const F3 cj = vel_j - crossf(bj, vrot_j);
  const F3 vel = cj - ci;
  const float vn = dotf(vel, n);
  const float fne = -p.kn * dn;
  const float fnv = damp * vn;
  const float fnval = (fnv + fne > 0.0f) ? fnv + fne : 0.0f;
  const F3 fn = fnval * n;
  const F3 vt = vel - vn * n;
  ft = ft + p.kt * (dt * vt);
  const float thr = fabsf(p.mu * (p.kn * dn));
  const float ft2 = dotf(ft, ft);
  if (ft2 > 0.0f && ft2 > thr * thr) ft = ft * (thr / sqrtf(ft2));
  const F3 f = fn + ft;
  if (KR) {
    mom = mom + (p.kr * (vrot_j - vrot_i)) * dt;
    F3 branch = bi;
    const float r = dotf(branch, vrot_i) / dotf(vrot_i, vrot_i);
    branch = branch - r * vrot_i;
    const float thr_m = fabsf(0.0f * sqrtf(dotf(branch, branch)) * sqrtf(dotf(fn, fn)));
    const float m2 = dotf(mom, mom);
    if (m2 > 0.0f && m2 > thr_m * thr_m) mom = mom * (thr_m / sqrtf(m2));
  }
  return f;
}

struct ForceArgs32 {
  int64_t n;
  PState s;
  NbrList nl;
  WallList wl;
  const Surface* walls;
  int nw;
  int64_t wstride;
  Grid g;
  ContactParams cpd;  // walls stay FP64
  ContactParams cp64;
  ParamsF cp;
  double dt;
  float dtf;
  V3 grav;
  int gravity_on;
  int32_t* flags;
  int32_t step_tag;
  int32_t trig_in;  // flags slot of the rebuild trigger this pass obeys
  // integrating pass only (see ForceArgs in contact_sphere.hip): next records, kick dt/2, squared
  // displacement threshold, trigger slot raised
  Core* core_out;
  double kdt_2;
  double thr2;
  int32_t trig_out;
  // terms of the triangle-mesh contacts, evaluated by the mesh pass just before (ncand = 0: none)
  MeshTerms mt;
};


struct Q4g {
  double w, x, y, z;
};
__device__ inline V3 qrot_g(Q4g q, V3 v) {
  const V3 qv{q.x, q.y, q.z};
  V3 uv = cross(qv, v);
  V3 uuv = cross(qv, uv);
  uv = uv * (2.0 * q.w);
  uuv = uuv * 2.0;
  return v + uv + uuv;
}
__device__ inline V3 angular_accel_g(Q4g Q, V3 mom, V3 vrot, V3 I) {
  const Q4g Qi{Q.w, -Q.x, -Q.y, -Q.z};
  const V3 om = qrot_g(Qi, vrot);
  const V3 M = qrot_g(Qi, mom);
  const V3 dom{(M.x - (I.z - I.y) * om.y * om.z) / I.x, (M.y - (I.x - I.z) * om.z * om.x) / I.y,
               (M.z - (I.y - I.x) * om.x * om.y) / I.z};
  return qrot_g(Q, dom);
}

struct Entry32 {
  uint32_t wd;
  d4 p;
  d4 v;
  d4 w;
  F3 ft;
  F3 mint;
};

template <bool KR>
__device__ inline void fetch32(const ForceArgs32& a, uint64_t sl, uint32_t wd, Entry32& e, const Core* tile,
                               int64_t b0) {
  const Core* __restrict__ core = a.s.core;
  const uint32_t y = wd & kIdxMask;
  e.wd = wd;
  partner_global(core, b0, y, e.p, e.v, e.w);
  if (wd & kActiveBit) {
    e.ft = {a.nl.fx32[sl], a.nl.fy32[sl], a.nl.fz32[sl]};
    if (KR) e.mint = {a.nl.mfx32[sl], a.nl.mfy32[sl], a.nl.mfz32[sl]};
    else e.mint = {0.0f, 0.0f, 0.0f};
  } else {
    e.ft = {0.0f, 0.0f, 0.0f};
    e.mint = {0.0f, 0.0f, 0.0f};
  }
}

// INTEG (with FUSED): as k_force_fp64's integrating pass, the kick and next prolog of each particle
// go to core_out right after its force.
// Four wavefronts per SIMD: the pass needs 130 VGPRs unconstrained (three wavefronts); capped at 128
// it spills nothing and hides more of the partner-gather latency (integrating pass 164.9 -> 154.9 us
// at 1M spheres).
// MESH: add the mesh terms of the pass (the particle's index among those with candidates is read ahead)
template <bool FUSED, bool KR, bool WALLS, bool GUARD, bool INTEG = false, bool MESH = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_force_fp32(ForceArgs32 a) {
  if (GUARD) {
    if (a.flags[1]) return;
    if (a.flags[a.trig_in]) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.flags[1] = 1;
        a.flags[2] = a.step_tag;
      }
      return;
    }
  }
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t b0 = blk * blockDim.x;
  // the block's own 256 records in LDS: partners inside the block are read from there
  __shared__ Core tile[256];
  stage_tile(tile, a.s.core, b0, a.n);
  __syncthreads();
  const int64_t x = b0 + threadIdx.x;
  if (x >= a.n) return;
  // mesh terms: only waves of 64 particles that have candidates read their particles' index (one
  // scalar load per wave)
  const bool mwave = MESH && a.mt.ncand && a.mt.wcnt[__builtin_amdgcn_readfirstlane((int)(x >> 6))] != 0u;
  const uint2 midx = mwave ? a.mt.idx[x] : uint2{0u, 0u};
  const PState& s = a.s;
  uint32_t* __restrict__ word = a.nl.word;
  const uint32_t cnt = a.nl.cnt[x];
  const uint64_t b = a.nl.base[x >> 6] + (uint64_t)(x & 63);
  Entry32 nxt;
  uint32_t wd2 = 0;
  if (cnt > 0) fetch32<KR>(a, b, word[b], nxt, tile, b0);
  if (cnt > 1) wd2 = word[b + kWave];
  const Core me = tile[threadIdx.x];
  const V3 rx{me.xr.x, me.xr.y, me.xr.z};
  const double Rx = me.xr.w, mx = me.vm.w;
  const F3 vx = to_f3(me.vm.x, me.vm.y, me.vm.z);
  const F3 wx = to_f3(me.w.x, me.w.y, me.w.z);
  const float mxf = (float)mx;
  CSum Fx{0, 0}, Fy{0, 0}, Fz{0, 0}, Mx{0, 0}, My{0, 0}, Mz{0, 0};
  for (uint32_t k = 0; k < cnt; k++) {
    const uint64_t sl = b + (uint64_t)k * kWave;
    Entry32 e = nxt;
    if (k + 1 < cnt) {
      fetch32<KR>(a, sl + kWave, wd2, nxt, tile, b0);
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
    // FP64 overlap sign test: the contact set is the plain form's
    const V3 dist = ri - rj;
    const double dnorm = sqrt(dot(dist, dist));
    const double dn = dnorm - (Ri + Rj);
    if (dn < 0) {
      const F3 d = to_f3(dist.x, dist.y, dist.z);
      const float dnf = (float)dn;
      const float inv = 1.0f / (float)dnorm;
      const F3 nrm = d * inv;
      const F3 bi = -(nrm * ((float)Ri + 0.5f * dnf));  // cp - ri
      const F3 bj = bi + d;                            // cp - rj
      const F3 vy = to_f3(e.v.x, e.v.y, e.v.z);
      const F3 wy = to_f3(e.w.x, e.w.y, e.w.z);
      const float my = (float)e.v.w;
      const float mi = owner ? mxf : my, mj = owner ? my : mxf;
      const float meff = (mi * mj) / (mi + mj);
      F3 ft = e.ft;
      F3 mint = e.mint;
      const F3 f = owner ? hooke_force_f32<KR>(dnf, nrm, a.dtf, a.cp, meff, ft, mint, bi, bj, vx, wx, vy, wy)
                         : hooke_force_f32<KR>(dnf, nrm, a.dtf, a.cp, meff, ft, mint, bi, bj, vy, wy, vx, wx);
      const bool act = !(ft.x == 0 && ft.y == 0 && ft.z == 0 && mint.x == 0 && mint.y == 0 && mint.z == 0);
      if (act || (wd & kActiveBit)) {
        a.nl.fx32[sl] = ft.x;
        a.nl.fy32[sl] = ft.y;
        a.nl.fz32[sl] = ft.z;
        if (KR) {
          a.nl.mfx32[sl] = mint.x;
          a.nl.mfy32[sl] = mint.y;
          a.nl.mfz32[sl] = mint.z;
        }
      }
      const uint32_t nwd = act ? (wd | kActiveBit) : (wd & ~kActiveBit);
      if (nwd != wd) word[sl] = nwd;
      // Both contact arms are parallel to n, so the normal force has no moment about either centre:
      // (cp - r) x f = (cp - r) x ft exactly. The FP64 form evaluates the full product, whose normal
      // part there is round-off of ~1e-16 |b| |fn|; in FP32 the same round-off (~1e-7 |b| |fn|) would
      // swamp the friction moment, so only the tangential part enters.
      F3 fc, tc;
      if (owner) {
        fc = f;
        tc = crossf(bi, ft) + mint;
      } else {
        fc = -f;
        tc = crossf(bj, -ft) + (-mint);
      }
      cadd(Fx, fc.x);
      cadd(Fy, fc.y);
      cadd(Fz, fc.z);
      cadd(Mx, tc.x);
      cadd(My, tc.y);
      cadd(Mz, tc.z);
    } else if (wd & kActiveBit) {
      a.nl.fx32[sl] = 0.0f;
      a.nl.fy32[sl] = 0.0f;
      a.nl.fz32[sl] = 0.0f;
      if (KR) {
        a.nl.mfx32[sl] = 0.0f;
        a.nl.mfy32[sl] = 0.0f;
        a.nl.mfz32[sl] = 0.0f;
      }
      word[sl] = wd & ~kActiveBit;
    }
  }
  V3 F, M;
  if (FUSED) {
    F = {0.0, 0.0, 0.0};
    M = {0.0, 0.0, 0.0};
    if (a.gravity_on) F = {F.x + a.grav.x * mx, F.y + a.grav.y * mx, F.z + a.grav.z * mx};
  } else {
    F = {s.ax[x], s.ay[x], s.az[x]};
    M = {s.mx[x], s.my[x], s.mz[x]};
  }
  F = F + V3{cval(Fx), cval(Fy), cval(Fz)};
  M = M + V3{cval(Mx), cval(My), cval(Mz)};
  // the integration's own inputs (inertia, rebuild reference), loaded before the walls and the mesh terms
  // so that their latency overlaps with those
  V3 inert{0.0, 0.0, 0.0}, bref{0.0, 0.0, 0.0};
  if (FUSED && INTEG) {
    inert = {s.ix[x], s.iy[x], s.iz[x]};
    bref = {s.bx[x], s.by[x], s.bz[x]};
  }
  if (WALLS) {
    const V3 vx64{me.vm.x, me.vm.y, me.vm.z}, wx64{me.w.x, me.w.y, me.w.z};
    const uint32_t mask = a.nw ? a.wl.mask[x] : 0u;
    for (int w = 0; w < a.nw; w++) {
      if (!((mask >> w) & 1u)) continue;
      const Surface& sf = a.walls[w];
      const int64_t j = (int64_t)w * a.wstride + x;
      double dnw;
      V3 nrm, cpos;
      if (!driver_detect(sf, rx, Rx, dnw, nrm, cpos)) {
        a.wl.hx[j] = 0.0; a.wl.hy[j] = 0.0; a.wl.hz[j] = 0.0;
        a.wl.mhx[j] = 0.0; a.wl.mhy[j] = 0.0; a.wl.mhz[j] = 0.0;
        continue;
      }
      V3 ft{a.wl.hx[j], a.wl.hy[j], a.wl.hz[j]};
      V3 mint{0.0, 0.0, 0.0};
      if (KR) mint = {a.wl.mhx[j], a.wl.mhy[j], a.wl.mhz[j]};
      const V3 f =
          hooke_force<KR>(dnw, nrm, a.dt, a.cpd, mx, ft, mint, cpos, rx, vx64, wx64, sf.center, sf.vel, sf.vrot);
      a.wl.hx[j] = ft.x; a.wl.hy[j] = ft.y; a.wl.hz[j] = ft.z;
      if (KR) {
        a.wl.mhx[j] = mint.x; a.wl.mhy[j] = mint.y; a.wl.mhz[j] = mint.z;
      }
      F = F + f;
      M = M + (cross(cpos - rx, f) + mint);
    }
  }
  if (MESH && mwave) mesh_add_at(a.mt, midx, F, M);
  if (FUSED && INTEG) {
    const V3 al = angular_accel_g(Q4g{me.q.x, me.q.y, me.q.z, me.q.w}, M, V3{me.w.x, me.w.y, me.w.z}, inert);
    Core nx = me;
    kick_prolog<true>(nx.xr, nx.vm, nx.w, nx.q, F.x / mx, F.y / mx, F.z / mx, al.x, al.y, al.z, a.dt, a.kdt_2);
    a.core_out[x] = nx;
    const double dx = nx.xr.x - bref.x, dy = nx.xr.y - bref.y, dz = nx.xr.z - bref.z;
    if (dx * dx + dy * dy + dz * dz >= a.thr2) a.flags[a.trig_out] = 1;
  } else if (FUSED) {
    s.ax[x] = F.x / mx;
    s.ay[x] = F.y / mx;
    s.az[x] = F.z / mx;
    s.mx[x] = M.x;
    s.my[x] = M.y;
    s.mz[x] = M.z;
    const d4 q = me.q;
    const V3 al =
        angular_accel_g(Q4g{q.x, q.y, q.z, q.w}, M, V3{me.w.x, me.w.y, me.w.z}, V3{s.ix[x], s.iy[x], s.iz[x]});
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

// history representation switch (FP64 <-> FP32) over every slot of the current list
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void k_hist_convert(int64_t slots, const TI* __restrict__ ix, const TI* __restrict__ iy,
                                                      const TI* __restrict__ iz, TO* __restrict__ ox, TO* __restrict__ oy,
                                                      TO* __restrict__ oz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= slots) return;
  ox[i] = (TO)ix[i];
  oy[i] = (TO)iy[i];
  oz[i] = (TO)iz[i];
}

int set_history_form(xdem_ctx* c, bool fp32) {
  if (c->hist_fp32 == fp32) return XDEM_OK;
  NbrList& L = c->nl[c->nl_cur];
  if (c->nl_valid && L.slots > 0) {
    const dim3 grid(grid_blocks(L.slots, 256)), block(256);
    if (fp32) {
      hipLaunchKernelGGL((k_hist_convert<double, float>), grid, block, 0, c->stream, L.slots, L.hx, L.hy, L.hz,
                         L.fx32, L.fy32, L.fz32);
      hipLaunchKernelGGL((k_hist_convert<double, float>), grid, block, 0, c->stream, L.slots, L.mhx, L.mhy, L.mhz,
                         L.mfx32, L.mfy32, L.mfz32);
    } else {
      hipLaunchKernelGGL((k_hist_convert<float, double>), grid, block, 0, c->stream, L.slots, L.fx32, L.fy32,
                         L.fz32, L.hx, L.hy, L.hz);
      hipLaunchKernelGGL((k_hist_convert<float, double>), grid, block, 0, c->stream, L.slots, L.mfx32, L.mfy32,
                         L.mfz32, L.mhx, L.mhy, L.mhz);
    }
    XDEM_HIP(c, hipGetLastError());
  }
  c->hist_fp32 = fp32;
  return XDEM_OK;
}

template <bool FUSED, bool KR, bool WALLS>
static void launch32(xdem_ctx* c, const ForceArgs32& a, bool guarded) {
  const dim3 grid(grid_blocks(a.n, 256)), block(256);
  if (guarded) hipLaunchKernelGGL((k_force_fp32<FUSED, KR, WALLS, true, false, WALLS>), grid, block, 0, c->stream, a);
  else hipLaunchKernelGGL((k_force_fp32<FUSED, KR, WALLS, false, false, WALLS>), grid, block, 0, c->stream, a);
}

int launch_force_fp32(xdem_ctx* c, const ContactParams& cp, const ContactParams& cpd, double dt, bool fused,
                      bool gravity_on, V3 g, bool guarded, int32_t step_tag) {
  int rc = set_history_form(c, true);
  if (rc) return rc;
  if ((cp.kr != 0.0 || cpd.kr != 0.0) && (rc = use_kr(c))) return rc;
  const bool planes = c->nwalls > 0 && c->wl_valid;
  const bool meshed = c->has_mesh && c->ml_valid;
  const bool walls = planes || meshed;
  ForceArgs32 a{};
  a.n = c->n;
  a.s = live(c);
  a.nl = c->nl[c->nl_cur];
  if (planes) a.wl = c->wl[c->wl_cur];
  a.walls = walls_for_pass(c);
  a.nw = planes ? c->nwalls : 0;
  if (meshed && (rc = launch_mesh_pass(c, cpd, dt, guarded, c->trig_slot, nullptr, &a.mt))) return rc;
  a.wstride = c->cap;
  a.g = c->grid;
  a.cpd = cpd;
  a.cp64 = cp;
  a.cp = {(float)cp.kn, (float)cp.kt, (float)cp.kr, (float)cp.mu, (float)cp.damp_rate};
  a.dt = dt;
  a.dtf = (float)dt;
  a.grav = g;
  a.gravity_on = gravity_on ? 1 : 0;
  a.flags = c->d_flags;
  a.step_tag = step_tag;
  a.trig_in = c->trig_slot;
  const bool kr = c->kr_used;
  if (fused) {
    if (kr) walls ? launch32<true, true, true>(c, a, guarded) : launch32<true, true, false>(c, a, guarded);
    else walls ? launch32<true, false, true>(c, a, guarded) : launch32<true, false, false>(c, a, guarded);
  } else {
    if (kr) walls ? launch32<false, true, true>(c, a, guarded) : launch32<false, true, false>(c, a, guarded);
    else walls ? launch32<false, false, true>(c, a, guarded) : launch32<false, false, false>(c, a, guarded);
  }
  XDEM_HIP(c, hipGetLastError());
  c->cs_valid = false;  // the FP32 form does not record
  return XDEM_OK;
}

// FP32-form engine step in one kernel (see launch_step)
int launch_step_fp32(xdem_ctx* c, const ContactParams& cp, const ContactParams& cpd, double dt, bool gravity_on, V3 g,
                     double thr2, int32_t step_tag, Core** input) {
  int rc = set_history_form(c, true);
  if (rc) return rc;
  if ((cp.kr != 0.0 || cpd.kr != 0.0) && (rc = use_kr(c))) return rc;
  const bool planes = c->nwalls > 0 && c->wl_valid;
  const bool meshed = c->has_mesh && c->ml_valid;
  ForceArgs32 a{};
  a.n = c->n;
  a.s = live(c);
  a.nl = c->nl[c->nl_cur];
  if (planes) a.wl = c->wl[c->wl_cur];
  a.walls = walls_for_pass(c);
  a.nw = planes ? c->nwalls : 0;
  if (meshed && (rc = launch_mesh_pass(c, cpd, dt, true, c->trig_slot, nullptr, &a.mt))) return rc;
  a.wstride = c->cap;
  a.g = c->grid;
  a.cpd = cpd;
  a.cp64 = cp;
  a.cp = {(float)cp.kn, (float)cp.kt, (float)cp.kr, (float)cp.mu, (float)cp.damp_rate};
  a.dt = dt;
  a.dtf = (float)dt;
  a.grav = g;
  a.gravity_on = gravity_on ? 1 : 0;
  a.flags = c->d_flags;
  a.step_tag = step_tag;
  a.trig_in = c->trig_slot;
  a.trig_out = c->trig_slot == 0 ? 3 : 0;
  a.core_out = c->core_alt;
  a.kdt_2 = dt * 0.5;
  a.thr2 = thr2;
  const dim3 grid(grid_blocks(a.n, 256)), block(256);
#define STEP(KRV, W, M) hipLaunchKernelGGL((k_force_fp32<true, KRV, W, true, true, M>), grid, block, 0, c->stream, a)
  if (c->kr_used) {
    if (planes) { if (meshed) STEP(true, true, true); else STEP(true, true, false); }
    else { if (meshed) STEP(true, false, true); else STEP(true, false, false); }
  } else {
    if (planes) { if (meshed) STEP(false, true, true); else STEP(false, true, false); }
    else { if (meshed) STEP(false, false, true); else STEP(false, false, false); }
  }
#undef STEP
  XDEM_HIP(c, hipGetLastError());
  *input = live(c).core;
  std::swap(live(c).core, c->core_alt);
  c->trig_slot = a.trig_out;
  c->cs_valid = false;  // the FP32 form does not record
  return XDEM_OK;
}

}  // namespace xdem
