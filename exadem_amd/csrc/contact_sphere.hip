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
#include <algorithm>
#include <cmath>

#include "contact_law.h"
#include "mesh.h"
#include "step.h"

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
  // flags slot of the rebuild trigger this pass obeys (raised by the pass before it)
  int32_t trig_in;
  // integrating pass only: the next state goes to core_out (the records of this pass are still being
  // read as partners), kick dt/2 of the finished step, squared displacement threshold, and the
  // trigger slot it raises (another slot than trig_in: blocks of this pass that start late must not
  // see the trigger of this pass)
  Core* core_out;
  double kdt_2;
  double thr2;
  int32_t trig_out;
  // passes outside the integrating step, when recording: contact state per owner slot and per
  // (wall, particle), kCStateLen doubles each (null: not recorded)
  double* cstate;
  double* wstate;
  // terms of the triangle-mesh contacts, evaluated by the mesh pass just before (ncand = 0: none)
  MeshTerms mt;
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

// INTEG (with FUSED): after the particle's force, its epilog kick and the prolog of the next step
// are applied and the next record is written to core_out, with the displacement test of the rebuild
// trigger; accelerations, moments and angular accelerations are then not stored (the next pass
// recomputes them).
// MESH: add the mesh terms of the pass (the particle's index among those with candidates is read ahead)
template <bool FUSED, bool KR, bool WALLS, bool GUARD, bool INTEG = false, bool MESH = false>
__global__ __launch_bounds__(256) void k_force_fp64(ForceArgs a) {
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
  // the block's own 256 records (32 KB, coalesced 16-byte copies): partners inside the block are
  // read from LDS, the others from global memory
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
      if (!INTEG && a.cstate && owner) {
        // the tuple the reference's functor returns for analyses: dn, contact point, f - ft, ft
        // (contact_sphere.hpp:175-202), with the two positions the interaction dump writes
        const V3 fnv = f - ft;
        const double rec[kCStateLen] = {dn,   cpos.x, cpos.y, cpos.z, fnv.x, fnv.y, fnv.z, ft.x,
                                        ft.y, ft.z,   ri.x,   ri.y,   ri.z,   rj.x,  rj.y,  rj.z};
        double* r = a.cstate + (uint64_t)kCStateLen * sl;
#pragma unroll
        for (int q = 0; q < kCStateLen; q++) r[q] = rec[q];
      }
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
    } else {
      if (wd & kActiveBit) {
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
      if (!INTEG && a.cstate && owner) {
        // no contact: dn reported 0, no forces (contact_sphere.hpp:197-200); the dump skips the entry
        double* r = a.cstate + (uint64_t)kCStateLen * sl;
#pragma unroll
        for (int q = 0; q < kCStateLen; q++) r[q] = 0.0;
      }
    }
  }
  // the integration's own inputs (inertia, rebuild reference), loaded before the walls and the mesh terms
  // so that their latency overlaps with those
  V3 inert{0.0, 0.0, 0.0}, bref{0.0, 0.0, 0.0};
  if (FUSED && INTEG) {
    inert = {s.ix[x], s.iy[x], s.iz[x]};
    bref = {s.bx[x], s.by[x], s.bz[x]};
  }
  if (WALLS) {
    // ContactLawDriverFunc<Surface / Cylinder / Ball> (contact_sphere.hpp:216-297) with the driver's
    // detector (driver_detect)
    const uint32_t mask = a.nw ? a.wl.mask[x] : 0u;
    for (int w = 0; w < a.nw; w++) {
      if (!((mask >> w) & 1u)) continue;
      const Surface& sf = a.walls[w];
      const int64_t j = (int64_t)w * a.wstride + x;
      double dn;
      V3 nrm, cpos;
      if (!driver_detect(sf, rx, Rx, dn, nrm, cpos)) {
        a.wl.hx[j] = 0.0; a.wl.hy[j] = 0.0; a.wl.hz[j] = 0.0;
        a.wl.mhx[j] = 0.0; a.wl.mhy[j] = 0.0; a.wl.mhz[j] = 0.0;
        if (!INTEG && a.wstate) {
          double* r = a.wstate + (uint64_t)kCStateLen * j;
#pragma unroll
          for (int q = 0; q < kCStateLen; q++) r[q] = 0.0;
        }
        continue;
      }
      V3 ft{a.wl.hx[j], a.wl.hy[j], a.wl.hz[j]};
      V3 mint{0.0, 0.0, 0.0};
      if (KR) mint = {a.wl.mhx[j], a.wl.mhy[j], a.wl.mhz[j]};
      const V3 f = hooke_force<KR>(dn, nrm, a.dt, a.cpd, mx, ft, mint, cpos, rx, vx, wx, sf.center, sf.vel, sf.vrot);
      if (!INTEG && a.wstate) {
        const V3 fnv = f - ft;
        const double rec[kCStateLen] = {dn,   cpos.x, cpos.y, cpos.z, fnv.x, fnv.y, fnv.z, ft.x,
                                        ft.y, ft.z,   rx.x,   rx.y,   rx.z,   0.0,   0.0,   0.0};
        double* r = a.wstate + (uint64_t)kCStateLen * j;
#pragma unroll
        for (int q = 0; q < kCStateLen; q++) r[q] = rec[q];
      }
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
    const V3 al = angular_accel_d(Q4f{me.q.x, me.q.y, me.q.z, me.q.w}, M, wx, inert);
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

// operator path: the driver variant (WALLS) also adds the mesh terms
template <bool FUSED, bool KR, bool WALLS>
static void launch_g(xdem_ctx* c, const ForceArgs& a, bool guarded) {
  const dim3 grid(grid_blocks(a.n, 256)), block(256);
  if (guarded) hipLaunchKernelGGL((k_force_fp64<FUSED, KR, WALLS, true, false, WALLS>), grid, block, 0, c->stream, a);
  else hipLaunchKernelGGL((k_force_fp64<FUSED, KR, WALLS, false, false, WALLS>), grid, block, 0, c->stream, a);
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

// One step of the engine in one kernel: force pass of the current state, then each particle's kick
// and next prolog into the other record buffer (c->core_alt), which becomes the live one. *input
// receives the record buffer the pass read (it holds this step's state if the pass stalls).
int launch_step_fp32(xdem_ctx* c, const ContactParams& cp, const ContactParams& cpd, double dt, bool gravity_on, V3 g,
                     double thr2, int32_t step_tag, Core** input);

int launch_step(xdem_ctx* c, const ContactParams& cp, const ContactParams& cpd, double dt, int form, bool gravity_on,
                V3 g, double thr2, int32_t step_tag, Core** input) {
  *input = nullptr;
  if (c->n == 0) return XDEM_OK;  // the kernels stage records n-1 and below
  if (!c->hist_ready) return set_err(c, XDEM_ERR_STATE, "contact: run nbh_sphere after chunk_neighbors_contact");
  if (form == XDEM_FORM_FP32) return launch_step_fp32(c, cp, cpd, dt, gravity_on, g, thr2, step_tag, input);
  int rc = set_history_form(c, false);
  if (rc) return rc;
  if ((cp.kr != 0.0 || cpd.kr != 0.0) && (rc = use_kr(c))) return rc;
  const bool planes = c->nwalls > 0 && c->wl_valid;
  const bool meshed = c->has_mesh && c->ml_valid;
  ForceArgs a{};
  a.n = c->n;
  a.s = live(c);
  a.nl = c->nl[c->nl_cur];
  if (planes) a.wl = c->wl[c->wl_cur];
  a.walls = walls_for_pass(c);
  a.nw = planes ? c->nwalls : 0;
  if (meshed && (rc = launch_mesh_pass(c, cpd, dt, true, c->trig_slot, nullptr, &a.mt))) return rc;
  a.wstride = c->cap;
  a.g = c->grid;
  a.cp = cp;
  a.cpd = cpd;
  a.dt = dt;
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
  // engine step: drivers and mesh terms each only when present
#define STEP(KRV, W, M) hipLaunchKernelGGL((k_force_fp64<true, KRV, W, true, true, M>), grid, block, 0, c->stream, a)
  if (c->kr_used) {
    if (planes) { if (meshed) STEP(true, true, true); else STEP(true, true, false); }
    else { if (meshed) STEP(true, false, true); else STEP(true, false, false); }
  } else {
    if (planes) { if (meshed) STEP(false, true, true); else STEP(false, true, false); }
    else { if (meshed) STEP(false, false, true); else STEP(false, false, false); }
  }
#undef STEP
  XDEM_HIP(c, hipGetLastError());
  c->cs_valid = false;  // the integrating pass does not record
  *input = live(c).core;
  std::swap(live(c).core, c->core_alt);
  c->trig_slot = a.trig_out;
  return XDEM_OK;
}

int launch_force(xdem_ctx* c, const ContactParams& cp, const ContactParams& cpd, double dt, int form, bool fused,
                 bool gravity_on, V3 g, bool guarded, int32_t step_tag) {
  if (c->n == 0) return XDEM_OK;  // the kernels stage records n-1 and below
  if (!c->hist_ready) return set_err(c, XDEM_ERR_STATE, "contact: run nbh_sphere after chunk_neighbors_contact");
  if (form == XDEM_FORM_FP32) return launch_force_fp32(c, cp, cpd, dt, fused, gravity_on, g, guarded, step_tag);
  int rc = set_history_form(c, false);
  if (rc) return rc;
  if ((cp.kr != 0.0 || cpd.kr != 0.0) && (rc = use_kr(c))) return rc;
  const bool planes = c->nwalls > 0 && c->wl_valid;
  const bool meshed = c->has_mesh && c->ml_valid;
  const bool walls = planes || meshed;
  ForceArgs a{};
  a.n = c->n;
  a.s = live(c);
  a.nl = c->nl[c->nl_cur];
  if (planes) a.wl = c->wl[c->wl_cur];
  a.walls = walls_for_pass(c);
  a.nw = planes ? c->nwalls : 0;
  double* mstate = nullptr;
  a.wstride = c->cap;
  a.g = c->grid;
  a.cp = cp;
  a.cpd = cpd;
  a.dt = dt;
  a.grav = g;
  a.gravity_on = gravity_on ? 1 : 0;
  a.flags = c->d_flags;
  a.step_tag = step_tag;
  a.trig_in = c->trig_slot;
  if (c->record_cs) {
    const NbrList& L = c->nl[c->nl_cur];
    if (L.slots > c->cap_cstate) {
      const int64_t cap = (int64_t)(L.slots * 1.25) + 4096;
      if ((rc = dev_realloc(c, &c->cstate, cap * kCStateLen))) return rc;
      c->cap_cstate = cap;
    }
    a.cstate = c->cstate;
    if (planes) {
      const int64_t m = (int64_t)c->nwalls * c->cap;
      if (m > c->cap_wstate) {
        if ((rc = dev_realloc(c, &c->wstate, m * kCStateLen))) return rc;
        c->cap_wstate = m;
      }
      a.wstate = c->wstate;
    }
    if (meshed) {
      const int64_t m = (int64_t)c->mesh_cap * c->cap;
      if (m > c->cap_mstate) {
        if ((rc = dev_realloc(c, &c->mstate, m * kCStateLen))) return rc;
        c->cap_mstate = m;
      }
      mstate = c->mstate;
    }
  }
  if (meshed && (rc = launch_mesh_pass(c, cpd, dt, guarded, c->trig_slot, mstate, &a.mt))) return rc;
  if (fused) launch_f<true>(c, a, c->kr_used, walls, guarded);
  else launch_f<false>(c, a, c->kr_used, walls, guarded);
  XDEM_HIP(c, hipGetLastError());
  c->cs_valid = c->record_cs;
  c->cs_epoch = c->nl_epoch;
  return XDEM_OK;
}

}  // namespace xdem

namespace xdem {
// Rows of the recorded contact state that the reference's interaction dump writes
// (src/interaction/include/exaDEM/itools/writer.hpp:35-75): owner-side interactions with a non-zero
// force or a negative dn; particle pairs in device order, then walls.
static int contact_state_rows(xdem_ctx* c, std::vector<uint64_t>& ii, std::vector<uint64_t>& jj,
                              std::vector<uint32_t>& si, std::vector<uint32_t>& sj, std::vector<uint16_t>& tt,
                              std::vector<double>& vals) {
  if (!c->record_cs) return set_err(c, XDEM_ERR_STATE, "contact state: recording is off (xdem_set_contact_record)");
  if (!c->nl_valid || c->n == 0) return XDEM_OK;
  if (!c->cs_valid || c->cs_epoch != c->nl_epoch)
    return set_err(c, XDEM_ERR_STATE, "contact state: the last force pass was not a recorded FP64 pass");
  if (c->nshapes > 0) return poly_contact_state_rows(c, ii, jj, si, sj, tt, vals);
  const int64_t n = c->n;
  const NbrList& L = c->nl[c->nl_cur];
  const int64_t nw = (n + kWave - 1) / kWave;
  std::vector<uint32_t> cnt(n), word(L.slots);
  std::vector<uint64_t> base(nw + 1), ids(n);
  std::vector<double> cs((size_t)L.slots * kCStateLen);
  XDEM_HIP(c, hipMemcpyAsync(cnt.data(), L.cnt, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(base.data(), L.base, sizeof(uint64_t) * (nw + 1), hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(ids.data(), live(c).id, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, c->stream));
  if (L.slots > 0) {
    XDEM_HIP(c, hipMemcpyAsync(word.data(), L.word, sizeof(uint32_t) * L.slots, hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipMemcpyAsync(cs.data(), c->cstate, sizeof(double) * kCStateLen * L.slots, hipMemcpyDeviceToHost,
                               c->stream));
  }
  const bool walls = c->nwalls > 0 && c->wl_valid && c->wstate;
  std::vector<uint8_t> mask;
  std::vector<double> ws;
  if (walls) {
    mask.resize(n);
    ws.resize((size_t)c->nwalls * c->cap * kCStateLen);
    XDEM_HIP(c, hipMemcpyAsync(mask.data(), c->wl[c->wl_cur].mask, n, hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipMemcpyAsync(ws.data(), c->wstate, sizeof(double) * ws.size(), hipMemcpyDeviceToHost, c->stream));
  }
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  auto keep = [](const double* v) {
    const double f = std::sqrt(v[4] * v[4] + v[5] * v[5] + v[6] * v[6]) + std::sqrt(v[7] * v[7] + v[8] * v[8] + v[9] * v[9]);
    return f != 0 || v[0] < 0;
  };
  for (int64_t x = 0; x < n; x++) {
    const uint64_t b = base[x >> 6] + (uint64_t)(x & 63);
    for (uint32_t e = 0; e < cnt[x]; e++) {
      const uint64_t s = b + (uint64_t)e * kWave;
      if (!(word[s] & kOwnerBit)) continue;
      const double* v = cs.data() + (size_t)kCStateLen * s;
      if (!keep(v)) continue;
      ii.push_back(ids[x]);
      jj.push_back(ids[word[s] & kIdxMask]);
      si.push_back(0);
      sj.push_back(0);
      tt.push_back(0);
      vals.insert(vals.end(), v, v + kCStateLen);
    }
  }
  if (walls)
    for (int64_t x = 0; x < n; x++)
      for (int w = 0; w < c->nwalls; w++) {
        if (!((mask[x] >> w) & 1)) continue;
        const double* v = ws.data() + (size_t)kCStateLen * ((int64_t)w * c->cap + x);
        if (!keep(v)) continue;
        ii.push_back(ids[x]);
        jj.push_back((uint64_t)w);
        si.push_back(0);
        sj.push_back(0);
        tt.push_back(driver_type(c->walls[w]));
        vals.insert(vals.end(), v, v + kCStateLen);
      }
  if (c->has_mesh && c->ml_valid && c->mstate) {  // mesh features: type 7 + kind, partner = driver index
    const MeshList& ML = c->ml[c->ml_cur];
    const int64_t st = c->cap;
    std::vector<uint32_t> mcnt(n);
    XDEM_HIP(c, hipMemcpyAsync(mcnt.data(), ML.cnt, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipStreamSynchronize(c->stream));
    uint32_t kmax = 0;
    for (auto v : mcnt) kmax = std::max(kmax, std::min(v, (uint32_t)c->mesh_cap));
    const int64_t tot = (int64_t)kmax * st;
    std::vector<uint32_t> mw(tot);
    std::vector<double> ms((size_t)tot * kCStateLen);
    if (tot) {
      XDEM_HIP(c, hipMemcpyAsync(mw.data(), ML.word, sizeof(uint32_t) * tot, hipMemcpyDeviceToHost, c->stream));
      XDEM_HIP(c, hipMemcpyAsync(ms.data(), c->mstate, sizeof(double) * ms.size(), hipMemcpyDeviceToHost, c->stream));
      XDEM_HIP(c, hipStreamSynchronize(c->stream));
    }
    for (int64_t x = 0; x < n; x++)
      for (uint32_t k = 0; k < std::min(mcnt[x], (uint32_t)c->mesh_cap); k++) {
        const int64_t j = (int64_t)k * st + x;
        const double* v = ms.data() + (size_t)kCStateLen * j;
        if (!keep(v)) continue;
        ii.push_back(ids[x]);
        jj.push_back((uint64_t)c->mesh_id);
        si.push_back(0);
        sj.push_back(mw[j] & kMeshIdxMask);
        tt.push_back((uint16_t)(7 + (mw[j] >> kMeshKindShift)));
        vals.insert(vals.end(), v, v + kCStateLen);
      }
  }
  return XDEM_OK;
}
}  // namespace xdem

using namespace xdem;

extern "C" int xdem_set_contact_record(xdem_ctx* c, int32_t on) {
  if (!c || on < 0) return XDEM_ERR_ARG;
  c->record_cs = on != 0;
  if (!c->record_cs) c->cs_valid = false;
  return XDEM_OK;
}

extern "C" int xdem_contact_state_count(xdem_ctx* c, int64_t* n) {
  if (!c || !n) return XDEM_ERR_ARG;
  *n = 0;
  std::vector<uint64_t> ii, jj;
  std::vector<uint32_t> si, sj;
  std::vector<uint16_t> tt;
  std::vector<double> v;
  const int rc = contact_state_rows(c, ii, jj, si, sj, tt, v);
  if (rc) return rc;
  *n = (int64_t)ii.size();
  return XDEM_OK;
}

extern "C" int xdem_contact_state_dump(xdem_ctx* c, int64_t cap, uint64_t* id_i, uint64_t* id_j, uint32_t* sub_i,
                                       uint32_t* sub_j, uint16_t* type, double* values, int64_t* written) {
  if (!c || !id_i || !id_j || !sub_i || !sub_j || !type || !values || cap < 0) return XDEM_ERR_ARG;
  if (written) *written = 0;
  std::vector<uint64_t> ii, jj;
  std::vector<uint32_t> si, sj;
  std::vector<uint16_t> tt;
  std::vector<double> v;
  const int rc = contact_state_rows(c, ii, jj, si, sj, tt, v);
  if (rc) return rc;
  if ((int64_t)ii.size() > cap) return set_err(c, XDEM_ERR_CAPACITY, "contact_state_dump: cap too small");
  // one block per interaction type in ascending type, device order inside each, as the reference's
  // writer walks the classifier's waves (itools/writer.hpp:35-75)
  std::vector<int64_t> order(ii.size());
  for (size_t k = 0; k < order.size(); k++) order[k] = (int64_t)k;
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return tt[a] < tt[b]; });
  for (size_t k = 0; k < order.size(); k++) {
    const int64_t r = order[k];
    id_i[k] = ii[r];
    id_j[k] = jj[r];
    sub_i[k] = si[r];
    sub_j[k] = sj[r];
    type[k] = tt[r];
    std::copy(v.begin() + r * kCStateLen, v.begin() + (r + 1) * kCStateLen, values + k * kCStateLen);
  }
  if (written) *written = (int64_t)ii.size();
  return XDEM_OK;
}
