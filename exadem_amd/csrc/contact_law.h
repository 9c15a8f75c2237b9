// contact_law.h — Hooke normal force with viscous damping, Coulomb-capped tangential history and
// rolling-moment history, in FP64 (device functions shared by the sphere and polyhedron kernels).
//
//   detection_vertex_vertex_core  src/shape/include/exaDEM/shape_detection_core.hpp:57-81
//   contact_force_core<Hooke,None> src/forcefield/include/exaDEM/forcefield/contact_force.hpp:146-235
//   common kernels                 src/forcefield/include/exaDEM/forcefield/common_kernels.hpp:27-145
// The expression order follows the reference so that results are reproducible bit for bit
// against the CPU restatement when both are compiled without floating-point contraction.
#pragma once

#include "xdem_internal.h"

namespace xdem {

// rolling-moment update and its clip; mur = 0 in the reference, so any non-zero moment is cleared
// whenever the branch is well defined (|vrot_i| > 0), and kept otherwise (NaN threshold).
__device__ inline void rolling_moment(const ContactParams& p, double dt, V3 cp, V3 pos_i, V3 vrot_i, V3 vrot_j,
                                      V3 fn, V3& mom) {
  mom = mom + (p.kr * (vrot_j - vrot_i)) * dt;
  V3 branch = cp - pos_i;
  const double r = dot(branch, vrot_i) / dot(vrot_i, vrot_i);
  branch = branch - r * vrot_i;
  const double mur = 0;
  const double thr_m = fabs(mur * norm(branch) * norm(fn));
  const double m2 = dot(mom, mom);
  if (m2 > 0.0 && m2 > thr_m * thr_m) mom = mom * (thr_m / sqrt(m2));
}

// Returns f_i; updates ft (friction history) and, when KR, mom (rolling-moment history).
template <bool KR>
__device__ inline V3 hooke_force(double dn, V3 n, double dt, const ContactParams& p, double meff, V3& ft, V3& mom,
                                 V3 cp, V3 pos_i, V3 vel_i, V3 vrot_i, V3 pos_j, V3 vel_j, V3 vrot_j) {
  const double damp = p.damp_rate * 2.0 * sqrt(p.kn * meff);
  const V3 ci = vel_i - cross(cp - pos_i, vrot_i);
  const V3 cj = vel_j - cross(cp - pos_j, vrot_j);
  const V3 vel = cj - ci;
  const double vn = dot(vel, n);
  const double fne = -p.kn * dn;
  const double fnv = damp * vn;
  const double fnval = (fnv + fne > 0) ? fnv + fne : 0;
  const V3 fn = fnval * n;
  const V3 vt = vel - vn * n;
  ft = ft + p.kt * (dt * vt);
  const double thr = fabs(p.mu * (p.kn * dn));
  const double ft2 = dot(ft, ft);
  if (ft2 > 0.0 && ft2 > thr * thr) ft = ft * (thr / sqrt(ft2));
  const V3 f = fn + ft;
  if (KR) rolling_moment(p, dt, cp, pos_i, vrot_i, vrot_j, fn, mom);
  return f;
}

}  // namespace xdem
