// step.h — per-particle velocity-Verlet pieces shared by the integrator kernels (integrate.hip) and
// the force pass that integrates each particle right after its force (contact_sphere.hip): one
// definition, so both paths run the same operation sequence.
//
//   prolog  combined_compute_prolog (src/numericalscheme/combined_compute_prolog.cu:35-47,94-116):
//           r += v dt + a dt^2/2 ; v += a dt/2 ; Q <- normalize(Q + dot(Q,w) dt) ; w += alpha dt/2
//   kick    the velocity updates of combined_compute_epilog (combined_compute_epilog.cu:94-113):
//           w += alpha dt/2 ; v += a dt/2
#pragma once

#include "xdem_internal.h"

namespace xdem {

struct Q4 {
  double w, x, y, z;
};

// quat::dot (src/shape/include/exaDEM/type/quat.hpp:92-96): 0.5*(0,w)*q
__device__ inline Q4 qdot(Q4 q, V3 w) {
  return {-0.5 * (w.x * q.x + w.y * q.y + w.z * q.z), 0.5 * (w.y * q.z - w.z * q.y + w.x * q.w),
          0.5 * (w.z * q.x - w.x * q.z + w.y * q.w), 0.5 * (w.x * q.y - w.y * q.x + w.z * q.w)};
}
__device__ inline Q4 qnormalize(Q4 q) {
  const double n = sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
  return {q.w / n, q.x / n, q.y / n, q.z / n};
}

// Kick of the finished step (KICK) then the prolog of the next one, on one particle's record
// (xr = position + radius, vm = velocity + mass, w = angular velocity + homothety, qq = orientation
// stored {w, x, y, z} in {x, y, z, w}); a and al are the acceleration and angular acceleration of
// the finished force pass. kdt_2 = dt_kick / 2.
template <bool KICK>
__device__ inline void kick_prolog(d4& xr, d4& vm, d4& w, d4& qq, double ax, double ay, double az, double lx,
                                   double ly, double lz, double dt, double kdt_2) {
  const double dt_2 = dt * 0.5;
  const double dt2_2 = dt * dt * 0.5;
  if (KICK) {
    // epilog of the previous step: angular velocity, then push_f_v
    w.x = w.x + lx * kdt_2;
    w.y = w.y + ly * kdt_2;
    w.z = w.z + lz * kdt_2;
    vm.x = vm.x + ax * kdt_2;
    vm.y = vm.y + ay * kdt_2;
    vm.z = vm.z + az * kdt_2;
  }
  // push_f_v_r, push_f_v
  xr.x = xr.x + (vm.x * dt + ax * dt2_2);
  xr.y = xr.y + (vm.y * dt + ay * dt2_2);
  xr.z = xr.z + (vm.z * dt + az * dt2_2);
  vm.x = vm.x + ax * dt_2;
  vm.y = vm.y + ay * dt_2;
  vm.z = vm.z + az * dt_2;
  // push_to_quaternion
  Q4 q{qq.x, qq.y, qq.z, qq.w};
  const Q4 d = qdot(q, V3{w.x, w.y, w.z});
  q = {q.w + d.w * dt, q.x + d.x * dt, q.y + d.y * dt, q.z + d.z * dt};
  q = qnormalize(q);
  w.x = w.x + dt_2 * lx;
  w.y = w.y + dt_2 * ly;
  w.z = w.z + dt_2 * lz;
  qq = d4{q.w, q.x, q.y, q.z};
}

}  // namespace xdem
