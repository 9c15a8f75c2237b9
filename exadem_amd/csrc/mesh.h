// mesh.h — sphere against a stationary triangle-mesh driver: candidate filters and contact detection
// (the sphere is a vertex of radius R; the mesh features carry the Minkowski radius rm).
//
//   candidates  detection_sphere_driver (src/sphere/include/exaDEM/sphere/nbh_sphere.hpp:25-86): per driver
//               vertex, edge and face passing its filter with rVerlet = rcut_inc
//   filters     filter_vertex_vertex, filter_vertex_edge_core, filter_vertex_face
//               (src/shape/include/exaDEM/shape_detection_core.hpp:43-53, 83-101;
//               shape_detection_api.hpp:246-309)
//   detection   detect<7>, detect<8>, detect<9> (shape_detection_api.hpp:357-386) over
//               detection_vertex_vertex_core, detection_vertex_edge_core, detection_vertex_face_core
//               (shape_detection_core.hpp:57-81, 122-148, 397-466)
#pragma once

#include "contact_law.h"

namespace xdem {

constexpr uint32_t kMeshKindShift = 30;
constexpr uint32_t kMeshIdxMask = (1u << 30) - 1u;

__device__ inline V3 mesh_unit(V3 v) { return v / norm(v); }

// Vertex against a triangle: distance to its plane (normal oriented towards the vertex) and whether the
// vertex projects inside (crossing count of the projected point over the triangle's edges). early: the
// distance exceeds limit.
__device__ inline bool mesh_vf_core(V3 vi, const V3 (&t)[3], double limit, double& dist_out, V3& n_out,
                                    bool& early) {
  const V3 va = t[0], vb = t[1], vc = t[2];
  const V3 v = vi - va;
  V3 v1 = vb - va;
  V3 v2 = vc - va;
  v1 = mesh_unit(v1);
  V3 n = cross(v1, v2);
  n = mesh_unit(n);
  double dist = dot(n, v);
  if (dist < 0) {
    n = n * (-1);
    dist = -dist;
  }
  early = false;
  if (dist > limit) {
    early = true;
    return false;
  }
  const V3 P = vi - n * dist;
  int odd = 0;
  v2 = cross(n, v1);
  const double ori1 = dot(P, v1), ori2 = dot(P, v2);
#pragma unroll
  for (int iva = 0; iva < 3; ++iva) {
    const int ivb = iva == 2 ? 0 : iva + 1;
    const V3 A = t[iva], B = t[ivb];
    const double pa1 = dot(A, v1), pb1 = dot(B, v1), pa2 = dot(A, v2), pb2 = dot(B, v2);
    if ((pa2 < ori2 && pb2 >= ori2) || (pb2 < ori2 && pa2 >= ori2)) {
      if (pa1 + (ori2 - pa2) / (pb2 - pa2) * (pb1 - pa1) < ori1) odd = 1 - odd;
    }
  }
  dist_out = dist;
  n_out = n;
  return odd == 1;
}

__device__ inline void mesh_tri(const MeshView& m, uint32_t k, V3 (&t)[3]) {
  const uint4 f = m.f[k];
  t[0] = m.v[f.x];
  t[1] = m.v[f.y];
  t[2] = m.v[f.z];
}

// Candidate filter of feature word wd for a sphere at p of radius R (rVerlet = rv).
__device__ inline bool mesh_filter(const MeshView& m, uint32_t wd, V3 p, double R, double rv) {
  const uint32_t kind = wd >> kMeshKindShift, k = wd & kMeshIdxMask;
  if (kind == 0) {
    const double Rm = R + m.rm + rv;
    const V3 d = p - m.v[k];
    return dot(d, d) <= Rm * Rm;
  }
  if (kind == 1) {
    const uint2 e = m.e[k];
    const V3 vf = m.v[e.x], vs = m.v[e.y];
    const V3 fs = vs - vf;
    const V3 fi = p - vf;
    const double r = dot(fs, fi) / dot(fs, fs);
    if (r < -rv || r > 1.0 + rv) return false;
    const V3 d = fi - fs * r;
    const double Rmax = R + m.rm + rv;
    return dot(d, d) < Rmax * Rmax;
  }
  V3 t[3];
  mesh_tri(m, k, t);
  double dist;
  V3 n;
  bool early;
  return mesh_vf_core(p, t, R + m.rm + rv, dist, n, early);
}

// Geometry of feature word wd: kind 0 a vertex t[0], 1 an edge t[0] -> t[1], 2 a triangle t[0..2]
// (unused points zero). Returns the kind.
__device__ inline uint32_t mesh_feature(const MeshView& m, uint32_t wd, V3 (&t)[3]) {
  const uint32_t kind = wd >> kMeshKindShift, k = wd & kMeshIdxMask;
  t[1] = V3{0.0, 0.0, 0.0};
  t[2] = V3{0.0, 0.0, 0.0};
  if (kind == 0) {
    t[0] = m.v[k];
  } else if (kind == 1) {
    const uint2 e = m.e[k];
    t[0] = m.v[e.x];
    t[1] = m.v[e.y];
  } else {
    mesh_tri(m, k, t);
  }
  return kind;
}

// Contact of a mesh feature (kind, geometry t as mesh_feature gives it, Minkowski radius rm) with the
// sphere (p, R): overlap dn, normal n (from the mesh to the sphere) and contact point cp.
__device__ inline bool mesh_detect_t(uint32_t kind, const V3 (&t)[3], double rm, V3 p, double R, double& dn, V3& n,
                                     V3& cp) {
  if (kind == 0) {
    const double Rs = R + rm;
    const V3 dist = p - t[0];
    const double dist_norm = norm(dist);
    const double inv = 1.0 / dist_norm;
    dn = dist_norm - Rs;
    n = dist * inv;
    cp = p - n * (R + 0.5 * dn);
    return dn < 0;
  }
  if (kind == 1) {
    const V3 vf = t[0], vs = t[1];
    const V3 fs = vs - vf, fi = p - vf;
    const double r = dot(fs, fi) / dot(fs, fs);
    if (r <= 0 || r >= 1.0) return false;
    V3 nn = fi - fs * r;
    dn = norm(nn) - (R + rm);
    if (dn > 0) return false;
    n = mesh_unit(nn);
    cp = p - n * (R + 0.5 * dn);
    return true;
  }
  double dist;
  bool early;
  const bool in = mesh_vf_core(p, t, R + rm, dist, n, early);
  if (early) return false;
  dn = dist - (R + rm);
  cp = p - n * (R + 0.5 * dn);
  return in;
}

// Contact of feature word wd with the sphere (p, R).
__device__ inline bool mesh_detect(const MeshView& m, uint32_t wd, V3 p, double R, double& dn, V3& n, V3& cp) {
  V3 t[3];
  const uint32_t kind = mesh_feature(m, wd, t);
  return mesh_detect_t(kind, t, m.rm, p, R, dn, n, cp);
}

// ContactLawRShapeDriverFunc (contact_sphere.hpp:298-396) over sphere x's mesh candidates in list order,
// in FP64 (both force forms): the driver sits at the mesh centre, stationary; no contact resets the
// history. Adds the forces and moments to F and M. cs (null: not recorded): contact state per slot,
// kCStateLen doubles (dn, contact point, f - ft, ft, sphere centre, zeros; zeros without contact).
// Candidate jc of the compact lists l (its feature: kind and geometry tg) of a sphere at rx (radius Rx, mass
// mx, velocity vx, spin wx): the contact law against the mesh (ContactLawDriverFunc,
// contact_sphere.hpp:298-396) with its friction and rolling history at jc, and the terms it adds to the
// sphere's force and moment written to out[0..5]; out of contact the history is cleared and the terms
// are -0.0, which adds nothing to any value. cs: the recorded contact state of the candidate's list slot
// j (null: not recorded).
// ft0 / mint0: the history at jc, read by the caller ahead of the particle's record.
template <bool KR>
__device__ inline void mesh_candidate(const MeshView& m, const MeshList& l, int64_t jc, int64_t j, uint32_t kind,
                                      const V3 (&tg)[3], V3 rx, double Rx, double mx, V3 vx, V3 wx, double dt,
                                      const ContactParams& cpd, double* __restrict__ out, double* cs, V3 ft0,
                                      V3 mint0) {
  double dn;
  V3 nrm, cpos;
  if (!mesh_detect_t(kind, tg, m.rm, rx, Rx, dn, nrm, cpos)) {
    l.hx[jc] = 0.0; l.hy[jc] = 0.0; l.hz[jc] = 0.0;
    l.mhx[jc] = 0.0; l.mhy[jc] = 0.0; l.mhz[jc] = 0.0;
    if (cs) {
      double* r = cs + (int64_t)kCStateLen * j;
#pragma unroll
      for (int q = 0; q < kCStateLen; q++) r[q] = 0.0;
    }
    reinterpret_cast<double2*>(out)[0] = double2{-0.0, -0.0};
    reinterpret_cast<double2*>(out)[1] = double2{-0.0, -0.0};
    reinterpret_cast<double2*>(out)[2] = double2{-0.0, -0.0};
    return;
  }
  V3 ft = ft0;
  V3 mint{0.0, 0.0, 0.0};
  if (KR) mint = mint0;
  const V3 zero{0.0, 0.0, 0.0};
  const V3 f = hooke_force<KR>(dn, nrm, dt, cpd, mx, ft, mint, cpos, rx, vx, wx, m.center, zero, zero);
  l.hx[jc] = ft.x; l.hy[jc] = ft.y; l.hz[jc] = ft.z;
  if (KR) {
    l.mhx[jc] = mint.x; l.mhy[jc] = mint.y; l.mhz[jc] = mint.z;
  }
  if (cs) {
    const V3 fnv = f - ft;
    const double rec[kCStateLen] = {dn,   cpos.x, cpos.y, cpos.z, fnv.x, fnv.y, fnv.z, ft.x,
                                    ft.y, ft.z,   rx.x,   rx.y,   rx.z,   0.0,   0.0,   0.0};
    double* r = cs + (int64_t)kCStateLen * j;
#pragma unroll
    for (int q = 0; q < kCStateLen; q++) r[q] = rec[q];
  }
  const V3 mt = cross(cpos - rx, f) + mint;
  reinterpret_cast<double2*>(out)[0] = double2{f.x, f.y};
  reinterpret_cast<double2*>(out)[1] = double2{f.z, mt.x};
  reinterpret_cast<double2*>(out)[2] = double2{mt.y, mt.z};
}

// The mesh terms of a particle (mi = its first flat candidate and count) added to its force and moment
// in candidate order, after the walls (the same sums, in the same order, as evaluating each candidate
// here). A particle's terms are contiguous, 48 bytes per candidate.
__device__ inline void mesh_add_at(const MeshTerms& t, uint2 mi, V3& F, V3& M) {
  const uint32_t c = mi.y;
  const double2* p = reinterpret_cast<const double2*>(t.terms) + (int64_t)mi.x * 3;
  // batches of four candidates, the loads of the next batch in flight while the current one is summed
  // (in order)
  double2 r[4][3] = {}, nx[4][3] = {};
#pragma unroll
  for (int q = 0; q < 4; q++)
    if (q < (int)c) {
      r[q][0] = p[q * 3];
      r[q][1] = p[q * 3 + 1];
      r[q][2] = p[q * 3 + 2];
    }
  for (uint32_t k0 = 0; k0 < c; k0 += 4) {
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (k0 + 4 + q < c) {
        nx[q][0] = p[(k0 + 4 + q) * 3];
        nx[q][1] = p[(k0 + 4 + q) * 3 + 1];
        nx[q][2] = p[(k0 + 4 + q) * 3 + 2];
      }
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (k0 + q < c) {
        F = F + V3{r[q][0].x, r[q][0].y, r[q][1].x};
        M = M + V3{r[q][1].y, r[q][2].x, r[q][2].y};
      }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      r[q][0] = nx[q][0];
      r[q][1] = nx[q][1];
      r[q][2] = nx[q][2];
    }
  }
}
// The same sums with four loads in flight per batch and no pipelining across batches (fewer registers:
// the polyhedra gather, where the pipelined form lowers occupancy).
__device__ inline void mesh_add(const MeshTerms& t, int64_t x, V3& F, V3& M) {
  const uint2 mi = t.idx[x];
  const uint32_t c = mi.y;
  const double2* p = reinterpret_cast<const double2*>(t.terms) + (int64_t)mi.x * 3;
  for (uint32_t k0 = 0; k0 < c; k0 += 4) {
    double2 r[4][3] = {};
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (k0 + q < c) {
        r[q][0] = p[(k0 + q) * 3];
        r[q][1] = p[(k0 + q) * 3 + 1];
        r[q][2] = p[(k0 + q) * 3 + 2];
      }
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (k0 + q < c) {
        F = F + V3{r[q][0].x, r[q][0].y, r[q][1].x};
        M = M + V3{r[q][1].y, r[q][2].x, r[q][2].y};
      }
  }
}

}  // namespace xdem
