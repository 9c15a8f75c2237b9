// poly.hip — convex sphero-polyhedra: world vertices, feature-interaction lists with history carry-over,
// and the atomic-free contact pass.
//
//   vertices   shape::get_vertex p + Q*(h v) (src/shape/include/exaDEM/shape.hpp:244-262),
//              compute_vertices.hpp:50-67; vertex_displ_over (src/polyhedron/vertex_displ_over.cu:74-133)
//   lists      nbh_polyhedron (src/polyhedron/nbh_polyhedron.cpp:223-376): owner = lower id, OBB test with both
//              boxes enlarged by rVerlet/2 (then again), per-vertex OBB gate, VV, VE, VF, then EE, then the
//              swap pass VE/VF from the partner's vertices; vertex-wall candidates
//              (nbh_polyhedron_driver.hpp:216-236)
//   filters    src/shape/include/exaDEM/shape_detection_core.hpp:43-53, 83-101, 245-309, 468-504
//   detection  shape_detection_core.hpp:57-81, 122-148, 397-466, 520-566
//   contact    src/polyhedron/include/exaDEM/polyhedron/contact.hpp:124-211 (symmetric), 219-297 (walls)
//
// Contact pass without atomics: one thread per feature evaluates it and writes its terms of the owner
// force, owner moment and partner moment (the partner's force term is the exact negation). A second
// kernel walks, per particle and in stream order, its entries (each slot holds its pair's feature range,
// the owner slot's for a partner entry), forms each pair's subtotal over the features in list order,
// adds it, then the vertex-wall contacts, then finishes force_to_accel and the angular acceleration.
#include <algorithm>
#include <unordered_map>

#include "contact_law.h"
#include "mesh.h"

namespace xdem {

constexpr uint32_t kFeatTypeMask = 3u;
constexpr uint32_t kFeatSwap = 4u;
constexpr uint32_t kFeatKeyMask = (1u << 11) - 1u;  // type, swap, owner sub, partner sub
constexpr uint32_t kFeatActive = 1u << 30;

__device__ inline uint32_t feat_word(int type, int swap, int osub, int psub) {
  return (uint32_t)type | (swap ? kFeatSwap : 0u) | ((uint32_t)osub << 3) | ((uint32_t)psub << 7);
}

struct OBBd {
  V3 c, e1, e2, e3, ext;
};
__device__ inline void obb_enlarge(OBBd& b, double more) {
  b.ext.x += more;
  b.ext.y += more;
  b.ext.z += more;
}
__device__ inline double vcomp(const V3& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// OBB::intersect (src/shape/include/exaDEM/type/OBB.hpp:90-224), tolerance FLT_EPSILON
__device__ inline bool obb_intersect(const OBBd& a, const OBBd& b) {
  const double tol = 1.1920928955078125e-07;
  const V3 ea[3] = {a.e1, a.e2, a.e3}, eb[3] = {b.e1, b.e2, b.e3};
  const double ae[3] = {a.ext.x, a.ext.y, a.ext.z}, be[3] = {b.ext.x, b.ext.y, b.ext.z};
  const V3 tt = a.c - b.c;
  const double t[3] = {dot(tt, a.e1), dot(tt, a.e2), dot(tt, a.e3)};
  double R[3][3], A[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      R[i][j] = dot(ea[i], eb[j]);
      A[i][j] = fabs(R[i][j]) + tol;
    }
  double ra, rb;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    ra = ae[i];
    rb = be[0] * A[i][0] + be[1] * A[i][1] + be[2] * A[i][2];
    if (fabs(t[i]) > ra + rb) return false;
  }
#pragma unroll
  for (int j = 0; j < 3; j++) {
    ra = ae[0] * A[0][j] + ae[1] * A[1][j] + ae[2] * A[2][j];
    rb = be[j];
    if (fabs(t[0] * R[0][j] + t[1] * R[1][j] + t[2] * R[2][j]) > ra + rb) return false;
  }
  ra = ae[1] * A[2][0] + ae[2] * A[1][0];
  rb = be[1] * A[0][2] + be[2] * A[0][1];
  if (fabs(t[2] * R[1][0] - t[1] * R[2][0]) > ra + rb) return false;
  ra = ae[1] * A[2][1] + ae[2] * A[1][1];
  rb = be[0] * A[0][2] + be[2] * A[0][0];
  if (fabs(t[2] * R[1][1] - t[1] * R[2][1]) > ra + rb) return false;
  ra = ae[1] * A[2][2] + ae[2] * A[1][2];
  rb = be[0] * A[0][1] + be[1] * A[0][0];
  if (fabs(t[2] * R[1][2] - t[1] * R[2][2]) > ra + rb) return false;
  ra = ae[0] * A[2][0] + ae[2] * A[0][0];
  rb = be[1] * A[1][2] + be[2] * A[1][1];
  if (fabs(t[0] * R[2][0] - t[2] * R[0][0]) > ra + rb) return false;
  ra = ae[0] * A[2][1] + ae[2] * A[0][1];
  rb = be[0] * A[1][2] + be[2] * A[1][0];
  if (fabs(t[0] * R[2][1] - t[2] * R[0][1]) > ra + rb) return false;
  ra = ae[0] * A[2][2] + ae[2] * A[0][2];
  rb = be[0] * A[1][1] + be[1] * A[1][0];
  if (fabs(t[0] * R[2][2] - t[2] * R[0][2]) > ra + rb) return false;
  ra = ae[0] * A[1][0] + ae[1] * A[0][0];
  rb = be[1] * A[2][2] + be[2] * A[2][1];
  if (fabs(t[1] * R[0][0] - t[0] * R[1][0]) > ra + rb) return false;
  ra = ae[0] * A[1][1] + ae[1] * A[0][1];
  rb = be[0] * A[2][2] + be[2] * A[2][0];
  if (fabs(t[1] * R[0][1] - t[0] * R[1][1]) > ra + rb) return false;
  ra = ae[0] * A[1][2] + ae[1] * A[0][2];
  rb = be[0] * A[2][1] + be[1] * A[2][0];
  if (fabs(t[1] * R[0][2] - t[0] * R[1][2]) > ra + rb) return false;
  return true;
}

// compute_obb (src/shape/include/exaDEM/basic_types.hpp:37-49)
__device__ inline OBBd compute_obb(const xdem_shape& sh, V3 p, const d4& q, double h) {
  OBBd b;
  b.c = {0, 0, 0};
  b.e1 = qrot4(q, V3{1, 0, 0});
  b.e2 = qrot4(q, V3{0, 1, 0});
  b.e3 = qrot4(q, V3{0, 0, 1});
  b.c = qrot4(q, b.c);
  b.ext = {sh.obb_extent[0], sh.obb_extent[1], sh.obb_extent[2]};
  if (h != 1.0) {
    b.ext = b.ext * h;
    b.c = b.c * h;
  }
  b.c = b.c + p;
  return b;
}

__device__ inline V3 vnormalize(V3 v) { return v / norm(v); }

// Vertices of one particle read in place from the world-vertex array (plus a periodic image shift),
// so that no per-lane vertex array is built (dynamically indexed arrays live in scratch memory).
struct Verts {
  const d4* __restrict__ p;
  V3 sh;
  __device__ V3 operator[](int k) const {
    const d4 a = p[k];
    return {a.x + sh.x, a.y + sh.y, a.z + sh.z};
  }
};

__device__ inline Verts verts_of(const d4* __restrict__ vtx, int64_t p, V3 sh) {
  return Verts{vtx + p * XDEM_MAX_SHAPE_VERTICES, sh};
}

__device__ inline bool filter_vv(double rv, V3 vi, double ri, V3 vj, double rj) {
  const double R = ri + rj + rv;
  const V3 d = vi - vj;
  return dot(d, d) <= R * R;
}
__device__ inline bool filter_ve(double rv, V3 vi, double ri, V3 vf, V3 vs, double rj) {
  const V3 fs = vs - vf;
  const V3 fi = vi - vf;
  const double r = dot(fs, fi) / dot(fs, fs);
  if (r < -rv || r > 1.0 + rv) return false;
  const V3 d = fi - fs * r;
  const double Rmax = ri + rj + rv;
  return dot(d, d) < Rmax * Rmax;
}
// vertex-face core shared by the filter (limit = ri + rj + rVerlet) and the detection (limit = ri + rj)
__device__ inline bool vf_core(V3 vi, const Verts& vj, const xdem_shape& shj, int face, double limit, double& dist_out,
                               V3& n_out, bool& early) {
  const int nf = shj.face_size[face];
  const int* f = shj.faces[face];
  const V3 va = vj[f[0]];
  const V3 vb = vj[f[1]];
  const V3 vc = vj[f[nf - 1]];
  const V3 v = vi - va;
  V3 v1 = vb - va;
  V3 v2 = vc - va;
  v1 = vnormalize(v1);
  V3 n = cross(v1, v2);
  n = vnormalize(n);
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
  for (int iva = 0; iva < nf; ++iva) {
    int ivb = iva + 1;
    if (ivb == nf) ivb = 0;
    const V3 A = vj[f[iva]], B = vj[f[ivb]];
    const double pa1 = dot(A, v1), pb1 = dot(B, v1), pa2 = dot(A, v2), pb2 = dot(B, v2);
    if ((pa2 < ori2 && pb2 >= ori2) || (pb2 < ori2 && pa2 >= ori2)) {
      if (pa1 + (ori2 - pa2) / (pb2 - pa2) * (pb1 - pa1) < ori1) odd = 1 - odd;
    }
  }
  dist_out = dist;
  n_out = n;
  return odd == 1;
}
// The vertex-independent part of vf_core for one face (the same operations): its first vertex, the unit
// edge v1, the unit normal n0, v2 = n0 x v1 and the projections of the face's vertices on v1 and v2.
struct FaceFrame {
  V3 va, v1, n0, v2;
  double q1[XDEM_MAX_FACE_VERTICES], q2[XDEM_MAX_FACE_VERTICES];
  int nf;
};
__device__ inline FaceFrame face_frame(const Verts& vj, const xdem_shape& shj, int face) {
  FaceFrame fr;
  const int nf = shj.face_size[face];
  const int* f = shj.faces[face];
  const V3 va = vj[f[0]];
  const V3 vb = vj[f[1]];
  const V3 vc = vj[f[nf - 1]];
  V3 v1 = vb - va;
  const V3 v2 = vc - va;
  v1 = vnormalize(v1);
  V3 n = cross(v1, v2);
  n = vnormalize(n);
  fr.va = va;
  fr.v1 = v1;
  fr.n0 = n;
  fr.v2 = cross(n, v1);
  fr.nf = nf;
#pragma unroll
  for (int k = 0; k < XDEM_MAX_FACE_VERTICES; k++) {
    fr.q1[k] = 0.0;
    fr.q2[k] = 0.0;
    if (k < nf) {
      const V3 A = vj[f[k]];
      fr.q1[k] = dot(A, v1);
      fr.q2[k] = dot(A, fr.v2);
    }
  }
  return fr;
}
// vf_core's verdict from a face frame. Where the normal flips (n = n0 (-1)), v2 = n x v1 and the
// projections on it are the frame's negated, which is exact (negating a factor negates each product, and
// rounding to nearest is symmetric): every comparison sees the values vf_core computes.
__device__ inline bool vf_frame(V3 vi, const FaceFrame& fr, double limit) {
  const V3 v = vi - fr.va;
  V3 n = fr.n0;
  double dist = dot(n, v);
  bool flip = false;
  if (dist < 0) {
    n = n * (-1);
    dist = -dist;
    flip = true;
  }
  if (dist > limit) return false;
  const V3 P = vi - n * dist;
  const double ori1 = dot(P, fr.v1), o2 = dot(P, fr.v2), ori2 = flip ? -o2 : o2;
  int odd = 0;
#pragma unroll
  for (int iva = 0; iva < XDEM_MAX_FACE_VERTICES; ++iva) {
    if (iva < fr.nf) {
      const bool last = iva + 1 == fr.nf;
      const double pa1 = fr.q1[iva], pb1 = last ? fr.q1[0] : fr.q1[(iva + 1) % XDEM_MAX_FACE_VERTICES];
      const double qa2 = fr.q2[iva], qb2 = last ? fr.q2[0] : fr.q2[(iva + 1) % XDEM_MAX_FACE_VERTICES];
      const double pa2 = flip ? -qa2 : qa2, pb2 = flip ? -qb2 : qb2;
      if ((pa2 < ori2 && pb2 >= ori2) || (pb2 < ori2 && pa2 >= ori2)) {
        if (pa1 + (ori2 - pa2) / (pb2 - pa2) * (pb1 - pa1) < ori1) odd = 1 - odd;
      }
    }
  }
  return odd == 1;
}

__device__ inline bool filter_ee(double rv, V3 vfi, V3 vsi, double ri, V3 vfj, V3 vsj, double rj) {
  const double R = ri + rj + rv;
  const V3 Ei = vsi - vfi, Ej = vsj - vfj, v = vfi - vfj;
  const double c = dot(Ei, Ei), d = dot(Ej, Ej), e = dot(Ei, Ej);
  double f = (c * d) - (e * e);
  if (fabs(f) > 1.0e-12) {
    f = 1.0 / f;
    const double a = dot(Ei, v), b = dot(Ej, v);
    const double s = (e * b - a * d) * f, t = (c * b - e * a) * f;
    if (s <= 0.0 || s >= 1.0 || t <= 0.0 || t >= 1.0) return false;
    const V3 pi = vfi + Ei * s, pj = vfj + Ej * t;
    const V3 n = pi - pj;
    return dot(n, n) <= R * R;
  }
  return false;
}

struct Det {
  bool contact;
  double dn;
  V3 n, cp;
};
__device__ inline Det det_none() { return {false, 0, {0, 0, 0}, {0, 0, 0}}; }
__device__ inline Det detect_vv(V3 pi, double ri, V3 pj, double rj) {
  const double R = ri + rj;
  const V3 dist = pi - pj;
  const double dnorm = norm(dist);
  const double inv = 1.0 / dnorm;
  const double dn = dnorm - R;
  const V3 n = dist * inv;
  return {dn < 0, dn, n, pi - n * (ri + 0.5 * dn)};
}
__device__ inline Det detect_ve(V3 vi, double ri, V3 vf, V3 vs, double rj) {
  const V3 fs = vs - vf, fi = vi - vf;
  const double r = dot(fs, fi) / dot(fs, fs);
  if (r <= 0 || r >= 1.0) return det_none();
  V3 n = fi - fs * r;
  const double dn = norm(n) - (ri + rj);
  if (dn > 0) return det_none();
  n = vnormalize(n);
  return {true, dn, n, vi - n * (ri + 0.5 * dn)};
}
__device__ inline Det detect_vf(V3 vi, double ri, const Verts& vj, const xdem_shape& shj, int face, double rj) {
  double dist;
  V3 n;
  bool early;
  const bool in = vf_core(vi, vj, shj, face, ri + rj, dist, n, early);
  if (early) return det_none();
  const double dn = dist - (ri + rj);
  return {in, dn, n, vi - n * (ri + 0.5 * dn)};
}
__device__ inline Det detect_ee(V3 vfi, V3 vsi, double ri, V3 vfj, V3 vsj, double rj) {
  const double R = ri + rj;
  const V3 Ei = vsi - vfi, Ej = vsj - vfj, v = vfi - vfj;
  const double c = dot(Ei, Ei), d = dot(Ej, Ej), e = dot(Ei, Ej);
  double f = (c * d) - (e * e);
  if (fabs(f) > 1.0e-12) {
    f = 1.0 / f;
    const double a = dot(Ei, v), b = dot(Ej, v);
    const double s = (e * b - a * d) * f, t = (c * b - e * a) * f;
    if (s <= 0.0 || s >= 1.0 || t <= 0.0 || t >= 1.0) return det_none();
    const V3 pi = vfi + Ei * s, pj = vfj + Ej * t;
    V3 n = pi - pj;
    const double dn = norm(n) - R;
    if (dn > 0) return det_none();
    n = vnormalize(n);
    return {dn <= 0, dn, n, pi - n * (ri + 0.5 * dn)};
  }
  return det_none();
}

// ------------------------------------------------------------------------------------------ vertices
__global__ __launch_bounds__(256) void k_poly_vertices(int64_t n, PState s, const xdem_shape* __restrict__ shapes,
                                                       d4* __restrict__ vtx, int displ, double thr2,
                                                       int32_t* __restrict__ flags, int guarded) {
  if (guarded && flags[1]) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Core c = s.core[i];
  const xdem_shape& sh = shapes[s.type[i]];
  const V3 p{c.xr.x, c.xr.y, c.xr.z};
  const double h = c.w.w;
  bool moved = false;
  const V3 pb{s.bx[i], s.by[i], s.bz[i]};
  const d4 qb = s.bq[i];
  for (int k = 0; k < sh.nv; k++) {
    const V3 v = get_vertex(sh, k, p, h, c.q);
    vtx[i * XDEM_MAX_SHAPE_VERTICES + k] = {v.x, v.y, v.z, 0.0};
    if (displ) {
      const V3 a = get_vertex(sh, k, pb, h, qb);
      const V3 d = v - a;
      if (dot(d, d) >= thr2) moved = true;
    }
  }
  if (moved) flags[0] = 1;
}

int poly_vtx_alloc(xdem_ctx* c) {
  if (c->cap_vtx < c->cap) {
    if (c->vtx) (void)hipFree(c->vtx);
    c->vtx = nullptr;
    XDEM_HIP(c, hipMalloc((void**)&c->vtx, sizeof(d4) * XDEM_MAX_SHAPE_VERTICES * (size_t)c->cap));
    c->cap_vtx = c->cap;
  }
  return XDEM_OK;
}

int poly_compute_vertices(xdem_ctx* c, bool displ, double thr2, bool guarded) {
  if (c->n == 0) return XDEM_OK;
  int rc = poly_vtx_alloc(c);
  if (rc) return rc;
  hipLaunchKernelGGL(k_poly_vertices, dim3(grid_blocks(c->n, 256)), dim3(256), 0, c->stream, c->n, live(c),
                     c->d_shapes, c->vtx, displ ? 1 : 0, thr2, c->d_flags, guarded ? 1 : 0);
  XDEM_HIP(c, hipGetLastError());
  return XDEM_OK;
}

// ------------------------------------------------------------------------------------------ feature lists
struct PolyNbhArgs {
  int64_t n;
  PState s;
  Grid g;
  const xdem_shape* shapes;
  const d4* vtx;
  NbrList nl;
  double rv;
};

// Enumerates the feature interactions of the pair (x owner, y partner seen through image `code`) in the
// order of nbh_polyhedron; fn(type, swap, owner_sub, partner_sub).
template <typename F>
__device__ inline void for_each_feature(const PolyNbhArgs& a, int64_t x, uint32_t y, int code, F&& fn) {
  const Core cx = a.s.core[x], cy = a.s.core[y];
  const xdem_shape& sa = a.shapes[a.s.type[x]];
  const xdem_shape& sb = a.shapes[a.s.type[y]];
  const double ha = cx.w.w, hb = cy.w.w;
  const double ra = sa.minkowski * ha, rb = sb.minkowski * hb;
  const V3 sh = shift_vec(a.g, code);
  OBBd obb_i = compute_obb(sa, V3{cx.xr.x, cx.xr.y, cx.xr.z}, cx.q, ha);
  OBBd obb_j = compute_obb(sb, V3{cy.xr.x, cy.xr.y, cy.xr.z} + sh, cy.q, hb);
  obb_enlarge(obb_i, 0.5 * a.rv);
  obb_enlarge(obb_j, 0.5 * a.rv);
  if (!obb_intersect(obb_i, obb_j)) return;
  obb_enlarge(obb_i, 0.5 * a.rv);
  obb_enlarge(obb_j, 0.5 * a.rv);
  const Verts va = verts_of(a.vtx, x, V3{0, 0, 0});
  const Verts vb = verts_of(a.vtx, y, sh);
  const double rv = a.rv;
  // the per-vertex box gates first (each vertex's box against the other particle's enlarged box), so
  // that the two boxes are dead during the feature loops
  uint32_t gate_a = 0, gate_b = 0;
  for (int i = 0; i < sa.nv; i++) {
    OBBd ov;
    ov.c = va[i];
    ov.e1 = {1, 0, 0};
    ov.e2 = {0, 1, 0};
    ov.e3 = {0, 0, 1};
    ov.ext = {0, 0, 0};
    obb_enlarge(ov, ra);
    if (obb_intersect(obb_j, ov)) gate_a |= 1u << i;
  }
  for (int j = 0; j < sb.nv; j++) {
    OBBd ov;
    ov.c = vb[j];
    ov.e1 = {1, 0, 0};
    ov.e2 = {0, 1, 0};
    ov.e3 = {0, 0, 1};
    ov.ext = {0, 0, 0};
    obb_enlarge(ov, rb);
    if (obb_intersect(obb_i, ov)) gate_b |= 1u << j;
  }
  // vertex-face filters face by face (each face's frame once, its gated vertices against it), as bits
  // 8 (vertex) + face read back in the reference's loop order below
  uint64_t hit_a = 0, hit_b = 0;
  if (gate_a) {
    const double lim = ra + rb + rv;
    for (int j = 0; j < sb.nf; j++) {
      const FaceFrame fr = face_frame(vb, sb, j);
      for (int i = 0; i < sa.nv; i++)
        if (((gate_a >> i) & 1u) && vf_frame(va[i], fr, lim)) hit_a |= 1ull << (8 * i + j);
    }
  }
  if (gate_b) {
    const double lim = rb + ra + rv;
    for (int i = 0; i < sa.nf; i++) {
      const FaceFrame fr = face_frame(va, sa, i);
      for (int j = 0; j < sb.nv; j++)
        if (((gate_b >> j) & 1u) && vf_frame(vb[j], fr, lim)) hit_b |= 1ull << (8 * j + i);
    }
  }
  for (int i = 0; i < sa.nv; i++) {
    if (!((gate_a >> i) & 1u)) continue;
    for (int j = 0; j < sb.nv; j++)
      if (filter_vv(rv, va[i], ra, vb[j], rb)) fn(0, 0, i, j);
    for (int j = 0; j < sb.ne; j++)
      if (filter_ve(rv, va[i], ra, vb[sb.edges[j][0]], vb[sb.edges[j][1]], rb)) fn(1, 0, i, j);
    for (int j = 0; j < sb.nf; j++)
      if ((hit_a >> (8 * i + j)) & 1ull) fn(2, 0, i, j);
  }
  for (int i = 0; i < sa.ne; i++)
    for (int j = 0; j < sb.ne; j++)
      if (filter_ee(rv, va[sa.edges[i][0]], va[sa.edges[i][1]], ra, vb[sb.edges[j][0]], vb[sb.edges[j][1]],
                    rb))
        fn(3, 0, i, j);
  for (int j = 0; j < sb.nv; j++) {
    if (!((gate_b >> j) & 1u)) continue;
    for (int i = 0; i < sa.ne; i++)
      if (filter_ve(rv, vb[j], rb, va[sa.edges[i][0]], va[sa.edges[i][1]], ra)) fn(1, 1, i, j);
    for (int i = 0; i < sa.nf; i++)
      if ((hit_b >> (8 * j + i)) & 1ull) fn(2, 1, i, j);
  }
}

// The owner slots of the list (one per pair, the lower id's entry; both entries of a pair seen through a
// periodic image), compacted so that the feature enumeration runs one owner pair per thread, without the
// partner entries' idle lanes. A block takes 16 x 256 slots and places them with one atomic (one atomic
// per wavefront on the one counter serialised to 0.38 ms at 200k polyhedra); their order within the
// compacted list places nothing that depends on it.
constexpr int kOwnChunk = 16;
__global__ __launch_bounds__(256) void k_owner_slots(NbrList nl, uint32_t* __restrict__ oslot,
                                                     uint32_t* __restrict__ count) {
  __shared__ uint32_t wtot[4];
  __shared__ uint32_t bbase;
  const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;
  const int64_t c0 = (int64_t)blockIdx.x * 256 * kOwnChunk + threadIdx.x;
  const unsigned long long below = (1ull << lane) - 1ull;
  uint32_t ownbits = 0, wcount = 0;
#pragma unroll
  for (int k = 0; k < kOwnChunk; k++) {
    const int64_t sl = c0 + (int64_t)k * 256;
    const bool own = sl < nl.slots && nl.sx[sl] != 0xffffffffu && (nl.word[sl] & kOwnerBit);
    ownbits |= own ? (1u << k) : 0u;
    wcount += (uint32_t)__popcll(__ballot(own));
  }
  if (lane == 0) wtot[wv] = wcount;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t tot = wtot[0] + wtot[1] + wtot[2] + wtot[3];
    bbase = tot ? atomicAdd(count, tot) : 0u;
  }
  __syncthreads();
  uint32_t pos = bbase;
  for (int w = 0; w < wv; w++) pos += wtot[w];
#pragma unroll
  for (int k = 0; k < kOwnChunk; k++) {
    const bool own = (ownbits >> k) & 1u;
    const unsigned long long bal = __ballot(own);
    if (own) oslot[pos + (uint32_t)__popcll(bal & below)] = (uint32_t)(c0 + (int64_t)k * 256);
    pos += (uint32_t)__popcll(bal);
  }
}

// One thread per owner slot: the pair's features enumerated once, the count and up to kFeatPerSlot
// feature words kept in the slot's region of tmp.
constexpr int kFeatPerSlot = 24;
// Features past kFeatPerSlot go to a chunk of kPoolChunk words claimed from a shared pool (counter
// nown[2], capacity pool_cap) at the first one; fpool[sl] = the chunk, or ~0u when the pool is exhausted
// or the pair has more than kFeatPerSlot + kPoolChunk features (then k_feat_overflow enumerates it again).
constexpr int kEnumBlock = 128;
constexpr uint32_t kPoolChunk = 232;
__global__ __launch_bounds__(kEnumBlock) void k_feat_enum(PolyNbhArgs a, const uint32_t* __restrict__ oslot,
                                                          uint32_t* __restrict__ nown, uint32_t* __restrict__ fcount,
                                                          uint32_t* __restrict__ tmp, uint32_t* __restrict__ pool,
                                                          uint32_t pool_cap, uint32_t* __restrict__ fpool) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nown[0]) return;
  const uint32_t sl = oslot[t];
  const uint32_t x = a.nl.sx[sl], wd = a.nl.word[sl];
  uint32_t nf = 0, pb = ~0u;
  uint32_t* out = tmp + (int64_t)sl * kFeatPerSlot;
  for_each_feature(a, x, wd & kIdxMask, (int)((wd >> kShiftShift) & 31u), [&](int ty, int sw, int si, int sj) {
    const uint32_t w = sw ? feat_word(ty, 1, si, sj) : feat_word(ty, 0, si, sj);
    if (nf < (uint32_t)kFeatPerSlot) {
      out[nf] = w;
    } else {
      if (nf == (uint32_t)kFeatPerSlot) {
        const uint32_t b = atomicAdd(nown + 2, kPoolChunk);
        pb = b <= pool_cap - kPoolChunk ? b : ~0u;
      }
      if (pb != ~0u && nf < (uint32_t)kFeatPerSlot + kPoolChunk) pool[pb + (nf - kFeatPerSlot)] = w;
    }
    nf++;
  });
  fcount[sl] = nf;
  if (nf > (uint32_t)kFeatPerSlot) fpool[sl] = nf <= (uint32_t)kFeatPerSlot + kPoolChunk ? pb : ~0u;
}

// One thread per owner slot: its feature words copied to their place in the list (slot order, from the
// scan of the counts), the first kFeatPerSlot from tmp and the rest from its pool chunk; a pair without a
// chunk goes to the overflow list (count in nown[1]) for k_feat_overflow.
__global__ __launch_bounds__(256) void k_feat_place(NbrList nl, const uint32_t* __restrict__ oslot,
                                                    uint32_t* __restrict__ nown, const uint32_t* __restrict__ fcount,
                                                    const uint32_t* __restrict__ tmp, const uint32_t* __restrict__ pool,
                                                    const uint32_t* __restrict__ fpool,
                                                    const uint64_t* __restrict__ start, uint32_t* __restrict__ feat,
                                                    uint32_t* __restrict__ fslot, uint32_t* __restrict__ fown,
                                                    uint32_t* __restrict__ ovf) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nown[0]) return;
  const uint32_t sl = oslot[t];
  const uint32_t nf = fcount[sl];
  if (nf == 0) return;
  const uint32_t pb = nf > (uint32_t)kFeatPerSlot ? fpool[sl] : 0u;
  if (pb == ~0u) {
    ovf[atomicAdd(nown + 1, 1u)] = sl;
    return;
  }
  const uint32_t x = nl.sx[sl];
  const uint64_t o = start[sl];
  for (uint32_t q = 0; q < nf; q++) {
    feat[o + q] = q < (uint32_t)kFeatPerSlot ? tmp[(int64_t)sl * kFeatPerSlot + q] : pool[pb + (q - kFeatPerSlot)];
    fslot[o + q] = sl;
    fown[o + q] = x;
  }
}

// The overflowing pairs (rare) enumerated again, straight into place; a fixed grid walks the list.
constexpr int kOvfBlocks = 1024;
__global__ __launch_bounds__(kEnumBlock) void k_feat_overflow(PolyNbhArgs a, const uint32_t* __restrict__ nown,
                                                              const uint32_t* __restrict__ ovf,
                                                              const uint64_t* __restrict__ start,
                                                              uint32_t* __restrict__ feat, uint32_t* __restrict__ fslot,
                                                              uint32_t* __restrict__ fown) {
  const uint32_t novf = nown[1];
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (int64_t)novf;
       t += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t sl = ovf[t];
    const uint32_t x = a.nl.sx[sl], wd = a.nl.word[sl];
    uint64_t o = start[sl];
    for_each_feature(a, x, wd & kIdxMask, (int)((wd >> kShiftShift) & 31u), [&](int ty, int sw, int si, int sj) {
      feat[o] = sw ? feat_word(ty, 1, si, sj) : feat_word(ty, 0, si, sj);
      fslot[o] = sl;
      fown[o] = x;
      o++;
    });
  }
}

// Feature range of every list slot: an owner entry's own range; a partner entry gets the range of the
// owner's entry pointing back with the mirrored image code, or x = ~0u when there is none. The force
// gather then reads one 8-byte record per entry instead of following slot -> owner slot -> range.
__global__ __launch_bounds__(256) void k_feat_range(NbrList nl, const uint64_t* __restrict__ start,
                                                    uint2* __restrict__ frange) {
  const int64_t sl = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sl >= nl.slots) return;
  const uint32_t x = nl.sx[sl];
  if (x == 0xffffffffu) return;
  const uint32_t wd = nl.word[sl];
  if (wd & kOwnerBit) {
    frange[sl] = uint2{(uint32_t)start[sl], (uint32_t)(start[sl + 1] - start[sl])};
    return;
  }
  const uint32_t y = wd & kIdxMask;
  const uint32_t code = (wd >> kShiftShift) & 31u;
  const uint32_t want = 26u - code;
  const uint32_t cy = nl.cnt[y];
  const uint64_t by = nl.base[y >> 6] + (uint64_t)(y & 63);
  uint32_t found = 0xffffffffu;
  for (uint32_t k2 = 0; k2 < cy; k2++) {
    const uint64_t s2 = by + (uint64_t)k2 * kWave;
    const uint32_t w2 = nl.word[s2];
    if ((w2 & kOwnerBit) && (w2 & kIdxMask) == x && ((w2 >> kShiftShift) & 31u) == want) {
      found = (uint32_t)s2;
      break;
    }
  }
  frange[sl] = found == 0xffffffffu ? uint2{0xffffffffu, 0u}
                                    : uint2{(uint32_t)start[found], (uint32_t)(start[found + 1] - start[found])};
}

struct FeatCarryArgs {
  int64_t n;
  NbrList nl, ol;
  PolyList pl, opl;
  const uint32_t* perm;
  int have_old;
};

// history of each new feature: the old owner pair (old owner, old partner) and the same feature key
__global__ __launch_bounds__(256) void k_feat_carry(FeatCarryArgs a) {
  const int64_t sl = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sl >= a.nl.slots) return;
  const uint32_t x = a.nl.sx[sl];
  if (x == 0xffffffffu) return;
  {
    const uint32_t wd = a.nl.word[sl];
    if (!(wd & kOwnerBit)) return;
    const uint64_t f0 = a.pl.start[sl], f1 = a.pl.start[sl + 1];
    uint64_t o0 = 0, o1 = 0;
    if (a.have_old) {
      const uint32_t ox = a.perm[x], oy = a.perm[wd & kIdxMask];
      const uint32_t ocnt = a.ol.cnt[ox];
      const uint64_t ob = a.ol.base[ox >> 6] + (uint64_t)(ox & 63);
      for (uint32_t k2 = 0; k2 < ocnt; k2++) {
        const uint64_t s2 = ob + (uint64_t)k2 * kWave;
        const uint32_t w2 = a.ol.word[s2];
        if ((w2 & kOwnerBit) && (w2 & kIdxMask) == oy) {
          o0 = a.opl.start[s2];
          o1 = a.opl.start[s2 + 1];
          break;
        }
      }
    }
    for (uint64_t f = f0; f < f1; f++) {
      const uint32_t key = a.pl.feat[f] & kFeatKeyMask;
      double hx = 0, hy = 0, hz = 0, mx = 0, my = 0, mz = 0;
      bool act = false;
      for (uint64_t g = o0; g < o1; g++) {
        const uint32_t w2 = a.opl.feat[g];
        if ((w2 & kFeatKeyMask) == key) {
          if (w2 & kFeatActive) {
            hx = a.opl.hx[g]; hy = a.opl.hy[g]; hz = a.opl.hz[g];
            mx = a.opl.mhx[g]; my = a.opl.mhy[g]; mz = a.opl.mhz[g];
            act = true;
          }
          break;
        }
      }
      a.pl.hx[f] = hx; a.pl.hy[f] = hy; a.pl.hz[f] = hz;
      a.pl.mhx[f] = mx; a.pl.mhy[f] = my; a.pl.mhz[f] = mz;
      a.pl.feat[f] = key | (act ? kFeatActive : 0u);
    }
  }
}

// vertex-driver candidates (the driver's filter with rVerlet + Minkowski radius) and their history
__global__ __launch_bounds__(256) void k_poly_wall(int64_t n, PState s, const xdem_shape* __restrict__ shapes,
                                                   const d4* __restrict__ vtx, int nw, const Surface* __restrict__ walls,
                                                   double rv, PolyWall nwl, PolyWall owl, const uint32_t* __restrict__ perm,
                                                   int have_old, int64_t stride) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const xdem_shape& sh = shapes[s.type[x]];
  const double r = sh.minkowski * s.core[x].w.w;
  const uint32_t o = have_old ? perm[x] : 0u;
  const uint64_t omask = have_old ? owl.mask[o] : 0ull;
  uint64_t mask = 0;
  for (int w = 0; w < nw; w++)
    for (int k = 0; k < sh.nv; k++) {
      const d4 v4 = vtx[x * XDEM_MAX_SHAPE_VERTICES + k];
      const V3 p{v4.x, v4.y, v4.z};
      const Surface& sf = walls[w];
      const int bit = w * XDEM_MAX_SHAPE_VERTICES + k;
      const int64_t j = (int64_t)bit * stride + x;
      double hx = 0, hy = 0, hz = 0, mx = 0, my = 0, mz = 0;
      if (driver_filter(sf, p, rv + r)) {
        mask |= 1ull << bit;
        if ((omask >> bit) & 1ull) {
          const int64_t jo = (int64_t)bit * stride + o;
          hx = owl.hx[jo]; hy = owl.hy[jo]; hz = owl.hz[jo];
          mx = owl.mhx[jo]; my = owl.mhy[jo]; mz = owl.mhz[jo];
        }
      }
      nwl.hx[j] = hx; nwl.hy[j] = hy; nwl.hz[j] = hz;
      nwl.mhx[j] = mx; nwl.mhy[j] = my; nwl.mhz[j] = mz;
    }
  nwl.mask[x] = mask;
}

// ------------------------------------------------------------------------------------------ mesh driver
// Polyhedron against the stationary triangle mesh (RShape driver). A candidate is one 32-bit word:
//   bits 31..29  type - 7: 7 particle vertex / mesh vertex, 8 particle vertex / mesh edge, 9 particle
//                vertex / mesh face, 10 particle edge / mesh edge, 11 particle edge / mesh vertex,
//                12 particle face / mesh vertex (interaction_enum.hpp:40-46)
//   bits 28..25  particle sub-feature
//   bits 24..0   mesh feature index
constexpr uint32_t kPmTypeShift = 29, kPmSubShift = 25, kPmIdxMask = (1u << 25) - 1u;
__host__ __device__ inline uint32_t pm_word(int type, int sub, uint32_t k) {
  return ((uint32_t)(type - 7) << kPmTypeShift) | ((uint32_t)sub << kPmSubShift) | k;
}
__host__ __device__ inline int pm_type(uint32_t wd) { return 7 + (int)(wd >> kPmTypeShift); }
__host__ __device__ inline int pm_sub(uint32_t wd) { return (int)((wd >> kPmSubShift) & 15u); }

// Candidates of one particle in the loop order of add_driver_interaction for RShapeDriver
// (nbh_polyhedron_driver.hpp:96-200), over the mesh features of the particle's bin [q0, q1), which hold
// in ascending word order (vertices, edges, faces, each by index) every feature that can pass a filter:
// vertices against mesh vertices, edges, faces; edges against mesh edges; then the mesh vertices whose box
// (the vertex enlarged by r_m) meets obb_i against the particle's edges and faces. The face boxes that
// gate the vertex-face pass contain every point that passes its filter, so that gate is implied.
template <typename F>
__device__ inline void for_each_mesh_feature(const MeshView& m, const xdem_shape& sa, const Verts& va, double ra,
                                             const OBBd& obb_i, uint32_t q0, uint32_t q1, double rv, F&& fn) {
  uint32_t qe = q0;
  while (qe < q1 && (m.feat[qe] >> kMeshKindShift) == 0u) qe++;
  uint32_t qf = qe;
  while (qf < q1 && (m.feat[qf] >> kMeshKindShift) == 1u) qf++;
  const double rm = m.rm;
  for (int i = 0; i < sa.nv; i++) {
    const V3 v = va[i];
    for (uint32_t q = q0; q < qe; q++) {
      const uint32_t k = m.feat[q] & kMeshIdxMask;
      if (filter_vv(rv, v, ra, m.v[k], rm)) fn(7, i, k);
    }
    for (uint32_t q = qe; q < qf; q++) {
      const uint32_t k = m.feat[q] & kMeshIdxMask;
      const uint2 e = m.e[k];
      if (filter_ve(rv, v, ra, m.v[e.x], m.v[e.y], rm)) fn(8, i, k);
    }
    for (uint32_t q = qf; q < q1; q++) {
      const uint32_t k = m.feat[q] & kMeshIdxMask;
      const uint4 f = m.f[k];
      const V3 t[3] = {m.v[f.x], m.v[f.y], m.v[f.z]};
      double dist;
      V3 nn;
      bool early;
      if (mesh_vf_core(v, t, ra + rm + rv, dist, nn, early)) fn(9, i, k);
    }
  }
  for (int i = 0; i < sa.ne; i++) {
    const V3 a = va[sa.edges[i][0]], b = va[sa.edges[i][1]];
    for (uint32_t q = qe; q < qf; q++) {
      const uint32_t k = m.feat[q] & kMeshIdxMask;
      const uint2 e = m.e[k];
      if (filter_ee(rv, a, b, ra, m.v[e.x], m.v[e.y], rm)) fn(10, i, k);
    }
  }
  for (uint32_t q = q0; q < qe; q++) {
    const uint32_t k = m.feat[q] & kMeshIdxMask;
    const V3 mv = m.v[k];
    OBBd ov;
    ov.c = mv;
    ov.e1 = {1, 0, 0};
    ov.e2 = {0, 1, 0};
    ov.e3 = {0, 0, 1};
    ov.ext = {0, 0, 0};
    obb_enlarge(ov, rm);
    if (!obb_intersect(ov, obb_i)) continue;
    for (int i = 0; i < sa.ne; i++)
      if (filter_ve(rv, mv, rm, va[sa.edges[i][0]], va[sa.edges[i][1]], ra)) fn(11, i, k);
    for (int i = 0; i < sa.nf; i++) {
      double dist;
      V3 nn;
      bool early;
      if (vf_core(mv, va, sa, i, rm + ra + rv, dist, nn, early)) fn(12, i, k);
    }
  }
}

// Candidate lists of every particle against the mesh (slot k of particle x at k * stride + x), the
// history of each carried from the previous lists by word (the previous particle through perm). The
// particle's box for the mesh-vertex gate is the shape box rotated and moved without the homothety and
// enlarged by rVerlet, as add_driver_interaction builds it.
__global__ __launch_bounds__(256) void k_poly_mesh_carry(int64_t n, PState s, const xdem_shape* __restrict__ shapes,
                                                         const d4* __restrict__ vtx, MeshView m, double rv,
                                                         MeshList nl, MeshList ol, const uint32_t* __restrict__ perm,
                                                         int have_old, int32_t* __restrict__ over) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const Core cx = s.core[x];
  const V3 p{cx.xr.x, cx.xr.y, cx.xr.z};
  const double fx = (p.x - m.lo.x) * m.inv_h, fy = (p.y - m.lo.y) * m.inv_h, fz = (p.z - m.lo.z) * m.inv_h;
  uint32_t cnt = 0;
  if (fx >= 0.0 && fy >= 0.0 && fz >= 0.0 && fx < (double)m.dim[0] && fy < (double)m.dim[1] &&
      fz < (double)m.dim[2]) {
    const int64_t b = ((int64_t)(int)fz * m.dim[1] + (int)fy) * m.dim[0] + (int)fx;
    const uint32_t q0 = m.start[b], q1 = m.start[b + 1];
    if (q1 > q0) {
      const xdem_shape& sa = shapes[s.type[x]];
      const double ra = sa.minkowski * cx.w.w;
      const Verts va = verts_of(vtx, x, V3{0, 0, 0});
      OBBd obb_i = compute_obb(sa, p, cx.q, 1.0);
      obb_enlarge(obb_i, rv);
      const uint32_t o = have_old ? perm[x] : 0u;
      const uint32_t oc = have_old ? min(ol.cnt[o], (uint32_t)m.cap) : 0u;
      for_each_mesh_feature(m, sa, va, ra, obb_i, q0, q1, rv, [&](int t, int i, uint32_t k) {
        const uint32_t wd = pm_word(t, i, k);
        if (cnt < (uint32_t)m.cap) {
          const int64_t j = (int64_t)cnt * m.stride + x;
          double hx = 0, hy = 0, hz = 0, mx = 0, my = 0, mz = 0;
          for (uint32_t kk = 0; kk < oc; kk++) {
            const int64_t jo = (int64_t)kk * m.stride + o;
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
      });
    }
  }
  nl.cnt[x] = cnt;
  if (cnt > (uint32_t)m.cap) atomicMax(over, (int32_t)cnt);
}

int launch_poly_mesh_carry(xdem_ctx* c, double rv, const MeshList& nl, const MeshList& ol, bool have_old,
                           int32_t* over) {
  hipLaunchKernelGGL(k_poly_mesh_carry, dim3(grid_blocks(c->n, 256)), dim3(256), 0, c->stream, c->n, live(c),
                     c->d_shapes, c->vtx, c->mesh, rv, nl, ol, c->perm, have_old ? 1 : 0, over);
  XDEM_HIP(c, hipGetLastError());
  return XDEM_OK;
}

// Detection of one candidate (detect<7..12>, shape_detection_api.hpp:357-406): 7-10 with the particle
// feature as i, 11-12 with the mesh vertex as i.
__device__ inline Det poly_mesh_detect(const MeshView& m, uint32_t wd, const xdem_shape& sa, const Verts& va,
                                       double ra) {
  const int t = pm_type(wd), sub = pm_sub(wd);
  const uint32_t k = wd & kPmIdxMask;
  const double rm = m.rm;
  switch (t) {
    case 7:
      return detect_vv(va[sub], ra, m.v[k], rm);
    case 8: {
      const uint2 e = m.e[k];
      return detect_ve(va[sub], ra, m.v[e.x], m.v[e.y], rm);
    }
    case 9: {
      const uint4 f = m.f[k];
      const V3 tr[3] = {m.v[f.x], m.v[f.y], m.v[f.z]};
      const V3 vi = va[sub];
      double dist;
      V3 nn;
      bool early;
      const bool in = mesh_vf_core(vi, tr, ra + rm, dist, nn, early);
      if (early) return det_none();
      const double dn = dist - (ra + rm);
      return {in, dn, nn, vi - nn * (ra + 0.5 * dn)};
    }
    case 10: {
      const uint2 e = m.e[k];
      return detect_ee(va[sa.edges[sub][0]], va[sa.edges[sub][1]], ra, m.v[e.x], m.v[e.y], rm);
    }
    case 11:
      return detect_ve(m.v[k], rm, va[sa.edges[sub][0]], va[sa.edges[sub][1]], ra);
    default:
      return detect_vf(m.v[k], rm, va, sa, sub, ra);
  }
}

struct PolyMeshArgs {
  MeshView m;
  MeshList l;
  MeshTerms t;
  PState s;
  const xdem_shape* shapes;
  const d4* vtx;
  double dt;
  ContactParams cpd;
  const int32_t* flags;
  double* cs;  // recorded contact state per candidate slot (null: not recorded)
};

// One thread per flat candidate: ContactLawRShapeDriverFunc
// (polyhedron/contact.hpp:300-432) against the stationary mesh (its centre, no velocity, no spin). 7-10:
// the law with the particle as i, its force f and moment (cp - r) x f + m on the particle. 11-12: the law
// with the mesh vertex as i; the particle takes -f and (cp - r) x (-f) - m, and the friction is stored
// negated, as the reference does. The terms go to t.terms for the gather; -0.0 without contact.
template <bool KR, bool GUARD>
__global__ __launch_bounds__(256) void k_poly_mesh_pass(PolyMeshArgs a) {
  if (GUARD && (a.flags[1] || a.flags[0])) return;
  const int64_t jc = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // compact copy (word, history, terms)
  if (jc >= a.t.ncand) return;
  const uint2 ck = a.t.cand[jc];
  const int64_t x = ck.x;
  const int64_t j = (int64_t)ck.y * a.m.stride + x;  // list slot (recorded state)
  const MeshList& l = a.l;
  double* out = a.t.terms + jc * 6;
  const uint32_t wd = l.word[jc];
  const Core cx = a.s.core[x];
  const xdem_shape& sa = a.shapes[a.s.type[x]];
  const double ra = sa.minkowski * cx.w.w;
  const Verts va = verts_of(a.vtx, x, V3{0, 0, 0});
  const Det d = poly_mesh_detect(a.m, wd, sa, va, ra);
  if (!d.contact) {
    l.hx[jc] = 0.0; l.hy[jc] = 0.0; l.hz[jc] = 0.0;
    l.mhx[jc] = 0.0; l.mhy[jc] = 0.0; l.mhz[jc] = 0.0;
    if (a.cs) {
      double* r = a.cs + (int64_t)kCStateLen * j;
#pragma unroll
      for (int q = 0; q < kCStateLen; q++) r[q] = 0.0;
    }
    reinterpret_cast<double2*>(out)[0] = double2{-0.0, -0.0};
    reinterpret_cast<double2*>(out)[1] = double2{-0.0, -0.0};
    reinterpret_cast<double2*>(out)[2] = double2{-0.0, -0.0};
    return;
  }
  const V3 rx{cx.xr.x, cx.xr.y, cx.xr.z}, vx{cx.vm.x, cx.vm.y, cx.vm.z}, wx{cx.w.x, cx.w.y, cx.w.z};
  const double mx = cx.vm.w;
  const V3 zero{0.0, 0.0, 0.0};
  V3 ft{l.hx[jc], l.hy[jc], l.hz[jc]};
  V3 mint{0.0, 0.0, 0.0};
  if (KR) mint = {l.mhx[jc], l.mhy[jc], l.mhz[jc]};
  V3 fp, mp, fnv, fst;
  if (pm_type(wd) <= 10) {
    const V3 f = hooke_force<KR>(d.dn, d.n, a.dt, a.cpd, mx, ft, mint, d.cp, rx, vx, wx, a.m.center, zero, zero);
    fp = f;
    mp = cross(d.cp - rx, f) + mint;
    fnv = f - ft;
    fst = ft;
  } else {
    const V3 f = hooke_force<KR>(d.dn, d.n, a.dt, a.cpd, mx, ft, mint, d.cp, a.m.center, zero, zero, rx, vx, wx);
    fp = -f;
    mp = cross(d.cp - rx, fp) + (-mint);
    fnv = ft - f;
    fst = -ft;
  }
  l.hx[jc] = fst.x; l.hy[jc] = fst.y; l.hz[jc] = fst.z;
  if (KR) {
    l.mhx[jc] = mint.x; l.mhy[jc] = mint.y; l.mhz[jc] = mint.z;
  }
  if (a.cs) {
    const double rec[kCStateLen] = {d.dn, d.cp.x, d.cp.y, d.cp.z, fnv.x, fnv.y, fnv.z, fst.x,
                                    fst.y, fst.z, rx.x,   rx.y,   rx.z,   0.0,   0.0,   0.0};
    double* r = a.cs + (int64_t)kCStateLen * j;
#pragma unroll
    for (int q = 0; q < kCStateLen; q++) r[q] = rec[q];
  }
  reinterpret_cast<double2*>(out)[0] = double2{fp.x, fp.y};
  reinterpret_cast<double2*>(out)[1] = double2{fp.z, mp.x};
  reinterpret_cast<double2*>(out)[2] = double2{mp.y, mp.z};
}

template <typename T>
static int galloc(xdem_ctx* c, T** p, int64_t count) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  XDEM_HIP(c, hipMalloc((void**)p, sizeof(T) * (size_t)std::max<int64_t>(count, 1)));
  return XDEM_OK;
}

int poly_nbh(xdem_ctx* c, double rcut_inc) {
  if (!c->nl_valid) return set_err(c, XDEM_ERR_STATE, "nbh_polyhedron: run chunk_neighbors_contact first");
  const int64_t n = c->n;
  int rc;
  if ((rc = poly_compute_vertices(c, false, 0.0, false))) return rc;
  const NbrList& L = c->nl[c->nl_cur];
  const int nxt = 1 - c->pl_cur;
  PolyList& P = c->pl[nxt];
  if (L.slots + 1 > P.cap_slots) {
    const int64_t cap = (int64_t)(L.slots * 1.25) + 4096;
    if ((rc = galloc(c, &P.start, cap + 1)) || (rc = galloc(c, &P.fcount, cap))) return rc;
    P.cap_slots = cap;
  }
  if (L.slots > c->cap_frange) {
    const int64_t cap = (int64_t)(L.slots * 1.25) + 4096;
    if ((rc = galloc(c, &c->frange, cap))) return rc;
    c->cap_frange = cap;
  }
  if (L.slots > c->cap_feat_tmp) {
    const int64_t cap = (int64_t)(L.slots * 1.25) + 4096;
    if ((rc = galloc(c, &c->feat_tmp, cap * kFeatPerSlot))) return rc;
    c->cap_feat_tmp = cap;
  }
  if (L.slots > c->cap_oslot) {
    const int64_t cap = (int64_t)(L.slots * 1.25) + 4096;
    // the pool holds about one chunk per 128 slots (pairs past it are enumerated again);
    // XDEM_FEAT_POOL_WORDS sets another capacity (the tests shrink it to take the re-enumeration path)
    int64_t pcap = std::min<int64_t>(cap / 128 * kPoolChunk + 64 * kPoolChunk, 0x7fffffffLL);
    if (const char* e = std::getenv("XDEM_FEAT_POOL_WORDS"))
      pcap = std::min<int64_t>(std::max<int64_t>(std::atoll(e), kPoolChunk), 0x7fffffffLL);
    if ((rc = galloc(c, &c->oslot, cap)) || (rc = galloc(c, &c->feat_ovf, cap)) ||
        (rc = galloc(c, &c->feat_fpool, cap)) || (rc = galloc(c, &c->feat_pool, pcap)))
      return rc;
    c->cap_oslot = cap;
    c->cap_feat_pool = pcap;
  }
  if (!c->d_nown && (rc = galloc(c, &c->d_nown, 4))) return rc;
  XDEM_HIP(c, hipMemsetAsync(c->d_nown, 0, 4 * sizeof(uint32_t), c->stream));
  XDEM_HIP(c, hipMemsetAsync(P.fcount, 0, sizeof(uint32_t) * std::max<int64_t>(L.slots, 1), c->stream));
  PolyNbhArgs a{n, live(c), c->grid, c->d_shapes, c->vtx, L, rcut_inc};
  const dim3 gsl(grid_blocks(std::max<int64_t>(L.slots, 1), 256)), bsl(256);
  if (n > 0) {
    hipLaunchKernelGGL(k_owner_slots, dim3(grid_blocks(std::max<int64_t>(L.slots, 1), 256 * kOwnChunk)), dim3(256), 0,
                       c->stream, L, c->oslot, c->d_nown);
    hipLaunchKernelGGL(k_feat_enum, dim3(grid_blocks(std::max<int64_t>(L.slots, 1), kEnumBlock)), dim3(kEnumBlock), 0,
                       c->stream, a, c->oslot, c->d_nown, P.fcount, c->feat_tmp, c->feat_pool,
                       (uint32_t)c->cap_feat_pool, c->feat_fpool);
    if ((rc = scan_u64_from_u32(c, P.fcount, P.start, L.slots, 1u))) return rc;
  }
  uint64_t nf = 0;
  if (n > 0 && (rc = read_u64(c, P.start + L.slots, &nf))) return rc;
  if ((int64_t)nf > P.cap_feat) {
    const int64_t cap = (int64_t)(nf * 1.25) + 4096;
    if ((rc = galloc(c, &P.feat, cap)) || (rc = galloc(c, &P.hx, cap)) || (rc = galloc(c, &P.hy, cap)) ||
        (rc = galloc(c, &P.hz, cap)) || (rc = galloc(c, &P.mhx, cap)) || (rc = galloc(c, &P.mhy, cap)) ||
        (rc = galloc(c, &P.mhz, cap)) || (rc = galloc(c, &P.fslot, cap)) || (rc = galloc(c, &P.fown, cap)))
      return rc;
    P.cap_feat = cap;
  }
  if ((int64_t)nf > c->cap_contrib) {
    const int64_t cap = (int64_t)(nf * 1.25) + 4096;
    if ((rc = galloc(c, &c->fcontrib, cap * 9)) || (rc = galloc(c, &c->fcontact, cap)) ||
        (rc = galloc(c, &c->fdn, cap)))
      return rc;
    c->cap_contrib = cap;
  }
  if (nf >= 0xffffffffull) return set_err(c, XDEM_ERR_CAPACITY, "nbh_polyhedron: more than 2^32 - 2 features");
  P.nfeat = (int64_t)nf;
  if (n > 0) {
    hipLaunchKernelGGL(k_feat_place, gsl, bsl, 0, c->stream, L, c->oslot, c->d_nown, P.fcount, c->feat_tmp,
                       c->feat_pool, c->feat_fpool, P.start, P.feat, P.fslot, P.fown, c->feat_ovf);
    hipLaunchKernelGGL(k_feat_overflow, dim3(kOvfBlocks), dim3(kEnumBlock), 0, c->stream, a, c->d_nown, c->feat_ovf,
                       P.start, P.feat, P.fslot, P.fown);
    hipLaunchKernelGGL(k_feat_range, dim3(grid_blocks(L.slots, 256)), dim3(256), 0, c->stream, L, P.start,
                       c->frange);
    const bool have_old = c->nl_prev_valid && c->perm_valid && c->pl_valid;
    FeatCarryArgs fa{n, L, c->nl[1 - c->nl_cur], P, c->pl[c->pl_cur], c->perm, have_old ? 1 : 0};
    hipLaunchKernelGGL(k_feat_carry, dim3(grid_blocks(L.slots, 256)), dim3(256), 0, c->stream, fa);
    XDEM_HIP(c, hipGetLastError());
    // vertex-wall candidates
    if (c->nwalls > 0) {
      const int pn = 1 - c->pw_cur;
      const int64_t m = (int64_t)c->nwalls * XDEM_MAX_SHAPE_VERTICES * c->cap;
      for (int k = 0; k < 2; k++) {
        PolyWall& W = c->pw[k];
        if (W.cap < m) {
          if ((rc = galloc(c, &W.mask, c->cap)) || (rc = galloc(c, &W.hx, m)) || (rc = galloc(c, &W.hy, m)) ||
              (rc = galloc(c, &W.hz, m)) || (rc = galloc(c, &W.mhx, m)) || (rc = galloc(c, &W.mhy, m)) ||
              (rc = galloc(c, &W.mhz, m)))
            return rc;
          W.cap = m;
          if (k == c->pw_cur) c->pw_valid = false;
        }
      }
      const bool wold = have_old && c->pw_valid;
      hipLaunchKernelGGL(k_poly_wall, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, n, live(c), c->d_shapes,
                         c->vtx, c->nwalls, c->d_walls, rcut_inc, c->pw[pn], c->pw[c->pw_cur], c->perm,
                         wold ? 1 : 0, c->cap);
      XDEM_HIP(c, hipGetLastError());
      c->pw_cur = pn;
      c->pw_valid = true;
    }
    if ((rc = mesh_carry(c, rcut_inc, have_old, true))) return rc;
  }
  c->pl_cur = nxt;
  c->pl_valid = true;
  c->pl_epoch = c->nl_epoch;
  c->nl_prev_valid = false;
  c->perm_valid = false;
  c->hist_ready = true;
  return XDEM_OK;
}

// ------------------------------------------------------------------------------------------ classifier
// classify_interactions (src/interaction/classify_interactions.cpp:52-136, classifier.hpp:33-271): the
// feature interactions split by type (VV, VE, VF, EE), each type keeping the list order. A wave entry
// carries what the contact pass needs of a feature, as the reference's per-type containers carry the
// interaction's fields (classifier_container.hpp:87-300): the feature's index in the lists, its owner,
// the owner's list word (partner and image code) and the feature word.
// The split is made per window of 1024 consecutive features: each window's features of one type are
// contiguous, padded (index ~0u) to whole wavefronts, so every wavefront of the contact pass evaluates one
// type while all the writes of a window (history, contact terms) stay within the window's cache lines. The
// wave of a type is its segments across the windows, in window order.
constexpr int kClassPad = 64;         // wavefront: a padded segment is a whole number of them
constexpr int kClassPerBlock = 1024;  // window of the partition (256 threads x 4 features)

__device__ inline int feat_type_or_none(int64_t f, int64_t nf, const uint32_t* feat) {
  return f < nf ? (int)(feat[f] & kFeatTypeMask) : -1;
}

__global__ __launch_bounds__(256) void k_class_count(int64_t nf, const uint32_t* __restrict__ feat,
                                                     uint32_t* __restrict__ blk) {
  __shared__ uint32_t cnt[4];
  if (threadIdx.x < 4) cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t f0 = (int64_t)blockIdx.x * kClassPerBlock + 4 * (int64_t)threadIdx.x;
  uint32_t loc[4] = {0, 0, 0, 0};
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int t = feat_type_or_none(f0 + u, nf, feat);
#pragma unroll
    for (int q = 0; q < 4; q++) loc[q] += t == q ? 1u : 0u;
  }
#pragma unroll
  for (int q = 0; q < 4; q++)
    if (loc[q]) atomicAdd(&cnt[q], loc[q]);
  __syncthreads();
  if (threadIdx.x < 4) blk[(int64_t)blockIdx.x * 4 + threadIdx.x] = cnt[threadIdx.x];
}

// blk: per window and type counts -> start of the window's segment of that type, in (window, type)
// order; info: [4] padded total, [5..8] features per type. One block of 1024 threads: each sums a
// contiguous run of the (window, type) counts, a block scan of the run sums, then each writes its
// run's starts.
__global__ __launch_bounds__(1024) void k_class_scan(int64_t nblk, uint32_t* __restrict__ blk,
                                                     int64_t* __restrict__ info) {
  __shared__ int64_t part[1024];
  __shared__ int64_t tsz[4][1024];
  const int t = (int)threadIdx.x;
  const int64_t m = nblk * 4;
  const int64_t chunk = (m + 1023) / 1024;
  const int64_t lo = min(m, (int64_t)t * chunk), hi = min(m, lo + chunk);
  int64_t sum = 0, sz[4] = {0, 0, 0, 0};
  for (int64_t e = lo; e < hi; e++) {
    const uint32_t v = blk[e];
    sum += (v + kClassPad - 1) / kClassPad * kClassPad;
    sz[e & 3] += v;
  }
  part[t] = sum;
  for (int q = 0; q < 4; q++) tsz[q][t] = sz[q];
  __syncthreads();
  // inclusive scan of part (Hillis-Steele)
  for (int o = 1; o < 1024; o <<= 1) {
    const int64_t y = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += y;
    __syncthreads();
  }
  int64_t run = part[t] - sum;
  for (int64_t e = lo; e < hi; e++) {
    const uint32_t v = blk[e];
    blk[e] = (uint32_t)run;
    run += (v + kClassPad - 1) / kClassPad * kClassPad;
  }
  // per-type totals (tree reduction)
  for (int o = 512; o > 0; o >>= 1) {
    if (t < o)
      for (int q = 0; q < 4; q++) tsz[q][t] += tsz[q][t + o];
    __syncthreads();
  }
  if (t == 0) {
    for (int q = 0; q < 4; q++) {
      info[q] = 0;
      info[5 + q] = tsz[q][0];
    }
    info[4] = part[1023];
  }
}

__global__ __launch_bounds__(256) void k_class_scatter(int64_t nf, const uint32_t* __restrict__ feat,
                                                       const uint32_t* __restrict__ fslot,
                                                       const uint32_t* __restrict__ fown,
                                                       const uint32_t* __restrict__ nlword,
                                                       const uint32_t* __restrict__ blk,
                                                       const int64_t* __restrict__ info,
                                                       uint4* __restrict__ wave) {
  __shared__ uint32_t wsum[4][4];  // [type][wavefront of the block]
  const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;
  const int64_t f0 = (int64_t)blockIdx.x * kClassPerBlock + 4 * (int64_t)threadIdx.x;
  int ty[4];
  uint32_t loc[4] = {0, 0, 0, 0};
#pragma unroll
  for (int u = 0; u < 4; u++) {
    ty[u] = feat_type_or_none(f0 + u, nf, feat);
#pragma unroll
    for (int q = 0; q < 4; q++) loc[q] += ty[u] == q ? 1u : 0u;
  }
  // rank of this thread's first feature of each type among the block's features of that type
  uint32_t excl[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    uint32_t inc = loc[q];
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(inc, o, 64);
      if (lane >= o) inc += y;
    }
    excl[q] = inc - loc[q];
    if (lane == 63) wsum[q][wv] = inc;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; q++)
    for (int w = 0; w < wv; w++) excl[q] += wsum[q][w];
  const int64_t b = blockIdx.x;
#pragma unroll
  for (int u = 0; u < 4; u++) {
    int64_t pos = -1;
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (ty[u] == q) {
        pos = (int64_t)blk[b * 4 + q] + (int64_t)excl[q];
        excl[q]++;
      }
    if (pos >= 0) {
      const int64_t f = f0 + u;
      wave[pos] = uint4{(uint32_t)f, fown[f], nlword[fslot[f]] & ~kActiveBit, feat[f]};
    }
  }
}

int poly_classify(xdem_ctx* c) {
  if (!c->pl_valid || c->pl_epoch != c->nl_epoch)
    return set_err(c, XDEM_ERR_STATE, "classify_interactions: run nbh_polyhedron first");
  const PolyList& P = c->pl[c->pl_cur];
  const int64_t nf = P.nfeat;
  const int64_t nblk = std::max<int64_t>(1, (nf + kClassPerBlock - 1) / kClassPerBlock);
  int rc;
  if (nblk * 4 > c->cap_fclass_blk) {
    if ((rc = dev_realloc(c, &c->fclass_blk, nblk * 4))) return rc;
    c->cap_fclass_blk = nblk * 4;
  }
  if (!c->fclass_info && (rc = dev_realloc(c, &c->fclass_info, 9))) return rc;
  XDEM_HIP(c, hipMemsetAsync(c->fclass_blk, 0, sizeof(uint32_t) * nblk * 4, c->stream));
  if (nf > 0)
    hipLaunchKernelGGL(k_class_count, dim3((uint32_t)nblk), dim3(256), 0, c->stream, nf, P.feat, c->fclass_blk);
  hipLaunchKernelGGL(k_class_scan, dim3(1), dim3(1024), 0, c->stream, nblk, c->fclass_blk, c->fclass_info);
  XDEM_HIP(c, hipGetLastError());
  int64_t info[9];
  XDEM_HIP(c, hipMemcpyAsync(info, c->fclass_info, sizeof(info), hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  const int64_t total = info[4];
  if (total > c->cap_forder) {
    const int64_t cap = (int64_t)(total * 1.25) + 4096;
    if ((rc = dev_realloc(c, &c->forder, cap))) return rc;
    c->cap_forder = cap;
  }
  if (total > 0) {
    XDEM_HIP(c, hipMemsetAsync(c->forder, 0xff, sizeof(uint4) * total, c->stream));
    if (nf > 0)
      hipLaunchKernelGGL(k_class_scatter, dim3((uint32_t)nblk), dim3(256), 0, c->stream, nf, P.feat, P.fslot, P.fown,
                         c->nl[c->nl_cur].word, c->fclass_blk, c->fclass_info, c->forder);
    XDEM_HIP(c, hipGetLastError());
  }
  for (int t = 0; t < 4; t++) {
    c->fwave_start[t] = info[t];
    c->fwave_size[t] = info[5 + t];
  }
  c->fwave_start[4] = total;
  c->pl_class_epoch = c->nl_epoch;
  return XDEM_OK;
}

// ------------------------------------------------------------------------------------------ forces
struct PolyForceArgs {
  int64_t n;
  PState s;
  Grid g;
  const xdem_shape* shapes;
  const d4* vtx;
  NbrList nl;
  PolyList pl;
  PolyWall pw;
  const uint2* frange;
  uint4* wave;  // classifier waves: {feature, owner, owner list word, feature word}; feature ~0u pads
  int64_t norder;
  // when recording: contact state per feature and per (wall, vertex, particle), kCStateLen doubles
  // each, in the evaluation orientation (null: not recorded)
  double* cstate;
  double* wstate;
  double* contrib;    // [nfeat * 9]
  uint8_t* contact;   // [nfeat]
  double* fdn;        // [nfeat] dn at this pass (0 without contact), for simulation_state
  const Surface* walls;
  int nw;
  int64_t wstride;
  ContactParams cp, cpd;
  double dt;
  V3 grav;
  int gravity_on;
  int32_t* flags;
  int32_t step_tag;
  MeshTerms mt;  // terms of the mesh contacts, evaluated by k_poly_mesh_pass just before (ncand = 0: none)
};

__device__ inline bool poly_guard(const PolyForceArgs& a) {
  if (a.flags[1]) return true;
  if (a.flags[0]) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      a.flags[1] = 1;
      a.flags[2] = a.step_tag;
    }
    return true;
  }
  return false;
}

// One thread per feature: detection, Hooke law and history of that feature; its terms of the owner
// force, owner moment and partner moment go to contrib (9 doubles) when it is in contact. The pair
// subtotals are formed by k_poly_gather in feature order.
template <bool KR, bool GUARD>
__global__ __launch_bounds__(128) void k_poly_feat(PolyForceArgs a) {
  if (GUARD && poly_guard(a)) return;
  // XCD-aware block order: consecutive wave entries (owners close in the order) share an XCD's L2. Every
  // wavefront lies inside one type segment of the classifier, so it evaluates one feature type.
  const int64_t t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= a.norder) return;
  const uint4 we = a.wave[t];
  if (we.x == 0xffffffffu) return;
  const int64_t f = we.x;
  const PState& s = a.s;
  const uint32_t x = we.y;
  const uint32_t wd = we.z;
  const uint32_t y = wd & kIdxMask;
  const V3 sh = shift_vec(a.g, (int)((wd >> kShiftShift) & 31u));
  // detection needs only the homotheties; the rest of the two records is read for contacts only
  const double hx = s.core[x].w.w, hy = s.core[y].w.w;
  const xdem_shape& sa = a.shapes[s.type[x]];
  const xdem_shape& sb = a.shapes[s.type[y]];
  const double ra = sa.minkowski * hx, rb = sb.minkowski * hy;
  const Verts va = verts_of(a.vtx, x, V3{0, 0, 0});
  const Verts vb = verts_of(a.vtx, y, sh);
  const uint32_t fw = we.w;
  const int type = (int)(fw & kFeatTypeMask);
  const bool swap = (fw & kFeatSwap) != 0;
  const int osub = (int)((fw >> 3) & 15u), psub = (int)((fw >> 7) & 15u);
  // evaluation particle i is the vertex side: the owner unless swap
  const Verts& vi = swap ? vb : va;
  const Verts& vj = swap ? va : vb;
  const xdem_shape& si = swap ? sb : sa;
  const xdem_shape& sj = swap ? sa : sb;
  const double ri = swap ? rb : ra, rj = swap ? ra : rb;
  const int fi = swap ? psub : osub, fj = swap ? osub : psub;
  Det d;
  if (type == 0) d = detect_vv(vi[fi], ri, vj[fj], rj);
  else if (type == 1) d = detect_ve(vi[fi], ri, vj[sj.edges[fj][0]], vj[sj.edges[fj][1]], rj);
  else if (type == 2) d = detect_vf(vi[fi], ri, vj, sj, fj, rj);
  else d = detect_ee(vi[si.edges[fi][0]], vi[si.edges[fi][1]], ri, vj[sj.edges[fj][0]], vj[sj.edges[fj][1]], rj);
  if (d.contact) {
    const Core cx = s.core[x], cy = s.core[y];
    const V3 rx{cx.xr.x, cx.xr.y, cx.xr.z}, vx{cx.vm.x, cx.vm.y, cx.vm.z}, wx{cx.w.x, cx.w.y, cx.w.z};
    const V3 ry{cy.xr.x + sh.x, cy.xr.y + sh.y, cy.xr.z + sh.z}, vy{cy.vm.x, cy.vm.y, cy.vm.z},
        wy{cy.w.x, cy.w.y, cy.w.z};
    const double mx = cx.vm.w, my = cy.vm.w;
    const V3 r_i = swap ? ry : rx, r_j = swap ? rx : ry;
    const V3 v_i = swap ? vy : vx, v_j = swap ? vx : vy;
    const V3 w_i = swap ? wy : wx, w_j = swap ? wx : wy;
    const double mi = swap ? my : mx, mj = swap ? mx : my;
    const double meff = (mi * mj) / (mi + mj);
    V3 ft{0, 0, 0}, mint{0, 0, 0};
    if (fw & kFeatActive) {
      ft = {a.pl.hx[f], a.pl.hy[f], a.pl.hz[f]};
      if (KR) mint = {a.pl.mhx[f], a.pl.mhy[f], a.pl.mhz[f]};
    }
    const V3 fc = hooke_force<KR>(d.dn, d.n, a.dt, a.cp, meff, ft, mint, d.cp, r_i, v_i, w_i, r_j, v_j, w_j);
    const bool act = !(ft.x == 0 && ft.y == 0 && ft.z == 0 && mint.x == 0 && mint.y == 0 && mint.z == 0);
    if (act || (fw & kFeatActive)) {
      a.pl.hx[f] = ft.x; a.pl.hy[f] = ft.y; a.pl.hz[f] = ft.z;
      if (KR) {
        a.pl.mhx[f] = mint.x; a.pl.mhy[f] = mint.y; a.pl.mhz[f] = mint.z;
      }
    }
    const uint32_t nfw = act ? (fw | kFeatActive) : (fw & ~kFeatActive);
    if (nfw != fw) {
      a.pl.feat[f] = nfw;
      a.wave[t].w = nfw;
    }
    if (a.cstate) {
      // the reference functor's analysis tuple (polyhedron/contact.hpp:160-210) with the positions:
      // i is the vertex side, as the reference's swapped pair stores it
      const V3 fnv = fc - ft;
      const double rec[kCStateLen] = {d.dn,  d.cp.x, d.cp.y, d.cp.z, fnv.x, fnv.y, fnv.z, ft.x,
                                      ft.y,  ft.z,   r_i.x,  r_i.y,  r_i.z,  r_j.x, r_j.y, r_j.z};
      double* rc = a.cstate + (uint64_t)kCStateLen * f;
#pragma unroll
      for (int q = 0; q < kCStateLen; q++) rc[q] = rec[q];
    }
    const V3 nf = -fc;
    V3 Fo, Mo, Mp;
    if (!swap) {  // owner is i
      Fo = fc;
      Mo = cross(d.cp - r_i, fc) + mint;
      Mp = cross(d.cp - r_j, nf) + (-mint);
    } else {  // owner is j
      Fo = nf;
      Mo = cross(d.cp - r_j, nf) + (-mint);
      Mp = cross(d.cp - r_i, fc) + mint;
    }
    a.fdn[f] = d.dn;
    double* pr = a.contrib + 9 * f;
    pr[0] = Fo.x; pr[1] = Fo.y; pr[2] = Fo.z;
    pr[3] = Mo.x; pr[4] = Mo.y; pr[5] = Mo.z;
    pr[6] = Mp.x; pr[7] = Mp.y; pr[8] = Mp.z;
    a.contact[f] = 1;
  } else {
    a.contact[f] = 0;
    a.fdn[f] = 0.0;
    if (a.cstate) {
      double* rc = a.cstate + (uint64_t)kCStateLen * f;
#pragma unroll
      for (int q = 0; q < kCStateLen; q++) rc[q] = 0.0;
    }
    if (fw & kFeatActive) {
      a.pl.hx[f] = 0.0; a.pl.hy[f] = 0.0; a.pl.hz[f] = 0.0;
      a.pl.mhx[f] = 0.0; a.pl.mhy[f] = 0.0; a.pl.mhz[f] = 0.0;
      a.pl.feat[f] = fw & ~kFeatActive;
      a.wave[t].w = fw & ~kFeatActive;
    }
  }
}

struct Q4p {
  double w, x, y, z;
};
__device__ inline V3 angular_accel_p(const d4& q, V3 mom, V3 vrot, V3 I) {
  const d4 qi{q.x, -q.y, -q.z, -q.w};
  const V3 om = qrot4(qi, vrot);
  const V3 M = qrot4(qi, mom);
  const V3 dom{(M.x - (I.z - I.y) * om.y * om.z) / I.x, (M.y - (I.x - I.z) * om.z * om.x) / I.y,
               (M.z - (I.y - I.x) * om.x * om.y) / I.z};
  return qrot4(q, dom);
}

template <bool FUSED, bool KR, bool GUARD>
__global__ __launch_bounds__(256) void k_poly_gather(PolyForceArgs a) {
  // the stall decision was taken by k_poly_feat of the same step
  if (GUARD && a.flags[1]) return;
  // XCD-aware block order (partner contributions are read through the owner slots' feature ranges)
  const int64_t b0 = xcd_block(blockIdx.x, gridDim.x) * blockDim.x;
  // the block's 256 records, staged in LDS with coalesced 16-byte copies: the mass is read from there
  // before the list walk and the rest of the record after it, without holding it in registers
  __shared__ Core tile[256];
  stage_tile(tile, a.s.core, b0, a.n);
  __syncthreads();
  const int64_t x = b0 + threadIdx.x;
  if (x >= a.n) return;
  const PState& s = a.s;
  const double mx = tile[threadIdx.x].vm.w;
  V3 F, M;
  if (FUSED) {
    F = {0.0, 0.0, 0.0};
    M = {0.0, 0.0, 0.0};
    if (a.gravity_on) F = {F.x + a.grav.x * mx, F.y + a.grav.y * mx, F.z + a.grav.z * mx};
  } else {
    F = {s.ax[x], s.ay[x], s.az[x]};
    M = {s.mx[x], s.my[x], s.mz[x]};
  }
  const uint32_t cnt = a.nl.cnt[x];
  const uint64_t b = a.nl.base[x >> 6] + (uint64_t)(x & 63);
  // per entry, the pair subtotal in feature order (as the owner's pass forms it), then the sum; the
  // word and feature range of entry k+1 are loaded while entry k is summed
  uint32_t wd_n = 0;
  uint2 fr_n{0u, 0u};
  if (cnt > 0) {
    wd_n = a.nl.word[b];
    fr_n = a.frange[b];
  }
  for (uint32_t k = 0; k < cnt; k++) {
    const uint64_t sl = b + (uint64_t)k * kWave;
    const uint32_t wd = wd_n;
    const uint2 fr = fr_n;
    if (k + 1 < cnt) {
      wd_n = a.nl.word[sl + kWave];
      fr_n = a.frange[sl + kWave];
    }
    const bool own = (wd & kOwnerBit) != 0;
    if (fr.x == 0xffffffffu) continue;
    V3 Fp{0, 0, 0}, Mp{0, 0, 0};
    const uint64_t f0 = fr.x, f1 = (uint64_t)fr.x + fr.y;
    // the pair's contact flags as aligned 32-bit words, four words (16 flags) in flight at a time (the
    // flag array has at least 4096 bytes of padding), then its features in contact in feature order
    const uint32_t* __restrict__ cw = reinterpret_cast<const uint32_t*>(a.contact);
    const uint64_t wend = (f1 + 3) >> 2;
    for (uint64_t w0 = f0 >> 2; w0 < wend; w0 += 4) {
      uint32_t v[4];
#pragma unroll
      for (int q = 0; q < 4; q++) v[q] = w0 + q < wend ? cw[w0 + q] : 0u;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        // the word's features in contact; for each half word in which any lane of the wavefront has one,
        // the two features' terms are loaded together (a position without a hit reads row 0, which exists
        // whenever some feature is in contact, and discards it), then added in feature order
        uint32_t hits = 0;
#pragma unroll
        for (int byte = 0; byte < 4; byte++) {
          const uint64_t f = ((w0 + q) << 2) + byte;
          if (((v[q] >> (8 * byte)) & 0xffu) != 0u && f >= f0 && f < f1) hits |= 1u << byte;
        }
#pragma unroll
        for (int h = 0; h < 4; h += 2) {
          const uint32_t hh = (hits >> h) & 3u;
          if (__builtin_amdgcn_ballot_w64(hh != 0u) == 0ull) continue;
          V3 tf[2], tm[2];
#pragma unroll
          for (int e = 0; e < 2; e++) {
            const uint64_t f = ((hh >> e) & 1u) ? ((w0 + q) << 2) + h + e : 0u;
            const double* pr = a.contrib + 9 * f;
            tf[e] = V3{pr[0], pr[1], pr[2]};
            tm[e] = own ? V3{pr[3], pr[4], pr[5]} : V3{pr[6], pr[7], pr[8]};
          }
#pragma unroll
          for (int e = 0; e < 2; e++) {
            if (!((hh >> e) & 1u)) continue;
            if (own) Fp = Fp + tf[e];
            else Fp = Fp + V3{-tf[e].x, -tf[e].y, -tf[e].z};
            Mp = Mp + tm[e];
          }
        }
      }
    }
    F = F + Fp;
    M = M + Mp;
  }
  // the record's fields are read from the tile where they are used below
  const Core& cx = tile[threadIdx.x];
  if (a.nw > 0) {
    const xdem_shape& sh = a.shapes[s.type[x]];
    const double r = sh.minkowski * cx.w.w;
    const uint64_t mask = a.pw.mask[x];
    const V3 rx{cx.xr.x, cx.xr.y, cx.xr.z}, vx{cx.vm.x, cx.vm.y, cx.vm.z}, wx{cx.w.x, cx.w.y, cx.w.z};
    for (int w = 0; w < a.nw; w++)
      for (int k = 0; k < sh.nv; k++) {
        const int bit = w * XDEM_MAX_SHAPE_VERTICES + k;
        if (!((mask >> bit) & 1ull)) continue;
        const int64_t j = (int64_t)bit * a.wstride + x;
        const d4 v4 = a.vtx[x * XDEM_MAX_SHAPE_VERTICES + k];
        const V3 p{v4.x, v4.y, v4.z};
        const Surface& sf = a.walls[w];
        double dn;
        V3 nrm, cpos;
        if (!driver_detect(sf, p, r, dn, nrm, cpos)) {
          a.pw.hx[j] = 0.0; a.pw.hy[j] = 0.0; a.pw.hz[j] = 0.0;
          a.pw.mhx[j] = 0.0; a.pw.mhy[j] = 0.0; a.pw.mhz[j] = 0.0;
          if (a.wstate) {
            double* rc = a.wstate + (uint64_t)kCStateLen * j;
#pragma unroll
            for (int q = 0; q < kCStateLen; q++) rc[q] = 0.0;
          }
          continue;
        }
        V3 ft{a.pw.hx[j], a.pw.hy[j], a.pw.hz[j]};
        V3 mint{0.0, 0.0, 0.0};
        if (KR) mint = {a.pw.mhx[j], a.pw.mhy[j], a.pw.mhz[j]};
        const V3 f = hooke_force<KR>(dn, nrm, a.dt, a.cpd, mx, ft, mint, cpos, rx, vx, wx, sf.center, sf.vel, sf.vrot);
        if (a.wstate) {
          const V3 fnv = f - ft;
          const double rec[kCStateLen] = {dn,   cpos.x, cpos.y, cpos.z, fnv.x, fnv.y, fnv.z, ft.x,
                                          ft.y, ft.z,   rx.x,   rx.y,   rx.z,   0.0,   0.0,   0.0};
          double* rc = a.wstate + (uint64_t)kCStateLen * j;
#pragma unroll
          for (int q = 0; q < kCStateLen; q++) rc[q] = rec[q];
        }
        a.pw.hx[j] = ft.x; a.pw.hy[j] = ft.y; a.pw.hz[j] = ft.z;
        if (KR) {
          a.pw.mhx[j] = mint.x; a.pw.mhy[j] = mint.y; a.pw.mhz[j] = mint.z;
        }
        F = F + f;
        M = M + (cross(cpos - rx, f) + mint);
      }
  }
  if (a.mt.ncand) mesh_add(a.mt, x, F, M);
  if (FUSED) {
    s.ax[x] = F.x / mx;
    s.ay[x] = F.y / mx;
    s.az[x] = F.z / mx;
    s.mx[x] = M.x;
    s.my[x] = M.y;
    s.mz[x] = M.z;
    const V3 al = angular_accel_p(cx.q, M, V3{cx.w.x, cx.w.y, cx.w.z}, V3{s.ix[x], s.iy[x], s.iz[x]});
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

int poly_force(xdem_ctx* c, const ContactParams& cp, const ContactParams& cpd, double dt, bool fused, bool gravity_on,
               V3 g, bool guarded, int32_t step_tag) {
  if (!c->pl_valid || !c->hist_ready || c->pl_epoch != c->nl_epoch)
    return set_err(c, XDEM_ERR_STATE, "contact_polyhedron: run nbh_polyhedron after chunk_neighbors_contact");
  if (c->pl_class_epoch != c->nl_epoch)
    return set_err(c, XDEM_ERR_STATE, "contact_polyhedron: run classify_interactions after nbh_polyhedron");
  // feature and vertex-wall moment histories are written in full by every neighbour build
  if (cp.kr != 0.0 || cpd.kr != 0.0) c->kr_used = true;
  if (c->n == 0) return XDEM_OK;
  PolyForceArgs a{};
  a.n = c->n;
  a.s = live(c);
  a.g = c->grid;
  a.shapes = c->d_shapes;
  a.vtx = c->vtx;
  a.nl = c->nl[c->nl_cur];
  a.pl = c->pl[c->pl_cur];
  a.frange = c->frange;
  a.wave = c->forder;
  a.norder = c->fwave_start[4];
  a.contrib = c->fcontrib;
  a.contact = c->fcontact;
  a.fdn = c->fdn;
  a.nw = (c->nwalls > 0 && c->pw_valid) ? c->nwalls : 0;
  if (a.nw) a.pw = c->pw[c->pw_cur];
  a.walls = walls_for_pass(c);
  a.wstride = c->cap;
  a.cp = cp;
  a.cpd = cpd;
  a.dt = dt;
  a.grav = g;
  a.gravity_on = gravity_on ? 1 : 0;
  a.flags = c->d_flags;
  a.step_tag = step_tag;
  int rc;
  if (c->record_cs) {
    if (a.pl.nfeat > c->cap_cstate) {
      const int64_t cap = (int64_t)(a.pl.nfeat * 1.25) + 4096;
      if ((rc = dev_realloc(c, &c->cstate, cap * kCStateLen))) return rc;
      c->cap_cstate = cap;
    }
    a.cstate = c->cstate;
    if (a.nw) {
      const int64_t m = (int64_t)c->nwalls * XDEM_MAX_SHAPE_VERTICES * c->cap;
      if (m > c->cap_wstate) {
        if ((rc = dev_realloc(c, &c->wstate, m * kCStateLen))) return rc;
        c->cap_wstate = m;
      }
      a.wstate = c->wstate;
    }
  }
  if (c->has_mesh && c->ml_valid && c->mesh_ncand > 0) {
    if (!c->mesh_poly)
      return set_err(c, XDEM_ERR_STATE, "mesh contacts of polyhedra: the mesh lists were built by nbh_sphere");
    double* mcs = nullptr;
    if (c->record_cs) {
      const int64_t m = (int64_t)c->mesh_cap * c->cap;
      if (m > c->cap_mstate) {
        if ((rc = dev_realloc(c, &c->mstate, m * kCStateLen))) return rc;
        c->cap_mstate = m;
      }
      mcs = c->mstate;
    }
    if ((rc = mesh_compact(c))) return rc;
    c->mc_dirty = true;
    PolyMeshArgs pm{c->mesh, c->mc, MeshTerms{}, a.s, c->d_shapes, c->vtx, dt, cpd, c->d_flags, mcs};
    pm.t.cand = c->mesh_cand;
    pm.t.idx = c->mesh_idx;
    pm.t.wcnt = c->mesh_wcnt;
    pm.t.terms = c->mesh_terms;
    pm.t.ncand = c->mesh_ncand;
    a.mt = pm.t;
    const dim3 gm(grid_blocks(c->mesh_ncand, 256)), bm(256);
    if (c->kr_used) {
      if (guarded) hipLaunchKernelGGL((k_poly_mesh_pass<true, true>), gm, bm, 0, c->stream, pm);
      else hipLaunchKernelGGL((k_poly_mesh_pass<true, false>), gm, bm, 0, c->stream, pm);
    } else {
      if (guarded) hipLaunchKernelGGL((k_poly_mesh_pass<false, true>), gm, bm, 0, c->stream, pm);
      else hipLaunchKernelGGL((k_poly_mesh_pass<false, false>), gm, bm, 0, c->stream, pm);
    }
    XDEM_HIP(c, hipGetLastError());
  }
  const dim3 gp(grid_blocks(std::max<int64_t>(a.norder, 1), 128)), bp(128), gg(grid_blocks(a.n, 256)),
      bg(256);
  const bool kr = c->kr_used;
#define PAIRS(KRV, G) hipLaunchKernelGGL((k_poly_feat<KRV, G>), gp, bp, 0, c->stream, a)
#define GATHER(F, KRV, G) hipLaunchKernelGGL((k_poly_gather<F, KRV, G>), gg, bg, 0, c->stream, a)
  if (kr) {
    if (guarded) PAIRS(true, true); else PAIRS(true, false);
  } else {
    if (guarded) PAIRS(false, true); else PAIRS(false, false);
  }
  if (fused) {
    if (kr) { if (guarded) GATHER(true, true, true); else GATHER(true, true, false); }
    else { if (guarded) GATHER(true, false, true); else GATHER(true, false, false); }
  } else {
    if (kr) { if (guarded) GATHER(false, true, true); else GATHER(false, true, false); }
    else { if (guarded) GATHER(false, false, true); else GATHER(false, false, false); }
  }
#undef PAIRS
#undef GATHER
  XDEM_HIP(c, hipGetLastError());
  c->cs_valid = c->record_cs;
  c->cs_epoch = c->nl_epoch;
  c->poly_pass_epoch = c->nl_epoch;
  return XDEM_OK;
}

}  // namespace xdem

using namespace xdem;

// ------------------------------------------------------------------------------------------ C ABI
extern "C" int xdem_set_shapes(xdem_ctx* c, const xdem_shape* shapes, int32_t n) {
  if (!c || (n > 0 && !shapes)) return XDEM_ERR_ARG;
  if (n < 0 || n > XDEM_MAX_SHAPES) return set_err(c, XDEM_ERR_CAPACITY, "at most 8 shapes");
  for (int k = 0; k < n; k++) {
    const xdem_shape& s = shapes[k];
    if (s.nv < 1 || s.nv > XDEM_MAX_SHAPE_VERTICES || s.ne < 0 || s.ne > XDEM_MAX_SHAPE_EDGES || s.nf < 0 ||
        s.nf > XDEM_MAX_SHAPE_FACES)
      return set_err(c, XDEM_ERR_CAPACITY, "shape: too many vertices, edges or faces");
    for (int f = 0; f < s.nf; f++)
      if (s.face_size[f] < 3 || s.face_size[f] > XDEM_MAX_FACE_VERTICES)
        return set_err(c, XDEM_ERR_CAPACITY, "shape: faces need 3 or 4 vertices");
    for (int e = 0; e < s.ne; e++)
      if (s.edges[e][0] < 0 || s.edges[e][0] >= s.nv || s.edges[e][1] < 0 || s.edges[e][1] >= s.nv)
        return set_err(c, XDEM_ERR_ARG, "shape: edge vertex out of range");
    for (int f = 0; f < s.nf; f++)
      for (int v = 0; v < s.face_size[f]; v++)
        if (s.faces[f][v] < 0 || s.faces[f][v] >= s.nv) return set_err(c, XDEM_ERR_ARG, "shape: face vertex out of range");
  }
  c->nshapes = n;
  for (int k = 0; k < n; k++) c->shapes[k] = shapes[k];
  if (!c->d_shapes) XDEM_HIP(c, hipMalloc((void**)&c->d_shapes, sizeof(xdem_shape) * XDEM_MAX_SHAPES));
  XDEM_HIP(c, hipMemcpyAsync(c->d_shapes, c->shapes, sizeof(xdem_shape) * XDEM_MAX_SHAPES, hipMemcpyHostToDevice,
                             c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  c->nl_valid = false;
  c->pl_valid = false;
  return XDEM_OK;
}

static int poly_ready(xdem_ctx* c) {
  if (c->nshapes == 0) return set_err(c, XDEM_ERR_STATE, "no shapes registered (xdem_set_shapes)");
  if (c->n > 0 && c->max_type >= (uint32_t)c->nshapes)
    return set_err(c, XDEM_ERR_ARG, "a particle type has no registered shape");
  return XDEM_OK;
}

extern "C" int xdem_compute_vertices(xdem_ctx* c) {
  if (!c) return XDEM_ERR_ARG;
  int rc = poly_ready(c);
  if (rc) return rc;
  return poly_compute_vertices(c, false, 0.0, false);
}

extern "C" int xdem_vertex_displ_over(xdem_ctx* c, double threshold, int32_t* result) {
  if (!c || !result) return XDEM_ERR_ARG;
  int rc = poly_ready(c);
  if (rc) return rc;
  *result = 0;
  XDEM_HIP(c, hipMemsetAsync(c->d_flags, 0, sizeof(int32_t), c->stream));
  if ((rc = poly_compute_vertices(c, true, threshold * threshold, false))) return rc;
  XDEM_HIP(c, hipMemcpyAsync(c->h_flags, c->d_flags, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  *result = c->h_flags[0] ? 1 : 0;
  XDEM_HIP(c, hipMemsetAsync(c->d_flags, 0, sizeof(int32_t), c->stream));
  return XDEM_OK;
}

extern "C" int xdem_nbh_polyhedron(xdem_ctx* c, double rcut_inc) {
  if (!c) return XDEM_ERR_ARG;
  int rc = poly_ready(c);
  if (rc) return rc;
  return poly_nbh(c, rcut_inc);
}

extern "C" int xdem_contact_polyhedron(xdem_ctx* c, const xdem_contact_params* cp, const xdem_contact_params* cpd,
                                       double dt) {
  if (!c || !cp) return XDEM_ERR_ARG;
  int rc = poly_ready(c);
  if (rc) return rc;
  if (cp->dncut > 0 || cp->fc != 0 || cp->gamma > 0)
    return set_err(c, XDEM_ERR_UNSUPPORTED, "contact_polyhedron: cohesive and DMT laws are not provided");
  if (c->nwalls > 0 && !cpd)
    return set_err(c, XDEM_ERR_ARG, "contact_polyhedron: config_driver is required when surfaces are registered");
  ContactParams a{cp->kn, cp->kt, cp->kr, cp->mu, cp->damp_rate, cp->fc, cp->dncut, cp->gamma};
  ContactParams b{};
  if (cpd) b = {cpd->kn, cpd->kt, cpd->kr, cpd->mu, cpd->damp_rate, cpd->fc, cpd->dncut, cpd->gamma};
  return poly_force(c, a, b, dt, false, false, V3{0, 0, 0}, false, 0);
}

// The rows of xdem_poly_interactions_dump with, per row, the friction history or (moments) the
// rolling-moment history.
// Host copy of the current mesh candidate lists (slot k of particle x at k * stride + x); h: friction then
// moment history (6 arrays), when asked for. Empty when there is no valid mesh list.
struct HostMeshLists {
  int64_t stride = 0;
  uint32_t cap = 0;
  std::vector<uint32_t> cnt, word;
  std::vector<double> h[6];
  uint32_t count(int64_t x) const { return std::min(cnt[x], cap); }
};
static int read_mesh_lists(xdem_ctx* c, HostMeshLists& m, bool hist) {
  if (!c->has_mesh || !c->ml_valid || c->n == 0) return XDEM_OK;
  if (hist) {
    const int rc = mesh_sync(c);
    if (rc) return rc;
  }
  const MeshList& L = c->ml[c->ml_cur];
  const int64_t n = c->n;
  m.stride = c->mesh.stride;
  m.cap = (uint32_t)c->mesh.cap;
  m.cnt.resize(n);
  XDEM_HIP(c, hipMemcpyAsync(m.cnt.data(), L.cnt, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  uint32_t kmax = 0;
  for (int64_t x = 0; x < n; x++) kmax = std::max(kmax, m.count(x));
  const int64_t tot = (int64_t)kmax * m.stride;
  m.word.resize(tot);
  if (tot == 0) return XDEM_OK;
  XDEM_HIP(c, hipMemcpyAsync(m.word.data(), L.word, sizeof(uint32_t) * tot, hipMemcpyDeviceToHost, c->stream));
  if (hist) {
    double* src[6] = {L.hx, L.hy, L.hz, L.mhx, L.mhy, L.mhz};
    for (int q = 0; q < 6; q++) {
      m.h[q].resize(tot);
      XDEM_HIP(c, hipMemcpyAsync(m.h[q].data(), src[q], sizeof(double) * tot, hipMemcpyDeviceToHost, c->stream));
    }
  }
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  return XDEM_OK;
}

static int poly_interactions_rows(xdem_ctx* c, int64_t cap, uint64_t* id_i, uint64_t* id_j, uint16_t* type,
                                  uint8_t* swap, uint32_t* sub_i, uint32_t* sub_j, double* friction, bool moments,
                                  int64_t* written) {
  if (written) *written = 0;
  if (!c->pl_valid || c->n == 0) return XDEM_OK;
  if (c->pl_epoch != c->nl_epoch) return set_err(c, XDEM_ERR_STATE, "poly dump: run nbh_polyhedron first");
  const int64_t n = c->n;
  const NbrList& L = c->nl[c->nl_cur];
  const PolyList& P = c->pl[c->pl_cur];
  const int64_t nw = (n + kWave - 1) / kWave;
  std::vector<uint32_t> cnt(n), word(L.slots), feat(P.nfeat);
  std::vector<uint64_t> base(nw + 1), ids(n), start(L.slots + 1);
  std::vector<double> hx(P.nfeat), hy(P.nfeat), hz(P.nfeat);
  std::vector<uint32_t> types(n);
  XDEM_HIP(c, hipMemcpyAsync(cnt.data(), L.cnt, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(base.data(), L.base, sizeof(uint64_t) * (nw + 1), hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(ids.data(), live(c).id, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(types.data(), live(c).type, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
  if (L.slots > 0) {
    XDEM_HIP(c, hipMemcpyAsync(word.data(), L.word, sizeof(uint32_t) * L.slots, hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipMemcpyAsync(start.data(), P.start, sizeof(uint64_t) * (L.slots + 1), hipMemcpyDeviceToHost,
                               c->stream));
  }
  if (P.nfeat > 0) {
    XDEM_HIP(c, hipMemcpyAsync(feat.data(), P.feat, sizeof(uint32_t) * P.nfeat, hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipMemcpyAsync(hx.data(), moments ? P.mhx : P.hx, sizeof(double) * P.nfeat, hipMemcpyDeviceToHost,
                               c->stream));
    XDEM_HIP(c, hipMemcpyAsync(hy.data(), moments ? P.mhy : P.hy, sizeof(double) * P.nfeat, hipMemcpyDeviceToHost,
                               c->stream));
    XDEM_HIP(c, hipMemcpyAsync(hz.data(), moments ? P.mhz : P.hz, sizeof(double) * P.nfeat, hipMemcpyDeviceToHost,
                               c->stream));
  }
  std::vector<uint64_t> wmask;
  std::vector<double> whx, why, whz;
  const bool walls = c->nwalls > 0 && c->pw_valid;
  if (walls) {
    const PolyWall& W = c->pw[c->pw_cur];
    const int64_t m = (int64_t)c->nwalls * XDEM_MAX_SHAPE_VERTICES * c->cap;
    wmask.resize(n);
    whx.resize(m); why.resize(m); whz.resize(m);
    XDEM_HIP(c, hipMemcpyAsync(wmask.data(), W.mask, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipMemcpyAsync(whx.data(), moments ? W.mhx : W.hx, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipMemcpyAsync(why.data(), moments ? W.mhy : W.hy, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipMemcpyAsync(whz.data(), moments ? W.mhz : W.hz, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
  }
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  HostMeshLists ml;
  int rc = read_mesh_lists(c, ml, true);
  if (rc) return rc;
  int64_t k = 0;
  for (int64_t x = 0; x < n; x++) {
    if (walls)
      for (int w = 0; w < c->nwalls; w++)
        for (int v = 0; v < c->shapes[types[x]].nv; v++) {
          const int bit = w * XDEM_MAX_SHAPE_VERTICES + v;
          if (!((wmask[x] >> bit) & 1ull)) continue;
          if (k >= cap) return set_err(c, XDEM_ERR_CAPACITY, "poly dump: cap too small");
          const int64_t j = (int64_t)bit * c->cap + x;
          id_i[k] = ids[x]; id_j[k] = (uint64_t)w; type[k] = 5; swap[k] = 0;
          sub_i[k] = (uint32_t)v; sub_j[k] = 0xffffffffu;
          friction[3 * k] = whx[j]; friction[3 * k + 1] = why[j]; friction[3 * k + 2] = whz[j];
          k++;
        }
    if (!ml.cnt.empty())
      for (uint32_t e = 0; e < ml.count(x); e++) {  // mesh candidates (type 7-12, partner = the mesh id)
        if (k >= cap) return set_err(c, XDEM_ERR_CAPACITY, "poly dump: cap too small");
        const int64_t j = (int64_t)e * ml.stride + x;
        const uint32_t wd = ml.word[j];
        id_i[k] = ids[x]; id_j[k] = (uint64_t)c->mesh_id; type[k] = (uint16_t)pm_type(wd); swap[k] = 0;
        sub_i[k] = (uint32_t)pm_sub(wd); sub_j[k] = wd & kPmIdxMask;
        const int q0 = moments ? 3 : 0;
        for (int q = 0; q < 3; q++) friction[3 * k + q] = ml.h[q0 + q][j];
        k++;
      }
    const uint64_t b = base[x >> 6] + (uint64_t)(x & 63);
    for (uint32_t e = 0; e < cnt[x]; e++) {
      const uint64_t s = b + (uint64_t)e * kWave;
      const uint32_t wd = word[s];
      if (!(wd & kOwnerBit)) continue;
      for (uint64_t f = start[s]; f < start[s + 1]; f++) {
        if (k >= cap) return set_err(c, XDEM_ERR_CAPACITY, "poly dump: cap too small");
        const uint32_t fw = feat[f];
        id_i[k] = ids[x];
        id_j[k] = ids[wd & kIdxMask];
        type[k] = (uint16_t)(fw & kFeatTypeMask);
        swap[k] = (fw & kFeatSwap) ? 1 : 0;
        sub_i[k] = (fw >> 3) & 15u;
        sub_j[k] = (fw >> 7) & 15u;
        friction[3 * k] = hx[f]; friction[3 * k + 1] = hy[f]; friction[3 * k + 2] = hz[f];
        k++;
      }
    }
  }
  if (written) *written = k;
  return XDEM_OK;
}

extern "C" int xdem_poly_interactions_dump(xdem_ctx* c, int64_t cap, uint64_t* id_i, uint64_t* id_j, uint16_t* type,
                                           uint8_t* swap, uint32_t* sub_i, uint32_t* sub_j, double* friction,
                                           int64_t* written) {
  if (!c || !id_i || !id_j || !type || !swap || !sub_i || !sub_j || !friction) return XDEM_ERR_ARG;
  return poly_interactions_rows(c, cap, id_i, id_j, type, swap, sub_i, sub_j, friction, false, written);
}

extern "C" int xdem_poly_interactions_moments(xdem_ctx* c, int64_t cap, double* moment, int64_t* written) {
  if (!c || !moment || cap < 0) return XDEM_ERR_ARG;
  const size_t m = (size_t)std::max<int64_t>(cap, 1);
  std::vector<uint64_t> a(m), b(m);
  std::vector<uint16_t> t(m);
  std::vector<uint8_t> sw(m);
  std::vector<uint32_t> si(m), sj(m);
  return poly_interactions_rows(c, cap, a.data(), b.data(), t.data(), sw.data(), si.data(), sj.data(), moment, true,
                                written);
}

extern "C" int xdem_poly_interactions_count(xdem_ctx* c, int64_t* n) {
  if (!c || !n) return XDEM_ERR_ARG;
  *n = 0;
  if (!c->pl_valid) return XDEM_OK;
  if (c->pl_epoch != c->nl_epoch) return set_err(c, XDEM_ERR_STATE, "poly dump: run nbh_polyhedron first");
  int64_t walls = 0;
  if (c->nwalls > 0 && c->pw_valid) {
    std::vector<uint64_t> m(c->n);
    XDEM_HIP(c, hipMemcpyAsync(m.data(), c->pw[c->pw_cur].mask, sizeof(uint64_t) * c->n, hipMemcpyDeviceToHost,
                               c->stream));
    XDEM_HIP(c, hipStreamSynchronize(c->stream));
    for (auto v : m) walls += __builtin_popcountll(v);
  }
  HostMeshLists ml;
  const int rc = read_mesh_lists(c, ml, false);
  if (rc) return rc;
  int64_t mesh = 0;
  for (int64_t x = 0; x < (int64_t)ml.cnt.size(); x++) mesh += ml.count(x);
  *n = c->pl[c->pl_cur].nfeat + walls + mesh;
  return XDEM_OK;
}

// Classifier wave sizes per interaction type (interaction_enum.hpp:24-57): polyhedra 0-3 from the
// classified feature waves, 4-6 vertex-driver candidates by driver kind; spheres 0 the particle pairs
// (owner entries), 4-6 driver candidates by kind, 7-9 mesh candidates by kind.
extern "C" int xdem_classifier_waves(xdem_ctx* c, int64_t sizes[13]) {
  if (!c || !sizes) return XDEM_ERR_ARG;
  for (int t = 0; t < 13; t++) sizes[t] = 0;
  const int64_t n = c->n;
  if (n == 0) return XDEM_OK;
  if (c->nshapes > 0) {
    if (!c->pl_valid || c->pl_class_epoch != c->nl_epoch)
      return set_err(c, XDEM_ERR_STATE, "classifier_waves: run classify_interactions after nbh_polyhedron");
    for (int t = 0; t < 4; t++) sizes[t] = c->fwave_size[t];
    if (c->nwalls > 0 && c->pw_valid) {
      std::vector<uint64_t> m(n);
      XDEM_HIP(c, hipMemcpyAsync(m.data(), c->pw[c->pw_cur].mask, sizeof(uint64_t) * n, hipMemcpyDeviceToHost,
                                 c->stream));
      XDEM_HIP(c, hipStreamSynchronize(c->stream));
      for (int w = 0; w < c->nwalls; w++) {
        const uint64_t sel = 0xffull << (w * XDEM_MAX_SHAPE_VERTICES);
        int64_t k = 0;
        for (auto v : m) k += __builtin_popcountll(v & sel);
        sizes[driver_type(c->walls[w])] += k;
      }
    }
    HostMeshLists ml;
    const int rc = read_mesh_lists(c, ml, false);
    if (rc) return rc;
    for (int64_t x = 0; x < (int64_t)ml.cnt.size(); x++)
      for (uint32_t e = 0; e < ml.count(x); e++) sizes[pm_type(ml.word[(int64_t)e * ml.stride + x])]++;
    return XDEM_OK;
  }
  if (!c->nl_valid || !c->hist_ready)
    return set_err(c, XDEM_ERR_STATE, "classifier_waves: run nbh_sphere and classify_interactions first");
  int rc = nl_pairs(c, &sizes[0]);
  if (rc) return rc;
  if (c->nwalls > 0 && c->wl_valid) {
    std::vector<uint8_t> m(n);
    XDEM_HIP(c, hipMemcpyAsync(m.data(), c->wl[c->wl_cur].mask, n, hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipStreamSynchronize(c->stream));
    for (int w = 0; w < c->nwalls; w++) {
      int64_t k = 0;
      for (auto v : m) k += (v >> w) & 1u;
      sizes[driver_type(c->walls[w])] += k;
    }
  }
  if (c->has_mesh && c->ml_valid && c->mesh.cap > 0) {
    const MeshList& M = c->ml[c->ml_cur];
    std::vector<uint32_t> cnt(n), word((size_t)c->mesh.cap * c->mesh.stride);
    XDEM_HIP(c, hipMemcpyAsync(cnt.data(), M.cnt, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipMemcpyAsync(word.data(), M.word, sizeof(uint32_t) * word.size(), hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipStreamSynchronize(c->stream));
    for (int64_t x = 0; x < n; x++) {
      const uint32_t mc = std::min<uint32_t>(cnt[x], (uint32_t)c->mesh.cap);
      for (uint32_t k = 0; k < mc; k++) sizes[7 + (word[(size_t)k * c->mesh.stride + x] >> 30)]++;
    }
  }
  return XDEM_OK;
}

// The rows of one polyhedron classifier wave (type 0-3) in wave order: owner id, partner id, swap,
// owner and partner sub-features (the key fields of a ClassifierContainer, classifier_container.hpp:87-300)
// and the feature's index in the interaction lists.
extern "C" int xdem_classifier_wave_dump(xdem_ctx* c, int32_t type, int64_t cap, uint64_t* id_i, uint64_t* id_j,
                                         uint8_t* swap, uint32_t* sub_i, uint32_t* sub_j, uint32_t* index,
                                         int64_t* written) {
  if (!c || !id_i || !id_j || !swap || !sub_i || !sub_j || !index) return XDEM_ERR_ARG;
  if (written) *written = 0;
  if (c->nshapes == 0) return set_err(c, XDEM_ERR_UNSUPPORTED, "classifier_wave_dump: polyhedron feature waves only");
  if (type < 0 || type > 3) return set_err(c, XDEM_ERR_ARG, "classifier_wave_dump: type must be 0 (VV) to 3 (EE)");
  if (c->n == 0) return XDEM_OK;
  if (!c->pl_valid || c->pl_class_epoch != c->nl_epoch)
    return set_err(c, XDEM_ERR_STATE, "classifier_wave_dump: run classify_interactions after nbh_polyhedron");
  const int64_t m = c->fwave_size[type];
  if (m > cap) return set_err(c, XDEM_ERR_CAPACITY, "classifier_wave_dump: cap too small");
  if (m == 0) return XDEM_OK;
  const PolyList& P = c->pl[c->pl_cur];
  const NbrList& L = c->nl[c->nl_cur];
  const int64_t total = c->fwave_start[4];
  std::vector<uint4> all(total), ord;
  std::vector<uint32_t> feat(P.nfeat), fslot(P.nfeat), fown(P.nfeat), word(L.slots);
  std::vector<uint64_t> ids(c->n);
  XDEM_HIP(c, hipMemcpyAsync(all.data(), c->forder, sizeof(uint4) * total, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(feat.data(), P.feat, sizeof(uint32_t) * P.nfeat, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(fslot.data(), P.fslot, sizeof(uint32_t) * P.nfeat, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(fown.data(), P.fown, sizeof(uint32_t) * P.nfeat, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(word.data(), L.word, sizeof(uint32_t) * L.slots, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(ids.data(), live(c).id, sizeof(uint64_t) * c->n, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  // the wave of the type: its segments across the windows, in window order; padding only ends segments
  for (int64_t q = 0; q < total; q++) {
    const uint32_t f = all[q].x;
    if (f == 0xffffffffu) {
      if (q % kClassPad == 0) return set_err(c, XDEM_ERR_STATE, "classifier_wave_dump: a segment of padding only");
      continue;
    }
    if (f >= (uint32_t)P.nfeat) return set_err(c, XDEM_ERR_STATE, "classifier_wave_dump: bad feature index");
    if ((int)(feat[f] & kFeatTypeMask) == type) ord.push_back(all[q]);
    const int64_t seg = q - q % kClassPad;  // a wavefront holds one type
    if ((all[seg].w & kFeatTypeMask) != (feat[f] & kFeatTypeMask))
      return set_err(c, XDEM_ERR_STATE, "classifier_wave_dump: a wavefront mixes types");
  }
  if ((int64_t)ord.size() != m) return set_err(c, XDEM_ERR_STATE, "classifier_wave_dump: wave size mismatch");
  for (int64_t k = 0; k < m; k++) {
    const uint32_t f = ord[k].x;
    if (f >= (uint32_t)P.nfeat) return set_err(c, XDEM_ERR_STATE, "classifier_wave_dump: padding inside a wave");
    if (ord[k].y != fown[f] || ord[k].z != (word[fslot[f]] & ~kActiveBit) || ord[k].w != feat[f])
      return set_err(c, XDEM_ERR_STATE, "classifier_wave_dump: a wave entry differs from its feature");
    const uint32_t fw = feat[f];
    id_i[k] = ids[fown[f]];
    id_j[k] = ids[word[fslot[f]] & kIdxMask];
    swap[k] = (fw & kFeatSwap) ? 1 : 0;
    sub_i[k] = (fw >> 3) & 15u;
    sub_j[k] = (fw >> 7) & 15u;
    index[k] = f;
  }
  if (written) *written = m;
  return XDEM_OK;
}

namespace xdem {
// Rows of the recorded polyhedron contact state, as the reference's interaction dump writes them
// (itools/writer.hpp:35-75): features in contact (non-zero force or dn < 0) in feature order, then
// vertex-wall contacts. A feature's first particle is the vertex side (the partner for a swapped
// VE/VF feature, as the reference's pair stores pi), with its sub-feature first.
int poly_contact_state_rows(xdem_ctx* c, std::vector<uint64_t>& ii, std::vector<uint64_t>& jj,
                            std::vector<uint32_t>& si, std::vector<uint32_t>& sj, std::vector<uint16_t>& tt,
                            std::vector<double>& vals) {
  const int64_t n = c->n;
  const PolyList& P = c->pl[c->pl_cur];
  const NbrList& L = c->nl[c->nl_cur];
  const int64_t nf = P.nfeat;
  std::vector<uint32_t> feat(nf), fslot(nf), fown(nf), word(L.slots);
  std::vector<uint64_t> ids(n);
  std::vector<double> cs((size_t)nf * kCStateLen);
  XDEM_HIP(c, hipMemcpyAsync(ids.data(), live(c).id, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, c->stream));
  if (nf > 0) {
    XDEM_HIP(c, hipMemcpyAsync(feat.data(), P.feat, sizeof(uint32_t) * nf, hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipMemcpyAsync(fslot.data(), P.fslot, sizeof(uint32_t) * nf, hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipMemcpyAsync(fown.data(), P.fown, sizeof(uint32_t) * nf, hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipMemcpyAsync(cs.data(), c->cstate, sizeof(double) * kCStateLen * nf, hipMemcpyDeviceToHost,
                               c->stream));
  }
  if (L.slots > 0)
    XDEM_HIP(c, hipMemcpyAsync(word.data(), L.word, sizeof(uint32_t) * L.slots, hipMemcpyDeviceToHost, c->stream));
  const bool walls = c->nwalls > 0 && c->pw_valid && c->wstate;
  std::vector<uint64_t> mask;
  std::vector<double> ws;
  if (walls) {
    mask.resize(n);
    ws.resize((size_t)c->nwalls * XDEM_MAX_SHAPE_VERTICES * c->cap * kCStateLen);
    XDEM_HIP(c, hipMemcpyAsync(mask.data(), c->pw[c->pw_cur].mask, sizeof(uint64_t) * n, hipMemcpyDeviceToHost,
                               c->stream));
    XDEM_HIP(c, hipMemcpyAsync(ws.data(), c->wstate, sizeof(double) * ws.size(), hipMemcpyDeviceToHost, c->stream));
  }
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  auto keep = [](const double* v) {
    const double f =
        std::sqrt(v[4] * v[4] + v[5] * v[5] + v[6] * v[6]) + std::sqrt(v[7] * v[7] + v[8] * v[8] + v[9] * v[9]);
    return f != 0 || v[0] < 0;
  };
  for (int64_t f = 0; f < nf; f++) {
    const double* v = cs.data() + (size_t)kCStateLen * f;
    if (!keep(v)) continue;
    const uint32_t fw = feat[f];
    const bool swap = (fw & kFeatSwap) != 0;
    const uint64_t own = ids[fown[f]], par = ids[word[fslot[f]] & kIdxMask];
    const uint32_t osub = (fw >> 3) & 15u, psub = (fw >> 7) & 15u;
    ii.push_back(swap ? par : own);
    jj.push_back(swap ? own : par);
    si.push_back(swap ? psub : osub);
    sj.push_back(swap ? osub : psub);
    tt.push_back((uint16_t)(fw & kFeatTypeMask));
    vals.insert(vals.end(), v, v + kCStateLen);
  }
  if (walls)
    for (int64_t x = 0; x < n; x++)
      for (int bit = 0; bit < c->nwalls * XDEM_MAX_SHAPE_VERTICES; bit++) {
        if (!((mask[x] >> bit) & 1ull)) continue;
        const double* v = ws.data() + (size_t)kCStateLen * ((int64_t)bit * c->cap + x);
        if (!keep(v)) continue;
        ii.push_back(ids[x]);
        jj.push_back((uint64_t)(bit / XDEM_MAX_SHAPE_VERTICES));
        si.push_back((uint32_t)(bit % XDEM_MAX_SHAPE_VERTICES));
        sj.push_back(0xffffffffu);  // the driver side's sub-feature is -1 (nbh_polyhedron_driver.hpp:220)
        tt.push_back(driver_type(c->walls[bit / XDEM_MAX_SHAPE_VERTICES]));
        vals.insert(vals.end(), v, v + kCStateLen);
      }
  if (c->mstate) {  // mesh candidates: partner = the mesh id, sub_j = the mesh feature
    HostMeshLists ml;
    const int rc = read_mesh_lists(c, ml, false);
    if (rc) return rc;
    if (!ml.word.empty()) {
      std::vector<double> ms(ml.word.size() * kCStateLen);
      XDEM_HIP(c, hipMemcpyAsync(ms.data(), c->mstate, sizeof(double) * ms.size(), hipMemcpyDeviceToHost, c->stream));
      XDEM_HIP(c, hipStreamSynchronize(c->stream));
      for (int64_t x = 0; x < n; x++)
        for (uint32_t e = 0; e < ml.count(x); e++) {
          const int64_t j = (int64_t)e * ml.stride + x;
          const double* v = ms.data() + (size_t)kCStateLen * j;
          if (!keep(v)) continue;
          const uint32_t wd = ml.word[j];
          ii.push_back(ids[x]);
          jj.push_back((uint64_t)c->mesh_id);
          si.push_back((uint32_t)pm_sub(wd));
          sj.push_back(wd & kPmIdxMask);
          tt.push_back((uint16_t)pm_type(wd));
          vals.insert(vals.end(), v, v + kCStateLen);
        }
    }
  }
  return XDEM_OK;
}
}  // namespace xdem

// Restart of polyhedra with interaction history: saved rows (the layout of xdem_poly_interactions_dump)
// set the history of the current feature and vertex-wall interactions with the same key.
namespace {
struct PKey {
  uint64_t i, j;
  uint32_t si, sj;
  uint16_t t;
  uint8_t sw;
  bool operator==(const PKey& o) const {
    return i == o.i && j == o.j && si == o.si && sj == o.sj && t == o.t && sw == o.sw;
  }
};
struct PKeyHash {
  size_t operator()(const PKey& k) const {
    uint64_t h = k.i * 0x9e3779b97f4a7c15ull;
    h ^= k.j + 0x7f4a7c159e3779b9ull + (h << 6) + (h >> 2);
    h ^= ((uint64_t)k.si << 32 | k.sj) + (h << 6) + (h >> 2);
    h ^= ((uint64_t)k.t << 8 | k.sw) + (h << 6) + (h >> 2);
    return (size_t)h;
  }
};
}  // namespace

extern "C" int xdem_poly_interactions_load(xdem_ctx* c, int64_t m, const uint64_t* id_i, const uint64_t* id_j,
                                           const uint16_t* type, const uint8_t* swap, const uint32_t* sub_i,
                                           const uint32_t* sub_j, const double* friction, const double* moment,
                                           int64_t* matched) {
  if (!c || m < 0 || (m > 0 && (!id_i || !id_j || !type || !swap || !sub_i || !sub_j || !friction)))
    return XDEM_ERR_ARG;
  if (matched) *matched = 0;
  if (!c->pl_valid || c->pl_epoch != c->nl_epoch)
    return set_err(c, XDEM_ERR_STATE, "poly_interactions_load: run nbh_polyhedron first");
  const int64_t n = c->n;
  if (n == 0) return XDEM_OK;
  std::unordered_map<PKey, int64_t, PKeyHash> saved;
  saved.reserve((size_t)m * 2);
  for (int64_t k = 0; k < m; k++)
    if (!saved.emplace(PKey{id_i[k], id_j[k], sub_i[k], sub_j[k], type[k], swap[k]}, k).second)
      return set_err(c, XDEM_ERR_ARG, "poly_interactions_load: duplicate interaction key");
  PolyList& P = c->pl[c->pl_cur];
  const NbrList& L = c->nl[c->nl_cur];
  const int64_t nf = P.nfeat;
  std::vector<uint32_t> feat(nf), fslot(nf), fown(nf), word(L.slots);
  std::vector<uint64_t> ids(n);
  XDEM_HIP(c, hipMemcpyAsync(ids.data(), live(c).id, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, c->stream));
  if (nf > 0) {
    XDEM_HIP(c, hipMemcpyAsync(feat.data(), P.feat, sizeof(uint32_t) * nf, hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipMemcpyAsync(fslot.data(), P.fslot, sizeof(uint32_t) * nf, hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipMemcpyAsync(fown.data(), P.fown, sizeof(uint32_t) * nf, hipMemcpyDeviceToHost, c->stream));
  }
  if (L.slots > 0)
    XDEM_HIP(c, hipMemcpyAsync(word.data(), L.word, sizeof(uint32_t) * L.slots, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<double> h[6];
  for (auto& a : h) a.assign((size_t)nf, 0.0);
  int64_t hit = 0;
  for (int64_t f = 0; f < nf; f++) {
    const uint32_t fw = feat[f];
    const PKey key{ids[fown[f]], ids[word[fslot[f]] & kIdxMask], (fw >> 3) & 15u, (fw >> 7) & 15u,
                   (uint16_t)(fw & kFeatTypeMask), (uint8_t)((fw & kFeatSwap) ? 1 : 0)};
    const auto it = saved.find(key);
    bool act = false;
    if (it != saved.end()) {
      for (int q = 0; q < 3; q++) {
        h[q][f] = friction[3 * it->second + q];
        h[3 + q][f] = moment ? moment[3 * it->second + q] : 0.0;
        act = act || h[q][f] != 0.0 || h[3 + q][f] != 0.0;
      }
      hit++;
    }
    feat[f] = act ? (fw | kFeatActive) : (fw & ~kFeatActive);
  }
  if (nf > 0) {
    double* dst[6] = {P.hx, P.hy, P.hz, P.mhx, P.mhy, P.mhz};
    for (int q = 0; q < 6; q++)
      XDEM_HIP(c, hipMemcpyAsync(dst[q], h[q].data(), sizeof(double) * nf, hipMemcpyHostToDevice, c->stream));
    XDEM_HIP(c, hipMemcpyAsync(P.feat, feat.data(), sizeof(uint32_t) * nf, hipMemcpyHostToDevice, c->stream));
  }
  if (c->nwalls > 0 && c->pw_valid) {
    PolyWall& W = c->pw[c->pw_cur];
    const int64_t mw = (int64_t)c->nwalls * XDEM_MAX_SHAPE_VERTICES * c->cap;
    std::vector<uint64_t> mask(n);
    XDEM_HIP(c, hipMemcpyAsync(mask.data(), W.mask, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<double> wh[6];
    for (auto& a : wh) a.assign((size_t)mw, 0.0);
    for (int64_t x = 0; x < n; x++)
      for (int bit = 0; bit < c->nwalls * XDEM_MAX_SHAPE_VERTICES; bit++) {
        if (!((mask[x] >> bit) & 1ull)) continue;
        const auto it = saved.find(PKey{ids[x], (uint64_t)(bit / XDEM_MAX_SHAPE_VERTICES),
                                        (uint32_t)(bit % XDEM_MAX_SHAPE_VERTICES), 0xffffffffu, 5, 0});
        if (it == saved.end()) continue;
        const int64_t j = (int64_t)bit * c->cap + x;
        for (int q = 0; q < 3; q++) {
          wh[q][j] = friction[3 * it->second + q];
          wh[3 + q][j] = moment ? moment[3 * it->second + q] : 0.0;
        }
        hit++;
      }
    double* dst[6] = {W.hx, W.hy, W.hz, W.mhx, W.mhy, W.mhz};
    for (int q = 0; q < 6; q++)
      XDEM_HIP(c, hipMemcpyAsync(dst[q], wh[q].data(), sizeof(double) * mw, hipMemcpyHostToDevice, c->stream));
  }
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  {
    HostMeshLists ml;
    const int rc = read_mesh_lists(c, ml, true);
    if (rc) return rc;
    if (!ml.word.empty()) {
      for (int64_t x = 0; x < n; x++)
        for (uint32_t e = 0; e < ml.count(x); e++) {
          const int64_t j = (int64_t)e * ml.stride + x;
          const uint32_t wd = ml.word[j];
          const auto it = saved.find(PKey{ids[x], (uint64_t)c->mesh_id, (uint32_t)pm_sub(wd), wd & kPmIdxMask,
                                          (uint16_t)pm_type(wd), 0});
          for (int q = 0; q < 6; q++) ml.h[q][j] = 0.0;
          if (it == saved.end()) continue;
          for (int q = 0; q < 3; q++) {
            ml.h[q][j] = friction[3 * it->second + q];
            ml.h[3 + q][j] = moment ? moment[3 * it->second + q] : 0.0;
          }
          hit++;
        }
      const MeshList& L = c->ml[c->ml_cur];
      double* dst[6] = {L.hx, L.hy, L.hz, L.mhx, L.mhy, L.mhz};
      for (int q = 0; q < 6; q++)
        XDEM_HIP(c, hipMemcpyAsync(dst[q], ml.h[q].data(), sizeof(double) * ml.h[q].size(), hipMemcpyHostToDevice,
                                   c->stream));
      XDEM_HIP(c, hipStreamSynchronize(c->stream));
      c->mc_valid = false;  // the passes gather the loaded histories again
    }
  }
  if (moment)
    for (int64_t k = 0; k < 3 * m && !c->kr_used; k++) c->kr_used = moment[k] != 0.0;
  if (matched) *matched = hit;
  // the classifier waves carry the feature words: refresh them with the loaded active bits
  if (c->pl_class_epoch == c->nl_epoch) return poly_classify(c);
  return XDEM_OK;
}

namespace xdem {
// simulation_state counts of polyhedra (itools/counter.hpp:140-153): features of the lists (particle
// pairs, counted twice under symetric: true) and vertex-driver candidates (once); active: in contact at
// the last pass or with history; the smallest dn over the active ones (0 when none overlaps).
__device__ inline void block_reduce3(double tot, double act, double mn, double* part) {
  __shared__ double r[3][256];
  r[0][threadIdx.x] = tot;
  r[1][threadIdx.x] = act;
  r[2][threadIdx.x] = mn;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      r[0][threadIdx.x] += r[0][threadIdx.x + st];
      r[1][threadIdx.x] += r[1][threadIdx.x + st];
      r[2][threadIdx.x] = fmin(r[2][threadIdx.x], r[2][threadIdx.x + st]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 3) part[3 * (int64_t)blockIdx.x + threadIdx.x] = r[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_poly_sim_feat(int64_t nf, const uint32_t* __restrict__ feat,
                                                       const uint8_t* __restrict__ contact,
                                                       const double* __restrict__ fdn, double* __restrict__ part) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double tot = 0, act = 0, mn = 0;
  if (f < nf) {
    tot = 2.0;
    const bool in = contact[f] != 0;
    if (in || (feat[f] & kFeatActive)) {
      act = 2.0;
      mn = in && fdn[f] < 0.0 ? fdn[f] : 0.0;
    }
  }
  block_reduce3(tot, act, mn, part);
}

// vertex-driver candidates, then mesh candidates (once each, the driver rule of counter.hpp:145-151)
__global__ __launch_bounds__(256) void k_poly_sim_wall(int64_t n, PState s, const xdem_shape* __restrict__ shapes,
                                                       const d4* __restrict__ vtx, PolyWall pw,
                                                       const Surface* __restrict__ walls, int nw, int64_t stride,
                                                       MeshView m, MeshList ml, double* __restrict__ part) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double tot = 0, act = 0, mn = 0;
  if (x < n && m.cap > 0) {
    const xdem_shape& sa = shapes[s.type[x]];
    const double ra = sa.minkowski * s.core[x].w.w;
    const Verts va = verts_of(vtx, x, V3{0, 0, 0});
    const uint32_t mc = min(ml.cnt[x], (uint32_t)m.cap);
    for (uint32_t k = 0; k < mc; k++) {
      const int64_t j = (int64_t)k * m.stride + x;
      const Det d = poly_mesh_detect(m, ml.word[j], sa, va, ra);
      tot += 1.0;
      if (d.contact || ml.hx[j] != 0.0 || ml.hy[j] != 0.0 || ml.hz[j] != 0.0) {
        act += 1.0;
        mn = fmin(mn, d.contact ? d.dn : 0.0);
      }
    }
  }
  if (x < n && nw > 0) {
    const xdem_shape& sh = shapes[s.type[x]];
    const double r = sh.minkowski * s.core[x].w.w;
    const uint64_t mask = pw.mask[x];
    for (int w = 0; w < nw; w++)
      for (int k = 0; k < sh.nv; k++) {
        const int bit = w * XDEM_MAX_SHAPE_VERTICES + k;
        if (!((mask >> bit) & 1ull)) continue;
        const d4 v4 = vtx[x * XDEM_MAX_SHAPE_VERTICES + k];
        double dn;
        V3 nrm, cpos;
        const bool in = driver_detect(walls[w], V3{v4.x, v4.y, v4.z}, r, dn, nrm, cpos);
        const int64_t j = (int64_t)bit * stride + x;
        tot += 1.0;
        if (in || pw.hx[j] != 0.0 || pw.hy[j] != 0.0 || pw.hz[j] != 0.0) {
          act += 1.0;
          mn = fmin(mn, in ? dn : 0.0);
        }
      }
  }
  block_reduce3(tot, act, mn, part);
}

int poly_sim_counts(xdem_ctx* c, double* tot, double* act, double* min_dn, double* fixed, int64_t fixed_cap) {
  *tot = *act = *min_dn = 0.0;
  if (!c->pl_valid || c->pl_epoch != c->nl_epoch || c->poly_pass_epoch != c->nl_epoch)
    return set_err(c, XDEM_ERR_STATE, "simulation_state: run contact_polyhedron after nbh_polyhedron first");
  const PolyList& P = c->pl[c->pl_cur];
  const int64_t nf = P.nfeat, n = c->n;
  const int bf = grid_blocks(std::max<int64_t>(nf, 1), 256), bw = grid_blocks(std::max<int64_t>(n, 1), 256);
  const int64_t need = 3 * (int64_t)(bf + bw);
  double* part = fixed;
  if (fixed) {
    if (need + 3 > fixed_cap)
      return set_err(c, XDEM_ERR_CAPACITY, "simulation_state_noalloc: ceil(features / 256) + ceil(particles / 256) "
                                           "exceeds " + std::to_string((fixed_cap - 3) / 3) + " blocks");
  } else {
    if (need > c->cap_simpart) {
      int rc = dev_realloc(c, &c->simpart, need);
      if (rc) return rc;
      c->cap_simpart = need;
    }
    part = c->simpart;
  }
  const int nw = (c->nwalls > 0 && c->pw_valid) ? c->nwalls : 0;
  hipLaunchKernelGGL(k_poly_sim_feat, dim3(bf), dim3(256), 0, c->stream, nf, P.feat, c->fcontact, c->fdn, part);
  PolyWall pw{};
  if (nw) pw = c->pw[c->pw_cur];
  MeshView mv{};
  MeshList ml{};
  if (c->has_mesh && c->ml_valid) {
    int rc = mesh_sync(c);
    if (rc) return rc;
    mv = c->mesh;
    ml = c->ml[c->ml_cur];
  }
  hipLaunchKernelGGL(k_poly_sim_wall, dim3(bw), dim3(256), 0, c->stream, n, live(c), c->d_shapes, c->vtx, pw,
                     c->d_walls, nw, c->cap, mv, ml, part + 3 * (int64_t)bf);
  XDEM_HIP(c, hipGetLastError());
  if (fixed) {
    double res[3];
    int rc = fold_partials(c, part, bf + bw, 3, part + need);
    if (rc) return rc;
    XDEM_HIP(c, hipMemcpyAsync(res, part + need, sizeof(res), hipMemcpyDeviceToHost, c->stream));
    XDEM_HIP(c, hipStreamSynchronize(c->stream));
    *tot = res[0];
    *act = res[1];
    *min_dn = res[2];
    return XDEM_OK;
  }
  std::vector<double> h((size_t)need);
  XDEM_HIP(c, hipMemcpyAsync(h.data(), part, sizeof(double) * need, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < bf + bw; k++) {
    *tot += h[3 * (size_t)k];
    *act += h[3 * (size_t)k + 1];
    *min_dn = std::min(*min_dn, h[3 * (size_t)k + 2]);
  }
  return XDEM_OK;
}
}  // namespace xdem
