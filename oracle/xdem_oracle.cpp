// xdem_oracle.cpp — CPU oracle for exaDEM's per-timestep sphere path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker the HIP path is compared
// against; it is loaded by tests/, by __graft_entry__.smoke() and by the
// cpu_baseline leg of bench.py, and by nothing in the product.
//
// It restates, in plain FP64 C++ (OpenMP over particles), the reference
// algorithm for one time step (paths relative to the exaDEM source tree):
//   integration      src/numericalscheme/combined_compute_prolog.cu:35-47,94-116
//                    src/numericalscheme/combined_compute_epilog.cu:35-46,94-113
//                    src/numericalscheme/include/exaDEM/{push_to_quaternion,angular_acceleration,
//                      angular_velocity,force_to_accel}.hpp
//                    quaternion helpers: src/shape/include/exaDEM/type/quat.hpp:76-106
//   gravity / reset  src/forcefield/include/exaDEM/forcefield/gravity_force.hpp:38-46,
//                    src/numericalscheme/reset_force_moment.cu:69-70
//   neighbours       src/grid/chunk_neighbors_contact.cpp:45-57 (filter), periodic images as
//                    exaNBody ghost cells (one layer, positions r + shift)
//   interactions     src/sphere/nbh_sphere.cu:64-309 (sym mode), interaction_manager.hpp:88-112,205-217,
//                    placeholder_interaction.hpp:398-408, interaction_pair.hpp:186-249
//   contact law      src/forcefield/include/exaDEM/forcefield/common_kernels.hpp:27-145,
//                    contact_force.hpp:146-235, src/sphere/include/exaDEM/sphere/contact_sphere.hpp:121-297,
//                    src/shape/include/exaDEM/shape_detection_core.hpp:57-81,
//                    src/driver/include/exaDEM/surface.hpp:320-350, cylinder.hpp:225-272, ball.hpp:292-327
//   polyhedra        src/polyhedron/nbh_polyhedron.cpp:75-384, polyhedron/contact.hpp:124-432,
//                    polyhedron/include/exaDEM/polyhedron/nbh_polyhedron_driver.hpp:60-200 (mesh driver),
//                    src/shape/include/exaDEM/shape_detection_core.hpp:43-566, shape_detection_api.hpp:38-406
//   operator order   data/config/main-config.msp:55-65 (first_iteration),
//                    data/config/config_numerical_schemes.msp:6-28, config_move_particles.msp:141-195
//
// Choices where the reference leaves the result open (exaNBody/onika are not in the tree):
//   * order of particles inside a cell: ascending id;
//   * neighbour stream order: neighbour cells dz,dy,dx in -1..1 (dx fastest), particles in cell order;
//   * force/moment summation order (the reference uses atomics, so any order is a valid result):
//     per particle, gravity, then particle-particle contributions in its stream order, then walls
//     in driver index order;
//   * PushVec3SecondOrder: r += v*dt + f*dt^2/2; PushVec3FirstOrder: v += f*dt/2;
//   * periodic wrap at rebuild: one shift by +-L per axis; displacement test |r - r_backup|^2 >= max_displ^2.
// Build: see Makefile in this directory (g++ -O2 -ffp-contract=off -fopenmp).

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../include/xdem.h"

namespace orc {

struct V3 {
  double x, y, z;
};
static inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
static inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
static inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
static inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
static inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
static inline double norm(V3 a) { return std::sqrt(dot(a, a)); }

struct Q4 {
  double w, x, y, z;
};

// Mutation switch used by the oracle's own self-tests: a non-zero value plants one known
// mistake so that the closed-form checks can show they catch it. Always 0 in normal use.
//   1 damping sign, 2 friction increment sign, 3 no Coulomb cap, 4 history dropped at rebuild,
//   5 polyhedron swap side ignored in the force assembly, 6 the vertex side's moment arm taken from
//   the vertex instead of the contact point, 7 polyhedron meff = m_i, 8 shaker evaluated at the end of
//   the step, 9 the friction negation of mesh-vertex items (11, 12) dropped, 10 the mesh-vertex box gate
//   of those items dropped.
static int g_mutation = 0;

// ---------------------------------------------------------------- quaternion helpers
// quat::dot(omega): time derivative 0.5*(0,omega)*q   (type/quat.hpp:92-96)
static inline Q4 qdot(Q4 q, V3 w) {
  return {-0.5 * (w.x * q.x + w.y * q.y + w.z * q.z), 0.5 * (w.y * q.z - w.z * q.y + w.x * q.w),
          0.5 * (w.z * q.x - w.x * q.z + w.y * q.w), 0.5 * (w.x * q.y - w.y * q.x + w.z * q.w)};
}
static inline Q4 qnormalize(Q4 q) {
  const double n = std::sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
  return {q.w / n, q.x / n, q.y / n, q.z / n};
}
// quat * vec3 rotation (type/quat.hpp:84-91)
static inline V3 qrot(Q4 q, V3 v) {
  V3 qv{q.x, q.y, q.z};
  V3 uv = cross(qv, v);
  V3 uuv = cross(qv, uv);
  uv = uv * (2.0 * q.w);
  uuv = uuv * 2.0;
  return v + uv + uuv;
}
static inline Q4 qconj(Q4 q) { return {q.w, -q.x, -q.y, -q.z}; }

// ---------------------------------------------------------------- contact law
struct CP {
  double kn, kt, kr, mu, damp_rate, fc, dncut, gamma;
};

// detection_vertex_vertex_core (shape_detection_core.hpp:57-81)
struct Contact {
  bool contact;
  double dn;
  V3 n, cp;
};
static inline Contact detect_vv(V3 pi, double ri, V3 pj, double rj) {
  const double R = ri + rj;
  const V3 dist = pi - pj;
  const double dist_norm = norm(dist);
  const double inv = 1.0 / dist_norm;
  const double dn = dist_norm - R;
  const V3 n = dist * inv;
  const V3 cp = pi - n * (ri + 0.5 * dn);
  return {dn < 0, dn, n, cp};
}

// contact_force_core<Hooke, None> (contact_force.hpp:146-235) with common_kernels.hpp:27-145
static inline void contact_force_core(double dn, V3 n, double dt, const CP& p, double meff, V3& ft, V3 cp,
                                      V3 pos_i, V3 vel_i, V3& f_i, V3& mom_i, V3 vrot_i, V3 pos_j, V3 vel_j,
                                      V3 vrot_j) {
  const double damp = p.damp_rate * 2.0 * std::sqrt(p.kn * meff);
  const V3 ci = vel_i - cross(cp - pos_i, vrot_i);
  const V3 cj = vel_j - cross(cp - pos_j, vrot_j);
  const V3 vel = cj - ci;
  const double vn = dot(vel, n);
  // normal_force
  const double fne = -p.kn * dn;
  const double fnv = (g_mutation == 1 ? -damp : damp) * vn;
  const double fnval = (fnv + fne > 0) ? fnv + fne : 0;
  const V3 fn = fnval * n;
  // tangential force: kt * (dt * (vel - vn*n))
  const V3 vt = vel - vn * n;
  const V3 inc = (g_mutation == 2 ? -p.kt : p.kt) * (dt * vt);
  ft = ft + inc;
  // compute_threshold_ft + fit_tangential_force
  const double thr = std::fabs(p.mu * (p.kn * dn));
  const double ft2 = dot(ft, ft);
  if (g_mutation != 3 && ft2 > 0.0 && ft2 > thr * thr) ft = ft * (thr / std::sqrt(ft2));
  f_i = fn + ft;
  // rolling moment history and its clip (mur = 0)
  mom_i = mom_i + (p.kr * (vrot_j - vrot_i)) * dt;
  V3 branch = cp - pos_i;
  const double r = dot(branch, vrot_i) / dot(vrot_i, vrot_i);
  branch = branch - r * vrot_i;
  const double mur = 0;
  const double thr_m = std::fabs(mur * norm(branch) * norm(fn));
  const double m2 = dot(mom_i, mom_i);
  if (m2 > 0.0 && m2 > thr_m * thr_m) mom_i = mom_i * (thr_m / std::sqrt(m2));
}

// compute_moments (common_kernels.hpp:134-142)
static inline V3 compute_moments(V3 cp, V3 p, V3 f, V3 m) { return cross(cp - p, f) + m; }

// Stationary drivers: Surface (surface.hpp:320-350), Cylinder (cylinder.hpp:225-272), Ball
// (ball.hpp:292-327). kind: 0 surface, 1 cylinder, 2 ball.
// Kinematic motion of a driver (Driver_params, driver_params.hpp:33-60; MotionType values of
// motion_type.hpp:23-35): 0 STATIONARY, 1 LINEAR_MOTION, 7 SHAKER.
struct Motion {
  int type = 0;
  V3 mvec{0, 0, 0};
  double const_vel = 0, t_start = 0, t_end = 1e300, amplitude = 0, omega = 0;
  V3 shaker_dir{0, 0, 1};
};
struct Surf {
  double offset;
  V3 normal, center, vel, vrot, center_proj;
  int kind;
  double radius;
  V3 axis;
  Motion mo;
};

// is_motion_triggered (driver_params.hpp:84-86)
static inline bool motion_on(const Motion& m, double time) { return time >= m.t_start && time <= m.t_end; }
// shaker_signal / shaker_velocity (driver_params.hpp:346-356)
static inline double shaker_signal(const Motion& m, double time) {
  time -= m.t_start;
  return m.amplitude * std::sin(m.omega * time);
}
static inline V3 shaker_velocity(const Motion& m, double time) {
  time -= m.t_start;
  return m.amplitude * m.omega * std::cos(m.omega * time) * m.shaker_dir;
}
// push_f_v_r of one driver: Surface (surface.hpp:260-299; acc = 0 for kinematic motions), Cylinder
// (cylinder.hpp:160-173), Ball (ball.hpp:203-213)
static void driver_push_f_v_r(Surf& s, double time, double dt) {
  const Motion& m = s.mo;
  if (m.type == 0) return;
  if (s.kind == 0) {
    const double acc = 0.0;  // kinematic motions carry no acceleration
    double displ = dt * dot(s.vel, s.normal) + 0.5 * dt * dt * acc;
    if (m.type == 7) {
      const double next = shaker_signal(m, time + dt), cur = shaker_signal(m, time);
      displ = (next - cur) * dot(m.shaker_dir, s.normal);
      s.vel = shaker_velocity(m, time + dt);
    }
    s.center = s.center + displ * s.normal;
    s.offset += displ;
    s.center_proj = s.center_proj + displ * s.normal;
  } else if (s.kind == 1) {
    if (m.type == 7) {
      const double displ = shaker_signal(m, time + dt) - shaker_signal(m, time);
      s.center = s.center + displ * m.shaker_dir;
      s.vel = shaker_velocity(m, time + dt);
    }
  } else {
    s.center = s.center + dt * s.vel;
  }
}
// push_f_v of one driver (surface.hpp:234-252, ball.hpp:219-234; Cylinder: nothing, cylinder.hpp:149-151)
static void driver_push_f_v(Surf& s) {
  const Motion& m = s.mo;
  if (s.kind == 0) {
    if (m.type == 0) s.vel = {0, 0, 0};
    else if (m.type == 1) s.vel = m.const_vel * m.mvec;
  } else if (s.kind == 2) {
    if (m.type == 1) s.vel = m.mvec * m.const_vel;
  }
}
// driver_displ_over against the drivers at the last rebuild (driver_displ_over.cu:59-130): projected
// centre of a surface, centre of a ball; a cylinder never triggers.
static bool drivers_displ_over(const std::vector<Surf>& now, const std::vector<Surf>& bk, double thr) {
  const double r2 = thr * thr;
  for (size_t w = 0; w < now.size() && w < bk.size(); w++) {
    V3 d{0, 0, 0};
    if (now[w].kind == 0) d = now[w].center_proj - bk[w].center_proj;
    else if (now[w].kind == 2) d = now[w].center - bk[w].center;
    else continue;
    if (dot(d, d) >= r2) return true;
  }
  return false;
}
static inline V3 mask_mul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
static inline bool surf_filter(const Surf& s, double rcut, V3 p) {
  if (s.kind == 1) {  // Cylinder::filter
    const V3 proj = mask_mul(p, s.axis);
    const V3 dir = proj - s.center;
    const double d = norm(dir);
    const double dn = s.radius - (d + rcut);
    return dn <= 0;
  }
  if (s.kind == 2) {  // Ball::filter
    const V3 dist = s.center - p;
    const double d = s.radius - norm(dist);
    return std::fabs(d) <= rcut;
  }
  const V3 proj = dot(p, s.normal) * s.normal;
  const double d = norm(proj - s.center_proj);
  return d <= rcut;
}
static inline Contact surf_detect(const Surf& s, double rcut, V3 p) {
  if (s.kind == 1) {  // Cylinder::detector
    const V3 proj = mask_mul(p, s.axis);
    const V3 dir = s.center - proj;
    const double d = norm(dir);
    const double dn = s.radius - (rcut + d);
    if (dn > 0) return {false, dn, {0, 0, 0}, {0, 0, 0}};
    const V3 n = dir / d;
    return {true, dn, n, p - n * (rcut + 0.5 * dn)};
  }
  if (s.kind == 2) {  // Ball::detector
    const V3 ptc = s.center - p;
    const double d = norm(ptc);
    V3 n = ptc / d;
    double dn;
    if (d > s.radius) {
      dn = d - s.radius - rcut;
      n = (-1.0) * n;
    } else {
      dn = s.radius - d - rcut;
    }
    if (dn > 0) return {false, dn, {0, 0, 0}, {0, 0, 0}};
    return {true, dn, n, p - n * (rcut + 0.5 * dn)};
  }
  const V3 proj = dot(p, s.normal) * s.normal;
  const V3 s2p = -(s.center_proj - proj);
  const double d = norm(s2p);
  const double dn = d - rcut;
  if (dn > 0) return {false, dn, {0, 0, 0}, {0, 0, 0}};
  const V3 n = s2p / d;
  const V3 cp = p - n * (rcut + 0.5 * dn);
  return {true, dn, n, cp};
}

// ---------------------------------------------------------------- interactions
// InteractionPair key (interaction_pair.hpp:53-62,186-249) plus payload (interaction.hpp:33-35)
struct Item {
  uint64_t id_i, id_j;  // owner id, partner id (driver index for walls)
  uint32_t sub_i, sub_j;
  uint16_t type;  // 0 VertexVertex, 5 VertexSurface
  uint8_t swap, ghost;
  int32_t partner;  // sorted index of the partner particle (-1 for walls)
  int32_t shift;    // periodic image code of the partner, 13 = none
  V3 friction, moment;
  // results of the last contact evaluation (for the canonical accumulation)
  bool in_contact;
  V3 f, cp, mint;
};
static inline bool key_less(const Item& a, const Item& b) {  // operator< (swap = 0 throughout)
  if (a.id_i != b.id_i) return a.id_i < b.id_i;
  if (a.id_j != b.id_j) return a.id_j < b.id_j;
  if (a.type != b.type) return a.type < b.type;
  if (a.swap != b.swap) return a.swap < b.swap;
  if (a.sub_i != b.sub_i) return a.sub_i < b.sub_i;
  return a.sub_j < b.sub_j;
}
static inline bool key_eq(const Item& a, const Item& b) {
  return a.id_i == b.id_i && a.id_j == b.id_j && a.sub_i == b.sub_i && a.sub_j == b.sub_j && a.swap == b.swap &&
         a.type == b.type;
}
static inline bool item_active(const Item& it) {
  const bool zf = it.friction.x == 0 && it.friction.y == 0 && it.friction.z == 0;
  const bool zm = it.moment.x == 0 && it.moment.y == 0 && it.moment.z == 0;
  return !(zf && zm);
}

struct Stream {  // one neighbour of a particle in its stream
  int32_t j;
  int32_t shift;
};

struct Oracle {
  int64_t n = 0;
  // particle arrays (current, cell-sorted order)
  std::vector<uint64_t> id;
  std::vector<uint32_t> type, group;
  std::vector<V3> r, v, f, mom, vrot, arot, inertia, backup;
  std::vector<double> mass, radius, homothety;
  std::vector<Q4> q;
  // domain / grid
  double lo[3] = {0, 0, 0}, L[3] = {1, 1, 1}, h[3] = {1, 1, 1};
  int per[3] = {0, 0, 0};
  int dims[3] = {1, 1, 1};
  std::vector<int64_t> cell_start;  // size ncells+1
  std::vector<Surf> walls;
  std::vector<Surf> walls_bk;  // drivers at the last rebuild (backup_drivers)
  double time = 0;             // physical_time
  // stationary triangle-mesh driver (register_stl_mesh.cpp:30-140), world-space features
  struct Mesh {
    bool on = false;
    uint64_t id = 0;
    double rm = 0;
    V3 center{0, 0, 0};
    std::vector<V3> v;
    std::vector<std::array<uint32_t, 2>> e;
    std::vector<std::array<uint32_t, 3>> f;
  } mesh;
  // neighbour streams and interaction lists per particle
  std::vector<std::vector<Stream>> streams;
  std::vector<std::vector<Item>> items;  // per owner particle: walls first, then particle pairs
  bool built = false;
  int threads = 0;
  std::string err;
  // polyhedra
  bool poly = false;
  std::vector<xdem_shape> shapes;
  std::vector<std::array<V3, 8>> verts;  // world vertices per particle
  std::vector<Q4> backup_q;              // orientation at the last rebuild (vertex_displ_over)
};

static inline V3 shift_vec(const Oracle& o, int code) {
  const int sx = code % 3 - 1, sy = (code / 3) % 3 - 1, sz = code / 9 - 1;
  return {sx * o.L[0], sy * o.L[1], sz * o.L[2]};
}

static void set_domain(Oracle& o, const xdem_domain* d) {
  for (int a = 0; a < 3; a++) {
    o.lo[a] = d->lo[a];
    o.L[a] = d->hi[a] - d->lo[a];
    o.per[a] = d->periodic[a] ? 1 : 0;
    int nd = (int)std::floor(o.L[a] / d->cell_size);
    if (nd < 1) nd = 1;
    o.dims[a] = nd;
    o.h[a] = o.L[a] / nd;
  }
}

// migrate + backup_r: wrap, bin, reorder (cells in index order, particles by id)
static void move_particles(Oracle& o) {
  o.walls_bk = o.walls;  // backup_drivers (config_move_particles.msp:149-153)
  const int64_t n = o.n;
  std::vector<int64_t> cell(n);
  for (int64_t i = 0; i < n; i++) {
    double c[3] = {o.r[i].x, o.r[i].y, o.r[i].z};
    int ci[3];
    for (int a = 0; a < 3; a++) {
      if (o.per[a]) {
        if (c[a] < o.lo[a]) c[a] += o.L[a];
        else if (c[a] >= o.lo[a] + o.L[a]) c[a] -= o.L[a];
      }
      int k = (int)std::floor((c[a] - o.lo[a]) / o.h[a]);
      if (k < 0) k = 0;
      if (k >= o.dims[a]) k = o.dims[a] - 1;
      ci[a] = k;
    }
    o.r[i] = {c[0], c[1], c[2]};
    cell[i] = ((int64_t)ci[2] * o.dims[1] + ci[1]) * o.dims[0] + ci[0];
  }
  std::vector<int64_t> perm(n);
  for (int64_t i = 0; i < n; i++) perm[i] = i;
  std::sort(perm.begin(), perm.end(), [&](int64_t a, int64_t b) {
    if (cell[a] != cell[b]) return cell[a] < cell[b];
    return o.id[a] < o.id[b];
  });
  auto apply = [&](auto& vec) {
    auto tmp = vec;
    for (int64_t i = 0; i < n; i++) vec[i] = tmp[perm[i]];
  };
  apply(o.id); apply(o.type); apply(o.group);
  apply(o.r); apply(o.v); apply(o.f); apply(o.mom); apply(o.vrot); apply(o.arot); apply(o.inertia);
  apply(o.mass); apply(o.radius); apply(o.homothety); apply(o.q);
  if (o.poly) apply(o.backup_q);
  // interaction storage travels with its particle
  {
    auto tmp = std::move(o.items);
    o.items.assign(n, {});
    if ((int64_t)tmp.size() == n)
      for (int64_t i = 0; i < n; i++) o.items[i] = std::move(tmp[perm[i]]);
  }
  const int64_t nc = (int64_t)o.dims[0] * o.dims[1] * o.dims[2];
  o.cell_start.assign(nc + 1, 0);
  for (int64_t i = 0; i < n; i++) o.cell_start[cell[perm[i]] + 1]++;
  for (int64_t c = 0; c < nc; c++) o.cell_start[c + 1] += o.cell_start[c];
  o.backup = o.r;
  if (o.poly) o.backup_q = o.q;
}

// chunk_neighbors_contact: full (both direction) streams with the DEM filter.
static void chunk_neighbors(Oracle& o, double rcut_inc, double nbh_dist) {
  const int64_t n = o.n;
  o.streams.assign(n, {});
  const double rcut2 = nbh_dist * nbh_dist;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t x = 0; x < n; x++) {
    const V3 rx = o.r[x];
    int ci[3];
    // the cell of x is recomputed from its (wrapped) position at rebuild
    double c[3] = {rx.x, rx.y, rx.z};
    for (int a = 0; a < 3; a++) {
      int k = (int)std::floor((c[a] - o.lo[a]) / o.h[a]);
      if (k < 0) k = 0;
      if (k >= o.dims[a]) k = o.dims[a] - 1;
      ci[a] = k;
    }
    auto& st = o.streams[x];
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          int nb[3] = {ci[0] + dx, ci[1] + dy, ci[2] + dz};
          int s[3] = {0, 0, 0};
          bool ok = true;
          for (int a = 0; a < 3; a++) {
            if (nb[a] < 0) {
              if (!o.per[a]) { ok = false; break; }
              nb[a] += o.dims[a];
              s[a] = -1;
            } else if (nb[a] >= o.dims[a]) {
              if (!o.per[a]) { ok = false; break; }
              nb[a] -= o.dims[a];
              s[a] = 1;
            }
          }
          if (!ok) continue;
          const int code = (s[0] + 1) + 3 * (s[1] + 1) + 9 * (s[2] + 1);
          const V3 sh = shift_vec(o, code);
          const int64_t cb = ((int64_t)nb[2] * o.dims[1] + nb[1]) * o.dims[0] + nb[0];
          for (int64_t y = o.cell_start[cb]; y < o.cell_start[cb + 1]; y++) {
            if (y == x && code == 13) continue;
            const V3 ry = o.r[y] + sh;  // ghost copy position
            const V3 d = rx - ry;
            const double d2 = dot(d, d);
            const double rmax = o.radius[x] + o.radius[y] + rcut_inc;
            if (d2 > 0.0 && d2 < rmax * rmax && d2 < rcut2) st.push_back({(int32_t)y, code});
          }
        }
  }
}

// nbh_sphere, symetric mode: rebuild per-particle item lists carrying history by key.
// Sphere against the mesh (the sphere is a vertex of radius R; features carry the Minkowski radius rm):
// filters and detection of detect<7>, detect<8>, detect<9> (shape_detection_api.hpp:246-309, 357-386;
// shape_detection_core.hpp:43-53, 57-81, 83-101, 122-148, 397-466), restated for triangles.
static inline V3 unit3(V3 v) { return v / norm(v); }
static bool mesh_vf(V3 vi, const V3 (&t)[3], double limit, double& dist_out, V3& n_out, bool& early) {
  const V3 va = t[0], vb = t[1], vc = t[2];
  const V3 v = vi - va;
  V3 v1 = vb - va;
  V3 v2 = vc - va;
  v1 = unit3(v1);
  V3 n = unit3(cross(v1, v2));
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
static void mesh_tri(const Oracle& o, uint32_t k, V3 (&t)[3]) {
  for (int a = 0; a < 3; a++) t[a] = o.mesh.v[o.mesh.f[k][a]];
}
// kind 0 vertex, 1 edge, 2 face
static bool mesh_filter(const Oracle& o, int kind, uint32_t k, V3 p, double R, double rv) {
  const double rm = o.mesh.rm;
  if (kind == 0) {
    const double Rm = R + rm + rv;
    const V3 d = p - o.mesh.v[k];
    return dot(d, d) <= Rm * Rm;
  }
  if (kind == 1) {
    const V3 vf = o.mesh.v[o.mesh.e[k][0]], vs = o.mesh.v[o.mesh.e[k][1]];
    const V3 fs = vs - vf, fi = p - vf;
    const double r = dot(fs, fi) / dot(fs, fs);
    if (r < -rv || r > 1.0 + rv) return false;
    const V3 d = fi - fs * r;
    const double Rmax = R + rm + rv;
    return dot(d, d) < Rmax * Rmax;
  }
  V3 t[3];
  mesh_tri(o, k, t);
  double dist;
  V3 n;
  bool early;
  return mesh_vf(p, t, R + rm + rv, dist, n, early);
}
static Contact mesh_detect(const Oracle& o, int kind, uint32_t k, V3 p, double R) {
  const double rm = o.mesh.rm;
  if (kind == 0) return detect_vv(p, R, o.mesh.v[k], rm);
  if (kind == 1) {
    const V3 vf = o.mesh.v[o.mesh.e[k][0]], vs = o.mesh.v[o.mesh.e[k][1]];
    const V3 fs = vs - vf, fi = p - vf;
    const double r = dot(fs, fi) / dot(fs, fs);
    if (r <= 0 || r >= 1.0) return {false, 0, {0, 0, 0}, {0, 0, 0}};
    V3 n = fi - fs * r;
    const double dn = norm(n) - (R + rm);
    if (dn > 0) return {false, 0, {0, 0, 0}, {0, 0, 0}};
    n = unit3(n);
    return {true, dn, n, p - n * (R + 0.5 * dn)};
  }
  V3 t[3];
  mesh_tri(o, k, t);
  double dist;
  V3 n;
  bool early;
  const bool in = mesh_vf(p, t, R + rm, dist, n, early);
  if (early) return {false, 0, {0, 0, 0}, {0, 0, 0}};
  const double dn = dist - (R + rm);
  return {in, dn, n, p - n * (R + 0.5 * dn)};
}

static void nbh_sphere(Oracle& o, double rcut_inc) {
  const int64_t n = o.n;
  if ((int64_t)o.items.size() != n) o.items.assign(n, {});
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t x = 0; x < n; x++) {
    // extract_history + stable_sort (interaction_manager.hpp:205-217; nbh_sphere.cu:122-124)
    std::vector<Item> hist;
    for (auto& it : o.items[x])
      if (item_active(it)) hist.push_back(it);
    std::stable_sort(hist.begin(), hist.end(), key_less);
    std::vector<Item> lst;
    Item base{};
    base.id_i = o.id[x];
    base.sub_i = base.sub_j = 0;
    base.swap = 0;
    base.ghost = 0;
    base.friction = {0, 0, 0};
    base.moment = {0, 0, 0};
    // drivers first (nbh_sphere.cu:137-233): Surface filter with rad + rcut_inc
    for (size_t w = 0; w < o.walls.size(); w++) {
      if (surf_filter(o.walls[w], o.radius[x] + rcut_inc, o.r[x])) {
        Item it = base;
        it.type = 5;
        it.id_j = w;
        it.partner = -1;
        it.shift = 13;
        lst.push_back(it);
      }
    }
    // mesh driver (detection_sphere_driver, nbh_sphere.hpp:25-86, rVerlet = rcut_inc): vertices, edges,
    // faces, each by index
    if (o.mesh.on) {
      const size_t cnt[3] = {o.mesh.v.size(), o.mesh.e.size(), o.mesh.f.size()};
      for (int kind = 0; kind < 3; kind++)
        for (size_t k = 0; k < cnt[kind]; k++) {
          if (!mesh_filter(o, kind, (uint32_t)k, o.r[x], o.radius[x], rcut_inc)) continue;
          Item it = base;
          it.type = (uint16_t)(7 + kind);
          it.id_j = o.mesh.id;
          it.sub_j = (uint32_t)k;
          it.partner = -1;
          it.shift = 13;
          lst.push_back(it);
        }
    }
    // particle pairs in stream order, owner = lower id unless the partner is a periodic image
    for (auto& s : o.streams[x]) {
      const uint64_t idn = o.id[s.j];
      if (o.id[x] >= idn && s.shift == 13) continue;
      Item it = base;
      it.type = 0;
      it.id_j = idn;
      it.partner = s.j;
      it.shift = s.shift;
      lst.push_back(it);
    }
    // update(): lower_bound in history, copy payload on key equality (placeholder_interaction.hpp:398-408)
    if (g_mutation != 4) {
      for (auto& it : lst) {
        auto lb = std::lower_bound(hist.begin(), hist.end(), it, key_less);
        if (lb != hist.end() && key_eq(it, *lb)) {
          it.friction = lb->friction;
          it.moment = lb->moment;
        }
      }
    }
    o.items[x] = std::move(lst);
  }
  o.built = true;
}

static void rebuild(Oracle& o, double rcut_inc) {
  move_particles(o);
  double rmax = 0;
  for (int64_t i = 0; i < o.n; i++) rmax = std::max(rmax, 2 * o.radius[i]);
  chunk_neighbors(o, rcut_inc, rmax + rcut_inc);
  nbh_sphere(o, rcut_inc);
}

// contact_sphere: evaluate every interaction (owner side), then accumulate per particle.
static void contact(Oracle& o, const CP& cp, const CP& cpd, double dt) {
  const int64_t n = o.n;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t x = 0; x < n; x++) {
    for (auto& it : o.items[x]) {
      it.in_contact = false;
      if (it.type == 0) {
        const int64_t y = it.partner;
        const V3 ri = o.r[x];
        const V3 rj = o.r[y] + shift_vec(o, it.shift);
        const Contact c = detect_vv(ri, o.radius[x], rj, o.radius[y]);
        if (c.contact) {
          const double mi = o.mass[x], mj = o.mass[y];
          const double meff = (mi * mj) / (mi + mj);
          V3 f{0, 0, 0};
          contact_force_core(c.dn, c.n, dt, cp, meff, it.friction, c.cp, ri, o.v[x], f, it.moment, o.vrot[x], rj,
                             o.v[y], o.vrot[y]);
          it.in_contact = true;
          it.f = f;
          it.cp = c.cp;
          it.mint = it.moment;
        } else {
          it.friction = {0, 0, 0};
          it.moment = {0, 0, 0};
        }
      } else if (it.type >= 7) {  // mesh (ContactLawRShapeDriverFunc, contact_sphere.hpp:298-396)
        const V3 ri = o.r[x];
        const Contact c = mesh_detect(o, it.type - 7, it.sub_j, ri, o.radius[x]);
        if (c.contact) {
          V3 f{0, 0, 0};
          const V3 zero{0, 0, 0};
          contact_force_core(c.dn, c.n, dt, cpd, o.mass[x], it.friction, c.cp, ri, o.v[x], f, it.moment, o.vrot[x],
                             o.mesh.center, zero, zero);
          it.in_contact = true;
          it.f = f;
          it.cp = c.cp;
          it.mint = it.moment;
        } else {
          it.friction = {0, 0, 0};
          it.moment = {0, 0, 0};
        }
      } else {  // wall (ContactLawDriverFunc<Surface>, contact_sphere.hpp:216-297)
        const Surf& s = o.walls[it.id_j];
        const V3 ri = o.r[x];
        const Contact c = surf_detect(s, o.radius[x], ri);
        if (c.contact) {
          V3 f{0, 0, 0};
          contact_force_core(c.dn, c.n, dt, cpd, o.mass[x], it.friction, c.cp, ri, o.v[x], f, it.moment, o.vrot[x],
                             s.center, s.vel, s.vrot);
          it.in_contact = true;
          it.f = f;
          it.cp = c.cp;
          it.mint = it.moment;
        } else {
          it.friction = {0, 0, 0};
          it.moment = {0, 0, 0};
        }
      }
    }
  }
  // canonical accumulation: stream order, then walls, then mesh features
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t x = 0; x < n; x++) {
    V3 fx = o.f[x], mx = o.mom[x];
    const auto& own = o.items[x];
    size_t k = 0;
    while (k < own.size() && own[k].type != 0) k++;  // skip wall items (listed first)
    for (auto& s : o.streams[x]) {
      const uint64_t idn = o.id[s.j];
      const bool owner = !(o.id[x] >= idn && s.shift == 13);
      if (owner) {
        const Item& it = own[k++];
        if (it.in_contact) {
          fx = fx + it.f;
          mx = mx + compute_moments(it.cp, o.r[x], it.f, it.mint);
        }
      } else {
        // the partner owns the interaction (interior pair): find its item for x
        const auto& ol = o.items[s.j];
        for (const auto& it : ol) {
          if (it.type == 0 && it.partner == x && it.shift == 13) {
            if (it.in_contact) {
              const V3 nf = -it.f;
              fx = fx + nf;
              mx = mx + compute_moments(it.cp, o.r[x], nf, -it.mint);
            }
            break;
          }
        }
      }
    }
    for (auto& it : own) {  // drivers in list order: walls, then mesh features
      if (it.type == 0 || !it.in_contact) continue;
      fx = fx + it.f;
      mx = mx + compute_moments(it.cp, o.r[x], it.f, it.mint);
    }
    o.f[x] = fx;
    o.mom[x] = mx;
  }
}

static void reset_force_moment(Oracle& o) {
  for (int64_t i = 0; i < o.n; i++) {
    o.f[i] = {0, 0, 0};
    o.mom[i] = {0, 0, 0};
  }
}
static void gravity(Oracle& o, V3 g) {
  for (int64_t i = 0; i < o.n; i++) {
    const double m = o.mass[i];
    o.f[i] = {o.f[i].x + g.x * m, o.f[i].y + g.y * m, o.f[i].z + g.z * m};
  }
}
static void force_to_accel(Oracle& o) {
  for (int64_t i = 0; i < o.n; i++) {
    const double m = o.mass[i];
    o.f[i] = {o.f[i].x / m, o.f[i].y / m, o.f[i].z / m};
  }
}
static void prolog(Oracle& o, double dt) {
  const double dt_2 = dt * 0.5, dt2_2 = dt * dt * 0.5;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < o.n; i++) {
    V3& r = o.r[i];
    V3& v = o.v[i];
    const V3 a = o.f[i];
    r.x = r.x + (v.x * dt + a.x * dt2_2);
    r.y = r.y + (v.y * dt + a.y * dt2_2);
    r.z = r.z + (v.z * dt + a.z * dt2_2);
    v.x = v.x + a.x * dt_2;
    v.y = v.y + a.y * dt_2;
    v.z = v.z + a.z * dt_2;
    // push_to_quaternion (push_to_quaternion.hpp:31-37)
    Q4 q = o.q[i];
    const Q4 d = qdot(q, o.vrot[i]);
    q = {q.w + d.w * dt, q.x + d.x * dt, q.y + d.y * dt, q.z + d.z * dt};
    o.q[i] = qnormalize(q);
    o.vrot[i] = o.vrot[i] + dt_2 * o.arot[i];
  }
}
static inline V3 angular_accel(Q4 Q, V3 mom, V3 vrot, V3 I) {  // angular_acceleration.hpp:28-42
  const Q4 Qi = qconj(Q);
  const V3 om = qrot(Qi, vrot);
  const V3 M = qrot(Qi, mom);
  const V3 dom{(M.x - (I.z - I.y) * om.y * om.z) / I.x, (M.y - (I.x - I.z) * om.z * om.x) / I.y,
               (M.z - (I.y - I.x) * om.x * om.y) / I.z};
  return qrot(Q, dom);
}
static void epilog(Oracle& o, double dt) {
  const double dt_2 = dt * 0.5;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < o.n; i++) {
    o.arot[i] = angular_accel(o.q[i], o.mom[i], o.vrot[i], o.inertia[i]);
    o.vrot[i] = o.vrot[i] + o.arot[i] * dt_2;
    o.v[i] = {o.v[i].x + o.f[i].x * dt_2, o.v[i].y + o.f[i].y * dt_2, o.v[i].z + o.f[i].z * dt_2};
  }
}
static bool displ_over(const Oracle& o, double thr) {
  const double t2 = thr * thr;
  for (int64_t i = 0; i < o.n; i++) {
    const V3 d = o.r[i] - o.backup[i];
    if (dot(d, d) >= t2) return true;
  }
  return false;
}

static CP to_cp(const xdem_contact_params* p) {
  CP c{};
  if (p) c = {p->kn, p->kt, p->kr, p->mu, p->damp_rate, p->fc, p->dncut, p->gamma};
  return c;
}


// ============================================================================ polyhedra
// Restates, for convex sphero-polyhedra (paths in the exaDEM source tree):
//   vertices        shape::get_vertex p + Q*(h v) (src/shape/include/exaDEM/shape.hpp:244-262),
//                   compute_vertices.hpp:50-67
//   OBB             type/OBB.hpp:62-224 (enlarge, rotate, 15-axis separating test with FLT_EPSILON),
//                   compute_obb (basic_types.hpp:37-49)
//   filters         shape_detection_core.hpp:43-53 (VV), 83-101 (VE), 245-309 (VF), 468-504 (EE)
//   detection       shape_detection_core.hpp:57-81, 122-148, 397-466, 520-566
//   interactions    src/polyhedron/nbh_polyhedron.cpp:223-376 (owner = lower id, OBB prefilter enlarged twice
//                   by rVerlet/2, per-vertex OBB gate, VV/VE/VF then EE, then the swap pass VE/VF B->A),
//                   nbh_polyhedron_driver.hpp:216-236 (vertex-wall candidates)
//   contact         src/polyhedron/include/exaDEM/polyhedron/contact.hpp:124-211 (symmetric) and
//                   contact.hpp:219-297 (vertex-wall)
//   trigger         vertex_displ_over (src/polyhedron/vertex_displ_over.cu:74-133, backup_dem.hpp:65-110)
// Accumulation order (free in the reference, which uses atomics): per particle, gravity, then for every
// neighbour of its stream one subtotal over that pair's feature interactions in list order, then walls.

struct OBBo {
  V3 c, e1, e2, e3, ext;
};

static inline void obb_enlarge(OBBo& b, double more) {
  b.ext.x += more;
  b.ext.y += more;
  b.ext.z += more;
}

// OBB::intersect (type/OBB.hpp:90-224)
static bool obb_intersect(const OBBo& a, const OBBo& b) {
  const double tol = FLT_EPSILON;
  double ra, rb;
  double R[3][3], A[3][3];
  const V3 ea[3] = {a.e1, a.e2, a.e3}, eb[3] = {b.e1, b.e2, b.e3};
  const double ae[3] = {a.ext.x, a.ext.y, a.ext.z}, be[3] = {b.ext.x, b.ext.y, b.ext.z};
  const V3 tt = a.c - b.c;
  const double t[3] = {dot(tt, a.e1), dot(tt, a.e2), dot(tt, a.e3)};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      R[i][j] = dot(ea[i], eb[j]);
      A[i][j] = std::fabs(R[i][j]) + tol;
    }
  for (int i = 0; i < 3; i++) {
    ra = ae[i];
    rb = be[0] * A[i][0] + be[1] * A[i][1] + be[2] * A[i][2];
    if (std::fabs(t[i]) > ra + rb) return false;
  }
  for (int j = 0; j < 3; j++) {
    ra = ae[0] * A[0][j] + ae[1] * A[1][j] + ae[2] * A[2][j];
    rb = be[j];
    if (std::fabs(t[0] * R[0][j] + t[1] * R[1][j] + t[2] * R[2][j]) > ra + rb) return false;
  }
  // cross axes a_i x b_j
  ra = ae[1] * A[2][0] + ae[2] * A[1][0];
  rb = be[1] * A[0][2] + be[2] * A[0][1];
  if (std::fabs(t[2] * R[1][0] - t[1] * R[2][0]) > ra + rb) return false;
  ra = ae[1] * A[2][1] + ae[2] * A[1][1];
  rb = be[0] * A[0][2] + be[2] * A[0][0];
  if (std::fabs(t[2] * R[1][1] - t[1] * R[2][1]) > ra + rb) return false;
  ra = ae[1] * A[2][2] + ae[2] * A[1][2];
  rb = be[0] * A[0][1] + be[1] * A[0][0];
  if (std::fabs(t[2] * R[1][2] - t[1] * R[2][2]) > ra + rb) return false;
  ra = ae[0] * A[2][0] + ae[2] * A[0][0];
  rb = be[1] * A[1][2] + be[2] * A[1][1];
  if (std::fabs(t[0] * R[2][0] - t[2] * R[0][0]) > ra + rb) return false;
  ra = ae[0] * A[2][1] + ae[2] * A[0][1];
  rb = be[0] * A[1][2] + be[2] * A[1][0];
  if (std::fabs(t[0] * R[2][1] - t[2] * R[0][1]) > ra + rb) return false;
  ra = ae[0] * A[2][2] + ae[2] * A[0][2];
  rb = be[0] * A[1][1] + be[1] * A[1][0];
  if (std::fabs(t[0] * R[2][2] - t[2] * R[0][2]) > ra + rb) return false;
  ra = ae[0] * A[1][0] + ae[1] * A[0][0];
  rb = be[1] * A[2][2] + be[2] * A[2][1];
  if (std::fabs(t[1] * R[0][0] - t[0] * R[1][0]) > ra + rb) return false;
  ra = ae[0] * A[1][1] + ae[1] * A[0][1];
  rb = be[0] * A[2][2] + be[2] * A[2][0];
  if (std::fabs(t[1] * R[0][1] - t[0] * R[1][1]) > ra + rb) return false;
  ra = ae[0] * A[1][2] + ae[1] * A[0][2];
  rb = be[0] * A[2][1] + be[1] * A[2][0];
  if (std::fabs(t[1] * R[0][2] - t[0] * R[1][2]) > ra + rb) return false;
  return true;
}

// compute_obb: body-frame box rotated by Q, scaled by h, centred on the particle
static OBBo compute_obb(const xdem_shape& sh, V3 p, Q4 q, double h) {
  OBBo b;
  b.c = {0, 0, 0};
  b.e1 = qrot(q, V3{1, 0, 0});
  b.e2 = qrot(q, V3{0, 1, 0});
  b.e3 = qrot(q, V3{0, 0, 1});
  b.c = qrot(q, b.c);
  b.ext = {sh.obb_extent[0], sh.obb_extent[1], sh.obb_extent[2]};
  if (h != 1.0) {
    b.ext = b.ext * h;
    b.c = b.c * h;
  }
  b.c = b.c + p;
  return b;
}

static inline V3 shape_vertex(const xdem_shape& sh, int k) {
  return {sh.vertices[k][0], sh.vertices[k][1], sh.vertices[k][2]};
}
static inline V3 get_vertex(const xdem_shape& sh, int k, V3 p, double h, Q4 q) {
  return p + qrot(q, h * shape_vertex(sh, k));
}
static inline V3 vnormalize(V3 v) { return v / norm(v); }

static inline bool filter_vv(double rv, V3 vi, double ri, V3 vj, double rj) {
  const double R = ri + rj + rv;
  const V3 d = vi - vj;
  return dot(d, d) <= R * R;
}
static inline bool filter_ve(double rv, V3 vi, double ri, V3 vf, V3 vs, double rj) {
  const V3 fs = vs - vf;
  const V3 fi = vi - vf;
  const double r = dot(fs, fi) / dot(fs, fs);
  if (r < -rv || r > 1.0 + rv) return false;
  const V3 d = fi - fs * r;
  const double Rmax = ri + rj + rv;
  return dot(d, d) < Rmax * Rmax;
}

// shared body of the vertex-face filter and detection; returns (inside polygon, dist, n)
static inline bool vf_core(V3 vi, const std::array<V3, 8>& vj, const xdem_shape& shj, int face, double limit,
                           double& dist_out, V3& n_out, bool& early_out) {
  const int nf = shj.face_size[face];
  const int* f = shj.faces[face];
  V3 va = vj[f[0]];
  V3 vb = vj[f[1]];
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
  early_out = false;
  if (dist > limit) {
    early_out = true;
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

static inline bool filter_ee(double rv, V3 vfi, V3 vsi, double ri, V3 vfj, V3 vsj, double rj) {
  const double R = ri + rj + rv;
  const V3 Ei = vsi - vfi, Ej = vsj - vfj, v = vfi - vfj;
  const double c = dot(Ei, Ei), d = dot(Ej, Ej), e = dot(Ei, Ej);
  double f = (c * d) - (e * e);
  if (std::fabs(f) > 1.0e-12) {
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

static inline Contact detect_ve(V3 vi, double ri, V3 vf, V3 vs, double rj) {
  const V3 fs = vs - vf, fi = vi - vf;
  const double r = dot(fs, fi) / dot(fs, fs);
  if (r <= 0 || r >= 1.0) return {false, 0, {0, 0, 0}, {0, 0, 0}};
  V3 n = fi - fs * r;
  const double dn = norm(n) - (ri + rj);
  if (dn > 0) return {false, 0, {0, 0, 0}, {0, 0, 0}};
  n = vnormalize(n);
  return {true, dn, n, vi - n * (ri + 0.5 * dn)};
}

static inline Contact detect_vf(V3 vi, double ri, const std::array<V3, 8>& vj, const xdem_shape& shj, int face,
                                double rj) {
  double dist;
  V3 n;
  bool early;
  const bool in = vf_core(vi, vj, shj, face, ri + rj, dist, n, early);
  if (early) return {false, 0, {0, 0, 0}, {0, 0, 0}};
  const double dn = dist - (ri + rj);
  return {in, dn, n, vi - n * (ri + 0.5 * dn)};
}

static inline Contact detect_ee(V3 vfi, V3 vsi, double ri, V3 vfj, V3 vsj, double rj) {
  const double R = ri + rj;
  const V3 Ei = vsi - vfi, Ej = vsj - vfj, v = vfi - vfj;
  const double c = dot(Ei, Ei), d = dot(Ej, Ej), e = dot(Ei, Ej);
  double f = (c * d) - (e * e);
  if (std::fabs(f) > 1.0e-12) {
    f = 1.0 / f;
    const double a = dot(Ei, v), b = dot(Ej, v);
    const double s = (e * b - a * d) * f, t = (c * b - e * a) * f;
    if (s <= 0.0 || s >= 1.0 || t <= 0.0 || t >= 1.0) return {false, 0, {0, 0, 0}, {0, 0, 0}};
    const V3 pi = vfi + Ei * s, pj = vfj + Ej * t;
    V3 n = pi - pj;
    const double dn = norm(n) - R;
    if (dn > 0) return {false, 0, {0, 0, 0}, {0, 0, 0}};
    n = vnormalize(n);
    return {dn <= 0, dn, n, pi - n * (ri + 0.5 * dn)};
  }
  return {false, 0, {0, 0, 0}, {0, 0, 0}};
}

static void compute_vertices(Oracle& o) {
  o.verts.resize(o.n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < o.n; i++) {
    const xdem_shape& sh = o.shapes[o.type[i]];
    for (int k = 0; k < sh.nv; k++) o.verts[i][k] = get_vertex(sh, k, o.r[i], o.homothety[i], o.q[i]);
  }
}

// feature evaluation of item `it` owned by x (returns contact + force on the item's particle i)
struct PolyEval {
  bool in_contact;
  V3 f, cp, mint;
  V3 arm_i;     // point whose arm gives the vertex side's moment (the contact point)
  bool x_is_i;  // the list holder is the evaluation's particle i (swap = 0)
  V3 ri, rj;    // centre of i and j in the owner's frame
};

static std::array<V3, 8> shifted(const std::array<V3, 8>& v, int nv, V3 s) {
  std::array<V3, 8> out{};
  for (int k = 0; k < nv; k++) out[k] = v[k] + s;
  return out;
}

// Polyhedron against the triangle-mesh driver (RShape): add_driver_interaction for RShapeDriver
// (nbh_polyhedron_driver.hpp:60-200, rVerlet = rcut_inc), in its loop order with the mesh features by
// index: particle vertices against mesh vertices (7), edges (8) and faces (9); particle edges against mesh
// edges (10); then the mesh vertices whose box (the vertex enlarged by the mesh's Minkowski radius) meets
// the particle's box -- the shape box rotated and moved but not scaled by the homothety, enlarged by
// rVerlet -- against the particle's edges (11) and faces (12). sub_i is the particle's feature, sub_j the
// mesh's. The face boxes that gate (9) contain every point that passes the face filter, so the gate is
// not restated.
static void poly_mesh_items(const Oracle& o, int64_t x, const Item& base, double rv, std::vector<Item>& lst) {
  const xdem_shape& sa = o.shapes[o.type[x]];
  const double ra = sa.minkowski * o.homothety[x], rm = o.mesh.rm;
  const std::array<V3, 8>& va = o.verts[x];
  const auto& mv = o.mesh.v;
  const auto& me = o.mesh.e;
  auto add = [&](int type, int sub_i, size_t k) {
    Item it = base;
    it.type = (uint16_t)type;
    it.id_j = o.mesh.id;
    it.sub_i = (uint32_t)sub_i;
    it.sub_j = (uint32_t)k;
    it.partner = -1;
    it.shift = 13;
    lst.push_back(it);
  };
  for (int i = 0; i < sa.nv; i++) {
    for (size_t k = 0; k < mv.size(); k++)
      if (filter_vv(rv, va[i], ra, mv[k], rm)) add(7, i, k);
    for (size_t k = 0; k < me.size(); k++)
      if (filter_ve(rv, va[i], ra, mv[me[k][0]], mv[me[k][1]], rm)) add(8, i, k);
    for (size_t k = 0; k < o.mesh.f.size(); k++) {
      V3 t[3];
      mesh_tri(o, (uint32_t)k, t);
      double dist;
      V3 nn;
      bool early;
      if (mesh_vf(va[i], t, ra + rm + rv, dist, nn, early)) add(9, i, k);
    }
  }
  for (int i = 0; i < sa.ne; i++)
    for (size_t k = 0; k < me.size(); k++)
      if (filter_ee(rv, va[sa.edges[i][0]], va[sa.edges[i][1]], ra, mv[me[k][0]], mv[me[k][1]], rm)) add(10, i, k);
  OBBo obb_i = compute_obb(sa, o.r[x], o.q[x], 1.0);
  obb_enlarge(obb_i, rv);
  for (size_t k = 0; k < mv.size(); k++) {
    OBBo ov;
    ov.c = mv[k];
    ov.e1 = {1, 0, 0};
    ov.e2 = {0, 1, 0};
    ov.e3 = {0, 0, 1};
    ov.ext = {0, 0, 0};
    obb_enlarge(ov, rm);
    if (g_mutation != 10 && !obb_intersect(ov, obb_i)) continue;  // planted mistake 10: the gate dropped
    for (int i = 0; i < sa.ne; i++)
      if (filter_ve(rv, mv[k], rm, va[sa.edges[i][0]], va[sa.edges[i][1]], ra)) add(11, i, k);
    for (int i = 0; i < sa.nf; i++) {
      double dist;
      V3 nn;
      bool early;
      if (vf_core(mv[k], va, sa, i, rm + ra + rv, dist, nn, early)) add(12, i, k);
    }
  }
}

// nbh_polyhedron: rebuild the per-particle item lists (owner = lower id), carrying history by key.
static void nbh_polyhedron(Oracle& o, double rv) {
  const int64_t n = o.n;
  if ((int64_t)o.items.size() != n) o.items.assign(n, {});
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t x = 0; x < n; x++) {
    std::vector<Item> hist;
    for (auto& it : o.items[x])
      if (item_active(it)) hist.push_back(it);
    std::stable_sort(hist.begin(), hist.end(), key_less);
    std::vector<Item> lst;
    Item base{};
    base.id_i = o.id[x];
    const xdem_shape& sa = o.shapes[o.type[x]];
    const double ha = o.homothety[x];
    const double ra = sa.minkowski * ha;
    // drivers first: one vertex-wall candidate per vertex passing Surface::filter(rVerlet + r, vertex)
    for (size_t w = 0; w < o.walls.size(); w++)
      for (int k = 0; k < sa.nv; k++)
        if (surf_filter(o.walls[w], rv + ra, o.verts[x][k])) {
          Item it = base;
          it.type = 5;
          it.id_j = w;
          it.partner = -1;
          it.shift = 13;
          it.sub_i = (uint32_t)k;
          it.sub_j = (uint32_t)-1;
          lst.push_back(it);
        }
    if (o.mesh.on) poly_mesh_items(o, x, base, rv, lst);
    OBBo obb_i0 = compute_obb(sa, o.r[x], o.q[x], ha);
    for (auto& s : o.streams[x]) {
      const int64_t y = s.j;
      if (o.id[x] >= o.id[y]) continue;
      const V3 sh = shift_vec(o, s.shift);
      const xdem_shape& sb = o.shapes[o.type[y]];
      const double hb = o.homothety[y];
      const double rb = sb.minkowski * hb;
      const std::array<V3, 8> vb = shifted(o.verts[y], sb.nv, sh);
      const std::array<V3, 8>& va = o.verts[x];
      OBBo obb_i = obb_i0;
      OBBo obb_j = compute_obb(sb, o.r[y] + sh, o.q[y], hb);
      obb_enlarge(obb_i, 0.5 * rv);
      obb_enlarge(obb_j, 0.5 * rv);
      if (!obb_intersect(obb_i, obb_j)) continue;
      obb_enlarge(obb_i, 0.5 * rv);
      obb_enlarge(obb_j, 0.5 * rv);
      Item it = base;
      it.id_j = o.id[y];
      it.partner = (int32_t)y;
      it.shift = s.shift;
      auto add = [&](int type, int swap, int sub_i, int sub_j) {
        it.type = (uint16_t)type;
        it.swap = (uint8_t)swap;
        // key subs are owner().sub_, partner().sub_
        it.sub_i = (uint32_t)(swap ? sub_j : sub_i);
        it.sub_j = (uint32_t)(swap ? sub_i : sub_j);
        lst.push_back(it);
      };
      for (int i = 0; i < sa.nv; i++) {
        OBBo ov;
        ov.c = va[i];
        ov.e1 = {1, 0, 0};
        ov.e2 = {0, 1, 0};
        ov.e3 = {0, 0, 1};
        ov.ext = {0, 0, 0};
        obb_enlarge(ov, ra);
        if (!obb_intersect(obb_j, ov)) continue;
        for (int j = 0; j < sb.nv; j++)
          if (filter_vv(rv, va[i], ra, vb[j], rb)) add(0, 0, i, j);
        for (int j = 0; j < sb.ne; j++)
          if (filter_ve(rv, va[i], ra, vb[sb.edges[j][0]], vb[sb.edges[j][1]], rb)) add(1, 0, i, j);
        for (int j = 0; j < sb.nf; j++) {
          double dist;
          V3 nn;
          bool early;
          if (vf_core(va[i], vb, sb, j, ra + rb + rv, dist, nn, early)) add(2, 0, i, j);
        }
      }
      for (int i = 0; i < sa.ne; i++)
        for (int j = 0; j < sb.ne; j++)
          if (filter_ee(rv, va[sa.edges[i][0]], va[sa.edges[i][1]], ra, vb[sb.edges[j][0]], vb[sb.edges[j][1]], rb))
            add(3, 0, i, j);
      for (int j = 0; j < sb.nv; j++) {
        OBBo ov;
        ov.c = vb[j];
        ov.e1 = {1, 0, 0};
        ov.e2 = {0, 1, 0};
        ov.e3 = {0, 0, 1};
        ov.ext = {0, 0, 0};
        obb_enlarge(ov, rb);
        if (!obb_intersect(obb_i, ov)) continue;
        for (int i = 0; i < sa.ne; i++)
          if (filter_ve(rv, vb[j], rb, va[sa.edges[i][0]], va[sa.edges[i][1]], ra)) add(1, 1, j, i);
        for (int i = 0; i < sa.nf; i++) {
          double dist;
          V3 nn;
          bool early;
          if (vf_core(vb[j], va, sa, i, rb + ra + rv, dist, nn, early)) add(2, 1, j, i);
        }
      }
    }
    if (g_mutation != 4) {
      for (auto& it : lst) {
        auto lb = std::lower_bound(hist.begin(), hist.end(), it, key_less);
        if (lb != hist.end() && key_eq(it, *lb)) {
          it.friction = lb->friction;
          it.moment = lb->moment;
        }
      }
    }
    o.items[x] = std::move(lst);
  }
  o.built = true;
}

// evaluation of one particle-particle feature item held by x
static PolyEval eval_pp(Oracle& o, int64_t x, Item& it, const CP& cp, double dt) {
  PolyEval e{};
  const int64_t y = it.partner;
  const V3 sh = shift_vec(o, it.shift);
  const xdem_shape& sa = o.shapes[o.type[x]];
  const xdem_shape& sb = o.shapes[o.type[y]];
  const double ha = o.homothety[x], hb = o.homothety[y];
  const std::array<V3, 8> vb = shifted(o.verts[y], sb.nv, sh);
  const std::array<V3, 8>& va = o.verts[x];
  const bool swap = it.swap != 0;
  // evaluation particle i: the vertex side; j: the other one
  const xdem_shape& si = swap ? sb : sa;
  const xdem_shape& sj = swap ? sa : sb;
  const std::array<V3, 8>& vi = swap ? vb : va;
  const std::array<V3, 8>& vj = swap ? va : vb;
  const double ri = si.minkowski * (swap ? hb : ha), rj = sj.minkowski * (swap ? ha : hb);
  const int fi = (int)(swap ? it.sub_j : it.sub_i), fj = (int)(swap ? it.sub_i : it.sub_j);
  Contact c{false, 0, {0, 0, 0}, {0, 0, 0}};
  switch (it.type) {
    case 0: {
      const Contact d = detect_vv(vi[fi], ri, vj[fj], rj);
      c = d;
      break;
    }
    case 1:
      c = detect_ve(vi[fi], ri, vj[sj.edges[fj][0]], vj[sj.edges[fj][1]], rj);
      break;
    case 2:
      c = detect_vf(vi[fi], ri, vj, sj, fj, rj);
      break;
    case 3:
      c = detect_ee(vi[si.edges[fi][0]], vi[si.edges[fi][1]], ri, vj[sj.edges[fj][0]], vj[sj.edges[fj][1]], rj);
      break;
  }
  const int64_t pi = swap ? y : x, pj = swap ? x : y;
  const V3 r_i = swap ? o.r[y] + sh : o.r[x];
  const V3 r_j = swap ? o.r[x] : o.r[y] + sh;
  e.x_is_i = !swap;
  e.ri = r_i;
  e.rj = r_j;
  if (c.contact) {
    const double mi = o.mass[pi], mj = o.mass[pj];
    const double meff = g_mutation == 7 ? mi : (mi * mj) / (mi + mj);
    V3 f{0, 0, 0};
    contact_force_core(c.dn, c.n, dt, cp, meff, it.friction, c.cp, r_i, o.v[pi], f, it.moment, o.vrot[pi], r_j,
                       o.v[pj], o.vrot[pj]);
    e.in_contact = true;
    e.f = f;
    e.cp = c.cp;
    e.mint = it.moment;
    // planted mistake 6: the moment arm of the vertex side taken from the vertex, not the contact point
    e.arm_i = (g_mutation == 6 && it.type != 3) ? vi[fi] : c.cp;
  } else {
    it.friction = {0, 0, 0};
    it.moment = {0, 0, 0};
  }
  return e;
}

// One mesh item of x (ContactLawRShapeDriverFunc, polyhedron/contact.hpp:300-432; detect<7..12>,
// shape_detection_api.hpp:357-406): 7-10 with the particle as i and the mesh as j, 11-12 with the mesh
// vertex as i (the particle then takes -f and -moment, and the stored friction is negated after the law,
// as the reference does). The mesh is stationary: position its centre, no velocity, no spin.
static PolyEval eval_mesh(Oracle& o, int64_t x, Item& it, const CP& cpd, double dt) {
  PolyEval e{};
  const xdem_shape& sa = o.shapes[o.type[x]];
  const double ra = sa.minkowski * o.homothety[x], rm = o.mesh.rm;
  const std::array<V3, 8>& va = o.verts[x];
  const auto& mv = o.mesh.v;
  const uint32_t i = it.sub_i, k = it.sub_j;
  Contact c{false, 0, {0, 0, 0}, {0, 0, 0}};
  switch (it.type) {
    case 7:
      c = detect_vv(va[i], ra, mv[k], rm);
      break;
    case 8:
      c = detect_ve(va[i], ra, mv[o.mesh.e[k][0]], mv[o.mesh.e[k][1]], rm);
      break;
    case 9:
      c = mesh_detect(o, 2, k, va[i], ra);
      break;
    case 10:
      c = detect_ee(va[sa.edges[i][0]], va[sa.edges[i][1]], ra, mv[o.mesh.e[k][0]], mv[o.mesh.e[k][1]], rm);
      break;
    case 11:
      c = detect_ve(mv[k], rm, va[sa.edges[i][0]], va[sa.edges[i][1]], ra);
      break;
    case 12:
      c = detect_vf(mv[k], rm, va, sa, (int)i, ra);
      break;
  }
  const V3 zero{0, 0, 0};
  if (c.contact) {
    V3 f{0, 0, 0};
    e.in_contact = true;
    e.cp = c.cp;
    if (it.type <= 10) {
      contact_force_core(c.dn, c.n, dt, cpd, o.mass[x], it.friction, c.cp, o.r[x], o.v[x], f, it.moment, o.vrot[x],
                         o.mesh.center, zero, zero);
      e.f = f;
      e.mint = it.moment;
    } else {
      contact_force_core(c.dn, c.n, dt, cpd, o.mass[x], it.friction, c.cp, o.mesh.center, zero, f, it.moment, zero,
                         o.r[x], o.v[x], o.vrot[x]);
      e.f = -f;
      e.mint = -it.moment;
      if (g_mutation != 9) it.friction = -it.friction;  // planted mistake 9: the negation dropped
    }
  } else {
    it.friction = {0, 0, 0};
    it.moment = {0, 0, 0};
  }
  return e;
}

static void contact_poly(Oracle& o, const CP& cp, const CP& cpd, double dt) {
  const int64_t n = o.n;
  std::vector<std::vector<PolyEval>> ev(n);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t x = 0; x < n; x++) {
    ev[x].resize(o.items[x].size());
    for (size_t k = 0; k < o.items[x].size(); k++) {
      Item& it = o.items[x][k];
      if (it.type == 5) {  // vertex - wall
        const Surf& s = o.walls[it.id_j];
        const xdem_shape& sa = o.shapes[o.type[x]];
        const double r = sa.minkowski * o.homothety[x];
        const Contact c = surf_detect(s, r, o.verts[x][it.sub_i]);
        PolyEval e{};
        if (c.contact) {
          V3 f{0, 0, 0};
          contact_force_core(c.dn, c.n, dt, cpd, o.mass[x], it.friction, c.cp, o.r[x], o.v[x], f, it.moment,
                             o.vrot[x], s.center, s.vel, s.vrot);
          e.in_contact = true;
          e.f = f;
          e.cp = c.cp;
          e.mint = it.moment;
        } else {
          it.friction = {0, 0, 0};
          it.moment = {0, 0, 0};
        }
        ev[x][k] = e;
      } else if (it.type >= 7) {
        ev[x][k] = eval_mesh(o, x, it, cpd, dt);
      } else {
        ev[x][k] = eval_pp(o, x, it, cp, dt);
      }
    }
  }
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t x = 0; x < n; x++) {
    V3 F = o.f[x], M = o.mom[x];
    for (auto& s : o.streams[x]) {
      const int64_t y = s.j;
      V3 Fp{0, 0, 0}, Mp{0, 0, 0};
      if (o.id[x] < o.id[y]) {
        for (size_t k = 0; k < o.items[x].size(); k++) {
          const Item& it = o.items[x][k];
          if (it.type == 5 || it.partner != y || it.shift != s.shift) continue;
          const PolyEval& e = ev[x][k];
          if (!e.in_contact) continue;
          if (e.x_is_i || g_mutation == 5) {  // planted mistake 5: the swap side ignored
            Fp = Fp + e.f;
            Mp = Mp + compute_moments(e.arm_i, e.ri, e.f, e.mint);
          } else {
            const V3 nf = -e.f;
            Fp = Fp + nf;
            Mp = Mp + compute_moments(e.cp, e.rj, nf, -e.mint);
          }
        }
      } else {
        const int mirror = 26 - s.shift;
        for (size_t k = 0; k < o.items[y].size(); k++) {
          const Item& it = o.items[y][k];
          if (it.type == 5 || it.partner != x || it.shift != mirror) continue;
          const PolyEval& e = ev[y][k];
          if (!e.in_contact) continue;
          if (e.x_is_i || g_mutation == 5) {  // y is i, x is j
            const V3 nf = -e.f;
            Fp = Fp + nf;
            Mp = Mp + compute_moments(e.cp, e.rj, nf, -e.mint);
          } else {  // x is i (vertex side of a swap item held by y)
            Fp = Fp + e.f;
            Mp = Mp + compute_moments(e.arm_i, e.ri, e.f, e.mint);
          }
        }
      }
      F = F + Fp;
      M = M + Mp;
    }
    for (size_t k = 0; k < o.items[x].size(); k++) {
      const Item& it = o.items[x][k];
      if (it.type != 5 || !ev[x][k].in_contact) continue;
      const PolyEval& e = ev[x][k];
      F = F + e.f;
      M = M + compute_moments(e.cp, o.r[x], e.f, e.mint);
    }
    // mesh contacts after the walls, in list order (f and mint already on the particle's side)
    for (size_t k = 0; k < o.items[x].size(); k++) {
      const Item& it = o.items[x][k];
      if (it.type < 7 || !ev[x][k].in_contact) continue;
      const PolyEval& e = ev[x][k];
      F = F + e.f;
      M = M + compute_moments(e.cp, o.r[x], e.f, e.mint);
    }
    o.f[x] = F;
    o.mom[x] = M;
  }
}

static void rebuild_poly(Oracle& o, double rcut_inc) {
  move_particles(o);
  compute_vertices(o);
  double rmax = 0;
  for (int64_t i = 0; i < o.n; i++) rmax = std::max(rmax, 2 * o.radius[i]);
  chunk_neighbors(o, rcut_inc, rmax + rcut_inc);
  nbh_polyhedron(o, rcut_inc);
}

// vertex_displ_over: any vertex moved by at least thr since the last rebuild
static bool vertex_displ_over(const Oracle& o, double thr) {
  const double t2 = thr * thr;
  for (int64_t i = 0; i < o.n; i++) {
    const xdem_shape& sh = o.shapes[o.type[i]];
    for (int k = 0; k < sh.nv; k++) {
      const V3 a = get_vertex(sh, k, o.backup[i], o.homothety[i], o.backup_q[i]);
      const V3 b = get_vertex(sh, k, o.r[i], o.homothety[i], o.q[i]);
      const V3 d = b - a;
      if (dot(d, d) >= t2) return true;
    }
  }
  return false;
}

}  // namespace orc

using namespace orc;

// ------------------------------------------------------------------ C interface (ctypes)
extern "C" {

void* xo_create(const xdem_particles* p, const xdem_domain* dom) {
  Oracle* o = new Oracle();
  const int64_t n = p->n;
  o->n = n;
  o->id.assign(p->id, p->id + n);
  o->type.assign(n, 0);
  o->group.assign(n, 0);
  if (p->type) o->type.assign(p->type, p->type + n);
  if (p->group) o->group.assign(p->group, p->group + n);
  auto v3 = [&](double* a, double* b, double* c) {
    std::vector<V3> out(n);
    for (int64_t i = 0; i < n; i++) out[i] = {a ? a[i] : 0.0, b ? b[i] : 0.0, c ? c[i] : 0.0};
    return out;
  };
  o->r = v3(p->rx, p->ry, p->rz);
  o->v = v3(p->vx, p->vy, p->vz);
  o->f = v3(p->fx, p->fy, p->fz);
  o->mom = v3(p->momx, p->momy, p->momz);
  o->vrot = v3(p->vrotx, p->vroty, p->vrotz);
  o->arot = v3(p->arotx, p->aroty, p->arotz);
  o->inertia = v3(p->inx, p->iny, p->inz);
  o->mass.assign(p->mass, p->mass + n);
  o->radius.assign(p->radius, p->radius + n);
  o->homothety.assign(n, 1.0);
  if (p->homothety) o->homothety.assign(p->homothety, p->homothety + n);
  o->q.resize(n);
  for (int64_t i = 0; i < n; i++) o->q[i] = {p->qw ? p->qw[i] : 1.0, p->qx ? p->qx[i] : 0.0, p->qy ? p->qy[i] : 0.0,
                                              p->qz ? p->qz[i] : 0.0};
  o->backup = o->r;
  set_domain(*o, dom);
  return o;
}

void xo_destroy(void* h) { delete (Oracle*)h; }

void xo_set_mutation(int m) { g_mutation = m; }

void xo_set_surfaces(void* h, const xdem_surface* w, int nw) {
  Oracle* o = (Oracle*)h;
  o->walls.clear();
  for (int k = 0; k < nw; k++) {
    Surf s{};
    s.kind = w[k].kind;
    s.radius = w[k].radius;
    s.axis = {w[k].axis[0], w[k].axis[1], w[k].axis[2]};
    s.offset = w[k].offset;
    s.normal = {w[k].normal[0], w[k].normal[1], w[k].normal[2]};
    s.center = {w[k].center[0], w[k].center[1], w[k].center[2]};
    if (s.kind == 0 && s.center.x == 0 && s.center.y == 0 && s.center.z == 0) s.center = s.offset * s.normal;
    if (s.kind == 1) s.center = mask_mul(s.axis, s.center);  // Cylinder::initialize (cylinder.hpp:104)
    s.vel = {w[k].vel[0], w[k].vel[1], w[k].vel[2]};
    s.vrot = {w[k].vrot[0], w[k].vrot[1], w[k].vrot[2]};
    s.center_proj = s.offset * s.normal;  // Surface::initialize (surface.hpp:145-150)
    o->walls.push_back(s);
  }
}

// Mesh driver: world vertex = center + quat (w, x, y, z) applied to vertex k (shape::get_vertex with
// homothety 1, shape.hpp:259-262). nv = 0 removes it.
void xo_set_mesh(void* h, uint64_t id, const double* vertices, int64_t nv, const uint32_t* edges, int64_t ne,
                 const uint32_t* faces, int64_t nf, double rm, const double* center, const double* quat) {
  Oracle* o = (Oracle*)h;
  o->mesh = Oracle::Mesh{};
  if (nv == 0) return;
  o->mesh.on = true;
  o->mesh.id = id;
  o->mesh.rm = rm;
  o->mesh.center = {center[0], center[1], center[2]};
  const Q4 q{quat[0], quat[1], quat[2], quat[3]};
  for (int64_t k = 0; k < nv; k++)
    o->mesh.v.push_back(o->mesh.center + qrot(q, V3{vertices[3 * k], vertices[3 * k + 1], vertices[3 * k + 2]}));
  for (int64_t k = 0; k < ne; k++) o->mesh.e.push_back({edges[2 * k], edges[2 * k + 1]});
  for (int64_t k = 0; k < nf; k++) o->mesh.f.push_back({faces[3 * k], faces[3 * k + 1], faces[3 * k + 2]});
}

// Mesh items (owner id, feature, type 7/8/9, friction, moment), sorted by (owner id, type, feature).
int64_t xo_mesh_interactions(void* h, int64_t cap, uint64_t* id_i, uint32_t* sub_j, uint16_t* type, double* fr,
                             double* mo) {
  Oracle& o = *(Oracle*)h;
  std::vector<const Item*> rows;
  for (auto& l : o.items)
    for (auto& it : l)
      if (it.type >= 7) rows.push_back(&it);
  std::sort(rows.begin(), rows.end(), [](const Item* a, const Item* b) {
    if (a->id_i != b->id_i) return a->id_i < b->id_i;
    if (a->type != b->type) return a->type < b->type;
    return a->sub_j < b->sub_j;
  });
  int64_t k = 0;
  for (const Item* it : rows) {
    if (k >= cap) break;
    id_i[k] = it->id_i;
    sub_j[k] = it->sub_j;
    type[k] = it->type;
    fr[3 * k] = it->friction.x;
    fr[3 * k + 1] = it->friction.y;
    fr[3 * k + 2] = it->friction.z;
    mo[3 * k] = it->moment.x;
    mo[3 * k + 1] = it->moment.y;
    mo[3 * k + 2] = it->moment.z;
    k++;
  }
  return (int64_t)rows.size();
}

void xo_rebuild(void* h, double rcut_inc) { rebuild(*(Oracle*)h, rcut_inc); }
void xo_move_particles(void* h) { move_particles(*(Oracle*)h); }
void xo_chunk_neighbors(void* h, double rcut_inc, double nbh_dist) { chunk_neighbors(*(Oracle*)h, rcut_inc, nbh_dist); }
void xo_nbh_sphere(void* h, double rcut_inc) { nbh_sphere(*(Oracle*)h, rcut_inc); }
void xo_reset_force_moment(void* h) { reset_force_moment(*(Oracle*)h); }
void xo_gravity(void* h, const double* g) { gravity(*(Oracle*)h, {g[0], g[1], g[2]}); }
void xo_contact(void* h, const xdem_contact_params* cp, const xdem_contact_params* cpd, double dt) {
  contact(*(Oracle*)h, to_cp(cp), to_cp(cpd), dt);
}
void xo_force_to_accel(void* h) { force_to_accel(*(Oracle*)h); }
void xo_prolog(void* h, double dt) { prolog(*(Oracle*)h, dt); }
void xo_epilog(void* h, double dt) { epilog(*(Oracle*)h, dt); }
int xo_displ_over(void* h, double thr) { return displ_over(*(Oracle*)h, thr) ? 1 : 0; }

// push_f_v_r_driver then push_f_v_driver (config_numerical_schemes.msp:13-14): every driver whose
// motion is active at time + dt moves by one step, then the kinematic velocities are set.
static void drivers_prolog(Oracle& o, double dt) {
  for (auto& w : o.walls)
    if (motion_on(w.mo, o.time + dt)) driver_push_f_v_r(w, g_mutation == 8 ? o.time + dt : o.time, dt);
  for (auto& w : o.walls) driver_push_f_v(w);
}

// first_iteration (main-config.msp:55-65): init_particles (rebuild), reset_force_moment,
// compute_force (gravity + contact), next_time_step; force_to_accel is not part of it.
void xo_first_iteration(void* h, const xdem_scheme* s) {
  Oracle& o = *(Oracle*)h;
  rebuild(o, s->rcut_inc);
  reset_force_moment(o);
  if (s->gravity_on) gravity(o, {s->gravity[0], s->gravity[1], s->gravity[2]});
  contact(o, to_cp(&s->cp), to_cp(&s->cpd), s->dt);
  o.time += s->dt;
}

// numerical_scheme_verlet_macro then next_time_step, nsteps times (config_numerical_schemes.msp:6-28:
// prolog, driver motion, rebuild when a particle or a driver moved by rcut_inc/2, forces, epilog).
// Returns the number of rebuilds.
int64_t xo_run_steps(void* h, const xdem_scheme* s, int64_t nsteps) {
  Oracle& o = *(Oracle*)h;
  int64_t rebuilds = 0;
  for (int64_t k = 0; k < nsteps; k++) {
    prolog(o, s->dt);
    drivers_prolog(o, s->dt);
    const bool moved = displ_over(o, 0.5 * s->rcut_inc);
    if (moved || drivers_displ_over(o.walls, o.walls_bk, 0.5 * s->rcut_inc)) {
      rebuild(o, s->rcut_inc);
      rebuilds++;
    }
    reset_force_moment(o);
    if (s->gravity_on) gravity(o, {s->gravity[0], s->gravity[1], s->gravity[2]});
    contact(o, to_cp(&s->cp), to_cp(&s->cpd), s->dt);
    force_to_accel(o);
    for (auto& w : o.walls) driver_push_f_v(w);
    epilog(o, s->dt);
    o.time += s->dt;
  }
  return rebuilds;
}

// Motion of each driver, as register_* checks it: check_motion_coherence (driver_params.hpp:108-214:
// LINEAR motion vectors and shaker directions normalised when |v|^2 - 1 >= 1e-14) and, for a linear
// surface, the motion vector replaced by the normal unless it is +-normal (surface.hpp:183-195).
// Returns 0, or -1 for a motion the reference rejects (zero motion vector, amplitude or omega <= 0).
int xo_set_driver_motion(void* h, const xdem_driver_motion* m, int n) {
  Oracle& o = *(Oracle*)h;
  for (int k = 0; k < n && k < (int)o.walls.size(); k++) {
    Motion mo;
    mo.type = m[k].motion_type;
    mo.mvec = {m[k].motion_vector[0], m[k].motion_vector[1], m[k].motion_vector[2]};
    mo.const_vel = m[k].const_vel;
    mo.t_start = m[k].motion_start_threshold;
    mo.t_end = m[k].motion_end_threshold;
    mo.amplitude = m[k].amplitude;
    mo.omega = m[k].omega;
    mo.shaker_dir = {m[k].shaker_dir[0], m[k].shaker_dir[1], m[k].shaker_dir[2]};
    if (mo.type == 7) {
      if (mo.amplitude <= 0.0 || mo.omega <= 0.0) return -1;
      if (dot(mo.shaker_dir, mo.shaker_dir) - 1 >= 1e-14) mo.shaker_dir = mo.shaker_dir / norm(mo.shaker_dir);
    }
    if (mo.type == 1) {
      if (mo.mvec.x == 0 && mo.mvec.y == 0 && mo.mvec.z == 0) return -1;
      if (dot(mo.mvec, mo.mvec) - 1 >= 1e-14) mo.mvec = mo.mvec / norm(mo.mvec);
      Surf& w = o.walls[k];
      if (w.kind == 0) {
        const V3 nn = w.normal, mn = -1.0 * nn;
        const bool same = mo.mvec.x == nn.x && mo.mvec.y == nn.y && mo.mvec.z == nn.z;
        const bool opp = mo.mvec.x == mn.x && mo.mvec.y == mn.y && mo.mvec.z == mn.z;
        if (!same && !opp) mo.mvec = nn;
      }
    }
    o.walls[k].mo = mo;
  }
  return 0;
}
void xo_set_time(void* h, double t) { ((Oracle*)h)->time = t; }
double xo_time(void* h) { return ((Oracle*)h)->time; }
// drivers as xdem_surface rows (offset, normal, centre, velocity, spin, kind, radius, axis)
void xo_get_drivers(void* h, xdem_surface* out, int n) {
  Oracle& o = *(Oracle*)h;
  for (int k = 0; k < n && k < (int)o.walls.size(); k++) {
    const Surf& w = o.walls[k];
    xdem_surface r{};
    r.offset = w.offset;
    const V3* src[4] = {&w.normal, &w.center, &w.vel, &w.vrot};
    double* dst[4] = {r.normal, r.center, r.vel, r.vrot};
    for (int q = 0; q < 4; q++) {
      dst[q][0] = src[q]->x;
      dst[q][1] = src[q]->y;
      dst[q][2] = src[q]->z;
    }
    r.kind = w.kind;
    r.radius = w.radius;
    r.axis[0] = w.axis.x;
    r.axis[1] = w.axis.y;
    r.axis[2] = w.axis.z;
    out[k] = r;
  }
}

int64_t xo_num_particles(void* h) { return ((Oracle*)h)->n; }

void xo_get_particles(void* h, xdem_particles* p) {
  Oracle& o = *(Oracle*)h;
  const int64_t n = o.n;
  for (int64_t i = 0; i < n; i++) {
    if (p->id) p->id[i] = o.id[i];
    if (p->type) p->type[i] = o.type[i];
    if (p->group) p->group[i] = o.group[i];
#define PUT3(A, B, C, F) \
  if (p->A) p->A[i] = o.F[i].x; \
  if (p->B) p->B[i] = o.F[i].y; \
  if (p->C) p->C[i] = o.F[i].z;
    PUT3(rx, ry, rz, r)
    PUT3(vx, vy, vz, v)
    PUT3(fx, fy, fz, f)
    PUT3(momx, momy, momz, mom)
    PUT3(vrotx, vroty, vrotz, vrot)
    PUT3(arotx, aroty, arotz, arot)
    PUT3(inx, iny, inz, inertia)
#undef PUT3
    if (p->mass) p->mass[i] = o.mass[i];
    if (p->radius) p->radius[i] = o.radius[i];
    if (p->homothety) p->homothety[i] = o.homothety[i];
    if (p->qw) p->qw[i] = o.q[i].w;
    if (p->qx) p->qx[i] = o.q[i].x;
    if (p->qy) p->qy[i] = o.q[i].y;
    if (p->qz) p->qz[i] = o.q[i].z;
  }
}

int64_t xo_interactions_count(void* h) {
  Oracle& o = *(Oracle*)h;
  int64_t c = 0;
  for (auto& l : o.items) c += (int64_t)l.size();
  return c;
}

// owner id, partner id (wall index), type, friction (3 per item), moment (3 per item)
int64_t xo_interactions_dump(void* h, int64_t cap, uint64_t* id_i, uint64_t* id_j, uint16_t* type, double* fr,
                             double* mo) {
  Oracle& o = *(Oracle*)h;
  int64_t k = 0;
  for (auto& l : o.items)
    for (auto& it : l) {
      if (k >= cap) return k;
      id_i[k] = it.id_i;
      id_j[k] = it.id_j;
      type[k] = it.type;
      fr[3 * k] = it.friction.x;
      fr[3 * k + 1] = it.friction.y;
      fr[3 * k + 2] = it.friction.z;
      if (mo) {
        mo[3 * k] = it.moment.x;
        mo[3 * k + 1] = it.moment.y;
        mo[3 * k + 2] = it.moment.z;
      }
      k++;
    }
  return k;
}

// Restart with interaction history (sim_dump_writer_interaction.cpp:25-46 writes, per particle, its
// interaction storage with the friction and moment payload; the restarted run's first nbh_sphere /
// nbh_polyhedron extracts the active ones and hands them to the new interactions by key,
// interaction_manager.hpp:88-112,205-217). Saved rows become the storage of their owner particle (id_i)
// before the caller's rebuild; rows whose owner is not a particle of this oracle are ignored. swap,
// sub_i, sub_j and mo may be null (zero).
int64_t xo_interactions_restore(void* h, int64_t m, const uint64_t* id_i, const uint64_t* id_j, const uint16_t* type,
                                const uint8_t* swap, const uint32_t* sub_i, const uint32_t* sub_j, const double* fr,
                                const double* mo) {
  Oracle& o = *(Oracle*)h;
  std::unordered_map<uint64_t, int64_t> at;
  at.reserve((size_t)o.n * 2);
  for (int64_t x = 0; x < o.n; x++) at[o.id[x]] = x;
  o.items.assign(o.n, {});
  int64_t used = 0;
  for (int64_t k = 0; k < m; k++) {
    const auto it = at.find(id_i[k]);
    if (it == at.end()) continue;
    Item t{};
    t.id_i = id_i[k];
    t.id_j = id_j[k];
    t.type = type[k];
    t.swap = swap ? swap[k] : 0;
    t.sub_i = sub_i ? sub_i[k] : 0;
    t.sub_j = sub_j ? sub_j[k] : 0;
    t.partner = -1;
    t.shift = 13;
    t.friction = {fr[3 * k], fr[3 * k + 1], fr[3 * k + 2]};
    t.moment = mo ? V3{mo[3 * k], mo[3 * k + 1], mo[3 * k + 2]} : V3{0, 0, 0};
    o.items[it->second].push_back(t);
    used++;
  }
  return used;
}

// Stream (neighbour list) dump: particle id, neighbour id, image code; full both-direction streams.
int64_t xo_streams_count(void* h) {
  Oracle& o = *(Oracle*)h;
  int64_t c = 0;
  for (auto& s : o.streams) c += (int64_t)s.size();
  return c;
}
int64_t xo_streams_dump(void* h, int64_t cap, uint64_t* a, uint64_t* b, int32_t* code) {
  Oracle& o = *(Oracle*)h;
  int64_t k = 0;
  for (int64_t x = 0; x < o.n; x++)
    for (auto& s : o.streams[x]) {
      if (k >= cap) return k;
      a[k] = o.id[x];
      b[k] = o.id[s.j];
      code[k] = s.shift;
      k++;
    }
  return k;
}

// Contact-law unit entry: one sphere-sphere evaluation (owner side) for closed-form checks.
// out: [0..2] f_i, [3..5] friction after, [6] dn, [7..9] n, [10..12] cp, [13] contact flag
void xo_pair_force(const double* pi, const double* vi, const double* wi, double ri, double mi, const double* pj,
                   const double* vj, const double* wj, double rj, double mj, const xdem_contact_params* cp, double dt,
                   double* friction_io, double* out) {
  const CP p = to_cp(cp);
  const V3 Pi{pi[0], pi[1], pi[2]}, Pj{pj[0], pj[1], pj[2]};
  const Contact c = detect_vv(Pi, ri, Pj, rj);
  V3 ft{friction_io[0], friction_io[1], friction_io[2]};
  V3 mom{0, 0, 0}, f{0, 0, 0};
  if (c.contact) {
    const double meff = (mi * mj) / (mi + mj);
    contact_force_core(c.dn, c.n, dt, p, meff, ft, c.cp, Pi, {vi[0], vi[1], vi[2]}, f, mom, {wi[0], wi[1], wi[2]}, Pj,
                       {vj[0], vj[1], vj[2]}, {wj[0], wj[1], wj[2]});
  } else {
    ft = {0, 0, 0};
  }
  friction_io[0] = ft.x;
  friction_io[1] = ft.y;
  friction_io[2] = ft.z;
  const double o[14] = {f.x, f.y, f.z, ft.x, ft.y, ft.z, c.dn, c.n.x, c.n.y, c.n.z, c.cp.x, c.cp.y, c.cp.z,
                        c.contact ? 1.0 : 0.0};
  std::memcpy(out, o, sizeof(o));
}

// Rigid-body helpers for closed-form checks of the integrator pieces.
void xo_angular_accel(const double* q, const double* mom, const double* vrot, const double* I, double* out) {
  const V3 a = angular_accel({q[0], q[1], q[2], q[3]}, {mom[0], mom[1], mom[2]}, {vrot[0], vrot[1], vrot[2]},
                             {I[0], I[1], I[2]});
  out[0] = a.x;
  out[1] = a.y;
  out[2] = a.z;
}
void xo_quat_rotate(const double* q, const double* v, double* out) {
  const V3 a = qrot({q[0], q[1], q[2], q[3]}, {v[0], v[1], v[2]});
  out[0] = a.x;
  out[1] = a.y;
  out[2] = a.z;
}

// ---------------------------------------------------------------- polyhedra entry points
void xo_set_shapes(void* h, const xdem_shape* shapes, int n) {
  Oracle& o = *(Oracle*)h;
  o.shapes.assign(shapes, shapes + n);
  o.poly = true;
  o.backup_q = o.q;
}
void xo_compute_vertices(void* h) { compute_vertices(*(Oracle*)h); }
void xo_rebuild_poly(void* h, double rcut_inc) { rebuild_poly(*(Oracle*)h, rcut_inc); }
void xo_contact_poly(void* h, const xdem_contact_params* cp, const xdem_contact_params* cpd, double dt) {
  contact_poly(*(Oracle*)h, to_cp(cp), to_cp(cpd), dt);
}
int xo_vertex_displ_over(void* h, double thr) { return vertex_displ_over(*(Oracle*)h, thr) ? 1 : 0; }

// first_iteration + nsteps of the polyhedra scheme (config_polyhedra.msp: compute_vertices on the fast
// path, vertex_displ_over as the trigger, contact_polyhedron in compute_force).
void xo_first_iteration_poly(void* h, const xdem_scheme* s) {
  Oracle& o = *(Oracle*)h;
  rebuild_poly(o, s->rcut_inc);
  reset_force_moment(o);
  if (s->gravity_on) gravity(o, {s->gravity[0], s->gravity[1], s->gravity[2]});
  contact_poly(o, to_cp(&s->cp), to_cp(&s->cpd), s->dt);
  o.time += s->dt;
}
int64_t xo_run_steps_poly(void* h, const xdem_scheme* s, int64_t nsteps) {
  Oracle& o = *(Oracle*)h;
  int64_t rebuilds = 0;
  for (int64_t k = 0; k < nsteps; k++) {
    prolog(o, s->dt);
    drivers_prolog(o, s->dt);
    const bool moved = vertex_displ_over(o, 0.5 * s->rcut_inc);
    if (moved || drivers_displ_over(o.walls, o.walls_bk, 0.5 * s->rcut_inc)) {
      rebuild_poly(o, s->rcut_inc);
      rebuilds++;
    } else {
      compute_vertices(o);
    }
    reset_force_moment(o);
    if (s->gravity_on) gravity(o, {s->gravity[0], s->gravity[1], s->gravity[2]});
    contact_poly(o, to_cp(&s->cp), to_cp(&s->cpd), s->dt);
    force_to_accel(o);
    for (auto& w : o.walls) driver_push_f_v(w);
    epilog(o, s->dt);
    o.time += s->dt;
  }
  return rebuilds;
}

// interaction dump with feature keys: owner id, partner id, type, swap, owner sub, partner sub, friction
int64_t xo_poly_interactions_dump(void* h, int64_t cap, uint64_t* id_i, uint64_t* id_j, uint16_t* type, uint8_t* swap,
                                  uint32_t* sub_i, uint32_t* sub_j, double* fr) {
  Oracle& o = *(Oracle*)h;
  int64_t k = 0;
  for (auto& l : o.items)
    for (auto& it : l) {
      if (k >= cap) return k;
      id_i[k] = it.id_i;
      id_j[k] = it.id_j;
      type[k] = it.type;
      swap[k] = it.swap;
      sub_i[k] = it.sub_i;
      sub_j[k] = it.sub_j;
      fr[3 * k] = it.friction.x;
      fr[3 * k + 1] = it.friction.y;
      fr[3 * k + 2] = it.friction.z;
      k++;
    }
  return k;
}

// One feature evaluation for closed-form checks: two polyhedra (shape a at pa/qa, shape b at pb/qb),
// feature type/subs with i = a; returns contact flag, dn, n, cp.
int xo_poly_detect(const xdem_shape* sa, const double* pa, const double* qa, const xdem_shape* sb, const double* pb,
                   const double* qb, int type, int sub_i, int sub_j, double* out) {
  std::array<V3, 8> va{}, vb{};
  const Q4 QA{qa[0], qa[1], qa[2], qa[3]}, QB{qb[0], qb[1], qb[2], qb[3]};
  for (int k = 0; k < sa->nv; k++) va[k] = get_vertex(*sa, k, {pa[0], pa[1], pa[2]}, 1.0, QA);
  for (int k = 0; k < sb->nv; k++) vb[k] = get_vertex(*sb, k, {pb[0], pb[1], pb[2]}, 1.0, QB);
  const double ri = sa->minkowski, rj = sb->minkowski;
  Contact c{false, 0, {0, 0, 0}, {0, 0, 0}};
  if (type == 0) c = detect_vv(va[sub_i], ri, vb[sub_j], rj);
  if (type == 1) c = detect_ve(va[sub_i], ri, vb[sb->edges[sub_j][0]], vb[sb->edges[sub_j][1]], rj);
  if (type == 2) c = detect_vf(va[sub_i], ri, vb, *sb, sub_j, rj);
  if (type == 3)
    c = detect_ee(va[sa->edges[sub_i][0]], va[sa->edges[sub_i][1]], ri, vb[sb->edges[sub_j][0]],
                  vb[sb->edges[sub_j][1]], rj);
  const double o8[8] = {c.contact ? 1.0 : 0.0, c.dn, c.n.x, c.n.y, c.n.z, c.cp.x, c.cp.y, c.cp.z};
  std::memcpy(out, o8, sizeof(o8));
  return c.contact ? 1 : 0;
}

int xo_obb_intersect(const xdem_shape* sa, const double* pa, const double* qa, const xdem_shape* sb, const double* pb,
                     const double* qb, double enlarge) {
  OBBo a = compute_obb(*sa, {pa[0], pa[1], pa[2]}, {qa[0], qa[1], qa[2], qa[3]}, 1.0);
  OBBo b = compute_obb(*sb, {pb[0], pb[1], pb[2]}, {qb[0], qb[1], qb[2], qb[3]}, 1.0);
  obb_enlarge(a, enlarge);
  obb_enlarge(b, enlarge);
  return obb_intersect(a, b) ? 1 : 0;
}

}  // extern "C"
