// xdem_internal.h — device data layout, math helpers and the context of libxdem_hip.so.
//
// Layout in HBM (all arrays in cell order along a Morton curve, rebuilt at every neighbour rebuild):
//   core[i]  one 128-byte record = one L2 line, everything a neighbour or the integrator touches:
//            xr = (rx, ry, rz, radius), vm = (vx, vy, vz, mass), w = (vrotx, vroty, vrotz, homothety),
//            q = (qw, qx, qy, qz)
//   a?/mom?/alpha?/in?/bk?  3 separate double arrays each (own particle only, coalesced)
// Neighbour list: per wavefront of 64 particles, an ELL block of K_w x 64 words
// (slot = base[w] + k*64 + lane), one 32-bit word per entry:
//   bits 0..24 partner index, 25..29 periodic image code (13 = none), 30 history-nonzero, 31 owner.
// Friction history lives in the same slots (3 double arrays), one copy per direction; both copies
// are computed in the owner's orientation and stay bit-identical.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/xdem.h"

namespace xdem {

constexpr uint32_t kIdxBits = 25;
constexpr uint32_t kIdxMask = (1u << kIdxBits) - 1u;
constexpr uint32_t kShiftShift = 25;
constexpr uint32_t kActiveBit = 1u << 30;
constexpr uint32_t kOwnerBit = 1u << 31;
constexpr int kNoShift = 13;
constexpr int kWave = 64;
// recorded contact state per interaction: dn, contact point (3), normal force f - ft (3), tangential
// force ft (3), owner position (3), partner position (3; the periodic image across a boundary, zero
// for walls)
constexpr int kCStateLen = 16;

struct alignas(32) d4 {
  double x, y, z, w;
};

struct V3 {
  double x, y, z;
};

__host__ __device__ inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
__host__ __device__ inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__host__ __device__ inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
__host__ __device__ inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline double norm(V3 a) { return sqrt(dot(a, a)); }

struct ContactParams {
  double kn, kt, kr, mu, damp_rate, fc, dncut, gamma;
};

// A stationary driver: plane (kind 0), cylinder (1) or ball (2), see xdem_surface.
struct Surface {
  V3 normal, center, vel, vrot, center_proj;
  double offset;
  int kind;
  double radius;
  V3 axis;  // cylinder: 0/1 mask of the coordinates the projection keeps
};

// Kinematic motion of a driver (host side; the force passes read the driver's state of each step):
// type 0 STATIONARY, 1 LINEAR_MOTION, 7 SHAKER (motion_type.hpp:23-35).
struct Motion {
  int type = 0;
  V3 mvec{0, 0, 0};
  double const_vel = 0, t_start = 0, t_end = 1e300, amplitude = 0, omega = 0;
  V3 shaker_dir{0, 0, 1};
};

// Interaction type of a driver contact (interaction_enum.hpp:36-38): 4 cylinder, 5 surface, 6 ball.
inline uint16_t driver_type(const Surface& s) { return s.kind == 1 ? 4 : (s.kind == 2 ? 6 : 5); }
// Candidate test of the interaction update for a point p (a sphere centre or a polyhedron vertex)
// with cut-off rc: Surface::filter (surface.hpp:320-324), Cylinder::filter (cylinder.hpp:225-234),
// Ball::filter (ball.hpp:292-296).
__host__ __device__ inline bool driver_filter(const Surface& s, V3 p, double rc) {
  if (s.kind == XDEM_DRIVER_CYLINDER) {
    const V3 proj{p.x * s.axis.x, p.y * s.axis.y, p.z * s.axis.z};
    const V3 dir = proj - s.center;
    const double d = norm(dir);
    const double dn = s.radius - (d + rc);
    return dn <= 0;
  }
  if (s.kind == XDEM_DRIVER_BALL) {
    const V3 dist = s.center - p;
    const double d = s.radius - norm(dist);
    return fabs(d) <= rc;
  }
  const V3 proj = dot(p, s.normal) * s.normal;
  const double d = norm(proj - s.center_proj);
  return d <= rc;
}

// Detection of a point p with radius rc against the driver: Surface::detector (surface.hpp:337-350),
// Cylinder::detector (cylinder.hpp:250-272), Ball::detector (ball.hpp:309-327). Returns the contact
// flag; dn is set in every case, n and the contact point cp on contact.
__host__ __device__ inline bool driver_detect(const Surface& s, V3 p, double rc, double& dn, V3& n, V3& cp) {
  if (s.kind == XDEM_DRIVER_CYLINDER) {
    const V3 proj{p.x * s.axis.x, p.y * s.axis.y, p.z * s.axis.z};
    const V3 dir = s.center - proj;
    const double d = norm(dir);
    dn = s.radius - (rc + d);
    if (dn > 0) return false;
    n = dir / d;
    cp = p - n * (rc + 0.5 * dn);
    return true;
  }
  if (s.kind == XDEM_DRIVER_BALL) {
    const V3 ptc = s.center - p;
    const double d = norm(ptc);
    n = ptc / d;
    if (d > s.radius) {
      dn = d - s.radius - rc;
      n = -1.0 * n;
    } else {
      dn = s.radius - d - rc;
    }
    if (dn > 0) return false;
    cp = p - n * (rc + 0.5 * dn);
    return true;
  }
  const V3 proj = dot(p, s.normal) * s.normal;
  const V3 s2p = -(s.center_proj - proj);
  const double d = norm(s2p);
  dn = d - rc;
  if (dn > 0) return false;
  n = s2p / d;
  cp = p - n * (rc + 0.5 * dn);
  return true;
}

// Domain and grid as seen by kernels.
struct Grid {
  double lo[3];
  double L[3];
  double h[3];
  double inv_h[3];
  int dims[3];
  int per[3];
  int64_t ncells;
};

// quat * vec3 (src/shape/include/exaDEM/type/quat.hpp:84-91); q stored as d4 {w, x, y, z} in {x, y, z, w}
__host__ __device__ inline V3 qrot4(const d4& q, V3 v) {
  const V3 qv{q.y, q.z, q.w};
  V3 uv = cross(qv, v);
  V3 uuv = cross(qv, uv);
  uv = uv * (2.0 * q.x);
  uuv = uuv * 2.0;
  return v + uv + uuv;
}

// shape::get_vertex (src/shape/include/exaDEM/shape.hpp:244-262): p + Q (h v_k)
__host__ __device__ inline V3 get_vertex(const xdem_shape& sh, int k, V3 p, double h, const d4& q) {
  const V3 v{sh.vertices[k][0], sh.vertices[k][1], sh.vertices[k][2]};
  return p + qrot4(q, h * v);
}

__host__ __device__ inline V3 shift_vec(const Grid& g, int code) {
  const int sx = code % 3 - 1, sy = (code / 3) % 3 - 1, sz = code / 9 - 1;
  return {sx * g.L[0], sy * g.L[1], sz * g.L[2]};
}

struct alignas(128) Core {
  d4 xr;
  d4 vm;
  d4 w;
  d4 q;
};
static_assert(sizeof(Core) == 128, "one particle record per 128-byte line");

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a contiguous range of particles so
// that neighbour gathers of consecutive blocks share that XCD's L2.
__device__ inline int64_t xcd_block(int64_t b, int64_t nb) {
  const int64_t q = nb >> 3, r = nb & 7;
  const int64_t x = b & 7, k = b >> 3;
  return x * q + (x < r ? x : r) + k;
}

// Partner records are read either from a block's LDS copy or from global memory. The two reads go
// through pointers typed with their address space: through plain (generic) pointers the compiler
// may fold the two branches into one load through a selected pointer, i.e. a flat load, which also
// counts on the LDS counter; every later wait for an LDS read then waits for the gathers in flight
// as well, which undoes the software pipeline of the contact loops.
typedef const double __attribute__((address_space(3)))* LdsF64;
typedef const double __attribute__((address_space(1)))* GlbF64;
__device__ inline d4 ld4(LdsF64 p) { return d4{p[0], p[1], p[2], p[3]}; }
__device__ inline d4 ld4(GlbF64 p) { return d4{p[0], p[1], p[2], p[3]}; }
__device__ inline void rec_fields(LdsF64 r, d4& p, d4& v, d4& w) {
  p = ld4(r);
  v = ld4(r + 4);
  w = ld4(r + 8);
}
__device__ inline void rec_fields(GlbF64 r, d4& p, d4& v, d4& w) {
  p = ld4(r);
  v = ld4(r + 4);
  w = ld4(r + 8);
}
// Partner y's position, velocity and spin in two halves: the prefetch reads partners outside the
// tile of records b0 .. b0+255 from `core` (global memory, issued one entry ahead), and the use reads
// partners inside the tile from LDS. Were both reads in the prefetch, the LDS read into the registers
// of a pending global load would wait for that load at once.
__device__ inline void partner_global(const Core* core, int64_t b0, uint32_t y, d4& p, d4& v, d4& w) {
  if ((uint64_t)((int64_t)y - b0) >= 256) rec_fields((GlbF64)(const double*)(core + y), p, v, w);
}
__device__ inline void partner_tile(const Core* tile, int64_t b0, uint32_t y, d4& p, d4& v, d4& w) {
  const uint64_t rel = (uint64_t)((int64_t)y - b0);
  if (rel < 256) rec_fields((LdsF64)(const double*)(tile + rel), p, v, w);
}
// Copies records b0 .. b0+255 into the tile with coalesced 16-byte loads, all eight issued before
// the first LDS write; indexes past n read record n-1 (never used).
__device__ inline void stage_tile(Core* tile, const Core* core, int64_t b0, int64_t n) {
  uint4 v[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int k = (int)threadIdx.x + 256 * q;
    const int64_t src = b0 + (k >> 3) < n ? b0 + (k >> 3) : n - 1;
    v[q] = ((const uint4*)(core + src))[k & 7];
  }
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int k = (int)threadIdx.x + 256 * q;
    ((uint4*)(tile + (k >> 3)))[k & 7] = v[q];
  }
}

// Device particle state (one buffer set; two sets are kept and swapped at each reorder).
struct PState {
  Core* core;
  double *ax, *ay, *az;
  double *mx, *my, *mz;
  double *alx, *aly, *alz;
  double *ix, *iy, *iz;
  double *bx, *by, *bz;
  d4* bq;  // orientation at the last rebuild (vertex displacement trigger of polyhedra)
  uint64_t* id;
  uint32_t* type;
  uint32_t* group;
};

// Neighbour list + per-slot history.
struct NbrList {
  uint32_t* word = nullptr;   // [slots]
  uint32_t* sx = nullptr;     // [slots] particle of each slot (~0u for the padding of a wave's ELL block)
  double* hx = nullptr;       // friction history [slots]
  double* hy = nullptr;
  double* hz = nullptr;
  double* mhx = nullptr;      // rolling-moment history [slots] (allocated when kr != 0 is used)
  double* mhy = nullptr;
  double* mhz = nullptr;
  float* fx32 = nullptr;      // friction history of the FP32-force form [slots]
  float* fy32 = nullptr;
  float* fz32 = nullptr;
  float* mfx32 = nullptr;
  float* mfy32 = nullptr;
  float* mfz32 = nullptr;
  uint32_t* cnt = nullptr;    // entries per particle [n]
  uint32_t* kmax = nullptr;   // per wave max count [nwaves]
  uint64_t* base = nullptr;   // per wave first slot [nwaves+1]
  int64_t slots = 0;
  int64_t cap_slots = 0;
  int64_t pairs = 0;          // half-list pairs (owner entries)
};

// Wall (Surface driver) interactions: candidate mask per particle, history per (wall, particle).
struct WallList {
  uint8_t* mask = nullptr;  // [n] bit w set if wall w is a candidate
  double* hx = nullptr;     // [nw*n]
  double* hy = nullptr;
  double* hz = nullptr;
  double* mhx = nullptr;
  double* mhy = nullptr;
  double* mhz = nullptr;
};

// Polyhedron feature interactions, grouped by owner list slot (CSR over the ELL slots):
// features of slot s are [start[s], start[s+1]); one 32-bit word each:
//   bits 0..1 type (0 VV, 1 VE, 2 VF, 3 EE), bit 2 swap, bits 3..6 owner sub, bits 7..10 partner sub,
//   bit 30 history non-zero. Friction history in three FP64 arrays (moment history alongside).
struct PolyList {
  uint64_t* start = nullptr;  // [cap_slots + 1]
  uint32_t* feat = nullptr;   // [cap_feat]
  double *hx = nullptr, *hy = nullptr, *hz = nullptr;
  double *mhx = nullptr, *mhy = nullptr, *mhz = nullptr;
  uint32_t* fcount = nullptr;  // [cap_slots] scratch: features per slot
  uint32_t* fslot = nullptr;   // [cap_feat] list slot of each feature (owner entry)
  uint32_t* fown = nullptr;    // [cap_feat] owner particle of each feature
  int64_t cap_slots = 0;
  int64_t cap_feat = 0;
  int64_t nfeat = 0;
};
// Vertex-wall candidates: per particle a 64-bit mask (bit w*8 + vertex), history per (wall, vertex, particle).
struct PolyWall {
  uint64_t* mask = nullptr;
  double *hx = nullptr, *hy = nullptr, *hz = nullptr;
  double *mhx = nullptr, *mhy = nullptr, *mhz = nullptr;
  int64_t cap = 0;
};

struct KernelTimer {
  std::vector<hipEvent_t> ev;  // pairs (start, stop)
  int64_t count = 0;
  double total_ms = 0;
};

}  // namespace xdem

namespace xdem {
// Stationary triangle-mesh driver (register_stl_mesh, src/driver/register_stl_mesh.cpp:30-140): world
// vertices, edges (two per entry, shared edges listed once per face as build_shape lists them), triangles,
// and a uniform grid of bins over the mesh: bin b lists, as words sorted ascending, the features whose
// bounding box enlarged by `reach` overlaps it. Word: kind << 30 | index, kind 0 vertex (interaction
// type 7), 1 edge (8), 2 face (9).
struct MeshView {
  const V3* v = nullptr;
  const uint2* e = nullptr;
  const uint4* f = nullptr;
  double rm = 0;  // Minkowski radius
  V3 center{0, 0, 0};
  V3 lo{0, 0, 0};
  double inv_h = 0;
  int dim[3] = {0, 0, 0};
  const uint32_t* start = nullptr;
  const uint32_t* feat = nullptr;
  int cap = 0;  // candidate slots per particle (0: no mesh)
  int64_t stride = 0;
};
// Per particle candidates: slot k of particle x at k * stride + x; cnt may exceed the capacity after a
// failed build (readers clamp).
struct MeshList {
  uint32_t* cnt = nullptr;
  uint32_t* word = nullptr;
  double *hx = nullptr, *hy = nullptr, *hz = nullptr, *mhx = nullptr, *mhy = nullptr, *mhz = nullptr;
};
// Mesh contacts of a force pass, evaluated one candidate per thread just before it (k_mesh_pass): the
// candidates of all particles in one flat array, each particle's contiguous and in rank order; the
// force and moment terms of candidate t at t * 6 (-0.0 when not in contact), which the force pass adds in
// candidate order. Spheres: each candidate's feature geometry and kind cached at geo + t * 10 (three
// points, then the kind), since the mesh is stationary.
struct MeshTerms {
  const uint2* cand = nullptr;  // per candidate t: (particle x, rank k among x's candidates)
  const uint2* idx = nullptr;   // per particle: (first candidate t, count); count 0: none
  const uint32_t* wcnt = nullptr;  // per wave of 64 particles (x >> 6): its candidates
  double* terms = nullptr;      // 6 per candidate: force, then moment
  const double* geo = nullptr;  // spheres: 10 per candidate (null for polyhedra)
  int64_t ncand = 0;
};
}  // namespace xdem

struct xdem_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  int64_t n = 0;
  int64_t cap = 0;
  xdem::PState ps[2] = {};
  int cur = 0;  // index of the live particle buffer set
  bool have_domain = false;
  xdem::Grid grid{};
  int nwalls = 0;
  xdem::Surface walls[XDEM_MAX_SURFACES];
  xdem::Surface* d_walls = nullptr;
  // driver motion: per driver its motion, the drivers at the last neighbour build (backup_drivers), and
  // for the engine a ring of per-step driver states (pinned host + device, one set of
  // XDEM_MAX_SURFACES per step of a poll chunk); d_walls_pass, when set, is the state the next force
  // pass reads
  xdem::Motion motion[XDEM_MAX_SURFACES];
  bool drivers_move = false;  // some driver has a motion other than STATIONARY
  xdem::Surface walls_bk[XDEM_MAX_SURFACES];
  double physical_time = 0;
  xdem::Surface* h_walls_ring = nullptr;
  xdem::Surface* d_walls_ring = nullptr;
  int64_t cap_ring = 0;  // steps
  const xdem::Surface* d_walls_pass = nullptr;

  // binning scratch
  uint32_t* cell_of = nullptr;      // [cap]
  uint32_t* cell_count = nullptr;   // [ncells]
  uint32_t* cell_start = nullptr;   // [ncells+1], indexed by Morton rank of the cell
  uint32_t* cell_rank = nullptr;    // [ncells] rank of each cell along the Morton curve
  uint64_t* d_rmax = nullptr;       // largest radius as IEEE bits, reduced by the binning pass
  xdem::d4* mxr = nullptr;          // [cap] positions and radii in Morton order (binning -> neighbour build)
  uint64_t* mid = nullptr;          // [cap] ids in Morton order
  int64_t cap_morton = 0;
  uint32_t* blk_u32 = nullptr;      // per-block partial counts
  int64_t cap_blk = 0;
  uint32_t* nbh_tmp = nullptr;      // candidate rows of the neighbour build, nbh_tmp_cap per particle
  uint32_t nbh_tmp_cap = 0;
  int64_t nbh_tmp_n = 0;            // particles the rows were allocated for
  uint32_t* perm = nullptr;         // [cap] new index -> old index (last reorder)
  uint32_t* perm_tmp = nullptr;     // [cap] scratch for composing permutations
  uint64_t* sort_key = nullptr;     // [cap]
  int64_t cap_cells = 0;
  void* scan_tmp = nullptr;
  int64_t scan_tmp_bytes = 0;

  // neighbour lists (two: current and previous for history carry-over)
  xdem::NbrList nl[2];
  int nl_cur = 0;
  bool nl_valid = false;       // a list exists for the current particle order
  bool nl_prev_valid = false;  // previous list is aligned with perm
  bool perm_valid = false;
  bool cells_valid = false;     // cell_start matches the current positions and order
  bool hist_ready = false;      // history of the current list has been carried over (nbh_sphere)
  uint32_t max_type = 0;        // largest particle type of the last upload
  bool ids_dense = false;       // the uploaded ids are a permutation of 0..n-1
  uint64_t nl_epoch = 0;        // number of neighbour builds so far
  uint64_t pl_epoch = ~0ull;    // neighbour build the feature lists were made from
  bool wl_valid = false;        // wall list exists (for the order it was built in)
  bool kr_used = false;         // a rolling stiffness was used: moment histories are live
  bool hist_fp32 = false;       // the live history of the list is the FP32 copy
  int nl_max_rows = 0;          // largest per-wavefront row count of the current list
  xdem::Core* core_alt = nullptr;  // second record buffer of the integrating force pass (ping-pong)
  int32_t trig_slot = 0;        // d_flags slot of the pending rebuild trigger (0, or 3 after an integrating pass)
  // contact state recorded by FP64 sphere force passes outside the integrating engine step (analysis
  // dumps, xdem_set_contact_record): per list slot and per (wall, particle) kCStateLen doubles
  bool record_cs = false;
  bool cs_valid = false;        // cstate/wstate hold the last force pass of list epoch cs_epoch
  uint64_t cs_epoch = 0;
  double* cstate = nullptr;
  int64_t cap_cstate = 0;       // slots
  double* wstate = nullptr;
  int64_t cap_wstate = 0;       // (wall, particle) entries
  double* simpart = nullptr;    // block partials of xdem_simulation_state
  int64_t cap_simpart = 0;
  double rcut_inc = 0;
  double nbh_dist = 0;
  double rmax_host = -1;            // largest radius of the uploaded particles (-1: none yet)
  uint64_t* d_pairs = nullptr;      // [2] half-list pair count of each neighbour list (read on demand)
  xdem::WallList wl[2];
  int wl_cur = 0;
  int wl_nw = 0;  // wall count used to size wl

  // rebuild trigger / stall flags on device: [0] trigger, [1] stall, [2] stall step (low), [3] pad
  int32_t* d_flags = nullptr;
  int32_t* h_flags = nullptr;  // pinned
  uint32_t* d_u32tmp = nullptr;
  uint32_t* h_u32tmp = nullptr;  // pinned

  // fused-engine state: whether the epilog kicks of the previous force pass are pending
  bool pending_kick = false;
  double pending_dt = 0;

  bool timing = false;
  int32_t timing_every = 1;  // time the launches of every timing_every-th step
  xdem::KernelTimer timers[2];

  // polyhedra (set when shapes are registered)
  int nshapes = 0;
  xdem_shape shapes[XDEM_MAX_SHAPES];
  xdem_shape* d_shapes = nullptr;
  xdem::d4* vtx = nullptr;          // [cap * 8] world vertices (x, y, z, pad)
  int64_t cap_vtx = 0;
  xdem::PolyList pl[2];             // feature lists, current and previous
  int pl_cur = 0;
  bool pl_valid = false;
  double* fcontrib = nullptr;       // [cap_contrib * 9] per feature: F_owner, M_owner, M_partner terms
  uint8_t* fcontact = nullptr;      // [cap_contrib] 1 when the feature is in contact
  double* fdn = nullptr;            // [cap_contrib] dn of the feature at the last pass (0 without contact)
  uint64_t poly_pass_epoch = ~0ull; // neighbour build of the last polyhedron force pass
  int64_t cap_contrib = 0;
  uint32_t* feat_tmp = nullptr;     // [cap_feat_tmp] per list slot, up to kFeatPerSlot feature words of the
  int64_t cap_feat_tmp = 0;         // one-pass enumeration (nbh_polyhedron)
  uint32_t* oslot = nullptr;        // [cap_oslot] owner slots of the list (nbh_polyhedron)
  uint32_t* feat_ovf = nullptr;     // [cap_oslot] owner slots whose features are enumerated again
  uint32_t* feat_fpool = nullptr;   // [cap_oslot] per slot: its chunk of feat_pool (features past the tmp region)
  uint32_t* feat_pool = nullptr;    // [cap_feat_pool]
  int64_t cap_feat_pool = 0;
  int64_t cap_oslot = 0;
  uint32_t* d_nown = nullptr;       // their number
  uint2* frange = nullptr;          // [cap_frange] per list slot: first feature and count of its pair
                                    // (through the owner entry for partner entries); x = ~0u: no owner entry
  int64_t cap_frange = 0;
  uint4* forder = nullptr;          // [cap_forder] classifier waves by type (feature, owner, list word,
                                    // feature word), padded
  int64_t cap_forder = 0;
  int64_t fwave_start[5] = {0, 0, 0, 0, 0};  // padded start of the VV, VE, VF, EE waves; [4] padded total
  int64_t fwave_size[4] = {0, 0, 0, 0};      // features per wave
  uint64_t pl_class_epoch = ~0ull;  // neighbour build the waves were classified from
  uint32_t* fclass_blk = nullptr;   // per block of 1024 features and type: count, then offset
  int64_t cap_fclass_blk = 0;
  int64_t* fclass_info = nullptr;   // [9] wave starts, padded total, wave sizes
  xdem::PolyWall pw[2];             // vertex-wall candidates and history, current and previous
  int pw_cur = 0;
  bool pw_valid = false;

  bool has_mesh = false;            // a triangle-mesh driver is registered
  int32_t mesh_id = 0;              // its driver index (partner id of its interactions)
  int mesh_cap = 0;                 // candidate slots per particle
  xdem::MeshView mesh{};            // device view (bins filled at the neighbour build)
  std::vector<xdem::V3> mesh_hv;    // host copies for binning
  std::vector<uint2> mesh_he;
  std::vector<uint4> mesh_hf;
  double mesh_reach = -1;           // reach the bins were built for (-1: none)
  double mesh_rv = -1;              // rcut_inc the bins were built for
  xdem::V3* d_mesh_v = nullptr;
  uint2* d_mesh_e = nullptr;
  uint4* d_mesh_f = nullptr;
  uint32_t* d_mesh_start = nullptr;
  uint32_t* d_mesh_feat = nullptr;
  xdem::MeshList ml[2];             // candidates and history, current and previous
  int ml_cur = 0;
  bool ml_valid = false;
  int64_t ml_alloc = 0;             // particle capacity ml was allocated for
  int32_t* d_mesh_over = nullptr;   // [0..1] max radius bits, [2] largest overflowing count
  // The candidates of the current lists in one flat array, each particle's candidates contiguous and in
  // rank order (MeshTerms): cand and terms have room for cap_mesh_cand candidates, idx for the particle
  // capacity; mesh_ncand candidates in all
  uint2* mesh_cand = nullptr;
  uint2* mesh_idx = nullptr;
  uint32_t* mesh_wcnt = nullptr;    // candidates per wave of 64 particles (MeshTerms::wcnt)
  uint32_t* mesh_cc = nullptr;      // [cap] clamped candidate counts, scanned into mesh_off [cap + 1]
  uint32_t* mesh_off = nullptr;
  double* mesh_terms = nullptr;
  double* mesh_geo = nullptr;       // spheres: feature geometry per candidate (MeshTerms::geo)
  bool mesh_poly = false;           // the lists hold polyhedron words (built by nbh_polyhedron)
  int64_t cap_mesh_cand = 0;
  int64_t mesh_ncand = 0;
  // compact copy of the current candidate words and histories at the flat index t, which the mesh passes
  // read and write; mc_valid: gathered from the lists since their last change; mc_dirty: newer than the
  // lists (mesh_sync writes them back)
  xdem::MeshList mc;
  bool mc_valid = false;
  bool mc_dirty = false;
  double* mstate = nullptr;         // recorded contact state per mesh slot (kCStateLen doubles each)
  int64_t cap_mstate = 0;
};

// helpers shared between translation units
namespace xdem {
int set_err(xdem_ctx* c, int code, const std::string& msg);
int hip_check(xdem_ctx* c, hipError_t e, const char* what);
#define XDEM_HIP(c, call)                               \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) return xdem::hip_check(c, e_, #call); \
  } while (0)

inline PState& live(xdem_ctx* c) { return c->ps[c->cur]; }
// drivers read by a force pass: the engine's state of that step, else the current drivers
inline const Surface* walls_for_pass(const xdem_ctx* c) { return c->d_walls_pass ? c->d_walls_pass : c->d_walls; }
// driver motion of one step on the host (push_f_v_r then push_f_v) and the driver trigger
void drivers_advance(Surface* w, const Motion* m, int nw, double time, double dt);
void drivers_push_f_v(Surface* w, const Motion* m, int nw);
bool drivers_displaced(const Surface* now, const Surface* bk, int nw, double thr);
int ensure_capacity(xdem_ctx* c, int64_t n);
int ensure_cells(xdem_ctx* c, int64_t ncells);
int scan_u32(xdem_ctx* c, const uint32_t* in, uint32_t* out, int64_t n);  // exclusive, out[n] = total
int scan_u64_from_u32(xdem_ctx* c, const uint32_t* in, uint64_t* out, int64_t n, uint32_t mul);
int read_u64(xdem_ctx* c, const uint64_t* d, uint64_t* h);

inline int grid_blocks(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

// (Re)allocates a device array of count elements; the old contents are dropped.
template <typename T>
int dev_realloc(xdem_ctx* c, T** p, int64_t count) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  XDEM_HIP(c, hipMalloc((void**)p, sizeof(T) * (size_t)(count > 1 ? count : 1)));
  return XDEM_OK;
}

// kernels launched from api code
int launch_prolog(xdem_ctx* c, double dt, bool with_kick, double kick_dt, bool displ, double thr2, bool guarded);
int launch_epilog(xdem_ctx* c, double dt);
int launch_kick(xdem_ctx* c, double dt);
int launch_reset(xdem_ctx* c);
int launch_gravity(xdem_ctx* c, V3 g);
int launch_to_accel(xdem_ctx* c);
int launch_displ(xdem_ctx* c, double thr2);
int do_move_particles(xdem_ctx* c);
int do_chunk_neighbors(xdem_ctx* c, double rcut_inc, double nbh_dist);
// Half-list pairs of the current neighbour list (the device count is read the first time it is asked for).
int nl_pairs(xdem_ctx* c, int64_t* out);
int do_nbh_sphere(xdem_ctx* c, double rcut_inc);
int launch_force(xdem_ctx* c, const ContactParams& cp, const ContactParams& cpd, double dt, int form, bool fused,
                 bool gravity_on, V3 g, bool guarded, int32_t step_tag);
// engine step of spheres in one kernel (force pass + kick + next prolog into c->core_alt, which
// becomes live), FP64 or FP32 form; *input = the record buffer the pass read
int launch_step(xdem_ctx* c, const ContactParams& cp, const ContactParams& cpd, double dt, int form, bool gravity_on,
                V3 g, double thr2, int32_t step_tag, Core** input);
int alloc_walls(xdem_ctx* c, int nw);
// mesh candidates + history carry-over (poly: the polyhedron candidates of the particles' vertices, edges
// and faces, built by launch_poly_mesh_carry)
int mesh_carry(xdem_ctx* c, double rcut_inc, bool have_old, bool poly = false);
int launch_poly_mesh_carry(xdem_ctx* c, double rv, const MeshList& nl, const MeshList& ol, bool have_old,
                           int32_t* over);
void mesh_free(xdem_ctx* c);
// Mesh contacts of the coming force pass (k_mesh_pass over the live records): histories updated and the
// terms written for the pass to add; guarded: a no-op when the pass would be (stall or trigger in
// flags[trig_in]); cs: recorded contact state per mesh slot (null: not recorded). *t receives the view
// the force pass reads (t->ncand = 0: no mesh contacts to add).
int launch_mesh_pass(xdem_ctx* c, const ContactParams& cpd, double dt, bool guarded, int32_t trig_in, double* cs,
                     MeshTerms* t);
// The compact copy the passes use (gathered from the lists when they changed), and the write-back of
// its histories into the lists before anything else reads or replaces them.
int mesh_compact(xdem_ctx* c);
int mesh_sync(xdem_ctx* c);
int set_history_form(xdem_ctx* c, bool fp32);
// a rolling stiffness comes into use (clears the sphere list's moment histories the first time)
int use_kr(xdem_ctx* c);
int engine_rebuild(xdem_ctx* c, double rcut_inc);
int engine_force_pass(xdem_ctx* c, const xdem_scheme* s);  // fused force pass at the current state
int engine_flush_kick(xdem_ctx* c);
// polyhedra
int poly_compute_vertices(xdem_ctx* c, bool displ, double thr2, bool guarded);
int poly_vtx_alloc(xdem_ctx* c);
// The prolog of polyhedra fused with compute_vertices and the vertex displacement test (k_prolog_verts).
int launch_prolog_verts(xdem_ctx* c, double dt, bool with_kick, double kick_dt, double thr2, bool guarded);
int poly_nbh(xdem_ctx* c, double rcut_inc);
// classify_interactions of polyhedra: the feature waves by type (forder, fwave_start, fwave_size)
int poly_classify(xdem_ctx* c);
int poly_force(xdem_ctx* c, const ContactParams& cp, const ContactParams& cpd, double dt, bool fused, bool gravity_on,
               V3 g, bool guarded, int32_t step_tag);
// interaction counts of the polyhedron lists for simulation_state: total, active, smallest dn; block
// partials in the library's buffer (fixed = null) or in the fixed buffer of fixed_cap doubles
int poly_sim_counts(xdem_ctx* c, double* tot, double* act, double* min_dn, double* fixed, int64_t fixed_cap);
// folds nblocks x nvals block partials in block order into out_dev[0..nvals) on the device (sums, the
// last value a minimum), as xdem_simulation_state folds them on the host
int fold_partials(xdem_ctx* c, const double* part, int64_t nblocks, int nvals, double* out_dev);
// device address of this GPU's copy of the no-allocation scratch array (once per device, device current)
int noalloc_scratch_init(int device);
int poly_contact_state_rows(xdem_ctx* c, std::vector<uint64_t>& ii, std::vector<uint64_t>& jj,
                            std::vector<uint32_t>& si, std::vector<uint32_t>& sj, std::vector<uint16_t>& tt,
                            std::vector<double>& vals);
}  // namespace xdem
