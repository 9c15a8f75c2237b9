// deck.cpp — the operator graph of exaNBody decks over the C ABI of libxdem_hip.so.
//
// Execution follows the deck semantics of the reference (onika operator graph, as used by
// data/config/main-config.msp and the example decks): every top-level key of the merged deck names
// an operator; a list is a batch run in order, whose items are names or {name: slot values}; a
// scalar is an alias of another operator ("nop" does nothing); a mapping with a "body" is a batch
// (with "loop: true" it repeats while its condition slot holds). Leaf operators are native: their
// slot values come from the deck's mapping of the same name, overridden by the call site.
//
// Native operators and the reference operator each one replaces:
//   input      particle_type, set_fields (src/field/set_fields.cu), init_rsa (src/rsa/init_rsa.cpp),
//              read_xyz (example/spheres/movable-wall input format), polydisperse_pack (this library's
//              seeded RSA + relaxation packing of the benchmark inputs)
//   drivers    register_surface, register_cylinder, register_ball, register_stl_mesh (src/driver/register_*.cpp;
//              STATIONARY, LINEAR_MOTION and SHAKER motions, a spinning cylinder or ball through vrot);
//              push_f_v_r_driver, push_f_v_driver, driver_displ_over in check_and_update_particles
//   scheme     combined_compute_prolog, combined_compute_epilog, reset_force_moment, gravity_force,
//              force_to_accel, particle_displ_over, vertex_displ_over, compute_vertices
//   lists      migrate_cell_particles_interaction / move_particles (with backup_r),
//              chunk_neighbors_contact, nbh_sphere, classify_interactions, nbh_polyhedron
//   forces     contact_sphere and its generated aliases (src/sphere/CMakeLists.txt:19-123),
//              contact_polyhedron and aliases (src/polyhedron/CMakeLists.txt:19-122)
//   control    check_and_update_particles (trigger then update_particles_full), init_particles,
//              compute_loop (numerical_scheme until simulation_end_iteration)
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <algorithm>
#include <array>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/xdem_host.h"
#include "msp.hpp"

using xdem_host::Node;

namespace {

struct Fail : std::runtime_error {
  int code;
  Fail(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

struct HostParticles {
  std::vector<double> rx, ry, rz, vx, vy, vz, radius, mass, inx, iny, inz, qw, qx, qy, qz, wx, wy, wz, h;
  std::vector<uint64_t> id;
  std::vector<uint32_t> type, group;
  size_t size() const { return rx.size(); }
  void add(double x, double y, double z, uint32_t t, double r) {
    const uint64_t k = rx.size();
    rx.push_back(x); ry.push_back(y); rz.push_back(z);
    vx.push_back(0); vy.push_back(0); vz.push_back(0);
    radius.push_back(r); mass.push_back(1); inx.push_back(1); iny.push_back(1); inz.push_back(1);
    qw.push_back(1); qx.push_back(0); qy.push_back(0); qz.push_back(0);
    wx.push_back(0); wy.push_back(0); wz.push_back(0); h.push_back(1);
    id.push_back(k); type.push_back(t); group.push_back(0);
  }
};

// operators outside the hot path that the runner recognises and skips
const std::set<std::string> kSkipped = {
    "message", "print_logo_banner", "hw_device_init", "hw_device_finalize", "mpi_comm_world", "init_cuda",
    "finalize_cuda", "io_tree", "io_config", "dump_data_paraview", "final_dump", "default_thermodynamic_state",
    "default_print_thermodynamic_state", "print_simulation_state", "dt_critical",
    "start_iteration", "end_iteration", "update_stress_tensor", "freeze", "reine_des_neiges",
    "push_av_to_quat_driver", "reset_force_driver", "force_to_accel_driver",
    "init_drivers", "backup_drivers", "driver_vertices", "init_particle_regions", "grid_flavor", "simulation_epilog",
    "compute_loop_prolog", "compute_loop_epilog", "update_interaction_ghost", "rebuild_amr", "backup_r",
    "dem_rcut_max", "nbh_dist", "check_rcut", "reduce_species_after_read", "cost_model", "save_shapes"};

class Runner {
 public:
  Node deck;
  std::string error;
  xdem_deck_stats stats{};

  Runner() { register_natives(); }

  void run(xdem_ctx* c) {
    ctx_ = c;
    stats = {};
    timestep_ = (int64_t)gdouble("timestep", 0);
    const auto t0 = std::chrono::steady_clock::now();
    exec("simulation", nullptr, 0);
    stats.particles = (int64_t)P_.size();
    stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

 private:
  using Fn = std::function<void(const Node&)>;
  std::map<std::string, Fn> natives_;
  xdem_ctx* ctx_ = nullptr;
  HostParticles P_;
  std::vector<std::string> type_names_;
  std::vector<xdem_surface> walls_;
  std::vector<xdem_driver_motion> motions_;  // params slot of each register_* (parallel to walls_)
  bool uploaded_ = false, poly_ = false;
  double bmin_[3] = {0, 0, 0}, bmax_[3] = {0, 0, 0};
  bool have_bounds_ = false;
  uint64_t seed_ = 0;
  int64_t timestep_ = 0;  // the deck's timestep slot: 0 in first_iteration, k in the k-th loop iteration

  // ---------------------------------------------------------------- helpers
  [[noreturn]] void fail(int code, const std::string& m) { throw Fail(code, m); }
  void check(int rc, const char* what) {
    if (rc == XDEM_OK) return;
    char buf[512];
    xdem_last_error(ctx_, buf, sizeof(buf));
    fail(rc, std::string(what) + ": " + buf);
  }
  const Node& global() {
    const Node* g = deck.get("global");
    if (!g) fail(XDEM_ERR_ARG, "deck has no 'global' block");
    return *g;
  }
  double gdouble(const char* k, double dflt) {
    const Node* v = global().get(k);
    return v ? xdem_host::to_double(*v) : dflt;
  }
  static const Node* slot(const Node& s, const char* k) { return s.is_map() ? s.get(k) : nullptr; }
  static double sdouble(const Node& s, const char* k, double dflt) {
    const Node* v = slot(s, k);
    return v ? xdem_host::to_double(*v) : dflt;
  }
  uint32_t type_index(const Node& n) {
    if (n.is_scalar()) {
      for (size_t k = 0; k < type_names_.size(); k++)
        if (type_names_[k] == n.s) return (uint32_t)k;
      char* e = nullptr;
      const long v = std::strtol(n.s.c_str(), &e, 10);
      if (e && *e == 0 && v >= 0) return (uint32_t)v;
    }
    fail(XDEM_ERR_ARG, "unknown particle type '" + xdem_host::dump(n) + "'");
  }
  double rcut_inc() { return gdouble("rcut_inc", 0.0); }
  double dt() { return gdouble("dt", 1e-4); }

  xdem_contact_params contact_params(const Node& c) {
    xdem_contact_params p{};
    p.kn = sdouble(c, "kn", 0);
    p.kt = sdouble(c, "kt", 0);
    p.kr = sdouble(c, "kr", 0);
    p.mu = sdouble(c, "mu", 0);
    p.damp_rate = sdouble(c, "damp_rate", 0);
    p.fc = sdouble(c, "fc", 0);
    p.dncut = sdouble(c, "dncut", 0);
    p.gamma = sdouble(c, "gamma", 0);
    return p;
  }

  // ---------------------------------------------------------------- graph execution
  void exec(const std::string& name, const Node* params, int depth) {
    if (depth > 64) fail(XDEM_ERR_ARG, "operator graph too deep at '" + name + "'");
    auto nat = natives_.find(name);
    const Node* def = deck.get(name);
    if (nat != natives_.end()) {
      Node slots;
      slots.kind = Node::Map;
      if (def && def->is_map() && !def->get("body")) xdem_host::merge_into(slots, *def);
      if (params && params->is_map()) xdem_host::merge_into(slots, *params);
      stats.operator_calls++;
      nat->second(slots);
      return;
    }
    if (!def) {
      if (kSkipped.count(name)) return;
      fail(XDEM_ERR_UNSUPPORTED, "operator '" + name + "' is not provided by this library");
    }
    if (def->is_scalar()) {
      if (def->s == "nop") return;
      exec(def->s, params, depth + 1);
      return;
    }
    if (def->is_seq()) {
      batch(*def, depth);
      return;
    }
    if (def->is_map() && def->get("body")) {
      const Node& body = *def->get("body");
      const Node* loop = def->get("loop");
      if (loop && xdem_host::to_bool(*loop)) fail(XDEM_ERR_UNSUPPORTED, "generic loop operator '" + name + "'");
      if (body.is_seq()) batch(body, depth);
      else if (body.is_scalar()) exec(body.s, nullptr, depth + 1);
      return;
    }
    if (kSkipped.count(name)) return;
    fail(XDEM_ERR_UNSUPPORTED, "operator '" + name + "' is not provided by this library");
  }

  void batch(const Node& list, int depth) {
    for (const Node& it : list.seq) {
      if (it.is_scalar()) exec(it.s, nullptr, depth + 1);
      else if (it.is_map() && it.map.size() == 1) exec(it.map[0].first, &it.map[0].second, depth + 1);
      else fail(XDEM_ERR_ARG, "bad batch item at line " + std::to_string(it.line));
    }
  }

  // ---------------------------------------------------------------- natives
  void register_natives() {
    auto& N = natives_;
    // ---- input
    N["particle_type"] = [this](const Node& s) {
      const Node* t = slot(s, "type");
      if (!t || !t->is_seq()) fail(XDEM_ERR_ARG, "particle_type: 'type' must be a list of names");
      type_names_.clear();
      for (auto& x : t->seq) type_names_.push_back(x.s);
    };
    N["init_rsa"] = [this](const Node& s) { init_rsa(s); };
    N["read_xyz"] = [this](const Node& s) { read_xyz(s); };
    N["polydisperse_pack"] = [this](const Node& s) { polydisperse_pack(s); };
    N["set_fields"] = [this](const Node& s) { set_fields(s); };
    // ---- drivers
    N["register_surface"] = [this](const Node& s) { register_surface(s); };
    // driver motion (config_numerical_schemes.msp:13-24): push_f_v_r_driver at the physical time of the
    // step, push_f_v_driver; dt_scale multiplies dt as in the reference operators
    N["push_f_v_r_driver"] = [this](const Node& s) {
      double t = 0;
      check(xdem_physical_time(ctx_, &t), "push_f_v_r_driver");
      check(xdem_push_f_v_r_driver(ctx_, t, dt() * sdouble(s, "dt_scale", 1.0)), "push_f_v_r_driver");
    };
    N["push_f_v_driver"] = [this](const Node& s) {
      check(xdem_push_f_v_driver(ctx_, dt() * sdouble(s, "dt_scale", 1.0)), "push_f_v_driver");
    };
    N["next_time_step"] = [this](const Node&) { check(xdem_next_time_step(ctx_, dt()), "next_time_step"); };
    N["register_cylinder"] = [this](const Node& s) { register_round(s, XDEM_DRIVER_CYLINDER, "register_cylinder"); };
    N["register_ball"] = [this](const Node& s) { register_round(s, XDEM_DRIVER_BALL, "register_ball"); };
    N["register_stl_mesh"] = [this](const Node& s) { register_stl_mesh(s); };
    // ---- scheme operators
    N["combined_compute_prolog"] = [this](const Node&) { check(xdem_combined_compute_prolog(ctx_, dt()), "prolog"); };
    N["combined_compute_epilog"] = [this](const Node&) { check(xdem_combined_compute_epilog(ctx_, dt()), "epilog"); };
    N["reset_force_moment"] = [this](const Node&) { check(xdem_reset_force_moment(ctx_), "reset_force_moment"); };
    N["gravity_force"] = [this](const Node& s) {
      double g[3] = {0.0, 0.0, -9.807};  // default of the reference operator (src/forcefield/gravity_force.cu:40)
      if (const Node* v = slot(s, "gravity")) {
        const auto d = xdem_host::to_doubles(*v);
        if (d.size() != 3) fail(XDEM_ERR_ARG, "gravity_force: gravity needs 3 values");
        for (int a = 0; a < 3; a++) g[a] = d[a];
      }
      check(xdem_gravity_force(ctx_, g), "gravity_force");
    };
    N["force_to_accel"] = [this](const Node&) { check(xdem_force_to_accel(ctx_), "force_to_accel"); };
    N["compute_vertices"] = [this](const Node&) {
      if (poly_) check(xdem_compute_vertices(ctx_), "compute_vertices");
    };
    // ---- lists
    auto move = [this](const Node&) { check(xdem_move_particles(ctx_), "move_particles"); };
    N["move_particles"] = move;
    N["migrate_cell_particles_interaction"] = move;
    N["chunk_neighbors_contact"] = [this](const Node&) {
      check(xdem_chunk_neighbors_contact(ctx_, rcut_inc(), -1.0), "chunk_neighbors_contact");
    };
    N["nbh_sphere"] = [this](const Node& s) {
      const Node* sym = slot(s, "symetric");
      check(xdem_nbh_sphere(ctx_, rcut_inc(), sym ? (xdem_host::to_bool(*sym) ? 1 : 0) : 1), "nbh_sphere");
    };
    N["classify_interactions"] = [this](const Node&) { check(xdem_classify_interactions(ctx_), "classify_interactions"); };
    N["nbh_polyhedron"] = [this](const Node&) { check(xdem_nbh_polyhedron(ctx_, rcut_inc()), "nbh_polyhedron"); };
    // ---- forces
    const char* sphere_names[] = {"contact_sphere", "contact_singlemat_sphere", "contact_singlemat_sphere_hooke",
                                  "contact_singlemat_sphere_hooke_none", "contact_singlemat_sphere_hooke_cohesive",
                                  "contact_singlemat_sphere_hooke_dmt"};
    for (const char* nm : sphere_names) N[nm] = [this](const Node& s) { contact(s, false); };
    const char* poly_names[] = {"contact_polyhedron", "contact_singlemat_polyhedron",
                                "contact_singlemat_polyhedron_hooke", "contact_singlemat_polyhedron_hooke_none",
                                "contact_singlemat_polyhedron_hooke_cohesive"};
    for (const char* nm : poly_names) N[nm] = [this](const Node& s) { contact(s, true); };
    // ---- control
    N["init_particles"] = [this](const Node&) { init_particles(); };
    N["check_and_update_particles"] = [this](const Node&) { check_and_update(); };
    N["compute_loop"] = [this](const Node&) { compute_loop(); };
  }

  // ---------------------------------------------------------------- input operators
  void set_bounds(const double* lo, const double* hi) {
    for (int a = 0; a < 3; a++) {
      bmin_[a] = lo[a];
      bmax_[a] = hi[a];
    }
    have_bounds_ = true;
  }

  void init_rsa(const Node& s) {
    const Node* b = slot(s, "bounds");
    if (!b || !b->is_seq() || b->seq.size() != 2) fail(XDEM_ERR_ARG, "init_rsa: bounds: [[xmin,ymin,zmin],[xmax,ymax,zmax]]");
    const auto lo = xdem_host::to_doubles(b->seq[0]), hi = xdem_host::to_doubles(b->seq[1]);
    const double r = sdouble(s, "radius", 0);
    if (!(r > 0) || lo.size() != 3 || hi.size() != 3) fail(XDEM_ERR_ARG, "init_rsa: radius > 0 and 3-d bounds required");
    const uint32_t t = slot(s, "type") ? type_index(*slot(s, "type")) : 0;
    // seeded random sequential addition of equal spheres in the box (cell list, 6000 consecutive
    // rejected draws end the filling); the reference calls an external RSA package for this
    std::vector<double> cand(3);
    const double d2 = 4.0 * r * r, cell = 2.0 * r;
    int dims[3];
    for (int a = 0; a < 3; a++) dims[a] = std::max(1, (int)std::floor((hi[a] - lo[a]) / cell));
    std::vector<std::vector<int64_t>> grid((size_t)dims[0] * dims[1] * dims[2]);
    auto cidx = [&](double x, int a) {
      int k = (int)((x - lo[a]) / (hi[a] - lo[a]) * dims[a]);
      return std::min(std::max(k, 0), dims[a] - 1);
    };
    std::vector<double> u(3 * 4096);
    uint64_t drawn = 0;
    int fails = 0;
    const size_t first = P_.size();
    while (fails < 6000) {
      if (drawn % 4096 == 0) {
        for (int a = 0; a < 3; a++) xdem_gen_uniform(seed_ + 17, (uint64_t)a + 1 + 4 * (drawn / 4096), 4096,
                                                     lo[a] + r, hi[a] - r, u.data() + a * 4096);
      }
      const double x = u[drawn % 4096], y = u[4096 + drawn % 4096], z = u[8192 + drawn % 4096];
      drawn++;
      const int cx = cidx(x, 0), cy = cidx(y, 1), cz = cidx(z, 2);
      bool ok = true;
      for (int dz = -1; dz <= 1 && ok; dz++)
        for (int dy = -1; dy <= 1 && ok; dy++)
          for (int dx = -1; dx <= 1 && ok; dx++) {
            const int X = cx + dx, Y = cy + dy, Z = cz + dz;
            if (X < 0 || Y < 0 || Z < 0 || X >= dims[0] || Y >= dims[1] || Z >= dims[2]) continue;
            for (int64_t q : grid[((size_t)Z * dims[1] + Y) * dims[0] + X]) {
              const double ex = P_.rx[q] - x, ey = P_.ry[q] - y, ez = P_.rz[q] - z;
              if (ex * ex + ey * ey + ez * ez < d2) {
                ok = false;
                break;
              }
            }
          }
      if (!ok) {
        fails++;
        continue;
      }
      fails = 0;
      grid[((size_t)cz * dims[1] + cy) * dims[0] + cx].push_back((int64_t)P_.size());
      P_.add(x, y, z, t, r);
    }
    (void)first;
    set_bounds(lo.data(), hi.data());
  }

  // "N", then a comment line, then "type x y z" per particle
  void read_xyz(const Node& s) {
    const Node* f = slot(s, "file");
    if (!f) fail(XDEM_ERR_ARG, "read_xyz: file is required");
    std::string path = f->s;
    std::ifstream in(path);
    if (!in) {
      for (auto& d : search_dirs) {
        in.open(d + "/" + path);
        if (in) break;
      }
    }
    if (!in) fail(XDEM_ERR_ARG, "read_xyz: cannot open " + path);
    long long n = 0;
    std::string line;
    std::getline(in, line);
    n = std::atoll(line.c_str());
    std::getline(in, line);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (long long k = 0; k < n; k++) {
      std::string tname;
      double x, y, z;
      if (!(in >> tname >> x >> y >> z)) fail(XDEM_ERR_ARG, "read_xyz: short file " + path);
      Node tn = Node::scalar(tname);
      P_.add(x, y, z, type_index(tn), 0.5);
      const double p[3] = {x, y, z};
      for (int a = 0; a < 3; a++) {
        lo[a] = std::min(lo[a], p[a]);
        hi[a] = std::max(hi[a], p[a]);
      }
    }
    const double enl = sdouble(s, "enlarge_bounds", 0.0);
    for (int a = 0; a < 3; a++) {
      lo[a] -= enl;
      hi[a] += enl;
    }
    set_bounds(lo, hi);
  }

  // n spheres with radius uniform in [rmin, rmax], packed at solid fraction phi in a periodic box
  // [0, L)^2 x [0, Lz) (the polydisperse benchmark input: RSA at scale 0.8 plus relaxation)
  void polydisperse_pack(const Node& s) {
    const int64_t n = (int64_t)sdouble(s, "n", 0);
    const double rmin = sdouble(s, "rmin", 0.5), rmax = sdouble(s, "rmax", 1.0), phi = sdouble(s, "phi", 0.55);
    const uint64_t seed = (uint64_t)sdouble(s, "seed", 0);
    const int64_t sweeps = (int64_t)sdouble(s, "relax_sweeps", 150);
    if (n <= 0) fail(XDEM_ERR_ARG, "polydisperse_pack: n > 0 required");
    const uint32_t t = slot(s, "type") ? type_index(*slot(s, "type")) : 0;
    std::vector<double> r(n), x(n), y(n), z(n);
    xdem_gen_uniform(seed, 1, n, rmin, rmax, r.data());
    double vol = 0;
    for (double v : r) vol += 4.0 / 3.0 * M_PI * v * v * v;
    const double aspect = sdouble(s, "aspect_z", 0.5);
    const double Lxy = std::cbrt(vol / phi / aspect), Lz = Lxy * aspect;
    const double lo[3] = {0, 0, 0}, hi[3] = {Lxy, Lxy, Lz};
    double mo = 0;
    if (xdem_gen_pack_periodic(seed + 2, n, r.data(), lo, hi, 0.8, sweeps, 1e-3 * rmin, x.data(), y.data(), z.data(),
                               &mo) < 0)
      fail(XDEM_ERR_ARG, "polydisperse_pack: packing failed");
    for (int64_t i = 0; i < n; i++) P_.add(x[i], y[i], z[i], t, r[i]);
    set_bounds(lo, hi);
  }

  // per-type fields (src/field/set_fields.cu): radius, density -> mass = rho 4/3 pi r^3 and inertia
  // 0.4 m r^2 (set_fields.cu:323-327), velocity, Gaussian velocity spread
  void set_fields(const Node& s) {
    if (const Node* poly = slot(s, "polyhedra"))
      if (xdem_host::to_bool(*poly)) fail(XDEM_ERR_UNSUPPORTED, "set_fields: polyhedra decks need shape files");
    const Node* types = slot(s, "type");
    if (!types || !types->is_seq()) fail(XDEM_ERR_ARG, "set_fields: type list required");
    auto per_type = [&](const char* key, size_t k) -> const Node* {
      const Node* v = slot(s, key);
      if (!v) return nullptr;
      if (!v->is_seq() || v->seq.size() != types->seq.size()) fail(XDEM_ERR_ARG, std::string("set_fields: ") + key);
      return &v->seq[k];
    };
    for (size_t k = 0; k < types->seq.size(); k++) {
      const uint32_t t = type_index(types->seq[k]);
      const Node* rad = per_type("radius", k);
      const Node* den = per_type("density", k);
      const Node* vel = per_type("velocity", k);
      const Node* sig = per_type("sigma_velocity", k);
      const Node* grp = per_type("group", k);
      std::vector<size_t> idx;
      for (size_t i = 0; i < P_.size(); i++)
        if (P_.type[i] == t) idx.push_back(i);
      std::vector<double> nv(3 * idx.size());
      if (sig) xdem_gen_normal(seed_ + 29 + t, 7, (int64_t)nv.size(), 0.0, xdem_host::to_double(*sig), nv.data());
      const double rho = den ? xdem_host::to_double(*den) : 1.0;
      std::vector<double> v0 = vel ? xdem_host::to_doubles(*vel) : std::vector<double>{0, 0, 0};
      if (v0.size() != 3) fail(XDEM_ERR_ARG, "set_fields: velocity needs 3 values");
      for (size_t q = 0; q < idx.size(); q++) {
        const size_t i = idx[q];
        if (rad) P_.radius[i] = xdem_host::to_double(*rad);
        const double r = P_.radius[i];
        const double V = (4.0 / 3.0) * M_PI * r * r * r;
        P_.mass[i] = V * rho;
        P_.inx[i] = P_.iny[i] = P_.inz[i] = 0.4 * P_.mass[i] * r * r;
        P_.vx[i] = v0[0] + (sig ? nv[3 * q] : 0.0);
        P_.vy[i] = v0[1] + (sig ? nv[3 * q + 1] : 0.0);
        P_.vz[i] = v0[2] + (sig ? nv[3 * q + 2] : 0.0);
        if (grp) P_.group[i] = (uint32_t)xdem_host::to_int(*grp);
      }
    }
  }

  // params slot of register_* (Driver_params decode, driver_params.hpp:420-535): STATIONARY, LINEAR_MOTION
  // {motion_vector, const_vel} and SHAKER {amplitude, omega, shaker_dir (default [0, 0, 1])}, with
  // motion_start_threshold / motion_end_threshold; other motion types stop the run
  xdem_driver_motion motion_params(const Node* pr, const char* what) {
    xdem_driver_motion m{};
    m.motion_end_threshold = 1e300;
    m.shaker_dir[2] = 1.0;
    if (!pr) return m;
    const Node* t = slot(*pr, "motion_type");
    const std::string type = t ? t->s : std::string("STATIONARY");
    auto vec3 = [&](const char* k, double* out) {
      if (const Node* v = slot(*pr, k)) {
        const auto x = xdem_host::to_doubles(*v);
        if (x.size() != 3) fail(XDEM_ERR_ARG, std::string(what) + ": " + k + " needs 3 values");
        for (int a = 0; a < 3; a++) out[a] = x[a];
        return true;
      }
      return false;
    };
    if (type == "STATIONARY") {
      m.motion_type = XDEM_MOTION_STATIONARY;
    } else if (type == "LINEAR_MOTION") {
      m.motion_type = XDEM_MOTION_LINEAR;
      if (!vec3("motion_vector", m.motion_vector)) fail(XDEM_ERR_ARG, std::string(what) + ": motion_vector required");
      m.const_vel = sdouble(*pr, "const_vel", 0.0);
    } else if (type == "SHAKER") {
      m.motion_type = XDEM_MOTION_SHAKER;
      if (!slot(*pr, "omega") || !slot(*pr, "amplitude"))
        fail(XDEM_ERR_ARG, std::string(what) + ": SHAKER needs omega and amplitude");
      m.omega = sdouble(*pr, "omega", 0.0);
      m.amplitude = sdouble(*pr, "amplitude", 0.0);
      vec3("shaker_dir", m.shaker_dir);
    } else {
      fail(XDEM_ERR_UNSUPPORTED, std::string(what) + ": motion type " + type +
                                     " is not provided (STATIONARY, LINEAR_MOTION, SHAKER)");
    }
    m.motion_start_threshold = sdouble(*pr, "motion_start_threshold", 0.0);
    m.motion_end_threshold = sdouble(*pr, "motion_end_threshold", 1e300);
    return m;
  }

  // driver ids in use (surfaces, cylinders, balls and the mesh share one numbering: the partner id of
  // their interactions)
  std::vector<size_t> driver_ids_;
  void claim_driver_id(size_t id, const std::string& what) {
    if (std::find(driver_ids_.begin(), driver_ids_.end(), id) != driver_ids_.end())
      fail(XDEM_ERR_ARG, what + ": driver id " + std::to_string(id) + " is already registered");
    driver_ids_.push_back(id);
  }

  void put_driver(size_t id, const xdem_surface& w, const xdem_driver_motion& m) {
    claim_driver_id(id, "register driver");
    if (walls_.size() <= id) {
      walls_.resize(id + 1);
      xdem_driver_motion still{};
      still.motion_end_threshold = 1e300;
      still.shaker_dir[2] = 1.0;
      motions_.resize(id + 1, still);
    }
    walls_[id] = w;
    motions_[id] = m;
  }

  void register_surface(const Node& s) {
    const Node* st = slot(s, "state");
    const xdem_driver_motion mo = motion_params(slot(s, "params"), "register_surface");
    if (!st) fail(XDEM_ERR_ARG, "register_surface: state {normal, offset} required");
    xdem_surface w{};
    const Node* nrm = slot(*st, "normal");
    if (!nrm) fail(XDEM_ERR_ARG, "register_surface: normal required");
    const auto n = xdem_host::to_doubles(*nrm);
    if (n.size() != 3) fail(XDEM_ERR_ARG, "register_surface: normal needs 3 values");
    const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int a = 0; a < 3; a++) w.normal[a] = n[a] / len;
    w.offset = sdouble(*st, "offset", 0.0);
    for (int a = 0; a < 3; a++) w.center[a] = w.normal[a] * w.offset;
    if (const Node* v = slot(*st, "vel")) {
      const auto x = xdem_host::to_doubles(*v);
      if (x.size() != 3) fail(XDEM_ERR_ARG, "register_surface: vel needs 3 values");
      for (int a = 0; a < 3; a++) w.vel[a] = x[a];
    }
    const size_t id = (size_t)sdouble(s, "id", (double)walls_.size());
    if (id >= XDEM_MAX_SURFACES) fail(XDEM_ERR_CAPACITY, "register_surface: at most 8 surfaces");
    put_driver(id, w, mo);
  }

  // register_cylinder / register_ball (src/driver/register_cylinder.cpp, register_ball.cpp): state
  // {center, radius[, axis (cylinder), vel, vrot]} as CylinderFields / BallFields decode it
  // (cylinder.hpp:40-70, ball.hpp:44-80); STATIONARY motion only
  void register_round(const Node& s, int kind, const char* what) {
    const Node* st = slot(s, "state");
    const xdem_driver_motion mo = motion_params(slot(s, "params"), what);
    if (!st) fail(XDEM_ERR_ARG, std::string(what) + ": state required");
    xdem_surface w{};
    w.kind = kind;
    auto vec3 = [&](const char* k, double* out, bool required) {
      const Node* v = slot(*st, k);
      if (!v) {
        if (required) fail(XDEM_ERR_ARG, std::string(what) + ": " + k + " required");
        return;
      }
      const auto x = xdem_host::to_doubles(*v);
      if (x.size() != 3) fail(XDEM_ERR_ARG, std::string(what) + ": " + k + " needs 3 values");
      for (int a = 0; a < 3; a++) out[a] = x[a];
    };
    vec3("center", w.center, true);
    if (kind == XDEM_DRIVER_CYLINDER) vec3("axis", w.axis, true);
    vec3("vel", w.vel, false);
    vec3("vrot", w.vrot, false);
    if (!slot(*st, "radius")) fail(XDEM_ERR_ARG, std::string(what) + ": radius required");
    w.radius = sdouble(*st, "radius", 0.0);
    const size_t id = (size_t)sdouble(s, "id", (double)walls_.size());
    if (id >= XDEM_MAX_SURFACES) fail(XDEM_ERR_CAPACITY, std::string(what) + ": at most 8 drivers");
    put_driver(id, w, mo);
  }

  // register_stl_mesh (src/driver/register_stl_mesh.cpp:30-140) for a stationary .stl mesh acting on
  // spheres: STLMeshReader (stl_mesh_reader.hpp:63-135) and build_shape (stl_mesh_to_driver.hpp:25-126) as it
  // runs on one thread (vertices sorted and unique, faces in file order, three edges per face with
  // consecutive repeats removed); scale multiplies the vertices. One mesh; STATIONARY with zero vel and
  // vrot only; .shp meshes and deform are not provided.
  struct MeshIn {
    bool on = false;
    int32_t id = 0;
    std::vector<double> v;
    std::vector<uint32_t> e, f;
    double rm = 0, center[3] = {0, 0, 0}, quat[4] = {1, 0, 0, 0};
    int cap = 32;
  } mesh_;

  static std::vector<std::array<double, 3>> read_stl(const std::string& path, bool binary, const std::string& what) {
    std::vector<std::array<double, 3>> t;
    if (binary) {
      std::ifstream in(path, std::ios::binary);
      char head[80];
      uint32_t n = 0;
      if (!in.read(head, 80) || !in.read((char*)&n, 4)) throw std::runtime_error(what + ": truncated binary STL " + path);
      for (uint32_t k = 0; k < n; k++) {
        float rec[12];
        uint16_t attr;
        if (!in.read((char*)rec, 48) || !in.read((char*)&attr, 2))
          throw std::runtime_error(what + ": truncated binary STL " + path);
        for (int q = 1; q < 4; q++) t.push_back({(double)rec[3 * q], (double)rec[3 * q + 1], (double)rec[3 * q + 2]});
      }
      return t;
    }
    std::ifstream in(path);
    std::string tok;
    bool inside = false;
    std::vector<std::array<double, 3>> cur;
    while (in >> tok) {
      if (tok == "outer") {
        inside = true;
        cur.clear();
      } else if (tok == "vertex" && inside) {
        std::array<double, 3> x;
        if (!(in >> x[0] >> x[1] >> x[2])) throw std::runtime_error(what + ": bad vertex line in " + path);
        cur.push_back(x);
      } else if (tok == "endloop" && inside) {
        if (cur.size() != 3) throw std::runtime_error(what + ": a face with " + std::to_string(cur.size()) +
                                                      " vertices in " + path + " (triangles only)");
        t.insert(t.end(), cur.begin(), cur.end());
        inside = false;
      }
    }
    return t;
  }

  void register_stl_mesh(const Node& s) {
    const char* what = "register_stl_mesh";
    if (mesh_.on) fail(XDEM_ERR_CAPACITY, "register_stl_mesh: one mesh per run");
    if (const Node* pr = slot(s, "params"))
      if (const Node* m = slot(*pr, "motion_type"))
        if (m->s != "STATIONARY") fail(XDEM_ERR_UNSUPPORTED, "register_stl_mesh: only STATIONARY meshes are provided");
    if (slot(s, "deform")) fail(XDEM_ERR_UNSUPPORTED, "register_stl_mesh: deform is not provided");
    const Node* f = slot(s, "filename");
    if (!f) fail(XDEM_ERR_ARG, "register_stl_mesh: filename is required");
    if (f->s.find(".stl") == std::string::npos)
      fail(XDEM_ERR_UNSUPPORTED, "register_stl_mesh: only .stl meshes are provided (" + f->s + ")");
    if (!slot(s, "minkowski")) fail(XDEM_ERR_ARG, "register_stl_mesh: minkowski is required");
    mesh_.rm = sdouble(s, "minkowski", 0.0);
    // id is required (register_stl_mesh.cpp:37) and must differ from the other drivers' ids
    if (!slot(s, "id")) fail(XDEM_ERR_ARG, "register_stl_mesh: id is required");
    const double idv = sdouble(s, "id", 0.0);
    if (!(idv >= 0) || idv != (double)(int32_t)idv) fail(XDEM_ERR_ARG, "register_stl_mesh: id must be a non-negative integer");
    mesh_.id = (int32_t)idv;
    claim_driver_id((size_t)mesh_.id, what);
    // candidate slots per sphere (no counterpart in the reference, whose lists grow): max_candidates
    mesh_.cap = 32;
    if (slot(s, "max_candidates")) {
      const double mc = sdouble(s, "max_candidates", 32.0);
      if (!(mc >= 1) || mc > 1024 || mc != (double)(int)mc)
        fail(XDEM_ERR_ARG, "register_stl_mesh: max_candidates must be an integer in 1..1024");
      mesh_.cap = (int)mc;
    }
    if (const Node* st = slot(s, "state")) {
      auto vec = [&](const char* k, double* out, size_t nval) {
        if (const Node* v = slot(*st, k)) {
          const auto x = xdem_host::to_doubles(*v);
          if (x.size() != nval) fail(XDEM_ERR_ARG, std::string(what) + ": " + k + " needs " + std::to_string(nval) + " values");
          for (size_t a = 0; a < nval; a++) out[a] = x[a];
        }
      };
      vec("center", mesh_.center, 3);
      vec("quat", mesh_.quat, 4);
      double vel[3] = {0, 0, 0}, vrot[3] = {0, 0, 0};
      vec("vel", vel, 3);
      vec("vrot", vrot, 3);
      for (int a = 0; a < 3; a++)
        if (vel[a] != 0 || vrot[a] != 0) fail(XDEM_ERR_UNSUPPORTED, "register_stl_mesh: moving meshes are not provided");
    }
    std::string path = f->s;
    if (!std::ifstream(path)) {
      for (auto& d : search_dirs)
        if (std::ifstream(d + "/" + path)) {
          path = d + "/" + path;
          break;
        }
    }
    if (!std::ifstream(path)) fail(XDEM_ERR_ARG, "register_stl_mesh: cannot open " + f->s);
    const Node* b = slot(s, "binary");
    std::vector<std::array<double, 3>> t;
    try {
      t = read_stl(path, b && xdem_host::to_bool(*b), what);
    } catch (const std::exception& ex) {
      fail(XDEM_ERR_ARG, ex.what());
    }
    // build_shape
    std::vector<std::array<double, 3>> v(t);
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    std::vector<uint32_t> fi(t.size());
    for (size_t k = 0; k < t.size(); k++) fi[k] = (uint32_t)(std::lower_bound(v.begin(), v.end(), t[k]) - v.begin());
    std::vector<std::array<uint32_t, 2>> e;
    for (size_t k = 0; k < t.size(); k += 3)
      for (int a = 0; a < 3; a++) {
        const std::array<uint32_t, 2> ed{fi[k + a], fi[k + (a + 1) % 3]};
        if (e.empty() || e.back() != ed) e.push_back(ed);
      }
    double scale = 1.0;
    if (slot(s, "scale")) {
      scale = sdouble(s, "scale", 1.0);
      if (!(scale > 0)) fail(XDEM_ERR_ARG, "register_stl_mesh: scale <= 0");
    }
    mesh_.v.clear();
    for (auto& x : v)
      for (int a = 0; a < 3; a++) mesh_.v.push_back(scale == 1.0 ? x[a] : x[a] * scale);
    mesh_.e.clear();
    for (auto& ed : e) {
      mesh_.e.push_back(ed[0]);
      mesh_.e.push_back(ed[1]);
    }
    mesh_.f = fi;
    mesh_.on = true;
  }

  // ---------------------------------------------------------------- domain + upload
  void setup_domain() {
    const Node* d = deck.get("domain");
    xdem_domain dom{};
    if (!have_bounds_) fail(XDEM_ERR_STATE, "domain: no input operator set the particle bounds");
    for (int a = 0; a < 3; a++) {
      dom.lo[a] = bmin_[a];
      dom.hi[a] = bmax_[a];
    }
    if (d) {
      if (const Node* p = d->get("periodic")) {
        if (!p->is_seq() || p->seq.size() != 3) fail(XDEM_ERR_ARG, "domain: periodic needs 3 booleans");
        for (int a = 0; a < 3; a++) dom.periodic[a] = xdem_host::to_bool(p->seq[a]) ? 1 : 0;
      }
      if (const Node* b = d->get("bounds")) {
        if (!b->is_seq() || b->seq.size() != 2) fail(XDEM_ERR_ARG, "domain: bounds [[lo],[hi]]");
        const auto lo = xdem_host::to_doubles(b->seq[0]), hi = xdem_host::to_doubles(b->seq[1]);
        for (int a = 0; a < 3; a++) {
          dom.lo[a] = lo[a];
          dom.hi[a] = hi[a];
        }
      }
      dom.cell_size = d->get("cell_size") ? xdem_host::to_double(*d->get("cell_size")) : 0.0;
    }
    double rmax = 0;
    for (double r : P_.radius) rmax = std::max(rmax, r);
    // dem_rcut_max (src/field/dem_rcut_max.cpp:55-65): the cell must hold 2 rmax + rcut_inc
    dom.cell_size = std::max(dom.cell_size, 2.0 * rmax + rcut_inc());
    check(xdem_set_domain(ctx_, &dom), "domain");
  }

  void upload() {
    xdem_particles p{};
    const int64_t n = (int64_t)P_.size();
    p.n = n;
    p.id = P_.id.data();
    p.type = P_.type.data();
    p.group = P_.group.data();
    p.rx = P_.rx.data(); p.ry = P_.ry.data(); p.rz = P_.rz.data();
    p.vx = P_.vx.data(); p.vy = P_.vy.data(); p.vz = P_.vz.data();
    p.mass = P_.mass.data(); p.radius = P_.radius.data(); p.homothety = P_.h.data();
    p.qw = P_.qw.data(); p.qx = P_.qx.data(); p.qy = P_.qy.data(); p.qz = P_.qz.data();
    p.vrotx = P_.wx.data(); p.vroty = P_.wy.data(); p.vrotz = P_.wz.data();
    p.inx = P_.inx.data(); p.iny = P_.iny.data(); p.inz = P_.inz.data();
    check(xdem_particles_upload(ctx_, &p), "upload");
    if (!walls_.empty()) {
      check(xdem_set_surfaces(ctx_, walls_.data(), (int32_t)walls_.size()), "surfaces");
      check(xdem_set_driver_motion(ctx_, motions_.data(), (int32_t)motions_.size()), "driver motion");
    }
    if (mesh_.on)
      check(xdem_set_mesh(ctx_, mesh_.id, mesh_.v.data(), (int64_t)mesh_.v.size() / 3, mesh_.e.data(),
                          (int64_t)mesh_.e.size() / 2, mesh_.f.data(), (int64_t)mesh_.f.size() / 3, mesh_.rm,
                          mesh_.center, mesh_.quat, mesh_.cap),
            "mesh");
    uploaded_ = true;
  }

  // init_particles: the grid, then the full update (update_particles_full)
  void init_particles() {
    setup_domain();
    upload();
    exec("update_particles_full", nullptr, 1);
  }

  void check_and_update() {
    int32_t over = 0;
    const double thr = 0.5 * rcut_inc();
    if (poly_) check(xdem_vertex_displ_over(ctx_, thr, &over), "vertex_displ_over");
    else check(xdem_particle_displ_over(ctx_, thr, &over), "particle_displ_over");
    // trigger_move_particles = particle_displ or driver_displ (config_move_particles.msp:22-27)
    int32_t dover = 0;
    check(xdem_driver_displ_over(ctx_, thr, &dover), "driver_displ_over");
    if (over || dover) {
      stats.rebuilds++;
      exec("update_particles_full", nullptr, 1);
    } else if (poly_) {
      check(xdem_compute_vertices(ctx_), "compute_vertices");
    }
  }

  void contact(const Node& s, bool poly) {
    if (poly != poly_) fail(XDEM_ERR_ARG, "contact operator does not match the particle kind of the deck");
    if (const Node* sym = slot(s, "symetric"))
      if (!xdem_host::to_bool(*sym)) fail(XDEM_ERR_UNSUPPORTED, "contact: symetric: false is not provided");
    const Node* cfg = slot(s, "config");
    if (!cfg) fail(XDEM_ERR_ARG, "contact: config block required");
    const xdem_contact_params cp = contact_params(*cfg);
    const Node* cfd = slot(s, "config_driver");
    xdem_contact_params cpd{};
    if (cfd) cpd = contact_params(*cfd);
    int form = XDEM_FORM_FP64;
    if (const Node* f = slot(s, "precision")) {
      if (f->s == "fp32") form = XDEM_FORM_FP32;
      else if (f->s != "fp64") fail(XDEM_ERR_ARG, "contact: precision is fp64 or fp32");
    }
    // analysis (contact_force_sphere.cu.in:62-66, 190-199, and the same slots of contact_polyhedron):
    // every analysis_interaction_dump_frequency
    // steps the interactions of this pass go to <dir_name>/ExaDEMAnalyses/<basename><timestep>/
    const Node* fq = slot(s, "analysis_interaction_dump_frequency");
    const int64_t freq =
        fq ? (int64_t)xdem_host::to_double(*fq) : (int64_t)gdouble("analysis_interaction_dump_frequency", -1);
    const bool dump = freq > 0 && timestep_ % freq == 0;
    if (dump) check(xdem_set_contact_record(ctx_, 1), "contact_sphere");
    if (poly) check(xdem_contact_polyhedron(ctx_, &cp, cfd ? &cpd : nullptr, dt()), "contact_polyhedron");
    else check(xdem_contact_sphere(ctx_, &cp, cfd ? &cpd : nullptr, dt(), form), "contact_sphere");
    if (dump) {
      write_interaction_dump(sstring(s, "dir_name", "ExaDEMOutputDir"),
                             sstring(s, "interaction_basename", "InteractionOutputDir-"));
      check(xdem_set_contact_record(ctx_, 0), "contact_sphere");
    }
  }

  std::string sstring(const Node& s, const char* k, const char* dflt) {
    if (const Node* v = slot(s, k)) return v->s;
    if (const Node* v = global().get(k)) return v->s;
    return dflt;
  }

  // itools::create_buffer + write_file (src/interaction/include/exaDEM/itools/writer.hpp:35-115): one
  // row per interaction in contact, owner side, numbers as a default std::ostream writes them
  void write_interaction_dump(const std::string& dir, const std::string& base) {
    int64_t n = 0;
    check(xdem_contact_state_count(ctx_, &n), "interaction dump");
    std::vector<uint64_t> ii((size_t)std::max<int64_t>(n, 1)), jj(ii.size());
    std::vector<uint32_t> si(ii.size()), sj(ii.size());
    std::vector<uint16_t> tt(ii.size());
    std::vector<double> v(16 * ii.size());
    int64_t w = 0;
    if (n > 0)
      check(xdem_contact_state_dump(ctx_, n, ii.data(), jj.data(), si.data(), sj.data(), tt.data(), v.data(), &w),
            "interaction dump");
    const std::string name = base + std::to_string(timestep_);
    const std::string sub = dir + "/ExaDEMAnalyses/" + name;
    std::error_code ec;
    std::filesystem::create_directories(sub, ec);
    if (ec) fail(XDEM_ERR_ARG, "interaction dump: cannot create " + sub);
    std::ofstream out(sub + "/" + name + "_0.txt");
    if (!out) fail(XDEM_ERR_ARG, "interaction dump: cannot write in " + sub);
    for (int64_t k = 0; k < w; k++) {
      out << ii[k] << "," << jj[k] << "," << si[k] << "," << sj[k] << "," << tt[k];
      for (int q = 0; q < 16; q++) out << "," << v[16 * k + q];
      out << "\n";
    }
    stats.operator_calls++;
  }

  void compute_loop() {
    const int64_t end = (int64_t)gdouble("simulation_end_iteration", 0);
    int64_t t = (int64_t)gdouble("timestep", 0);
    for (; t < end; t++) {
      timestep_ = t + 1;  // next_time_step ran after first_iteration and after each iteration
      exec("numerical_scheme", nullptr, 1);
      check(xdem_next_time_step(ctx_, dt()), "next_time_step");
      stats.iterations++;
    }
  }

 public:
  std::vector<std::string> search_dirs;
};

}  // namespace

struct xdem_deck {
  Runner r;
};

extern "C" int xdem_deck_load(const char* path, const char* const* search_dirs, int32_t n_dirs, xdem_deck** out) {
  if (!path || !out || (n_dirs > 0 && !search_dirs)) return XDEM_ERR_ARG;
  *out = nullptr;
  xdem_deck* d = new xdem_deck();
  std::vector<std::string> dirs;
  for (int k = 0; k < n_dirs; k++) dirs.push_back(search_dirs[k]);
  d->r.search_dirs = dirs;
  std::string p(path);
  const size_t k = p.find_last_of('/');
  d->r.search_dirs.push_back(k == std::string::npos ? std::string(".") : p.substr(0, k));
  try {
    Node base;
    base.kind = Node::Map;
    bool found = false;
    for (auto& dir : dirs) {
      std::ifstream f(dir + "/main-config.msp");
      if (f) {
        base = xdem_host::load_msp(dir + "/main-config.msp", dirs);
        found = true;
        break;
      }
    }
    if (!found) throw xdem_host::ParseError("main-config.msp not found in the search directories");
    xdem_host::merge_into(base, xdem_host::load_msp(path, dirs));
    d->r.deck = base;
  } catch (std::exception& e) {
    d->r.error = e.what();
    *out = d;
    return XDEM_ERR_ARG;
  }
  *out = d;
  return XDEM_OK;
}

extern "C" int xdem_deck_set_global(xdem_deck* d, const char* name, const char* value) {
  if (!d || !name || !value) return XDEM_ERR_ARG;
  Node* g = d->r.deck.get("global");
  if (!g) {
    Node m;
    m.kind = Node::Map;
    d->r.deck.set("global", m);
    g = d->r.deck.get("global");
  }
  g->set(name, Node::scalar(value));
  return XDEM_OK;
}

extern "C" int xdem_deck_run(xdem_deck* d, xdem_ctx* ctx, xdem_deck_stats* stats) {
  if (!d || !ctx) return XDEM_ERR_ARG;
  try {
    d->r.run(ctx);
  } catch (Fail& f) {
    d->r.error = f.what();
    if (stats) *stats = d->r.stats;
    return f.code;
  } catch (std::exception& e) {
    d->r.error = e.what();
    if (stats) *stats = d->r.stats;
    return XDEM_ERR_ARG;
  }
  if (stats) *stats = d->r.stats;
  return XDEM_OK;
}

extern "C" int xdem_deck_last_error(const xdem_deck* d, char* buf, int64_t n) {
  if (!d || !buf || n <= 0) return XDEM_ERR_ARG;
  std::snprintf(buf, (size_t)n, "%s", d->r.error.c_str());
  return XDEM_OK;
}

extern "C" void xdem_deck_destroy(xdem_deck* d) { delete d; }
