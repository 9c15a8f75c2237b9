/*
 * xdem.h — C ABI of libxdem_hip.so, the MI355X (gfx950) implementation of
 * exaDEM's per-timestep hot path: cell-grid neighbour build, interaction
 * update with friction-history carry-over, Hooke contact with Coulomb
 * friction (spheres and walls), and velocity-Verlet integration.
 *
 * Rules of the boundary
 *   - plain C types only: POD structs, raw host pointers, sizes;
 *   - every call takes an opaque xdem_ctx that owns the device buffers and one
 *     HIP stream (default priority) on one GPU;
 *   - every call returns 0 (XDEM_OK) or a negative XDEM_ERR_* code; the text
 *     of the last error is available through xdem_last_error();
 *   - device work is queued on the context stream; calls that return host
 *     scalars or fill host arrays synchronise that stream first;
 *   - one host thread per context.
 *
 * Each operator entry point names the reference operator it replaces
 * (paths relative to the exaDEM source tree).
 */
#ifndef XDEM_H
#define XDEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ codes */
#define XDEM_OK 0
#define XDEM_ERR_ARG (-1)         /* invalid argument or slot value           */
#define XDEM_ERR_HIP (-2)         /* HIP runtime error                        */
#define XDEM_ERR_STATE (-3)       /* call made in the wrong state             */
#define XDEM_ERR_CAPACITY (-4)    /* a documented size limit was exceeded     */
#define XDEM_ERR_UNSUPPORTED (-5) /* feature not provided by this build      */

/* Largest particle count a context accepts (neighbour words hold 25 bits). */
#define XDEM_MAX_PARTICLES 33554431
/* Largest number of planar walls (Surface drivers). */
#define XDEM_MAX_SURFACES 8

/* Contact-force forms accepted by xdem_contact_sphere / xdem_run_steps. */
#define XDEM_FORM_FP64 0 /* plain form: FP64 end to end                   */
#define XDEM_FORM_FP32 1 /* FP32 geometry/force on cell-relative offsets, */
                         /* FP64 overlap sign test, compensated FP32 sums */

typedef struct xdem_ctx xdem_ctx;

/* Particle fields, structure of arrays, one entry per particle.
 * Mirrors DEMFieldSet (include/exaDEM/fields.h:34-54): rx..rz implicit,
 * vx..vz, fx..fz, mass, homothety, radius, orient (w,x,y,z), mom, vrot,
 * arot, inertia, id, type, group. The stress tensor is not carried.
 * Pointers may be NULL in a download for fields the caller does not want. */
typedef struct xdem_particles {
  int64_t n;
  uint64_t* id;
  uint32_t* type;
  uint32_t* group;
  double *rx, *ry, *rz;
  double *vx, *vy, *vz;
  double *fx, *fy, *fz;
  double *mass, *radius, *homothety;
  double *qw, *qx, *qy, *qz;
  double *momx, *momy, *momz;
  double *vrotx, *vroty, *vrotz;
  double *arotx, *aroty, *arotz;
  double *inx, *iny, *inz;
} xdem_particles;

/* Simulation box. Axes flagged periodic wrap particles at each rebuild and
 * see periodic images of the particles near the opposite face (the role of
 * exaNBody's ghost layer on one rank). cell_size is the minimum cell edge;
 * the grid uses floor(L/cell_size) cells per axis (at least 3 on periodic
 * axes) and an edge of L/dims. */
typedef struct xdem_domain {
  double lo[3];
  double hi[3];
  int32_t periodic[3];
  double cell_size;
} xdem_domain;

/* ContactParams (src/forcefield/include/exaDEM/forcefield/contact_parameters.hpp:32-41),
 * same members in the same order. */
typedef struct xdem_contact_params {
  double kn, kt, kr, mu, damp_rate, fc, dncut, gamma;
} xdem_contact_params;

/* A stationary driver. kind XDEM_DRIVER_SURFACE: planar wall, SurfaceFields
 * (src/driver/include/exaDEM/surface.hpp:29-43), center defaults to offset*normal when all three
 * components are zero. XDEM_DRIVER_CYLINDER: CylinderFields (cylinder.hpp:28-36): radius, axis as
 * the 0/1 mask of the coordinates the projection keeps ((1, 0, 1) for a cylinder along y), center
 * (projected with the mask, Cylinder::initialize); particles inside. XDEM_DRIVER_BALL: BallFields
 * (ball.hpp:27-40): center and radius, contact from inside or outside. vel and vrot enter the contact
 * law as the driver's velocity and spin (a rotating drum); the driver does not move. offset and
 * normal are used by planes only, radius and axis by cylinders and balls. */
#define XDEM_DRIVER_SURFACE 0
#define XDEM_DRIVER_CYLINDER 1
#define XDEM_DRIVER_BALL 2
typedef struct xdem_surface {
  double offset;
  double normal[3];
  double center[3];
  double vel[3];
  double vrot[3];
  int32_t kind;
  int32_t pad;
  double radius;
  double axis[3];
} xdem_surface;

/* Convex sphero-polyhedron shape (src/shape/include/exaDEM/shape.hpp:50-466): vertices, edges and
 * faces of the core polyhedron at homothety 1, its Minkowski radius, volume and inertia over mass,
 * and its oriented bounding box in the body frame (axes x, y, z, centre 0, extents including the
 * Minkowski radius). Faces list their vertex indices in order around the face. */
#define XDEM_MAX_SHAPES 8
#define XDEM_MAX_SHAPE_VERTICES 8
#define XDEM_MAX_SHAPE_EDGES 12
#define XDEM_MAX_SHAPE_FACES 8
#define XDEM_MAX_FACE_VERTICES 4
typedef struct xdem_shape {
  int32_t nv, ne, nf, pad;
  double minkowski;
  double volume;
  double inertia_on_mass[3];
  double obb_extent[3];
  double vertices[XDEM_MAX_SHAPE_VERTICES][3];
  int32_t edges[XDEM_MAX_SHAPE_EDGES][2];
  int32_t face_size[XDEM_MAX_SHAPE_FACES];
  int32_t faces[XDEM_MAX_SHAPE_FACES][XDEM_MAX_FACE_VERTICES];
} xdem_shape;

/* One time step of numerical_scheme_verlet_macro
 * (data/config/config_numerical_schemes.msp:6-28) with
 * compute_force = [gravity_force, contact_sphere]. */
typedef struct xdem_scheme {
  double dt;
  double gravity[3];
  int32_t gravity_on;
  int32_t form;            /* XDEM_FORM_* */
  xdem_contact_params cp;  /* particle-particle law (slot config)          */
  xdem_contact_params cpd; /* particle-wall law (slot config_driver)       */
  double rcut_inc;         /* Verlet skin; rebuild when a particle moved   */
                           /* at least rcut_inc/2 since the last rebuild   */
  int32_t poll;            /* steps between host polls of the rebuild flag */
} xdem_scheme;

/* Per-call statistics of xdem_run_steps. */
typedef struct xdem_run_stats {
  int64_t steps;    /* steps executed                                   */
  int64_t rebuilds; /* neighbour rebuilds triggered inside the call     */
  int64_t polls;    /* host synchronisations made to read the flag      */
} xdem_run_stats;

/* --------------------------------------------------------- context / data */
int xdem_ctx_create(int32_t device, xdem_ctx** out);
void xdem_ctx_destroy(xdem_ctx* ctx);
/* Copies the last error text (NUL terminated, truncated to n bytes). */
int xdem_last_error(const xdem_ctx* ctx, char* buf, int64_t n);
int xdem_synchronize(xdem_ctx* ctx);

/* domain operator (data/config/main-config.msp:160-163) + grid sizing. */
int xdem_set_domain(xdem_ctx* ctx, const xdem_domain* dom);
/* register_surface (src/driver/register_surface.cpp:27-58), whole list at once. */
int xdem_set_surfaces(xdem_ctx* ctx, const xdem_surface* walls, int32_t n);

/* Driver motion: the params slot of register_surface / register_cylinder / register_ball
 * (Driver_params, src/driver/include/exaDEM/driver_params.hpp:33-60; MotionType values of
 * motion_type.hpp:23-35). Kinematic motions are provided: STATIONARY (every driver), LINEAR_MOTION
 * (surfaces and balls) and SHAKER (surfaces and cylinders); others fail with XDEM_ERR_UNSUPPORTED. */
#define XDEM_MOTION_STATIONARY 0
#define XDEM_MOTION_LINEAR 1
#define XDEM_MOTION_SHAKER 7
typedef struct xdem_driver_motion {
  int32_t motion_type;
  int32_t pad;
  double motion_vector[3];
  double const_vel;
  double motion_start_threshold; /* motion acts while start <= time + dt <= end */
  double motion_end_threshold;
  double amplitude;              /* SHAKER: A sin(omega (t - start)) along shaker_dir */
  double omega;
  double shaker_dir[3];
} xdem_driver_motion;
/* Motion of each registered driver (n = the number registered by xdem_set_surfaces, in that order),
 * checked and normalised as register_* does (check_motion_coherence, driver_params.hpp:108-214;
 * Surface::initialize, surface.hpp:144-195: a linear surface moves along its normal). NULL with
 * n = 0 makes every driver stationary. Drivers' positions at the last neighbour build are the
 * reference of driver_displ_over. */
int xdem_set_driver_motion(xdem_ctx* ctx, const xdem_driver_motion* motion, int32_t n);
/* push_f_v_r_driver (src/driver/push_f_v_r_driver.cpp:28-65): every driver whose motion is active at
 * time + dt moves by one step (Surface::push_f_v_r surface.hpp:260-299, Cylinder cylinder.hpp:160-173,
 * Ball ball.hpp:203-213). */
int xdem_push_f_v_r_driver(xdem_ctx* ctx, double time, double dt);
/* push_f_v_driver (src/driver/push_f_v_driver.cpp): velocity update of the kinematic motions
 * (LINEAR_MOTION: vel = const_vel * motion_vector; a STATIONARY surface: vel = 0). */
int xdem_push_f_v_driver(xdem_ctx* ctx, double dt);
/* driver_displ_over (src/driver/driver_displ_over.cu:59-130): *result = 1 when a surface's projected
 * centre or a ball's centre moved by at least threshold since the last neighbour build. */
int xdem_driver_displ_over(xdem_ctx* ctx, double threshold, int32_t* result);
/* The drivers' current state (offset, centre, velocity, ...) in xdem_surface form. */
int xdem_drivers_download(xdem_ctx* ctx, xdem_surface* out, int32_t n);
/* physical_time of the time loop; next_time_step adds dt (main-config.msp:65,105), the fused engine
 * (xdem_run_steps) adds dt per step. */
int xdem_set_physical_time(xdem_ctx* ctx, double time);
int xdem_physical_time(xdem_ctx* ctx, double* time);
int xdem_next_time_step(xdem_ctx* ctx, double dt);
/* Replaces the particle set (input_data); ids must be unique. */
int xdem_particles_upload(xdem_ctx* ctx, const xdem_particles* p);
/* Fills the caller's arrays in the current device (cell-sorted) order. */
int xdem_particles_download(xdem_ctx* ctx, xdem_particles* p);
int64_t xdem_num_particles(const xdem_ctx* ctx);

/* ------------------------------------------------------------- operators */
/* combined_compute_prolog (src/numericalscheme/combined_compute_prolog.cu:94-116). */
int xdem_combined_compute_prolog(xdem_ctx* ctx, double dt);
/* combined_compute_epilog (src/numericalscheme/combined_compute_epilog.cu:94-113). */
int xdem_combined_compute_epilog(xdem_ctx* ctx, double dt);
/* reset_force_moment (src/numericalscheme/reset_force_moment.cu:55-72). */
int xdem_reset_force_moment(xdem_ctx* ctx);
/* gravity_force (src/forcefield/gravity_force.cu:38-87). */
int xdem_gravity_force(xdem_ctx* ctx, const double g[3]);
/* force_to_accel (src/numericalscheme/include/exaDEM/force_to_accel.hpp:26-33). */
int xdem_force_to_accel(xdem_ctx* ctx);
/* particle_displ_over (exaNBody; config_move_particles.msp:12-15): *result = 1
 * when a particle moved by at least `threshold` since the last rebuild. */
int xdem_particle_displ_over(xdem_ctx* ctx, double threshold, int32_t* result);

/* migrate_cell_particles_interaction + backup_r (config_spheres.msp:14-17):
 * wraps periodic coordinates, bins particles into cells and reorders every
 * field cell by cell (particles of a cell by ascending id). */
int xdem_move_particles(xdem_ctx* ctx);
/* chunk_neighbors_contact (src/grid/chunk_neighbors_contact.cpp:45-114):
 * pair (a,b) is a neighbour iff d2 > 0, d2 < (ra+rb+rcut_inc)^2 and
 * d2 < nbh_dist^2. nbh_dist <= 0 selects 2*max(radius)+rcut_inc. */
int xdem_chunk_neighbors_contact(xdem_ctx* ctx, double rcut_inc, double nbh_dist);
/* nbh_sphere (src/sphere/nbh_sphere.cu:64-309) with symetric=true: owner
 * choice, wall candidates, friction-history carry-over by key. */
int xdem_nbh_sphere(xdem_ctx* ctx, double rcut_inc, int32_t symetric);
/* classify_interactions (src/interaction/classify_interactions.cpp:52-136, classifier.hpp:33-271):
 * polyhedra: splits the feature interactions into one wave per type (VV, VE, VF, EE), each in list
 * order (a stable partition on the device); contact_polyhedron evaluates one type per wavefront and
 * requires a classification of the current lists (XDEM_ERR_STATE otherwise). Spheres: each
 * interaction type already has its own device structure (pairs in the neighbour list, driver and
 * mesh candidates in their arrays); validates state. */
int xdem_classify_interactions(xdem_ctx* ctx);
/* Classifier wave sizes, one per interaction type 0-12 (interaction_enum.hpp:24-57): polyhedra
 * 0-3 feature waves and 4-6 vertex-driver candidates by driver kind; spheres 0 particle pairs (one
 * per pair), 4-6 driver candidates by kind, 7-9 mesh candidates (vertex, edge, face). */
int xdem_classifier_waves(xdem_ctx* ctx, int64_t sizes[13]);
/* The rows of polyhedron wave `type` (0-3) in wave order: owner id, partner id, swap, owner and
 * partner sub-features (the key fields of ClassifierContainer, classifier_container.hpp:87-300) and
 * the feature's index in the interaction lists. XDEM_ERR_CAPACITY when cap is below the wave size. */
int xdem_classifier_wave_dump(xdem_ctx* ctx, int32_t type, int64_t cap, uint64_t* id_i, uint64_t* id_j,
                              uint8_t* swap, uint32_t* sub_i, uint32_t* sub_j, uint32_t* index, int64_t* written);
/* contact_sphere / contact_singlemat_sphere[_hooke[_none]]
 * (src/sphere/contact_force_sphere.cu.in:44-333): adds contact forces and
 * moments of every particle-particle and particle-wall interaction to f and
 * mom, and updates friction history. Deterministic, no atomics. */
int xdem_contact_sphere(xdem_ctx* ctx, const xdem_contact_params* cp, const xdem_contact_params* cp_driver,
                        double dt, int32_t form);

/* Fused engine: nsteps iterations of numerical_scheme_verlet_macro with
 * compute_force = [gravity_force, contact_sphere] (or contact_polyhedron once shapes are registered,
 * with compute_vertices and vertex_displ_over as in config_polyhedra.msp), rebuilding
 * (move_particles, chunk_neighbors_contact, nbh_sphere) whenever
 * particle_displ_over fires. Produces the same state as the operator-by-
 * operator sequence. Requires a prior neighbour build and force pass
 * (first_iteration). stats may be NULL. */
int xdem_run_steps(xdem_ctx* ctx, const xdem_scheme* scheme, int64_t nsteps, xdem_run_stats* stats);

/* ------------------------------------------------------------- polyhedra */
/* Shape table (ParticleTypeInt -> shape, as loaded by read_shape_file); particle `type` indexes it,
 * `radius` must hold the bounding radius h * (max |vertex| + minkowski) (set_fields.cu:296-310). */
int xdem_set_shapes(xdem_ctx* ctx, const xdem_shape* shapes, int32_t n);
/* compute_vertices (src/polyhedron/compute_vertices.cu:68-108): world vertices p + Q*(h v). */
int xdem_compute_vertices(xdem_ctx* ctx);
/* vertex_displ_over (src/polyhedron/vertex_displ_over.cu:43-143): *result = 1 when a vertex moved by
 * at least `threshold` since the last rebuild. */
int xdem_vertex_displ_over(xdem_ctx* ctx, double threshold, int32_t* result);
/* nbh_polyhedron (src/polyhedron/nbh_polyhedron.cpp:75-384): per owner pair (lower id) the OBB prefilter
 * and the VV/VE/VF/EE feature filters, vertex-wall candidates, friction history carried by key
 * (owner id, partner id, type, swap, owner sub, partner sub). Requires chunk_neighbors_contact. */
int xdem_nbh_polyhedron(xdem_ctx* ctx, double rcut_inc);
/* contact_polyhedron / contact_singlemat_polyhedron[_hooke[_none]]
 * (src/polyhedron/contact_force_polyhedron.cu.in:44-319): adds forces and moments of every feature
 * interaction (both particles) and vertex-wall interaction to f and mom. Deterministic, no atomics. */
int xdem_contact_polyhedron(xdem_ctx* ctx, const xdem_contact_params* cp, const xdem_contact_params* cp_driver,
                            double dt);
/* Feature interactions held (walls included) and their dump: owner id, partner id (wall index), type
 * (0 VV, 1 VE, 2 VF, 3 EE, 5 vertex-wall), swap, owner sub, partner sub, friction (3 per entry). */
int xdem_poly_interactions_count(xdem_ctx* ctx, int64_t* n);
int xdem_poly_interactions_dump(xdem_ctx* ctx, int64_t cap, uint64_t* id_i, uint64_t* id_j, uint16_t* type,
                                uint8_t* swap, uint32_t* sub_i, uint32_t* sub_j, double* friction, int64_t* written);
/* The rolling-moment history (3 per row) of the rows xdem_poly_interactions_dump writes, in the same
 * order (restart files of decks with kr != 0). */
int xdem_poly_interactions_moments(xdem_ctx* ctx, int64_t cap, double* moment, int64_t* written);

/* ------------------------------------------------------------ inspection */
/* Number of particle-particle interactions held (owner side, half list) and
 * of particle-wall interactions. */
int xdem_interactions_count(xdem_ctx* ctx, int64_t* n_pp, int64_t* n_pw);
/* Writes the interactions in device order: owner id, partner id (wall
 * index for walls), type (0 = VertexVertex, 5 = VertexSurface) and the
 * friction history (3 doubles per interaction). cap is the array length. */
int xdem_interactions_dump(xdem_ctx* ctx, int64_t cap, uint64_t* id_i, uint64_t* id_j, uint16_t* type,
                           double* friction, int64_t* written);
/* The rolling-moment history (3 per row) of the rows xdem_interactions_dump writes, in the same order
 * (interaction.hpp:33-35 keeps it beside the friction; restart files of decks with kr != 0). */
int xdem_interactions_moments(xdem_ctx* ctx, int64_t cap, double* moment, int64_t* written);
/* Restart with interaction history (replaces the history hand-over of the reference's restart: dumped
 * interactions re-enter the interaction update by key, interaction_manager.hpp:88-112 and 205-217,
 * sim_dump_writer_interaction.cpp:25-46). After nbh_sphere on the restored particles, sets the history
 * of every current interaction whose key (owner id, partner id, type: 0 sphere pair, 5 wall with the
 * wall index as partner id) is in the m saved rows (the layout of xdem_interactions_dump): friction
 * (3 per row) and, when moment is not NULL, the rolling-moment history (3 per row). Both list entries
 * of a pair take its history; current interactions without a saved row start from zero, saved rows
 * without a current interaction are dropped. *matched = rows applied. */
int xdem_interactions_load(xdem_ctx* ctx, int64_t m, const uint64_t* id_i, const uint64_t* id_j,
                           const uint16_t* type, const double* friction, const double* moment, int64_t* matched);
/* The same for polyhedra: after nbh_polyhedron on the restored particles, saved rows in the layout of
 * xdem_poly_interactions_dump (owner id, partner id or wall index, type, swap, owner and partner
 * sub-features, friction) set the friction history and, when moment is not NULL, the rolling-moment
 * history of the current feature and vertex-wall interactions with the same key. *matched = rows
 * applied. */
int xdem_poly_interactions_load(xdem_ctx* ctx, int64_t m, const uint64_t* id_i, const uint64_t* id_j,
                                const uint16_t* type, const uint8_t* swap, const uint32_t* sub_i,
                                const uint32_t* sub_j, const double* friction, const double* moment,
                                int64_t* matched);
/* Contact-state recording for analyses (replaces the ContactState the reference's contact operators
 * fill for every interaction: contact_force_sphere.cu.in:190-199 with sphere/contact_sphere.hpp:
 * 175-202 and 282-296; contact_force_polyhedron.cu.in with polyhedron/contact.hpp:160-210). With
 * on != 0, every FP64 force pass outside the engine's integrating sphere step (the operator path,
 * the last step of each xdem_run_steps call, every polyhedron pass) records per interaction dn, the
 * contact point, the normal force f - ft, the tangential force ft and the two positions. */
int xdem_set_contact_record(xdem_ctx* ctx, int32_t on);
/* The recorded rows the reference's interaction dump writes (itools/writer.hpp:35-75 with the
 * dump of contact_force_sphere.cu.in:190-199): interactions with a non-zero force or dn < 0,
 * particle pairs (sphere pairs owner side; polyhedron features in feature order) then walls. Per
 * row: first id, second id (wall index for walls), their sub-features, type (0 VV / sphere pair,
 * 1 VE, 2 VF, 3 EE, 5 wall) and 16 doubles: dn, contact point (3), f - ft (3), ft (3), first
 * position (3), second position (3; the periodic image across a boundary, zeros for walls). As in
 * the reference's pairs, a swapped polyhedron feature lists the vertex side first. Errors with
 * XDEM_ERR_STATE when recording is off or the last force pass did not record (FP32 form,
 * integrating engine step, a rebuild since). */
int xdem_contact_state_count(xdem_ctx* ctx, int64_t* n);
int xdem_contact_state_dump(xdem_ctx* ctx, int64_t cap, uint64_t* id_i, uint64_t* id_j, uint32_t* sub_i,
                            uint32_t* sub_j, uint16_t* type, double* values, int64_t* written);
/* Thermodynamic state (replaces simulation_state, src/io/simulation_state.cu:116-200, with
 * simulation_state.hpp:48-57 and itools/counter.hpp:140-153), reduced on the device in a fixed order:
 * kinetic energy sum 1/2 m v_c^2 and rotation energy sum 1/2 m w_c^2 per component, mass, particle
 * count, domain volume; interactions of the current lists (particle pairs - sphere pairs or polyhedron
 * features - counted twice as under symetric: true, driver interactions once), the active ones (dn < 0
 * or non-zero friction history) and the smallest dn over them (0 when none overlaps). Spheres: dn at the
 * current positions. Polyhedra: dn and contact of the last contact_polyhedron pass (XDEM_ERR_STATE
 * before one). */
typedef struct xdem_sim_state {
  double kinetic_energy[3];
  double rotation_energy[3];
  double mass;
  double volume;
  double min_dn;
  int64_t n_particles;
  int64_t n_interactions;
  int64_t n_active;
} xdem_sim_state;
int xdem_simulation_state(xdem_ctx* ctx, xdem_sim_state* out);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): xdem_simulation_state with
 * the same result bit for bit, but its block partials and results live in a fixed 1.5 MiB device
 * array of the library, one per GPU, held until the process ends, and it calls no hipMalloc or
 * hipFree. Limits: at most 5,032,704 particles (19,659 blocks of 256); polyhedra, also
 * ceil(features / 256) + ceil(particles / 256) <= 65,535. Larger inputs fail with
 * XDEM_ERR_CAPACITY. Every call on one GPU shares that array: calls that overlap on one GPU (other
 * host threads, other contexts on the same device) race and must be serialised by the caller. */
int xdem_simulation_state_noalloc(xdem_ctx* ctx, xdem_sim_state* out);
/* Neighbour-list statistics of the last build: half-list pairs P, list slots
 * allocated (ELL with per-wave padding) and the grid dimensions. */
int xdem_nbh_stats(xdem_ctx* ctx, int64_t* pairs, int64_t* slots, int32_t dims[3]);

/* Full (both-direction) neighbour streams of the last build, in device order: particle id, partner
 * id, periodic image code (13 = none, else (sx+1)+3(sy+1)+9(sz+1)) and owner flag. */
int xdem_streams_dump(xdem_ctx* ctx, int64_t cap, uint64_t* id_a, uint64_t* id_b, int32_t* code, uint8_t* owner,
                      int64_t* written);

/* HIP-event timing of the integrate and force launches of xdem_run_steps: on = 0 disables it, on = k
 * >= 1 times the launches of every k-th step (an event pair costs the stream a few microseconds, so a
 * sampled timing leaves the step time of the run nearly unchanged). */
int xdem_set_timing(xdem_ctx* ctx, int32_t on);

/* Device time of the last kernel launches of the fused engine, in
 * milliseconds per launch, measured with HIP events on the context stream:
 * which = 0 integrate, 1 force. Returns the average over `count` launches
 * recorded since the last reset (reset when reset != 0). */
int xdem_kernel_timing(xdem_ctx* ctx, int32_t which, int32_t reset, double* avg_ms, int64_t* count);


/* ------------------------------------------------------------------------------------------------
 * Learned surrogate of K DEM steps. No counterpart in the reference (exaDEM has no learned part).
 *
 * An edge-message graph network on the contact neighbour list of the context: 3 message-passing
 * layers of width 128, BF16 weights and activations, FP32 accumulation. Node inputs (8): radius and
 * angular velocity; edge inputs (16, one per directed list entry i -> j): contact normal, relative
 * overlap, relative surface velocity at the contact point, radius asymmetry, and the contact-gated
 * overlap and normal-velocity vectors. Outputs per particle: the mean contact force and torque over
 * the next K DEM steps in units of m0 a0 and m0 a0 r0; the surrogate step turns them into linear and
 * angular accelerations with the particle's mass and inertia.
 * The flat parameter vector (XDEM_GNN_NPARAMS floats) is, in order: W_in[128][8], b_in[128], then per
 * layer W_AB[256][128], W_E[128][16], b_E[128], W_U[128][256], b_U[128], then W_out[6][128], b_out[6]
 * (matrices row-major, out x in). Spheres only (no registered shapes).
 * ---------------------------------------------------------------------------------------------- */
#define XDEM_GNN_HIDDEN 128
#define XDEM_GNN_LAYERS 3
#define XDEM_GNN_NPARAMS 205446

typedef struct xdem_gnn xdem_gnn;

typedef struct xdem_gnn_config {
  double r0;        /* radius scale of the inputs                            */
  double v0;        /* velocity scale of the inputs                          */
  double a0;        /* acceleration scale of the outputs                     */
  double m0;        /* mass scale of the outputs                             */
  double ov_scale;  /* factor of the relative-overlap input                  */
  int32_t substeps; /* K: DEM steps replaced by one surrogate step           */
  int32_t train;    /* non-zero: keep the activations of training             */
} xdem_gnn_config;

typedef struct xdem_surrogate_stats {
  int64_t surrogate_steps; /* surrogate steps accepted                              */
  int64_t dem_steps;       /* DEM steps run (drift checks and fallback)               */
  int64_t checks;          /* drift checks made                                      */
  int64_t fallbacks;       /* checks that failed (the DEM result was kept, DEM ran on) */
  int64_t rebuilds;        /* neighbour rebuilds                                     */
  double max_err;          /* largest checked position error / radius                */
  double max_err_free;     /* same check for free flight under gravity (no network)   */
  double rms_err;          /* largest (over checks) RMS position error / radius       */
  double rms_err_free;     /* same for free flight                                    */
} xdem_surrogate_stats;

/* Creates a network bound to a context (parameters zero until set). */
int xdem_gnn_create(xdem_ctx* ctx, const xdem_gnn_config* cfg, xdem_gnn** out);
int xdem_gnn_destroy(xdem_gnn* gnn);
/* Parameters and gradients as the flat FP32 vector described above (n = XDEM_GNN_NPARAMS). */
int xdem_gnn_set_params(xdem_gnn* gnn, const float* params, int64_t n);
int xdem_gnn_get_params(xdem_gnn* gnn, float* params, int64_t n);
int xdem_gnn_get_grads(xdem_gnn* gnn, float* grads, int64_t n);
/* Inputs and forward pass at the current state of the context and its neighbour list. out (host,
 * may be NULL): [n][6] in device order. */
int xdem_gnn_forward(xdem_gnn* gnn, float* out);
/* Training targets, host [n][6] in device order (in the units of the outputs). */
int xdem_gnn_set_target(xdem_gnn* gnn, const float* target);
/* Backward pass of the mean squared error of the last forward pass (train mode); loss may be NULL. */
int xdem_gnn_backward(xdem_gnn* gnn, double* loss);
/* One Adam update with the gradients of the last backward pass. */
int xdem_gnn_adam(xdem_gnn* gnn, double lr, double beta1, double beta2, double eps);
/* One training sample from the DEM engine: forward at the current state, K fused DEM steps, targets
 * m ((v(t+K dt) - v(t)) / (K dt) - g) and I (w(t+K dt) - w(t)) / (K dt), backward, Adam. *used = 0 when
 * a neighbour rebuild fell inside the window (the DEM steps stand, no update is made). */
int xdem_gnn_train_window(xdem_gnn* gnn, const xdem_scheme* scheme, double lr, double* loss, int32_t* used);
/* Graph of the last forward pass in device order: per entry source and target index and the 16
 * edge inputs (bfloat16 bit patterns); per particle the 8 node inputs. */
int xdem_gnn_dump_graph(xdem_gnn* gnn, int64_t cap, uint32_t* src, uint32_t* dst, uint16_t* edge_in,
                        uint16_t* node_in, int64_t* n_edges);
/* Surrogate-accelerated run of nsteps surrogate steps (nsteps * K DEM steps of simulated time). Every
 * check_every-th step is checked: the surrogate step and K DEM steps start from the same state, the
 * largest position difference over radius is compared with tol, and the DEM state is kept. A
 * failed check makes the next check_every steps plain DEM. Needs ids that are a permutation of
 * 0..n-1. */
int xdem_surrogate_run(xdem_gnn* gnn, const xdem_scheme* scheme, int64_t nsteps, int32_t check_every, double tol,
                       xdem_surrogate_stats* stats);

/* Stationary triangle-mesh driver acting on spheres (register_stl_mesh, src/driver/register_stl_mesh.cpp:30-140;
 * contact ContactLawRShapeDriverFunc, src/sphere/include/exaDEM/sphere/contact_sphere.hpp:298-396; candidates
 * detection_sphere_driver, src/sphere/include/exaDEM/sphere/nbh_sphere.hpp:25-86). vertices [nv][3] in the
 * mesh frame (world vertex = center + quat (w, x, y, z) applied to it), edges [ne][2] and triangles [nf][3]
 * as vertex indices, Minkowski radius, driver index id (partner id of its interactions, types 7 vertex,
 * 8 edge, 9 face). max_candidates: slots per sphere for candidate features (a neighbour build with more
 * fails with XDEM_ERR_CAPACITY). nv = 0 removes the mesh. Takes effect at the next nbh_sphere; both
 * force forms evaluate its contacts in FP64; spheres only. */
int xdem_set_mesh(xdem_ctx* ctx, int32_t id, const double* vertices, int64_t nv, const uint32_t* edges, int64_t ne,
                  const uint32_t* faces, int64_t nf, double minkowski, const double* center, const double* quat,
                  int32_t max_candidates);
/* Mesh candidates of the current neighbour list with their history, sorted by (owner id, type, feature):
 * owner id, feature index, type (7, 8, 9), friction [3], moment [3] per row; *written = row count (rows
 * beyond cap: XDEM_ERR_CAPACITY, nothing written). Null output arrays are skipped. */
int xdem_mesh_interactions(xdem_ctx* ctx, int64_t cap, uint64_t* id_i, uint32_t* sub_j, uint16_t* type,
                           double* friction, double* moment, int64_t* written);
/* Restart: m saved mesh rows (as xdem_mesh_interactions writes them; moment may be null) handed to the
 * current candidate lists by key (owner id, type, feature) after a neighbour build; *matched = rows that
 * found their candidate. No counterpart call in the reference, whose restart dump carries the history
 * (sim_dump_writer_interaction.cpp:25-46). */
int xdem_mesh_interactions_load(xdem_ctx* ctx, int64_t m, const uint64_t* id_i, const uint32_t* sub_j,
                                const uint16_t* type, const double* friction, const double* moment, int64_t* matched);

#ifdef __cplusplus
}
#endif

#endif /* XDEM_H */
