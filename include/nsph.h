/*
 * nsph.h -- C ABI of libnsph.so, the gfx950 (MI355X) SPH particle step.
 *
 * Every entry point returns an int status (NSPH_OK == 0).  On failure the
 * message is available from nsph_last_error().  No C++ exception crosses
 * this boundary.  Pointers named d_* are device pointers (HBM) owned by the
 * caller; the library owns only its scratch, allocated in the *_create and
 * *_reserve calls and never inside nsph_solver_step / nsph_grid_build after
 * a matching reserve.  All work is enqueued on the context's stream.
 *
 * Reference interfaces replaced (paths relative to the Newton source tree):
 *   HashGrid      -> warp.HashGrid as created in newton/_src/sim/builder.py:9851-9855,
 *                    reserved in newton/_src/solvers/xpbd/solver_xpbd.py:136-139,
 *                    built in solver_xpbd.py:315-319, consumed through
 *                    hash_grid_point_id/hash_grid_query/hash_grid_query_next in
 *                    newton/_src/solvers/semi_implicit/kernels_contact.py:54-73 and
 *                    newton/_src/solvers/xpbd/kernels.py:244-264.
 *   Solver step   -> SolverBase.step(state_in, state_out, control, contacts, dt)
 *                    (newton/_src/solvers/solver.py:301-316) with the particle
 *                    integrator of solver.py:10-47 (integrate_particles).
 */
#ifndef NSPH_H
#define NSPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NSPH_ABI_VERSION 3

enum nsph_status {
    NSPH_OK = 0,
    NSPH_ERR_INVALID = 1,       /* bad argument or handle                              */
    NSPH_ERR_HIP = 2,           /* HIP runtime error                                   */
    NSPH_ERR_CAPACITY = 3,      /* n above the reserved capacity / neighbour overflow  */
    NSPH_ERR_NOT_BUILT = 4,     /* grid queried before build                           */
    NSPH_ERR_NONFINITE = 5,     /* non-finite value in the step output                 */
    NSPH_ERR_NOT_CONVERGED = 6  /* pressure solve hit max_iters in at least one world  */
};

/* Particle flag bits, same values as ParticleFlags (newton/_src/geometry/flags.py:8-14). */
#define NSPH_PARTICLE_ACTIVE 1

typedef struct nsph_ctx nsph_ctx;
typedef struct nsph_grid nsph_grid;
typedef struct nsph_solver nsph_solver;
typedef struct nsph_mlp nsph_mlp;

/* ---- context --------------------------------------------------------------------- */

/* device: HIP ordinal; stream: hipStream_t (NULL = default stream). */
int nsph_ctx_create(int device, void* stream, nsph_ctx** out);
int nsph_ctx_set_stream(nsph_ctx* ctx, void* stream);
int nsph_ctx_synchronize(nsph_ctx* ctx);
int nsph_ctx_destroy(nsph_ctx* ctx);
/* Last error message of the calling thread ("" if none). */
const char* nsph_last_error(void);
int nsph_abi_version(void);

/* ---- HashGrid (replaces warp.HashGrid; builder.py:9853, solver_xpbd.py:139,319) ---- */

/* A periodic hash grid of dim_x*dim_y*dim_z cells: cell = floor(x / radius) mod dim,
 * as the reference's HashGrid(128,128,128).  Each dim must be >= 3. */
int nsph_grid_create(nsph_ctx* ctx, int dim_x, int dim_y, int dim_z, nsph_grid** out);
/* Pre-allocate for n points (solver_xpbd.py:139 particle_grid.reserve). */
int nsph_grid_reserve(nsph_grid* grid, int n);
/* Hash d_points (n x 3 float32, particle order) with cell width `radius`
 * (solver_xpbd.py:319 particle_grid.build(points, radius)).  The sort is stable:
 * points of one cell keep ascending id order. Grows the reservation if n exceeds it. */
int nsph_grid_build(nsph_grid* grid, const float* d_points, int n, float radius);
/* d_ids[k] = id of the k-th point in cell-sorted order (hash_grid_point_id). */
int nsph_grid_point_ids(const nsph_grid* grid, int* d_ids);
int nsph_grid_num_points(const nsph_grid* grid, int* n);
/* Neighbour lists of grid points: for each point i with (flags[i] & ACTIVE) (all points if
 * d_flags is NULL), every j != i with ACTIVE flag and
 *     dx*dx + dy*dy + dz*dz < radius*radius     (FP32, x_i - x_j, left-to-right, no FMA)
 * in the candidate order of the grid query.  Lists are slot-major: d_idx[k*n + i],
 * k < max_neighbors; d_counts[i] is the true count (may exceed max_neighbors, in which
 * case the call returns NSPH_ERR_CAPACITY after filling the first max_neighbors).
 * Mirrors the consumer loop of kernels_contact.py:68-73 / xpbd/kernels.py:257-264. */
int nsph_grid_neighbors(const nsph_grid* grid, const int* d_flags, float radius,
                        int max_neighbors, int* d_counts, int* d_idx);
/* Same for arbitrary query points d_q (nq x 3): all grid points j with |q - x_j|^2 < r^2. */
int nsph_grid_query(const nsph_grid* grid, const float* d_q, int nq, float radius,
                    int max_neighbors, int* d_counts, int* d_idx);
int nsph_grid_destroy(nsph_grid* grid);
/* Device view of the last build, for in-kernel queries (include/nsph_device.h: point_id,
 * query, query_next, replacing hash_grid_point_id / hash_grid_query / hash_grid_query_next
 * of kernels_contact.py:54-73 and xpbd/kernels.py:244-264).  Valid until the next build. */
typedef struct nsph_grid_device {
    const int* cells;          /* 2 per cell: first and end sorted position; cell index
                                  (cz * dim_y + cy) * dim_x + cx of the periodic hash */
    const unsigned int* ids;   /* point id at each sorted position                        */
    const float* points;       /* 4 per sorted position: x, y, z of the build, 0          */
    int dim[3];
    float inv_cell;            /* 1 / cell width (the build radius)                       */
    int n;                     /* points of the build                                     */
} nsph_grid_device;
int nsph_grid_device_view(const nsph_grid* grid, nsph_grid_device* out);
/* A consumer written on nsph_grid_device.h: d_counts[k] = number of grid points within
 * `radius` of query point k (FP32 (dx*dx + dy*dy) + dz*dz < radius^2). */
int nsph_grid_count(const nsph_grid* grid, const float* d_q, int nq, float radius, int* d_counts);

/* ---- the reference's particle kernels on the HIP grid ------------------------------ */

/* integrate_particles (newton/_src/solvers/solver.py:10-47, launched at solver.py:266-299):
 * d_f may be NULL (zero force); d_gravity is a device array of n_gravity vec3 (model.gravity). */
int nsph_particles_integrate(nsph_ctx* ctx, int n, const float* d_x, const float* d_v, const float* d_f,
                             const float* d_inv_mass, const int* d_flags, const int* d_world,
                             const float* d_gravity, int n_gravity, float dt, float v_max,
                             float* d_x_out, float* d_v_out);
/* solve_particle_particle_contacts (newton/_src/solvers/xpbd/kernels.py:225-289, launched at
 * solver_xpbd.py:420-441) over a grid built from the same n particles; adds each active
 * particle's correction into d_deltas[i] (the caller zeroes it, solver_xpbd.py:379-383). */
int nsph_xpbd_particle_contacts(nsph_grid* grid, int n, const float* d_x, const float* d_v,
                                const float* d_inv_mass, const float* d_radius, const int* d_flags,
                                float mu, float cohesion, float max_radius, float dt, float relaxation,
                                float* d_deltas);
/* apply_particle_deltas (newton/_src/solvers/xpbd/kernels.py:776-806). */
int nsph_apply_particle_deltas(nsph_ctx* ctx, int n, const float* d_x_orig, const float* d_x_pred,
                               const int* d_flags, const float* d_deltas, float dt, float v_max,
                               float* d_x_out, float* d_v_out);
/* eval_particle_contact (newton/_src/solvers/semi_implicit/kernels_contact.py:35-84, launched at
 * kernels_contact.py:559-579) over a grid built from the same n particles: overwrites d_f[i]
 * of active particles, as the reference does. */
int nsph_eval_particle_contact(nsph_grid* grid, int n, const float* d_x, const float* d_v, const float* d_radius,
                               const int* d_flags, float ke, float kd, float kf, float mu, float cohesion,
                               float max_radius, float* d_f);

/* ---- particle-shape soft contacts, static shapes (body = -1) ------------------------ */

/* Workspace of nsph_soft_contacts for n particles, in ints: a count per particle and one per
 * workgroup of 256 particles. */
#define NSPH_SOFT_WORK_INTS(n) ((n) + ((n) + 255) / 256)

/* create_soft_contacts (newton/_src/geometry/kernels.py:1000-1141, launched at
 * collide.py:1194-1230) for plane, sphere, capsule, cylinder and box shapes.  Shapes are
 * d_shape_xform[7 s] = (px, py, pz, qx, qy, qz, qw), d_shape_scale[3 s], type (GeoType
 * values), world and flags (ShapeFlags).  Contacts are ordered by (particle, shape);
 * d_start[n + 1] receives each particle's first contact slot, *d_count the total (which may
 * exceed `capacity`; slots at or beyond it are not written).  d_work: NSPH_SOFT_WORK_INTS(n)
 * ints of scratch. */
int nsph_soft_contacts(nsph_ctx* ctx, int n, const float* d_x, const float* d_radius, const int* d_flags,
                       const int* d_world, int n_shapes, const float* d_shape_xform, const int* d_shape_type,
                       const float* d_shape_scale, const int* d_shape_world, const int* d_shape_flags,
                       float margin, int capacity, int* d_work, int* d_start, int* d_count, int* d_particle,
                       int* d_shape, float* d_body_pos, float* d_body_vel, float* d_normal);
/* nsph_soft_contacts without a workspace argument and without any allocation.  No
 * counterpart in the reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The workspace is a
 * library-owned device array of NSPH_SOFT_WORK_INTS(NSPH_SOFT_STATIC_MAX_PARTICLES) ints
 * (about 16 MB), one per GPU, resolved by nsph_ctx_create and held until the process ends;
 * n above NSPH_SOFT_STATIC_MAX_PARTICLES fails with NSPH_ERR_CAPACITY.  Every call on a GPU
 * shares that GPU's array (each call writes all of the workspace it reads), so calls that
 * overlap on one GPU (different streams or threads) race.  Results equal nsph_soft_contacts'. */
#define NSPH_SOFT_STATIC_MAX_PARTICLES 4000000
int nsph_soft_contacts_static(nsph_ctx* ctx, int n, const float* d_x, const float* d_radius, const int* d_flags,
                              const int* d_world, int n_shapes, const float* d_shape_xform, const int* d_shape_type,
                              const float* d_shape_scale, const int* d_shape_world, const int* d_shape_flags,
                              float margin, int capacity, int* d_start, int* d_count, int* d_particle, int* d_shape,
                              float* d_body_pos, float* d_body_vel, float* d_normal);
/* solve_particle_shape_contacts (newton/_src/solvers/xpbd/kernels.py:114-222, launched at
 * solver_xpbd.py:386-418) for static shapes: adds each active particle's corrections over
 * its contacts into d_deltas[i]. */
int nsph_xpbd_shape_contacts(nsph_ctx* ctx, int n, const float* d_x, const float* d_v, const float* d_inv_mass,
                             const float* d_radius, const int* d_flags, const float* d_shape_mu, float particle_mu,
                             float particle_ka, const int* d_start, int capacity, const int* d_shape,
                             const float* d_body_pos, const float* d_body_vel, const float* d_normal, float dt,
                             float relaxation, float* d_deltas);
/* eval_particle_body_contact (newton/_src/solvers/semi_implicit/kernels_contact.py:144-262,
 * launched at kernels_contact.py:644-687) for static shapes: subtracts each active particle's
 * contact forces from d_f[i]. */
int nsph_eval_shape_contacts(nsph_ctx* ctx, int n, const float* d_x, const float* d_v, const float* d_radius,
                             const int* d_flags, const float* d_shape_ke, const float* d_shape_kd,
                             const float* d_shape_kf, const float* d_shape_mu, float ke, float kd, float kf, float mu,
                             float ka, const int* d_start, int capacity, const int* d_shape, const float* d_body_pos,
                             const float* d_body_vel, const float* d_normal, float* d_f);

/* Shape corrections after an SPH step (no reference counterpart: SolverSPH accepts contacts
 * the way SolverXPBD does, solver_xpbd.py:386-418): particles with a nonzero correction move
 * by it and gain correction / dt of velocity; the others are untouched. */
int nsph_apply_shape_deltas(nsph_ctx* ctx, int n, const int* d_flags, const float* d_deltas, float dt, float v_max,
                            float* d_x, float* d_v);

/* ---- SPH solver (behind SolverBase.step, solver.py:301-316) ------------------------ */

enum nsph_warm_start { NSPH_WARM_ZERO = 0, NSPH_WARM_PREV = 1, NSPH_WARM_MLP = 2 };

typedef struct nsph_model_desc {
    int n_particles;          /* model.particle_count                                 */
    int n_worlds;             /* model.world_count (>= 1); world -1 maps to world 0   */
    const float* d_q;         /* model.particle_q   (N x 3) positions at solver creation */
    const float* d_qd;        /* model.particle_qd  (N x 3)                           */
    const float* d_mass;      /* model.particle_mass (N)                              */
    const float* d_inv_mass;  /* model.particle_inv_mass (N)                          */
    const int* d_flags;       /* model.particle_flags (N)                             */
    const int* d_world;       /* model.particle_world (N), NULL = all world 0        */
    const float* h_gravity;   /* host, n_worlds x 3 (model.gravity)                   */
    float v_max;              /* model.particle_max_velocity                          */
} nsph_model_desc;

typedef struct nsph_sph_params {
    const double* h_support;        /* host, n_worlds: kernel support radius H [m]        */
    const double* h_boundary_volume;/* host, n_worlds: rest volume of a boundary particle */
    const double* h_domain_lo;      /* host, n_worlds x 3: cell-table box lower corner    */
    const double* h_domain_hi;      /* host, n_worlds x 3: cell-table box upper corner    */
    double rest_density;            /* rho_0 [kg/m^3]                                     */
    double viscosity;               /* kinematic viscosity nu [m^2/s]                     */
    double omega;                   /* relaxed-Jacobi factor (IISPH: 0.5)                 */
    double tolerance;               /* mean compression / rho_0 stop test (1e-3 = 0.1%)   */
    int min_iters;                  /* IISPH: 2                                           */
    int max_iters;
    int max_neighbors;              /* fluid-fluid list capacity per particle             */
    int warm_start;                 /* enum nsph_warm_start                               */
    double skin;                    /* 0: lists searched every step; else in [0.01, 1]:
                                       candidates within H (1 + skin) of each particle's
                                       search position are kept, and a world searches
                                       again when one of its particles is farther than
                                       about H skin / 2 from its search position          */
    int max_skin_neighbors;         /* capacity of the kept candidates (skin > 0), in
                                       [max_neighbors, 8192]                              */
} nsph_sph_params;

typedef struct nsph_step_report {
    int status;               /* NSPH_OK / NSPH_ERR_CAPACITY / NSPH_ERR_NOT_CONVERGED / NSPH_ERR_NONFINITE */
    int max_iterations;       /* max over worlds of Jacobi updates applied            */
    int sum_iterations;       /* sum over worlds                                      */
    int unconverged_worlds;   /* worlds that stopped at max_iters                     */
    int max_neighbor_count;   /* max fluid-fluid neighbours of any particle           */
    int neighbor_overflow;    /* 1 if any list (or kept-candidate set) exceeded its capacity */
    int searched_worlds;      /* worlds whose neighbour candidates were searched anew   */
    double max_density_error; /* max over worlds of the final mean compression / rho_0 */
} nsph_step_report;

/* Create the solver: classifies particles (fluid = ACTIVE and inv_mass > 0; static
 * boundary = ACTIVE, inv_mass == 0 and zero velocity in the model; moving boundary = ACTIVE,
 * inv_mass == 0 and a velocity in the model; the rest inactive) and hashes the static
 * boundary once.  Moving boundary particles are integrated like every zero-mass particle
 * (x + v dt, solver.py:24-46) and hashed from state_in at each step, where their velocity
 * enters the source term as psi sum_b (v_i - v_b) . gradW_ib. */
int nsph_solver_create(nsph_ctx* ctx, const nsph_model_desc* model, const nsph_sph_params* params,
                       nsph_solver** out);
/* The particle arrays of one State (newton/_src/sim/state.py:101-108), particle order, device
 * pointers.  State in: */
typedef struct nsph_state_in {
    const float* d_q;   /* particle_q  (N x 3)                                                 */
    const float* d_qd;  /* particle_qd (N x 3)                                                 */
    const float* d_f;   /* particle_f  (N x 3) or NULL: forces the caller applied to the
                           particles, integrated together with the SPH force as
                           integrate_particles integrates particle_f (solver.py:24-46)        */
    const float* d_p;   /* sph_pressure (N) or NULL: previous-step pressure, the input of the
                           "prev" and "mlp" warm starts (NULL: the solver's own copy of its last
                           step's pressure); entries of non-fluid particles are not read        */
    const float* d_xb;  /* sph_search_q (N x 3) or NULL: positions around which the kept
                           neighbour candidates were searched (NULL: the solver's own copy of
                           its last step's; NaN entries: none).  A world whose particles all
                           lie within the rebuild distance of them uses the candidates of
                           these positions, else it searches around the current positions,
                           so the step's result is a function of (q, qd, f, p, xb) alone      */
} nsph_state_in;

/* State out: */
typedef struct nsph_state_out {
    float* d_q;   /* particle_q  (N x 3)                                                       */
    float* d_qd;  /* particle_qd (N x 3)                                                       */
    float* d_f;   /* NULL, or receives the SPH force (viscosity + pressure, N) of fluid
                     particles and 0 elsewhere                                                 */
    float* d_p;   /* NULL, or receives the step's converged pressure (sph_pressure; 0 for
                     non-fluid particles)                                                      */
    float* d_xb;  /* NULL, or receives the step's search positions (sph_search_q; the
                     position of a non-fluid particle); not written when skin = 0            */
} nsph_state_out;

/* One step state_in -> state_out (SolverBase.step, solver.py:301-316).  Inactive particles:
 * q_out = q_in, qd_out untouched (solver.py:27-29).  state_in and state_out must not alias.
 * check_every > 0: after min_iters Jacobi iterations the step's end is enqueued guarded by
 * "every world converged" and the host synchronises once; only if a world has not converged
 * does it iterate on, reading the convergence count every check_every iterations, and enqueue
 * the end again (iteration counts are decided per world on the device, so results do not
 * depend on this schedule).  check_every = 0: launch max_iters early-exit iterations without
 * any host read; the step is then capture-safe.  report may be NULL; with check_every = 0 and
 * no report the step does not synchronise. */
int nsph_solver_step(nsph_solver* s, const nsph_state_in* in, const nsph_state_out* out, float dt,
                     int check_every, nsph_step_report* report);
/* Capture one step from state_in to state_out (fixed pointers and dt) into HIP graphs:
 * prologue + iter_slots Jacobi iterations, and the epilogue (under NSPH_SOLVE_WORLDS the
 * schedule's pass_iters iterations and the persistent launch instead; iter_slots is then
 * ignored). Capture runs on a private stream; nothing is allocated on the device during
 * capture or replay. *handle names the graph pair for nsph_solver_graph_step.  The
 * warm-start mode, MLP and solve schedule in force at capture are part of the graph. */
int nsph_solver_graph_capture(nsph_solver* s, const nsph_state_in* in, const nsph_state_out* out, float dt,
                              int iter_slots, int* handle);
/* Replay a captured step on the context stream. Under NSPH_SOLVE_PASSES, when
 * iter_slots < max_iters the host reads the still-iterating world count once after the slots
 * and, if needed, finishes the solve with eager iterations before the epilogue; under
 * NSPH_SOLVE_WORLDS nothing is read.  Results equal nsph_solver_step's.  Fails with
 * NSPH_ERR_INVALID when the solver's warm-start mode, MLP or solve schedule differs from the
 * capture's. */
int nsph_solver_graph_step(nsph_solver* s, int handle, nsph_step_report* report);
/* Free a captured step (its handle is not reused). */
int nsph_solver_graph_destroy(nsph_solver* s, int handle);
/* Collect the report of the last step (synchronises the stream). */
int nsph_solver_report(nsph_solver* s, nsph_step_report* report);
/* Per-particle fields of the last step in particle order (any pointer may be NULL):
 * density rho, pressure p, fluid-fluid neighbour count. Non-fluid entries are 0. */
int nsph_solver_fields(nsph_solver* s, float* d_rho, float* d_p, int* d_nbr_count);
/* Per-world Jacobi iterations and final mean compression of the last step (host arrays). */
int nsph_solver_world_stats(nsph_solver* s, int* h_iters, double* h_err, int* h_state);
/* Fluid-fluid neighbour lists of the last step in particle ids, slot-major
 * d_idx[k*N + id] for k < max_neighbors, d_counts[id] (0 for non-fluid). */
int nsph_solver_neighbors(nsph_solver* s, int* d_counts, int* d_idx);
int nsph_solver_set_warm_start(nsph_solver* s, int mode);
/* Search every world's candidates around the current positions at the next eager step
 * (captured steps never force). */
int nsph_solver_rebuild_lists(nsph_solver* s);
/* Forget the previous-step pressure (warm start input). */
int nsph_solver_reset_pressure(nsph_solver* s);
/* Warm-start pressure p0 of the last step in particle order (the pressure the Jacobi
 * iterations started from: 0, 0.5 p_prev or the MLP's prediction; 0 for non-fluid). */
int nsph_solver_warm_pressure(nsph_solver* s, float* d_p0);
int nsph_solver_num_fluid(const nsph_solver* s, int* n_fluid, int* n_boundary);
/* Workgroup size of the gather kernels (64, 128 or 256 threads; a 256-particle chunk is split
 * over 256 / gather_block workgroups). Scheduling only: results are identical for every value.
 * Set by the autotuner; captured graphs keep the value they were captured with. */
int nsph_solver_set_launch(nsph_solver* s, int gather_block);
/* Variant of the neighbour-search kernel: list quads per particle staged in LDS, which also
 * sets the workgroups per CU (six of 256 up to 6 quads, five at 8).  Compiled variants: 6 (the
 * default), 4, 8, 2; anything else fails with NSPH_ERR_INVALID.  Scheduling only: results are
 * identical for every variant.  Set by the autotuner. */
int nsph_solver_set_search(nsph_solver* s, int lds_quads);
/* Schedule of the pressure solve (replaces the per-pass launches behind SolverBase.step,
 * solver.py:301-316; batched worlds as model.py:170-187 groups them).
 * NSPH_SOLVE_PASSES (the default): one launch per pass over all worlds, the host reading the
 * still-iterating count once per step.  NSPH_SOLVE_WORLDS: one persistent launch, each
 * workgroup (`block` = 256, 512 or 1024 threads) taking whole worlds from a device queue,
 * largest first, iterates every world to its own stop test.  With pass_iters = 0 it runs
 * advection, source, the Jacobi iterations and the forces (with the learned warm start,
 * advection and the network run as their own launches over all worlds before it); with
 * pass_iters = k > 0 the first k iterations run as per-pass launches over all worlds and the
 * persistent launch only finishes the worlds still iterating.  Either way no host read, and a
 * captured step holds no iteration slots beyond the k.
 * Scheduling only: results are identical for both.  Set by the autotuner; captured graphs
 * must be captured again after a change (nsph_solver_graph_step fails otherwise). */
enum nsph_solve_mode { NSPH_SOLVE_PASSES = 0, NSPH_SOLVE_WORLDS = 1 };
int nsph_solver_set_solve(nsph_solver* s, int mode, int block, int pass_iters);
/* Per-kernel-class GPU timing with HIP events recorded on the context stream around each
 * launch group of the classes whose bit is set in `enable` (bit k = class k; -1 = all; 0 = off,
 * the default). Changing it resets the totals. nsph_solver_timing
 * synchronises and returns accumulated milliseconds and launch counts per class
 * (nsph_kernel_count() classes named by nsph_kernel_name). No reference counterpart:
 * the reference's timers are wall-clock ScopedTimer brackets (solver_xpbd.py:302,372). */
int nsph_solver_set_timing(nsph_solver* s, int enable);
int nsph_solver_timing(nsph_solver* s, double* h_ms, int* h_launches);
int nsph_kernel_count(void);
const char* nsph_kernel_name(int k);
/* Jacobi iterations launched by the last step (>= the max over worlds of iterations
 * applied when check_every > 1 or 0; the extra ones exit at once); -1 after a whole-world
 * solve, which iterates inside one launch. */
int nsph_solver_iterations_launched(const nsph_solver* s, int* n);
int nsph_solver_destroy(nsph_solver* s);

/* ---- learned warm start: per-particle MLP 6 -> 64 -> 64 -> 1 (BF16 MFMA, FP32 accumulate) ----
 * No counterpart in the reference (which has no pressure solve).  The closest analog is the
 * ONNX MLP of ControllerNeuralMLP (newton/_src/actuators/controllers/controller_neural_mlp.py:131,
 * run by newton/_src/utils/onnx_runtime.py:69-87).
 * Parameters, flat FP32 (nsph_mlp_num_params() = 4673): W1[6][64], b1[64], W2[64][64], b2[64],
 * w3[64], b3.  Feature rows are 8 floats (6 features, 2 ignored). */
int nsph_mlp_num_params(void);
int nsph_mlp_create(nsph_ctx* ctx, nsph_mlp** out);
int nsph_mlp_destroy(nsph_mlp* m);
/* Host parameter copy in/out (set also resets the Adam moments). */
int nsph_mlp_set_params(nsph_mlp* m, const float* h_params);
int nsph_mlp_get_params(nsph_mlp* m, float* h_params);
/* d_y[i] = MLP(d_x[8 i .. 8 i + 5]) for i < n. */
int nsph_mlp_forward(nsph_mlp* m, const float* d_x, int n, float* d_y);
/* Mean squared error of MLP(x) against d_t and its gradient (kept for nsph_mlp_adam; also
 * copied to d_grad[0 .. 4672] when non-NULL, d_grad[4673] = sum of squared errors).
 * *h_loss (optional, synchronises) receives the mean squared error. Deterministic. */
int nsph_mlp_grad(nsph_mlp* m, const float* d_x, const float* d_t, int n, float* h_loss,
                  float* d_grad);
/* One Adam step (FP32 master weights) with the last gradient; refreshes the BF16 copies. */
int nsph_mlp_adam(nsph_mlp* m, float lr, float beta1, float beta2, float eps);
/* One 16x16x32 BF16 MFMA product C = A B (A 16x32, B 32x16 row-major host arrays, rounded to
 * BF16) through the operand layout the MLP kernels use; a self-check of that layout. */
int nsph_mfma_selftest(nsph_ctx* ctx, const float* h_a, const float* h_b, float* h_c);
/* Attach an MLP to a solver (NULL detaches); warm_start NSPH_WARM_MLP then predicts p0 from
 * the particle features before the Jacobi iterations, which still run to the tolerance. */
int nsph_solver_set_mlp(nsph_solver* s, nsph_mlp* m);
/* Features (n_fluid x 8) and training target (p - 0.5 p_prev) / p_unit (n_fluid) of the last
 * step, cell-sorted order; the MLP warm start is p0 = max(0.5 p_prev + y p_unit, 0). */
int nsph_solver_mlp_features(nsph_solver* s, float* d_x, float* d_target);

#ifdef __cplusplus
}
#endif

#endif /* NSPH_H */
