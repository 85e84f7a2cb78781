/*
 * rbh.h -- C ABI of librbh, the MI355X (gfx950) implementation of the
 * prox_rigid_bodies time step of RAINBOW (diku-dk/RAINBOW).
 *
 * The reference is pure Python; its plug-in point for this path is the
 * engine's stepper object (rainbow/simulators/prox_rigid_bodies/api.py:42-53,
 * called by api.simulate at api.py:676-693).  These entry points are what a
 * ctypes stepper bound to that plug-in point calls; each comment names the
 * reference code whose work the call takes over.
 *
 * Conventions: every function returns 0 on success or a negative RBH_E_*
 * code; the message of the last failure is rbh_last_error(ctx).  Arrays are
 * caller-owned host memory, C order, FP64 unless stated.  Body indices are
 * local to their scene (insertion order of engine.bodies).  A context owns
 * one HIP stream on one device and is not thread-safe.
 */
#ifndef RBH_H
#define RBH_H

#ifdef __cplusplus
extern "C" {
#endif

#define RBH_OK 0
#define RBH_E_ARG -1        /* invalid argument (ValueError / RuntimeError in the reference API) */
#define RBH_E_HIP -2        /* HIP runtime failure */
#define RBH_E_MEMORY -3     /* device memory or HBM budget exhausted */
#define RBH_E_CAPACITY -4   /* internal capacity exceeded and could not grow */
#define RBH_E_COLORS -5     /* colouring needed more than rbh_max_colors() colours */
#define RBH_E_STATE -6      /* call out of order (e.g. contacts of a scene never stepped) */

/* proximal_solver schemes (proximal_contact/prox_solvers.py:6-37) */
#define RBH_SCHEME_GAUSS_SEIDEL 0          /* serial, canonical contact order (contacts.py:66-107) */
#define RBH_SCHEME_PARALLEL_GAUSS_SEIDEL 1 /* random-colour order, colours swept in parallel */
#define RBH_SCHEME_PARALLEL_JACOBI 2       /* jacobi_solver.py:41-57 */
#define RBH_SCHEME_PARALLEL_JACOBI_HYBRID 3/* jacobi_solver.py:60-79 */

/* friction models */
#define RBH_FRICTION_ISOTROPIC 0           /* prox_operators.py:14-34 with mu[0] (contacts.py:255) */
#define RBH_FRICTION_ANISOTROPIC 1         /* proximal_contact/proximal_operators.py:42-139 */

/* rbh_step flags */
#define RBH_STEP_COUNT_OVERLAPS 1          /* count overlapping triangle pairs (profile key number_of_overlaps) */
#define RBH_STEP_TIMING 2                  /* record per-stage GPU times for rbh_get_timings */
#define RBH_STEP_GRAPH 4                   /* capture the launch sequence in a hipGraph and replay it */
#define RBH_STEP_FORCE_SINGLE 8            /* solve each scene on one workgroup */
#define RBH_STEP_FORCE_MULTI 16            /* solve a lone scene on many workgroups (islands in parallel) */

typedef struct rbh_ctx rbh_ctx;

/* Numerical parameters of one scene; field for field the reference's
 * types.Parameters (types.py:356-418) plus the knobs of this build. */
typedef struct rbh_params {
  double time_step;                       /* Parameters.time_step */
  double envelope;                        /* Parameters.envelope */
  double absolute_tolerance;              /* Parameters.absolute_tolerance */
  double relative_tolerance;              /* Parameters.relative_tolerance */
  double nu_reduce;                       /* Parameters.nu_reduce */
  double nu_increase;                     /* Parameters.nu_increase */
  double too_small_merit_change;          /* Parameters.too_small_merit_change */
  double contact_optimization_tolerance;  /* Parameters.contact_optimization_tolerance */
  double gap_reduction;                   /* Parameters.gap_reduction */
  double min_gap_value;                   /* Parameters.min_gap_value */
  double max_gap_value;                   /* Parameters.max_gap_value */
  double ellipsoid_tolerance;             /* Parameters.ellipsoid_tolerance */
  double ellipsoid_expansion;             /* Parameters.ellipsoid_expansion */
  int max_iterations;                     /* Parameters.max_iterations */
  int contact_optimization_max_iterations;/* Parameters.contact_optimization_max_iterations */
  int ellipsoid_max_iterations;           /* Parameters.ellipsoid_max_iterations */
  int use_bounce;                         /* Parameters.use_bounce */
  int use_pre_stabilization;              /* Parameters.use_pre_stabilization */
  int proximal_solver;                    /* RBH_SCHEME_* (soft-body Parameters.proximal_solver) */
  int friction_model;                     /* RBH_FRICTION_* */
  int use_post_stabilization;             /* Parameters.use_post_stabilization (steppers.py:15-36, 90-91) */
  unsigned long long coloring_seed;       /* seed of the random colouring */
} rbh_params;

/* One scene (one reference Engine): bodies in engine.bodies order. */
typedef struct rbh_scene_desc {
  int n_bodies;
  const int* shape;            /* shape id (rbh_shape_upload) per body */
  const int* fixed;            /* RigidBody.is_fixed */
  const int* material;         /* index into the interaction table */
  const double* mass;          /* RigidBody.mass */
  const double* inertia;       /* 3 per body: RigidBody.inertia (principal, body frame) */
  const double* gravity;       /* 3 per body: sum over the body's Gravity forces of g * up (types.py:204-221) */
  const double* damping;       /* 2 per body: sums of Damping alpha and beta (types.py:239-256) */
  int n_materials;
  const double* interactions;  /* n_materials^2 x 4: mu[0], mu[1], mu[2], epsilon (types.py:20-53) */
  rbh_params params;
} rbh_scene_desc;

/* Per-scene record of the last step (the profile dict keys of steppers.py:95-106,
 * collision_detection.py:80-83 and gauss_seidel_solver.py:74-101). */
typedef struct rbh_step_log {
  int step_index;              /* steps committed so far */
  int contact_points;          /* len(engine.contact_points) */
  int iterations;              /* solver 'iterations' */
  int exitcode;                /* solver 'exitcode' (0 max iterations, 1 absolute, 2 relative) */
  int rejects;                 /* rejected (rolled back) sweeps */
  int sweeps;                  /* sweeps executed, rejected ones included */
  int colors;                  /* colours used by the sweep order */
  int solved;                  /* 1 if the contact problem was solved this step */
  unsigned long long number_of_overlaps; /* overlapping triangle pairs (RBH_STEP_COUNT_OVERLAPS) */
  unsigned long long fw_cell_evals;      /* trilinear cell evaluations in Frank-Wolfe */
  double natural_residual;     /* ||x - prox(x - r0 (A x + b))||_inf / max(1, ||x||_inf), FP64 */
  double max_penetration;      /* Contacts.get_largest_penetration */
  double final_merit;          /* last accepted merit ||x - sol||_inf */
  int post_iterations;         /* post-stabilisation solve 'post_stabilization_iterations' (0: not run) */
  int post_exitcode;           /* 'post_stabilization_exitcode'; -1 when the step ran no post-stabilisation */
} rbh_step_log;

/* Launch shapes of the step (0 = default).  No reference counterpart: these
 * are the knobs an autotuner sets. */
typedef struct rbh_tuning {
  int unit_size;          /* contact positions per colouring unit window (default 64) */
  int lanes_per_scene;    /* sweep lanes of a batched scene, about one per wavefront (default 4) */
  int resident_contacts;  /* most contacts of a solve group kept in LDS (multiple of 64; at most 384, FP64 192) */
  int resident_bodies;    /* most bodies of a solve group kept in LDS (multiple of 64; same bounds) */
  int lanes_per_group;    /* a single scene's cooperative grid: lanes per workgroup (default 4) */
  int waves_per_group;    /* wavefronts per workgroup of the batched solve: 4 (default) or 8 */
  int color_lds_contacts; /* most contacts of a colouring unit whose colouring state is kept in LDS
                             (default and upper bound 4096; larger units keep it in global memory) */
  int color_body_bits;    /* most words of each per-body colour bitmap array of a unit of at most 256
                             contacts (default and upper bound 512; a unit needing more finds its
                             contacts' neighbours through its bodies' contact lists) */
  int solve_groups_per_cu;/* batched solve: 1 = full LDS layout (one workgroup per CU), 2 = half layout
                             (FP32 192 resident contacts; two per CU); 0 = 2 when the batch has more
                             scenes than the device has CUs, else 1 */
} rbh_tuning;

typedef struct rbh_device_info {
  char name[256];
  int compute_units;
  int lds_per_block;
  unsigned long long hbm_bytes;
  int precision_bytes;
  int max_colors;
} rbh_device_info;

/* Library identity; no reference counterpart. */
int rbh_version(void);

/* Create a context on `device` computing in FP32 (precision_bytes = 4) or
 * FP64 (8).  hbm_budget_bytes = 0: no budget.  Replaces the engine-level
 * state the reference keeps in Python objects (types.py:421-445). */
int rbh_ctx_create(int device, int precision_bytes, unsigned long long hbm_budget_bytes, rbh_ctx** out);
int rbh_ctx_destroy(rbh_ctx* ctx);
const char* rbh_last_error(rbh_ctx* ctx);
int rbh_device_query(rbh_ctx* ctx, rbh_device_info* out);
/* Device bytes held by the context (arenas + step workspace). */
int rbh_memory_usage(rbh_ctx* ctx, unsigned long long* bytes);

/* Signed distance of the res I x J x K node grid spanning [min_coord, max_coord]
 * to the closed mesh (V, T): replaces igl.signed_distance inside
 * GRID.create_signed_distance (grid3.py:320-339).  values: I*J*K, index
 * i + I*(j + J*k). */
int rbh_sdf_build(rbh_ctx* ctx, const double* V, int nv, const int* T, int nt, const double* min_coord,
                  const double* max_coord, int I, int J, int K, double* values);

/* rbh_sdf_build without device allocation.  No counterpart in the reference:
 * it is for callers that cannot call hipMalloc (serving loops and harnesses
 * that forbid it).  Same arguments and the same values, bit for bit; the
 * inputs and the grid are staged in a fixed 32 MiB device array per GPU
 * (held until the process ends) instead of buffers allocated per call.  The
 * staging needs 24 nv + 12 nt + 72 nt + 8 I J K bytes, each part rounded up
 * to 256; a larger call fails with RBH_E_CAPACITY (e.g. a 128^3 grid does not
 * fit, a 96^3 grid of 10^4 triangles does).  Every context on one GPU shares
 * that array: calls that overlap on one GPU (two contexts, two threads) race. */
int rbh_sdf_build_noalloc(rbh_ctx* ctx, const double* V, int nv, const int* T, int nt, const double* min_coord,
                          const double* max_coord, int I, int J, int K, double* values);

/* Upload one shape (types.py:82-108): body-frame mesh (at least 2 triangles)
 * and its SDF grid. */
int rbh_shape_upload(rbh_ctx* ctx, const double* V, int nv, const int* T, int nt, const double* sdf_values, int I,
                     int J, int K, const double* min_coord, const double* max_coord, int* shape_id);

/* Scenes: device-resident bodies of one engine.  rbh_scene_create takes the
 * static data StateStorage reads each step (state_storage.py:108-185). */
int rbh_scene_create(rbh_ctx* ctx, const rbh_scene_desc* desc, int* scene_id);
int rbh_scene_destroy(rbh_ctx* ctx, int scene_id);
int rbh_scene_set_params(rbh_ctx* ctx, int scene_id, const rbh_params* params);
/* The scene's hinge joints, replacing any set before (n = 0 removes them):
 * Engine.hinges in creation order with the sockets api.set_hinge computes
 * (api.py:86-156, types.py:329-353) -- parent and child as scene body
 * indices, arm_p / arm_c (3n) the joint origin and axis_p / axis_c (3n) the
 * joint axis in the parent's and the child's body frames.  Each step then
 * solves the hinge problem of constraints/hinges.py:16-218 together with the
 * contacts (gauss_seidel_solver.py:14-101).  A scene holds at most 1024
 * hinges joining at most 1024 bodies; a hinge needs two distinct bodies. */
int rbh_scene_set_hinges(rbh_ctx* ctx, int scene_id, int n, const int* parent, const int* child, const double* arm_p,
                         const double* arm_c, const double* axis_p, const double* axis_c);
/* Body state: r (3n), q (4n, s x y z), v (3n), w (3n); replaces
 * StateStorage.copy_from_engine / copy_to_engine (state_storage.py:43-68). */
int rbh_scene_set_state(rbh_ctx* ctx, int scene_id, const double* r, const double* q, const double* v,
                        const double* w);
int rbh_scene_get_state(rbh_ctx* ctx, int scene_id, double* r, double* q, double* v, double* w);
/* Number of committed steps of the scene; it keys the colouring's random
 * draws, so a scene restored from a snapshot replays the same colourings.
 * No reference counterpart (the reference draws from numpy's global RNG). */
int rbh_scene_set_step_index(rbh_ctx* ctx, int scene_id, int step_index);
/* Device bytes a scene of this size would take (admission control). */
int rbh_scene_footprint(rbh_ctx* ctx, const rbh_scene_desc* desc, unsigned long long* bytes);

/* Advance the listed scenes together by n_steps steps of SemiImplicitStepper.step
 * (steppers.py:47-106): collision detection (collision_detection.py:293-318),
 * contact problem (contacts.py:31-55), prox solve (gauss_seidel_solver.py:14-101)
 * and integration (state_storage.py:239-302).  dt > 0 is used for every scene,
 * dt <= 0 means each scene's time_step. */
int rbh_step(rbh_ctx* ctx, const int* scene_ids, int n_scenes, int n_steps, double dt, int flags);
int rbh_sync(rbh_ctx* ctx);

/* Log of the scene's last step; residuals (cap entries) receives the per-sweep
 * merit array of gauss_seidel_solver.py:31-33,76 (inf past the last sweep) in
 * entries [0, max_iterations), and with cap >= 2 max_iterations the
 * post-stabilisation solve's ('post_stabilization_residuals', steppers.py:33-34)
 * in [max_iterations, 2 max_iterations) (inf when the step ran none). */
int rbh_get_log(rbh_ctx* ctx, int scene_id, rbh_step_log* log, double* residuals, int cap);

/* Contacts of the scene's last step in the reference's order
 * (collision_detection.py:204-290); ia = ContactPoint.bodyA (SDF owner),
 * ib = ContactPoint.bodyB (triangle owner), tri = triangle of ib, p/n (3 each),
 * gap, impulse (4 each: normal, s, t, torsion; the solved x of contacts.py:57-64),
 * colour.  Any output pointer may be NULL.  *count receives the total. */
int rbh_get_contacts(rbh_ctx* ctx, int scene_id, int cap, int* count, int* ia, int* ib, int* tri, double* p,
                     double* n, double* gap, double* impulse, int* color);

/* GPU milliseconds per stage, averaged over the steps of the last rbh_step
 * call made with RBH_STEP_TIMING (its last 512 steps at most):
 * [0] body prep + refit (update_bvh), [1] broad phase + candidates (narrow_phase),
 * [2] Frank-Wolfe, whose kernel also moves the triangles into the SDF's body
 * frame and writes the contact points (contact_determination), [3] assembly,
 * [4] colouring + order, [5] solve (solver_time), [6] integration,
 * [7] total (stepper_time), [8] exact-position reduction and contact order
 * (contact_point_reduction), [9] post-stabilisation rows, solve and position
 * update (post_stabilization_solver_time; 0 when none ran).  With
 * RBH_STEP_GRAPH only [7] is measured (per replay).  The stage names in
 * parentheses are the profile keys of steppers.py:55-100,
 * collision_detection.py:10-83, 193-201, 204-290 and
 * gauss_seidel_solver.py:98-101.  Returns the number of entries written. */
int rbh_get_timings(rbh_ctx* ctx, double* ms, int cap);

/* Batch-level counters of the last step: [0] body pairs, [1] candidate slots,
 * [2] candidate triangles, [3] contacts. */
int rbh_get_counters(rbh_ctx* ctx, long long* out, int cap);

/* Set / read the launch shapes (rbh_tuning).  Takes effect at the next
 * rbh_step; a captured graph is dropped.  No reference counterpart. */
int rbh_set_tuning(rbh_ctx* ctx, const rbh_tuning* tuning);
int rbh_get_tuning(rbh_ctx* ctx, rbh_tuning* tuning);

/* Running totals of a scene since its creation or the last clear (clear != 0
 * zeroes them after reading): [0] committed steps, [1] contact-iterations
 * (contacts x sweeps, rejected sweeps included), [2] contacts, [3] candidate
 * triangles, [4] trilinear cell evaluations in Frank-Wolfe, [5] sweeps.
 * Kept on the device, so reading them is the only synchronisation.  No
 * reference counterpart: the reference's profiler (steppers.py:95-106) keeps
 * per-step values only.  Returns the number of entries written. */
int rbh_scene_totals(rbh_ctx* ctx, int scene_id, unsigned long long* out, int cap, int clear);

#ifdef __cplusplus
}
#endif
#endif
