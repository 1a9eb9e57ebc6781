/*
 * alens_capi.h -- C ABI of the MI355X collision-constraint engine (libalens.so).
 *
 * One context owns one rod system on one GPU: its device buffers, a HIP stream, the Brownian key
 * and the previous step's contact impulses.  Every entry point is plain C: host pointers are
 * borrowed for the duration of the call and copied; outputs are caller-allocated; errors come back
 * as alens_status with a message from alens_last_error().  The library never exits the process.
 * A context is used by one thread at a time; distinct contexts are independent.
 *
 * The reference reaches this path through SylinderSystem (SimToolbox, absent from the reference
 * tree): TubuleSystem calls rodSystem.prepareStep() (SRC/TubuleSystem.cpp:120), pushes protein
 * constraint blocks into the per-thread pool (SRC/TubuleSystem.cpp:694-745), then
 * rodSystem.runStep() (SRC/TubuleSystem.cpp:223) runs Brownian + non-constraint velocity, collision
 * collection, the BCQP solve and the write-back, and stepEuler() (SRC/TubuleSystem.cpp:55) moves
 * the rods.  Each function below names the call it replaces.
 */
#ifndef ALENS_CAPI_H
#define ALENS_CAPI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct alens_ctx alens_ctx;

typedef enum {
    ALENS_OK = 0,
    ALENS_E_ARG = 1,        /* invalid argument or configuration */
    ALENS_E_HIP = 2,        /* HIP runtime error (message names the call) */
    ALENS_E_OOM = 3,        /* device allocation failed */
    ALENS_E_NONFINITE = 4,  /* NaN/Inf in the rod state or in the solver iterate */
    ALENS_E_NOCONV = 5,     /* BBPGD reached max_ite above tol (results are still written) */
    ALENS_E_BOX = 6,        /* periodic box shorter than twice the largest interaction range */
    ALENS_E_STATE = 7       /* call order violated (e.g. step before set_rods) */
} alens_status;

/* Run configuration: the hot-path subset of SylinderConfig (docs/source/quickstart.md:316-358). */
typedef struct {
    double box_lo[3], box_hi[3]; /* simBoxLow / simBoxHigh */
    int32_t pbc[3];              /* simBoxPBC */
    int32_t monolayer;           /* monolayer: motion restricted to the x-y plane */
    double col_buf;              /* sylinderColBuf, absolute (um) */
    double viscosity;            /* viscosity, pN s / um^2 */
    double kT;                   /* KBT, pN um; <= 0 disables Brownian motion */
    double dt;                   /* dt, s */
    double tol;                  /* conResTol */
    double diam_col_ratio;       /* sylinderDiameterColRatio */
    double len_col_ratio;        /* sylinderLengthColRatio */
    double cell_size;            /* cell-list edge (um); <= 0 picks a quarter of the mean collision length */
    int32_t max_ite;             /* conMaxIte */
    int32_t solver;              /* conSolverChoice; 0 = BBPGD (the only solver provided) */
    int32_t precision;           /* 0 = FP64; 1 = FP32 iterations + FP64 residual and refinement */
    int32_t warm_start;          /* 0 = cold; 1 = previous-step gamma matched by rod pair;
                                    2 = learned network (alens_mlp_*) fed with the previous gamma */
    uint64_t seed;               /* rngSeed */
    int32_t device;              /* HIP device ordinal */
    int32_t reserved;
} alens_config;

/* One constraint, field for field the ConstraintBlock built at SRC/TubuleSystem.cpp:727-742. */
typedef struct {
    double delta0, gamma, kappa, gamma_lb;
    int32_t gid_i, gid_j, gidx_i, gidx_j;
    double norm_i[3], norm_j[3], pos_i[3], pos_j[3], lab_i[3], lab_j[3];
    double stress[9];
    int32_t one_side, bilateral;
} alens_con_block;

typedef struct {
    int32_t iters;          /* BBPGD iterations of the final (FP64) solve */
    int32_t matvecs;        /* FP64 applications of A */
    int32_t converged;
    int32_t fallback_used;  /* warm start rejected (raised the initial residual), cold restart */
    int32_t status;         /* alens_status of the solve */
    int32_t iters_fp32;     /* FP32 iterations before the FP64 refinement (precision = 1) */
    int64_t n_con;          /* constraints collected this step */
    int64_t n_entries;      /* cell-list entries */
    double res0, res;       /* residual at the start point used, final FP64 residual */
    double t_collect_ms;    /* device time: U_nc + cell list + pair sweep + rows */
    double t_solve_ms;      /* device time of the BBPGD kernel(s) */
    double t_step_ms;       /* device time of the whole step_constraints call */
} alens_solve_report;

/* Create / destroy a context (replaces SylinderSystem construction + initialize(),
 * SRC/TubuleSystem.cpp:79, minus file IO). */
alens_status alens_create(const alens_config *cfg, alens_ctx **out);
void alens_destroy(alens_ctx *ctx);
const char *alens_last_error(const alens_ctx *ctx);
const char *alens_version(void);

/* Load the rod state (Sylinder fields copied at SRC/TubuleBind.hpp:54-67: gid, radius, length,
 * pos, direction = q e_z).  center/dir are 3n (x,y,z per rod), dir unit; immovable may be NULL.
 * Rod order is kept; constraints are indexed by it. */
alens_status alens_set_rods(alens_ctx *ctx, int32_t n, const int32_t *gid, const double *center,
                            const double *dir, const double *length, const double *radius,
                            const uint8_t *immovable);
alens_status alens_get_rods(alens_ctx *ctx, double *center, double *dir);

/* Many independent rod systems in one context (the batched form): system s owns rods
 * [rod_offset[s], rod_offset[s+1]) of the concatenated arrays and its own box / periodicity /
 * monolayer flag; all other settings come from the context's alens_config.  One step collects and
 * solves every system together (one persistent solver launch with per-system step lengths and
 * convergence); constraints never couple systems.  alens_set_rods is the one-system case. */
typedef struct {
    double box_lo[3], box_hi[3];
    int32_t pbc[3];
    int32_t monolayer;
} alens_box;
alens_status alens_set_systems(alens_ctx *ctx, int32_t nsys, const int32_t *rod_offset, const alens_box *boxes,
                               const int32_t *gid, const double *center, const double *dir, const double *length,
                               const double *radius, const uint8_t *immovable);
alens_status alens_get_num_systems(const alens_ctx *ctx, int32_t *nsys);
/* per-system solve reports of the last step (nsys entries; n_con = the system's rows; iters_fp32 filled
 * for precision = 1) */
alens_status alens_get_system_reports(alens_ctx *ctx, alens_solve_report *reps);

/* Reorder the rods on the device by the Morton code of their centre's cell so that neighbours sit
 * close in memory (the reference keeps its particle arrays in FDPS tree order for the same reason;
 * repartitioning happens in prepareStep, SRC/TubuleSystem.cpp:120-125).  perm_out (n, may be NULL)
 * receives, for each new index i, the rod's previous index.  Rod and constraint indices refer to the
 * new order afterwards; the previous-step impulse cache is cleared. */
alens_status alens_sort_rods(alens_ctx *ctx, int32_t *perm_out);  /* within each system */

/* Step counter keying the Brownian stream (SylinderSystem's stepCount, SRC/TubuleSystem.cpp:114). */
alens_status alens_set_step(alens_ctx *ctx, int64_t step);
int64_t alens_get_step(const alens_ctx *ctx);

/* Caller-supplied constraint blocks for the next step: the reference's constraint pool, filled
 * through getConstraintPoolNonConst() by setProteinConstraints (SRC/TubuleSystem.cpp:694-745) and
 * consumed by runStep().  Blocks accumulate until the next alens_step_constraints, which appends
 * them after the collision rows in the order given and then clears them.  Rods are named by gid_i /
 * gid_j (gidx_* are ignored on input); norm_j must equal -norm_i; pos_i / pos_j are the arms from
 * the rod centres and lab_i / lab_j / stress are passed through.  Bilateral rows are unbounded and
 * add 1 / kappa (kappa > 0) to the operator's diagonal (a spring: gamma = -kappa * separation);
 * unilateral rows are bounded below by gamma_lb.  The solve starts from each block's gamma.
 * One-system contexts only. */
alens_status alens_add_constraints(alens_ctx *ctx, int64_t m, const alens_con_block *blocks);
alens_status alens_clear_constraints(alens_ctx *ctx);

/* Graph mode (no counterpart in the reference): steps run as one captured HIP graph.  The first step
 * sizes every buffer exactly (reading the counts of cell entries, pairs and adjacency slots back
 * between its phases); the next step sets capacities with head room, records the whole step (U_nc,
 * cell list, pair sweep, rows, adjacency, chunk table, start point, the persistent solver, write-back)
 * into a graph once with every count kept on the device, and each later step replays it with one
 * launch and one read-back.  A replay whose counts outgrow the capacities is redone exactly and the
 * graph is recorded again.  Results are bit-identical to plain steps.  Steps with caller constraint
 * blocks run plain.  The graph is dropped when rods, systems, the network or the schedule change.
 * A replay's report times collection and solve from device clock readings taken inside the graph. */
alens_status alens_set_graph(alens_ctx *ctx, int32_t enable);
typedef struct {
    int32_t captured;         /* a recorded graph is ready for the next step */
    int32_t reserved;
    int64_t replays;          /* steps run as one graph replay */
    int64_t captures;         /* graphs recorded */
    int64_t capture_failures; /* recordings that failed (those steps ran plain) */
    int64_t overflows;        /* replays redone plainly because a count outgrew its capacity */
    char last_failure[192];   /* message of the last failed recording, or empty */
} alens_graph_info;
alens_status alens_get_graph_info(const alens_ctx *ctx, alens_graph_info *info);

/* One constraint step (rodSystem.runStep(), SRC/TubuleSystem.cpp:223): U_nc = M F_ext + U_B,
 * collision collection, BBPGD to tol, write-back F_con = D gamma, U = U_nc + M F_con.
 * f_ext is 6n (force, torque per rod) or NULL. */
alens_status alens_step_constraints(alens_ctx *ctx, const double *f_ext, alens_solve_report *rep);

/* Move rods with the last step's velocity (stepEuler(), SRC/TubuleSystem.cpp:55) and advance the
 * step counter.  The update is enqueued on the context's stream and the call returns without waiting;
 * later calls on the context are ordered after it (those that copy to the host wait for it). */
alens_status alens_step_euler(alens_ctx *ctx);

/* Constraints of the last step: the collision rows in canonical order (ascending rod index I,
 * then J; I < J), then the caller's rows in the order they were added.  get_constraints fills
 * blocks (gamma and stress set); get_constraint_arrays fills any non-NULL SoA array (I, J: m; img:
 * 3m periodic image of J, zero on caller rows; delta0, b: m; nrm, pos_i, pos_j, r_i, r_j: 3m). */
alens_status alens_get_num_constraints(alens_ctx *ctx, int64_t *m);
alens_status alens_get_constraints(alens_ctx *ctx, int64_t cap, alens_con_block *out, int64_t *m);
alens_status alens_get_constraint_arrays(alens_ctx *ctx, int32_t *I, int32_t *J, int32_t *img,
                                         double *delta0, double *nrm, double *pos_i, double *pos_j,
                                         double *r_i, double *r_j, double *b);
alens_status alens_get_gamma(alens_ctx *ctx, double *gamma);
/* Constraint stress of the last step (the sums behind calcConStress, SRC/TubuleSystem.cpp:224):
 * row-major 3 x 3 sums, in row order, of gamma_k * S_k over the collision rows (S_k = (lab_i -
 * lab_j) (x) norm_i, the virial of the unit force pair) and over the caller's rows (S_k = the block's
 * stress).  Not divided by a volume.  Either pointer may be NULL. */
alens_status alens_get_con_stress(alens_ctx *ctx, double collision[9], double external[9]);
/* u, u_nc, f_con: 6n (vx,vy,vz,wx,wy,wz per rod); any may be NULL. mob: 6n {t, 1/zperp,
 * 1/zpar - 1/zperp, 1/zrot} or NULL. */
alens_status alens_get_velocity(alens_ctx *ctx, double *u, double *u_nc, double *f_con, double *mob);

/* Solver autotuner (no counterpart in the reference, whose solver runs on Trilinos over MPI):
 * times cold solves of the last step's rows under candidate schedules - adjacency entries in flight
 * per rod step (largest, 2), grid-barrier variant (grouped by XCC, grouped, flat, grouped with longer polls),
 * XCD-contiguous work mapping, persistent grid size (all co-resident, 3/4, 1/2), reduction-group
 * width of the constraint phase (automatic = 256, 128 lanes), cache hint of the constraint phase's
 * row streams (automatic, cached, non-temporal) - one axis at a time from the defaults (a value
 * replaces the current one only when at least 1% faster), keeps the fastest in the context for later
 * steps, and reports the choice.  A tuned width of 128 lanes holds while a step's row count stays
 * within 10% of the tuned step's; other steps use 256 (the solver checks the device's row count, so
 * replayed graph steps follow it).
 * Results are bit-identical under every schedule; only the time changes.  Needs a step with
 * constraints; leaves that step's outputs untouched.  Environment variables ALENS_SOLVER_ENTRIES,
 * ALENS_BARRIER_MODE, ALENS_XCD_MAP, ALENS_SOLVER_MAX_BLOCKS, ALENS_GROUP_WIDTH and ALENS_STREAM_HINT
 * override the choice outside tuning. */
typedef struct {
    int32_t entries;     /* 0 = largest instantiated (3 in FP64, 4 in FP32), or 2 */
    int32_t bar_mode;    /* 0 flat, 1 grouped, 2 grouped with longer polls, 3 grouped by XCC with one
                            L2 write-back per XCC */
    int32_t xcd_map;     /* 1 = blocks of one XCD take consecutive work */
    int32_t max_blocks;  /* 0 = all co-resident blocks */
    int32_t candidates;  /* schedules timed */
    float default_ms;    /* solve time of the default schedule */
    float best_ms;       /* solve time of the chosen schedule */
    int32_t group_width; /* phase-B reduction-group width: 0 = automatic (256 lanes), 128 (held for row
                            counts within 10% of the tuned step's) */
    int32_t stream_hint; /* phase-B row streams: 0 = automatic (non-temporal loads when an iteration's
                            bytes exceed the 256 MiB last-level cache), 1 = cached, 2 = non-temporal */
} alens_tune_report;
alens_status alens_autotune(alens_ctx *ctx, alens_tune_report *rep);

/* Stand-alone BCQP solve (BCQPSolver::solveBBPGD on A = dt D^T M D, b; reached in the reference
 * through ConstraintSolver, CMakeLists.txt:109-112).  Rows: I, J (m, I < J not required), nrm, r_i,
 * r_j (3m), b (m); mob (6n) as in alens_get_velocity; x0 (m) or NULL; gamma_out (m). */
alens_status alens_bcqp_solve(alens_ctx *ctx, int32_t n, int64_t m, const int32_t *I,
                              const int32_t *J, const double *nrm, const double *r_i,
                              const double *r_j, const double *b, const double *mob,
                              const double *x0, double *gamma_out, alens_solve_report *rep);

/* Allocation-free variant of alens_bcqp_solve (no counterpart in the reference): same arguments and
 * results, bit-identical gamma and report, but every device buffer it uses (staged rows, sliced
 * adjacency, iterate, velocities, partial sums, barrier and report words) lives in one static device
 * array of the library, so the call makes no hipMalloc / hipFree.  It is for callers that cannot
 * allocate device memory: serving loops with a fixed footprint, harnesses that forbid allocation,
 * code that must not grow the device heap.  Limits: m <= 131072 rows, n <= 65536 rods and at most
 * 524288 padded adjacency slots (64 rods per slice times the slice's largest degree); larger problems
 * are refused with ALENS_E_ARG.  The array (about 50 MB) is held on every GPU the library runs on until
 * the process ends.  All calls on one GPU share it: calls that overlap on one GPU (from several
 * threads or contexts) race and must be serialised by the caller.  The call waits for its result. */
alens_status alens_bcqp_solve_static(alens_ctx *ctx, int32_t n, int64_t m, const int32_t *I,
                                     const int32_t *J, const double *nrm, const double *r_i,
                                     const double *r_j, const double *b, const double *mob,
                                     const double *x0, double *gamma_out, alens_solve_report *rep);

/* ---- learned warm start (no counterpart in the reference: the reference starts its BBPGD from the
 * constraint blocks' gamma guess, SRC/TubuleSystem.cpp:722,728) ----
 * Network 12 -> 64 -> 64 -> 64 -> 1, SiLU hidden layers, softplus output, BF16 MFMA with FP32
 * accumulation.  params: ALENS_MLP_PARAMS floats (W1 64x16 row-major with input columns 12..15 zero,
 * b1 64, W2 64x64, b2 64, W3 64x64, b3 64, w4 64, b4); norm: ALENS_MLP_NORM floats (feature mean[12],
 * 1/std[12], gamma scale).  Features X are 16 floats per constraint (12 used): delta0, normal(3),
 * s_I/(L_I/2), s_J/(L_J/2), dt * relative contact velocity of U_nc (3), L_I, L_J, previous gamma. */
#define ALENS_MLP_PARAMS 9473
#define ALENS_MLP_NORM 25
alens_status alens_mlp_set(alens_ctx *ctx, const float *params, const float *norm);
alens_status alens_mlp_get(alens_ctx *ctx, float *params, float *norm);
/* features of the last step's constraints (16 m floats) */
alens_status alens_get_features(alens_ctx *ctx, float *X);
/* gamma0 = gamma_scale * network(X) for m feature rows */
alens_status alens_mlp_predict(alens_ctx *ctx, int64_t m, const float *X, float *gamma0);
/* full-batch gradient of mean((network(X)/gamma_scale - target)^2); m a multiple of 16 */
alens_status alens_mlp_gradient(alens_ctx *ctx, int64_t m, const float *X, const float *target, float *grad,
                                float *loss);
/* Adam training on the GPU on (X, gamma) pairs; sets the normalisation from the data and initialises
 * the weights when no network is loaded.  batch is a multiple of 16; loss_per_epoch may be NULL. */
alens_status alens_mlp_train(alens_ctx *ctx, int64_t m, const float *X, const float *gamma, int32_t epochs,
                             int32_t batch, float lr, uint64_t seed, float *loss_per_epoch);
/* Training on the quantity that sets the solver's work: the BCQP energy of the network's start point
 * on the last step's collision rows (the step must have no caller blocks), E = 1/2 g0' A g0 + b' g0 with
 * g0 = network(features), A = dt D^T M D and b of the step.  dE/dg0 = A g0 + b is one start-point pass
 * of the solver; it is backpropagated through the network's BF16-MFMA training kernels.
 * alens_mlp_energy: E and (grad non-NULL, ALENS_MLP_PARAMS floats) dE/dparameters.
 * alens_mlp_train_energy: iters Adam steps of rate lr on E (moments kept in the context across calls,
 * reset by alens_mlp_set); energy (iters + 1 values or NULL): E before each step and after the last.
 * Both overwrite the solver's iterate buffer; the step's outputs (gamma, velocities, rows) stay. */
alens_status alens_mlp_energy(alens_ctx *ctx, double *energy, float *grad);
alens_status alens_mlp_train_energy(alens_ctx *ctx, int32_t iters, float lr, double *energy);

/* Device copy probe (no counterpart in the reference; a measurement aid for benchmark scripts): on
 * HIP device `device`, one kernel (k_copy_probe) copies nbytes between two fresh buffers, reps times;
 * *gbs = (read + written bytes) / time of those launches, in GB/s.  The buffers are freed on return.
 * A single launch of known traffic also calibrates profiler byte counters. */
alens_status alens_copy_probe(int32_t device, int64_t nbytes, int32_t reps, double *gbs);

#ifdef __cplusplus
}
#endif
#endif
