/*
 * brownian.h -- C-ABI of libbrownian: the MPCD (SRD) step for active Brownian
 * colloids on one MI355X (gfx950), hand-written HIP kernels behind the entry
 * points of the reference C++ program shashankgandham/brownian (cpp/).
 *
 * The reference declares its entry points as argument-free functions over
 * global state (cpp/parameters.hpp:12-14, globals defined at
 * cpp/active_brownian.cpp:3-7).  Here each keeps its name, takes the context
 * that owns that state and returns a status code (BW_OK or a negative BW_ERR_*).
 * Calls are asynchronous on the context's HIP stream; calls that return data to
 * the host synchronise.  One context per host thread.
 *
 * Units and indices: box edge L cells of size a = 1; particle, colloid and cell
 * indices are 0-based (the reference is 1-based); positions, velocities and
 * colloid quantities cross the boundary as FP64, n x 3, row-major.
 */
#ifndef BROWNIAN_H
#define BROWNIAN_H

#ifdef __cplusplus
extern "C" {
#endif

#define BW_API __attribute__((visibility("default")))

/* status codes */
#define BW_OK 0
#define BW_ERR_ARG (-1)      /* bad argument or parameter                          */
#define BW_ERR_HIP (-2)      /* HIP runtime error                                  */
#define BW_ERR_CAPACITY (-3) /* a fixed-capacity device buffer overflowed           */
#define BW_ERR_DRAWS (-4)    /* injected draws missing or not matching the events   */
#define BW_ERR_STATE (-5)    /* call out of order (e.g. maps not built)            */
#define BW_ERR_PHYSICS (-6)  /* particle outran its tile, zero-variance cell, NaN   */

/* device status bits (bw_status) */
#define BW_FLAG_EVENT_OVERFLOW 0x1u    /* > event capacity for one colloid in a substep */
#define BW_FLAG_CHAIN_OVERFLOW 0x2u    /* > serial-replay capacity (shared particles)    */
#define BW_FLAG_TILE_OVERFLOW 0x4u     /* a tile received more particles than its slots  */
#define BW_FLAG_LIST_OVERFLOW 0x8u     /* colloid list of a tile / MD list overflow       */
#define BW_FLAG_DRIFT 0x10u            /* particle moved more than a tile in one step     */
#define BW_FLAG_DRAW_MISSING 0x20u     /* injected draw absent for a needed cell/event    */
#define BW_FLAG_DRAW_MISMATCH 0x40u    /* injected reflection list differs from events    */
#define BW_FLAG_ZERO_VARIANCE 0x80u    /* thermostat on a cell with zero relative KE      */
#define BW_FLAG_NEIGH_OVERFLOW 0x100u  /* neigh_fl capacity exceeded                      */
#define BW_FLAG_MARGIN 0x200u          /* a colloid moved beyond the blocked-streaming margin
                                          within one step (use BW_STREAM_PER_SUBSTEP)    */

enum bw_precision { BW_FP32 = 0, BW_FP64 = 1 };
enum bw_vel_law { BW_VEL_REFERENCE = 0, BW_VEL_MAXWELL = 1 };
enum bw_orient_law { BW_ORIENT_REFERENCE = 0, BW_ORIENT_SPHERE = 1 };
enum bw_draw_source { BW_DRAWS_DEVICE = 0, BW_DRAWS_INJECTED = 1 };
/* Fused-step streaming.  BLOCKED: particles that cannot reach a colloid within
 * the step are streamed through all substeps by the rotation pass (same
 * arithmetic, same results); PER_SUBSTEP: every particle every substep. */
enum bw_streaming { BW_STREAM_BLOCKED = 0, BW_STREAM_PER_SUBSTEP = 1 };
/* initialize_fluid.  HOST: the reference's generator on the host, bit-exact
 * with cpp/initialize.cpp:62-83 (and the draws of everything after it);
 * DEVICE: the same laws with counter-based draws on the GPU (keyed by seed and
 * particle id), the state binned into tiles on the GPU -- for large systems
 * and ensembles, where the host start-up dominates. */
enum bw_fluid_init { BW_INIT_HOST = 0, BW_INIT_DEVICE = 1 };

/* Run parameters.  Mirrors the reference's global defaults
 * (cpp/active_brownian.cpp:3-7) with every value exposed; see
 * bw_default_params for the reference defaults. */
typedef struct bw_params {
    int L;               /* cells per box edge (reference len = 30)                  */
    int density;         /* fluid particles per cell (reference argv[1])             */
    int n_colloid;       /* reference argv[2]                                        */
    int n_sub;           /* streaming substeps per MPCD step (reference n = 10)      */
    double ndt;          /* MPCD step (0.1)                                          */
    double kbt;          /* kT of fluid and colloid initialisation (1)               */
    double mass_fl;      /* 1                                                        */
    double mass_colloid; /* 654.1                                                    */
    double sig_colloid;  /* LJ sigma (5)                                             */
    double eps;          /* LJ epsilon (1)                                           */
    double sigma;        /* colloid diameter (reference 0.8*sig_colloid)             */
    double theta_deg;    /* SRD rotation angle (reference hard-codes 130)            */
    double v0;           /* active speed (reference 0); per colloid via bw_set_v0    */
    int seed;            /* generator seed (77777)                                   */
    int precision;       /* BW_FP32: int cell + FP32 offset/velocity; BW_FP64        */
    int vel_law;         /* initial fluid/colloid velocity law                       */
    int orient_law;      /* initial orientation law                                  */
    int draws;           /* BW_DRAWS_DEVICE (counter-based) or BW_DRAWS_INJECTED     */
    int tile[3];         /* tile shape in cells; 0 = choose                          */
    int streaming;       /* bw_streaming of the fused step                           */
    int fluid_init;      /* bw_fluid_init of initialize_fluid                        */
    int reserved[6];
} bw_params;

typedef struct bw_ctx bw_ctx;

/* ---------------------------------------------------------------- context */
/* bw_ctx_create*: on an error after the context was allocated, *out still
 * holds it so that bw_last_error can say what failed; the caller then
 * releases it with bw_ctx_destroy (*out is NULL when nothing was allocated). */
BW_API void bw_default_params(bw_params *p);
BW_API int bw_ctx_create(const bw_params *p, bw_ctx **out);
/* n_replica independent systems stepped together (one launch per kernel) */
BW_API int bw_ctx_create_ensemble(const bw_params *p, int n_replica, bw_ctx **out);
BW_API int bw_ctx_destroy(bw_ctx *ctx);
BW_API const char *bw_last_error(bw_ctx *ctx);
BW_API int bw_n_replica(bw_ctx *ctx);
/* per replica: n_fluid, n_colloid, L, tile x/y/z, slots per tile */
BW_API int bw_replica_info(bw_ctx *ctx, int replica, int *out7);
BW_API int bw_sync(bw_ctx *ctx);
/* device status bits accumulated since the last call; clears them */
BW_API int bw_status(bw_ctx *ctx, unsigned int *flags);

/* ---------------------------------------- reference entry points (cpp/) */
/* cpp/initialize.cpp:3-29 -- (re)initialise state and generators           */
BW_API int initialize(bw_ctx *ctx);
/* cpp/initialize.cpp:31-60 -- colloid placement, velocities, spin (host)  */
BW_API int initialize_colloid(bw_ctx *ctx);
/* cpp/initialize.cpp:62-83 -- fluid placement and velocities (host)        */
BW_API int initialize_fluid(bw_ctx *ctx);
/* cpp/neighbour_list.cpp:12-25 -- the 7^3-1 stencil is arithmetic here     */
BW_API int create_box(bw_ctx *ctx);
/* cpp/neighbour_list.cpp:41-58 -- cell lists and per-colloid neigh_fl      */
BW_API int neighbour_list_mpcd(bw_ctx *ctx);
/* cpp/neighbour_list.cpp:27-39 -- colloid pair list (cell list)            */
BW_API int neighbour_list_md(bw_ctx *ctx);
/* cpp/update.cpp:14-32 -- WCA force                                        */
BW_API int compute_force_md(bw_ctx *ctx);
/* cpp/run.cpp:3-8 -- initial orientations (host)                          */
BW_API int tumble(bw_ctx *ctx);
/* cpp/rotation_mpcd.cpp:3-50 -- shifted-grid SRD rotation and thermostat   */
BW_API int rotation_mpcd(bw_ctx *ctx);
/* cpp/run.cpp:10-26 -- active kick with rear-shell momentum removal        */
BW_API int run(bw_ctx *ctx);
/* cpp/run.cpp:28-50 -- rear-shell and near-surface slip velocities behind
 * each colloid, a diagnostic (the reference discards its results; the fused
 * step, whose trajectory it does not change, leaves it out).  Needs the
 * neigh_fl lists of neighbour_list_mpcd.  Results via bw_get_updown.      */
BW_API int updown_velocity(bw_ctx *ctx);
/* after updown_velocity, per colloid of replica r: rear-shell and
 * near-surface counts (cnt, up_cnt) and the reference's vel and up_vel after
 * that colloid (n_colloid x 3 each; up_vel carries over between colloids as
 * in the reference) */
BW_API int bw_get_updown(bw_ctx *ctx, int replica, int *cnt, int *up_cnt, double *vel, double *up_vel);
/* cpp/update.cpp:48-54 -- colloid Verlet position                          */
BW_API int update_pos_md(bw_ctx *ctx);
/* cpp/update.cpp:56-60 -- fluid streaming                                  */
BW_API int update_pos_mpcd(bw_ctx *ctx);
/* cpp/fluid_colloid_collision.cpp:26-49 -- stochastic reflection            */
BW_API int fluid_colloid_collision(bw_ctx *ctx);
/* cpp/update.cpp:34-46 -- orientation update                               */
BW_API int update_activity_direction(bw_ctx *ctx);
/* cpp/update.cpp:62-67 -- colloid Verlet velocity                          */
BW_API int update_velocity_colloid(bw_ctx *ctx);

/* ------------------------------------------------------------- driver */
/* cpp/active_brownian.cpp:28 (old_force <- f)                             */
BW_API int bw_copy_old_force(bw_ctx *ctx);
/* cpp/active_brownian.cpp:14-21: the start-up sequence                     */
BW_API int bw_startup(bw_ctx *ctx);
/* nsteps MPCD steps, each equal to cpp/active_brownian.cpp:25-38, fused and
 * replayed from a captured HIP graph */
BW_API int bw_step(bw_ctx *ctx, int nsteps);
/* per-kernel timing of nsteps steps (HIP events on the context stream):
 * out[k] = total ms of kernel family k (see BW_KERNEL_*), launches[k] counts */
#define BW_KERNEL_STREAM 0
#define BW_KERNEL_COLLIDE 1
#define BW_KERNEL_DEPART 2
#define BW_KERNEL_SRD 3
#define BW_KERNEL_COLLOID 4
#define BW_KERNEL_OTHER 5
#define BW_KERNEL_COUNT 6
BW_API int bw_profile_steps(bw_ctx *ctx, int nsteps, double *ms, int *launches);
/* time nsteps graph-replayed steps with HIP events on the context stream */
BW_API int bw_time_steps(bw_ctx *ctx, int nsteps, double *ms);

/* ---------------------------------------------------------------- state */
/* per-colloid active speed of one replica (NULL: every colloid uses p.v0) */
BW_API int bw_set_v0(bw_ctx *ctx, int replica, const double *v0);
/* fluid: pos, vel (n_fluid x 3, particle order); colloid block: pos, vel,
 * omega, f, old_f, ra (each n_colloid x 3) -- any pointer may be NULL */
BW_API int bw_upload_fluid(bw_ctx *ctx, int replica, const double *pos, const double *vel);
BW_API int bw_download_fluid(bw_ctx *ctx, int replica, double *pos, double *vel, int *cell);
BW_API int bw_upload_colloid(bw_ctx *ctx, int replica, const double *pos, const double *vel,
                             const double *omega, const double *f, const double *old_f, const double *ra);
BW_API int bw_download_colloid(bw_ctx *ctx, int replica, double *pos, double *vel, double *omega,
                               double *f, double *old_f, double *ra);
/* unwrapped colloid displacement since start-up (n_colloid x 3) */
BW_API int bw_download_displacement(bw_ctx *ctx, int replica, double *disp);
/* generator state of the host-side reference generator (36 ints) */
BW_API int bw_rng_state(bw_ctx *ctx, int replica, int *out36);

/* ---------------------------------------------------------------- draws */
/* Injected draws (BW_DRAWS_INJECTED), keyed by what consumes them:
 *   shift: the 3 raw uniforms of cpp/rotation_mpcd.cpp:9 for the next step;
 *   axes:  L^3 x 2 raw uniforms (rho, phi draws of cpp/rotation_mpcd.cpp:19)
 *          per shifted cell, NaN where no draw was made;
 *   reflections for substep l (0-based) of the next step: n tuples
 *          (colloid, particle) with the raw uniforms (U_e, t.x, t.y, t.z, a, b)
 *          of cpp/fluid_colloid_collision.cpp:8-19 (a, b = accepted polar pair),
 *          in the order the reference consumes them.                          */
BW_API int bw_inject_rotation(bw_ctx *ctx, int replica, const double *shift3, const double *axes);
BW_API int bw_inject_reflections(bw_ctx *ctx, int replica, int substep, int n, const int *colloid,
                                 const int *particle, const double *vals6);
BW_API int bw_clear_injected(bw_ctx *ctx);

/* ----------------------------------------------------------------- maps */
/* after neighbour_list_mpcd: per-cell counts (L^3) and cell lists
 * (n_fluid ids, cell-major, ascending id within a cell) -- box_part/fluid_no */
BW_API int bw_get_cell_lists(bw_ctx *ctx, int replica, int *counts, int *list);
/* after neighbour_list_mpcd: neigh_fl counts (n_colloid) and concatenated
 * lists; cap = size of list; returns total length or BW_ERR_CAPACITY */
BW_API int bw_get_neigh_fl(bw_ctx *ctx, int replica, int *counts, int *list, int cap);
/* after neighbour_list_md: partner counts and lists (j > i, ascending) */
BW_API int bw_get_md_list(bw_ctx *ctx, int replica, int *counts, int *list, int cap);
/* after rotation_mpcd: the shifted cell of every particle (n_fluid) */
BW_API int bw_get_shifted_cells(bw_ctx *ctx, int replica, int *cells);
/* after run: rear-shell counts cnt (n_colloid) of cpp/run.cpp:15-20 */
BW_API int bw_get_run_counts(bw_ctx *ctx, int replica, int *cnt);

/* ---------------------------------------------------------------- audit */
/* FP64 audit of cpp/active_brownian.cpp:39-55: ke_colloid, ang_ke_colloid,
 * potential, ke_fluid, total momentum (3), fluid momentum (3),
 * sum |m v| of all particles (momentum scale) -> 11 doubles */
BW_API int bw_audit(bw_ctx *ctx, int replica, double *out11);

/* ------------------------------------------------------------- host only */
/* Reference start-up state computed on the host, without a GPU: runs
 * initialize_colloid, initialize_fluid and tumble of the reference with its
 * generator.  Outputs n x 3 arrays (may be NULL); used by CPU tests. */
BW_API int bw_host_initialize(const bw_params *p, double *pos_fl, double *vel_fl, double *pos_c,
                              double *vel_c, double *omega, double *ra, int *rng_out36);

/* Calibration: copy n bytes device to device with a plain vector kernel. */
BW_API int bw_debug_copy(bw_ctx *ctx, long long nbytes, int reps, double *ms);
/* Test hook: the device's periodic-geometry helpers on n given values --
 * img(a, L) and mod(a, L) of cpp/parameters.hpp:74-76, ia mod Li, and ia / b
 * for 0 <= ia < 1024, 1 <= b <= 63 (-1 elsewhere). */
BW_API int bw_debug_geometry(bw_ctx *ctx, int n, const double *a, double L, const int *ia, int Li, int b,
                             double *img_out, double *mod_out, int *wrap_out, int *div_out);

/* Test hook: the device draw source's Philox4x32-10 (Salmon et al., SC'11) on
 * n (counter, key) pairs: ctr n x 4 words, key n x 2 words -> out n x 4. */
BW_API int bw_debug_philox(bw_ctx *ctx, int n, const unsigned *ctr, const unsigned *key, unsigned *out);
/* Test hook: n device-drawn reflection tuples of replica 0 -- step `step`,
 * substep `substep`, its first colloid, particle ids 0..n-1 -- and the
 * stochastic reflection (cpp/fluid_colloid_collision.cpp:7-24) each gives at
 * particle position rf3 with surface vector rs3: vals6 n x 6 raw uniforms
 * (U_e, t.x, t.y, t.z, a, b), vel3 n x 3 velocities. */
BW_API int bw_debug_reflections(bw_ctx *ctx, int n, long long step, int substep, const double *rf3,
                                const double *rs3, double *vals6, double *vel3);

/* ----------------------------------------------------------- observables */
/* No counterpart in the reference (it computes energies and prints nothing,
 * cpp/active_brownian.cpp:39-55).  bw_observables_reset sets the time origin
 * of every replica: colloid displacements restart from zero and the current
 * orientations become ra(0).  bw_observables then returns for replica r:
 *   out[0] MPCD steps since the reset
 *   out[1] colloid mean squared displacement <|x(t)-x(0)|^2> (unwrapped)
 *   out[2] orientation autocorrelation <ra(t).ra(0)>
 *   out[3] fluid temperature m (sum|v|^2 - |sum v|^2/N) / (3 (N-1))
 *   out[4] colloid translational temperature m_c <|v|^2> / 3
 *   out[5] colloid rotational temperature I <|w|^2> / 3                    */
BW_API int bw_observables_reset(bw_ctx *ctx);
BW_API int bw_observables(bw_ctx *ctx, int replica, double *out6);

/* ------------------------------------------------------------ surrogate */
/* No counterpart in the reference.  A learned closure for the colloid-fluid
 * collisions (cpp/fluid_colloid_collision.cpp:26-49): a network predicts each
 * colloid's impulse over one MPCD step (mean and log-variance of momentum and
 * angular momentum transfer, 6 each) from 144 features (27 coarse-cell fluid
 * moments around the colloid and its velocity, angular velocity, orientation).
 *
 * Network: 144 -> 256 -> 256 -> 256 -> 12, SiLU; BF16 matrix-core products with
 * FP32 accumulation; FP32 master weights; Gaussian NLL; Adam. */
typedef struct bw_mlp bw_mlp;
BW_API int bw_mlp_create(int seed, bw_mlp **out);
BW_API int bw_mlp_destroy(bw_mlp *m);
BW_API const char *bw_mlp_last_error(bw_mlp *m);
/* parameters, logical layout: W1 (144x256, row-major), b1, W2 (256x256), b2,
 * W3, b3, W4 (256x12), b4 */
BW_API long long bw_mlp_n_params(void);
BW_API int bw_mlp_get_params(bw_mlp *m, float *out);
BW_API int bw_mlp_set_params(bw_mlp *m, const float *in);
/* standardisation of features (144) and targets (6) */
BW_API int bw_mlp_set_normalisation(bw_mlp *m, const float *x_mean, const float *x_std, const float *t_mean,
                                    const float *t_std);
/* standardised outputs (n x 12) for raw features x (n x 144) */
BW_API int bw_mlp_forward(bw_mlp *m, const float *x, int n, float *y);
/* mean loss over n samples (targets t: n x 6 raw) and its parameter gradient */
BW_API int bw_mlp_loss_grad(bw_mlp *m, const float *x, const float *t, int n, double *loss, float *grad);
/* Adam over shuffled mini-batches; loss_out[epoch] = mean loss of the epoch */
BW_API int bw_mlp_train(bw_mlp *m, const float *x, const float *t, int n, int epochs, int batch, float lr,
                        int seed, double *loss_out);
/* No counterpart in the reference: bw_mlp_forward for callers that cannot
 * allocate (tight serving loops, harnesses that forbid hipMalloc).  x (n x 144
 * raw features) and y (n x 12, the bw_mlp_forward outputs) are DEVICE
 * pointers on the network's GPU; the activations live in a fixed device
 * array of the library (at most 1024 rows; larger n is BW_ERR_ARG, about
 * 7.9 MiB held by each GPU until the process ends), so the call allocates
 * nothing.  It synchronises the network's stream before returning.  Calls
 * that overlap on one GPU share that array and race. */
BW_API int bw_mlp_forward_noalloc(bw_mlp *m, const float *x, int n, float *y);
/* the rotation pass also writes coarse-cell moments (needs even edges) */
BW_API int bw_set_moments(bw_ctx *ctx, int on);
/* drive colloids by the network's sampled impulses wherever every feature lies
 * in [lo, hi] (144 each); elsewhere, and in the first step, the explicit
 * collisions are kept.  Sampled impulses are clamped to [t_lo, t_hi] (6 each,
 * the training range).  NULL detaches.  The network must stay alive and must
 * not be trained while attached. */
BW_API int bw_surrogate_attach(bw_ctx *ctx, bw_mlp *m, const float *lo, const float *hi, const float *t_lo,
                               const float *t_hi);
/* after a fused step: features at its start (n_colloid x 144), collision
 * impulse over it (n_colloid x 6) and mode (0 explicit, 1 surrogate) */
BW_API int bw_get_surrogate_data(bw_ctx *ctx, int replica, float *feat, double *imp, int *mode);
/* after a fused step, per colloid of the replica that the network drove: the
 * fluid particles in its shell, which paid its impulse (count, n_colloid), and
 * the sums the payment was formed from (n_colloid x 15, fluid-mass units): the
 * mirrored fluid's momentum change (3) and angular momentum change about the
 * centre (3), the shell's first (3) and second (xx yy zz xy xz yz) moments of
 * position about the centre.  Zero for explicitly colliding colloids. */
BW_API int bw_get_surrogate_shells(bw_ctx *ctx, int replica, int *count, double *sums);

/* Diagnostics (no counterpart in the reference): particles of replica r that
 * the last fused step's per-substep kernels touched ("near" a colloid under
 * blocked streaming; all particles under per-substep streaming). */
BW_API int bw_near_particles(bw_ctx *ctx, int replica, long long *n_near);

#ifdef __cplusplus
}
#endif
#endif /* BROWNIAN_H */
