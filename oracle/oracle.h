/* CPU oracle for the MPCD step of shashankgandham/brownian (cpp/ tree).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a serial FP64 restatement of the
 * reference's algorithm, used by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg as the checker.  The product (libbrownian)
 * never links, loads or calls it.
 *
 * Arrays are 0-based here (the reference is 1-based); particle and colloid
 * ids in maps and logs are 0-based too.  Every arithmetic expression keeps
 * the reference's operation order so that, with the same inputs and the same
 * generator stream, results are bit-identical to the compiled reference.
 */
#ifndef BROWNIAN_ORACLE_H
#define BROWNIAN_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int L;             /* cells per box edge (a = 1), len = (L,L,L)            */
    int density;       /* fluid particles per cell                             */
    int n_colloid;     /* no_of_colloid                                        */
    int n_sub;         /* substeps per MPCD step (reference n)                 */
    double ndt;        /* MPCD step                                            */
    double kbt;        /* kbt (fluid) and kbt1 (colloid init) are equal        */
    double mass_fl;
    double mass_colloid;
    double sig_colloid;/* LJ sigma                                             */
    double eps;
    double sigma;      /* colloid diameter (reference: 0.8*sig_colloid)        */
    double theta_deg;  /* SRD angle (reference hard-codes 130)                 */
    double v0;         /* active speed, used when v0_per_colloid is not set    */
    int seed;          /* generator seed                                       */
} orc_params;

typedef struct orc_state orc_state;

/* draw-log kinds */
enum { ORC_DRAW_SHIFT = 1, ORC_DRAW_AXIS = 2, ORC_DRAW_REFLECT = 3, ORC_DRAW_TUMBLE = 4 };

orc_state *orc_create(const orc_params *p);
void orc_destroy(orc_state *s);
int orc_n_fluid(const orc_state *s);
int orc_n_colloid(const orc_state *s);

/* generator (cpp/parameters.hpp:16-42), fresh-process start state */
void orc_rng_reset(orc_state *s, int seed);
double orc_ran(orc_state *s);
/* copy generator state in/out: 36 ints (seed, idum, iy, iv[1..32], pad) */
void orc_rng_get(const orc_state *s, int *out36);
void orc_rng_set(orc_state *s, const int *in36);

/* per-colloid active speed (NULL restores the scalar v0) */
void orc_set_v0(orc_state *s, const double *v0);

/* state access: positions/velocities as n x 3 doubles */
void orc_get_fluid(const orc_state *s, double *pos, double *vel);
void orc_set_fluid(orc_state *s, const double *pos, const double *vel);
void orc_get_colloid(const orc_state *s, double *pos, double *vel, double *omega, double *f,
                     double *old_f, double *ra);
void orc_set_colloid(orc_state *s, const double *pos, const double *vel, const double *omega,
                     const double *f, const double *old_f, const double *ra);
double orc_potential(const orc_state *s);

/* reference entry points (cpp/parameters.hpp:12-14) */
void orc_initialize_colloid(orc_state *s);
void orc_initialize_fluid(orc_state *s);
void orc_create_box(orc_state *s);
void orc_neighbour_list_mpcd(orc_state *s);
void orc_neighbour_list_md(orc_state *s);
void orc_compute_force_md(orc_state *s);
void orc_tumble(orc_state *s);
void orc_rotation_mpcd(orc_state *s);
void orc_run(orc_state *s);
void orc_update_pos_md(orc_state *s);
void orc_update_pos_mpcd(orc_state *s);
void orc_fluid_colloid_collision(orc_state *s);
void orc_update_activity_direction(orc_state *s);
void orc_update_velocity_colloid(orc_state *s);
void orc_copy_old_force(orc_state *s);
/* start-up sequence (cpp/active_brownian.cpp:14-21) and loop body (:25-38) */
void orc_startup(orc_state *s);
void orc_step(orc_state *s);
/* substep body only (cpp/active_brownian.cpp:28-37) */
void orc_substep(orc_state *s);

/* maps */
int orc_nbox(const orc_state *s);
void orc_get_box_neigh(const orc_state *s, int *out);          /* cells x nbox, 0-based */
void orc_get_fluid_no(const orc_state *s, int *out);           /* L^3, last binning     */
void orc_get_cell_no(const orc_state *s, int *out);            /* L^3, shifted binning  */
int orc_get_cell_lists(const orc_state *s, int which, int *out);/* 0 box_part, 1 cell_part */
int orc_get_neigh_fl(const orc_state *s, int *counts, int *out);
int orc_get_md_list(const orc_state *s, int *counts, int *out);
int orc_get_run_lists(const orc_state *s, int *counts, int *out);

/* cpp/run.cpp:28-50 (a diagnostic the reference computes and discards):
 * rear/up shell counts, the up-shell lists, and per colloid the reference's
 * vel and up_vel after that colloid (n x 3 each) */
void orc_updown_velocity(orc_state *s);
void orc_get_updown(const orc_state *s, int *cnt, int *up_cnt, double *vel, double *up_vel);
int orc_get_up_lists(const orc_state *s, int *counts, int *out);

/* draw log: every draw consumed since the last orc_drawlog_clear */
void orc_drawlog_clear(orc_state *s);
int orc_drawlog_count(const orc_state *s);
/* entry i: kind, key0, key1, key2 (ints) and up to 8 doubles */
void orc_drawlog_get(const orc_state *s, int *kinds, int *keys /* n x 3 */, double *vals /* n x 8 */);

/* energies of the audit block at cpp/active_brownian.cpp:39-55 */
void orc_audit(const orc_state *s, double *out /* ke_colloid, ang_ke, potential, ke_fluid, px, py, pz */);

#ifdef __cplusplus
}
#endif
#endif
