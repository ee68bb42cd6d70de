// Test-infrastructure harness around the reference C++ program (cpp/ of
// shashankgandham/brownian).  It is compiled together with the reference's own
// source files, read in place under the reference tree, by oracle/Makefile, and
// the result lands in oracle/_ref/.  Nothing in the product links it.
//
// The reference keeps every quantity in extern globals (cpp/parameters.hpp:8-10,72,
// defined at cpp/active_brownian.cpp:3-7) and its main() is built as ref_main so
// those definitions stay.  This file only sets the globals, calls the reference
// entry points in the order of cpp/active_brownian.cpp:14-38 and copies arrays
// out (1-based in the reference, 0-based here).
#include "parameters.hpp"
#include <pthread.h>

// cpp/update.cpp:3-6 derives these from sig_colloid at static-init time; they
// must be re-derived whenever the harness changes sig_colloid.
extern double r_cutoff, fc, ufc, sig_colloid12, sig_colloid6;

extern "C" {

// Set the reference globals for a run.  Mirrors the defaults block at
// cpp/active_brownian.cpp:3-7 with every parameter exposed.
void ref_setup(int L, int density, int ncol, int nsub, double ndt_, double kbt_,
               double mass_fl_, double mass_colloid_, double sig_colloid_, double eps_,
               double sigma_, double v0_, int seed_) {
    len = point(L, L, L);
    no_of_fluid = (int)len.prod() * density;
    no_of_colloid = ncol;
    n = nsub;
    ndt = ndt_;
    dt = ndt / (double)n;
    kbt = kbt_;
    kbt1 = kbt_;
    mass_fl = mass_fl_;
    mass_colloid = mass_colloid_;
    sig_colloid = sig_colloid_;
    eps = eps_;
    v0 = v0_;
    sigma = sigma_;
    I_colloid = 0.1 * mass_colloid * sigma * sigma;
    seed = seed_;
    r_cutoff = pow(2, 1.0 / 6.0) * sig_colloid;
    fc = 4.0 * eps * (12.0 * (pow(sig_colloid, 12) / pow(r_cutoff, 13)) - 6.0 * (pow(sig_colloid, 6) / pow(r_cutoff, 7)));
    ufc = 4.0 * eps * (pow(sig_colloid / r_cutoff, 12) - pow(sig_colloid / r_cutoff, 6)) + fc * r_cutoff;
    sig_colloid12 = pow(sig_colloid, 12);
    sig_colloid6 = pow(sig_colloid, 6);
}

double ref_ran(void) { return ran(); }

void ref_initialize(void) { initialize(); }
void ref_initialize_colloid(void) { initialize_colloid(); }
void ref_initialize_fluid(void) { initialize_fluid(); }
void ref_create_box(void) { create_box(); }
void ref_neighbour_list_mpcd(void) { neighbour_list_mpcd(); }
void ref_neighbour_list_md(void) { neighbour_list_md(); }
void ref_compute_force_md(void) { compute_force_md(); }
void ref_tumble(void) { tumble(); }
void ref_rotation_mpcd(void) { rotation_mpcd(); }
void ref_run(void) { run(); }
void ref_update_pos_md(void) { update_pos_md(); }
void ref_update_pos_mpcd(void) { update_pos_mpcd(); }
void ref_fluid_colloid_collision(void) { fluid_colloid_collision(); }
void ref_update_activity_direction(void) { update_activity_direction(); }
void ref_update_velocity_colloid(void) { update_velocity_colloid(); }
void ref_copy_old_force(void) { std::copy(f, f + no_of_colloid + 2, old_force); }

// The whole start-up sequence of cpp/active_brownian.cpp:14-21.
void ref_startup(void) {
    initialize();
    initialize_colloid();
    initialize_fluid();
    create_box();
    neighbour_list_mpcd();
    neighbour_list_md();
    compute_force_md();
    tumble();
}

// One MPCD step: the body of the loop at cpp/active_brownian.cpp:23-38.
// updown_velocity (cpp/run.cpp:28-50) is left out: it only writes scratch lists
// that run() rebuilds, and is reached only after step 10000.
static void one_step(void) {
    rotation_mpcd();
    run();
    for (int l = 1; l <= n; l++) {
        std::copy(f, f + no_of_colloid + 2, old_force);
        update_pos_md();
        neighbour_list_md();
        update_pos_mpcd();
        neighbour_list_mpcd();
        fluid_colloid_collision();
        update_activity_direction();
        compute_force_md();
        update_velocity_colloid();
    }
}

struct steps_arg { int nsteps; };
static void *steps_thread(void *p) {
    steps_arg *a = (steps_arg *)p;
    for (int s = 0; s < a->nsteps; s++) one_step();
    return 0;
}

// rotation_mpcd keeps a per-cell array on the stack (cpp/rotation_mpcd.cpp:6),
// so large boxes run on a thread with a large stack.
int ref_steps(int nsteps) {
    pthread_attr_t attr;
    pthread_attr_init(&attr);
    size_t need = (size_t)(len.prod() + 16) * sizeof(point) * 2 + (64u << 20);
    pthread_attr_setstacksize(&attr, need);
    pthread_t th;
    steps_arg a = {nsteps};
    int rc = pthread_create(&th, &attr, steps_thread, &a);
    if (rc) return rc;
    pthread_join(th, 0);
    pthread_attr_destroy(&attr);
    return 0;
}

// Run only rotation_mpcd on the big-stack thread (same reason as above).
static void *rot_thread(void *) { rotation_mpcd(); return 0; }
int ref_rotation_mpcd_bigstack(void) {
    pthread_attr_t attr;
    pthread_attr_init(&attr);
    size_t need = (size_t)(len.prod() + 16) * sizeof(point) * 2 + (64u << 20);
    pthread_attr_setstacksize(&attr, need);
    pthread_t th;
    int rc = pthread_create(&th, &attr, rot_thread, 0);
    if (rc) return rc;
    pthread_join(th, 0);
    pthread_attr_destroy(&attr);
    return 0;
}

static void get_points(const point *src, int count, double *out) {
    for (int i = 0; i < count; i++) {
        out[3 * i + 0] = src[i + 1].x;
        out[3 * i + 1] = src[i + 1].y;
        out[3 * i + 2] = src[i + 1].z;
    }
}
static void set_points(point *dst, int count, const double *in) {
    for (int i = 0; i < count; i++) dst[i + 1] = point(in[3 * i], in[3 * i + 1], in[3 * i + 2]);
}

int ref_no_of_fluid(void) { return no_of_fluid; }
int ref_no_of_colloid(void) { return no_of_colloid; }
int ref_nbox(void) { return nbox; }
double ref_potential(void) { return potential_colloid; }

void ref_get_fluid(double *pos, double *vel) {
    get_points(pos_fl, no_of_fluid, pos);
    get_points(vel_fl, no_of_fluid, vel);
}
void ref_set_fluid(const double *pos, const double *vel) {
    set_points(pos_fl, no_of_fluid, pos);
    set_points(vel_fl, no_of_fluid, vel);
}
// colloid block: pos, vel, ang_vel, f, old_force, ra (each ncol x 3)
void ref_get_colloid(double *pos, double *vel, double *omega, double *force, double *oldf, double *ra_) {
    get_points(pos_colloid, no_of_colloid, pos);
    get_points(vel_colloid, no_of_colloid, vel);
    get_points(ang_vel_colloid, no_of_colloid, omega);
    get_points(f, no_of_colloid, force);
    get_points(old_force, no_of_colloid, oldf);
    get_points(ra, no_of_colloid, ra_);
}
void ref_set_colloid(const double *pos, const double *vel, const double *omega, const double *force,
                     const double *oldf, const double *ra_) {
    set_points(pos_colloid, no_of_colloid, pos);
    set_points(vel_colloid, no_of_colloid, vel);
    set_points(ang_vel_colloid, no_of_colloid, omega);
    set_points(f, no_of_colloid, force);
    set_points(old_force, no_of_colloid, oldf);
    set_points(ra, no_of_colloid, ra_);
}

// box_neigh[k][cell], k = 1..nbox, cell = 1..L^3 -> out[(cell-1)*nbox + (k-1)], 1-based cell ids
void ref_get_box_neigh(int *out) {
    int cells = (int)len.prod();
    for (int c = 1; c <= cells; c++)
        for (int k = 1; k <= nbox; k++) out[(size_t)(c - 1) * nbox + (k - 1)] = box_neigh[k][c];
}
// fluid_no[cell] for cell = 1..L^3
void ref_get_fluid_no(int *out) {
    int cells = (int)len.prod();
    for (int c = 1; c <= cells; c++) out[c - 1] = fluid_no[c];
}
// concatenated per-cell lists: box_part (which=0) or cell_part (which=1), 1-based particle ids
int ref_get_cell_lists(int which, int *out) {
    int cells = (int)len.prod(), w = 0;
    int **tab = which ? cell_part : box_part;
    for (int c = 1; c <= cells; c++)
        for (int j = 1; j <= fluid_no[c]; j++) out[w++] = tab[j][c];
    return w;
}
// neigh_fl: counts per colloid, then concatenated lists (1-based particle ids)
int ref_get_neigh_fl(int *counts, int *out) {
    int w = 0;
    for (int j = 1; j <= no_of_colloid; j++) {
        counts[j - 1] = no_neigh[j];
        for (int i = 1; i <= no_neigh[j]; i++) out[w++] = neigh_fl[j][i];
    }
    return w;
}
// MD neighbour list: counts per colloid, concatenated partner ids (1-based)
int ref_get_md_list(int *counts, int *out) {
    int w = 0;
    for (int i = 1; i <= no_of_colloid; i++) {
        counts[i - 1] = n_neighbour[i];
        for (int j = 1; j <= n_neighbour[i]; j++) out[w++] = neighbour[j][i];
    }
    return w;
}
// cpp/run.cpp:28-50 (rewrites cnt/nbr and up_cnt/up_nbr; its velocities are locals)
void ref_updown_velocity(void) { updown_velocity(); }
// updown_velocity's near-surface lists: up_cnt per colloid, then up_nbr entries
int ref_get_up_lists(int *counts, int *out) {
    int w = 0;
    for (int i = 1; i <= no_of_colloid; i++) {
        counts[i - 1] = up_cnt[i];
        for (int j = 1; j <= up_cnt[i]; j++) out[w++] = up_nbr[j][i];
    }
    return w;
}
// run() rear-shell lists: cnt per colloid, then concatenated nbr entries
int ref_get_run_lists(int *counts, int *out) {
    int w = 0;
    for (int i = 1; i <= no_of_colloid; i++) {
        counts[i - 1] = cnt[i];
        for (int j = 1; j <= cnt[i]; j++) out[w++] = nbr[j][i];
    }
    return w;
}

}  // extern "C"
