/* CPU oracle for the MPCD step of shashankgandham/brownian (cpp/ tree).
 *
 * TEST INFRASTRUCTURE ONLY: a serial FP64 restatement of the reference
 * algorithm.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, as the checker; libbrownian never links it.
 *
 * Each function cites the reference file:line it restates.  Expressions keep
 * the reference's evaluation order (point operators of cpp/parameters.hpp:44-71
 * are component-wise and left-associative), and the reference's libm calls
 * (pow, sqrt, log, cos, sin, asin, lround) are made in the same places, so that
 * with the same inputs and generator stream the results are bit-identical to the
 * compiled reference.  Build with -ffp-contract=off.
 *
 * Pinning: tests/test_oracle_golden.py compares every entry point, the
 * generator stream and whole multi-step trajectories against vectors written by
 * the compiled reference (tests/golden/).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double x, y, z; } v3;

static inline v3 V(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vdiv(v3 a, v3 b) { return V(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline v3 vscale(v3 a, double s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 vdivs(v3 a, double s) { return V(a.x / s, a.y / s, a.z / s); }
static inline double vsum(v3 a) { return a.x + a.y + a.z; }
/* cpp/parameters.hpp:74 round() with lround (half away from zero) */
static inline v3 vround(v3 a) { return V((double)lround(a.x), (double)lround(a.y), (double)lround(a.z)); }
/* cpp/parameters.hpp:75 mod(a,b) = a - b*round((a - b/2)/b) */
static inline v3 vmod(v3 a, v3 b) { return vsub(a, vmul(b, vround(vdiv(vsub(a, vdivs(b, 2)), b)))); }
/* cpp/parameters.hpp:76 img(a,b) = a - b*round(a/b) */
static inline v3 vimg(v3 a, v3 b) { return vsub(a, vmul(b, vround(vdiv(a, b)))); }
/* cpp/fluid_colloid_collision.cpp:3-5 */
static inline v3 vcross(v3 a, v3 b) {
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* cpp/parameters.hpp:59 point::cell, 0-based */
static inline int vcell(v3 a, v3 len) {
    return (int)((double)(int)a.x + len.x * (double)(int)a.y + len.x * len.y * (double)(int)a.z);
}

typedef struct {
    int kind, key[3];
    double val[8];
} drawrec;

struct orc_state {
    orc_params p;
    int L, nf, nc, cells;
    v3 len;
    double dt, I_colloid;
    double r_cutoff, fc, ufc, sig12, sig6;
    /* generator (cpp/parameters.hpp:16-18): seed is the global, idum/iy/iv the statics */
    int seed, idum, iy, iv[33];
    v3 *pos_fl, *vel_fl, *dump;
    v3 *pos_c, *vel_c, *omega, *f, *old_f, *ra;
    double *v0;  /* per colloid */
    double potential;
    /* unshifted binning (box_part) */
    int *fluid_no, *box_start, *box_list;
    /* shifted binning (cell_part) */
    int *cell_no, *cell_start, *cell_list;
    /* stencil */
    int nbox;
    int *box_neigh; /* cells x 343 */
    /* neigh_fl */
    int *no_neigh, *neigh_start, *neigh_list;
    size_t neigh_cap;
    /* md list */
    int *n_neighbour, *neighbour; /* nc x nc */
    /* run lists */
    int *cnt, *nbr_start, *nbr_list;
    size_t nbr_cap;
    /* updown_velocity: up-shell lists and the per-colloid results */
    int *up_cnt, *up_start, *up_list;
    size_t up_cap;
    v3 *ud_vel, *ud_up;
    /* draw log */
    drawrec *log;
    int nlog, caplog;
};

static void logdraw(orc_state *s, int kind, int k0, int k1, int k2, const double *v, int nv) {
    if (s->nlog == s->caplog) {
        s->caplog = s->caplog ? 2 * s->caplog : 1024;
        s->log = (drawrec *)realloc(s->log, sizeof(drawrec) * (size_t)s->caplog);
    }
    drawrec *d = &s->log[s->nlog++];
    memset(d, 0, sizeof(*d));
    d->kind = kind;
    d->key[0] = k0; d->key[1] = k1; d->key[2] = k2;
    for (int i = 0; i < nv; i++) d->val[i] = v[i];
}

/* ---------------------------------------------------------------- generator */
/* cpp/parameters.hpp:16-42: two L'Ecuyer LCGs (Schrage's method) combined
 * through a 32-slot Bays-Durham table.  A positive seed starts from the
 * static initial values: idum = 123456789, iy = 0, iv[] = 0 (calloc). */
void orc_rng_reset(orc_state *s, int seed) {
    s->seed = seed;
    s->idum = 123456789;
    s->iy = 0;
    memset(s->iv, 0, sizeof(s->iv));
}

double orc_ran(orc_state *s) {
    const int im1 = 2147483563, im2 = 2147483399, ia1 = 40014, ia2 = 40692, iq1 = 53668, iq2 = 52774;
    const int imm = im1 - 1, ir1 = 12211, ir2 = 3791, ntab = 32, ndiv = 1 + imm / ntab;
    const double eps = 1.2e-7, rnmx = 1 - eps, am = 1.0 / im1;
    int j, k;
    if (s->seed <= 0) {
        s->seed = (-s->seed > 1) ? -s->seed : 1;
        s->idum = s->seed;
        for (j = ntab + 8; j >= 1; j--) {
            k = s->seed / iq1;
            s->seed = ia1 * (s->seed - k * iq1) - k * ir1;
            if (s->seed < 0) s->seed += im1;
            if (j <= ntab) s->iv[j] = s->seed;
        }
        s->iy = s->iv[1];
    }
    k = s->seed / iq1;
    s->seed = ia1 * (s->seed - k * iq1) - k * ir1;
    if (s->seed < 0) s->seed += im1;
    k = s->idum / iq2;
    s->idum = ia2 * (s->idum - k * iq2) - k * ir2;
    if (s->idum < 0) s->idum = s->idum + im2;
    j = 1 + s->iy / ndiv;
    s->iy = s->iv[j] - s->idum;
    s->iv[j] = s->seed;
    if (s->iy < 1) s->iy = s->iy + imm;
    double r = am * s->iy;
    return r < rnmx ? r : rnmx;
}

void orc_rng_get(const orc_state *s, int *o) {
    o[0] = s->seed; o[1] = s->idum; o[2] = s->iy;
    for (int i = 1; i <= 32; i++) o[2 + i] = s->iv[i];
    o[35] = 0;
}
void orc_rng_set(orc_state *s, const int *in) {
    s->seed = in[0]; s->idum = in[1]; s->iy = in[2];
    for (int i = 1; i <= 32; i++) s->iv[i] = in[2 + i];
}

/* point::random (cpp/parameters.hpp:61-65): (ran,ran,ran)*mul - dec */
static inline v3 vrandom(orc_state *s, v3 dec, v3 mul, double *raw) {
    double x = orc_ran(s), y = orc_ran(s), z = orc_ran(s);
    if (raw) { raw[0] = x; raw[1] = y; raw[2] = z; }
    return vsub(vmul(V(x, y, z), mul), dec);
}

/* ----------------------------------------------------------------- set-up */
orc_state *orc_create(const orc_params *p) {
    orc_state *s = (orc_state *)calloc(1, sizeof(orc_state));
    s->p = *p;
    s->L = p->L;
    s->len = V(p->L, p->L, p->L);
    s->cells = p->L * p->L * p->L;
    s->nf = s->cells * p->density;
    s->nc = p->n_colloid;
    /* cpp/active_brownian.cpp:7 */
    s->dt = p->ndt / (double)p->n_sub;
    s->I_colloid = 0.1 * p->mass_colloid * p->sigma * p->sigma;
    /* cpp/update.cpp:3-6 */
    s->r_cutoff = pow(2, 1.0 / 6.0) * p->sig_colloid;
    s->fc = 4.0 * p->eps * (12.0 * (pow(p->sig_colloid, 12) / pow(s->r_cutoff, 13)) -
                            6.0 * (pow(p->sig_colloid, 6) / pow(s->r_cutoff, 7)));
    s->ufc = 4.0 * p->eps * (pow(p->sig_colloid / s->r_cutoff, 12) - pow(p->sig_colloid / s->r_cutoff, 6)) +
             s->fc * s->r_cutoff;
    s->sig12 = pow(p->sig_colloid, 12);
    s->sig6 = pow(p->sig_colloid, 6);
    orc_rng_reset(s, p->seed);
    size_t nf = (size_t)s->nf, nc = (size_t)(s->nc > 0 ? s->nc : 1), cells = (size_t)s->cells;
    s->pos_fl = (v3 *)calloc(nf, sizeof(v3));
    s->vel_fl = (v3 *)calloc(nf, sizeof(v3));
    s->dump = (v3 *)calloc(nf, sizeof(v3));
    s->pos_c = (v3 *)calloc(nc, sizeof(v3));
    s->vel_c = (v3 *)calloc(nc, sizeof(v3));
    s->omega = (v3 *)calloc(nc, sizeof(v3));
    s->f = (v3 *)calloc(nc, sizeof(v3));
    s->old_f = (v3 *)calloc(nc, sizeof(v3));
    s->ra = (v3 *)calloc(nc, sizeof(v3));
    s->v0 = (double *)calloc(nc, sizeof(double));
    for (size_t i = 0; i < nc; i++) s->v0[i] = p->v0;
    s->fluid_no = (int *)calloc(cells, sizeof(int));
    s->box_start = (int *)calloc(cells + 1, sizeof(int));
    s->box_list = (int *)calloc(nf + 1, sizeof(int));
    s->cell_no = (int *)calloc(cells, sizeof(int));
    s->cell_start = (int *)calloc(cells + 1, sizeof(int));
    s->cell_list = (int *)calloc(nf + 1, sizeof(int));
    s->box_neigh = (int *)calloc(cells * 343, sizeof(int));
    s->no_neigh = (int *)calloc(nc, sizeof(int));
    s->neigh_start = (int *)calloc(nc + 1, sizeof(int));
    s->neigh_cap = 1024;
    s->neigh_list = (int *)calloc(s->neigh_cap, sizeof(int));
    s->n_neighbour = (int *)calloc(nc, sizeof(int));
    s->neighbour = (int *)calloc(nc * nc, sizeof(int));
    s->cnt = (int *)calloc(nc, sizeof(int));
    s->nbr_start = (int *)calloc(nc + 1, sizeof(int));
    s->nbr_cap = 1024;
    s->nbr_list = (int *)calloc(s->nbr_cap, sizeof(int));
    s->up_cnt = (int *)calloc(nc, sizeof(int));
    s->up_start = (int *)calloc(nc + 1, sizeof(int));
    s->up_cap = 1024;
    s->up_list = (int *)calloc(s->up_cap, sizeof(int));
    s->ud_vel = (v3 *)calloc(nc, sizeof(v3));
    s->ud_up = (v3 *)calloc(nc, sizeof(v3));
    return s;
}

void orc_destroy(orc_state *s) {
    if (!s) return;
    free(s->pos_fl); free(s->vel_fl); free(s->dump);
    free(s->pos_c); free(s->vel_c); free(s->omega); free(s->f); free(s->old_f); free(s->ra); free(s->v0);
    free(s->fluid_no); free(s->box_start); free(s->box_list);
    free(s->cell_no); free(s->cell_start); free(s->cell_list);
    free(s->box_neigh); free(s->no_neigh); free(s->neigh_start); free(s->neigh_list);
    free(s->n_neighbour); free(s->neighbour);
    free(s->cnt); free(s->nbr_start); free(s->nbr_list);
    free(s->up_cnt); free(s->up_start); free(s->up_list); free(s->ud_vel); free(s->ud_up);
    free(s->log);
    free(s);
}

int orc_n_fluid(const orc_state *s) { return s->nf; }
int orc_n_colloid(const orc_state *s) { return s->nc; }
double orc_potential(const orc_state *s) { return s->potential; }

void orc_set_v0(orc_state *s, const double *v0) {
    for (int i = 0; i < s->nc; i++) s->v0[i] = v0 ? v0[i] : s->p.v0;
}

static void get3(const v3 *a, int n, double *o) {
    for (int i = 0; i < n; i++) { o[3 * i] = a[i].x; o[3 * i + 1] = a[i].y; o[3 * i + 2] = a[i].z; }
}
static void set3(v3 *a, int n, const double *in) {
    for (int i = 0; i < n; i++) a[i] = V(in[3 * i], in[3 * i + 1], in[3 * i + 2]);
}
void orc_get_fluid(const orc_state *s, double *pos, double *vel) {
    if (pos) get3(s->pos_fl, s->nf, pos);
    if (vel) get3(s->vel_fl, s->nf, vel);
}
void orc_set_fluid(orc_state *s, const double *pos, const double *vel) {
    if (pos) set3(s->pos_fl, s->nf, pos);
    if (vel) set3(s->vel_fl, s->nf, vel);
}
void orc_get_colloid(const orc_state *s, double *pos, double *vel, double *omega, double *f, double *old_f,
                     double *ra) {
    if (pos) get3(s->pos_c, s->nc, pos);
    if (vel) get3(s->vel_c, s->nc, vel);
    if (omega) get3(s->omega, s->nc, omega);
    if (f) get3(s->f, s->nc, f);
    if (old_f) get3(s->old_f, s->nc, old_f);
    if (ra) get3(s->ra, s->nc, ra);
}
void orc_set_colloid(orc_state *s, const double *pos, const double *vel, const double *omega, const double *f,
                     const double *old_f, const double *ra) {
    if (pos) set3(s->pos_c, s->nc, pos);
    if (vel) set3(s->vel_c, s->nc, vel);
    if (omega) set3(s->omega, s->nc, omega);
    if (f) set3(s->f, s->nc, f);
    if (old_f) set3(s->old_f, s->nc, old_f);
    if (ra) set3(s->ra, s->nc, ra);
}

/* ------------------------------------------------------------- initialise */
/* cpp/initialize.cpp:31-60.  The lattice pre-fill at :36-39 only writes slots
 * that the random placement overwrites (and the unused slot 0), so it has no
 * effect on the state and consumes no draws. */
void orc_initialize_colloid(orc_state *s) {
    int counter = 0, check;
    double kbt1 = s->p.kbt;
    double space_limit = 1.3 * s->p.sig_colloid, ang_vscale_colloid = sqrt(12.0 * kbt1 / s->I_colloid),
           vscale_colloid = sqrt(12.0 * kbt1 / s->p.mass_colloid);
    v3 avr_vel = V(0, 0, 0), t, temp;
    while (counter < s->nc) {
        t = vrandom(s, V(0, 0, 0), s->len, 0);
        check = 1;
        for (int j = 0; j < counter; j++) {
            temp = vimg(vsub(t, s->pos_c[j]), s->len);
            check = (sqrt(vsum(vmul(temp, temp))) < space_limit) ? 0 : check;
        }
        if (check) s->pos_c[counter++] = t;
    }
    for (int j = 0; j < s->nc; j++) {
        s->vel_c[j] = vscale(vrandom(s, V(0.5, 0.5, 0.5), V(1, 1, 1), 0), vscale_colloid);
        avr_vel = vadd(avr_vel, s->vel_c[j]);
    }
    avr_vel = vdivs(avr_vel, (double)s->nc);
    for (int j = 0; j < s->nc; j++) {
        s->vel_c[j] = vsub(s->vel_c[j], avr_vel);
        s->omega[j] = vscale(vrandom(s, V(0.5, 0.5, 0.5), V(1, 1, 1), 0), ang_vscale_colloid);
    }
}

/* cpp/initialize.cpp:62-83 */
void orc_initialize_fluid(orc_state *s) {
    int counter = 0, check;
    double vscale_fluid = sqrt(12.0 * s->p.kbt / s->p.mass_fl);
    v3 avr_vel = V(0, 0, 0), t, temp;
    while (counter < s->nf) {
        t = vrandom(s, V(0, 0, 0), s->len, 0);
        check = 1;
        for (int j = 0; j < s->nc; j++) {
            temp = vimg(vsub(t, s->pos_c[j]), s->len);
            check = (sqrt(vsum(vmul(temp, temp))) < s->p.sigma * 0.5) ? 0 : check;
        }
        if (check) s->pos_fl[counter++] = t;
    }
    for (int j = 0; j < s->nf; j++) {
        s->vel_fl[j] = vscale(vrandom(s, V(0.5, 0.5, 0.5), V(1, 1, 1), 0), vscale_fluid);
        avr_vel = vadd(avr_vel, s->vel_fl[j]);
    }
    avr_vel = vdivs(avr_vel, (double)s->nf);
    for (int j = 0; j < s->nf; j++) s->vel_fl[j] = vsub(s->vel_fl[j], avr_vel);
}

/* ------------------------------------------------------------ cell lists */
/* cpp/neighbour_list.cpp:3-11 cmod: wrap a 1-based cell coordinate into [1, L] */
static inline v3 cmod(v3 a, v3 b) {
    if (a.x <= 0) a.x += b.x;
    if (a.y <= 0) a.y += b.y;
    if (a.z <= 0) a.z += b.z;
    if (a.x > b.x) a.x -= b.x;
    if (a.y > b.y) a.y -= b.y;
    if (a.z > b.z) a.z -= b.z;
    return a;
}

/* cpp/neighbour_list.cpp:12-25: for every cell, the 7^3 stencil in
 * x-fastest order from -3 to +3 with periodic wrap, minus the cell itself. */
void orc_create_box(orc_state *s) {
    int L = s->L;
    for (int z = 1; z <= L; z++)
        for (int y = 1; y <= L; y++)
            for (int x = 1; x <= L; x++) {
                v3 iter = V(x, y, z);
                int box = vcell(vsub(iter, V(0, 1, 1)), s->len) - 1; /* 0-based */
                int nb = 0;
                for (int dz = -3; dz <= 3; dz++)
                    for (int dy = -3; dy <= 3; dy++)
                        for (int dx = -3; dx <= 3; dx++) {
                            v3 j = cmod(V(x + dx, y + dy, z + dz), s->len);
                            int tbox = vcell(vsub(j, V(0, 1, 1)), s->len) - 1;
                            if (tbox != box) s->box_neigh[(size_t)box * 343 + nb++] = tbox;
                        }
                s->nbox = nb;
            }
}

/* counting-sort binning that keeps ascending particle order within a cell,
 * i.e. the order the reference's appends produce */
static void bin_particles(int nf, int cells, const int *key, int *count, int *start, int *list) {
    memset(count, 0, sizeof(int) * (size_t)cells);
    for (int i = 0; i < nf; i++) count[key[i]]++;
    start[0] = 0;
    for (int c = 0; c < cells; c++) start[c + 1] = start[c] + count[c];
    int *fill = (int *)malloc(sizeof(int) * (size_t)cells);
    memcpy(fill, start, sizeof(int) * (size_t)cells);
    for (int i = 0; i < nf; i++) list[fill[key[i]]++] = i;
    free(fill);
}

/* cpp/neighbour_list.cpp:41-58 */
void orc_neighbour_list_mpcd(orc_state *s) {
    int *key = (int *)malloc(sizeof(int) * (size_t)(s->nf ? s->nf : 1));
    for (int i = 0; i < s->nf; i++) key[i] = vcell(s->pos_fl[i], s->len);
    bin_particles(s->nf, s->cells, key, s->fluid_no, s->box_start, s->box_list);
    free(key);
    size_t w = 0;
    for (int j = 0; j < s->nc; j++) {
        s->neigh_start[j] = (int)w;
        int cbox = vcell(s->pos_c[j], s->len);
        for (int k = 0; k < s->nbox; k++) {
            int mm = s->box_neigh[(size_t)cbox * 343 + k];
            int n = s->fluid_no[mm];
            if (w + (size_t)n > s->neigh_cap) {
                while (w + (size_t)n > s->neigh_cap) s->neigh_cap *= 2;
                s->neigh_list = (int *)realloc(s->neigh_list, sizeof(int) * s->neigh_cap);
            }
            memcpy(s->neigh_list + w, s->box_list + s->box_start[mm], sizeof(int) * (size_t)n);
            w += (size_t)n;
        }
        s->no_neigh[j] = (int)(w - (size_t)s->neigh_start[j]);
    }
    s->neigh_start[s->nc] = (int)w;
}

/* cpp/neighbour_list.cpp:27-39 */
void orc_neighbour_list_md(orc_state *s) {
    double neigh_cutoff = 3.0 * s->p.sig_colloid;
    v3 temp;
    memset(s->n_neighbour, 0, sizeof(int) * (size_t)(s->nc ? s->nc : 1));
    for (int i = 0; i < s->nc - 1; i++)
        for (int j = i + 1; j < s->nc; j++) {
            temp = vimg(vsub(s->pos_c[i], s->pos_c[j]), s->len);
            if (vsum(vmul(temp, temp)) < pow(neigh_cutoff, 2))
                s->neighbour[(size_t)i * s->nc + s->n_neighbour[i]++] = j;
        }
}

/* ------------------------------------------------------------------ forces */
/* cpp/update.cpp:8-13 */
static double power(double x, int r) {
    double ans = 1;
    for (int i = 1; i <= r; i++) ans *= x;
    return ans;
}

/* cpp/update.cpp:14-32 */
void orc_compute_force_md(orc_state *s) {
    v3 temp, ff;
    double r, t1, t2, mag_f, eps = s->p.eps, sig = s->p.sig_colloid;
    s->potential = 0;
    for (int i = 0; i < s->nc; i++) s->f[i] = V(0, 0, 0);
    for (int i = 0; i < s->nc; i++)
        for (int jj = 0; jj < s->n_neighbour[i]; jj++) {
            int j = s->neighbour[(size_t)i * s->nc + jj];
            temp = vimg(vsub(s->pos_c[i], s->pos_c[j]), s->len);
            r = sqrt(vsum(vmul(temp, temp)));
            if (r < s->r_cutoff) {
                s->potential += 4 * eps * (pow(sig / r, 12) - pow(sig / r, 6)) - s->ufc + s->fc * r;
                t1 = s->sig12 / power(r, 13), t2 = s->sig6 / power(r, 7);
                mag_f = 4.0 * eps * (12.0 * t1 - 6.0 * t2) - s->fc;
                ff = vdivs(vscale(temp, mag_f), r);
                s->f[i] = vadd(s->f[i], ff);
                s->f[j] = vsub(s->f[j], ff);
            }
        }
}

/* cpp/update.cpp:34-46 */
void orc_update_activity_direction(orc_state *s) {
    v3 m1, m2, m3, b, sb, cb;
    for (int i = 0; i < s->nc; i++) {
        b = vscale(s->omega[i], s->dt);
        sb = V(sin(b.x), sin(b.y), sin(b.z)), cb = V(cos(b.x), cos(b.y), cos(b.z));
        m1 = V(cb.y * cb.z, -cb.y * sb.z, sb.y);
        m2 = V(sb.x * sb.y * cb.z + cb.x * sb.z, -sb.x * sb.y * sb.z + cb.x * cb.z, -sb.x * cb.y);
        m3 = V(-cb.x * sb.y * cb.z + sb.x * sb.z, cb.x * sb.y * sb.z + sb.x * cb.z, cb.x * cb.y);
        s->ra[i] = V(vsum(vmul(m1, s->ra[i])), vsum(vmul(m2, s->ra[i])), vsum(vmul(m3, s->ra[i])));
    }
}

/* cpp/update.cpp:48-54 */
void orc_update_pos_md(orc_state *s) {
    double dt2 = s->dt * s->dt, ddt = 0.5 * dt2 / s->p.mass_colloid;
    for (int i = 0; i < s->nc; i++) {
        s->pos_c[i] = vadd(s->pos_c[i], vadd(vscale(s->vel_c[i], s->dt), vscale(s->f[i], ddt)));
        s->pos_c[i] = vmod(s->pos_c[i], s->len);
    }
}

/* cpp/update.cpp:56-60 */
void orc_update_pos_mpcd(orc_state *s) {
    for (int i = 0; i < s->nf; i++) s->pos_fl[i] = vmod(vadd(s->pos_fl[i], vscale(s->vel_fl[i], s->dt)), s->len);
}

/* cpp/update.cpp:62-67 */
void orc_update_velocity_colloid(orc_state *s) {
    double dtb2 = s->dt / (s->p.mass_colloid * 2);
    for (int i = 0; i < s->nc; i++) s->vel_c[i] = vadd(s->vel_c[i], vscale(vadd(s->old_f[i], s->f[i]), dtb2));
}

/* cpp/active_brownian.cpp:28 */
void orc_copy_old_force(orc_state *s) { memcpy(s->old_f, s->f, sizeof(v3) * (size_t)s->nc); }

/* ------------------------------------------------------------------- tumble */
/* cpp/run.cpp:3-8 */
void orc_tumble(orc_state *s) {
    for (int i = 0; i < s->nc; i++) {
        double raw[3];
        v3 r = vrandom(s, V(0, 0, 0), s->len, raw);
        logdraw(s, ORC_DRAW_TUMBLE, i, 0, 0, raw, 3);
        s->ra[i] = vimg(vsub(s->pos_c[i], r), s->len);
        s->ra[i] = vdivs(s->ra[i], sqrt(vsum(vmul(s->ra[i], s->ra[i]))));
    }
}

/* -------------------------------------------------------------------- SRD */
/* cpp/rotation_mpcd.cpp:3-50 */
void orc_rotation_mpcd(orc_state *s) {
    int cells = s->cells;
    double raw[3];
    v3 rr = vrandom(s, V(0.5, 0.5, 0.5), V(1, 1, 1), raw);
    logdraw(s, ORC_DRAW_SHIFT, 0, 0, 0, raw, 3);
    int *key = (int *)malloc(sizeof(int) * (size_t)(s->nf ? s->nf : 1));
    for (int i = 0; i < s->nf; i++) key[i] = vcell(vmod(vadd(s->pos_fl[i], rr), s->len), s->len);
    bin_particles(s->nf, cells, key, s->cell_no, s->cell_start, s->cell_list);
    free(key);
    v3 *cell_vel = (v3 *)calloc((size_t)cells, sizeof(v3));
    double theta_base = s->p.theta_deg / 180.0;
    for (int c = 0; c < cells; c++) {
        int N = s->cell_no[c];
        if (N > 1) {
            const int *lst = s->cell_list + s->cell_start[c];
            for (int j = 0; j < N; j++) cell_vel[c] = vadd(cell_vel[c], vdivs(s->vel_fl[lst[j]], (double)N));
            double u1 = orc_ran(s);
            double rho = 2 * u1 - 1;
            double u2 = orc_ran(s);
            double phi = 4.0 * asin(1.0) * u2, theta = 2 * asin(1) * theta_base;
            double ax[2] = {u1, u2};
            logdraw(s, ORC_DRAW_AXIS, c, 0, 0, ax, 2);
            double r1 = cos(phi) * sqrt(1 - rho * rho), r2 = sin(phi) * sqrt(1 - rho * rho), r3 = rho;
            double ct = cos(theta), st = sin(theta), ict = 1 - cos(theta);
            double ir1 = ict * r1, ir2 = ict * r2, ir3 = ict * r3;
            v3 rot1 = V(ir1 * r1 + ct, ir1 * r2 - st * r3, ir1 * r3 + st * r2);
            v3 rot2 = V(ir2 * r1 + st * r3, ir2 * r2 + ct, ir2 * r3 - st * r1);
            v3 rot3 = V(ir3 * r1 - st * r2, ir3 * r2 + st * r1, ir3 * r3 + ct);
            for (int j = 0; j < N; j++) {
                int k = lst[j];
                v3 del_v = vsub(s->vel_fl[k], cell_vel[c]);
                s->vel_fl[k] = vadd(cell_vel[c], V(vsum(vmul(rot1, del_v)), vsum(vmul(rot2, del_v)),
                                                   vsum(vmul(rot3, del_v))));
            }
        }
    }
    for (int c = 0; c < cells; c++) {
        int N = s->cell_no[c];
        double var = 0.0;
        if (N > 1) {
            const int *lst = s->cell_list + s->cell_start[c];
            for (int j = 0; j < N; j++) {
                v3 del_v = vsub(s->vel_fl[lst[j]], cell_vel[c]);
                var += vsum(vmul(del_v, del_v));
            }
            double scale = sqrt(3.0 * (N - 1) * s->p.kbt / (s->p.mass_fl * var));
            for (int j = 0; j < N; j++) {
                int k = lst[j];
                v3 del_v = vsub(s->vel_fl[k], cell_vel[c]);
                s->vel_fl[k] = vadd(cell_vel[c], vscale(del_v, scale));
            }
        }
    }
    free(cell_vel);
}

/* -------------------------------------------------------------- activity */
/* cpp/run.cpp:10-26 */
void orc_run(orc_state *s) {
    size_t w = 0;
    double lim = pow(s->p.sigma * 0.5 + 0.5, 2);
    for (int i = 0; i < s->nc; i++) {
        v3 del = vscale(s->ra[i], s->v0[i]);
        s->vel_c[i] = vadd(s->vel_c[i], vscale(s->ra[i], s->v0[i]));
        s->nbr_start[i] = (int)w;
        int c = 0;
        for (int jj = 0; jj < s->no_neigh[i]; jj++) {
            int l = s->neigh_list[s->neigh_start[i] + jj];
            v3 vector = vimg(vsub(s->pos_fl[l], s->pos_c[i]), s->len);
            if (vsum(vmul(vector, vector)) <= lim && vsum(vmul(vector, s->vel_c[i])) <= 0) {
                if (w + 1 > s->nbr_cap) {
                    s->nbr_cap *= 2;
                    s->nbr_list = (int *)realloc(s->nbr_list, sizeof(int) * s->nbr_cap);
                }
                s->nbr_list[w++] = l;
                c++;
            }
        }
        s->cnt[i] = c;
        for (int jj = 0; jj < c; jj++) {
            int l = s->nbr_list[s->nbr_start[i] + jj];
            s->vel_fl[l] = vsub(s->vel_fl[l], vscale(del, s->p.mass_colloid / (s->p.mass_fl * c)));
        }
    }
    s->nbr_start[s->nc] = (int)w;
}

/* cpp/run.cpp:28-50: rear-shell (|r| <= sigma/2 + 0.5) and near-surface
 * (|r| <= sigma/2 + 0.1) fluid behind each colloid (r . v_c <= 0) in neigh_fl
 * order, and their mean velocity relative to the colloid.  The reference
 * never resets up_vel between colloids and guards both divisions with
 * up_cnt > 0; kept as written.  ud_vel[i] / ud_up[i] are vel / up_vel after
 * colloid i (the reference discards them). */
void orc_updown_velocity(orc_state *s) {
    v3 up_vel = V(0, 0, 0);
    size_t w = 0, wu = 0;
    double lim = pow(s->p.sigma * 0.5 + 0.5, 2), lim_up = pow(s->p.sigma * 0.5 + 0.1, 2);
    for (int i = 0; i < s->nc; i++) {
        v3 vel = V(0, 0, 0);
        int c = 0, cu = 0;
        s->nbr_start[i] = (int)w;
        s->up_start[i] = (int)wu;
        for (int jj = 0; jj < s->no_neigh[i]; jj++) {
            int l = s->neigh_list[s->neigh_start[i] + jj];
            v3 vector = vimg(vsub(s->pos_fl[l], s->pos_c[i]), s->len);
            if (vsum(vmul(vector, vector)) <= lim && vsum(vmul(vector, s->vel_c[i])) <= 0.0) {
                if (w + 1 > s->nbr_cap) {
                    s->nbr_cap *= 2;
                    s->nbr_list = (int *)realloc(s->nbr_list, sizeof(int) * s->nbr_cap);
                }
                s->nbr_list[w++] = l;
                c++;
            }
            if (vsum(vmul(vector, vector)) <= lim_up && vsum(vmul(vector, s->vel_c[i])) <= 0.0) {
                if (wu + 1 > s->up_cap) {
                    s->up_cap *= 2;
                    s->up_list = (int *)realloc(s->up_list, sizeof(int) * s->up_cap);
                }
                s->up_list[wu++] = l;
                cu++;
            }
        }
        s->cnt[i] = c;
        s->up_cnt[i] = cu;
        for (int jj = 0; jj < c; jj++) vel = vadd(vel, s->vel_fl[s->nbr_list[s->nbr_start[i] + jj]]);
        for (int jj = 0; jj < cu; jj++) up_vel = vadd(up_vel, s->vel_fl[s->up_list[s->up_start[i] + jj]]);
        up_vel = cu > 0 ? vsub(vdivs(up_vel, cu), s->vel_c[i]) : up_vel;
        vel = cu > 0 ? vsub(vdivs(vel, c), s->vel_c[i]) : vel;
        s->ud_vel[i] = vel;
        s->ud_up[i] = up_vel;
    }
    s->nbr_start[s->nc] = (int)w;
    s->up_start[s->nc] = (int)wu;
}

void orc_get_updown(const orc_state *s, int *cnt, int *up_cnt, double *vel, double *up_vel) {
    for (int i = 0; i < s->nc; i++) {
        cnt[i] = s->cnt[i];
        up_cnt[i] = s->up_cnt[i];
    }
    get3(s->ud_vel, s->nc, vel);
    get3(s->ud_up, s->nc, up_vel);
}
int orc_get_up_lists(const orc_state *s, int *counts, int *out) {
    for (int i = 0; i < s->nc; i++) counts[i] = s->up_cnt[i];
    int w = s->up_start[s->nc];
    memcpy(out, s->up_list, sizeof(int) * (size_t)w);
    return w;
}

/* ------------------------------------------------------------- collisions */
/* cpp/fluid_colloid_collision.cpp:7-24 */
static v3 stochastic_reflection(orc_state *s, v3 rf, v3 rs, int j, int l) {
    double U_e = orc_ran(s);
    double m_beta = s->p.mass_fl / s->p.kbt, random_e = pow(1 - U_e, 2), val, x1 = 0, x2 = 0, z = 2;
    double kbt = s->p.kbt, mass_fl = s->p.mass_fl;
    v3 un, ut, t, n;
    n = vdivs(rs, sqrt(vsum(vmul(rs, rs))));
    val = sqrt(-log(random_e) / m_beta);
    un = vscale(n, val);
    double traw[3];
    t = vimg(vrandom(s, rf, s->len, traw), s->len);
    ut = vcross(un, t);
    ut = vdivs(ut, sqrt(vsum(vmul(ut, ut))));
    double a = 0, b = 0;
    int rounds = 0;
    while (z > 1) {
        a = orc_ran(s);
        x1 = 2.0 * a - 1;
        b = orc_ran(s);
        x2 = 2.0 * b - 1;
        z = x1 * x1 + x2 * x2;
        rounds++;
    }
    (void)x2;
    double vals[6] = {U_e, traw[0], traw[1], traw[2], a, b};
    logdraw(s, ORC_DRAW_REFLECT, j, l, rounds, vals, 6);
    z = sqrt((-2.0 * log(z)) / z);
    double v1 = x1 * z * sqrt(kbt / mass_fl);
    return vadd(vscale(ut, v1), un);
}

/* cpp/fluid_colloid_collision.cpp:26-49 */
void orc_fluid_colloid_collision(orc_state *s) {
    v3 rr, rs, u, om, vc;
    double lim = pow(s->p.sigma, 2) * 0.25, dt = s->dt;
    memcpy(s->dump, s->vel_fl, sizeof(v3) * (size_t)s->nf);
    for (int j = 0; j < s->nc; j++) {
        vc = om = V(0, 0, 0);
        for (int i = 0; i < s->no_neigh[j]; i++) {
            int l = s->neigh_list[s->neigh_start[j] + i];
            rr = vimg(vsub(s->pos_c[j], s->pos_fl[l]), s->len);
            if (vsum(vmul(rr, rr)) <= lim) {
                s->pos_fl[l] = vmod(vsub(s->pos_fl[l], vscale(vscale(s->vel_fl[l], dt), 0.5)), s->len);
                rs = vimg(vsub(s->pos_fl[l], s->pos_c[j]), s->len);
                u = stochastic_reflection(s, s->pos_fl[l], rs, j, l);
                s->vel_fl[l] = vadd(vadd(u, s->vel_c[j]), vcross(s->omega[j], rs));
                vc = vadd(vc, vsub(s->dump[l], s->vel_fl[l]));
                om = vadd(om, vcross(rs, vsub(s->dump[l], s->vel_fl[l])));
                s->pos_fl[l] = vmod(vadd(s->pos_fl[l], vscale(vscale(s->vel_fl[l], dt), 0.5)), s->len);
            }
        }
        s->vel_c[j] = vadd(s->vel_c[j], vdivs(vscale(vc, s->p.mass_fl), s->p.mass_colloid));
        s->omega[j] = vadd(s->omega[j], vdivs(vscale(om, s->p.mass_fl), s->I_colloid));
    }
}

/* ------------------------------------------------------------------ driver */
/* cpp/active_brownian.cpp:14-21 */
void orc_startup(orc_state *s) {
    orc_initialize_colloid(s);
    orc_initialize_fluid(s);
    orc_create_box(s);
    orc_neighbour_list_mpcd(s);
    orc_neighbour_list_md(s);
    orc_compute_force_md(s);
    orc_tumble(s);
}

/* cpp/active_brownian.cpp:28-37 (updown_velocity, :33, changes no state) */
void orc_substep(orc_state *s) {
    orc_copy_old_force(s);
    orc_update_pos_md(s);
    orc_neighbour_list_md(s);
    orc_update_pos_mpcd(s);
    orc_neighbour_list_mpcd(s);
    orc_fluid_colloid_collision(s);
    orc_update_activity_direction(s);
    orc_compute_force_md(s);
    orc_update_velocity_colloid(s);
}

/* cpp/active_brownian.cpp:25-38 */
void orc_step(orc_state *s) {
    orc_rotation_mpcd(s);
    orc_run(s);
    for (int l = 1; l <= s->p.n_sub; l++) orc_substep(s);
}

/* ------------------------------------------------------------------- maps */
int orc_nbox(const orc_state *s) { return s->nbox; }
void orc_get_box_neigh(const orc_state *s, int *out) {
    for (int c = 0; c < s->cells; c++)
        for (int k = 0; k < s->nbox; k++) out[(size_t)c * s->nbox + k] = s->box_neigh[(size_t)c * 343 + k];
}
void orc_get_fluid_no(const orc_state *s, int *out) { memcpy(out, s->fluid_no, sizeof(int) * (size_t)s->cells); }
void orc_get_cell_no(const orc_state *s, int *out) { memcpy(out, s->cell_no, sizeof(int) * (size_t)s->cells); }
int orc_get_cell_lists(const orc_state *s, int which, int *out) {
    const int *lst = which ? s->cell_list : s->box_list;
    memcpy(out, lst, sizeof(int) * (size_t)s->nf);
    return s->nf;
}
int orc_get_neigh_fl(const orc_state *s, int *counts, int *out) {
    memcpy(counts, s->no_neigh, sizeof(int) * (size_t)s->nc);
    int w = s->neigh_start[s->nc];
    memcpy(out, s->neigh_list, sizeof(int) * (size_t)w);
    return w;
}
int orc_get_md_list(const orc_state *s, int *counts, int *out) {
    int w = 0;
    for (int i = 0; i < s->nc; i++) {
        counts[i] = s->n_neighbour[i];
        for (int j = 0; j < s->n_neighbour[i]; j++) out[w++] = s->neighbour[(size_t)i * s->nc + j];
    }
    return w;
}
int orc_get_run_lists(const orc_state *s, int *counts, int *out) {
    memcpy(counts, s->cnt, sizeof(int) * (size_t)s->nc);
    int w = s->nbr_start[s->nc];
    memcpy(out, s->nbr_list, sizeof(int) * (size_t)w);
    return w;
}

void orc_drawlog_clear(orc_state *s) { s->nlog = 0; }
int orc_drawlog_count(const orc_state *s) { return s->nlog; }
void orc_drawlog_get(const orc_state *s, int *kinds, int *keys, double *vals) {
    for (int i = 0; i < s->nlog; i++) {
        kinds[i] = s->log[i].kind;
        for (int k = 0; k < 3; k++) keys[3 * i + k] = s->log[i].key[k];
        for (int k = 0; k < 8; k++) vals[8 * i + k] = s->log[i].val[k];
    }
}

/* cpp/active_brownian.cpp:39-55, with the momentum taken afresh each call
 * (the reference never resets its accumulator, :13) */
void orc_audit(const orc_state *s, double *out) {
    double ke_c = 0, ang = 0, ke_f = 0;
    v3 mom = V(0, 0, 0);
    for (int i = 0; i < s->nc; i++) {
        ke_c += vsum(vmul(s->vel_c[i], s->vel_c[i]));
        ang += vsum(vmul(s->omega[i], s->omega[i]));
        mom = vadd(mom, vscale(s->vel_c[i], s->p.mass_colloid));
    }
    for (int i = 0; i < s->nf; i++) {
        ke_f += vsum(vmul(s->vel_fl[i], s->vel_fl[i]));
        mom = vadd(mom, vscale(s->vel_fl[i], s->p.mass_fl));
    }
    out[0] = 0.5 * s->p.mass_colloid * ke_c;
    out[1] = 0.5 * s->I_colloid * ang;
    out[2] = s->potential;
    out[3] = 0.5 * ke_f * s->p.mass_fl;
    out[4] = mom.x; out[5] = mom.y; out[6] = mom.z;
}
