// C-ABI of libbrownian: contexts, the reference entry points, the fused
// graph-captured MPCD step, state transfer, injected draws, maps and audit.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/brownian.h"
#include "bw_common.h"
#include "host_init.h"
#include "mlp.h"

namespace bw {
// kernel launchers (kernels.hip)
size_t srd_lds_bytes(int ncell, int cap, bool fp64);
size_t srd_static_lds();
size_t srd_sur_lds();
bool srd_needs_wrap(int L, int bmax);
template <class Real> void launch_depart(const Phys &, const Dev &, const Fluid<Real> &, hipStream_t);
template <class Real>
void launch_srd(const Phys &, const Dev &, const Fluid<Real> &, const Fluid<Real> &, bool, bool, bool, size_t,
                hipStream_t);
template <class Real>
void launch_stream(const Phys &, const Dev &, const Fluid<Real> &, bool, bool, bool, bool, hipStream_t);
template <class Real>
void launch_stream_near(const Phys &, const Dev &, const Fluid<Real> &, bool, hipStream_t);
template <class Real>
void launch_collide(const Phys &, const Dev &, const Fluid<Real> &, int, const int2 *, int, hipStream_t);
template <class Real> void launch_run_count(const Phys &, const Dev &, const Fluid<Real> &, hipStream_t);
template <class Real>
void launch_maps(const Phys &, const Dev &, const Fluid<Real> &, long long, int *, const int *, hipStream_t);
template <class Real> void launch_audit(const Phys &, const Dev &, const Fluid<Real> &, hipStream_t);
template <class Real>
void launch_gather(const Phys &, const Dev &, const Fluid<Real> &, double *, double *, int *, hipStream_t);
void launch_step_begin(const Phys &, const Dev &, hipStream_t, bool clear_lists = false);
void launch_colloid(const Phys &, const Dev &, unsigned, hipStream_t);
void launch_colloid_update(const Phys &, const Dev &, unsigned, const int2 *, int, hipStream_t);
void launch_md_candidates(const Phys &, const Dev &, const int2 *, int, hipStream_t);
void launch_tile_colloids(const Phys &, const Dev &, hipStream_t, bool cleared = false);
void launch_debug_geometry(int n, const double *a, double L, const int *ia, int Li, int b, double *img_out,
                           double *mod_out, int *wrap_out, int *div_out, hipStream_t st);
void launch_copy(const void *, void *, long long, hipStream_t);
void launch_debug_philox(int n, const uint4 *ctr, const uint2 *key, uint4 *out, hipStream_t st);
template <class Real>
void launch_init_fluid(const Phys &ph, const Dev &d, const InitBins &b, int vel_law, long long n_total,
                       long long max_nf, double *pos, double *vel, double *mean, const Fluid<Real> &scratch,
                       const Fluid<Real> &out, int *fill, int *count, hipStream_t st);
void launch_updown(const Phys &ph, const Dev &d, int n_rep, const double *pos, const double *vel, double *sums,
                   int *cnt, int *up_cnt, double *v, double *up, hipStream_t st);
void launch_debug_reflect(const Phys &ph, const Dev &d, int n, long long step, int substep, double L,
                          const double *rf, const double *rs, double *vals6, double *vel3, hipStream_t st);
void launch_features(const Phys &, const Dev &, const int *, hipStream_t);
template <class Real> void launch_mark_phantoms(const Phys &, const Dev &, const Fluid<Real> &, hipStream_t);
template <class Real> void launch_sur_shell_kick(const Phys &, const Dev &, const Fluid<Real> &, hipStream_t);
hipError_t prepare_kernels(Phys &ph, bool fp64);
void launch_sur_sample(const Phys &, const Dev &, const int *, hipStream_t);
}  // namespace bw

using namespace bw;

namespace {

struct ReplicaHost {
    bw_params p;
    int L, nf, nc, cells;
    int B[3], nt[3], n_tiles, first_tile, first_colloid;
    long long first_cell, first_fluid;
    RefRng rng;
    std::vector<double> v0;
    // injected reflections per substep: (colloid, particle, 6 values)
    std::vector<std::vector<int>> inj_c, inj_p;
    std::vector<std::vector<double>> inj_v;
    double inj_shift[3];
    std::vector<double> inj_axes;
    std::vector<double> ra0;  // orientations at the observables' time origin
    long long obs_step0 = 0;
};

struct FluidAlloc {
    void *base = nullptr;
    Fluid<float> f32;
    Fluid<double> f64;
};

}  // namespace

struct bw_ctx {
    std::string err;
    bool fp64 = false;
    bool block = false;  // blocked streaming in the fused step
    int n_rep = 0;
    std::vector<ReplicaHost> reps;
    Phys ph{};
    Dev dv{};
    hipStream_t st = nullptr;
    long long steps_done = 0;      // MPCD steps taken (fused or through rotation_mpcd)
    int cur = 0;                   // fluid buffer holding the state
    int cb = 0;                    // colloid position/cell buffer holding the state
    int substep = 0;
    bool pending_prologue = true;  // next fused step needs the kick/POS prologue
    long long total_slots = 0, cells_total = 0, nf_total = 0;
    int nc_total = 0;
    FluidAlloc fl[2];
    int *counts[2] = {nullptr, nullptr};
    double *cposb[2] = {nullptr, nullptr};
    int *ccellb[2] = {nullptr, nullptr};
    std::vector<void *> allocs;
    int2 *wgt = nullptr;
    int nwg = 0;
    int *colloid_rep = nullptr;
    int *scan_bsum = nullptr;
    size_t srd_lds = 0;
    bool maps_alloc = false;
    double *stage_pos = nullptr, *stage_vel = nullptr;
    // updown_velocity results: per-colloid sums, counts and velocities
    double *ud_sums = nullptr, *ud_vel = nullptr, *ud_up = nullptr;
    int *ud_cnt = nullptr, *ud_upcnt = nullptr;
    int *stage_cell = nullptr;
    bool inj_dirty = false;
    bool moments_ok = false;        // every replica has an even edge and even tile edges
    bw_mlp *mlp = nullptr;          // attached surrogate network
    MlpBuf mbuf;                    // its activations for the step's inference
    hipGraphExec_t graphs[2][2][2] = {};  // [cur][cb][final]
};

namespace {

int fail(bw_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}
#define HIPCHK(ctx, x)                                                                        \
    do {                                                                                      \
        hipError_t e_ = (x);                                                                  \
        if (e_ != hipSuccess) return fail(ctx, BW_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
int dalloc(bw_ctx *c, T **p, size_t n) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) return fail(c, BW_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    // zero-filled before any later work on the context's (non-blocking) stream
    hipMemset(q, 0, std::max<size_t>(n, 1) * sizeof(T));
    hipDeviceSynchronize();
    c->allocs.push_back(q);
    *p = (T *)q;
    return BW_OK;
}

bool divides(int a, int b) { return b > 0 && a % b == 0; }

// slots per tile: mean occupancy + 15% (7 standard deviations of the Poisson
// count at 10 particles per cell and 256 cells) + 64, in multiples of 64
long long tile_slots(long long ncell, int density) {
    long long need = (ncell * density * 23 + 19) / 20 + 64;
    return (need + 63) / 64 * 64;
}

// tile colloid lists reach TILE_REACH + 4 cells beyond the tile box
constexpr int TILE_LIST_REACH_HOST = TILE_REACH + 4;

// a tile's region (its box grown by TILE_REACH per side, at most L) fits the
// per-axis masks
bool region_fits(int L, int bx, int by, int bz) {
    auto ext = [&](int b) { return std::min(b + 2 * TILE_REACH, L); };
    return ext(bx) <= AXIS_MAX && ext(by) <= AXIS_MAX && ext(bz) <= AXIS_MAX;
}

// tile shape: blocks of at least 4 cells per axis dividing L, about 256 cells
void choose_tile(int L, int density, bool fp64, int out[3]) {
    std::vector<int> divs;
    for (int b = 4; b <= std::min(L, 16); b++)
        if (L % b == 0) divs.push_back(b);
    if (divs.empty()) divs.push_back(L);
    int target = fp64 ? 128 : 256;
    double best = 1e30;
    out[0] = out[1] = out[2] = divs.back();
    for (int bx : divs)
        for (int by : divs)
            for (int bz : divs) {
                if (by > bx || bz > by) continue;
                int n = bx * by * bz;
                if (n > 512 || tile_slots(n, density) > MAX_CAP) continue;
                if (!region_fits(L, bx, by, bz)) continue;
                double score = fabs(log((double)n / target)) + 0.01 * (bx - bz);
                if (score < best) {
                    best = score;
                    out[0] = bx; out[1] = by; out[2] = bz;
                }
            }
}

template <class Real>
Fluid<Real> carve(void *base, long long n) {
    Fluid<Real> f;
    char *p = (char *)base;
    f.cell = (int *)p; p += sizeof(int) * n;
    f.pid = (int *)p; p += sizeof(int) * n;
    f.ox = (Real *)p; p += sizeof(Real) * n;
    f.oy = (Real *)p; p += sizeof(Real) * n;
    f.oz = (Real *)p; p += sizeof(Real) * n;
    f.vx = (Real *)p; p += sizeof(Real) * n;
    f.vy = (Real *)p; p += sizeof(Real) * n;
    f.vz = (Real *)p;
    return f;
}

// device view with fluid buffer `buf` and colloid position buffer `cb` current
Dev dev_at(bw_ctx *c, int buf, int cb) {
    Dev d = c->dv;
    d.count = c->counts[buf];
    d.count_next = c->counts[1 - buf];
    d.cpos = c->cposb[cb];
    d.cpos_nxt = c->cposb[1 - cb];
    d.ccell = c->ccellb[cb];
    d.ccell_nxt = c->ccellb[1 - cb];
    return d;
}
Dev dev_for(bw_ctx *c, int buf) { return dev_at(c, buf, c->cb); }

int check_status_code(unsigned s) {
    if (s & (BW_FLAG_DRAW_MISSING | BW_FLAG_DRAW_MISMATCH)) return BW_ERR_DRAWS;
    if (s & (BW_FLAG_DRIFT | BW_FLAG_ZERO_VARIANCE | BW_FLAG_MARGIN)) return BW_ERR_PHYSICS;
    return BW_ERR_CAPACITY;
}

void destroy_graphs(bw_ctx *c);

int sync_injected(bw_ctx *c) {
    if (!c->inj_dirty || c->ph.draws != BW_DRAWS_INJECTED) return BW_OK;
    HIPCHK(c, hipStreamSynchronize(c->st));  // in-flight steps read the old buffers
    int ns = c->ph.n_sub, nct = c->nc_total;
    std::vector<int> off((size_t)ns * nct + 1, 0), pids;
    std::vector<double> vals, shift(3 * c->n_rep), axes((size_t)c->cells_total * 2, NAN);
    // counts per (substep, colloid)
    std::vector<std::vector<int>> per((size_t)ns * nct);
    for (int r = 0; r < c->n_rep; r++) {
        ReplicaHost &R = c->reps[r];
        for (int l = 0; l < ns; l++)
            for (size_t i = 0; i < R.inj_c[l].size(); i++) {
                int g = R.first_colloid + R.inj_c[l][i];
                per[(size_t)l * nct + g].push_back((int)i);
            }
        for (int k = 0; k < 3; k++) shift[3 * r + k] = R.inj_shift[k];
        if (!R.inj_axes.empty())
            memcpy(axes.data() + 2 * R.first_cell, R.inj_axes.data(), sizeof(double) * 2 * R.cells);
    }
    for (int l = 0; l < ns; l++)
        for (int g = 0; g < nct; g++) {
            size_t b = (size_t)l * nct + g;
            off[b] = (int)pids.size();
            int r = 0;
            while (r + 1 < c->n_rep && c->reps[r + 1].first_colloid <= g) r++;
            ReplicaHost &R = c->reps[r];
            for (int i : per[b]) {
                pids.push_back(R.inj_p[l][i]);
                for (int k = 0; k < 6; k++) vals.push_back(R.inj_v[l][6 * i + k]);
            }
        }
    off[(size_t)ns * nct] = (int)pids.size();
    // (re)allocate device copies
    auto put = [&](const void *src, size_t bytes, const void **dst) -> int {
        void *q = nullptr;
        HIPCHK(c, hipMalloc(&q, std::max<size_t>(bytes, 8)));
        if (bytes) HIPCHK(c, hipMemcpy(q, src, bytes, hipMemcpyHostToDevice));
        if (*dst) hipFree((void *)*dst);
        *dst = q;
        return BW_OK;
    };
    int rc;
    if ((rc = put(shift.data(), sizeof(double) * shift.size(), (const void **)&c->dv.inj_shift))) return rc;
    if ((rc = put(axes.data(), sizeof(double) * axes.size(), (const void **)&c->dv.inj_axes))) return rc;
    if ((rc = put(off.data(), sizeof(int) * off.size(), (const void **)&c->dv.inj_off))) return rc;
    if ((rc = put(pids.data(), sizeof(int) * pids.size(), (const void **)&c->dv.inj_pid))) return rc;
    if ((rc = put(vals.data(), sizeof(double) * vals.size(), (const void **)&c->dv.inj_val))) return rc;
    c->inj_dirty = false;
    destroy_graphs(c);  // graphs hold the previous buffer addresses
    return BW_OK;
}

void destroy_graphs(bw_ctx *c) {
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int k = 0; k < 2; k++)
                if (c->graphs[a][b][k]) {
                    hipGraphExecDestroy(c->graphs[a][b][k]);
                    c->graphs[a][b][k] = nullptr;
                }
}

// ---------------------------------------------------------------- the step
template <class Real>
Fluid<Real> &fluid_of(bw_ctx *c, int buf);
template <>
Fluid<float> &fluid_of<float>(bw_ctx *c, int buf) { return c->fl[buf].f32; }
template <>
Fluid<double> &fluid_of<double>(bw_ctx *c, int buf) { return c->fl[buf].f64; }

// kernel families for bw_profile_steps
struct Prof {
    std::vector<hipEvent_t> ev;
    std::vector<int> fam;
    void mark(hipStream_t st, int family) {
        hipEvent_t e;
        hipEventCreate(&e);
        hipEventRecord(e, st);
        ev.push_back(e);
        fam.push_back(family);
    }
};

// One MPCD step (cpp/active_brownian.cpp:25-38) on fluid buffer `cur` and
// colloid buffer `cb`, fused; returns the colloid buffer current afterwards.
// Entry state: the run() colloid kick, old_force copy and first update_pos_md
// of this step done (by the previous step's last colloid update or the
// prologue).  `final`: the last colloid update stops after
// update_velocity_colloid.
template <class Real>
int enqueue_step(bw_ctx *c, int cur, int cb, bool final, Prof *prof) {
    hipStream_t st = c->st;
    const Phys &ph = c->ph;
    const bool col = c->nc_total > 0;
    Fluid<Real> &src = fluid_of<Real>(c, cur), &dst = fluid_of<Real>(c, 1 - cur);
    Dev d0 = dev_at(c, cur, cb);
    if (prof) prof->mark(st, BW_KERNEL_OTHER);
    if (col && ph.moments) {
        // surrogate inputs from the last rotation pass (its shift is still current)
        launch_features(ph, d0, c->colloid_rep, st);
        if (ph.surrogate) {
            mlp_forward_device(c->mlp, c->mbuf, d0.feat, c->nc_total, d0.sur_out, st);
            launch_sur_sample(ph, d0, c->colloid_rep, st);
        }
        if (prof) prof->mark(st, BW_KERNEL_COLLOID);
    }
    launch_step_begin(ph, d0, st, col);  // with colloids it also clears the tile lists
    if (col) {
        launch_md_candidates(ph, d0, c->wgt, c->nwg, st);
        launch_tile_colloids(ph, d0, st, true);
    }
    if (prof) prof->mark(st, BW_KERNEL_COLLOID);
    launch_depart<Real>(ph, d0, src, st);
    if (prof) prof->mark(st, BW_KERNEL_DEPART);
    launch_srd<Real>(ph, d0, src, dst, col, false, c->block, c->srd_lds + (ph.surrogate ? srd_sur_lds() : 0), st);
    if (prof) prof->mark(st, BW_KERNEL_SRD);
    if (col && ph.surrogate) launch_sur_shell_kick<Real>(ph, dev_at(c, 1 - cur, cb), dst, st);
    for (int l = 0; l < ph.n_sub; l++) {
        Dev d1 = dev_at(c, 1 - cur, cb);
        // blocked without colloids: the rotation pass did all the streaming
        if (col && c->block) launch_stream_near<Real>(ph, d1, dst, l == 0, st);  // near chunks only
        else if (col || !c->block) launch_stream<Real>(ph, d1, dst, col && l == 0, true, col, true, st);
        if (prof) prof->mark(st, BW_KERNEL_STREAM);
        if (!col) continue;
        launch_collide<Real>(ph, d1, dst, l, c->wgt, c->nwg, st);
        if (prof) prof->mark(st, BW_KERNEL_COLLIDE);
        unsigned flags = PH_APPLY | PH_ACTIVITY | PH_FORCE | PH_VELOCITY;
        if (l < ph.n_sub - 1) flags |= PH_OLDF | PH_POS | (c->block ? PH_MARGIN : 0u);
        else if (!final) flags |= PH_KICK | PH_OLDF | PH_POS;
        launch_colloid_update(ph, d1, flags, c->wgt, c->nwg, st);
        if (prof) prof->mark(st, BW_KERNEL_COLLOID);
        if (flags & PH_POS) cb ^= 1;
    }
    return cb;
}

// run() colloid kick, old_force copy and update_pos_md ahead of a first step
template <class Real>
void enqueue_prologue(bw_ctx *c) {
    if (c->nc_total == 0) return;
    launch_colloid(c->ph, dev_for(c, c->cur), PH_KICK | PH_OLDF | PH_POS, c->st);
    c->cb ^= 1;
}

template <class Real>
int build_graph(bw_ctx *c, int cur, int cb, int final) {
    hipGraph_t g;
    HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    enqueue_step<Real>(c, cur, cb, final != 0, nullptr);
    HIPCHK(c, hipStreamEndCapture(c->st, &g));
    HIPCHK(c, hipGraphInstantiate(&c->graphs[cur][cb][final], g, nullptr, nullptr, 0));
    hipGraphDestroy(g);
    return BW_OK;
}

// colloid buffer after one step from buffer cb: one flip per update_pos_md
int cb_after(const bw_ctx *c, int cb, bool final) {
    if (c->nc_total == 0) return cb;
    int npos = c->ph.n_sub - 1 + (final ? 0 : 1);
    return cb ^ (npos & 1);
}

template <class Real>
int do_steps(bw_ctx *c, int n) {
    if (n <= 0) return BW_OK;
    int rc = sync_injected(c);
    if (rc) return rc;
    if (c->pending_prologue) enqueue_prologue<Real>(c);
    c->pending_prologue = false;
    for (int s = 0; s < n; s++) {
        int fin = (s == n - 1) ? 1 : 0;
        if (!c->graphs[c->cur][c->cb][fin])
            if ((rc = build_graph<Real>(c, c->cur, c->cb, fin))) return rc;
        HIPCHK(c, hipGraphLaunch(c->graphs[c->cur][c->cb][fin], c->st));
        c->cb = cb_after(c, c->cb, fin != 0);
        c->cur ^= 1;
        c->steps_done++;
    }
    c->pending_prologue = true;
    c->substep = 0;
    return BW_OK;
}

template <class Real>
int do_profile(bw_ctx *c, int n, double *ms, int *launches) {
    int rc = sync_injected(c);
    if (rc) return rc;
    for (int k = 0; k < BW_KERNEL_COUNT; k++) { ms[k] = 0; if (launches) launches[k] = 0; }
    Prof prof;
    for (int s = 0; s < n; s++) {
        if (c->pending_prologue) enqueue_prologue<Real>(c);
        c->pending_prologue = false;
        c->cb = enqueue_step<Real>(c, c->cur, c->cb, s == n - 1, &prof);
        c->cur ^= 1;
        c->steps_done++;
    }
    prof.mark(c->st, BW_KERNEL_OTHER);
    c->pending_prologue = true;
    c->substep = 0;
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (size_t i = 0; i + 1 < prof.ev.size(); i++) {
        float t = 0;
        hipEventElapsedTime(&t, prof.ev[i], prof.ev[i + 1]);
        int fam = prof.fam[i + 1];
        ms[fam] += t;
        if (launches) launches[fam] += 1;
    }
    for (auto e : prof.ev) hipEventDestroy(e);
    return BW_OK;
}

// ------------------------------------------------------------------ create
int create_impl(const bw_params *ps, int n_rep, bw_ctx **out) {
    *out = nullptr;
    if (!ps || n_rep < 1) return BW_ERR_ARG;
    bw_ctx *c = new bw_ctx();
    const bw_params &p0 = ps[0];
    c->fp64 = p0.precision == BW_FP64;
    c->n_rep = n_rep;
    c->reps.resize(n_rep);
    // physics shared by all replicas
    for (int r = 1; r < n_rep; r++) {
        const bw_params &p = ps[r];
        if (p.n_sub != p0.n_sub || p.ndt != p0.ndt || p.kbt != p0.kbt || p.mass_fl != p0.mass_fl ||
            p.mass_colloid != p0.mass_colloid || p.sig_colloid != p0.sig_colloid || p.eps != p0.eps ||
            p.sigma != p0.sigma || p.theta_deg != p0.theta_deg || p.precision != p0.precision ||
            p.draws != p0.draws || p.streaming != p0.streaming || p.fluid_init != p0.fluid_init) {
            delete c;
            return BW_ERR_ARG;
        }
    }
    Phys &ph = c->ph;
    ph.kbt = p0.kbt;
    ph.mass_fl = p0.mass_fl;
    ph.mass_colloid = p0.mass_colloid;
    ph.sig_colloid = p0.sig_colloid;
    ph.eps = p0.eps;
    ph.sigma = p0.sigma;
    ph.I_colloid = 0.1 * p0.mass_colloid * p0.sigma * p0.sigma;  // cpp/active_brownian.cpp:7
    ph.ndt = p0.ndt;
    ph.n_sub = p0.n_sub;
    ph.dt = p0.ndt / (double)p0.n_sub;
    ph.theta = 2 * asin(1) * (p0.theta_deg / 180.0);  // cpp/rotation_mpcd.cpp:19
    ph.ct = cos(ph.theta);
    ph.st = sin(ph.theta);
    ph.ict = 1 - cos(ph.theta);
    ph.r_cutoff = pow(2, 1.0 / 6.0) * p0.sig_colloid;  // cpp/update.cpp:3-6
    ph.fc = 4.0 * p0.eps *
            (12.0 * (pow(p0.sig_colloid, 12) / pow(ph.r_cutoff, 13)) - 6.0 * (pow(p0.sig_colloid, 6) / pow(ph.r_cutoff, 7)));
    ph.ufc = 4.0 * p0.eps * (pow(p0.sig_colloid / ph.r_cutoff, 12) - pow(p0.sig_colloid / ph.r_cutoff, 6)) +
             ph.fc * ph.r_cutoff;
    ph.sig12 = pow(p0.sig_colloid, 12);
    ph.sig6 = pow(p0.sig_colloid, 6);
    ph.lim_coll = pow(p0.sigma, 2) * 0.25;              // cpp/fluid_colloid_collision.cpp:35
    ph.lim_shell = pow(p0.sigma * 0.5 + 0.5, 2);        // cpp/run.cpp:18
    ph.lim_md = pow(3.0 * p0.sig_colloid, 2);           // cpp/neighbour_list.cpp:28,34
    ph.lim_md_skin = pow(3.0 * p0.sig_colloid + MD_SKIN, 2);
    ph.vscale_thermal = sqrt(p0.kbt / p0.mass_fl);
    ph.draws = p0.draws;
    // blocked streaming: near radius sigma/2 + margin + travel (<= vcap_ndt),
    // run() shell radius sigma/2 + 0.5; 1e-3 covers FP32 offset rounding
    ph.margin = STREAM_MARGIN;
    ph.vcap_ndt = 0.9;  // < 1: a far particle moves less than a cell per substep
    ph.near_base = 0.5 * p0.sigma + ph.margin + 1e-3;
    ph.near_shell2 = pow(0.5 * p0.sigma + 0.5 + 1e-3, 2);
    ph.near_reach = (int)floor(ph.near_base + ph.vcap_ndt) + 1;
    c->block = p0.streaming == BW_STREAM_BLOCKED && ph.near_reach <= TILE_LIST_REACH_HOST - 1;
    ph.n_rep = n_rep;
    // replicas
    int cap = 0;
    long long tiles = 0, cells = 0, nf = 0;
    int nc = 0;
    for (int r = 0; r < n_rep; r++) {
        const bw_params &p = ps[r];
        ReplicaHost &R = c->reps[r];
        R.p = p;
        // particle ids and counts are int: the fluid of one replica (and of
        // all together) must stay below 2^31
        if (p.L < 7 || p.L >= (1 << CELL_BITS) || p.density < 1 || p.n_colloid < 0 || p.n_sub < 1 ||
            (long long)p.L * p.L * p.L * p.density > INT_MAX || nf + (long long)p.L * p.L * p.L * p.density > INT_MAX) {
            delete c;
            return BW_ERR_ARG;
        }
        R.L = p.L;
        R.cells = p.L * p.L * p.L;
        R.nf = R.cells * p.density;
        R.nc = p.n_colloid;
        int B[3] = {p.tile[0], p.tile[1], p.tile[2]};
        if (B[0] <= 0 || B[1] <= 0 || B[2] <= 0) choose_tile(p.L, p.density, c->fp64, B);
        for (int k = 0; k < 3; k++) {
            if (!divides(p.L, B[k]) || (B[k] < 4 && B[k] != p.L)) {
                delete c;
                return BW_ERR_ARG;
            }
            R.B[k] = B[k];
            R.nt[k] = p.L / B[k];
        }
        if (B[0] * B[1] * B[2] > 512 || !region_fits(p.L, B[0], B[1], B[2])) {
            delete c;
            return BW_ERR_ARG;
        }
        int need = (int)tile_slots(B[0] * B[1] * B[2], p.density);
        cap = std::max(cap, need);
        R.n_tiles = R.nt[0] * R.nt[1] * R.nt[2];
        R.first_tile = (int)tiles;
        R.first_colloid = nc;
        R.first_cell = cells;
        R.first_fluid = nf;
        tiles += R.n_tiles;
        cells += R.cells;
        nf += R.nf;
        nc += R.nc;
        R.rng.reset(p.seed);
        R.v0.assign(R.nc, p.v0);
        R.inj_c.assign(p.n_sub, {});
        R.inj_p.assign(p.n_sub, {});
        R.inj_v.assign(p.n_sub, {});
        R.inj_shift[0] = R.inj_shift[1] = R.inj_shift[2] = NAN;
    }
    if (cap > MAX_CAP) {
        delete c;
        return BW_ERR_ARG;
    }
    ph.cap = cap;
    // the SRD workgroup: the fewest slots (threads x slots per thread) that
    // hold the tile's capacity, 768 / 1024 / 1536 / 2048 / 3072 / 4096
    {
        static const int shape[6][2] = {{256, 3}, {256, 4}, {512, 3}, {512, 4}, {1024, 3}, {1024, 4}};
        int k = 0;
        while (k < 5 && cap > shape[k][0] * shape[k][1]) k++;
        ph.srd_threads = shape[k][0];
        ph.srd_pt = shape[k][1];
    }
    ph.near_wrap = 0;
    for (auto &R : c->reps)
        if (srd_needs_wrap(R.L, std::max(R.B[0], std::max(R.B[1], R.B[2])))) ph.near_wrap = 1;
    ph.n_tiles = (int)tiles;
    ph.n_colloid_total = nc;
    c->total_slots = tiles * cap;
    c->cells_total = cells;
    c->nf_total = nf;
    c->nc_total = nc;
    int maxcell = 0;
    for (auto &R : c->reps) maxcell = std::max(maxcell, R.B[0] * R.B[1] * R.B[2]);
    c->moments_ok = true;  // coarse cells of 2x2x2 need even edges everywhere
    for (auto &R : c->reps)
        if (R.L % 2 || R.B[0] % 2 || R.B[1] % 2 || R.B[2] % 2) c->moments_ok = false;
    if (c->moments_ok)
        for (auto &R : c->reps)
            if (R.first_cell % 8) c->moments_ok = false;
    c->srd_lds = srd_lds_bytes(maxcell, cap, c->fp64);
    if (c->srd_lds > 160 * 1024 - srd_static_lds()) {
        delete c;
        return BW_ERR_ARG;
    }
    int rc;
    hipError_t e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return BW_ERR_HIP;
    }
    if ((e = prepare_kernels(c->ph, c->fp64)) != hipSuccess) {
        c->err = std::string("kernel occupancy query: ") + hipGetErrorString(e);
        *out = c;
        return BW_ERR_HIP;
    }
    // device memory
    size_t rb = c->fp64 ? 8 : 4;
    for (int b = 0; b < 2; b++) {
        size_t bytes = (size_t)c->total_slots * (8 + 6 * rb);
        e = hipMalloc(&c->fl[b].base, bytes);
        if (e != hipSuccess) { c->err = "hipMalloc fluid"; *out = c; return BW_ERR_HIP; }
        hipMemset(c->fl[b].base, 0, bytes);
        c->allocs.push_back(c->fl[b].base);
        c->fl[b].f32 = carve<float>(c->fl[b].base, c->total_slots);
        c->fl[b].f64 = carve<double>(c->fl[b].base, c->total_slots);
        if ((rc = dalloc(c, &c->counts[b], tiles))) { *out = c; return rc; }
    }
    Dev &d = c->dv;
    if ((rc = dalloc(c, &d.near_cnt, tiles)) || (rc = dalloc(c, &d.chunk_cnt, 1)) ||
        (rc = dalloc(c, &d.chunk_tile, (size_t)tiles * ((cap + 63) / 64)))) {
        *out = c;
        return rc;
    }
    std::vector<Rep> hrep(n_rep);
    std::vector<int> trep(tiles), crep(std::max(nc, 1));
    for (int r = 0; r < n_rep; r++) {
        ReplicaHost &R = c->reps[r];
        Rep &q = hrep[r];
        q.L = R.L; q.nf = R.nf; q.nc = R.nc; q.first_colloid = R.first_colloid;
        for (int k = 0; k < 3; k++) { q.B[k] = R.B[k]; q.nt[k] = R.nt[k]; }
        q.first_tile = R.first_tile; q.n_tiles = R.n_tiles;
        q.first_cell = R.first_cell; q.first_fluid = R.first_fluid;
        q.first_coarse = R.first_cell / 8;
        unsigned long long s = (unsigned long long)(long long)R.p.seed;
        // key from the seed only, so a replica draws the same numbers alone or in an ensemble
        q.seed_lo = (uint32_t)(s * 0x9E3779B97F4A7C15ull >> 32) ^ 0x5bd1e995u;
        q.seed_hi = (uint32_t)s ^ 0x2545F491u;
        for (int t = 0; t < R.n_tiles; t++) trep[R.first_tile + t] = r;
        for (int j = 0; j < R.nc; j++) crep[R.first_colloid + j] = r;
    }
    Rep *drep;
    int *dtrep;
    if ((rc = dalloc(c, &drep, n_rep)) || (rc = dalloc(c, &dtrep, tiles)) || (rc = dalloc(c, &c->colloid_rep, crep.size()))) {
        *out = c;
        return rc;
    }
    hipMemcpy(drep, hrep.data(), sizeof(Rep) * n_rep, hipMemcpyHostToDevice);
    hipMemcpy(dtrep, trep.data(), sizeof(int) * tiles, hipMemcpyHostToDevice);
    hipMemcpy(c->colloid_rep, crep.data(), sizeof(int) * crep.size(), hipMemcpyHostToDevice);
    d.rep = drep;
    d.tile_rep = dtrep;
    size_t NC = std::max(nc, 1);
    if ((rc = dalloc(c, &d.dep_off, (size_t)tiles * 28)) || (rc = dalloc(c, &d.dep_idx, (size_t)c->total_slots)) ||
        (rc = dalloc(c, &c->cposb[0], 3 * NC)) || (rc = dalloc(c, &c->cposb[1], 3 * NC)) ||
        (rc = dalloc(c, &c->ccellb[0], 3 * NC)) || (rc = dalloc(c, &c->ccellb[1], 3 * NC)) ||
        (rc = dalloc(c, &d.md_stale, n_rep)) || (rc = dalloc(c, &d.tcl_stale, n_rep)) ||
        (rc = dalloc(c, &d.ovf_cnt, n_rep)) || (rc = dalloc(c, &d.ovf_list, (size_t)n_rep * OVF_CAP)) ||
        (rc = dalloc(c, &d.ev_total, n_rep)) ||
        (rc = dalloc(c, &d.cmom, c->moments_ok ? (size_t)cells / 8 * 5 : 1)) || (rc = dalloc(c, &d.mom_valid, 1)) ||
        (rc = dalloc(c, &d.cmode, NC)) || (rc = dalloc(c, &d.cmode_prev, NC)) || (rc = dalloc(c, &d.csur, 6 * NC)) || (rc = dalloc(c, &d.cimp, 6 * NC)) ||
        (rc = dalloc(c, &d.sshell, NC)) || (rc = dalloc(c, &d.sacc, SUR_ACC * NC)) ||
        (rc = dalloc(c, &d.feat, NC * N_FEAT)) || (rc = dalloc(c, &d.sur_out, NC * 12)) ||
        (rc = dalloc(c, &d.cvel, 3 * NC)) || (rc = dalloc(c, &d.comega, 3 * NC)) ||
        (rc = dalloc(c, &d.cf, 3 * NC)) || (rc = dalloc(c, &d.coldf, 3 * NC)) || (rc = dalloc(c, &d.cra, 3 * NC)) ||
        (rc = dalloc(c, &d.cdisp, 3 * NC)) || (rc = dalloc(c, &d.cpos_run, 3 * NC)) ||
        (rc = dalloc(c, &d.cdv, 3 * NC)) || (rc = dalloc(c, &d.cdw, 3 * NC)) || (rc = dalloc(c, &d.cpos_md, 3 * NC)) ||
        (rc = dalloc(c, &d.cv0, NC)) || (rc = dalloc(c, &d.crep, NC)) ||
        (rc = dalloc(c, &d.ccell_run, 3 * NC)) || (rc = dalloc(c, &d.ccell_md, 3 * NC)) ||
        (rc = dalloc(c, &d.tcl_cnt, tiles)) || (rc = dalloc(c, &d.tcl_list, (size_t)tiles * TCL_CAP)) ||
        (rc = dalloc(c, &d.tsl_cnt, tiles)) || (rc = dalloc(c, &d.tsl_list, (size_t)tiles * TCL_CAP)) ||
        (rc = dalloc(c, &d.ccnt, NC)) || (rc = dalloc(c, &d.md_cnt, NC)) ||
        (rc = dalloc(c, &d.md_list, NC * MD_CAP)) || (rc = dalloc(c, &d.md_f, NC * MD_CAP * 3)) || (rc = dalloc(c, &d.mdc_cnt, NC)) ||
        (rc = dalloc(c, &d.mdc_list, NC * MDC_CAP)) || (rc = dalloc(c, &d.cpot, n_rep)) ||
        (rc = dalloc(c, &d.ev_cnt, NC)) || (rc = dalloc(c, &d.ev_slot, NC * EV_CAP)) ||
        (rc = dalloc(c, &d.chain_cnt, n_rep)) || (rc = dalloc(c, &d.chain_slot, (size_t)n_rep * CHAIN_CAP)) ||
        (rc = dalloc(c, &d.cont_cnt, n_rep)) || (rc = dalloc(c, &d.cont_slot, (size_t)n_rep * CONT_CAP)) ||
        (rc = dalloc(c, &d.cont_from, (size_t)n_rep * CONT_CAP)) ||
        (rc = dalloc(c, &d.cont_dump, (size_t)n_rep * CONT_CAP * 3)) ||
        (rc = dalloc(c, &d.status, 1)) || (rc = dalloc(c, &d.step, 1)) ||
        (rc = dalloc(c, &d.shift, 3 * n_rep)) || (rc = dalloc(c, &d.audit_partial, (size_t)tiles * 8))) {
        *out = c;
        return rc;
    }
    hipMemcpy(d.crep, crep.data(), sizeof(int) * crep.size(), hipMemcpyHostToDevice);
    {
        std::vector<double> v0(NC, 0.0);
        for (auto &R : c->reps)
            for (int j = 0; j < R.nc; j++) v0[R.first_colloid + j] = R.p.v0;
        hipMemcpy(d.cv0, v0.data(), sizeof(double) * NC, hipMemcpyHostToDevice);
    }
    // workgroup table: groups of 4 colloids, never spanning replicas
    std::vector<int2> wgt;
    for (auto &R : c->reps)
        for (int j = 0; j < R.nc; j += COLLIDE_THREADS / 64) wgt.push_back(make_int2((int)(&R - &c->reps[0]), R.first_colloid + j));
    c->nwg = (int)wgt.size();
    if ((rc = dalloc(c, &c->wgt, std::max<size_t>(wgt.size(), 1)))) { *out = c; return rc; }
    if (!wgt.empty()) hipMemcpy(c->wgt, wgt.data(), sizeof(int2) * wgt.size(), hipMemcpyHostToDevice);
    // injected-draw buffers exist (empty) from the start
    c->inj_dirty = ph.draws == BW_DRAWS_INJECTED;
    if ((rc = sync_injected(c))) { *out = c; return rc; }
    hipDeviceSynchronize();
    *out = c;
    return BW_OK;
}

int alloc_maps(bw_ctx *c) {
    if (c->maps_alloc) return BW_OK;
    Dev &d = c->dv;
    int rc;
    int maxd = 1;
    for (auto &R : c->reps) maxd = std::max(maxd, R.p.density);
    d.neigh_cap = 342 * maxd * 4;
    size_t NC = std::max(c->nc_total, 1);
    if ((rc = dalloc(c, &d.map_count, c->cells_total + 1)) || (rc = dalloc(c, &d.map_start, c->cells_total + 1)) ||
        (rc = dalloc(c, &d.map_fill, c->cells_total + 1)) || (rc = dalloc(c, &d.map_list, c->nf_total + 1)) ||
        (rc = dalloc(c, &d.neigh_cnt, NC)) || (rc = dalloc(c, &d.neigh_list, NC * (size_t)d.neigh_cap)) ||
        (rc = dalloc(c, &d.shifted_cell, c->nf_total + 1)) ||
        (rc = dalloc(c, &c->scan_bsum, (c->cells_total + 1023) / 1024 + 1)))
        return rc;
    c->maps_alloc = true;
    destroy_graphs(c);  // Dev copies captured in graphs are stale
    return BW_OK;
}

ReplicaHost *rep_of(bw_ctx *c, int r) { return (r >= 0 && r < c->n_rep) ? &c->reps[r] : nullptr; }

InitParams init_params(const ReplicaHost &R) {
    InitParams ip;
    ip.L = R.L; ip.density = R.p.density; ip.n_colloid = R.nc;
    ip.kbt = R.p.kbt; ip.mass_fl = R.p.mass_fl; ip.mass_colloid = R.p.mass_colloid;
    ip.sig_colloid = R.p.sig_colloid; ip.sigma = R.p.sigma;
    ip.vel_law = R.p.vel_law; ip.orient_law = R.p.orient_law;
    return ip;
}

// place particles of replica r into tiles of the unshifted grid
template <class Real>
int upload_fluid_impl(bw_ctx *c, int r, const double *pos, const double *vel) {
    ReplicaHost &R = c->reps[r];
    const int L = R.L, cap = c->ph.cap;
    const long long nf = R.nf;
    std::vector<int> cell(nf), tile(nf);
    std::vector<Real> o(3 * nf);
    for (long long i = 0; i < nf; i++) {
        int ic[3];
        for (int k = 0; k < 3; k++) {
            double x = pos[3 * i + k];
            double fl = floor(x);
            long long q = (long long)fl;
            double off = x - fl;
            Real ro = (Real)off;
            if (ro >= Real(1)) { ro = Real(0); q += 1; }
            q %= L;
            if (q < 0) q += L;
            ic[k] = (int)q;
            o[3 * i + k] = ro;
        }
        cell[i] = ic[0] | (ic[1] << CELL_BITS) | (ic[2] << (2 * CELL_BITS));
        tile[i] = (ic[0] / R.B[0]) + R.nt[0] * ((ic[1] / R.B[1]) + R.nt[1] * (ic[2] / R.B[2]));
    }
    std::vector<int> cnt(R.n_tiles, 0);
    for (long long i = 0; i < nf; i++) cnt[tile[i]]++;
    for (int t = 0; t < R.n_tiles; t++)
        if (cnt[t] > cap) return fail(c, BW_ERR_CAPACITY, "tile capacity exceeded on upload");
    size_t ns = (size_t)R.n_tiles * cap;
    std::vector<int> hcell(ns, 0), hpid(ns, 0);
    std::vector<Real> hox(ns, 0), hoy(ns, 0), hoz(ns, 0), hvx(ns, 0), hvy(ns, 0), hvz(ns, 0);
    std::vector<int> fill(R.n_tiles, 0);
    for (long long i = 0; i < nf; i++) {
        int t = tile[i];
        size_t s = (size_t)t * cap + fill[t]++;
        hcell[s] = cell[i];
        hpid[s] = (int)i;
        hox[s] = o[3 * i]; hoy[s] = o[3 * i + 1]; hoz[s] = o[3 * i + 2];
        hvx[s] = (Real)vel[3 * i]; hvy[s] = (Real)vel[3 * i + 1]; hvz[s] = (Real)vel[3 * i + 2];
    }
    Fluid<Real> &f = fluid_of<Real>(c, c->cur);
    size_t off = (size_t)R.first_tile * cap;
    HIPCHK(c, hipMemcpy(f.cell + off, hcell.data(), sizeof(int) * ns, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(f.pid + off, hpid.data(), sizeof(int) * ns, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(f.ox + off, hox.data(), sizeof(Real) * ns, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(f.oy + off, hoy.data(), sizeof(Real) * ns, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(f.oz + off, hoz.data(), sizeof(Real) * ns, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(f.vx + off, hvx.data(), sizeof(Real) * ns, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(f.vy + off, hvy.data(), sizeof(Real) * ns, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(f.vz + off, hvz.data(), sizeof(Real) * ns, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->counts[c->cur] + R.first_tile, cnt.data(), sizeof(int) * R.n_tiles, hipMemcpyHostToDevice));
    return BW_OK;
}

template <class Real>
int download_fluid_impl(bw_ctx *c, int r, double *pos, double *vel, int *cell) {
    ReplicaHost &R = c->reps[r];
    if (!c->stage_pos) {
        int rc;
        if ((rc = dalloc(c, &c->stage_pos, 3 * (size_t)c->nf_total)) ||
            (rc = dalloc(c, &c->stage_vel, 3 * (size_t)c->nf_total)) || (rc = dalloc(c, &c->stage_cell, c->nf_total)))
            return rc;
    }
    launch_gather<Real>(c->ph, dev_for(c, c->cur), fluid_of<Real>(c, c->cur), c->stage_pos, c->stage_vel,
                        c->stage_cell, c->st);
    HIPCHK(c, hipStreamSynchronize(c->st));
    size_t off = (size_t)R.first_fluid;
    if (pos) HIPCHK(c, hipMemcpy(pos, c->stage_pos + 3 * off, sizeof(double) * 3 * R.nf, hipMemcpyDeviceToHost));
    if (vel) HIPCHK(c, hipMemcpy(vel, c->stage_vel + 3 * off, sizeof(double) * 3 * R.nf, hipMemcpyDeviceToHost));
    if (cell) HIPCHK(c, hipMemcpy(cell, c->stage_cell + off, sizeof(int) * R.nf, hipMemcpyDeviceToHost));
    return BW_OK;
}

int set_colloid_cells(bw_ctx *c, int r) {
    ReplicaHost &R = c->reps[r];
    if (R.nc == 0) return BW_OK;
    std::vector<double> pos(3 * R.nc);
    HIPCHK(c, hipMemcpy(pos.data(), c->cposb[c->cb] + 3 * R.first_colloid, sizeof(double) * pos.size(), hipMemcpyDeviceToHost));
    std::vector<int> cc(3 * R.nc);
    for (int i = 0; i < 3 * R.nc; i++) {
        int q = (int)pos[i];
        if (q >= R.L) q -= R.L;
        if (q < 0) q += R.L;
        cc[i] = q;
    }
    size_t o = 3 * (size_t)R.first_colloid;
    for (int b = 0; b < 2; b++)
        HIPCHK(c, hipMemcpy(c->ccellb[b] + o, cc.data(), sizeof(int) * cc.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->cposb[1 - c->cb] + o, pos.data(), sizeof(double) * pos.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->dv.ccell_run + o, cc.data(), sizeof(int) * cc.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->dv.ccell_md + o, cc.data(), sizeof(int) * cc.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->dv.cpos_run + o, pos.data(), sizeof(double) * pos.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->dv.cpos_md + o, pos.data(), sizeof(double) * pos.size(), hipMemcpyHostToDevice));
    return BW_OK;
}

}  // namespace

// ===================================================================== C-ABI
extern "C" {

BW_API void bw_default_params(bw_params *p) {
    memset(p, 0, sizeof(*p));
    // cpp/active_brownian.cpp:3-7
    p->L = 30;
    p->density = 10;
    p->n_colloid = 10;
    p->n_sub = 10;
    p->ndt = 0.1;
    p->kbt = 1.0;
    p->mass_fl = 1.0;
    p->mass_colloid = 654.1;
    p->sig_colloid = 5.0;
    p->eps = 1.0;
    p->sigma = 0.80 * 5.0;
    p->theta_deg = 130.0;
    p->v0 = 0.0;
    p->seed = 77777;
    p->precision = BW_FP64;
    p->vel_law = BW_VEL_REFERENCE;
    p->orient_law = BW_ORIENT_REFERENCE;
    p->draws = BW_DRAWS_DEVICE;
}

BW_API int bw_ctx_create(const bw_params *p, bw_ctx **out) { return create_impl(p, 1, out); }
BW_API int bw_ctx_create_ensemble(const bw_params *p, int n, bw_ctx **out) { return create_impl(p, n, out); }

BW_API int bw_ctx_destroy(bw_ctx *c) {
    if (!c) return BW_OK;
    if (c->st) hipStreamSynchronize(c->st);
    destroy_graphs(c);
    for (void *q : c->allocs) hipFree(q);
    mlp_buf_free(c->mbuf);
    const void *inj[] = {c->dv.inj_shift, c->dv.inj_axes, c->dv.inj_off, c->dv.inj_pid, c->dv.inj_val,
                         c->dv.sur_lo, c->dv.sur_hi};
    for (const void *q : inj)
        if (q) hipFree((void *)q);
    if (c->st) hipStreamDestroy(c->st);
    delete c;
    return BW_OK;
}

BW_API const char *bw_last_error(bw_ctx *c) { return c ? c->err.c_str() : "no context"; }
BW_API int bw_n_replica(bw_ctx *c) { return c ? c->n_rep : 0; }
BW_API int bw_replica_info(bw_ctx *c, int r, int *o) {
    ReplicaHost *R = rep_of(c, r);
    if (!R) return BW_ERR_ARG;
    o[0] = R->nf; o[1] = R->nc; o[2] = R->L; o[3] = R->B[0]; o[4] = R->B[1]; o[5] = R->B[2]; o[6] = c->ph.cap;
    return BW_OK;
}
BW_API int bw_sync(bw_ctx *c) {
    HIPCHK(c, hipStreamSynchronize(c->st));
    return BW_OK;
}
BW_API int bw_status(bw_ctx *c, unsigned *flags) {
    unsigned s = 0;
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, hipMemcpy(&s, c->dv.status, sizeof(unsigned), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemsetAsync(c->dv.status, 0, sizeof(unsigned), c->st));
    if (flags) *flags = s;
    if (s) {
        char buf[96];
        snprintf(buf, sizeof buf, "device status 0x%x", s);
        c->err = buf;
    }
    return s ? check_status_code(s) : BW_OK;
}

// ---------------------------------------------------------- entry points
BW_API int initialize(bw_ctx *c) {
    if (!c) return BW_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (auto &R : c->reps) R.rng.reset(R.p.seed);
    // ordered on the context's stream before any later launch
    for (int b = 0; b < 2; b++) HIPCHK(c, hipMemsetAsync(c->counts[b], 0, sizeof(int) * c->ph.n_tiles, c->st));
    HIPCHK(c, hipMemsetAsync(c->dv.step, 0, sizeof(long long), c->st));
    HIPCHK(c, hipMemsetAsync(c->dv.status, 0, sizeof(unsigned), c->st));
    size_t NC = std::max(c->nc_total, 1);
    double *arrs[] = {c->cposb[0], c->cposb[1], c->dv.cvel, c->dv.comega, c->dv.cf, c->dv.coldf,
                      c->dv.cra, c->dv.cdisp, c->dv.cdv, c->dv.cdw};
    for (double *a : arrs) HIPCHK(c, hipMemsetAsync(a, 0, sizeof(double) * 3 * NC, c->st));
    for (int b = 0; b < 2; b++) HIPCHK(c, hipMemsetAsync(c->ccellb[b], 0, sizeof(int) * 3 * NC, c->st));
    HIPCHK(c, hipMemsetAsync(c->dv.ev_cnt, 0, sizeof(int) * NC, c->st));
    int *per_rep[] = {c->dv.chain_cnt, c->dv.cont_cnt, c->dv.ovf_cnt, c->dv.ev_total, c->dv.md_stale, c->dv.tcl_stale};
    for (int *a : per_rep) HIPCHK(c, hipMemsetAsync(a, 0, sizeof(int) * c->n_rep, c->st));
    HIPCHK(c, hipMemsetAsync(c->dv.cimp, 0, sizeof(double) * 6 * NC, c->st));
    HIPCHK(c, hipMemsetAsync(c->dv.cmode, 0, sizeof(int) * NC, c->st));
    HIPCHK(c, hipMemsetAsync(c->dv.cmode_prev, 0, sizeof(int) * NC, c->st));
    HIPCHK(c, hipMemsetAsync(c->dv.sshell, 0, sizeof(int) * NC, c->st));
    HIPCHK(c, hipMemsetAsync(c->dv.sacc, 0, sizeof(unsigned long long) * SUR_ACC * NC, c->st));
    HIPCHK(c, hipMemsetAsync(c->dv.mom_valid, 0, sizeof(int), c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->cb = 0;
    c->pending_prologue = true;
    c->substep = 0;
    c->steps_done = 0;
    return BW_OK;
}

BW_API int initialize_colloid(bw_ctx *c) {
    if (!c) return BW_ERR_ARG;
    for (int r = 0; r < c->n_rep; r++) {
        ReplicaHost &R = c->reps[r];
        if (!R.nc) continue;
        std::vector<double> pos(3 * R.nc), vel(3 * R.nc), om(3 * R.nc);
        init_colloids(init_params(R), R.rng, pos.data(), vel.data(), om.data());
        int rc = bw_upload_colloid(c, r, pos.data(), vel.data(), om.data(), nullptr, nullptr, nullptr);
        if (rc) return rc;
    }
    return BW_OK;
}

}  // extern "C"

namespace {
// initialize_fluid on the GPU (BW_INIT_DEVICE): coarse colloid bins built on
// the host, then placement, mean removal and tile binning on the device
template <class Real>
int init_fluid_device(bw_ctx *c) {
    int rc;
    if (!c->stage_pos && ((rc = dalloc(c, &c->stage_pos, 3 * (size_t)c->nf_total)) ||
                          (rc = dalloc(c, &c->stage_vel, 3 * (size_t)c->nf_total)) ||
                          (rc = dalloc(c, &c->stage_cell, c->nf_total))))
        return rc;
    std::vector<int> G(c->n_rep), start(1, 0), list;
    std::vector<long long> first(c->n_rep);
    std::vector<double> cpos(3 * (size_t)std::max(c->nc_total, 1));
    if (c->nc_total)
        HIPCHK(c, hipMemcpy(cpos.data(), c->cposb[c->cb], sizeof(double) * 3 * c->nc_total, hipMemcpyDeviceToHost));
    for (int r = 0; r < c->n_rep; r++) {
        const ReplicaHost &R = c->reps[r];
        const double rad = 0.5 * R.p.sigma;
        int g = (int)floor(R.L / std::max(rad, 1.0));
        g = std::min(g, 32);
        if (g < 3) g = 1;  // edge >= sigma/2; one bin when fewer than 3 fit
        G[r] = g;
        first[r] = (long long)start.size() - 1;
        const double h = (double)R.L / g;
        std::vector<std::vector<int>> bins((size_t)g * g * g);
        for (int j = 0; j < R.nc; j++) {
            const int gj = R.first_colloid + j;
            int b3[3];
            for (int k = 0; k < 3; k++) b3[k] = std::min((int)(cpos[3 * gj + k] / h), g - 1);
            bins[(size_t)b3[0] + g * ((size_t)b3[1] + (size_t)g * b3[2])].push_back(gj);
        }
        for (auto &bl : bins) {
            list.insert(list.end(), bl.begin(), bl.end());
            start.push_back((int)list.size());
        }
    }
    // temporaries of the start-up, in one allocation freed before returning
    const size_t n_list = std::max<size_t>(list.size(), 1);
    long long max_nf = 1;
    for (const auto &R : c->reps) max_nf = std::max(max_nf, (long long)R.nf);
    const size_t n_mean = 3 * (size_t)c->n_rep * (1 + (size_t)init_mean_chunks(max_nf));  // means + chunk sums
    const size_t bytes = sizeof(double) * n_mean + sizeof(long long) * first.size() +
                         sizeof(int) * (G.size() + start.size() + n_list + (size_t)c->ph.n_tiles) + 64;
    char *tmp = nullptr;
    HIPCHK(c, hipMalloc((void **)&tmp, bytes));
    double *mean = (double *)tmp;
    long long *dfirst = (long long *)(mean + n_mean);
    int *dG = (int *)(dfirst + first.size());
    int *dstart = dG + G.size(), *dlist = dstart + start.size(), *fill = dlist + n_list;
    HIPCHK(c, hipMemsetAsync(fill, 0, sizeof(int) * (size_t)c->ph.n_tiles, c->st));
    HIPCHK(c, hipMemcpy(dG, G.data(), sizeof(int) * G.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(dfirst, first.data(), sizeof(long long) * first.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(dstart, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice));
    if (!list.empty()) HIPCHK(c, hipMemcpy(dlist, list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice));
    InitBins b{dG, dfirst, dstart, dlist, c->cposb[c->cb]};
    Dev d = dev_for(c, c->cur);
    HIPCHK(c, hipMemsetAsync(c->dv.status, 0, sizeof(unsigned), c->st));
    launch_init_fluid<Real>(c->ph, d, b, c->reps[0].p.vel_law, c->nf_total, max_nf, c->stage_pos, c->stage_vel, mean,
                            fluid_of<Real>(c, 1 - c->cur), fluid_of<Real>(c, c->cur), fill, c->counts[c->cur], c->st);
    HIPCHK(c, hipStreamSynchronize(c->st));
    hipFree(tmp);
    unsigned flags = 0;
    HIPCHK(c, hipMemcpy(&flags, c->dv.status, sizeof(unsigned), hipMemcpyDeviceToHost));
    if (flags & BW_FLAG_TILE_OVERFLOW) return fail(c, BW_ERR_CAPACITY, "tile capacity exceeded in the device start-up");
    if (flags) return fail(c, BW_ERR_PHYSICS, "device start-up could not place a particle outside the colloids");
    return BW_OK;
}
}  // namespace

extern "C" {

BW_API int initialize_fluid(bw_ctx *c) {
    if (!c) return BW_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (c->reps[0].p.fluid_init == BW_INIT_DEVICE)
        return c->fp64 ? init_fluid_device<double>(c) : init_fluid_device<float>(c);
    for (int r = 0; r < c->n_rep; r++) {
        ReplicaHost &R = c->reps[r];
        std::vector<double> cpos(3 * std::max(R.nc, 1));
        if (R.nc)
            HIPCHK(c, hipMemcpy(cpos.data(), c->cposb[c->cb] + 3 * R.first_colloid, sizeof(double) * 3 * R.nc,
                                hipMemcpyDeviceToHost));
        std::vector<double> pos(3 * (size_t)R.nf), vel(3 * (size_t)R.nf);
        init_fluid(init_params(R), R.rng, cpos.data(), pos.data(), vel.data());
        int rc = bw_upload_fluid(c, r, pos.data(), vel.data());
        if (rc) return rc;
    }
    return BW_OK;
}

BW_API int create_box(bw_ctx *c) {
    // The 7^3-1 stencil of cpp/neighbour_list.cpp:12-25 is evaluated
    // arithmetically wherever it is needed; there is no table to build.
    return c ? BW_OK : BW_ERR_ARG;
}

BW_API int tumble(bw_ctx *c) {
    if (!c) return BW_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (int r = 0; r < c->n_rep; r++) {
        ReplicaHost &R = c->reps[r];
        if (!R.nc) continue;
        std::vector<double> cpos(3 * R.nc), ra(3 * R.nc);
        HIPCHK(c, hipMemcpy(cpos.data(), c->cposb[c->cb] + 3 * R.first_colloid, sizeof(double) * 3 * R.nc,
                            hipMemcpyDeviceToHost));
        init_orientation(init_params(R), R.rng, cpos.data(), ra.data());
        HIPCHK(c, hipMemcpy(c->dv.cra + 3 * R.first_colloid, ra.data(), sizeof(double) * 3 * R.nc,
                            hipMemcpyHostToDevice));
    }
    return BW_OK;
}

BW_API int neighbour_list_mpcd(bw_ctx *c) {
    if (!c) return BW_ERR_ARG;
    int rc = alloc_maps(c);
    if (rc) return rc;
    if (c->fp64)
        launch_maps<double>(c->ph, dev_for(c, c->cur), c->fl[c->cur].f64, c->cells_total, c->scan_bsum, c->colloid_rep, c->st);
    else
        launch_maps<float>(c->ph, dev_for(c, c->cur), c->fl[c->cur].f32, c->cells_total, c->scan_bsum, c->colloid_rep, c->st);
    return BW_OK;
}

BW_API int neighbour_list_md(bw_ctx *c) {
    if (!c) return BW_ERR_ARG;
    if (c->nc_total == 0) return BW_OK;
    Dev d = dev_for(c, c->cur);
    launch_md_candidates(c->ph, d, c->wgt, c->nwg, c->st);
    launch_colloid(c->ph, d, PH_MDLIST, c->st);
    return BW_OK;
}

#define COLLOID_ENTRY(name, flags)                                   \
    BW_API int name(bw_ctx *c) {                                     \
        if (!c) return BW_ERR_ARG;                                   \
        launch_colloid(c->ph, dev_for(c, c->cur), (flags), c->st);   \
        return BW_OK;                                                \
    }
COLLOID_ENTRY(compute_force_md, PH_FORCE)
COLLOID_ENTRY(update_activity_direction, PH_ACTIVITY)
COLLOID_ENTRY(update_velocity_colloid, PH_VELOCITY)
COLLOID_ENTRY(bw_copy_old_force, PH_OLDF)

BW_API int update_pos_md(bw_ctx *c) {
    if (!c) return BW_ERR_ARG;
    if (c->nc_total == 0) return BW_OK;
    launch_colloid(c->ph, dev_for(c, c->cur), PH_POS, c->st);
    c->cb ^= 1;  // the new positions are in the other buffer
    return BW_OK;
}

BW_API int rotation_mpcd(bw_ctx *c) {
    if (!c) return BW_ERR_ARG;
    int rc = alloc_maps(c);
    if (rc) return rc;
    if ((rc = sync_injected(c))) return rc;
    Dev d = dev_for(c, c->cur);
    launch_step_begin(c->ph, d, c->st);
    if (c->fp64) {
        launch_depart<double>(c->ph, d, c->fl[c->cur].f64, c->st);
        launch_srd<double>(c->ph, d, c->fl[c->cur].f64, c->fl[1 - c->cur].f64, false, true, false, c->srd_lds, c->st);
    } else {
        launch_depart<float>(c->ph, d, c->fl[c->cur].f32, c->st);
        launch_srd<float>(c->ph, d, c->fl[c->cur].f32, c->fl[1 - c->cur].f32, false, true, false, c->srd_lds, c->st);
    }
    c->cur ^= 1;
    c->substep = 0;
    c->steps_done++;
    return BW_OK;
}

BW_API int run(bw_ctx *c) {
    if (!c) return BW_ERR_ARG;
    Dev d = dev_for(c, c->cur);
    launch_colloid(c->ph, d, PH_KICK, c->st);
    if (c->nc_total == 0) return BW_OK;
    launch_tile_colloids(c->ph, d, c->st);
    if (c->fp64) {
        launch_run_count<double>(c->ph, d, c->fl[c->cur].f64, c->st);
        launch_stream<double>(c->ph, d, c->fl[c->cur].f64, true, false, false, false, c->st);
    } else {
        launch_run_count<float>(c->ph, d, c->fl[c->cur].f32, c->st);
        launch_stream<float>(c->ph, d, c->fl[c->cur].f32, true, false, false, false, c->st);
    }
    return BW_OK;
}

// cpp/run.cpp:28-50 (the reference computes it and discards the results);
// needs the neigh_fl lists of neighbour_list_mpcd
BW_API int updown_velocity(bw_ctx *c) {
    if (!c) return BW_ERR_ARG;
    if (!c->maps_alloc) return fail(c, BW_ERR_STATE, "updown_velocity: call neighbour_list_mpcd first");
    int rc;
    const size_t NC = std::max(c->nc_total, 1);
    if (!c->stage_pos && ((rc = dalloc(c, &c->stage_pos, 3 * (size_t)c->nf_total)) ||
                          (rc = dalloc(c, &c->stage_vel, 3 * (size_t)c->nf_total)) ||
                          (rc = dalloc(c, &c->stage_cell, c->nf_total))))
        return rc;
    if (!c->ud_sums && ((rc = dalloc(c, &c->ud_sums, 8 * NC)) || (rc = dalloc(c, &c->ud_vel, 3 * NC)) ||
                        (rc = dalloc(c, &c->ud_up, 3 * NC)) || (rc = dalloc(c, &c->ud_cnt, NC)) ||
                        (rc = dalloc(c, &c->ud_upcnt, NC))))
        return rc;
    Dev d = dev_for(c, c->cur);
    if (c->fp64) launch_gather<double>(c->ph, d, c->fl[c->cur].f64, c->stage_pos, c->stage_vel, c->stage_cell, c->st);
    else launch_gather<float>(c->ph, d, c->fl[c->cur].f32, c->stage_pos, c->stage_vel, c->stage_cell, c->st);
    launch_updown(c->ph, d, c->n_rep, c->stage_pos, c->stage_vel, c->ud_sums, c->ud_cnt, c->ud_upcnt, c->ud_vel,
                  c->ud_up, c->st);
    return BW_OK;
}

BW_API int bw_get_updown(bw_ctx *c, int r, int *cnt, int *up_cnt, double *vel, double *up_vel) {
    ReplicaHost *R = rep_of(c, r);
    if (!R) return BW_ERR_ARG;
    if (!c->ud_sums) return fail(c, BW_ERR_STATE, "bw_get_updown: updown_velocity not run");
    HIPCHK(c, hipStreamSynchronize(c->st));
    const size_t g0 = (size_t)R->first_colloid, n = (size_t)R->nc;
    if (cnt) HIPCHK(c, hipMemcpy(cnt, c->ud_cnt + g0, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (up_cnt) HIPCHK(c, hipMemcpy(up_cnt, c->ud_upcnt + g0, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (vel) HIPCHK(c, hipMemcpy(vel, c->ud_vel + 3 * g0, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    if (up_vel) HIPCHK(c, hipMemcpy(up_vel, c->ud_up + 3 * g0, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    return BW_OK;
}

BW_API int update_pos_mpcd(bw_ctx *c) {
    if (!c) return BW_ERR_ARG;
    Dev d = dev_for(c, c->cur);
    if (c->fp64) launch_stream<double>(c->ph, d, c->fl[c->cur].f64, false, true, false, false, c->st);
    else launch_stream<float>(c->ph, d, c->fl[c->cur].f32, false, true, false, false, c->st);
    return BW_OK;
}

BW_API int fluid_colloid_collision(bw_ctx *c) {
    if (!c) return BW_ERR_ARG;
    int rc = sync_injected(c);
    if (rc) return rc;
    if (c->nc_total == 0) return BW_OK;
    Dev d = dev_for(c, c->cur);
    int l = c->substep % c->ph.n_sub;
    launch_tile_colloids(c->ph, d, c->st);
    if (c->fp64) {
        launch_stream<double>(c->ph, d, c->fl[c->cur].f64, false, false, true, false, c->st);
        launch_collide<double>(c->ph, d, c->fl[c->cur].f64, l, c->wgt, c->nwg, c->st);
    } else {
        launch_stream<float>(c->ph, d, c->fl[c->cur].f32, false, false, true, false, c->st);
        launch_collide<float>(c->ph, d, c->fl[c->cur].f32, l, c->wgt, c->nwg, c->st);
    }
    launch_colloid(c->ph, d, PH_APPLY, c->st);
    c->substep++;
    return BW_OK;
}

BW_API int bw_startup(bw_ctx *c) {
    int rc;
    // cpp/active_brownian.cpp:14-21
    if ((rc = initialize(c)) || (rc = initialize_colloid(c)) || (rc = initialize_fluid(c)) || (rc = create_box(c)) ||
        (rc = neighbour_list_mpcd(c)) || (rc = neighbour_list_md(c)) || (rc = compute_force_md(c)) || (rc = tumble(c)))
        return rc;
    return BW_OK;
}

BW_API int bw_step(bw_ctx *c, int nsteps) {
    if (!c) return BW_ERR_ARG;
    return c->fp64 ? do_steps<double>(c, nsteps) : do_steps<float>(c, nsteps);
}

BW_API int bw_profile_steps(bw_ctx *c, int nsteps, double *ms, int *launches) {
    if (!c || !ms) return BW_ERR_ARG;
    return c->fp64 ? do_profile<double>(c, nsteps, ms, launches) : do_profile<float>(c, nsteps, ms, launches);
}

BW_API int bw_time_steps(bw_ctx *c, int nsteps, double *ms) {
    if (!c || !ms) return BW_ERR_ARG;
    hipEvent_t a, b;
    HIPCHK(c, hipEventCreate(&a));
    HIPCHK(c, hipEventCreate(&b));
    HIPCHK(c, hipEventRecord(a, c->st));
    int rc = bw_step(c, nsteps);
    HIPCHK(c, hipEventRecord(b, c->st));
    HIPCHK(c, hipEventSynchronize(b));
    float t = 0;
    hipEventElapsedTime(&t, a, b);
    *ms = t;
    hipEventDestroy(a);
    hipEventDestroy(b);
    return rc;
}

// ------------------------------------------------------------------ state
BW_API int bw_set_v0(bw_ctx *c, int r, const double *v0) {
    ReplicaHost *R = rep_of(c, r);
    if (!R) return BW_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (int j = 0; j < R->nc; j++) R->v0[j] = v0 ? v0[j] : R->p.v0;
    if (R->nc)
        HIPCHK(c, hipMemcpy(c->dv.cv0 + R->first_colloid, R->v0.data(), sizeof(double) * R->nc, hipMemcpyHostToDevice));
    return BW_OK;
}

BW_API int bw_upload_fluid(bw_ctx *c, int r, const double *pos, const double *vel) {
    ReplicaHost *R = rep_of(c, r);
    if (!R || !pos || !vel) return BW_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return c->fp64 ? upload_fluid_impl<double>(c, r, pos, vel) : upload_fluid_impl<float>(c, r, pos, vel);
}

BW_API int bw_download_fluid(bw_ctx *c, int r, double *pos, double *vel, int *cell) {
    ReplicaHost *R = rep_of(c, r);
    if (!R) return BW_ERR_ARG;
    return c->fp64 ? download_fluid_impl<double>(c, r, pos, vel, cell) : download_fluid_impl<float>(c, r, pos, vel, cell);
}

BW_API int bw_upload_colloid(bw_ctx *c, int r, const double *pos, const double *vel, const double *omega,
                             const double *f, const double *old_f, const double *ra) {
    ReplicaHost *R = rep_of(c, r);
    if (!R) return BW_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    size_t o = 3 * (size_t)R->first_colloid, n = sizeof(double) * 3 * R->nc;
    if (!R->nc) return BW_OK;
    if (pos) HIPCHK(c, hipMemcpy(c->cposb[c->cb] + o, pos, n, hipMemcpyHostToDevice));
    if (vel) HIPCHK(c, hipMemcpy(c->dv.cvel + o, vel, n, hipMemcpyHostToDevice));
    if (omega) HIPCHK(c, hipMemcpy(c->dv.comega + o, omega, n, hipMemcpyHostToDevice));
    if (f) HIPCHK(c, hipMemcpy(c->dv.cf + o, f, n, hipMemcpyHostToDevice));
    if (old_f) HIPCHK(c, hipMemcpy(c->dv.coldf + o, old_f, n, hipMemcpyHostToDevice));
    if (ra) HIPCHK(c, hipMemcpy(c->dv.cra + o, ra, n, hipMemcpyHostToDevice));
    if (pos) {
        int rc = set_colloid_cells(c, r);
        if (rc) return rc;
    }
    return BW_OK;
}

BW_API int bw_download_colloid(bw_ctx *c, int r, double *pos, double *vel, double *omega, double *f, double *old_f,
                               double *ra) {
    ReplicaHost *R = rep_of(c, r);
    if (!R) return BW_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    size_t o = 3 * (size_t)R->first_colloid, n = sizeof(double) * 3 * R->nc;
    if (!R->nc) return BW_OK;
    if (pos) HIPCHK(c, hipMemcpy(pos, c->cposb[c->cb] + o, n, hipMemcpyDeviceToHost));
    if (vel) HIPCHK(c, hipMemcpy(vel, c->dv.cvel + o, n, hipMemcpyDeviceToHost));
    if (omega) HIPCHK(c, hipMemcpy(omega, c->dv.comega + o, n, hipMemcpyDeviceToHost));
    if (f) HIPCHK(c, hipMemcpy(f, c->dv.cf + o, n, hipMemcpyDeviceToHost));
    if (old_f) HIPCHK(c, hipMemcpy(old_f, c->dv.coldf + o, n, hipMemcpyDeviceToHost));
    if (ra) HIPCHK(c, hipMemcpy(ra, c->dv.cra + o, n, hipMemcpyDeviceToHost));
    return BW_OK;
}

BW_API int bw_download_displacement(bw_ctx *c, int r, double *disp) {
    ReplicaHost *R = rep_of(c, r);
    if (!R || !disp) return BW_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (R->nc)
        HIPCHK(c, hipMemcpy(disp, c->dv.cdisp + 3 * (size_t)R->first_colloid, sizeof(double) * 3 * R->nc,
                            hipMemcpyDeviceToHost));
    return BW_OK;
}

BW_API int bw_rng_state(bw_ctx *c, int r, int *out36) {
    ReplicaHost *R = rep_of(c, r);
    if (!R || !out36) return BW_ERR_ARG;
    R->rng.get(out36);
    return BW_OK;
}

// ------------------------------------------------------------------ draws
BW_API int bw_inject_rotation(bw_ctx *c, int r, const double *shift3, const double *axes) {
    ReplicaHost *R = rep_of(c, r);
    if (!R || c->ph.draws != BW_DRAWS_INJECTED) return BW_ERR_ARG;
    for (int k = 0; k < 3; k++) R->inj_shift[k] = shift3 ? shift3[k] : NAN;
    if (axes) R->inj_axes.assign(axes, axes + 2 * (size_t)R->cells);
    else R->inj_axes.assign(2 * (size_t)R->cells, NAN);
    c->inj_dirty = true;
    return BW_OK;
}

BW_API int bw_inject_reflections(bw_ctx *c, int r, int l, int n, const int *colloid, const int *particle,
                                 const double *vals6) {
    ReplicaHost *R = rep_of(c, r);
    if (!R || c->ph.draws != BW_DRAWS_INJECTED || l < 0 || l >= c->ph.n_sub || n < 0) return BW_ERR_ARG;
    R->inj_c[l].assign(colloid, colloid + n);
    R->inj_p[l].assign(particle, particle + n);
    R->inj_v[l].assign(vals6, vals6 + 6 * (size_t)n);
    for (int i = 0; i < n; i++)
        if (colloid[i] < 0 || colloid[i] >= R->nc) return BW_ERR_ARG;
    c->inj_dirty = true;
    return BW_OK;
}

BW_API int bw_clear_injected(bw_ctx *c) {
    if (!c) return BW_ERR_ARG;
    for (auto &R : c->reps) {
        for (int l = 0; l < c->ph.n_sub; l++) {
            R.inj_c[l].clear(); R.inj_p[l].clear(); R.inj_v[l].clear();
        }
        R.inj_axes.clear();
        R.inj_shift[0] = R.inj_shift[1] = R.inj_shift[2] = NAN;
    }
    c->inj_dirty = true;
    return BW_OK;
}

// ------------------------------------------------------------------- maps
BW_API int bw_get_cell_lists(bw_ctx *c, int r, int *counts, int *list) {
    ReplicaHost *R = rep_of(c, r);
    if (!R || !c->maps_alloc) return BW_ERR_STATE;
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (counts) HIPCHK(c, hipMemcpy(counts, c->dv.map_count + R->first_cell, sizeof(int) * R->cells, hipMemcpyDeviceToHost));
    if (list) HIPCHK(c, hipMemcpy(list, c->dv.map_list + R->first_fluid, sizeof(int) * R->nf, hipMemcpyDeviceToHost));
    return BW_OK;
}

BW_API int bw_get_neigh_fl(bw_ctx *c, int r, int *counts, int *list, int cap) {
    ReplicaHost *R = rep_of(c, r);
    if (!R || !c->maps_alloc) return BW_ERR_STATE;
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (!R->nc) return 0;
    std::vector<int> cnt(R->nc);
    HIPCHK(c, hipMemcpy(cnt.data(), c->dv.neigh_cnt + R->first_colloid, sizeof(int) * R->nc, hipMemcpyDeviceToHost));
    long long tot = 0;
    for (int v : cnt) tot += v;
    if (counts) memcpy(counts, cnt.data(), sizeof(int) * R->nc);
    if (!list) return (int)tot;
    if (tot > cap) return BW_ERR_CAPACITY;
    long long w = 0;
    for (int j = 0; j < R->nc; j++) {
        HIPCHK(c, hipMemcpy(list + w, c->dv.neigh_list + (size_t)(R->first_colloid + j) * c->dv.neigh_cap,
                            sizeof(int) * cnt[j], hipMemcpyDeviceToHost));
        w += cnt[j];
    }
    return (int)tot;
}

BW_API int bw_get_md_list(bw_ctx *c, int r, int *counts, int *list, int cap) {
    ReplicaHost *R = rep_of(c, r);
    if (!R) return BW_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (!R->nc) return 0;
    std::vector<int> cnt(R->nc), full((size_t)R->nc * MD_CAP);
    HIPCHK(c, hipMemcpy(cnt.data(), c->dv.md_cnt + R->first_colloid, sizeof(int) * R->nc, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(full.data(), c->dv.md_list + (size_t)R->first_colloid * MD_CAP, sizeof(int) * full.size(),
                        hipMemcpyDeviceToHost));
    int w = 0;
    for (int i = 0; i < R->nc; i++) {
        int g = R->first_colloid + i, k = 0;
        for (int m = 0; m < cnt[i]; m++) {
            int j = full[(size_t)i * MD_CAP + m];
            if (j > g) {
                if (list) {
                    if (w >= cap) return BW_ERR_CAPACITY;
                    list[w] = j - R->first_colloid;
                }
                w++;
                k++;
            }
        }
        if (counts) counts[i] = k;
    }
    return w;
}

BW_API int bw_get_shifted_cells(bw_ctx *c, int r, int *cells) {
    ReplicaHost *R = rep_of(c, r);
    if (!R || !c->maps_alloc || !cells) return BW_ERR_STATE;
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, hipMemcpy(cells, c->dv.shifted_cell + R->first_fluid, sizeof(int) * R->nf, hipMemcpyDeviceToHost));
    return BW_OK;
}

BW_API int bw_get_run_counts(bw_ctx *c, int r, int *cnt) {
    ReplicaHost *R = rep_of(c, r);
    if (!R || !cnt) return BW_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (R->nc) HIPCHK(c, hipMemcpy(cnt, c->dv.ccnt + R->first_colloid, sizeof(int) * R->nc, hipMemcpyDeviceToHost));
    return BW_OK;
}

// ------------------------------------------------------------------ audit
BW_API int bw_audit(bw_ctx *c, int r, double *out) {
    ReplicaHost *R = rep_of(c, r);
    if (!R || !out) return BW_ERR_ARG;
    Dev d = dev_for(c, c->cur);
    if (c->fp64) launch_audit<double>(c->ph, d, c->fl[c->cur].f64, c->st);
    else launch_audit<float>(c->ph, d, c->fl[c->cur].f32, c->st);
    launch_colloid(c->ph, d, PH_POTENTIAL, c->st);
    HIPCHK(c, hipStreamSynchronize(c->st));
    std::vector<double> part((size_t)R->n_tiles * 8);
    HIPCHK(c, hipMemcpy(part.data(), c->dv.audit_partial + (size_t)R->first_tile * 8, sizeof(double) * part.size(),
                        hipMemcpyDeviceToHost));
    double s[5] = {0, 0, 0, 0, 0};
    for (int t = 0; t < R->n_tiles; t++)
        for (int k = 0; k < 5; k++) s[k] += part[(size_t)t * 8 + k];
    std::vector<double> v(3 * R->nc), w(3 * R->nc);
    double pot = 0;
    if (R->nc) {
        HIPCHK(c, hipMemcpy(v.data(), c->dv.cvel + 3 * R->first_colloid, sizeof(double) * v.size(), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(w.data(), c->dv.comega + 3 * R->first_colloid, sizeof(double) * w.size(), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(&pot, c->dv.cpot + r, sizeof(double), hipMemcpyDeviceToHost));
    }
    double kec = 0, ang = 0, pc[3] = {0, 0, 0}, absm = 0;
    for (int i = 0; i < R->nc; i++) {
        kec += (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1]) + v[3 * i + 2] * v[3 * i + 2];
        ang += (w[3 * i] * w[3 * i] + w[3 * i + 1] * w[3 * i + 1]) + w[3 * i + 2] * w[3 * i + 2];
        for (int k = 0; k < 3; k++) {
            pc[k] += v[3 * i + k] * R->p.mass_colloid;
            absm += fabs(v[3 * i + k]) * R->p.mass_colloid;
        }
    }
    double mf = R->p.mass_fl;
    out[0] = 0.5 * R->p.mass_colloid * kec;
    out[1] = 0.5 * c->ph.I_colloid * ang;
    out[2] = pot;
    out[3] = 0.5 * s[3] * mf;
    for (int k = 0; k < 3; k++) {
        out[4 + k] = pc[k] + s[k] * mf;
        out[7 + k] = s[k] * mf;
    }
    out[10] = absm + s[4] * mf;
    return BW_OK;
}

// -------------------------------------------------------------- host only
BW_API int bw_host_initialize(const bw_params *p, double *pos_fl, double *vel_fl, double *pos_c, double *vel_c,
                              double *omega, double *ra, int *rng_out36) {
    if (!p || p->L < 1 || p->density < 1 || p->n_colloid < 0 || (long long)p->L * p->L * p->L * p->density > INT_MAX)
        return BW_ERR_ARG;
    ReplicaHost R;
    R.p = *p;
    R.L = p->L;
    R.nc = p->n_colloid;
    R.cells = p->L * p->L * p->L;
    R.nf = R.cells * p->density;
    R.rng.reset(p->seed);
    InitParams ip = init_params(R);
    std::vector<double> cp(3 * std::max(R.nc, 1)), cv(3 * std::max(R.nc, 1)), cw(3 * std::max(R.nc, 1)),
        cr(3 * std::max(R.nc, 1));
    std::vector<double> fp(3 * (size_t)R.nf), fv(3 * (size_t)R.nf);
    if (R.nc) init_colloids(ip, R.rng, cp.data(), cv.data(), cw.data());
    init_fluid(ip, R.rng, cp.data(), fp.data(), fv.data());
    if (R.nc) init_orientation(ip, R.rng, cp.data(), cr.data());
    if (pos_fl) memcpy(pos_fl, fp.data(), sizeof(double) * fp.size());
    if (vel_fl) memcpy(vel_fl, fv.data(), sizeof(double) * fv.size());
    size_t n = sizeof(double) * 3 * R.nc;
    if (pos_c) memcpy(pos_c, cp.data(), n);
    if (vel_c) memcpy(vel_c, cv.data(), n);
    if (omega) memcpy(omega, cw.data(), n);
    if (ra) memcpy(ra, cr.data(), n);
    if (rng_out36) R.rng.get(rng_out36);
    return BW_OK;
}

BW_API int bw_observables_reset(bw_ctx *c) {
    if (!c) return BW_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    size_t NC = std::max(c->nc_total, 1);
    HIPCHK(c, hipMemsetAsync(c->dv.cdisp, 0, sizeof(double) * 3 * NC, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (auto &R : c->reps) {
        R.ra0.assign(3 * (size_t)R.nc, 0.0);
        if (R.nc)
            HIPCHK(c, hipMemcpy(R.ra0.data(), c->dv.cra + 3 * (size_t)R.first_colloid, sizeof(double) * 3 * R.nc,
                                hipMemcpyDeviceToHost));
        R.obs_step0 = c->steps_done;
    }
    return BW_OK;
}

BW_API int bw_observables(bw_ctx *c, int r, double *out) {
    ReplicaHost *R = rep_of(c, r);
    if (!R || !out) return BW_ERR_ARG;
    double a[11];
    int rc = bw_audit(c, r, a);  // synchronises; FP64 fluid sums
    if (rc) return rc;
    const double m = R->p.mass_fl, N = (double)R->nf;
    const double p2 = a[7] * a[7] + a[8] * a[8] + a[9] * a[9];  // |m sum v|^2
    out[0] = (double)(c->steps_done - R->obs_step0);
    out[3] = R->nf > 1 ? (2.0 * a[3] - p2 / (N * m)) / (3.0 * (N - 1.0)) : 0.0;
    out[1] = out[2] = out[4] = out[5] = 0.0;
    if (R->nc) {
        size_t n3 = 3 * (size_t)R->nc, o = 3 * (size_t)R->first_colloid;
        std::vector<double> d(n3), ra(n3), v(n3), w(n3);
        HIPCHK(c, hipMemcpy(d.data(), c->dv.cdisp + o, sizeof(double) * n3, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(ra.data(), c->dv.cra + o, sizeof(double) * n3, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(v.data(), c->dv.cvel + o, sizeof(double) * n3, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(w.data(), c->dv.comega + o, sizeof(double) * n3, hipMemcpyDeviceToHost));
        if (R->ra0.size() != n3) R->ra0 = ra;  // no reset yet: the origin is now
        double msd = 0, oc = 0, tv = 0, tw = 0;
        for (int i = 0; i < R->nc; i++) {
            double dd = 0, rr = 0, vv = 0, ww = 0;
            for (int k = 0; k < 3; k++) {
                dd += d[3 * i + k] * d[3 * i + k];
                rr += ra[3 * i + k] * R->ra0[3 * i + k];
                vv += v[3 * i + k] * v[3 * i + k];
                ww += w[3 * i + k] * w[3 * i + k];
            }
            msd += dd; oc += rr; tv += vv; tw += ww;
        }
        out[1] = msd / R->nc;
        out[2] = oc / R->nc;
        out[4] = R->p.mass_colloid * tv / (3.0 * R->nc);
        out[5] = c->ph.I_colloid * tw / (3.0 * R->nc);
    }
    return BW_OK;
}

BW_API int bw_set_moments(bw_ctx *c, int on) {
    if (!c) return BW_ERR_ARG;
    if (on && !c->moments_ok) return fail(c, BW_ERR_ARG, "coarse moments need even box and tile edges");
    if (!on && c->ph.surrogate) return fail(c, BW_ERR_STATE, "the attached surrogate needs the moments");
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->ph.moments = on ? 1 : 0;
    HIPCHK(c, hipMemsetAsync(c->dv.mom_valid, 0, sizeof(int), c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    destroy_graphs(c);  // Phys is part of the captured launches
    return BW_OK;
}

BW_API int bw_surrogate_attach(bw_ctx *c, bw_mlp *m, const float *lo, const float *hi, const float *t_lo,
                               const float *t_hi) {
    if (!c) return BW_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    destroy_graphs(c);
    if (!m) {
        c->mlp = nullptr;
        // fluid inside the colloids that were phantoms collides with none
        // until it has left them
        if (c->ph.surrogate) {
            if (c->fp64) launch_mark_phantoms<double>(c->ph, dev_for(c, c->cur), c->fl[c->cur].f64, c->st);
            else launch_mark_phantoms<float>(c->ph, dev_for(c, c->cur), c->fl[c->cur].f32, c->st);
        }
        c->ph.surrogate = 0;
        HIPCHK(c, hipMemsetAsync(c->dv.cmode, 0, sizeof(int) * std::max(c->nc_total, 1), c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        return BW_OK;
    }
    if (!lo || !hi || !t_lo || !t_hi) return BW_ERR_ARG;
    if (!c->moments_ok) return fail(c, BW_ERR_ARG, "the surrogate needs even box and tile edges");
    if (c->srd_lds + srd_sur_lds() > 160 * 1024 - srd_static_lds())
        return fail(c, BW_ERR_ARG, "tiles too large for the rotation pass's surrogate data in LDS");
    if (c->nc_total == 0) return fail(c, BW_ERR_ARG, "no colloids");
    if (mlp_buf_alloc(c->mbuf, c->nc_total)) return fail(c, BW_ERR_HIP, "hipMalloc (surrogate activations)");
    for (int k = 0; k < 2; k++) {
        const float **dst = k ? &c->dv.sur_hi : &c->dv.sur_lo;
        if (!*dst) {
            void *q = nullptr;
            HIPCHK(c, hipMalloc(&q, sizeof(float) * (N_FEAT + 6)));
            *dst = (const float *)q;
        }
        HIPCHK(c, hipMemcpy((void *)*dst, k ? hi : lo, sizeof(float) * N_FEAT, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy((void *)(*dst + N_FEAT), k ? t_hi : t_lo, sizeof(float) * 6, hipMemcpyHostToDevice));
    }
    if (!c->ph.moments) {
        c->ph.moments = 1;
        HIPCHK(c, hipMemsetAsync(c->dv.mom_valid, 0, sizeof(int), c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
    }
    c->mlp = m;
    c->ph.surrogate = 1;
    return BW_OK;
}

// per colloid of replica r after a fused step: the features it started with
// (n_colloid x 144), its collision impulse over the step (n_colloid x 6:
// momentum and angular momentum in units of the fluid mass) and its mode
// (0 explicit, 1 surrogate)
BW_API int bw_get_surrogate_data(bw_ctx *c, int r, float *feat, double *imp, int *mode) {
    ReplicaHost *R = rep_of(c, r);
    if (!R) return BW_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (!R->nc) return BW_OK;
    size_t g0 = (size_t)R->first_colloid;
    if (feat)
        HIPCHK(c, hipMemcpy(feat, c->dv.feat + g0 * N_FEAT, sizeof(float) * N_FEAT * R->nc, hipMemcpyDeviceToHost));
    if (imp) HIPCHK(c, hipMemcpy(imp, c->dv.cimp + g0 * 6, sizeof(double) * 6 * R->nc, hipMemcpyDeviceToHost));
    if (mode) HIPCHK(c, hipMemcpy(mode, c->dv.cmode + g0, sizeof(int) * R->nc, hipMemcpyDeviceToHost));
    return BW_OK;
}

BW_API int bw_get_surrogate_shells(bw_ctx *c, int r, int *count, double *sums) {
    ReplicaHost *R = rep_of(c, r);
    if (!R) return BW_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (!R->nc) return BW_OK;
    size_t g0 = (size_t)R->first_colloid;
    if (count) HIPCHK(c, hipMemcpy(count, c->dv.sshell + g0, sizeof(int) * R->nc, hipMemcpyDeviceToHost));
    if (sums) {
        std::vector<unsigned long long> q((size_t)SUR_ACC * R->nc);
        HIPCHK(c, hipMemcpy(q.data(), c->dv.sacc + g0 * SUR_ACC, sizeof(unsigned long long) * q.size(),
                            hipMemcpyDeviceToHost));
        for (size_t i = 0; i < q.size(); i++) sums[i] = (double)(long long)q[i] * 0x1p-40;
    }
    return BW_OK;
}

BW_API int bw_near_particles(bw_ctx *c, int r, long long *n_near) {
    ReplicaHost *R = rep_of(c, r);
    if (!R || !n_near) return BW_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    std::vector<int> cnt(R->n_tiles);
    HIPCHK(c, hipMemcpy(cnt.data(), c->dv.near_cnt + R->first_tile, sizeof(int) * R->n_tiles, hipMemcpyDeviceToHost));
    long long s = 0;
    for (int v : cnt) s += v;
    *n_near = s;
    return BW_OK;
}

BW_API int bw_debug_geometry(bw_ctx *c, int n, const double *a, double L, const int *ia, int Li, int b,
                             double *img_out, double *mod_out, int *wrap_out, int *div_out) {
    if (!c || n < 1 || !a || !ia || !img_out || !mod_out || !wrap_out || !div_out || Li < 1 || b < 1 || b > 63)
        return BW_ERR_ARG;
    const size_t nd = sizeof(double) * n, ni = sizeof(int) * n;
    char *buf = nullptr;
    HIPCHK(c, hipMalloc((void **)&buf, 3 * nd + 3 * ni));
    double *da = (double *)buf, *dimg = da + n, *dmod = dimg + n;
    int *dia = (int *)(dmod + n), *dwrap = dia + n, *ddiv = dwrap + n;
    HIPCHK(c, hipMemcpyAsync(da, a, nd, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(dia, ia, ni, hipMemcpyHostToDevice, c->st));
    launch_debug_geometry(n, da, L, dia, Li, b, dimg, dmod, dwrap, ddiv, c->st);
    HIPCHK(c, hipMemcpyAsync(img_out, dimg, nd, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(mod_out, dmod, nd, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(wrap_out, dwrap, ni, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(div_out, ddiv, ni, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    hipFree(buf);
    return BW_OK;
}

BW_API int bw_debug_philox(bw_ctx *c, int n, const unsigned *ctr, const unsigned *key, unsigned *out) {
    if (!c || n < 1 || !ctr || !key || !out) return BW_ERR_ARG;
    char *buf = nullptr;
    const size_t b4 = 16 * (size_t)n, b2 = 8 * (size_t)n;
    HIPCHK(c, hipMalloc((void **)&buf, 2 * b4 + b2));
    uint4 *dc = (uint4 *)buf, *dout = (uint4 *)(buf + b4);
    uint2 *dk = (uint2 *)(buf + 2 * b4);
    HIPCHK(c, hipMemcpyAsync(dc, ctr, b4, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(dk, key, b2, hipMemcpyHostToDevice, c->st));
    launch_debug_philox(n, dc, dk, dout, c->st);
    HIPCHK(c, hipMemcpyAsync(out, dout, b4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    hipFree(buf);
    return BW_OK;
}

BW_API int bw_debug_reflections(bw_ctx *c, int n, long long step, int substep, const double *rf3, const double *rs3,
                                double *vals6, double *vel3) {
    if (!c || n < 1 || !rf3 || !rs3 || !vals6 || !vel3 || substep < 0 || substep > 255) return BW_ERR_ARG;
    if (c->reps[0].nc < 1) return fail(c, BW_ERR_ARG, "bw_debug_reflections: replica 0 has no colloid");
    Phys ph = c->ph;
    ph.draws = BW_DRAWS_DEVICE;  // the device source, whatever the context's draws
    double *buf = nullptr;
    HIPCHK(c, hipMalloc((void **)&buf, sizeof(double) * 9 * (size_t)n));
    launch_debug_reflect(ph, c->dv, n, step, substep, (double)c->reps[0].L, rf3, rs3, buf, buf + 6 * (size_t)n, c->st);
    HIPCHK(c, hipMemcpyAsync(vals6, buf, sizeof(double) * 6 * (size_t)n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(vel3, buf + 6 * (size_t)n, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    hipFree(buf);
    return BW_OK;
}

BW_API int bw_debug_copy(bw_ctx *c, long long nbytes, int reps, double *ms) {
    if (!c || nbytes <= 0 || reps < 1) return BW_ERR_ARG;
    void *a = nullptr, *b = nullptr;
    HIPCHK(c, hipMalloc(&a, nbytes));
    HIPCHK(c, hipMalloc(&b, nbytes));
    HIPCHK(c, hipMemsetAsync(a, 0, nbytes, c->st));
    launch_copy(a, b, nbytes, c->st);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventRecord(e0, c->st);
    for (int i = 0; i < reps; i++) launch_copy(a, b, nbytes, c->st);
    hipEventRecord(e1, c->st);
    hipEventSynchronize(e1);
    float t = 0;
    hipEventElapsedTime(&t, e0, e1);
    if (ms) *ms = t / reps;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    return BW_OK;
}

}  // extern "C"
