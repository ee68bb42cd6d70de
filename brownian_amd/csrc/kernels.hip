// HIP kernels of libbrownian for gfx950 (CDNA4, wave64).
//
// Fused MPCD step (one replica or many, one launch per kernel for all):
//   k_step_begin       step counter, grid shift of this step               (1 WG)
//   k_md_candidates    colloid pair candidates (cutoff + skin)              (1 wave / colloid)
//   k_tile_colloids    per tile: colloids that can touch it this step        (1 WG / tile)
//   k_depart           per tile: each particle's tile on the new shifted
//                      grid, stable departure lists per direction           (1 WG / tile)
//   k_srd              per tile: gather from 27 neighbour lists, LDS
//                      counting sort by shifted cell, FP64 per-cell sums,
//                      rotation + thermostat, run() shell counts             (1 WG / tile)
//   n_sub x { k_stream          per tile: run() kick (first substep), streaming,
//                                detection through LDS stencil masks
//             k_collide         1 wave / colloid: pair list, reflections;
//                                last WG of a replica: serial replays
//             k_colloid_update  1 wave / colloid: collision impulse, activity,
//                                WCA force, velocity, [kick], position }
#include "bw_device.h"

namespace bw {

// ------------------------------------------------------------- step begin
// Block 0: the step's shifts and counters.  clear_lists: every block also
// clears the per-tile colloid lists (k_tile_lists_clear, folded in).
__global__ void k_step_begin(Phys ph, Dev d, int clear_lists) {
    if (clear_lists) {
        for (int T = blockIdx.x * blockDim.x + threadIdx.x; T < ph.n_tiles; T += gridDim.x * blockDim.x) {
            d.tcl_cnt[T] = 0;
            if (ph.surrogate) d.tsl_cnt[T] = 0;
        }
        for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < ph.n_rep; r += gridDim.x * blockDim.x)
            d.tcl_stale[r] = 0;
    }
    if (blockIdx.x != 0) return;
    long long s = *d.step + 1;
    for (int r = threadIdx.x; r < ph.n_rep; r += blockDim.x) {
        double u[3];
        draw_shift_raw(d, ph, d.rep[r], r, s, u);
        // cpp/rotation_mpcd.cpp:9: rr = (ran,ran,ran) - 0.5
        for (int k = 0; k < 3; k++) {
            double v = u[k] * 1 - 0.5;
            if (!(v >= -0.5 && v < 0.5)) {  // injected draw absent (NaN) or out of range
                atomicOr(d.status, BW_FLAG_DRAW_MISSING);
                v = 0.0;
            }
            d.shift[3 * r + k] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        *d.step = s;
        *d.chunk_cnt = 0;
    }
}

// --------------------------------------------------------- colloid phases
__global__ void k_colloid(Phys ph, Dev d, unsigned flags) {
    int r = blockIdx.x;
    const Rep R = d.rep[r];
    colloid_phases(ph, d, R, r, flags);
}

// one wave per colloid: every partner within cutoff + skin, ascending
__global__ __launch_bounds__(256) void k_md_candidates(Phys ph, Dev d, const int2 *wgt) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int2 wt = wgt[blockIdx.x];
    const Rep R = d.rep[wt.x];
    const int g = wt.y + w;
    if (g >= R.first_colloid + R.nc) return;
    const double L = R.L;
    const double *pi = d.cpos + 3 * g;
    int cnt = 0;
    // four rounds of 64 partners per pass: their positions are requested
    // together, then tested and appended in ascending order
    constexpr int U = 4;
    for (int base = 0; base < R.nc; base += 64 * U) {
        double pj[U][3];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int m = base + 64 * u + lane;
            const int j = R.first_colloid + min(m, R.nc - 1);
            for (int k = 0; k < 3; k++) pj[u][k] = d.cpos[3 * j + k];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int m = base + 64 * u + lane, j = R.first_colloid + m;
            bool p = false;
            if (m < R.nc && j != g) {
                double tx = img1(pi[0] - pj[u][0], L), ty = img1(pi[1] - pj[u][1], L), tz = img1(pi[2] - pj[u][2], L);
                p = (tx * tx + ty * ty) + tz * tz < ph.lim_md_skin;
            }
            unsigned long long b = __ballot(p);
            int rank = __popcll(b & ((1ull << lane) - 1ull));
            if (p && cnt + rank < MDC_CAP) d.mdc_list[g * MDC_CAP + cnt + rank] = j;
            cnt += __popcll(b);
        }
    }
    if (lane == 0) {
        if (cnt > MDC_CAP) { atomicOr(d.status, BW_FLAG_LIST_OVERFLOW); cnt = MDC_CAP; }
        d.mdc_cnt[g] = cnt;
        for (int k = 0; k < 3; k++) {
            d.cpos_md[3 * g + k] = pi[k];
            d.ccell_md[3 * g + k] = d.ccell[3 * g + k];
        }
        d.md_stale[wt.x] = 0;
    }
}

// Per-tile colloid lists of the step, built colloid-first: the counters are
// cleared, then each colloid tests the tiles around its current and run()
// cells with the list's own membership rule and appends itself.  A list's
// order is then arbitrary; every user treats it as a set (masks, exact
// tests, counts).
__global__ __launch_bounds__(256) void k_tile_lists_clear(Phys ph, Dev d) {
    for (int T = blockIdx.x * blockDim.x + threadIdx.x; T < ph.n_tiles; T += gridDim.x * blockDim.x) {
        d.tcl_cnt[T] = 0;
        if (ph.surrogate) d.tsl_cnt[T] = 0;
    }
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < ph.n_rep; r += gridDim.x * blockDim.x) d.tcl_stale[r] = 0;
}
__global__ __launch_bounds__(256) void k_tile_colloids(Phys ph, Dev d) {
    // one workgroup per colloid, threads over its candidate tiles (at most 6
    // per axis when the two cells are neighbours: the window spans B + 15
    // cells, B >= 4)
    const int g = blockIdx.x;
    const Rep R = d.rep[d.crep[g]];
    const bool sur = ph.surrogate && d.cmode[g];
    int *list = sur ? d.tsl_list : d.tcl_list;
    int *cnt = sur ? d.tsl_cnt : d.tcl_cnt;
    int t0[3], nt[3];
    for (int a = 0; a < 3; a++) {  // tiles whose window can hold either cell, per axis
        const int B = R.B[a], c1 = d.ccell[3 * g + a], c2 = d.ccell_run[3 * g + a];
        const int lo = min(c1, c2) - TILE_LIST_REACH - (B - 1), hi = max(c1, c2) + TILE_LIST_REACH;
        const int tlo = lo >= 0 ? lo / B : -((-lo + B - 1) / B), thi = hi / B;
        t0[a] = wrapi(tlo, R.nt[a]);
        nt[a] = min(thi - tlo + 1, R.nt[a]);
    }
    for (int q = threadIdx.x; q < nt[0] * nt[1] * nt[2]; q += blockDim.x) {
        const int i = q % nt[0], j = (q / nt[0]) % nt[1], k = q / (nt[0] * nt[1]);
        const int tx = wrapi(t0[0] + i, R.nt[0]), ty = wrapi(t0[1] + j, R.nt[1]), tz = wrapi(t0[2] + k, R.nt[2]);
        const int T = R.first_tile + tx + R.nt[0] * (ty + R.nt[1] * tz);
        int lo3[3], ln3[3];
        tile_list_window(R, T, lo3, ln3);
        if (!tile_list_member(ph, d, R, g, lo3, ln3, sur)) continue;
        const int e = atomicAdd(&cnt[T], 1);
        if (e < TCL_CAP) list[T * TCL_CAP + e] = g;
        else atomicOr(d.status, BW_FLAG_LIST_OVERFLOW);
    }
}

// ------------------------------------------------------------ tile geometry
struct TileGeo {
    int r, tx, ty, tz, L;
};
__device__ __forceinline__ TileGeo tile_geo(const Dev &d, int T, Rep &R) {
    TileGeo t;
    t.r = d.tile_rep[T];
    R = d.rep[t.r];
    int lt = T - R.first_tile;
    t.tx = lt % R.nt[0];
    t.ty = (lt / R.nt[0]) % R.nt[1];
    t.tz = lt / (R.nt[0] * R.nt[1]);
    t.L = R.L;
    return t;
}
__device__ __forceinline__ int tile_delta(int tnew, int told, int nt) {
    int dlt = tnew - told;
    if (nt > 2) {
        if (dlt > nt / 2) dlt -= nt;
        if (dlt < -(nt / 2)) dlt += nt;
    }
    return dlt;
}

// --------------------------------------------------------------- departure
// For each particle of tile T, its tile on the SRD grid shifted by this step's
// shift, as a direction 0..26; stable (array-order) lists per direction.
template <class Real>
__global__ __launch_bounds__(DEPART_THREADS) void k_depart(Phys ph, Dev d, Fluid<Real> f) {
    constexpr int NW = DEPART_THREADS / 64, U = 4;  // waves; 64-particle groups whose loads fly together
    __shared__ unsigned char s_dir[MAX_CAP];
    __shared__ int s_cnt[NW][27], s_base[NW][27];
    const int T = blockIdx.x;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    Rep R;
    TileGeo tg = tile_geo(d, T, R);
    const int L = tg.L;
    const int n = min(d.count[T], ph.cap);
    const double sx = d.shift[3 * tg.r], sy = d.shift[3 * tg.r + 1], sz = d.shift[3 * tg.r + 2];
    const unsigned mx = div_small_magic(R.B[0]), my = div_small_magic(R.B[1]), mz = div_small_magic(R.B[2]);
    // each wave owns a contiguous range of the tile's slots, in groups of 64
    const int per = ((n + NW - 1) / NW + 63) & ~63;
    const int lo = min(n, w * per), hi = min(n, lo + per);
    for (int t = lane; t < 27; t += 64) s_cnt[w][t] = 0;
    if (w == NW - 1) {
        // the step's tile colloid lists, appended in arbitrary order, sorted:
        // every later pass meets a tile's colloids in ascending index, the
        // reference's colloid order (a particle in two run() shells is kicked
        // in that order, cpp/run.cpp:12-24)
        wave_sort_list(d.tcl_list + (size_t)T * TCL_CAP, min(d.tcl_cnt[T], TCL_CAP));
        if (ph.surrogate) wave_sort_list(d.tsl_list + (size_t)T * TCL_CAP, min(d.tsl_cnt[T], TCL_CAP));
    }
    bool drift = false;
    // pass 1: the departure direction of each particle, counts per wave (the
    // particles that stay in the tile, most of them, are counted by ballot
    // rather than by contended LDS atomics)
    int stay = 0;  // lane 0: direction 13 (same tile) of this wave
    for (int g0 = lo; g0 < hi; g0 += 64 * U) {
        int c[U];
        Real ox[U], oy[U], oz[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = g0 + 64 * u + lane;
            if (i < hi) {
                const int slot = T * ph.cap + i;
                c[u] = f.cell[slot];
                ox[u] = f.ox[slot]; oy[u] = f.oy[slot]; oz[u] = f.oz[slot];
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = g0 + 64 * u + lane;
            int dir = -1;
            if (i < hi) {
                int ix, iy, iz;
                unpack_cell(c[u], ix, iy, iz);
                int kx = shifted_coord(ix, ox[u], sx, L);
                int ky = shifted_coord(iy, oy[u], sy, L);
                int kz = shifted_coord(iz, oz[u], sz, L);
                int dx = tile_delta(div_small(kx, mx), tg.tx, R.nt[0]);
                int dy = tile_delta(div_small(ky, my), tg.ty, R.nt[1]);
                int dz = tile_delta(div_small(kz, mz), tg.tz, R.nt[2]);
                if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) {
                    drift = true;
                    dx = max(-1, min(1, dx)); dy = max(-1, min(1, dy)); dz = max(-1, min(1, dz));
                }
                dir = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
                s_dir[i] = (unsigned char)dir;
                if (dir != 13) atomicAdd(&s_cnt[w][dir], 1);
            }
            const unsigned long long m13 = __ballot(dir == 13);
            if (lane == 0) stay += __popcll(m13);
        }
    }
    if (drift) atomicOr(d.status, BW_FLAG_DRIFT);
    if (lane == 0) s_cnt[w][13] += stay;
    __syncthreads();
    // list offsets: direction-major, waves in order within a direction; lane
    // t of wave 0 owns direction t, the directions' totals are scanned across
    // lanes
    if (threadIdx.x < 64) {
        int c[NW], tot = 0;
#pragma unroll
        for (int v = 0; v < NW; v++) {
            c[v] = lane < 27 ? s_cnt[v][lane] : 0;
            tot += c[v];
        }
        int x = tot;  // inclusive prefix over directions
#pragma unroll
        for (int o = 1; o < 32; o <<= 1) {
            const int y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane < 27) {
            int acc = x - tot;
            d.dep_off[T * 28 + lane] = acc;
#pragma unroll
            for (int v = 0; v < NW; v++) {
                s_base[v][lane] = acc;
                acc += c[v];
            }
        }
        if (lane == 26) d.dep_off[T * 28 + 27] = x;
    }
    __syncthreads();
    // pass 2: each wave scatters its range in slot order (stable lists)
    const unsigned long long below = (1ull << lane) - 1ull;
    int run = lane < 27 ? s_base[w][lane] : 0;  // lane t: next offset of direction t
    for (int g0 = lo; g0 < hi; g0 += 64) {
        const int i = g0 + lane;
        const int dd = i < hi ? (int)s_dir[i] : -1;
        int off = 0;
        unsigned long long pending = __ballot(dd >= 0);
        while (pending) {  // one round per direction present in the group
            const int t = lane_value(dd, __ffsll((long long)pending) - 1);
            const unsigned long long m = __ballot(dd == t);
            const int rt = lane_value(run, t);
            if (dd == t) off = rt + __popcll(m & below);
            if (lane == t) run += __popcll(m);
            pending &= ~m;
        }
        if (dd >= 0) d.dep_idx[T * ph.cap + off] = i;
    }
}

// -------------------------------------------------------------------- SRD
// Destination tile T gathers the particles whose shifted cell lies in its box
// (27 departure lists of the neighbour tiles, fixed order), sorts them by
// shifted cell with an LDS counting sort, ranks each cell's particles by id
// (the reference's cell_part order), and per cell forms the FP64 mean velocity
// and the thermostat factor of cpp/rotation_mpcd.cpp:14-49: the per-particle
// terms are computed in parallel, each sum is taken serially in id order.
// COUNT_SHELL: also counts the run() rear shells (cpp/run.cpp:17-18).
// BLOCK (fused step): a particle that cannot meet a colloid during the whole
// step -- farther from every colloid than its radius plus the particle's own
// travel in the step plus the colloid displacement bound `margin` (checked
// on every colloid position update of the step), and outside every run()
// shell -- is streamed through all n_sub substeps here, in registers, with
// the arithmetic of the streaming kernel.  The others ("near") are stored
// first in the tile, and only they are touched by the per-substep kernels.
// x rounded to the nearest integer (ties to even), as __double2ll_rn does,
// for |x| < 2^51: x + 1.5 2^52 lies in [2^52, 2^53), where the doubles are
// the integers, so the addition's rounding is the integer rounding and the
// sum's bit pattern is that of 1.5 2^52 plus the integer
__device__ __forceinline__ long long rint_small(double x) {
    return __double_as_longlong(x + 0x1.8p52) - 0x4338000000000000LL;
}
// fixed-point (2^-40) difference a - b of two velocity components, exact for
// |a|, |b| < 1024 (larger ones are flagged by the caller and count as 0)
__device__ __forceinline__ unsigned long long fixq_diff(float a, float b) {
    const long long qa = fabsf(a) < 1024.0f ? rint_small((double)a * 0x1p40) : 0;
    const long long qb = fabsf(b) < 1024.0f ? rint_small((double)b * 0x1p40) : 0;
    return (unsigned long long)(qa - qb);
}
__device__ __forceinline__ unsigned long long fixq_diff(double a, double b) {
    const long long qa = fabs(a) < 1024.0 ? rint_small(a * 0x1p40) : 0;
    const long long qb = fabs(b) < 1024.0 ? rint_small(b * 0x1p40) : 0;
    return (unsigned long long)(qa - qb);
}
// a tile's surrogate shell counts and fixed-point sums into the per-colloid
// totals (integer sums: the same in any order)
__device__ inline void flush_surrogate_shells(const Dev &d, int ns, const int *sg, const int *scnt,
                                              const unsigned long long (*sacc)[SUR_ACC]) {
    for (int q = threadIdx.x; q < ns; q += blockDim.x) {
        const int g = sg[q];
        if (scnt[q]) atomicAdd(&d.sshell[g], scnt[q]);
        for (int k = 0; k < SUR_ACC; k++)
            if (sacc[q][k]) atomicAdd(&d.sacc[(size_t)SUR_ACC * g + k], sacc[q][k]);
    }
}
// bytes of the SRD pass's per-cell region (7 doubles per cell), which also
// holds the cap ids of the sort
__host__ __device__ inline size_t srd_cell_region(int ncell, int cap) {
    const size_t a = (size_t)ncell * 56, b = ((size_t)cap * 4 + 7) / 8 * 8;
    return a > b ? a : b;
}

// dynamic LDS of the SRD pass for tiles of ncell cells holding up to cap
// particles (the surrogate's colloid data follows, 16-byte aligned)
__host__ __device__ size_t srd_lds_bytes(int ncell, int cap, bool fp64) {
    if (!fp64)  // count, smallest id, exponent; 5 + 3 fixed-point sums, mean (3 FP64), matrix (9 FP32)
        return 8 * (((size_t)3 * ncell + 1) / 2) + (size_t)ncell * (64 + 24 + 36) + 16;
    return srd_cell_region(ncell, cap) + 3 * (size_t)cap * 8 + (size_t)(2 * ncell + 1) * 4 + (size_t)cap * 2 + 16;
}
// deepest entry into a surrogate-driven colloid that is mirrored back out
constexpr double MIRROR_DEPTH = 0.5;
// surrogate-driven colloids of a tile: position and velocity (6 FP64), the
// fixed-point sums of shell_pay (SUR_ACC), tile-local position (3 FP32), id,
// shell count
constexpr size_t SRD_SUR_LDS = (size_t)TCL_CAP * (48 + 8 * SUR_ACC + 12 + 4 + 4);

// WRAP: the near test applies the periodic wrap to its FP32 differences
// (srd_needs_wrap); otherwise they stay below L/2, where the wrap is the identity.
// PT: slots per thread (3 or 4; TH x PT slots hold the tile's capacity).
template <class Real, bool COUNT_SHELL, bool MAP, bool BLOCK, int TH, bool WRAP, int PT>
__global__ __launch_bounds__(TH) void k_srd(Phys ph, Dev d, Fluid<Real> src, Fluid<Real> dst) {
    constexpr bool NEAR_TEST = COUNT_SHELL && BLOCK;
    // FP32 storage: per-cell sums by exact fixed-point LDS atomics (no sort);
    // FP64 storage: the reference's id-ordered sums (bit parity)
    constexpr bool FAST = sizeof(Real) == 4;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_pre[28], s_src[27], s_base[27];
    __shared__ int s_wtot[16];
    __shared__ TileRun tc;
    __shared__ NearColl nc[1];  // used by the near test only
    __shared__ int s_shell[COUNT_SHELL ? TCL_CAP : 1];
    __shared__ AxisMask s_am[COUNT_SHELL ? 2 : 1];  // [0] run shells, [1] near test
    __shared__ int s_cc[PT][TH / 64];   // near particles per chunk and wave
    __shared__ int s_sn;
    const int T = blockIdx.x;
    Rep R;
    TileGeo tg = tile_geo(d, T, R);
    const int L = tg.L;
    const int cap = ph.cap;
    const int B0 = R.B[0], B1 = R.B[1], B2 = R.B[2];
    const int ncell = B0 * B1 * B2;
    const int ox0 = tg.tx * B0, oy0 = tg.ty * B1, oz0 = tg.tz * B2;  // tile origin (cells)
    // LDS carve-up.  The id list s_pk is dead once the ranks are known, before
    // the per-cell means are formed, so it shares the per-cell region.
    double *s_u = (double *)smem;                       // 3 * ncell
    double *s_ax = s_u + 3 * ncell;                     // 3 * ncell  (rotation axis)
    double *s_sc = s_ax + 3 * ncell;                    // ncell      (thermostat factor)
    int *s_pk = (int *)smem;                            // cap: ids in cell-sorted (unranked) order
    Real *s_v = (Real *)(smem + srd_cell_region(ncell, cap));  // 3 * cap; later the per-particle energy terms
    double *s_e2 = (double *)s_v;
    int *s_start = (int *)(s_v + 3 * cap);              // ncell + 1
    int *s_fill = s_start + ncell + 1;                  // ncell
    unsigned short *s_ord = (unsigned short *)(s_fill + ncell);  // cap: cell-sorted, id-ranked
    // FAST carve-up (the same buffer): count, smallest particle id, scale
    // exponent of the relative-energy sum; fixed-point sums (momentum 3,
    // |v|^2 1, |v - u|^2 1) and the momentum change of the rounded collision
    // (3); the mean (FP64) and the thermostat-scaled rotation (FP32)
    int *f_cnt = (int *)smem;                                                       // ncell
    int *f_minpid = f_cnt + ncell;                                                  // ncell
    int *f_exp = f_minpid + ncell;                                                  // ncell (+1 pad)
    unsigned long long *f_sum = (unsigned long long *)(smem + 8 * (((size_t)3 * ncell + 1) / 2));  // 5 ncell
    unsigned long long *f_dp = f_sum + 5 * ncell;                                   // 3 ncell
    double *f_u = (double *)(f_dp + 3 * ncell);                                     // 3 ncell
    float *f_m = (float *)(f_u + 3 * ncell);                                        // 9 ncell
    // surrogate-driven colloids (past the per-cell region, only while a
    // surrogate is attached): position and velocity, the fixed-point sums of
    // shell_pay, tile-local FP32 position, id, shell count
    unsigned char *sur = smem + ((srd_lds_bytes(ncell, cap, !FAST) + 15) & ~(size_t)15);
    double(*s_sp)[6] = (double(*)[6])sur;
    unsigned long long(*s_sacc)[SUR_ACC] = (unsigned long long(*)[SUR_ACC])(sur + (size_t)TCL_CAP * 48);
    float(*s_spf)[3] = (float(*)[3])(sur + (size_t)TCL_CAP * (48 + 8 * SUR_ACC));
    int *s_sg = (int *)(sur + (size_t)TCL_CAP * (60 + 8 * SUR_ACC));
    int *s_scnt = s_sg + TCL_CAP;

    // wave-uniform words through the scalar cache, so that no vector load
    // waits on them
    const long long step = ld_const(d.step);
    const double sx = ld_const(d.shift + 3 * tg.r), sy = ld_const(d.shift + 3 * tg.r + 1),
                 sz = ld_const(d.shift + 3 * tg.r + 2);

    // Every global request of the tile's set-up goes out before the first
    // barrier (the tile's colloid list and its staleness flag, the 27 source
    // lists' offsets), and the listed colloids' data goes out together with
    // the source-list indices below, so that their round trips overlap.  (The
    // step's list pass ran before this kernel; a list made stale since is
    // rebuilt below.)
    int cl_n = 0, cl_g = 0;
    bool cl_stale = false;
    if (COUNT_SHELL) {
        cl_stale = ld_const(d.tcl_stale + tg.r) != 0;
        cl_n = min(ld_const(d.tcl_cnt + T), TCL_CAP);
        if (threadIdx.x < cl_n) cl_g = d.tcl_list[T * TCL_CAP + threadIdx.x];
    }
    int dep_S = 0, dep_base = 0, dep_end = 0;
    if (threadIdx.x < 27) {
        const int t = threadIdx.x;
        const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
        const int ax = wrapi(tg.tx - dx, R.nt[0]), ay = wrapi(tg.ty - dy, R.nt[1]), az = wrapi(tg.tz - dz, R.nt[2]);
        dep_S = R.first_tile + ax + R.nt[0] * (ay + R.nt[1] * az);
        dep_base = d.dep_off[dep_S * 28 + t];
        dep_end = d.dep_off[dep_S * 28 + t + 1];
    }
    if constexpr (FAST) {
        for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
            f_cnt[c] = 0;
            f_minpid[c] = 0x7fffffff;
            for (int q = 0; q < 5; q++) f_sum[5 * c + q] = 0;
            f_dp[3 * c] = 0; f_dp[3 * c + 1] = 0; f_dp[3 * c + 2] = 0;
        }
    } else {
        for (int c = threadIdx.x; c < ncell + 1; c += blockDim.x) s_start[c] = 0;
    }
    if (COUNT_SHELL) {
        for (int k = threadIdx.x; k < TCL_CAP; k += blockDim.x) {
            s_shell[k] = 0;
            if (ph.surrogate) {
                s_scnt[k] = 0;
                for (int j = 0; j < SUR_ACC; j++) s_sacc[k][j] = 0;
            }
        }
        zero_axis_masks(s_am[0]);
        if (NEAR_TEST) zero_axis_masks(s_am[1]);
    }
    // gather offsets: an exclusive scan of the 27 list lengths across the
    // lanes of the first wave (the lists in direction order)
    if (threadIdx.x < 64) {
        const int dep_cnt = dep_end - dep_base;
        int incl = dep_cnt;
#pragma unroll
        for (int o = 1; o < 32; o <<= 1) {
            const int y = __shfl_up(incl, o, 64);
            if ((int)threadIdx.x >= o) incl += y;
        }
        if (threadIdx.x < 27) {
            s_src[threadIdx.x] = dep_S;
            s_base[threadIdx.x] = dep_base;
            s_pre[threadIdx.x] = incl - dep_cnt;
        }
        if (threadIdx.x == 26) s_pre[27] = incl;
    }
    __syncthreads();
    int n = s_pre[27];
    if (n > cap) {
        if (threadIdx.x == 0) atomicOr(d.status, BW_FLAG_TILE_OVERFLOW);
        n = cap;
    }
    // listed colloid threadIdx.x: requested before the source-list indices,
    // written to LDS once it is here (while those are in flight)
    const bool cl_mine = COUNT_SHELL && !cl_stale && (int)threadIdx.x < cl_n;
    double cpx = 0, cpy = 0, cpz = 0, crx = 0, cry = 0, crz = 0, cvx = 0, cvy = 0, cvz = 0;
    int ccx = 0, ccy = 0, ccz = 0, cqx = 0, cqy = 0, cqz = 0;
    if (cl_mine) {
        const int g = cl_g;
        cpx = d.cpos[3 * g]; cpy = d.cpos[3 * g + 1]; cpz = d.cpos[3 * g + 2];
        ccx = d.ccell[3 * g]; ccy = d.ccell[3 * g + 1]; ccz = d.ccell[3 * g + 2];
        crx = d.cpos_run[3 * g]; cry = d.cpos_run[3 * g + 1]; crz = d.cpos_run[3 * g + 2];
        cqx = d.ccell_run[3 * g]; cqy = d.ccell_run[3 * g + 1]; cqz = d.ccell_run[3 * g + 2];
        cvx = d.cvel[3 * g]; cvy = d.cvel[3 * g + 1]; cvz = d.cvel[3 * g + 2];
    }
    // gather: the source-list index of each slot (the list holding gather
    // index g is the last t with s_pre[t] <= g, by five fixed halvings).  A
    // slot past the tile's count reads the index and the state of the last
    // particle (of slot 0 of an in-range index for an empty tile): the loads
    // are straight runs, each with one wait after it, and the values are
    // never used
    int rcell[PT], rpid[PT], rkey[PT], rslot[PT];
    Real rox[PT], roy[PT], roz[PT], rvx[PT], rvy[PT], rvz[PT];
    {
        const int g_last = max(n - 1, 0), idx_last = ph.n_tiles * cap - 1;
#pragma unroll
        for (int k = 0; k < PT; k++) {
            const int g = min((int)threadIdx.x + k * TH, g_last);
            int lo = 0;
#pragma unroll
            for (int h = 16; h > 0; h >>= 1)
                if (lo + h <= 26 && s_pre[lo + h] <= g) lo += h;
            rkey[k] = s_src[lo] * cap;  // the source tile's first slot, until the key is known
            rslot[k] = min(rkey[k] + s_base[lo] + (g - s_pre[lo]), idx_last);
        }
#pragma unroll
        for (int k = 0; k < PT; k++) rslot[k] = d.dep_idx[rslot[k]];
#pragma unroll
        for (int k = 0; k < PT; k++) rslot[k] += rkey[k];
    }
#pragma unroll
    for (int k = 0; k < PT; k++) rkey[k] = -1;
    const double Ld = (double)L;
    if (cl_mine) {
        const int q = threadIdx.x;
        tc.g[q] = cl_g;
        tc.px[q] = cpx; tc.py[q] = cpy; tc.pz[q] = cpz;
        tc.cx[q] = ccx; tc.cy[q] = ccy; tc.cz[q] = ccz;
        tc.rx[q] = crx; tc.ry[q] = cry; tc.rz[q] = crz;
        tc.qx[q] = cqx; tc.qy[q] = cqy; tc.qz[q] = cqz;
        tc.vx[q] = cvx; tc.vy[q] = cvy; tc.vz[q] = cvz;
        if (NEAR_TEST) {
            nc[0].cx[q] = (float)img1(cpx - ox0, Ld);
            nc[0].cy[q] = (float)img1(cpy - oy0, Ld);
            nc[0].cz[q] = (float)img1(cpz - oz0, Ld);
            nc[0].rx[q] = (float)img1(crx - ox0, Ld);
            nc[0].ry[q] = (float)img1(cry - oy0, Ld);
            nc[0].rz[q] = (float)img1(crz - oz0, Ld);
        }
    }
    if (COUNT_SHELL && !cl_stale && threadIdx.x == 0) tc.n = cl_n;
#pragma unroll
    for (int k = 0; k < PT; k++) {
        const int slot = rslot[k];
        rcell[k] = src.cell[slot];
        rox[k] = src.ox[slot]; roy[k] = src.oy[slot]; roz[k] = src.oz[slot];
        rvx[k] = src.vx[slot]; rvy[k] = src.vy[slot]; rvz[k] = src.vz[slot];
        rpid[k] = src.pid[slot];
    }
    if constexpr (FAST) {
        // per cell, while the particles are on their way: the random axis
        // (rho = 2U - 1, phi = 2 pi U) and the rotation by theta about it
        // (cpp/rotation_mpcd.cpp:19-25), in FP64 (sqrt(1 - rho^2) cancels for
        // |rho| near 1), rounded once to FP32; the thermostat factor is folded
        // in once the relative energies are known.  A missing draw is marked in
        // the cell's exponent word, which the cell's first pass reads (only a
        // cell of two or more particles needs its draw)
        for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
            const int kx = ox0 + c % B0, ky = oy0 + (c / B0) % B1, kz = oz0 + c / (B0 * B1);
            double u1, u2;
            f_exp[c] = 0;
            if (!draw_axis_raw(d, ph, R, step, linear_cell(kx, ky, kz, L), u1, u2)) {
                f_exp[c] = 1;
                u1 = 0.5; u2 = 0.0;
            }
            const double rho = 2.0 * u1 - 1.0;
            const double phi = 6.283185307179586 * u2;
            const double sq = sqrt(1.0 - rho * rho);
            double sphi, cphi;
            sincos(phi, &sphi, &cphi);
            const double r1 = cphi * sq, r2 = sphi * sq, r3 = rho;
            const double ir1 = ph.ict * r1, ir2 = ph.ict * r2, ir3 = ph.ict * r3;
            const double rm[9] = {ir1 * r1 + ph.ct, ir1 * r2 - ph.st * r3, ir1 * r3 + ph.st * r2,
                                  ir2 * r1 + ph.st * r3, ir2 * r2 + ph.ct, ir2 * r3 - ph.st * r1,
                                  ir3 * r1 - ph.st * r2, ir3 * r2 + ph.st * r1, ir3 * r3 + ph.ct};
            float *m = f_m + 9 * c;
#pragma unroll
            for (int q = 0; q < 9; q++) m[q] = (float)rm[q];
        }
    }
    Region rg = tile_region(R, tg.tx, tg.ty, tg.tz);
    if (COUNT_SHELL && ph.surrogate) {
        const int ns = min(d.tsl_cnt[T], TCL_CAP);
        for (int q = threadIdx.x; q < ns; q += blockDim.x) {
            const int g = d.tsl_list[T * TCL_CAP + q];
            s_sg[q] = g;
            for (int k = 0; k < 3; k++) {
                s_sp[q][k] = d.cpos[3 * g + k];
                s_sp[q][3 + k] = d.cvel[3 * g + k];
            }
            s_spf[q][0] = (float)img1(s_sp[q][0] - ox0, Ld);
            s_spf[q][1] = (float)img1(s_sp[q][1] - oy0, Ld);
            s_spf[q][2] = (float)img1(s_sp[q][2] - oz0, Ld);
        }
        if (threadIdx.x == 0) s_sn = ns;
    }
    if (COUNT_SHELL) {
        if (cl_stale) {
            load_tile_colloids(d, ph, R, tg.r, T, true, tc);  // ends with a barrier
            if (NEAR_TEST)
                for (int q = threadIdx.x; q < tc.n; q += blockDim.x) {
                    nc[0].cx[q] = (float)img1(tc.px[q] - ox0, Ld);
                    nc[0].cy[q] = (float)img1(tc.py[q] - oy0, Ld);
                    nc[0].cz[q] = (float)img1(tc.pz[q] - oz0, Ld);
                    nc[0].rx[q] = (float)img1(tc.rx[q] - ox0, Ld);
                    nc[0].ry[q] = (float)img1(tc.ry[q] - oy0, Ld);
                    nc[0].rz[q] = (float)img1(tc.rz[q] - oz0, Ld);
                }
        } else {
            __syncthreads();
        }
        // masks zeroed before the first barrier
        if (NEAR_TEST) fill_axis_masks2(tc, rg, L, 3, s_am[0], ph.near_reach, s_am[COUNT_SHELL ? 1 : 0]);
        else fill_axis_masks(tc.n, tc.qx, tc.qy, tc.qz, rg, L, 3, s_am[0]);
    }
    // shifted cells and the per-cell histogram
    bool drift = false;
#pragma unroll
    for (int k = 0; k < PT; k++) {
        int g = threadIdx.x + k * TH;
        if (g < n) {
            int ix, iy, iz;
            unpack_cell(rcell[k], ix, iy, iz);
            int lx = shifted_coord(ix, rox[k], sx, L) - ox0;
            int ly = shifted_coord(iy, roy[k], sy, L) - oy0;
            int lz = shifted_coord(iz, roz[k], sz, L) - oz0;
            if (lx < 0 || lx >= B0 || ly < 0 || ly >= B1 || lz < 0 || lz >= B2) {
                drift = true;
                lx = max(0, min(B0 - 1, lx)); ly = max(0, min(B1 - 1, ly)); lz = max(0, min(B2 - 1, lz));
            }
            int key = lx + B0 * (ly + B1 * lz);
            rkey[k] = key;
            if constexpr (FAST) {
                // exact sums: velocities in units of 2^-40, |v|^2 in units of
                // 2^-28 (each particle's terms rounded once, the same whatever
                // the order of the additions; with |v| < 1024 per component and
                // at most 4096 particles a cell's sums stay inside int64).  A
                // component of 1024 or more (or NaN) is a particle that moves
                // over a hundred cells per step, past its tile: flagged
                const float vx = (float)rvx[k], vy = (float)rvy[k], vz = (float)rvz[k];
                const bool ok = fabsf(vx) < 1024.0f && fabsf(vy) < 1024.0f && fabsf(vz) < 1024.0f;
                if (!ok) drift = true;
                const double dx = ok ? (double)vx : 0.0, dy = ok ? (double)vy : 0.0, dz = ok ? (double)vz : 0.0;
                const double e = (dx * dx + dy * dy) + dz * dz;
                atomicAdd(&f_cnt[key], 1);
                atomicMin(&f_minpid[key], rpid[k]);
                atomicAdd(&f_sum[5 * key], (unsigned long long)rint_small(dx * 0x1p40));
                atomicAdd(&f_sum[5 * key + 1], (unsigned long long)rint_small(dy * 0x1p40));
                atomicAdd(&f_sum[5 * key + 2], (unsigned long long)rint_small(dz * 0x1p40));
                atomicAdd(&f_sum[5 * key + 3], (unsigned long long)rint_small(e * 0x1p28));
            } else {
                s_v[g] = rvx[k]; s_v[cap + g] = rvy[k]; s_v[2 * cap + g] = rvz[k];
                atomicAdd(&s_start[key], 1);
            }
        }
    }
    if (drift) atomicOr(d.status, BW_FLAG_DRIFT);
    __syncthreads();
    bool zero_var = false;
    if constexpr (FAST) {
        // per cell (cpp/rotation_mpcd.cpp:14-49).  The mean u = sum v / N in
        // FP64 from the exact sums (the random axis and the rotation R by theta
        // about it, FP32, were formed while the particles were on their way).
        // The relative energy sum |v - u|^2 (which the rotation leaves
        // unchanged) is sum |v|^2 - |sum v|^2 / N from the same sums, wherever
        // that difference is known to 2^-26 of itself: its error is bounded
        // (err, with a factor 2 of margin) by the sums' roundings (2^-29 per
        // |v|^2 term, 2^-41 per velocity term) and the FP64 operations.  A
        // cell where it is not (a cold cell, or one moving fast against its
        // spread) takes the exact second pass: per particle |v - u|^2 in FP64,
        // summed in the fixed point 2^(50 - e), where 2^e exceeds the one-pass
        // value plus twice its error bound and so every term and their sum
        // (each scaled term below 2^50 rounds to an integer by rint_small; the
        // resolution is 2^-50 of the cell's relative energy, whatever |u|).
        // The thermostat factor is folded into M = scale R.
        bool missing = false, second = false;
        for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
            const int N = f_cnt[c];
            if (N < 2) continue;
            if (f_exp[c]) missing = true;
            const double sx = (double)(long long)f_sum[5 * c] * 0x1p-40;
            const double sy = (double)(long long)f_sum[5 * c + 1] * 0x1p-40;
            const double sz = (double)(long long)f_sum[5 * c + 2] * 0x1p-40;
            const double E = (double)(long long)f_sum[5 * c + 3] * 0x1p-28;
            const double iN = 1.0 / N;  // the collision conserves the cell's momentum whatever u's rounding
            f_u[3 * c] = sx * iN; f_u[3 * c + 1] = sy * iN; f_u[3 * c + 2] = sz * iN;
            const double S2 = (sx * sx + sy * sy) + sz * sz;
            const double var = E - S2 * iN;
            const double err = N * 0x1p-28 + 0x1p-50 * E + 0x1p-38 * sqrt(S2) + 0x1p-49 * (S2 * iN);
            if (var > err * 0x1p26) {
                const double scale = sqrt(3.0 * (N - 1) * ph.kbt / (ph.mass_fl * var));
                float *m = f_m + 9 * c;
#pragma unroll
                for (int q = 0; q < 9; q++) m[q] = (float)((double)m[q] * scale);
                f_exp[c] = -1;  // done
            } else {
                int ex;
                frexp(fmax(var, 0.0) + 2.0 * err + 0x1p-40 * E + 0x1p-60, &ex);
                f_exp[c] = 50 - ex;
                second = true;
            }
        }
        if (missing) atomicOr(d.status, BW_FLAG_DRAW_MISSING);
        if (__syncthreads_or(second)) {  // some cell of the tile needs the exact pass
#pragma unroll
            for (int k = 0; k < PT; k++) {
                if (threadIdx.x + k * TH < n) {
                    const int c = rkey[k];
                    if (f_cnt[c] > 1 && f_exp[c] >= 0) {
                        const double ex = (double)rvx[k] - f_u[3 * c], ey = (double)rvy[k] - f_u[3 * c + 1],
                                     ez = (double)rvz[k] - f_u[3 * c + 2];
                        const double q = ldexp((ex * ex + ey * ey) + ez * ez, f_exp[c]);
                        if (q < 0x1p51)  // false for NaN (flagged with the velocity above)
                            atomicAdd(&f_sum[5 * c + 4], (unsigned long long)rint_small(q));
                    }
                }
            }
            __syncthreads();
            for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
                const int N = f_cnt[c];
                if (N < 2 || f_exp[c] < 0) continue;
                const double var = ldexp((double)(long long)f_sum[5 * c + 4], -f_exp[c]);
                double scale = sqrt(3.0 * (N - 1) * ph.kbt / (ph.mass_fl * var));
                if (!(var > 0)) {
                    zero_var = true;
                    scale = 1.0;
                }
                float *m = f_m + 9 * c;
#pragma unroll
                for (int q = 0; q < 9; q++) m[q] = (float)((double)m[q] * scale);
            }
        }
    } else {
    block_exclusive_scan(s_start, ncell, s_wtot);
    if (threadIdx.x == 0) s_start[ncell] = n;
    for (int c = threadIdx.x; c < ncell; c += blockDim.x) s_fill[c] = s_start[c];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PT; k++) {
        int g = threadIdx.x + k * TH;
        if (g < n) s_pk[atomicAdd(&s_fill[rkey[k]], 1)] = rpid[k];
    }
    __syncthreads();
    // rank of each particle in its cell by id: the cell's canonical order
#pragma unroll
    for (int k = 0; k < PT; k++) {
        int g = threadIdx.x + k * TH;
        if (g < n) {
            const int b = s_start[rkey[k]], e = s_start[rkey[k] + 1];
            int rank = 0;
#pragma unroll 4
            for (int j = b; j < e; j++) rank += s_pk[j] < rpid[k] ? 1 : 0;
            s_ord[b + rank] = (unsigned short)g;
        }
    }
    __syncthreads();
    // per cell: FP64 mean in id order (one thread per component) and the
    // rotation axis (a fourth thread)
    bool missing = false;
    for (int w = threadIdx.x; w < 4 * ncell; w += blockDim.x) {
        const int comp = w / ncell, c = w - comp * ncell;
        const int b = s_start[c], e = s_start[c + 1], N = e - b;
        if (N > 1 || (N == 1 && comp < 3 && ph.moments)) {  // one-particle means: moments only
            if (comp < 3) {
                const Real *sv = s_v + comp * cap;
                double u = 0;
#pragma unroll 4
                for (int i = b; i < e; i++) u = u + (double)sv[s_ord[i]] / N;
                s_u[3 * c + comp] = u;
            } else {
                int kx = ox0 + c % B0, ky = oy0 + (c / B0) % B1, kz = oz0 + c / (B0 * B1);
                double u1, u2;
                if (!draw_axis_raw(d, ph, R, step, linear_cell(kx, ky, kz, L), u1, u2)) {
                    missing = true;
                    u1 = 0.5; u2 = 0.0;
                }
                // cpp/rotation_mpcd.cpp:19-25
                double rho = 2 * u1 - 1;
                double phi = 4.0 * asin(1.0) * u2;
                s_ax[3 * c] = cos(phi) * sqrt(1 - rho * rho);
                s_ax[3 * c + 1] = sin(phi) * sqrt(1 - rho * rho);
                s_ax[3 * c + 2] = rho;
            }
        }
    }
    if (missing) atomicOr(d.status, BW_FLAG_DRAW_MISSING);
    __syncthreads();  // the velocities in LDS are no longer needed: energy terms overwrite them
    // per particle: rotated velocity relative to the mean, its squared norm (:37-41)
#pragma unroll
    for (int k = 0; k < PT; k++) {
        int g = threadIdx.x + k * TH;
        if (g < n) {
            const int c = rkey[k];
            if (s_start[c + 1] - s_start[c] > 1) {
                double ux = s_u[3 * c], uy = s_u[3 * c + 1], uz = s_u[3 * c + 2];
                double r1 = s_ax[3 * c], r2 = s_ax[3 * c + 1], r3 = s_ax[3 * c + 2];
                double ir1 = ph.ict * r1, ir2 = ph.ict * r2, ir3 = ph.ict * r3;
                double a11 = ir1 * r1 + ph.ct, a12 = ir1 * r2 - ph.st * r3, a13 = ir1 * r3 + ph.st * r2;
                double a21 = ir2 * r1 + ph.st * r3, a22 = ir2 * r2 + ph.ct, a23 = ir2 * r3 - ph.st * r1;
                double a31 = ir3 * r1 - ph.st * r2, a32 = ir3 * r2 + ph.st * r1, a33 = ir3 * r3 + ph.ct;
                double dx = (double)rvx[k] - ux, dy = (double)rvy[k] - uy, dz = (double)rvz[k] - uz;
                double v1x = ux + ((a11 * dx + a12 * dy) + a13 * dz);
                double v1y = uy + ((a21 * dx + a22 * dy) + a23 * dz);
                double v1z = uz + ((a31 * dx + a32 * dy) + a33 * dz);
                double ex = v1x - ux, ey = v1y - uy, ez = v1z - uz;
                s_e2[g] = (ex * ex + ey * ey) + ez * ez;
            }
        }
    }
    __syncthreads();
    // per cell: variance in id order, thermostat factor
    for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
        const int b = s_start[c], e = s_start[c + 1], N = e - b;
        if (N > 1) {
            double var = 0.0;
#pragma unroll 4
            for (int i = b; i < e; i++) var += s_e2[s_ord[i]];
            double scale = sqrt(3.0 * (N - 1) * ph.kbt / (ph.mass_fl * var));
            if (!(var > 0)) {
                zero_var = true;
                scale = 1.0;
            }
            s_sc[c] = scale;
        }
    }
    }  // FP64 path
    if (zero_var) atomicOr(d.status, BW_FLAG_ZERO_VARIANCE);
    if (ph.moments) {
        // coarse-cell moments (2x2x2 shifted cells) for the surrogate, from the
        // per-cell counts and means: momentum N u, and kinetic energy N |u|^2
        // plus the thermostat's relative energy 3 (N - 1) kT / m, summed in a
        // fixed order (even tile edges, checked at creation)
        const int C0 = B0 / 2, C1 = B1 / 2, C2 = B2 / 2, Lc = L / 2;
        for (int q = threadIdx.x; q < C0 * C1 * C2; q += blockDim.x) {
            const int qx = q % C0, qy = (q / C0) % C1, qz = q / (C0 * C1);
            double n = 0, px = 0, py = 0, pz = 0, en = 0;
            for (int o = 0; o < 8; o++) {
                const int c = (2 * qx + (o & 1)) + B0 * ((2 * qy + ((o >> 1) & 1)) + B1 * (2 * qz + (o >> 2)));
                const int N = FAST ? f_cnt[c] : s_start[c + 1] - s_start[c];
                if (N < 1) continue;
                double ux, uy, uz;
                if constexpr (FAST) {
                    ux = (double)(long long)f_sum[5 * c] * 0x1p-40 / N;
                    uy = (double)(long long)f_sum[5 * c + 1] * 0x1p-40 / N;
                    uz = (double)(long long)f_sum[5 * c + 2] * 0x1p-40 / N;
                } else {
                    ux = s_u[3 * c]; uy = s_u[3 * c + 1]; uz = s_u[3 * c + 2];
                }
                n += N;
                px += N * ux; py += N * uy; pz += N * uz;
                en += N * ((ux * ux + uy * uy) + uz * uz) + (N > 1 ? 3.0 * (N - 1) * ph.kbt / ph.mass_fl : 0.0);
            }
            const long long gq = R.first_coarse + (ox0 / 2 + qx) + (long long)Lc * ((oy0 / 2 + qy) + (long long)Lc * (oz0 / 2 + qz));
            float *mo = d.cmom + 5 * gq;
            mo[0] = (float)n; mo[1] = (float)px; mo[2] = (float)py; mo[3] = (float)pz; mo[4] = (float)en;
        }
        if (threadIdx.x == 0 && T == 0) *d.mom_valid = 1;
    }
    __syncthreads();
    // apply; count run() shells; near test
    unsigned nearbits = 0;
#pragma unroll
    for (int k = 0; k < PT; k++) {
        int g = threadIdx.x + k * TH;
        if (g < n) {
            int c = rkey[k];
            int N = FAST ? f_cnt[c] : s_start[c + 1] - s_start[c];
            if (FAST && N > 1) {
                // v - u in FP64 (then FP32): a cell of small relative velocities
                // multiplies it by a large thermostat factor
                const float *m = f_m + 9 * c;
                const double ux = f_u[3 * c], uy = f_u[3 * c + 1], uz = f_u[3 * c + 2];
                const float vx = (float)rvx[k], vy = (float)rvy[k], vz = (float)rvz[k];
                const float dx = (float)((double)vx - ux), dy = (float)((double)vy - uy), dz = (float)((double)vz - uz);
                const float nx = (float)ux + ((m[0] * dx + m[1] * dy) + m[2] * dz);
                const float ny = (float)uy + ((m[3] * dx + m[4] * dy) + m[5] * dz);
                const float nz = (float)uz + ((m[6] * dx + m[7] * dy) + m[8] * dz);
                // the particle's exact momentum change; the cell's sum, which the
                // collision makes zero up to the roundings, is taken back from
                // the cell's smallest id below
                atomicAdd(&f_dp[3 * c], fixq_diff(nx, vx));
                atomicAdd(&f_dp[3 * c + 1], fixq_diff(ny, vy));
                atomicAdd(&f_dp[3 * c + 2], fixq_diff(nz, vz));
                rvx[k] = (Real)nx; rvy[k] = (Real)ny; rvz[k] = (Real)nz;
            } else if (N > 1) {
                double ux = s_u[3 * c], uy = s_u[3 * c + 1], uz = s_u[3 * c + 2];
                double r1 = s_ax[3 * c], r2 = s_ax[3 * c + 1], r3 = s_ax[3 * c + 2];
                double ir1 = ph.ict * r1, ir2 = ph.ict * r2, ir3 = ph.ict * r3;
                double a11 = ir1 * r1 + ph.ct, a12 = ir1 * r2 - ph.st * r3, a13 = ir1 * r3 + ph.st * r2;
                double a21 = ir2 * r1 + ph.st * r3, a22 = ir2 * r2 + ph.ct, a23 = ir2 * r3 - ph.st * r1;
                double a31 = ir3 * r1 - ph.st * r2, a32 = ir3 * r2 + ph.st * r1, a33 = ir3 * r3 + ph.ct;
                double dx = (double)rvx[k] - ux, dy = (double)rvy[k] - uy, dz = (double)rvz[k] - uz;
                double v1x = ux + ((a11 * dx + a12 * dy) + a13 * dz);
                double v1y = uy + ((a21 * dx + a22 * dy) + a23 * dz);
                double v1z = uz + ((a31 * dx + a32 * dy) + a33 * dz);
                double sc = s_sc[c];
                rvx[k] = (Real)(ux + (v1x - ux) * sc);
                rvy[k] = (Real)(uy + (v1y - uy) * sc);
                rvz[k] = (Real)(uz + (v1z - uz) * sc);
            }
            bool near = !BLOCK, moved = false, shell = false;
            if (COUNT_SHELL) {
                if (ph.surrogate && s_sn) {
                    // fluid that entered a surrogate-driven colloid during the last
                    // step is mirrored back out through its surface, its velocity
                    // reversed relative to the colloid's; its momentum and angular
                    // momentum changes are summed per colloid, and the fluid in the
                    // colloid's shell (on the stored position, the test
                    // k_sur_shell_kick repeats) is counted with its moments of
                    // position: it pays them back with the colloid's sampled impulse
                    // and torque and its run() kick (shell_pay).  An FP32 test in
                    // tile-local coordinates (2e-2 slack) picks the colloids that get
                    // the exact tests.
                    int ix, iy, iz;
                    unpack_cell(rcell[k], ix, iy, iz);
                    unsigned long long cand = 0;
                    {
                        const float Lf = (float)L, iLf = 1.0f / Lf;
                        const float px = (float)pdiff(ix, ox0, L) + (float)rox[k];
                        const float py = (float)pdiff(iy, oy0, L) + (float)roy[k];
                        const float pz = (float)pdiff(iz, oz0, L) + (float)roz[k];
                        const float rs = sqrtf((float)ph.lim_shell) + 0.02f, rs2 = rs * rs;
                        for (int q = 0; q < s_sn; q++) {
                            float ax = px - s_spf[q][0], ay = py - s_spf[q][1], az = pz - s_spf[q][2];
                            if constexpr (WRAP) {
                                ax -= Lf * rintf(ax * iLf); ay -= Lf * rintf(ay * iLf); az -= Lf * rintf(az * iLf);
                            }
                            if ((ax * ax + ay * ay) + az * az <= rs2) cand |= 1ull << q;
                        }
                    }
                    if (cand) {
                        const double Ld = (double)L, R2 = 2.0 * sqrt(ph.lim_coll);  // colloid diameter
                        const double x = (double)ix + (double)rox[k], y = (double)iy + (double)roy[k],
                                     z = (double)iz + (double)roz[k];
                        for (unsigned long long m = cand; m; m &= m - 1) {
                            const int q = __ffsll((long long)m) - 1;
                            double ax = img1(x - s_sp[q][0], Ld), ay = img1(y - s_sp[q][1], Ld), az = img1(z - s_sp[q][2], Ld);
                            double r2 = (ax * ax + ay * ay) + az * az;
                            if (r2 >= ph.lim_coll) continue;
                            // a phantom collides with nothing; a particle deeper than
                            // MIRROR_DEPTH (one that crossed a colloid while it
                            // collided explicitly, or came in faster than the step
                            // allows) is left inside as a phantom; the others are
                            // reflected through the surface, radius rr -> 2R - rr
                            if (rcell[k] & PHANTOM) break;
                            double rr = sqrt(r2);
                            if (rr < 0.5 * R2 - MIRROR_DEPTH) {
                                rcell[k] |= PHANTOM;
                                break;
                            }
                            double sc = (R2 - rr) / rr;
                            double nx = s_sp[q][0] + ax * sc;
                            double ny = s_sp[q][1] + ay * sc, nz = s_sp[q][2] + az * sc;
                            nx = mod1(nx, Ld); ny = mod1(ny, Ld); nz = mod1(nz, Ld);
                            int jx = coord_cell(nx, L), jy = coord_cell(ny, L), jz = coord_cell(nz, L);
                            rcell[k] = pack_cell(jx, jy, jz) | (rcell[k] & PHANTOM);
                            rox[k] = (Real)(nx - jx); roy[k] = (Real)(ny - jy); roz[k] = (Real)(nz - jz);
                            if (!(rox[k] < Real(1))) rox[k] = Real(0);
                            if (!(roy[k] < Real(1))) roy[k] = Real(0);
                            if (!(roz[k] < Real(1))) roz[k] = Real(0);
                            const Real wx = (Real)(2.0 * s_sp[q][3] - (double)rvx[k]);
                            const Real wy = (Real)(2.0 * s_sp[q][4] - (double)rvy[k]);
                            const Real wz = (Real)(2.0 * s_sp[q][5] - (double)rvz[k]);
                            atomicAdd(&s_sacc[q][0], fixq_diff(wx, rvx[k]));
                            atomicAdd(&s_sacc[q][1], fixq_diff(wy, rvy[k]));
                            atomicAdd(&s_sacc[q][2], fixq_diff(wz, rvz[k]));
                            {  // angular momentum change about the centre, at the new position
                                const double dvx = (double)wx - (double)rvx[k], dvy = (double)wy - (double)rvy[k],
                                             dvz = (double)wz - (double)rvz[k];
                                const double px = img1(((double)jx + (double)rox[k]) - s_sp[q][0], Ld),
                                             py = img1(((double)jy + (double)roy[k]) - s_sp[q][1], Ld),
                                             pz = img1(((double)jz + (double)roz[k]) - s_sp[q][2], Ld);
                                atomicAdd(&s_sacc[q][3], fixq(py * dvz - pz * dvy));
                                atomicAdd(&s_sacc[q][4], fixq(pz * dvx - px * dvz));
                                atomicAdd(&s_sacc[q][5], fixq(px * dvy - py * dvx));
                            }
                            rvx[k] = wx; rvy[k] = wy; rvz[k] = wz;
                            moved = true;  // streamed per substep
                            break;
                        }
                        // a mirrored particle is tested against every listed colloid
                        if (moved) cand = s_sn >= 64 ? ~0ull : ((1ull << s_sn) - 1ull);
                        unpack_cell(rcell[k], ix, iy, iz);
                        const double x1 = (double)ix + (double)rox[k], y1 = (double)iy + (double)roy[k],
                                     z1 = (double)iz + (double)roz[k];
                        for (unsigned long long m = cand; m; m &= m - 1) {
                            const int q = __ffsll((long long)m) - 1;
                            double ax = img1(x1 - s_sp[q][0], Ld), ay = img1(y1 - s_sp[q][1], Ld), az = img1(z1 - s_sp[q][2], Ld);
                            if ((ax * ax + ay * ay) + az * az <= ph.lim_shell) {
                                atomicAdd(&s_scnt[q], 1);
                                atomicAdd(&s_sacc[q][6], fixq(ax));
                                atomicAdd(&s_sacc[q][7], fixq(ay));
                                atomicAdd(&s_sacc[q][8], fixq(az));
                                atomicAdd(&s_sacc[q][9], fixq(ax * ax));
                                atomicAdd(&s_sacc[q][10], fixq(ay * ay));
                                atomicAdd(&s_sacc[q][11], fixq(az * az));
                                atomicAdd(&s_sacc[q][12], fixq(ax * ay));
                                atomicAdd(&s_sacc[q][13], fixq(ax * az));
                                atomicAdd(&s_sacc[q][14], fixq(ay * az));
                                shell = true;  // kicked in the first substep
                            }
                        }
                    }
                }
                int ix, iy, iz;
                unpack_cell(rcell[k], ix, iy, iz);
                // the FP64 position, formed only for a colloid that passes the
                // mask and FP32 tests
                const Real sox = rox[k], soy = roy[k], soz = roz[k];
                auto shell_of = [&](int q) {
                    return shell_test(ph, tc, q, L, ix, iy, iz, (double)ix + (double)sox, (double)iy + (double)soy,
                                      (double)iz + (double)soz);
                };
                unsigned m = cell_mask(s_am[0], rg, ix, iy, iz, L);
                if (NEAR_TEST) {
                    // conservative test in FP32, tile-local coordinates (2e-2 cell slack)
                    const float Lf = (float)L, iLf = 1.0f / Lf;
                    const float px = (float)pdiff(ix, ox0, L) + (float)rox[k];
                    const float py = (float)pdiff(iy, oy0, L) + (float)roy[k];
                    const float pz = (float)pdiff(iz, oz0, L) + (float)roz[k];
                    const float vx = (float)rvx[k], vy = (float)rvy[k], vz = (float)rvz[k];
                    // the hardware square root (1 ulp): both uses below carry slack
                    const float tr = __builtin_amdgcn_sqrtf((vx * vx + vy * vy) + vz * vz) * (float)ph.ndt;
                    const bool inside01 = in_unit(rox[k]) && in_unit(roy[k]) && in_unit(roz[k]);
                    near = !(tr <= (float)ph.vcap_ndt) || !inside01;
                    const float rs = sqrtf((float)ph.near_shell2) + 0.02f, rs2 = rs * rs;
                    const float rc = (float)ph.near_base + tr + 0.02f, rc2 = rc * rc;
                    auto d2 = [&](float cx, float cy, float cz) {
                        float ax = px - cx, ay = py - cy, az = pz - cz;
                        if constexpr (WRAP) {
                            ax -= Lf * rintf(ax * iLf); ay -= Lf * rintf(ay * iLf); az -= Lf * rintf(az * iLf);
                        }
                        return (ax * ax + ay * ay) + az * az;
                    };
                    // run() shells: the FP32 test selects the candidates that get the
                    // exact test (it accepts every particle the exact test accepts)
                    // and makes them near
                    for_candidates(tc, m, [&](int q) {
                        if (q < tc.n && d2(nc[0].rx[q], nc[0].ry[q], nc[0].rz[q]) <= rs2) {
                            near = true;
                            if (shell_of(q)) atomicAdd(&s_shell[q], 1);
                        }
                    });
                    if (!near) {  // within reach of a collision during the step
                        unsigned m2 = cell_mask(s_am[COUNT_SHELL ? 1 : 0], rg, ix, iy, iz, L);
                        for_candidates(tc, m2, [&](int q) {
                            if (q < tc.n && !near) near = d2(nc[0].cx[q], nc[0].cy[q], nc[0].cz[q]) <= rc2;
                        });
                    }
                } else {
                    for_candidates(tc, m, [&](int q) {
                        if (q < tc.n && shell_of(q)) atomicAdd(&s_shell[q], 1);
                    });
                }
                if (ph.surrogate) {  // inside a colloid that was a phantom in the last step
                    const double Ld = (double)L;
                    for_candidates(tc, m, [&](int q) {
                        if (q >= tc.n || !d.cmode_prev[tc.g[q]]) return;
                        const double x = (double)ix + (double)sox, y = (double)iy + (double)soy, z = (double)iz + (double)soz;
                        double ax = img1(x - tc.px[q], Ld), ay = img1(y - tc.py[q], Ld), az = img1(z - tc.pz[q], Ld);
                        if ((ax * ax + ay * ay) + az * az <= ph.lim_coll) rcell[k] |= PHANTOM;
                    });
                }
            }
            if (near || moved || shell) nearbits |= 1u << k;
        }
    }
    // output slots: near particles first, each group in gather order
    int nrk[PT];  // near particles of lower gather index
    int n_near = n;
    if (BLOCK) {
        const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const unsigned long long lt = (1ull << lane) - 1ull;
        unsigned long long bal[PT];
#pragma unroll
        for (int k = 0; k < PT; k++) {
            bal[k] = __ballot((threadIdx.x + k * TH < n) && (nearbits & (1u << k)));
            if (lane == 0) s_cc[k][w] = __popcll(bal[k]);
        }
        __syncthreads();
        if (COUNT_SHELL) {  // the apply phase's shell counts are complete
            for (int q = threadIdx.x; q < tc.n; q += blockDim.x)
                if (s_shell[q]) atomicAdd(&d.ccnt[tc.g[q]], s_shell[q]);
            if (ph.surrogate) flush_surrogate_shells(d, s_sn, s_sg, s_scnt, s_sacc);
        }
        int base = 0;
#pragma unroll
        for (int k = 0; k < PT; k++) {
            int before = 0, tot = 0;
            for (int v = 0; v < TH / 64; v++) {
                int cnt = s_cc[k][v];
                before += v < w ? cnt : 0;
                tot += cnt;
            }
            nrk[k] = base + before + __popcll(bal[k] & lt);
            base += tot;
        }
        n_near = base;
    }
    if constexpr (FAST) {
        // per cell: the rounded collision's momentum residual (exact) is taken
        // from the particle of smallest id, so that the cell's momentum
        // changes by that particle's one rounding only (cpp/rotation_mpcd.cpp
        // conserves it exactly)
        if (!BLOCK) __syncthreads();
#pragma unroll
        for (int k = 0; k < PT; k++) {
            if (threadIdx.x + k * TH < n) {
                const int c = rkey[k];
                if (f_cnt[c] > 1 && rpid[k] == f_minpid[c]) {
                    rvx[k] = (Real)((double)rvx[k] - (double)(long long)f_dp[3 * c] * 0x1p-40);
                    rvy[k] = (Real)((double)rvy[k] - (double)(long long)f_dp[3 * c + 1] * 0x1p-40);
                    rvz[k] = (Real)((double)rvz[k] - (double)(long long)f_dp[3 * c + 2] * 0x1p-40);
                }
            }
        }
    }
    // The near prefix is published in chunks of 64 for the per-substep pass;
    // the chunk range is requested here and written after the far particles'
    // streaming, which hides the atomic's round trip.
    const bool publish = COUNT_SHELL && BLOCK && (threadIdx.x >> 6) == 0;
    int chunk0 = 0;
    if (publish && (threadIdx.x & 63) == 0 && n_near > 0) chunk0 = atomicAdd(d.chunk_cnt, (n_near + 63) >> 6);
    bool far_drift = false;
    if (BLOCK) {
        // far particles fly straight through the whole step.  FP64 storage:
        // every substep as the streaming kernel does it (normalise() is the
        // identity while an offset is in [0,1)).  FP32 storage: x + v ndt in
        // one rounding instead of n_sub (within a few FP32 ulps of the
        // substep sum, and closer to the exact flight).  A far particle moves
        // less than a cell in the step, so each axis crosses a cell face at
        // most once and the periodic wrap of the cell index comes once, after.
#pragma unroll
        for (int k = 0; k < PT; k++) {
            if (threadIdx.x + k * TH < n && !(nearbits & (1u << k))) {
                int ix, iy, iz;
                unpack_cell(rcell[k], ix, iy, iz);
                Real ox = rox[k], oy = roy[k], oz = roz[k];
                if constexpr (FAST) {
                    const float tn = (float)ph.ndt;
                    ox = ox + rvx[k] * tn;
                    oy = oy + rvy[k] * tn;
                    oz = oz + rvz[k] * tn;
                    normalise_nowrap(ix, ox);
                    normalise_nowrap(iy, oy);
                    normalise_nowrap(iz, oz);
                } else {
                    const Real dt = (Real)ph.dt;
                    const Real dx = rvx[k] * dt, dy = rvy[k] * dt, dz = rvz[k] * dt;
#pragma unroll 2
                    for (int l = 0; l < ph.n_sub; l++) {
                        ox = ox + dx;
                        oy = oy + dy;
                        oz = oz + dz;
                        if (!in_unit3(ox, oy, oz)) {
                            normalise_nowrap(ix, ox);
                            normalise_nowrap(iy, oy);
                            normalise_nowrap(iz, oz);
                        }
                    }
                }
                ix = wrap_once(ix, L); iy = wrap_once(iy, L); iz = wrap_once(iz, L);
                rcell[k] = pack_cell(ix, iy, iz) | (rcell[k] & PHANTOM);
                rox[k] = ox; roy[k] = oy; roz[k] = oz;
                // no reach check: the particle's shifted cell lies in the tile box
                // (binning outside it is flagged above) and differs from its cell
                // by floor(o + s) in {-1, 0, 1} (offset in [0,1), shift in
                // [-1/2, 1/2)); with less than a cell of travel it ends within two
                // cells of the box, inside the TILE_REACH the next step needs
            }
        }
    }
    if (publish) {
        chunk0 = lane_value(chunk0, 0);
        for (int j = threadIdx.x; j < (n_near + 63) >> 6; j += 64) d.chunk_tile[chunk0 + j] = (T << 6) | j;
    }
#pragma unroll
    for (int k = 0; k < PT; k++) {
        int g = threadIdx.x + k * TH;
        if (g < n) {
            int slot_in_tile = g;
            if (BLOCK) slot_in_tile = (nearbits & (1u << k)) ? nrk[k] : n_near + g - nrk[k];
            int o = T * cap + slot_in_tile;
            dst.cell[o] = rcell[k];
            dst.ox[o] = rox[k]; dst.oy[o] = roy[k]; dst.oz[o] = roz[k];
            dst.vx[o] = rvx[k]; dst.vy[o] = rvy[k]; dst.vz[o] = rvz[k];
            dst.pid[o] = rpid[k];
            if (MAP) {
                int c = rkey[k];
                int kx = ox0 + c % B0, ky = oy0 + (c / B0) % B1, kz = oz0 + c / (B0 * B1);
                d.shifted_cell[R.first_fluid + rpid[k]] = linear_cell(kx, ky, kz, L);
            }
        }
    }
    if (far_drift) atomicOr(d.status, BW_FLAG_DRIFT);
    if (threadIdx.x == 0) {
        d.count_next[T] = n;
        d.near_cnt[T] = n_near;
    }
    if (COUNT_SHELL && !BLOCK) {
        __syncthreads();
        for (int q = threadIdx.x; q < tc.n; q += blockDim.x)
            if (s_shell[q]) atomicAdd(&d.ccnt[tc.g[q]], s_shell[q]);
        if (ph.surrogate) flush_surrogate_shells(d, s_sn, s_sg, s_scnt, s_sacc);
    }
}

// ------------------------------------------------------------------ stream
// Per tile: [run() kick on the rear shells, cpp/run.cpp:21-24]
//           [streaming, cpp/update.cpp:56-60, with exact int-cell binning]
//           [collision detection, cpp/fluid_colloid_collision.cpp:32-35]
// Candidate colloids of a particle come from the tile's LDS stencil masks.
// DIRECT: no masks (null pointers); every listed colloid is a candidate and
// the cheap cell tests come first -- for a few dozen particles per tile.
template <class Real, bool KICK, bool STREAM, bool DETECT, class TC, bool DIRECT = false>
struct StreamOp {
    const Phys &ph;
    const Dev &d;
    Fluid<Real> &f;
    const TC &tc;
    const Region &rg;
    const AxisMask *mdet, *mkick;
    int L, r, lox, loy, loz, lnx, lny, lnz;
    bool drift, ch_over;
    // DIRECT: the listed colloids' current and run() positions relative to
    // the tile box's origin (bx0, by0, bz0), FP32, for a conservative
    // pre-test (2e-2 cell slack) ahead of the exact one; wrap when the box is
    // small enough for the periodic image to matter
    const NearColl *nf = nullptr;
    int bx0 = 0, by0 = 0, bz0 = 0;
    bool wrap = true;

    __device__ __forceinline__ float d2f(float px, float py, float pz, float cx, float cy, float cz) const {
        float ax = px - cx, ay = py - cy, az = pz - cz;
        if (wrap) {
            const float Lf = (float)L, iLf = 1.0f / Lf;
            ax -= Lf * rintf(ax * iLf); ay -= Lf * rintf(ay * iLf); az -= Lf * rintf(az * iLf);
        }
        return (ax * ax + ay * ay) + az * az;
    }

    __device__ __forceinline__ unsigned all_listed() const {
        return tc.n >= MASK_BITS ? 0xffffffffu : ((1u << tc.n) - 1u);
    }
    __device__ __forceinline__ void operator()(int slot, int c, Real ox, Real oy, Real oz, Real vx, Real vy, Real vz) {
        int ix, iy, iz;
        unpack_cell(c, ix, iy, iz);
        const int phantom = c & PHANTOM;
        const double Ld = (double)L;
        if constexpr (KICK) {
            unsigned m = DIRECT ? all_listed() : cell_mask(*mkick, rg, ix, iy, iz, L);
            if (m || tc.n > MASK_BITS) {
                double x = (double)ix + (double)ox, y = (double)iy + (double)oy, z = (double)iz + (double)oz;
                double wx = vx, wy = vy, wz = vz;
                bool hit = false;
                const float kpx = (float)pdiff(ix, bx0, L) + (float)ox, kpy = (float)pdiff(iy, by0, L) + (float)oy,
                            kpz = (float)pdiff(iz, bz0, L) + (float)oz;
                const float rs = sqrtf((float)ph.lim_shell) + 0.02f, rs2 = rs * rs;
                auto kick = [&](int k) {
                    if (DIRECT && nf && k < tc.n && d2f(kpx, kpy, kpz, nf->rx[k], nf->ry[k], nf->rz[k]) > rs2) return;
                    if (k < tc.n && shell_test(ph, tc, k, L, ix, iy, iz, x, y, z)) {
                        wx = wx - tc.kx[k];
                        wy = wy - tc.ky[k];
                        wz = wz - tc.kz[k];
                        hit = true;
                    }
                };
                if constexpr (DIRECT) for_listed(tc, kick);
                else for_candidates(tc, m, kick);
                if (hit) {
                    vx = (Real)wx; vy = (Real)wy; vz = (Real)wz;
                    f.vx[slot] = vx; f.vy[slot] = vy; f.vz[slot] = vz;
                }
            }
        }
        if (STREAM) {
            const Real dt = (Real)ph.dt;
            ox = ox + vx * dt;
            oy = oy + vy * dt;
            oz = oz + vz * dt;
            normalise(ix, ox, L);
            normalise(iy, oy, L);
            normalise(iz, oz, L);
            f.cell[slot] = pack_cell(ix, iy, iz) | phantom;
            f.ox[slot] = ox; f.oy[slot] = oy; f.oz[slot] = oz;
            if (!(near_interval(ix, lox, lnx, L) && near_interval(iy, loy, lny, L) && near_interval(iz, loz, lnz, L)))
                drift = true;
        }
        if (DETECT) {
            unsigned m = DIRECT ? all_listed() : cell_mask(*mdet, rg, ix, iy, iz, L);
            if (m == 0 && tc.n <= MASK_BITS) {
                if (phantom) f.cell[slot] = pack_cell(ix, iy, iz);
                return;
            }
            double x = (double)ix + (double)ox, y = (double)iy + (double)oy, z = (double)iz + (double)oz;
            int n_in = 0, gin = -1;
            bool inside_any = false;  // by distance alone (the stencil skips a colloid's own cell)
            const float dpx = (float)pdiff(ix, bx0, L) + (float)ox, dpy = (float)pdiff(iy, by0, L) + (float)oy,
                        dpz = (float)pdiff(iz, bz0, L) + (float)oz;
            const float rr = sqrtf((float)ph.lim_coll) + 0.02f, rr2 = rr * rr;
            auto test = [&](int k) {
                if (k >= tc.n) return;
                if (DIRECT) {
                    if (nf) {  // within the FP32 radius of the colloid (the exact test follows)
                        if (d2f(dpx, dpy, dpz, nf->cx[k], nf->cy[k], nf->cz[k]) > rr2) return;
                    } else {  // inside a colloid: within 3 cells of its cell on every axis
                        const int ax = pdiff(ix, tc.cx[k], L), ay = pdiff(iy, tc.cy[k], L), az = pdiff(iz, tc.cz[k], L);
                        if (ax < -3 || ax > 3 || ay < -3 || ay > 3 || az < -3 || az > 3) return;
                    }
                }
                double rx = img1(tc.px[k] - x, Ld), ry = img1(tc.py[k] - y, Ld), rz = img1(tc.pz[k] - z, Ld);
                if ((rx * rx + ry * ry) + rz * rz <= ph.lim_coll) {
                    inside_any = true;
                    if (in_stencil(ix, iy, iz, tc.cx[k], tc.cy[k], tc.cz[k], L)) {
                        n_in++;
                        gin = tc.g[k];
                    }
                }
            };
            if constexpr (DIRECT) for_listed(tc, test);
            else for_candidates(tc, m, test);
            if (phantom) {  // no events until it is inside no colloid
                if (!inside_any) f.cell[slot] = pack_cell(ix, iy, iz);
                return;
            }
            if (n_in > 0) {
                int e = n_in == 1 ? atomicAdd(&d.ev_cnt[gin], 1) : EV_CAP;
                if (e < EV_CAP) {
                    d.ev_slot[gin * EV_CAP + e] = slot;
                } else {  // inside several colloids, or beyond a full bucket: serial replay
                    int e2 = atomicAdd(&d.chain_cnt[r], 1);
                    if (e2 < CHAIN_CAP) d.chain_slot[r * CHAIN_CAP + e2] = slot;
                    else ch_over = true;
                }
            }
        }
    }
};

// near_only: the fused step's per-substep pass over the particles stored
// first in each tile by the SRD kernel (a few per cent of a tile: one wave per
// tile, so that every tile of a large box is resident at once); otherwise
// every particle.  Only the kicking pass stages the colloids' run() data.
template <class Real, bool KICK, bool STREAM, bool DETECT, int THREADS>
__global__ __launch_bounds__(THREADS) void k_stream(Phys ph, Dev d, Fluid<Real> f, int near_only) {
    using TC = typename std::conditional<KICK, TileColl, TileCore>::type;
    __shared__ TC tc;
    __shared__ AxisMask s_am[(KICK && DETECT) ? 2 : 1];
    const int T = blockIdx.x;
    Rep R;
    TileGeo tg = tile_geo(d, T, R);
    const int L = tg.L;
    const int n = min(near_only ? d.near_cnt[T] : d.count[T], ph.cap);
    Region rg = tile_region(R, tg.tx, tg.ty, tg.tz);
    const AxisMask *mkick = &s_am[0], *mdet = &s_am[(KICK && DETECT) ? 1 : 0];
    if (KICK || DETECT) {
        load_tile_colloids(d, ph, R, tg.r, T, KICK, tc);
        if constexpr (KICK) build_axis_masks(tc.n, tc.qx, tc.qy, tc.qz, rg, L, 3, s_am[0]);
        if (DETECT) build_axis_masks(tc.n, tc.cx, tc.cy, tc.cz, rg, L, 3, s_am[(KICK && DETECT) ? 1 : 0]);
    }
    StreamOp<Real, KICK, STREAM, DETECT, TC> op{ph, d, f, tc, rg, mdet, mkick, L, tg.r,
                                                tg.tx * R.B[0] - TILE_REACH, tg.ty * R.B[1] - TILE_REACH,
                                                tg.tz * R.B[2] - TILE_REACH, R.B[0] + 2 * TILE_REACH,
                                                R.B[1] + 2 * TILE_REACH, R.B[2] + 2 * TILE_REACH, false, false};
    const int base_slot = T * ph.cap;
    // two particles per thread per iteration: all loads first
    for (int i0 = threadIdx.x; i0 < n; i0 += 2 * THREADS) {
        const int i1 = i0 + THREADS;
        const bool has1 = i1 < n;
        const int s0 = base_slot + i0, s1 = base_slot + (has1 ? i1 : i0);
        int c0 = f.cell[s0], c1 = f.cell[s1];
        Real ox0 = f.ox[s0], oy0 = f.oy[s0], oz0 = f.oz[s0], vx0 = f.vx[s0], vy0 = f.vy[s0], vz0 = f.vz[s0];
        Real ox1 = f.ox[s1], oy1 = f.oy[s1], oz1 = f.oz[s1], vx1 = f.vx[s1], vy1 = f.vy[s1], vz1 = f.vz[s1];
        op(s0, c0, ox0, oy0, oz0, vx0, vy0, vz0);
        if (has1) op(s1, c1, ox1, oy1, oz1, vx1, vy1, vz1);
    }
    if (op.drift) atomicOr(d.status, BW_FLAG_DRIFT);
    if (op.ch_over) atomicOr(d.status, BW_FLAG_CHAIN_OVERFLOW);
}

// The fused step's per-substep pass over the near particles (stored first in
// each tile by the SRD kernel; a few per cent of the fluid, most of it around
// the colloids).  The SRD kernel lists them in chunks of 64 (tile, chunk);
// every wave of a fixed grid takes chunks in turn, stages its tile's colloids
// and masks and handles one particle per lane, so the work is spread evenly
// whatever the clustering.  Waves synchronise only within themselves.
template <class Real, bool KICK>
__global__ __launch_bounds__(256, KICK ? 1 : 8) void k_stream_near(Phys ph, Dev d, Fluid<Real> f) {
    static_assert(TCL_CAP == 64, "one lane per listed colloid");
    using TC = typename std::conditional<KICK, TileColl, TileCore>::type;
    __shared__ TC s_tc[4];
    __shared__ NearColl s_nf[4];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    TC &tc = s_tc[w];
    // the argument fields of the loop, fetched in one batch of scalar loads
    // rather than one dependent round trip per first use
    asm volatile("" ::"s"(d.chunk_tile), "s"(d.tile_rep), "s"(d.tcl_cnt), "s"(d.tcl_list), "s"(d.near_cnt),
                 "s"(d.rep), "s"(d.tcl_stale), "s"(d.cpos), "s"(d.ccell), "s"(d.status), "s"(d.ev_cnt),
                 "s"(d.ev_slot), "s"(d.chain_cnt), "s"(d.chain_slot));
    asm volatile("" ::"s"(f.cell), "s"(f.ox), "s"(f.oy), "s"(f.oz), "s"(f.vx), "s"(f.vy), "s"(f.vz), "s"(ph.cap),
                 "s"(ph.dt), "s"(ph.lim_coll));
    const int nchunk = ld_const(d.chunk_cnt);
    const int stride = gridDim.x * 4;
    bool drift = false, ch_over = false;
    int ch = blockIdx.x * 4 + w;
    int e = ch < nchunk ? ld_const(d.chunk_tile + ch) : 0;
    for (; ch < nchunk; ch += stride) {
        const int T = e >> 6, i = ((e & 63) << 6) + lane;
        if (ch + stride < nchunk) e = ld_const(d.chunk_tile + ch + stride);  // next chunk, ahead of its use
        const int r = ld_const(d.tile_rep + T);
        int cnt = min(ld_const(d.tcl_cnt + T), TCL_CAP);  // a stale list is rebuilt below
        int gl = d.tcl_list[T * TCL_CAP + lane];
        const int n = min(ld_const(d.near_cnt + T), ph.cap);
        const int slot = T * ph.cap + i;
        int c0 = 0;
        Real ox = 0, oy = 0, oz = 0, vx = 0, vy = 0, vz = 0;
        if (i < n) {
            c0 = f.cell[slot];
            ox = f.ox[slot]; oy = f.oy[slot]; oz = f.oz[slot];
            vx = f.vx[slot]; vy = f.vy[slot]; vz = f.vz[slot];
        }
        wave_sync();  // the previous chunk's lanes are done with the LDS lists
        // the listed colloids, loaded before the list is known to be current
        // (it is rebuilt and loaded again when a colloid left its margin)
        if (lane < cnt) load_tile_slot(d, ph, gl, lane, KICK, tc);
        const Rep R = ld_const(d.rep + r);
        const int L = R.L;
        if (ld_const(d.tcl_stale + r)) {
            wave_sync();
            cnt = build_tile_list_wave(ph, d, R, T, tc.g);
            gl = tc.g[lane];
            if (lane < cnt) load_tile_slot(d, ph, gl, lane, KICK, tc);
        }
        if (lane == 0) tc.n = cnt;
        const int lt = T - R.first_tile;
        const int tx = lt % R.nt[0], ty = (lt / R.nt[0]) % R.nt[1], tz = lt / (R.nt[0] * R.nt[1]);
        const int bx0 = tx * R.B[0], by0 = ty * R.B[1], bz0 = tz * R.B[2];
        if (lane < cnt) {  // this lane's colloid, relative to the tile box's origin
            const double Ld = (double)L;
            s_nf[w].cx[lane] = (float)img1(tc.px[lane] - bx0, Ld);
            s_nf[w].cy[lane] = (float)img1(tc.py[lane] - by0, Ld);
            s_nf[w].cz[lane] = (float)img1(tc.pz[lane] - bz0, Ld);
            if constexpr (KICK) {
                s_nf[w].rx[lane] = (float)img1(tc.rx[lane] - bx0, Ld);
                s_nf[w].ry[lane] = (float)img1(tc.ry[lane] - by0, Ld);
                s_nf[w].rz[lane] = (float)img1(tc.rz[lane] - bz0, Ld);
            }
        }
        wave_sync();
        const Region rg = tile_region(R, tx, ty, tz);
        StreamOp<Real, KICK, true, true, TC, true> op{ph, d, f, tc, rg, nullptr, nullptr, L, r,
                                                      bx0 - TILE_REACH, by0 - TILE_REACH,
                                                      bz0 - TILE_REACH, R.B[0] + 2 * TILE_REACH,
                                                      R.B[1] + 2 * TILE_REACH, R.B[2] + 2 * TILE_REACH, false, false,
                                                      &s_nf[w], bx0, by0, bz0, ph.near_wrap != 0};
        if (i < n) op(slot, c0, ox, oy, oz, vx, vy, vz);
        drift = drift || op.drift;
        ch_over = ch_over || op.ch_over;
    }
    if (drift) atomicOr(d.status, BW_FLAG_DRIFT);
    if (ch_over) atomicOr(d.status, BW_FLAG_CHAIN_OVERFLOW);
}

// --------------------------------------------------------------- collisions
// One event: back half-step, stochastic reflection with the colloid's rigid
// body velocity, forward half-step (cpp/fluid_colloid_collision.cpp:36-43).
// (bx,by,bz) is the offset in the binned cell (ix,iy,iz), updated in place.
__device__ inline void reflect_event(const Phys &ph, int L, int ix, int iy, int iz, double &bx, double &by,
                                     double &bz, double &vx, double &vy, double &vz, const double *pc,
                                     const double *vcol, const double *wcol, const ReflDraw &dr, double *rs,
                                     bool fp32) {
    const double Ld = (double)L, dt = ph.dt;
    bx = bx - vx * dt * 0.5;
    by = by - vy * dt * 0.5;
    bz = bz - vz * dt * 0.5;
    double ax = (double)ix + bx, ay = (double)iy + by, az = (double)iz + bz;
    double rsx = img1(ax - pc[0], Ld), rsy = img1(ay - pc[1], Ld), rsz = img1(az - pc[2], Ld);
    double ux, uy, uz;
    stochastic_reflection(ph, Ld, ax, ay, az, rsx, rsy, rsz, dr, ux, uy, uz);
    double nx = ux + vcol[0] + (wcol[1] * rsz - wcol[2] * rsy);
    double ny = uy + vcol[1] + (wcol[2] * rsx - wcol[0] * rsz);
    double nz = uz + vcol[2] + (wcol[0] * rsy - wcol[1] * rsx);
    if (fp32) {  // what the velocity store keeps
        nx = (double)(float)nx; ny = (double)(float)ny; nz = (double)(float)nz;
    }
    rs[0] = rsx; rs[1] = rsy; rs[2] = rsz;
    vx = nx; vy = ny; vz = nz;
    bx = bx + vx * dt * 0.5;
    by = by + vy * dt * 0.5;
    bz = bz + vz * dt * 0.5;
}

// momentum and angular-momentum transfer of one event: dump - vel and
// rs x (dump - vel) (cpp/fluid_colloid_collision.cpp:40-42)
__device__ __forceinline__ void transfer(const double *dump, double vx, double vy, double vz, const double *rs,
                                         double *c) {
    double ex = dump[0] - vx, ey = dump[1] - vy, ez = dump[2] - vz;
    c[0] = ex; c[1] = ey; c[2] = ez;
    c[3] = rs[1] * ez - rs[2] * ey;
    c[4] = rs[2] * ex - rs[0] * ez;
    c[5] = rs[0] * ey - rs[1] * ex;
}

// One wave per colloid g: its pair list for this substep, then the particles
// inside g and no other colloid, ranked in the reference's neigh_fl order
// (stencil position, then id) and reflected; the momentum/torque sums are taken
// in that order.  A particle a reflection moved into a later colloid continues
// the reference's colloid loop serially (k_serial).  The pair partners j > g
// (position and cell) are staged in LDS while the pair list is built, so the
// continuation test of every event reads no global memory.
struct PairColl {
    double px, py, pz;
    int cx, cy, cz;
};

template <class Real>
__global__ __launch_bounds__(COLLIDE_THREADS) void k_collide(Phys ph, Dev d, Fluid<Real> f, int substep,
                                                             const int2 *wgt) {
    __shared__ double s_c[COLLIDE_THREADS / 64][EV_CAP][6];
    __shared__ PairColl s_nb[COLLIDE_THREADS / 64][MD_CAP];
    __shared__ int s_ne[COLLIDE_THREADS / 64], s_nnb[COLLIDE_THREADS / 64];
    const bool fp32 = sizeof(Real) == 4;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int2 wt = ph.n_rep == 1 ? make_int2(0, blockIdx.x * (COLLIDE_THREADS / 64)) : wgt[blockIdx.x];
    const int r = wt.x;
    const int g = wt.y + w;
    // Loads that do not need the replica record go first, at an in-range
    // index (a workgroup's last colloids may lie past its replica).  Entries
    // of the candidate list and of the event bucket past their counts are
    // stale but valid indices, so the candidates' positions and the events'
    // particles are requested with the counts and discarded when unused.
    const int gs = min(g, ph.n_colloid_total - 1);
    const int mdc_n = d.mdc_cnt[gs], ne_s = d.ev_cnt[gs];
    const int js = d.mdc_list[gs * MDC_CAP + lane], slot_s = d.ev_slot[gs * EV_CAP + lane];
    double pg[3], pj[3], vg[3], wg[3];
    int cg[3], cj[3];
    for (int k = 0; k < 3; k++) {
        pg[k] = d.cpos[3 * gs + k];
        pj[k] = d.cpos[3 * js + k];
        vg[k] = d.cvel[3 * gs + k];
        wg[k] = d.comega[3 * gs + k];
        cg[k] = d.ccell[3 * gs + k];
        cj[k] = d.ccell[3 * js + k];
    }
    const int pid_s = f.pid[slot_s], cell_s = f.cell[slot_s];
    const Real pox_s = f.ox[slot_s], poy_s = f.oy[slot_s], poz_s = f.oz[slot_s];
    const Real pvx_s = f.vx[slot_s], pvy_s = f.vy[slot_s], pvz_s = f.vz[slot_s];
    const Rep R = d.rep[r];
    const int L = R.L;
    const double Ld = (double)L;
    const bool valid = g < R.first_colloid + R.nc;
    const long long step = *d.step;
    const bool full = d.md_stale[r] != 0;
    const int ncand = valid ? (full ? R.nc : mdc_n) : 0;
    int ne = 0;
    if (valid) {
        ne = ne_s;
        if (lane == 0) {
            d.ev_cnt[g] = 0;
            if (ne > EV_CAP) {
                int k = atomicAdd(&d.ovf_cnt[r], 1);
                if (k < OVF_CAP) d.ovf_list[r * OVF_CAP + k] = g;
                else atomicOr(d.status, BW_FLAG_EVENT_OVERFLOW);
            } else if (ne > 0 && ph.draws == BW_DRAWS_INJECTED) {
                atomicAdd(&d.ev_total[r], ne);
            }
        }
        if (ne > EV_CAP) ne = 0;  // full bucket: all its events are replayed serially
    }
    int slot = -1, pid = 0, ix = 0, iy = 0, iz = 0;
    long long key = 0x7fffffffffffffffll;
    Real pox = 0, poy = 0, poz = 0, pvx = 0, pvy = 0, pvz = 0;
    if (lane < ne) {
        slot = slot_s;
        pid = pid_s;
        unpack_cell(cell_s, ix, iy, iz);
        pox = pox_s; poy = poy_s; poz = poz_s;
        pvx = pvx_s; pvy = pvy_s; pvz = pvz_s;
    }
    if (valid) {  // pair list of g (cpp/neighbour_list.cpp:27-39), partners j > g into LDS
        const unsigned long long below = (1ull << lane) - 1ull;
        int nmd = 0, nnb = 0;
        for (int base = 0; base < ncand; base += 64) {
            const int m = base + lane;
            int j = -1;
            bool p = false;
            const bool pre = base == 0 && !full;  // this lane's candidate is js, loaded above
            if (m < ncand) {
                j = pre ? js : (full ? R.first_colloid + m : d.mdc_list[g * MDC_CAP + m]);
                p = j != g && (pre ? md_pair_at(ph, Ld, g, j, pg, pj) : md_pair(ph, d, Ld, g, j));
            }
            const unsigned long long b = __ballot(p);
            const int rank = __popcll(b & below);
            if (p && nmd + rank < MD_CAP) {
                const int e = g * MD_CAP + nmd + rank;
                d.md_list[e] = j;
                // the pair's force addend, summed in list order by k_colloid_update
                double pjv[3], t[3];
                for (int k = 0; k < 3; k++) pjv[k] = pre ? pj[k] : d.cpos[3 * j + k];
                force_addend(ph, Ld, g, j, pg, pjv, t);
                for (int k = 0; k < 3; k++) d.md_f[3 * e + k] = t[k];
            }
            nmd += __popcll(b);
            const bool q = p && j > g && !(ph.surrogate && d.cmode[j]);
            const unsigned long long bq = __ballot(q);
            const int rq = __popcll(bq & below);
            if (q && nnb + rq < MD_CAP) {
                PairColl &e = s_nb[w][nnb + rq];
                if (pre) {
                    e.px = pj[0]; e.py = pj[1]; e.pz = pj[2];
                    e.cx = cj[0]; e.cy = cj[1]; e.cz = cj[2];
                } else {
                    e.px = d.cpos[3 * j]; e.py = d.cpos[3 * j + 1]; e.pz = d.cpos[3 * j + 2];
                    e.cx = d.ccell[3 * j]; e.cy = d.ccell[3 * j + 1]; e.cz = d.ccell[3 * j + 2];
                }
            }
            nnb += __popcll(bq);
        }
        if (lane == 0) {
            if (nmd > MD_CAP) { atomicOr(d.status, BW_FLAG_LIST_OVERFLOW); nmd = MD_CAP; }
            d.md_cnt[g] = nmd;
            s_nnb[w] = min(nnb, MD_CAP);
        }
    }
    wave_sync();  // the wave's pair partners and list visible to its lanes (LDS is per wave)
    if (lane < ne) {
        int dx = pdiff(ix, cg[0], L), dy = pdiff(iy, cg[1], L), dz = pdiff(iz, cg[2], L);
        key = ((long long)stencil_index(dx, dy, dz) << 32) | (unsigned)pid;
    }
    int rank = 0;
    for (int m = 0; m < ne; m++) {
        long long km = lane_value(key, m);
        rank += (km < key) ? 1 : 0;
    }
    if (lane < ne) {
        double bx = pox, by = poy, bz = poz;
        double vx = pvx, vy = pvy, vz = pvz;
        ReflDraw dr;
        double cont[6] = {0, 0, 0, 0, 0, 0};
        const double dump[3] = {vx, vy, vz};
        if (draw_reflection(d, ph, R, step, substep, g, pid, dr)) {
            double rs[3];
            reflect_event(ph, L, ix, iy, iz, bx, by, bz, vx, vy, vz, pg, vg, wg, dr, rs, fp32);
            transfer(dump, vx, vy, vz, rs, cont);
            f.ox[slot] = (Real)bx; f.oy[slot] = (Real)by; f.oz[slot] = (Real)bz;
            f.vx[slot] = (Real)vx; f.vy[slot] = (Real)vy; f.vz[slot] = (Real)vz;
            // The reference goes on to test this particle, now moved, against
            // every later colloid (cpp/fluid_colloid_collision.cpp:30-35).  A
            // later colloid that now contains it is a pair partner of g; if
            // there is one, the rest of the colloid loop is replayed serially.
            const double ax = (double)ix + bx, ay = (double)iy + by, az = (double)iz + bz;
            const int nnb = s_nnb[w];
#pragma unroll 2
            for (int m = 0; m < nnb; m++) {
                const PairColl &e = s_nb[w][m];
                if (!in_stencil(ix, iy, iz, e.cx, e.cy, e.cz, L)) continue;
                double rx = img1(e.px - ax, Ld), ry = img1(e.py - ay, Ld), rz = img1(e.pz - az, Ld);
                if ((rx * rx + ry * ry) + rz * rz <= ph.lim_coll) {
                    int c = atomicAdd(&d.cont_cnt[r], 1);
                    if (c < CONT_CAP) {
                        int q = r * CONT_CAP + c;
                        d.cont_slot[q] = slot;
                        d.cont_from[q] = g + 1;
                        for (int kk = 0; kk < 3; kk++) d.cont_dump[3 * q + kk] = dump[kk];
                    } else {
                        atomicOr(d.status, BW_FLAG_CHAIN_OVERFLOW);
                    }
                    break;
                }
            }
        } else {
            atomicOr(d.status, BW_FLAG_DRAW_MISSING);
        }
        for (int k = 0; k < 6; k++) s_c[w][rank][k] = cont[k];
    }
    if (lane == 0) s_ne[w] = ne;
    wave_sync();
    if (lane == 0 && valid) {
        double a[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 2
        for (int e = 0; e < s_ne[w]; e++)
            for (int k = 0; k < 6; k++) a[k] = a[k] + s_c[w][e][k];
        for (int k = 0; k < 3; k++) {
            d.cdv[3 * g + k] = a[k];
            d.cdw[3 * g + k] = a[3 + k];
        }
    }
}

// One workgroup per replica, after k_collide: the serial list -- particles
// inside several colloids and the events of overflowed buckets (tested from
// the replica's first colloid on, dump = current velocity), and the
// continuations (from the colloid after the reflecting one, dump = velocity
// before that reflection) -- replayed in ascending particle id, each particle
// through the colloids whose stencil holds its binned cell in ascending order
// (cpp/fluid_colloid_collision.cpp:29-49).  Then the injected-draw totals are
// checked and the replica's counters cleared.
template <class Real>
__global__ __launch_bounds__(COLLIDE_THREADS) void k_serial(Phys ph, Dev d, Fluid<Real> f, int substep) {
    __shared__ int s_chain[SERIAL_CAP], s_src[SERIAL_CAP];
    __shared__ int s_nser, s_cand_n, s_serial_events, s_inj_tot;
    __shared__ int s_cand[TCL_CAP];
    __shared__ int s_wt[COLLIDE_THREADS / 64];
    const bool fp32 = sizeof(Real) == 4;
    const int r = blockIdx.x;
    const Rep R = d.rep[r];
    const int L = R.L;
    const double Ld = (double)L;
    const int gend = R.first_colloid + R.nc;
    const long long step = *d.step;
    const int nchain = d.chain_cnt[r], ncont = d.cont_cnt[r], novf = min(d.ovf_cnt[r], OVF_CAP);
    if (threadIdx.x == 0) {
        s_serial_events = 0;
        s_inj_tot = 0;
        s_nser = 0;
    }
    __syncthreads();
    if (nchain > 0 || ncont > 0 || novf > 0) {
        if (threadIdx.x == 0) {
            int n = 0;
            int nch = min(nchain, CHAIN_CAP);
            if (nchain > CHAIN_CAP) atomicOr(d.status, BW_FLAG_CHAIN_OVERFLOW);
            for (int i = 0; i < nch && n < SERIAL_CAP; i++, n++) { s_chain[n] = d.chain_slot[r * CHAIN_CAP + i]; s_src[n] = -1; }
            int nco = min(ncont, CONT_CAP);
            if (ncont > CONT_CAP) atomicOr(d.status, BW_FLAG_CHAIN_OVERFLOW);
            for (int i = 0; i < nco && n < SERIAL_CAP; i++, n++) { s_chain[n] = d.cont_slot[r * CONT_CAP + i]; s_src[n] = r * CONT_CAP + i; }
            for (int i = 0; i < novf; i++) {
                int gg = d.ovf_list[r * OVF_CAP + i];
                for (int e = 0; e < EV_CAP && n < SERIAL_CAP; e++, n++) { s_chain[n] = d.ev_slot[gg * EV_CAP + e]; s_src[n] = -1; }
            }
            if (n >= SERIAL_CAP) atomicOr(d.status, BW_FLAG_EVENT_OVERFLOW);
            for (int i = 1; i < n; i++) {
                int s = s_chain[i], q = s_src[i], p = f.pid[s], j = i - 1;
                while (j >= 0 && f.pid[s_chain[j]] > p) {
                    s_chain[j + 1] = s_chain[j];
                    s_src[j + 1] = s_src[j];
                    j--;
                }
                s_chain[j + 1] = s;
                s_src[j + 1] = q;
            }
            s_nser = n;
        }
        __syncthreads();
        const int nser = s_nser;
        for (int i = 0; i < nser; i++) {
            const int s = s_chain[i], q = s_src[i];
            int cx, cy, cz;
            unpack_cell(f.cell[s], cx, cy, cz);
            // colloids whose stencil holds the particle's binned cell, ascending
            const int from = q >= 0 ? d.cont_from[q] : R.first_colloid;
            if (threadIdx.x == 0) s_cand_n = 0;
            __syncthreads();
            for (int base = from; base < gend; base += blockDim.x) {
                int gg = base + threadIdx.x;
                bool p = gg < gend && !(ph.surrogate && d.cmode[gg]) &&
                         in_stencil(cx, cy, cz, d.ccell[3 * gg], d.ccell[3 * gg + 1], d.ccell[3 * gg + 2], L);
                int tot;
                int idx = block_compact(p, s_wt, tot);
                if (p && s_cand_n + idx < TCL_CAP) s_cand[s_cand_n + idx] = gg;
                __syncthreads();
                if (threadIdx.x == 0) {
                    // more candidate colloids than the list holds: the replay
                    // would miss collisions, so the capacity error is raised
                    if (s_cand_n + tot > TCL_CAP) atomicOr(d.status, BW_FLAG_LIST_OVERFLOW);
                    s_cand_n = min(s_cand_n + tot, TCL_CAP);
                }
                __syncthreads();
            }
            if (threadIdx.x == 0) {
                const int p = f.pid[s];
                double bx = f.ox[s], by = f.oy[s], bz = f.oz[s];
                double vx = f.vx[s], vy = f.vy[s], vz = f.vz[s];
                double dump[3] = {vx, vy, vz};
                if (q >= 0)
                    for (int kk = 0; kk < 3; kk++) dump[kk] = d.cont_dump[3 * q + kk];
                for (int m = 0; m < s_cand_n; m++) {
                    const int gg = s_cand[m];
                    const double *pc = d.cpos + 3 * gg;
                    double ax = (double)cx + bx, ay = (double)cy + by, az = (double)cz + bz;
                    double rx = img1(pc[0] - ax, Ld), ry = img1(pc[1] - ay, Ld), rz = img1(pc[2] - az, Ld);
                    if ((rx * rx + ry * ry) + rz * rz > ph.lim_coll) continue;
                    ReflDraw dr;
                    if (!draw_reflection(d, ph, R, step, substep, gg, p, dr)) {
                        atomicOr(d.status, BW_FLAG_DRAW_MISSING);
                        continue;
                    }
                    double rs[3], c6[6];
                    reflect_event(ph, L, cx, cy, cz, bx, by, bz, vx, vy, vz, pc, d.cvel + 3 * gg, d.comega + 3 * gg,
                                  dr, rs, fp32);
                    transfer(dump, vx, vy, vz, rs, c6);
                    for (int kk = 0; kk < 3; kk++) {
                        d.cdv[3 * gg + kk] += c6[kk];
                        d.cdw[3 * gg + kk] += c6[3 + kk];
                    }
                    s_serial_events++;
                }
                f.ox[s] = (Real)bx; f.oy[s] = (Real)by; f.oz[s] = (Real)bz;
                f.vx[s] = (Real)vx; f.vy[s] = (Real)vy; f.vz[s] = (Real)vz;
            }
            __syncthreads();
        }
    }
    if (ph.draws == BW_DRAWS_INJECTED) {
        // every GPU event found its (colloid, particle) tuple; equal totals
        // then mean equal sets
        int my_inj = 0;
        for (int i = threadIdx.x; i < R.nc; i += blockDim.x) {
            int base = substep * ph.n_colloid_total + R.first_colloid + i;
            my_inj += d.inj_off[base + 1] - d.inj_off[base];
        }
        atomicAdd(&s_inj_tot, my_inj);
        __syncthreads();
        if (threadIdx.x == 0 && s_inj_tot != d.ev_total[r] + s_serial_events)
            atomicOr(d.status, BW_FLAG_DRAW_MISMATCH);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        d.chain_cnt[r] = 0;
        d.cont_cnt[r] = 0;
        d.ovf_cnt[r] = 0;
        d.ev_total[r] = 0;
    }
}

// ----------------------------------------------------- colloid update (fused)
// One wave per colloid: collision impulse, orientation, WCA force (the pair
// addends k_collide stored with this substep's pair list, summed in list
// order by lane 0), velocity, [run() kick], old force, position into the
// other buffer.  Follows k_collide of the same substep.
__global__ __launch_bounds__(256) void k_colloid_update(Phys ph, Dev d, unsigned flags, const int2 *wgt) {
    __shared__ double s_f[4][MD_CAP][3];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int2 wt = ph.n_rep == 1 ? make_int2(0, blockIdx.x * 4) : wgt[blockIdx.x];
    const int r = wt.x;
    const int g = wt.y + w;
    // Loads that do not need the replica record go first, at an in-range
    // index (a workgroup's last colloids may lie past its replica); the first
    // 64 force addends are requested before their count arrives.
    const int gs = min(g, ph.n_colloid_total - 1);
    ColIn in;  // lane 0: the colloid's own state, in flight with the force terms
    if (lane == 0) colloid_load(ph, d, gs, flags, in);
    int nmd_s = 0;
    double tl[3] = {0, 0, 0};
    if (flags & PH_FORCE) {
        nmd_s = d.md_cnt[gs];
        for (int k = 0; k < 3; k++) tl[k] = d.md_f[3 * (gs * MD_CAP + lane) + k];
    }
    // lanes 0-5: sin (lanes 0-2) or cos (lanes 3-5) of the activity rotation
    // angle about axis lane % 3, from the angular velocity after the impulse,
    // formed as colloid_update forms them; lane 0 gathers them below
    double tv = 0.0;
    if ((flags & PH_ACTIVITY) && lane < 6) {
        const int k = lane % 3;
        double wk = d.comega[3 * gs + k];
        if (flags & PH_APPLY) {
            const double iw = (ph.surrogate && d.cmode[gs]) ? d.csur[6 * gs + 3 + k] : d.cdw[3 * gs + k];
            wk = wk + iw * ph.mass_fl / ph.I_colloid;
        }
        const double b = wk * ph.dt;
        tv = lane < 3 ? sin(b) : cos(b);
    }
    const Rep R = d.rep[r];
    const bool valid = g < R.first_colloid + R.nc;
    int nmd = 0;
    if (valid && (flags & PH_FORCE)) {
        nmd = nmd_s;
        for (int m = lane; m < nmd; m += 64)
            for (int k = 0; k < 3; k++) s_f[w][m][k] = m < 64 ? tl[k] : d.md_f[3 * (g * MD_CAP + m) + k];
    }
    double trig[6];
    for (int k = 0; k < 6; k++) trig[k] = __shfl(tv, k);
    __syncthreads();
    if (valid && lane == 0) {
        double fn[3] = {0, 0, 0};
        if (flags & PH_FORCE) {
            double f0 = 0, f1 = 0, f2 = 0;
#pragma unroll 4
            for (int m = 0; m < nmd; m++) {  // cpp/update.cpp:28 in list order
                f0 = f0 + s_f[w][m][0];
                f1 = f1 + s_f[w][m][1];
                f2 = f2 + s_f[w][m][2];
            }
            fn[0] = f0; fn[1] = f1; fn[2] = f2;
        }
        colloid_update(ph, d, R.L, r, g, flags, fn, in, (flags & PH_ACTIVITY) ? trig : nullptr);
    }
}

// ------------------------------------------------------- run() (standalone)
template <class Real>
__global__ __launch_bounds__(STREAM_THREADS) void k_run_count(Phys ph, Dev d, Fluid<Real> f) {
    __shared__ TileColl tc;
    __shared__ int s_shell[TCL_CAP];
    __shared__ AxisMask s_am;
    const int T = blockIdx.x;
    Rep R;
    TileGeo tg = tile_geo(d, T, R);
    const int L = tg.L;
    const int n = min(d.count[T], ph.cap);
    for (int k = threadIdx.x; k < TCL_CAP; k += blockDim.x) s_shell[k] = 0;
    Region rg = tile_region(R, tg.tx, tg.ty, tg.tz);
    load_tile_colloids(d, ph, R, tg.r, T, true, tc);
    build_axis_masks(tc, rg, L, true, 3, s_am);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int slot = T * ph.cap + i;
        int ix, iy, iz;
        unpack_cell(f.cell[slot], ix, iy, iz);
        unsigned m = cell_mask(s_am, rg, ix, iy, iz, L);
        double x = (double)ix + (double)f.ox[slot], y = (double)iy + (double)f.oy[slot],
               z = (double)iz + (double)f.oz[slot];
        for_candidates(tc, m, [&](int q) {
            if (q < tc.n && shell_test(ph, tc, q, L, ix, iy, iz, x, y, z)) atomicAdd(&s_shell[q], 1);
        });
    }
    __syncthreads();
    for (int q = threadIdx.x; q < tc.n; q += blockDim.x)
        if (s_shell[q]) atomicAdd(&d.ccnt[tc.g[q]], s_shell[q]);
}

// -------------------------------------------------------------------- maps
template <class Real>
__global__ void k_map_count(Phys ph, Dev d, Fluid<Real> f) {
    const int T = blockIdx.x;
    Rep R;
    tile_geo(d, T, R);
    const int n = min(d.count[T], ph.cap);
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        atomicAdd(&d.map_count[R.first_cell + linear_of_packed(f.cell[T * ph.cap + i], R.L)], 1);
}
template <class Real>
__global__ void k_map_fill(Phys ph, Dev d, Fluid<Real> f) {
    const int T = blockIdx.x;
    Rep R;
    tile_geo(d, T, R);
    const int n = min(d.count[T], ph.cap);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        int slot = T * ph.cap + i;
        int pos = atomicAdd(&d.map_fill[R.first_cell + linear_of_packed(f.cell[slot], R.L)], 1);
        d.map_list[pos] = f.pid[slot];
    }
}
// ascending id within each cell (the reference's box_part order)
__global__ void k_map_sort(Dev d, long long ncells) {
    long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    int b = d.map_start[c], e = b + d.map_count[c];
    for (int i = b + 1; i < e; i++) {
        int v = d.map_list[i], j = i - 1;
        while (j >= b && d.map_list[j] > v) {
            d.map_list[j + 1] = d.map_list[j];
            j--;
        }
        d.map_list[j + 1] = v;
    }
}
// exclusive scan, three passes (blocks of 1024)
__global__ void k_scan_block(const int *in, int *out, int *bsum, long long n) {
    __shared__ int s[1024];
    __shared__ int s_wtot[16];
    long long i = (long long)blockIdx.x * 1024 + threadIdx.x;
    s[threadIdx.x] = i < n ? in[i] : 0;
    __syncthreads();
    int tot = block_exclusive_scan(s, 1024, s_wtot);
    if (i < n) out[i] = s[threadIdx.x];
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}
__global__ void k_scan_sums(int *bsum, int nb) {
    __shared__ int s_wtot[16];
    __shared__ int s[1024];
    int carry = 0;
    for (int base = 0; base < nb; base += 1024) {
        int i = base + threadIdx.x;
        s[threadIdx.x] = i < nb ? bsum[i] : 0;
        __syncthreads();
        int tot = block_exclusive_scan(s, 1024, s_wtot);
        if (i < nb) bsum[i] = s[threadIdx.x] + carry;
        carry += tot;
        __syncthreads();
    }
}
__global__ void k_scan_add(int *out, const int *bsum, long long n) {
    long long i = (long long)blockIdx.x * 1024 + threadIdx.x;
    if (i < n) out[i] += bsum[blockIdx.x];
}
// neigh_fl of colloid g: stencil cells in the reference order (dx fastest,
// -3..3, centre skipped), each cell's particles in ascending id
__global__ void k_neigh_fl(Phys ph, Dev d, const int *colloid_rep) {
    __shared__ int s_off[343];
    __shared__ int s_wtot[16];
    const int g = blockIdx.x;
    const int r = colloid_rep[g];
    const Rep R = d.rep[r];
    const int L = R.L;
    const int cx = d.ccell[3 * g], cy = d.ccell[3 * g + 1], cz = d.ccell[3 * g + 2];
    for (int k = threadIdx.x; k < 342; k += blockDim.x) {
        int i = k < 171 ? k : k + 1;
        int dx = i % 7 - 3, dy = (i / 7) % 7 - 3, dz = i / 49 - 3;
        int c = linear_cell(wrapi(cx + dx, L), wrapi(cy + dy, L), wrapi(cz + dz, L), L);
        s_off[k] = d.map_count[R.first_cell + c];
    }
    __syncthreads();
    int tot = block_exclusive_scan(s_off, 342, s_wtot);
    if (threadIdx.x == 0) {
        if (tot > d.neigh_cap) atomicOr(d.status, BW_FLAG_NEIGH_OVERFLOW);
        d.neigh_cnt[g] = min(tot, d.neigh_cap);
    }
    for (int k = threadIdx.x; k < 342; k += blockDim.x) {
        int i = k < 171 ? k : k + 1;
        int dx = i % 7 - 3, dy = (i / 7) % 7 - 3, dz = i / 49 - 3;
        int c = linear_cell(wrapi(cx + dx, L), wrapi(cy + dy, L), wrapi(cz + dz, L), L);
        long long gc = R.first_cell + c;
        int b = d.map_start[gc], cnt = d.map_count[gc];
        for (int e = 0; e < cnt; e++) {
            int o = s_off[k] + e;
            if (o < d.neigh_cap) d.neigh_list[(long long)g * d.neigh_cap + o] = d.map_list[b + e];
        }
    }
}

// ------------------------------------------------------------ device start-up
// cpp/initialize.cpp:62-83 with counter-based draws keyed by (replica seed,
// particle id): particle i draws uniform positions in the box (attempt a in
// the counter) until none of its replica's colloids lies closer than sigma/2
// (the colloids of the 27 coarse bins around it: a skipped colloid is farther
// than a bin edge >= sigma/2); then its velocity, uniform in +-sqrt(12 kT/m)/2
// per component (the reference's law) or Maxwell-Boltzmann (Box-Muller).
__device__ __forceinline__ int rep_of_particle(const Dev &d, int n_rep, long long i) {
    int lo = 0, hi = n_rep - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (d.rep[mid].first_fluid <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
__global__ void k_init_place(Phys ph, Dev d, InitBins b, int vel_law, long long n_total, double *pos, double *vel) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_total) return;
    const int r = rep_of_particle(d, ph.n_rep, i);
    const Rep R = d.rep[r];
    const long long id = i - R.first_fluid;
    const double L = R.L, rad = 0.5 * ph.sigma;
    const int G = b.G[r];  // 1 (every colloid in one bin) or >= 3
    const double h = L / G;
    double t[3] = {0, 0, 0};
    bool placed = false;
    for (uint32_t a = 0; a < 100000u && !placed; a++) {
        const uint4 u = philox(make_uint4(KIND_INIT_POS | (a << 8), (uint32_t)id, (uint32_t)(id >> 32), 0u), rep_key(R));
        const uint4 w = philox(make_uint4(KIND_INIT_POS | (a << 8), (uint32_t)id, (uint32_t)(id >> 32), 1u), rep_key(R));
        t[0] = u01(u.x, u.y) * L; t[1] = u01(u.z, u.w) * L; t[2] = u01(w.x, w.y) * L;
        placed = true;
        if (R.nc > 0) {
            int bx = min((int)(t[0] / h), G - 1), by = min((int)(t[1] / h), G - 1), bz = min((int)(t[2] / h), G - 1);
            const int span = G < 3 ? 0 : 1;
            for (int dz = -span; dz <= span && placed; dz++)
                for (int dy = -span; dy <= span && placed; dy++)
                    for (int dx = -span; dx <= span && placed; dx++) {
                        const int X = G < 3 ? 0 : (bx + dx + G) % G, Y = G < 3 ? 0 : (by + dy + G) % G,
                                  Z = G < 3 ? 0 : (bz + dz + G) % G;
                        const long long bin = b.first[r] + X + (long long)G * (Y + (long long)G * Z);
                        for (int q = b.start[bin]; q < b.start[bin + 1]; q++) {
                            const double *c = b.cpos + 3 * b.list[q];
                            const double ax = img1(t[0] - c[0], L), ay = img1(t[1] - c[1], L), az = img1(t[2] - c[2], L);
                            if (sqrt((ax * ax + ay * ay) + az * az) < rad) { placed = false; break; }
                        }
                    }
        }
    }
    if (!placed) atomicOr(d.status, BW_FLAG_DRIFT);
    const uint4 u = philox(make_uint4(KIND_INIT_VEL, (uint32_t)id, (uint32_t)(id >> 32), 0u), rep_key(R));
    const uint4 w = philox(make_uint4(KIND_INIT_VEL, (uint32_t)id, (uint32_t)(id >> 32), 1u), rep_key(R));
    const double q[4] = {u01(u.x, u.y), u01(u.z, u.w), u01(w.x, w.y), u01(w.z, w.w)};
    double v[3];
    if (vel_law == BW_VEL_MAXWELL) {
        const uint4 e = philox(make_uint4(KIND_INIT_VEL, (uint32_t)id, (uint32_t)(id >> 32), 2u), rep_key(R));
        const double q2[2] = {u01(e.x, e.y), u01(e.z, e.w)};
        const double s = sqrt(ph.kbt / ph.mass_fl), tp = 6.283185307179586;
        const double r1 = sqrt(-2.0 * log(1.0 - q[0])), r2 = sqrt(-2.0 * log(1.0 - q[2]));
        v[0] = s * r1 * cos(tp * q[1]);
        v[1] = s * r1 * sin(tp * q[1]);
        v[2] = s * r2 * cos(tp * q2[0]);
    } else {
        const double vs = sqrt(12.0 * ph.kbt / ph.mass_fl);
        for (int k = 0; k < 3; k++) v[k] = (q[k] - 0.5) * vs;
    }
    for (int k = 0; k < 3; k++) {
        pos[3 * i + k] = t[k];
        vel[3 * i + k] = v[k];
    }
}
// per replica (one workgroup each): mean velocity, summed in a fixed order
// per replica mean velocity in two fixed-order levels: workgroup (p, r) sums
// particles [p MEAN_CHUNK, (p+1) MEAN_CHUNK) of replica r (thread-strided, then
// a tree), k_init_mean_final adds the chunks in order
__global__ __launch_bounds__(256) void k_init_mean_part(Dev d, const double *vel, int nchunk, double *part) {
    __shared__ double s[256][3];
    const int r = blockIdx.y, p = blockIdx.x;
    const Rep R = d.rep[r];
    const long long i0 = (long long)p * MEAN_CHUNK, i1 = min((long long)R.nf, i0 + MEAN_CHUNK);
    double a[3] = {0, 0, 0};
    for (long long i = i0 + threadIdx.x; i < i1; i += 256)
        for (int k = 0; k < 3; k++) a[k] += vel[3 * (R.first_fluid + i) + k];
    for (int k = 0; k < 3; k++) s[threadIdx.x][k] = a[k];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int k = 0; k < 3; k++) s[threadIdx.x][k] += s[threadIdx.x + h][k];
        __syncthreads();
    }
    if (threadIdx.x < 3) part[3 * ((long long)r * nchunk + p) + threadIdx.x] = s[0][threadIdx.x];
}
__global__ void k_init_mean_final(Dev d, const double *part, int nchunk, double *mean) {
    const int r = blockIdx.x, k = threadIdx.x;
    if (k >= 3) return;
    const Rep R = d.rep[r];
    double a = 0;
    for (int p = 0; p < nchunk; p++) a += part[3 * ((long long)r * nchunk + p) + k];
    mean[3 * r + k] = R.nf > 0 ? a / (double)R.nf : 0.0;
}
// mean removed; cell + offset; appended to its tile of buffer `f` (any order)
template <class Real>
__global__ void k_init_bin(Phys ph, Dev d, long long n_total, const double *pos, const double *vel,
                           const double *mean, Fluid<Real> f, int *fill) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_total) return;
    const int r = rep_of_particle(d, ph.n_rep, i);
    const Rep R = d.rep[r];
    int ic[3];
    Real o[3];
    for (int k = 0; k < 3; k++) {
        const double x = pos[3 * i + k];
        long long q = (long long)floor(x);
        Real ro = (Real)(x - (double)q);
        if (ro >= Real(1)) { ro = Real(0); q += 1; }
        q %= R.L;
        if (q < 0) q += R.L;
        ic[k] = (int)q;
        o[k] = ro;
    }
    const int T = R.first_tile + ic[0] / R.B[0] + R.nt[0] * (ic[1] / R.B[1] + R.nt[1] * (ic[2] / R.B[2]));
    const int e = atomicAdd(&fill[T], 1);
    if (e >= ph.cap) { atomicOr(d.status, BW_FLAG_TILE_OVERFLOW); return; }
    const size_t s = (size_t)T * ph.cap + e;
    f.cell[s] = pack_cell(ic[0], ic[1], ic[2]);
    f.ox[s] = o[0]; f.oy[s] = o[1]; f.oz[s] = o[2];
    f.vx[s] = (Real)(vel[3 * i] - mean[3 * r]);
    f.vy[s] = (Real)(vel[3 * i + 1] - mean[3 * r + 1]);
    f.vz[s] = (Real)(vel[3 * i + 2] - mean[3 * r + 2]);
    f.pid[s] = (int)(i - R.first_fluid);
}
// each tile's slots in ascending particle id (the upload's layout), from
// buffer `a` into buffer `b`
template <class Real>
__global__ __launch_bounds__(512) void k_init_sort(Phys ph, Dev d, const int *fill, Fluid<Real> a, Fluid<Real> b,
                                                   int *count) {
    __shared__ int s_pid[MAX_CAP];
    const int T = blockIdx.x;
    const int n = min(fill[T], ph.cap);
    const size_t base = (size_t)T * ph.cap;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s_pid[i] = a.pid[base + i];
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int p = s_pid[i];
        int rank = 0;
        for (int j = 0; j < n; j++) rank += s_pid[j] < p ? 1 : 0;
        const size_t src = base + i, dst = base + rank;
        b.cell[dst] = a.cell[src];
        b.ox[dst] = a.ox[src]; b.oy[dst] = a.oy[src]; b.oz[dst] = a.oz[src];
        b.vx[dst] = a.vx[src]; b.vy[dst] = a.vy[src]; b.vz[dst] = a.vz[src];
        b.pid[dst] = p;
    }
    if (threadIdx.x == 0) count[T] = n;
}

// ------------------------------------------------------------ updown_velocity
// cpp/run.cpp:28-50 on colloid g's neigh_fl list (built by neighbour_list_mpcd),
// with positions and velocities gathered by id: the fluid behind the colloid
// (r . v_c <= 0) within sigma/2 + 0.5 (rear shell) and sigma/2 + 0.1 (near
// surface), compacted in list order, and their velocity sums taken in that
// order.  out per colloid: [cnt, up_cnt, sum v rear (3), sum v near (3)].
__global__ __launch_bounds__(256) void k_updown(Phys ph, Dev d, const double *pos, const double *vel, double *out) {
    __shared__ int s_rear[1024], s_up[1024];
    __shared__ int s_wtot[16];
    __shared__ int s_nr, s_nu;
    const int g = blockIdx.x;
    const Rep R = d.rep[d.crep[g]];
    const double Ld = (double)R.L;
    const double lim_up = (0.5 * ph.sigma + 0.1) * (0.5 * ph.sigma + 0.1);
    const double *pc = d.cpos + 3 * g, *vc = d.cvel + 3 * g;
    const int n = d.neigh_cnt[g];
    if (threadIdx.x == 0) { s_nr = 0; s_nu = 0; }
    __syncthreads();
    for (int base = 0; base < n; base += blockDim.x) {
        const int i = base + threadIdx.x;
        bool rear = false, up = false;
        int id = 0;
        if (i < n) {
            id = d.neigh_list[(long long)g * d.neigh_cap + i];
            const double *p = pos + 3 * (R.first_fluid + id);
            const double ax = img1(p[0] - pc[0], Ld), ay = img1(p[1] - pc[1], Ld), az = img1(p[2] - pc[2], Ld);
            const double d2 = (ax * ax + ay * ay) + az * az, dot = (ax * vc[0] + ay * vc[1]) + az * vc[2];
            rear = d2 <= ph.lim_shell && dot <= 0.0;
            up = d2 <= lim_up && dot <= 0.0;
        }
        int tr, tu;
        const int kr = block_compact(rear, s_wtot, tr);
        if (rear && s_nr + kr < 1024) s_rear[s_nr + kr] = id;
        const int ku = block_compact(up, s_wtot, tu);
        if (up && s_nu + ku < 1024) s_up[s_nu + ku] = id;
        __syncthreads();
        if (threadIdx.x == 0) { s_nr += tr; s_nu += tu; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (s_nr > 1024 || s_nu > 1024) atomicOr(d.status, BW_FLAG_NEIGH_OVERFLOW);
        double vr[3] = {0, 0, 0}, vu[3] = {0, 0, 0};
        for (int j = 0; j < min(s_nr, 1024); j++)
            for (int k = 0; k < 3; k++) vr[k] += vel[3 * (R.first_fluid + s_rear[j]) + k];
        for (int j = 0; j < min(s_nu, 1024); j++)
            for (int k = 0; k < 3; k++) vu[k] += vel[3 * (R.first_fluid + s_up[j]) + k];
        double *o = out + 8 * g;
        o[0] = s_nr; o[1] = s_nu;
        for (int k = 0; k < 3; k++) { o[2 + k] = vr[k]; o[5 + k] = vu[k]; }
    }
}
// The reference's per-colloid results in colloid order (its up_vel is never
// reset between colloids; both divisions are guarded by up_cnt > 0).
__global__ void k_updown_finish(Dev d, const double *sums, int *cnt, int *up_cnt, double *vel, double *up_vel) {
    const int r = blockIdx.x;  // one thread per replica
    const Rep R = d.rep[r];
    double up[3] = {0, 0, 0};
    for (int j = 0; j < R.nc; j++) {
        const int g = R.first_colloid + j;
        const double *o = sums + 8 * g;
        const int c = (int)o[0], cu = (int)o[1];
        double v[3];
        for (int k = 0; k < 3; k++) {
            v[k] = o[2 + k];
            up[k] = up[k] + o[5 + k];
        }
        for (int k = 0; k < 3; k++) {
            if (cu > 0) {
                up[k] = up[k] / cu - d.cvel[3 * g + k];
                v[k] = v[k] / c - d.cvel[3 * g + k];
            }
            vel[3 * g + k] = v[k];
            up_vel[3 * g + k] = up[k];
        }
        cnt[g] = c;
        up_cnt[g] = cu;
    }
}

// ------------------------------------------------------------------- audit
template <class Real>
__global__ void k_audit(Phys ph, Dev d, Fluid<Real> f) {
    __shared__ double s[256][5];
    const int T = blockIdx.x;
    const int n = min(d.count[T], ph.cap);
    double a[5] = {0, 0, 0, 0, 0};
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        int slot = T * ph.cap + i;
        double vx = f.vx[slot], vy = f.vy[slot], vz = f.vz[slot];
        a[0] += vx; a[1] += vy; a[2] += vz;
        a[3] += (vx * vx + vy * vy) + vz * vz;
        a[4] += (fabs(vx) + fabs(vy)) + fabs(vz);
    }
    for (int k = 0; k < 5; k++) s[threadIdx.x][k] = a[k];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int k = 0; k < 5; k++) s[threadIdx.x][k] += s[threadIdx.x + h][k];
        __syncthreads();
    }
    if (threadIdx.x < 5) d.audit_partial[T * 8 + threadIdx.x] = s[0][threadIdx.x];
}

// ------------------------------------------------------- gather (download)
template <class Real>
__global__ void k_gather(Phys ph, Dev d, Fluid<Real> f, double *pos, double *vel, int *cell) {
    const int T = blockIdx.x;
    Rep R;
    TileGeo tg = tile_geo(d, T, R);
    const int L = tg.L;
    const int n = min(d.count[T], ph.cap);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        int slot = T * ph.cap + i;
        long long id = R.first_fluid + f.pid[slot];
        int ix, iy, iz;
        unpack_cell(f.cell[slot], ix, iy, iz);
        double p[3] = {(double)ix + (double)f.ox[slot], (double)iy + (double)f.oy[slot],
                       (double)iz + (double)f.oz[slot]};
        for (int k = 0; k < 3; k++) {
            if (p[k] >= (double)L) p[k] -= (double)L;
            if (p[k] < 0) p[k] += (double)L;
            pos[3 * id + k] = p[k];
        }
        vel[3 * id] = f.vx[slot]; vel[3 * id + 1] = f.vy[slot]; vel[3 * id + 2] = f.vz[slot];
        cell[id] = linear_cell(ix, iy, iz, L);
    }
}

// ---------------------------------------------------------- surrogate
// Features of colloid g at the start of a step, from the coarse moments of
// the last SRD pass (its shift is still in d.shift):
//   27 coarse cells around the colloid, x fastest: density n/(8 density),
//   mean velocity relative to the colloid (3), temperature (135 values);
//   then the colloid's velocity (3), angular velocity (3), orientation (3).
// The colloid is driven by the surrogate in this step when surrogate mode is
// on, the moments are valid and every feature lies in the training range;
// otherwise it collides explicitly.  Resets the impulse accumulator.
__global__ void k_features(Phys ph, Dev d, const int *colloid_rep) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ph.n_colloid_total) return;
    const int r = colloid_rep[g];
    const Rep R = d.rep[r];
    const int L = R.L, Lc = L / 2;
    const double Ld = (double)L;
    const bool valid = *d.mom_valid != 0;
    float *fo = d.feat + (long long)g * N_FEAT;
    bool in = ph.surrogate && valid;
    auto put = [&](int k, double v) {
        const float x = (float)v;
        fo[k] = x;
        if (ph.surrogate && !(x >= d.sur_lo[k] && x <= d.sur_hi[k])) in = false;
    };
    int c3[3];
    for (int k = 0; k < 3; k++) c3[k] = coord_cell(mod1(d.cpos[3 * g + k] + d.shift[3 * r + k], Ld), L) / 2;
    const double vcx = d.cvel[3 * g], vcy = d.cvel[3 * g + 1], vcz = d.cvel[3 * g + 2];
    const double rho0 = 8.0 * (double)R.nf / ((double)L * L * L);
    for (int o = 0; o < 27; o++) {
        const int qx = wrapi(c3[0] + o % 3 - 1, Lc), qy = wrapi(c3[1] + (o / 3) % 3 - 1, Lc),
                  qz = wrapi(c3[2] + o / 9 - 1, Lc);
        const float *mo = d.cmom + 5 * (R.first_coarse + qx + (long long)Lc * (qy + (long long)Lc * qz));
        const double n = valid ? mo[0] : 0.0;
        double ux = 0, uy = 0, uz = 0, t = 0;
        if (n >= 1) {
            ux = mo[1] / n; uy = mo[2] / n; uz = mo[3] / n;
            if (n >= 2) t = ph.mass_fl * (mo[4] - n * ((ux * ux + uy * uy) + uz * uz)) / (3.0 * (n - 1));
            ux -= vcx; uy -= vcy; uz -= vcz;
        }
        put(5 * o, n / rho0);
        put(5 * o + 1, ux); put(5 * o + 2, uy); put(5 * o + 3, uz);
        put(5 * o + 4, t);
    }
    for (int k = 0; k < 3; k++) {
        put(135 + k, d.cvel[3 * g + k]);
        put(138 + k, d.comega[3 * g + k]);
        put(141 + k, d.cra[3 * g + k]);
    }
    d.cmode_prev[g] = d.cmode[g];
    d.cmode[g] = in ? 1 : 0;
    for (int k = 0; k < 6; k++) d.cimp[6 * g + k] = 0.0;
}

// impulse of a surrogate-driven colloid over the coming step: mean + sd * xi,
// xi standard normal (Box-Muller on Philox draws keyed by step and colloid),
// clamped to the training range of the impulses, applied in n_sub equal parts
__global__ void k_sur_sample(Phys ph, Dev d, const int *colloid_rep) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ph.n_colloid_total) return;
    const Rep R = d.rep[colloid_rep[g]];
    const long long step = *d.step + 1;
    const uint32_t jl = (uint32_t)(g - R.first_colloid);
    double xi[6];
    for (int h = 0; h < 3; h++) {
        uint4 a = philox(make_uint4(KIND_SUR | ((uint32_t)h << 8), (uint32_t)step, jl, (uint32_t)(step >> 32)), rep_key(R));
        double u1 = u01(a.x, a.y), u2 = u01(a.z, a.w);
        double rr = sqrt(-2.0 * log(1.0 - u1)), ph2 = 4.0 * asin(1.0) * u2;
        xi[2 * h] = rr * cos(ph2);
        xi[2 * h + 1] = rr * sin(ph2);
    }
    for (int k = 0; k < 6; k++) {
        double imp = 0.0;
        if (d.cmode[g]) {
            const double mu = d.sur_out[(long long)g * 12 + k], lv = d.sur_out[(long long)g * 12 + 6 + k];
            // never more than the explicit collisions gave in training
            const double v = fmin(fmax(mu + exp(0.5 * lv) * xi[k], (double)d.sur_lo[N_FEAT + k]),
                                  (double)d.sur_hi[N_FEAT + k]);
            imp = v / ph.n_sub;
        }
        d.csur[6 * g + k] = imp;
    }
}

// First substep, surrogate-driven colloids: the fluid in a colloid's shell
// (distance from its centre <= sigma/2 + 0.5, counted by the rotation pass on
// the same stored positions) pays the colloid's momentum and angular momentum
// debt of the step (shell_pay): particle i at r_i from the centre gets
// dv_i = -u + alpha x r_i.  Shell particles were made near, so they are among
// the first near_cnt of their tile; kicks of several colloids are summed in
// FP64 and rounded once.  One workgroup per tile.
template <class Real>
__global__ __launch_bounds__(256) void k_sur_shell_kick(Phys ph, Dev d, Fluid<Real> f) {
    __shared__ double s_k[TCL_CAP][9];  // position, u, alpha
    const int T = blockIdx.x;
    const int ns = min(d.tsl_cnt[T], TCL_CAP);
    if (ns == 0) return;
    const int L = d.rep[d.tile_rep[T]].L;
    for (int q = threadIdx.x; q < ns; q += blockDim.x) {
        const int g = d.tsl_list[T * TCL_CAP + q];
        const ShellPay pay = shell_pay(ph, d, g);
        for (int k = 0; k < 3; k++) {
            s_k[q][k] = d.cpos[3 * g + k];
            s_k[q][3 + k] = pay.u[k];
            s_k[q][6 + k] = pay.a[k];
        }
    }
    __syncthreads();
    const int n = min(d.near_cnt[T], ph.cap);
    const double Ld = (double)L;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int slot = T * ph.cap + i;
        int ix, iy, iz;
        unpack_cell(f.cell[slot], ix, iy, iz);
        const double x = (double)ix + (double)f.ox[slot], y = (double)iy + (double)f.oy[slot],
                     z = (double)iz + (double)f.oz[slot];
        double vx = f.vx[slot], vy = f.vy[slot], vz = f.vz[slot];
        bool hit = false;
        for (int q = 0; q < ns; q++) {
            const double ax = img1(x - s_k[q][0], Ld), ay = img1(y - s_k[q][1], Ld), az = img1(z - s_k[q][2], Ld);
            if ((ax * ax + ay * ay) + az * az <= ph.lim_shell) {
                const double *a = s_k[q] + 6;
                vx = vx - s_k[q][3] + (a[1] * az - a[2] * ay);
                vy = vy - s_k[q][4] + (a[2] * ax - a[0] * az);
                vz = vz - s_k[q][5] + (a[0] * ay - a[1] * ax);
                hit = true;
            }
        }
        if (hit) {
            f.vx[slot] = (Real)vx; f.vy[slot] = (Real)vy; f.vz[slot] = (Real)vz;
        }
    }
}

// When the surrogate is detached: every particle inside a colloid that is a
// phantom now is marked, per tile (the tile's phantom colloids are listed as
// for the tile colloid lists)
template <class Real>
__global__ __launch_bounds__(256) void k_mark_phantoms(Phys ph, Dev d, Fluid<Real> f) {
    __shared__ int s_g[TCL_CAP];
    __shared__ int s_n, s_wtot[16];
    const int T = blockIdx.x;
    Rep R;
    TileGeo tg = tile_geo(d, T, R);
    const int L = tg.L;
    const double Ld = (double)L;
    const int reach = TILE_LIST_REACH;
    const int lox = tg.tx * R.B[0] - reach, loy = tg.ty * R.B[1] - reach, loz = tg.tz * R.B[2] - reach;
    const int lnx = R.B[0] + 2 * reach, lny = R.B[1] + 2 * reach, lnz = R.B[2] + 2 * reach;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    for (int base = 0; base < R.nc; base += blockDim.x) {
        const int i = base + threadIdx.x, g = R.first_colloid + i;
        const bool p = i < R.nc && d.cmode[g] && near_interval(d.ccell[3 * g], lox, lnx, L) &&
                       near_interval(d.ccell[3 * g + 1], loy, lny, L) && near_interval(d.ccell[3 * g + 2], loz, lnz, L);
        int tot;
        const int idx = block_compact(p, s_wtot, tot);
        if (p && s_n + idx < TCL_CAP) s_g[s_n + idx] = g;
        __syncthreads();
        if (threadIdx.x == 0) s_n = min(s_n + tot, TCL_CAP);
        __syncthreads();
    }
    const int n = min(d.count[T], ph.cap), ng = s_n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int slot = T * ph.cap + i;
        const int c = f.cell[slot];
        int ix, iy, iz;
        unpack_cell(c, ix, iy, iz);
        const double x = (double)ix + (double)f.ox[slot], y = (double)iy + (double)f.oy[slot],
                     z = (double)iz + (double)f.oz[slot];
        for (int q = 0; q < ng; q++) {
            const double *pc = d.cpos + 3 * s_g[q];
            const double ax = img1(x - pc[0], Ld), ay = img1(y - pc[1], Ld), az = img1(z - pc[2], Ld);
            if ((ax * ax + ay * ay) + az * az <= ph.lim_coll) {
                f.cell[slot] = c | PHANTOM;
                break;
            }
        }
    }
}

// plain copy for HBM counter calibration
// The periodic-geometry helpers on given inputs, for their test against the
// formulas they replace (img1/mod1 against a - L*round(a/L), wrapi against a
// mod n, div_small against a / b).
__global__ void k_debug_geometry(int n, const double *a, double L, const int *ia, int Li, int b, double *img_out,
                                 double *mod_out, int *wrap_out, int *div_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    img_out[i] = img1(a[i], L);
    mod_out[i] = mod1(a[i], L);
    wrap_out[i] = wrapi(ia[i], Li);
    div_out[i] = (ia[i] >= 0 && ia[i] < 1024) ? div_small(ia[i], div_small_magic(b)) : -1;
}
// Test hooks for the device draw source: Philox4x32-10 on given (counter,
// key) pairs, and device-drawn reflections of replica 0 applied at a fixed
// surface point (raw tuple and the reflected velocity per particle id).
__global__ void k_debug_philox(int n, const uint4 *ctr, const uint2 *key, uint4 *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = philox(ctr[i], key[i]);
}
__global__ void k_debug_reflect(Phys ph, Dev d, int n, long long step, int substep, double L, double rfx,
                                double rfy, double rfz, double rsx, double rsy, double rsz, double *vals6,
                                double *vel3) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Rep R = d.rep[0];
    ReflDraw dr;
    draw_reflection(d, ph, R, step, substep, R.first_colloid, i, dr);
    double ux, uy, uz;
    stochastic_reflection(ph, L, rfx, rfy, rfz, rsx, rsy, rsz, dr, ux, uy, uz);
    for (int k = 0; k < 6; k++) vals6[6 * i + k] = dr.v[k];
    vel3[3 * i] = ux;
    vel3[3 * i + 1] = uy;
    vel3[3 * i + 2] = uz;
}
__global__ void k_copy(const float4 *__restrict__ a, float4 *__restrict__ b, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) b[i] = a[i];
}

// -------------------------------------------------------- launch wrappers
#define LAUNCH(kern, grid, block, lds, st, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(block), lds, st, __VA_ARGS__)

// static LDS of the SRD kernel's largest variant (colloid lists, masks)
size_t srd_static_lds() { return sizeof(TileRun) + sizeof(NearColl) + 4 * TCL_CAP + 2 * sizeof(AxisMask) + 512; }
// dynamic LDS the SRD kernel adds while a surrogate is attached
size_t srd_sur_lds() { return SRD_SUR_LDS + 16; }

template <class Real>
void launch_depart(const Phys &ph, const Dev &d, const Fluid<Real> &f, hipStream_t st) {
    LAUNCH(k_depart<Real>, ph.n_tiles, DEPART_THREADS, 0, st, ph, d, f);
}
template <class Real, bool C, bool M, bool B, bool W>
static void srd_th(const Phys &ph, const Dev &d, const Fluid<Real> &src, const Fluid<Real> &dst, size_t lds,
                   hipStream_t st) {
    switch (ph.srd_threads * 8 + ph.srd_pt) {
        case 256 * 8 + 3: LAUNCH((k_srd<Real, C, M, B, 256, W, 3>), ph.n_tiles, 256, lds, st, ph, d, src, dst); break;
        case 256 * 8 + 4: LAUNCH((k_srd<Real, C, M, B, 256, W, 4>), ph.n_tiles, 256, lds, st, ph, d, src, dst); break;
        case 512 * 8 + 3: LAUNCH((k_srd<Real, C, M, B, 512, W, 3>), ph.n_tiles, 512, lds, st, ph, d, src, dst); break;
        case 512 * 8 + 4: LAUNCH((k_srd<Real, C, M, B, 512, W, 4>), ph.n_tiles, 512, lds, st, ph, d, src, dst); break;
        case 1024 * 8 + 3: LAUNCH((k_srd<Real, C, M, B, 1024, W, 3>), ph.n_tiles, 1024, lds, st, ph, d, src, dst); break;
        default: LAUNCH((k_srd<Real, C, M, B, 1024, W, 4>), ph.n_tiles, 1024, lds, st, ph, d, src, dst); break;
    }
}
template <class Real, bool C, bool M, bool B>
static void srd_one(const Phys &ph, const Dev &d, const Fluid<Real> &src, const Fluid<Real> &dst, size_t lds,
                    hipStream_t st) {
    if (C && B && !ph.near_wrap) srd_th<Real, C, M, B, false>(ph, d, src, dst, lds, st);
    else srd_th<Real, C, M, B, true>(ph, d, src, dst, lds, st);
}
// The SRD pass's FP32 near test compares a particle (within TILE_REACH + 1
// cells of its tile box) with listed colloids (within TILE_LIST_REACH + 1
// cells of it): every difference stays below L/2, where the periodic wrap is
// the identity, once L >= 4 (Bmax + TILE_LIST_REACH + TILE_REACH + 2).
bool srd_needs_wrap(int L, int bmax) { return L < 4 * (bmax + TILE_LIST_REACH + TILE_REACH + 2); }
// Dynamic LDS up to the CU's 160 KB is granted at launch on gfx950 without an
// opt-in attribute; the context checks the SRD pass's need against
// 160 KB minus its static LDS when it is created.
// The near-particle pass loops over its chunks: as many workgroups as the
// device holds at once (so that all start together), at most NEAR_BLOCKS.
hipError_t prepare_kernels(Phys &ph, bool fp64) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = fp64 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_stream_near<double, false>, 256, 0)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_stream_near<float, false>, 256, 0);
    if (e != hipSuccess) return e;
    ph.near_blocks = std::max(1, std::min(NEAR_BLOCKS, cus * std::max(per_cu, 1)));
    return hipSuccess;
}
// count_shell: run() shell counts (colloids present); map: shifted-cell map of
// every particle (the standalone rotation_mpcd); block: near/far split of the
// fused step
template <class Real>
void launch_srd(const Phys &ph, const Dev &d, const Fluid<Real> &src, const Fluid<Real> &dst, bool count_shell,
                bool map, bool block, size_t lds, hipStream_t st) {
    if (map) srd_one<Real, false, true, false>(ph, d, src, dst, lds, st);
    else if (count_shell && block) srd_one<Real, true, false, true>(ph, d, src, dst, lds, st);
    else if (count_shell) srd_one<Real, true, false, false>(ph, d, src, dst, lds, st);
    else if (block) srd_one<Real, false, false, true>(ph, d, src, dst, lds, st);
    else srd_one<Real, false, false, false>(ph, d, src, dst, lds, st);
}
template <class Real, int TH>
static void stream_one(const Phys &ph, const Dev &d, const Fluid<Real> &f, int m, int nf, hipStream_t st) {
    switch (m) {
        case 7: LAUNCH((k_stream<Real, true, true, true, TH>), ph.n_tiles, TH, 0, st, ph, d, f, nf); break;
        case 6: LAUNCH((k_stream<Real, true, true, false, TH>), ph.n_tiles, TH, 0, st, ph, d, f, nf); break;
        case 4: LAUNCH((k_stream<Real, true, false, false, TH>), ph.n_tiles, TH, 0, st, ph, d, f, nf); break;
        case 3: LAUNCH((k_stream<Real, false, true, true, TH>), ph.n_tiles, TH, 0, st, ph, d, f, nf); break;
        case 2: LAUNCH((k_stream<Real, false, true, false, TH>), ph.n_tiles, TH, 0, st, ph, d, f, nf); break;
        case 1: LAUNCH((k_stream<Real, false, false, true, TH>), ph.n_tiles, TH, 0, st, ph, d, f, nf); break;
        default: break;
    }
}
template <class Real>
void launch_stream(const Phys &ph, const Dev &d, const Fluid<Real> &f, bool kick, bool stream, bool detect,
                   bool near_only, hipStream_t st) {
    int m = (kick ? 4 : 0) | (stream ? 2 : 0) | (detect ? 1 : 0);
    stream_one<Real, STREAM_THREADS>(ph, d, f, m, near_only ? 1 : 0, st);
}
template <class Real>
void launch_stream_near(const Phys &ph, const Dev &d, const Fluid<Real> &f, bool kick, hipStream_t st) {
    if (kick) LAUNCH((k_stream_near<Real, true>), ph.near_blocks, 256, 0, st, ph, d, f);
    else LAUNCH((k_stream_near<Real, false>), ph.near_blocks, 256, 0, st, ph, d, f);
}
template <class Real>
void launch_collide(const Phys &ph, const Dev &d, const Fluid<Real> &f, int substep, const int2 *wgt, int nwg,
                    hipStream_t st) {
    LAUNCH(k_collide<Real>, nwg, COLLIDE_THREADS, 0, st, ph, d, f, substep, wgt);
    LAUNCH(k_serial<Real>, ph.n_rep, COLLIDE_THREADS, 0, st, ph, d, f, substep);
}
template <class Real>
void launch_run_count(const Phys &ph, const Dev &d, const Fluid<Real> &f, hipStream_t st) {
    LAUNCH(k_run_count<Real>, ph.n_tiles, STREAM_THREADS, 0, st, ph, d, f);
}
template <class Real>
void launch_maps(const Phys &ph, const Dev &d, const Fluid<Real> &f, long long ncells, int *bsum,
                 const int *colloid_rep, hipStream_t st) {
    (void)hipMemsetAsync(d.map_count, 0, sizeof(int) * ncells, st);
    LAUNCH(k_map_count<Real>, ph.n_tiles, 256, 0, st, ph, d, f);
    long long nb = (ncells + 1023) / 1024;
    LAUNCH(k_scan_block, nb, 1024, 0, st, d.map_count, d.map_start, bsum, ncells);
    LAUNCH(k_scan_sums, 1, 1024, 0, st, bsum, (int)nb);
    LAUNCH(k_scan_add, nb, 1024, 0, st, d.map_start, bsum, ncells);
    (void)hipMemcpyAsync(d.map_fill, d.map_start, sizeof(int) * ncells, hipMemcpyDeviceToDevice, st);
    LAUNCH(k_map_fill<Real>, ph.n_tiles, 256, 0, st, ph, d, f);
    LAUNCH(k_map_sort, (ncells + 255) / 256, 256, 0, st, d, ncells);
    if (ph.n_colloid_total > 0) LAUNCH(k_neigh_fl, ph.n_colloid_total, 256, 0, st, ph, d, colloid_rep);
}
template <class Real>
void launch_audit(const Phys &ph, const Dev &d, const Fluid<Real> &f, hipStream_t st) {
    LAUNCH(k_audit<Real>, ph.n_tiles, 256, 0, st, ph, d, f);
}
template <class Real>
void launch_gather(const Phys &ph, const Dev &d, const Fluid<Real> &f, double *pos, double *vel, int *cell,
                   hipStream_t st) {
    LAUNCH(k_gather<Real>, ph.n_tiles, 256, 0, st, ph, d, f, pos, vel, cell);
}
void launch_step_begin(const Phys &ph, const Dev &d, hipStream_t st, bool clear_lists) {
    LAUNCH(k_step_begin, clear_lists ? std::max(1, (ph.n_tiles + 255) / 256) : 1, 256, 0, st, ph, d, (int)clear_lists);
}
void launch_colloid(const Phys &ph, const Dev &d, unsigned flags, hipStream_t st) {
    LAUNCH(k_colloid, ph.n_rep, 256, 0, st, ph, d, flags);
}
void launch_colloid_update(const Phys &ph, const Dev &d, unsigned flags, const int2 *wgt, int nwg, hipStream_t st) {
    LAUNCH(k_colloid_update, nwg, 256, 0, st, ph, d, flags, wgt);
}
void launch_md_candidates(const Phys &ph, const Dev &d, const int2 *wgt, int nwg, hipStream_t st) {
    LAUNCH(k_md_candidates, nwg, 256, 0, st, ph, d, wgt);
}
void launch_tile_colloids(const Phys &ph, const Dev &d, hipStream_t st, bool cleared) {
    if (!cleared) LAUNCH(k_tile_lists_clear, (ph.n_tiles + 255) / 256, 256, 0, st, ph, d);
    if (ph.n_colloid_total > 0) LAUNCH(k_tile_colloids, ph.n_colloid_total, 256, 0, st, ph, d);
}
void launch_features(const Phys &ph, const Dev &d, const int *colloid_rep, hipStream_t st) {
    LAUNCH(k_features, (ph.n_colloid_total + 63) / 64, 64, 0, st, ph, d, colloid_rep);
}
void launch_sur_sample(const Phys &ph, const Dev &d, const int *colloid_rep, hipStream_t st) {
    LAUNCH(k_sur_sample, (ph.n_colloid_total + 63) / 64, 64, 0, st, ph, d, colloid_rep);
}
template <class Real>
void launch_mark_phantoms(const Phys &ph, const Dev &d, const Fluid<Real> &f, hipStream_t st) {
    LAUNCH(k_mark_phantoms<Real>, ph.n_tiles, 256, 0, st, ph, d, f);
}
template void launch_mark_phantoms<float>(const Phys &, const Dev &, const Fluid<float> &, hipStream_t);
template void launch_mark_phantoms<double>(const Phys &, const Dev &, const Fluid<double> &, hipStream_t);
template <class Real>
void launch_sur_shell_kick(const Phys &ph, const Dev &d, const Fluid<Real> &f, hipStream_t st) {
    LAUNCH(k_sur_shell_kick<Real>, ph.n_tiles, 256, 0, st, ph, d, f);
}
template void launch_sur_shell_kick<float>(const Phys &, const Dev &, const Fluid<float> &, hipStream_t);
template void launch_sur_shell_kick<double>(const Phys &, const Dev &, const Fluid<double> &, hipStream_t);
void launch_debug_geometry(int n, const double *a, double L, const int *ia, int Li, int b, double *img_out,
                           double *mod_out, int *wrap_out, int *div_out, hipStream_t st) {
    LAUNCH(k_debug_geometry, (n + 255) / 256, 256, 0, st, n, a, L, ia, Li, b, img_out, mod_out, wrap_out, div_out);
}
void launch_updown(const Phys &ph, const Dev &d, int n_rep, const double *pos, const double *vel, double *sums,
                   int *cnt, int *up_cnt, double *v, double *up, hipStream_t st) {
    if (ph.n_colloid_total > 0) LAUNCH(k_updown, ph.n_colloid_total, 256, 0, st, ph, d, pos, vel, sums);
    LAUNCH(k_updown_finish, n_rep, 1, 0, st, d, sums, cnt, up_cnt, v, up);
}
template <class Real>
void launch_init_fluid(const Phys &ph, const Dev &d, const InitBins &b, int vel_law, long long n_total,
                       long long max_nf, double *pos, double *vel, double *mean, const Fluid<Real> &scratch,
                       const Fluid<Real> &out, int *fill, int *count, hipStream_t st) {
    const unsigned g = (unsigned)((n_total + 255) / 256);
    LAUNCH(k_init_place, g, 256, 0, st, ph, d, b, vel_law, n_total, pos, vel);
    const int nchunk = init_mean_chunks(max_nf);
    double *part = mean + 3 * ph.n_rep;
    hipLaunchKernelGGL(k_init_mean_part, dim3(nchunk, ph.n_rep), dim3(256), 0, st, d, vel, nchunk, part);
    LAUNCH(k_init_mean_final, ph.n_rep, 64, 0, st, d, part, nchunk, mean);
    LAUNCH(k_init_bin<Real>, g, 256, 0, st, ph, d, n_total, pos, vel, mean, scratch, fill);
    LAUNCH(k_init_sort<Real>, ph.n_tiles, 512, 0, st, ph, d, fill, scratch, out, count);
}
template void launch_init_fluid<float>(const Phys &, const Dev &, const InitBins &, int, long long, long long,
                                       double *, double *, double *, const Fluid<float> &, const Fluid<float> &, int *,
                                       int *, hipStream_t);
template void launch_init_fluid<double>(const Phys &, const Dev &, const InitBins &, int, long long, long long,
                                        double *, double *, double *, const Fluid<double> &, const Fluid<double> &,
                                        int *, int *, hipStream_t);
void launch_debug_philox(int n, const uint4 *ctr, const uint2 *key, uint4 *out, hipStream_t st) {
    LAUNCH(k_debug_philox, (n + 255) / 256, 256, 0, st, n, ctr, key, out);
}
void launch_debug_reflect(const Phys &ph, const Dev &d, int n, long long step, int substep, double L,
                          const double *rf, const double *rs, double *vals6, double *vel3, hipStream_t st) {
    LAUNCH(k_debug_reflect, (n + 255) / 256, 256, 0, st, ph, d, n, step, substep, L, rf[0], rf[1], rf[2], rs[0],
           rs[1], rs[2], vals6, vel3);
}

void launch_copy(const void *a, void *b, long long nbytes, hipStream_t st) {
    long long n = nbytes / 16;
    LAUNCH(k_copy, 2048, 256, 0, st, (const float4 *)a, (float4 *)b, n);
}

#define INST(Real)                                                                                                 \
    template void launch_depart<Real>(const Phys &, const Dev &, const Fluid<Real> &, hipStream_t);               \
    template void launch_srd<Real>(const Phys &, const Dev &, const Fluid<Real> &, const Fluid<Real> &, bool,    \
                                   bool, bool, size_t, hipStream_t);                                               \
    template void launch_stream<Real>(const Phys &, const Dev &, const Fluid<Real> &, bool, bool, bool, bool,     \
                                      hipStream_t);                                                                \
    template void launch_collide<Real>(const Phys &, const Dev &, const Fluid<Real> &, int, const int2 *, int,    \
                                       hipStream_t);                                                               \
    template void launch_stream_near<Real>(const Phys &, const Dev &, const Fluid<Real> &, bool, hipStream_t);    \
    template void launch_run_count<Real>(const Phys &, const Dev &, const Fluid<Real> &, hipStream_t);            \
    template void launch_maps<Real>(const Phys &, const Dev &, const Fluid<Real> &, long long, int *,             \
                                    const int *, hipStream_t);                                                     \
    template void launch_audit<Real>(const Phys &, const Dev &, const Fluid<Real> &, hipStream_t);                \
    template void launch_gather<Real>(const Phys &, const Dev &, const Fluid<Real> &, double *, double *, int *,  \
                                      hipStream_t);
INST(float)
INST(double)

}  // namespace bw
