// Device helpers of libbrownian (gfx950).  Included by the kernel sources only.
#pragma once
#include <type_traits>
#include <hip/hip_runtime.h>

#include "../../include/brownian.h"
#include "bw_common.h"

namespace bw {

// A wave-uniform load through the scalar data cache (read-only use: the data
// must not change while the kernel runs).  Many waves reading one line at a
// kernel's start then share the scalar cache instead of each sending a vector
// request to L2.
template <class T>
__device__ __forceinline__ T ld_const(const T *p) {
    static_assert(sizeof(T) % 4 == 0, "whole 32-bit words");
    typedef const __attribute__((address_space(4))) unsigned cword;
    T v;
    unsigned *o = (unsigned *)&v;
    for (int i = 0; i < (int)(sizeof(T) / 4); i++) o[i] = ((cword *)p)[i];
    return v;
}

// Value of lane l (the same l in every lane) of a wave, through the scalar
// unit rather than a cross-lane LDS permute.
__device__ __forceinline__ int lane_value(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ long long lane_value(long long v, int l) {
    const unsigned long long u = (unsigned long long)v;
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return (long long)(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ double lane_value(double v, int l) {
    return __longlong_as_double(lane_value((long long)__double_as_longlong(v), l));
}

// ------------------------------------------------------------------ generator
// Counter-based generator for the device draw source: Philox4x32-10
// (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11).  A draw is a pure function of (seed, step, kind, index), so replicas,
// tiles and graph replays need no generator state.
__device__ __forceinline__ uint4 philox(uint4 c, uint2 k) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
        uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
        k.x += 0x9E3779B9u;
        k.y += 0xBB67AE85u;
    }
    return c;
}
// 53-bit uniform in [0,1) from two words
__device__ __forceinline__ double u01(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

enum : uint32_t { KIND_SHIFT = 1, KIND_AXIS = 2, KIND_REFL = 3, KIND_SUR = 5, KIND_INIT_POS = 6, KIND_INIT_VEL = 7 };

__device__ __forceinline__ uint2 rep_key(const Rep &R) { return make_uint2(R.seed_lo, R.seed_hi); }

// raw uniforms of the grid shift (cpp/rotation_mpcd.cpp:9 draws 3)
__device__ inline void draw_shift_raw(const Dev &d, const Phys &ph, const Rep &R, int r, long long step,
                                      double out[3]) {
    if (ph.draws == BW_DRAWS_INJECTED) {
        for (int k = 0; k < 3; k++) out[k] = d.inj_shift[3 * r + k];
        return;
    }
    uint4 a = philox(make_uint4(KIND_SHIFT, (uint32_t)step, 0u, (uint32_t)(step >> 32)), rep_key(R));
    uint4 b = philox(make_uint4(KIND_SHIFT | 0x100u, (uint32_t)step, 0u, (uint32_t)(step >> 32)), rep_key(R));
    out[0] = u01(a.x, a.y);
    out[1] = u01(a.z, a.w);
    out[2] = u01(b.x, b.y);
}

// raw uniforms (rho, phi draws of cpp/rotation_mpcd.cpp:19) for shifted cell `cell`
__device__ inline bool draw_axis_raw(const Dev &d, const Phys &ph, const Rep &R, long long step, int cell,
                                     double &u1, double &u2) {
    if (ph.draws == BW_DRAWS_INJECTED) {
        long long i = R.first_cell + cell;
        u1 = d.inj_axes[2 * i];
        u2 = d.inj_axes[2 * i + 1];
        return !(u1 != u1);  // NaN marks "no draw"
    }
    uint4 a = philox(make_uint4(KIND_AXIS, (uint32_t)step, (uint32_t)cell, (uint32_t)(step >> 32)), rep_key(R));
    u1 = u01(a.x, a.y);
    u2 = u01(a.z, a.w);
    return true;
}

// raw reflection tuple (U_e, t.x, t.y, t.z, a, b) of cpp/fluid_colloid_collision.cpp:8-19
struct ReflDraw {
    double v[6];
};
__device__ inline bool draw_reflection(const Dev &d, const Phys &ph, const Rep &R, long long step, int substep,
                                       int g, int pid, ReflDraw &o) {
    if (ph.draws == BW_DRAWS_INJECTED) {
        int base = substep * ph.n_colloid_total + g;
        int b = d.inj_off[base], e = d.inj_off[base + 1];
        for (int i = b; i < e; i++)
            if (d.inj_pid[i] == pid) {
                for (int k = 0; k < 6; k++) o.v[k] = d.inj_val[6 * i + k];
                return true;
            }
        return false;
    }
    // keyed by the replica-local colloid index: a replica draws the same
    // numbers alone or inside an ensemble
    uint32_t c0 = KIND_REFL | ((uint32_t)substep << 8);
    const uint32_t jl = (uint32_t)(g - R.first_colloid);
    uint4 a = philox(make_uint4(c0, (uint32_t)step, (uint32_t)pid, jl), rep_key(R));
    uint4 b = philox(make_uint4(c0 | (1u << 16), (uint32_t)step, (uint32_t)pid, jl), rep_key(R));
    o.v[0] = u01(a.x, a.y);
    o.v[1] = u01(a.z, a.w);
    o.v[2] = u01(b.x, b.y);
    o.v[3] = u01(b.z, b.w);
    // polar pair by rejection, as the reference's loop (:17-20); z = 0 also rejected
    for (uint32_t t = 2; t < 65535u; t++) {
        uint4 c = philox(make_uint4(c0 | (t << 16), (uint32_t)step, (uint32_t)pid, jl), rep_key(R));
        double aa = u01(c.x, c.y), bb = u01(c.z, c.w);
        double x1 = 2.0 * aa - 1, x2 = 2.0 * bb - 1, z = x1 * x1 + x2 * x2;
        if (z <= 1 && z > 0) {
            o.v[4] = aa;
            o.v[5] = bb;
            break;
        }
    }
    return true;
}

// -------------------------------------------------------------- geometry
// img(a,b) = a - b*round(a/b) with round half away from zero (cpp/parameters.hpp:74,76).
// Without the division where round(a/L) is certain: 0 for |a| < 0.4999 L
// (a + 0.0 turns -0 into +0, as a - L*(-0.0) does), +-1 for 0.5001 L < |a| <
// 1.4999 L (L*(+-1) = +-L exactly).  Bit-identical to the formula.
__device__ __forceinline__ double img1(double a, double L) {
    const double h = fabs(a);
    if (h < 0.4999 * L) return a + 0.0;
    if (h > 0.5001 * L && h < 1.4999 * L) return a - copysign(L, a);
    return a - L * round(a / L);
}
// mod(a,b) = a - b*round((a - b/2)/b) (cpp/parameters.hpp:75)
__device__ __forceinline__ double mod1(double a, double L) { return a - L * round((a - L / 2) / L); }

constexpr int CELL_MASK = (1 << CELL_BITS) - 1;
// bit 30 of the packed cell: the particle is inside a colloid that was a
// surrogate-driven phantom when it got there; it collides with no colloid
// until a detection finds it inside none
constexpr int PHANTOM = 1 << 30;
__device__ __forceinline__ void unpack_cell(int c, int &ix, int &iy, int &iz) {
    ix = c & CELL_MASK;
    iy = (c >> CELL_BITS) & CELL_MASK;
    iz = (c >> (2 * CELL_BITS)) & CELL_MASK;
}
__device__ __forceinline__ int pack_cell(int ix, int iy, int iz) {
    return ix | (iy << CELL_BITS) | (iz << (2 * CELL_BITS));
}
// reference cell numbering (cpp/parameters.hpp:59, 0-based)
__device__ __forceinline__ int linear_cell(int ix, int iy, int iz, int L) { return ix + L * (iy + L * iz); }
__device__ __forceinline__ int linear_of_packed(int c, int L) {
    int ix, iy, iz;
    unpack_cell(c, ix, iy, iz);
    return linear_cell(ix, iy, iz, L);
}
__device__ __forceinline__ int pdiff(int a, int b, int L) {
    int d = a - b;
    if (d > L / 2) d -= L;
    if (d < -(L / 2)) d += L;
    return d;
}
// a / b for 0 <= a < 1024 and 1 <= b <= 63 by one multiply: with
// m = ceil(2^16 / b), a*m/2^16 exceeds a/b by less than a/2^16 < 1/b, the
// least distance from a/b up to the next integer
__device__ __forceinline__ unsigned div_small_magic(int b) { return (65536u + (unsigned)b - 1u) / (unsigned)b; }
__device__ __forceinline__ int div_small(int a, unsigned m) { return (int)(((unsigned)a * m) >> 16); }
// a mod n in [0, n); one conditional step covers a in [-n, 2n)
__device__ __forceinline__ int wrapi(int a, int n) {
    if (a < 0) a += n;
    else if (a >= n) a -= n;
    if ((unsigned)a >= (unsigned)n) {
        a %= n;
        if (a < 0) a += n;
    }
    return a;
}
// cell (ix,iy,iz) belongs to the 7^3-1 stencil of cell (cx,cy,cz)
// (cpp/neighbour_list.cpp:12-25); requires L >= 7
__device__ __forceinline__ bool in_stencil(int ix, int iy, int iz, int cx, int cy, int cz, int L) {
    int dx = pdiff(ix, cx, L), dy = pdiff(iy, cy, L), dz = pdiff(iz, cz, L);
    return dx >= -3 && dx <= 3 && dy >= -3 && dy <= 3 && dz >= -3 && dz <= 3 && (dx | dy | dz) != 0;
}
// position of (dx,dy,dz) in the reference stencil order (x fastest, centre skipped)
__device__ __forceinline__ int stencil_index(int dx, int dy, int dz) {
    int i = (dx + 3) + 7 * (dy + 3) + 49 * (dz + 3);
    return i < 171 ? i : i - 1;
}
// cell of an absolute colloid position in [0, L] (cpp/parameters.hpp:59)
__device__ __forceinline__ int coord_cell(double x, int L) {
    int c = (int)x;
    if (c >= L) c -= L;
    if (c < 0) c += L;
    return c;
}

// shifted-cell coordinate: ix + floor(o + s) on one IEEE add, wrapped
template <class Real>
__device__ __forceinline__ int shifted_coord(int ix, Real o, double s, int L) {
    int k = ix + (int)floor((double)o + s);
    if (k >= L || k < 0) {  // a reflection may leave the offset far outside [0,1)
        k %= L;
        if (k < 0) k += L;
    }
    return k;
}

// bring an offset back into [0,1), moving the cell index (periodic)
template <class Real>
__device__ __forceinline__ void normalise(int &i, Real &o, int L) {
    Real f = floor(o);
    if (f != Real(0)) {
        o -= f;
        i += (int)f;
        if (o >= Real(1)) {  // o + 1 rounded up to 1 for a tiny negative o
            o = Real(0);
            i += 1;
        }
        if (i >= L || i < 0) {  // a particle may cross several box lengths in a substep
            i %= L;
            if (i < 0) i += L;
        }
    }
}

// o in [0,1) (+0 included, -0 and NaN excluded) with one integer compare
// normalise() for a particle that moves less than one cell in all: the cell
// index changes by at most one in the whole run of substeps, so its periodic
// wrap (which the offset does not depend on) is left to the caller, once
// after the run (i in [-1, L] then).
template <class Real>
__device__ __forceinline__ void normalise_nowrap(int &i, Real &o) {
    Real f = floor(o);
    if (f != Real(0)) {
        o -= f;
        i += (int)f;
        if (o >= Real(1)) {  // o + 1 rounded up to 1 for a tiny negative o
            o = Real(0);
            i += 1;
        }
    }
}
__device__ __forceinline__ int wrap_once(int i, int L) { return i >= L ? i - L : (i < 0 ? i + L : i); }

__device__ __forceinline__ bool in_unit(float o) { return __float_as_uint(o) < 0x3F800000u; }
__device__ __forceinline__ bool in_unit(double o) {
    return (unsigned long long)__double_as_longlong(o) < 0x3FF0000000000000ull;
}
// true only if all three lie in [0,1): the OR of the bit patterns is at least
// each of them, so any value outside (sign bit set, or >= 1.0) makes it false;
// it may also be false for three values inside, where the caller's
// normalise() is the identity
__device__ __forceinline__ bool in_unit3(float a, float b, float c) {
    return (__float_as_uint(a) | __float_as_uint(b) | __float_as_uint(c)) < 0x3F800000u;
}
__device__ __forceinline__ bool in_unit3(double a, double b, double c) {
    return ((unsigned long long)__double_as_longlong(a) | (unsigned long long)__double_as_longlong(b) |
            (unsigned long long)__double_as_longlong(c)) < 0x3FF0000000000000ull;
}


// ------------------------------------------------------------ reflection
// Stochastic (thermal-wall) reflection, cpp/fluid_colloid_collision.cpp:7-24:
// Rayleigh normal speed along rs, tangential Gaussian along un x t.
__device__ inline void stochastic_reflection(const Phys &ph, double L, double rfx, double rfy, double rfz,
                                             double rsx, double rsy, double rsz, const ReflDraw &dr,
                                             double &ux, double &uy, double &uz) {
    double m_beta = ph.mass_fl / ph.kbt;
    double re = (1 - dr.v[0]) * (1 - dr.v[0]);
    double nrm = sqrt((rsx * rsx + rsy * rsy) + rsz * rsz);
    double nx = rsx / nrm, ny = rsy / nrm, nz = rsz / nrm;
    double val = sqrt(-log(re) / m_beta);
    double unx = nx * val, uny = ny * val, unz = nz * val;
    double tx = img1(dr.v[1] * L - rfx, L), ty = img1(dr.v[2] * L - rfy, L), tz = img1(dr.v[3] * L - rfz, L);
    double utx = uny * tz - unz * ty, uty = unz * tx - unx * tz, utz = unx * ty - uny * tx;
    double un2 = sqrt((utx * utx + uty * uty) + utz * utz);
    utx /= un2; uty /= un2; utz /= un2;
    double x1 = 2.0 * dr.v[4] - 1, x2 = 2.0 * dr.v[5] - 1;
    double z = x1 * x1 + x2 * x2;
    z = sqrt((-2.0 * log(z)) / z);
    double v1 = x1 * z * ph.vscale_thermal;
    ux = utx * v1 + unx;
    uy = uty * v1 + uny;
    uz = utz * v1 + unz;
}

// ------------------------------------------------------- block utilities
// stable compaction index within the block for threads with pred set
__device__ inline int block_compact(bool pred, int *s_wtot, int &total) {
    const int nw = blockDim.x >> 6, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned long long m = __ballot(pred);
    int rank = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) s_wtot[w] = __popcll(m);
    __syncthreads();
    int off = 0, tot = 0;
    for (int i = 0; i < nw; i++) {
        int v = s_wtot[i];
        off += (i < w) ? v : 0;
        tot += v;
    }
    __syncthreads();
    total = tot;
    return off + rank;
}

// exclusive scan of s[0..n) in place (n <= 1024), returns the total
__device__ inline int block_exclusive_scan(int *s, int n, int *s_wtot) {
    const int nw = blockDim.x >> 6, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int total = 0;
    for (int base = 0; base < n; base += blockDim.x) {
        int i = base + threadIdx.x;
        int v = i < n ? s[i] : 0;
        int x = v;
        for (int o = 1; o < 64; o <<= 1) {
            int y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) s_wtot[w] = x;
        __syncthreads();
        int off = 0, tot = 0;
        for (int k = 0; k < nw; k++) {
            off += (k < w) ? s_wtot[k] : 0;
            tot += s_wtot[k];
        }
        if (i < n) s[i] = total + off + x - v;
        total += tot;
        __syncthreads();
    }
    return total;
}

// ------------------------------------------------ colloids listed per tile
__device__ __forceinline__ bool near_interval(int c, int lo, int len, int L) {
    if (len >= L) return true;
    return wrapi(c - lo, L) < len;
}

// Colloids that can touch a particle of a tile during one MPCD step: a
// particle stays within TILE_REACH cells of its tile box, the collision and
// run() stencils reach 3 cells and a colloid's cell moves by at most one cell
// within a step, so the colloid's cell at the start of the step lies within the
// tile box grown by TILE_REACH + 4 (per axis, periodic).  Built once per step
// (ascending g).  A colloid that leaves the one-cell margin marks its
// replica's lists stale and every workgroup then rebuilds its own list.
constexpr int TILE_LIST_REACH = TILE_REACH + 4;

// sur = false: the colloids that collide explicitly (tcl_*); sur = true: the
// surrogate-driven ones, whose fluid the rotation pass keeps out (tsl_*)
// cell window of tile T within which a colloid's current or run() cell puts
// it on the tile's list
__device__ __forceinline__ void tile_list_window(const Rep &R, int T, int *lo, int *ln) {
    const int lt = T - R.first_tile;
    const int t[3] = {lt % R.nt[0], (lt / R.nt[0]) % R.nt[1], lt / (R.nt[0] * R.nt[1])};
    for (int k = 0; k < 3; k++) {
        lo[k] = t[k] * R.B[k] - TILE_LIST_REACH;
        ln[k] = R.B[k] + 2 * TILE_LIST_REACH;
    }
}
__device__ __forceinline__ bool tile_list_member(const Phys &ph, const Dev &d, const Rep &R, int g, const int *lo,
                                                 const int *ln, bool sur) {
    if ((ph.surrogate && d.cmode[g]) != sur) return false;
    bool p = near_interval(d.ccell[3 * g], lo[0], ln[0], R.L) && near_interval(d.ccell[3 * g + 1], lo[1], ln[1], R.L) &&
             near_interval(d.ccell[3 * g + 2], lo[2], ln[2], R.L);
    return p || (near_interval(d.ccell_run[3 * g], lo[0], ln[0], R.L) &&
                 near_interval(d.ccell_run[3 * g + 1], lo[1], ln[1], R.L) &&
                 near_interval(d.ccell_run[3 * g + 2], lo[2], ln[2], R.L));
}

// sur = false: the colloids that collide explicitly (tcl_*); sur = true: the
// surrogate-driven ones, whose fluid the rotation pass keeps out (tsl_*)
__device__ inline void build_tile_list(const Phys &ph, const Dev &d, const Rep &R, int T, bool sur = false) {
    __shared__ int s_wtot[16];
    __shared__ int s_n;
    int *list = sur ? d.tsl_list : d.tcl_list;
    int lo[3], ln[3];
    tile_list_window(R, T, lo, ln);
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    for (int base = 0; base < R.nc; base += blockDim.x) {
        int i = base + threadIdx.x;
        int g = R.first_colloid + i;
        bool p = i < R.nc && tile_list_member(ph, d, R, g, lo, ln, sur);
        int tot;
        int idx = block_compact(p, s_wtot, tot);
        if (p) {
            if (s_n + idx < TCL_CAP) list[T * TCL_CAP + s_n + idx] = g;
            else atomicOr(d.status, BW_FLAG_LIST_OVERFLOW);
        }
        __syncthreads();
        if (threadIdx.x == 0) s_n = min(s_n + tot, TCL_CAP);
        __syncthreads();
    }
    if (threadIdx.x == 0) (sur ? d.tsl_cnt : d.tcl_cnt)[T] = s_n;
    __syncthreads();
}

// LDS written by some lanes of a wave is visible to the wave's other lanes
// afterwards
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// build_tile_list (explicit colloids) by one wave; the list also goes to
// lds_g.  Returns the list length.
// The calling wave sorts list[0..n) (n <= 64, distinct values) ascending in
// place: each lane's rank is the number of smaller values.
__device__ inline void wave_sort_list(int *list, int n) {
    const int lane = threadIdx.x & 63;
    const int v = lane < n ? list[lane] : 0x7fffffff;
    int rank = 0;
    for (int j = 0; j < n; j++) rank += lane_value(v, j) < v ? 1 : 0;
    if (lane < n) list[rank] = v;
}
__device__ inline int build_tile_list_wave(const Phys &ph, const Dev &d, const Rep &R, int T, int *lds_g) {
    const int lane = threadIdx.x & 63;
    int lo[3], ln[3];
    tile_list_window(R, T, lo, ln);
    int n = 0;
    for (int base = 0; base < R.nc; base += 64) {
        const int i = base + lane, g = R.first_colloid + i;
        const bool p = i < R.nc && tile_list_member(ph, d, R, g, lo, ln, false);
        const unsigned long long b = __ballot(p);
        const int rank = __popcll(b & ((1ull << lane) - 1ull));
        if (p) {
            if (n + rank < TCL_CAP) {
                d.tcl_list[T * TCL_CAP + n + rank] = g;
                lds_g[n + rank] = g;
            } else {
                atomicOr(d.status, BW_FLAG_LIST_OVERFLOW);
            }
        }
        n += __popcll(b);
    }
    n = min(n, TCL_CAP);
    if (lane == 0) d.tcl_cnt[T] = n;
    wave_sync();
    return n;
}

// colloids listed for a tile: id, position, cell (every per-tile pass) ...
struct TileCore {
    int n;
    int g[TCL_CAP];
    double px[TCL_CAP], py[TCL_CAP], pz[TCL_CAP];
    int cx[TCL_CAP], cy[TCL_CAP], cz[TCL_CAP];
};
// ... and their run() data (the passes that count shells)
struct TileRun : TileCore {
    double rx[TCL_CAP], ry[TCL_CAP], rz[TCL_CAP];    // position at run() time
    int qx[TCL_CAP], qy[TCL_CAP], qz[TCL_CAP];       // its cell
    double vx[TCL_CAP], vy[TCL_CAP], vz[TCL_CAP];    // velocity after the kick
};
// ... and the per-particle kick (the passes that kick)
struct TileColl : TileRun {
    double kx[TCL_CAP], ky[TCL_CAP], kz[TCL_CAP];
};

// listed colloids' current and run() positions relative to a tile's origin,
// FP32, for the conservative near test of blocked streaming
struct NearColl {
    float cx[TCL_CAP], cy[TCL_CAP], cz[TCL_CAP];
    float rx[TCL_CAP], ry[TCL_CAP], rz[TCL_CAP];
};

// listed colloid g into slot `slot` of the tile's LDS list
template <class TC>
__device__ __forceinline__ void load_tile_slot(const Dev &d, const Phys &ph, int g, int slot, bool kick_data, TC &tc) {
    tc.g[slot] = g;
    tc.px[slot] = d.cpos[3 * g]; tc.py[slot] = d.cpos[3 * g + 1]; tc.pz[slot] = d.cpos[3 * g + 2];
    tc.cx[slot] = d.ccell[3 * g]; tc.cy[slot] = d.ccell[3 * g + 1]; tc.cz[slot] = d.ccell[3 * g + 2];
    if constexpr (std::is_base_of<TileRun, TC>::value) {
        if (kick_data) {
            tc.rx[slot] = d.cpos_run[3 * g]; tc.ry[slot] = d.cpos_run[3 * g + 1]; tc.rz[slot] = d.cpos_run[3 * g + 2];
            tc.qx[slot] = d.ccell_run[3 * g]; tc.qy[slot] = d.ccell_run[3 * g + 1]; tc.qz[slot] = d.ccell_run[3 * g + 2];
            tc.vx[slot] = d.cvel[3 * g]; tc.vy[slot] = d.cvel[3 * g + 1]; tc.vz[slot] = d.cvel[3 * g + 2];
        }
    }
    if constexpr (std::is_same<TileColl, TC>::value) {
        if (kick_data) {
            // cpp/run.cpp:14,23: del*(mass_colloid/(mass_fl*cnt)) with del = ra*v0
            int c = d.ccnt[g];
            double fac = c > 0 ? ph.mass_colloid / (ph.mass_fl * c) : 0.0;
            double v0 = d.cv0[g];
            tc.kx[slot] = (d.cra[3 * g] * v0) * fac;
            tc.ky[slot] = (d.cra[3 * g + 1] * v0) * fac;
            tc.kz[slot] = (d.cra[3 * g + 2] * v0) * fac;
        }
    }
}
template <class TC>
__device__ inline void load_tile_colloids(const Dev &d, const Phys &ph, const Rep &R, int r, int T, bool kick_data,
                                          TC &tc) {
    if (d.tcl_stale[r]) build_tile_list(ph, d, R, T);
    const int n = min(d.tcl_cnt[T], TCL_CAP);
    for (int slot = threadIdx.x; slot < n; slot += blockDim.x)
        load_tile_slot(d, ph, d.tcl_list[T * TCL_CAP + slot], slot, kick_data, tc);
    if (threadIdx.x == 0) tc.n = n;
    __syncthreads();
}

// Candidate colloids of a cell through per-axis bit masks in LDS.  The 7^3-1
// stencil of a colloid is a cube of cells minus its centre, so "cell in the
// cube of colloid k" is the AND over the three axes of "coordinate within
// `reach` cells of the colloid's".  For the tile's region (its box grown by
// TILE_REACH per side, where all of its particles stay) bit k of
// m[a][u] is set when axis-a coordinate u lies within reach of listed colloid
// k (k < MASK_BITS).  Every use applies the exact test afterwards (stencil
// centre, distance), so the masks only prune.
struct Region {
    int lo[3], len[3];
};
__device__ inline Region tile_region(const Rep &R, int tx, int ty, int tz) {
    Region g;
    int t[3] = {tx, ty, tz};
    for (int k = 0; k < 3; k++) {
        g.len[k] = R.B[k] + 2 * TILE_REACH;
        g.lo[k] = t[k] * R.B[k] - TILE_REACH;
        if (g.len[k] >= R.L) {
            g.len[k] = R.L;
            g.lo[k] = 0;
        }
    }
    return g;
}
// index of coordinate i (in [0,L)) along axis a of the region, -1 outside
__device__ __forceinline__ int axis_index(const Region &g, int a, int i, int L) {
    int u = i - g.lo[a];
    if (u < 0) u += L;
    else if (u >= L) u -= L;
    return u < g.len[a] ? u : -1;
}
struct AxisMask {
    unsigned m[3][AXIS_MAX];
};
// candidate bits of cell (ix,iy,iz); all bits outside the region
__device__ __forceinline__ unsigned cell_mask(const AxisMask &am, const Region &g, int ix, int iy, int iz, int L) {
    int ux = axis_index(g, 0, ix, L), uy = axis_index(g, 1, iy, L), uz = axis_index(g, 2, iz, L);
    if ((ux | uy | uz) < 0) return 0xffffffffu;
    return am.m[0][ux] & am.m[1][uy] & am.m[2][uz];
}
__device__ inline void zero_axis_masks(AxisMask &am) {
    for (int i = threadIdx.x; i < 3 * AXIS_MAX; i += blockDim.x) (&am.m[0][0])[i] = 0u;
}
// masks of the cells (cx, cy, cz)[k] of the first n listed colloids within
// `reach` into masks zeroed before the last barrier; ends with a barrier
__device__ inline void fill_axis_masks(int n, const int *cx, const int *cy, const int *cz, const Region &g, int L,
                                       int reach, AxisMask &am) {
    // the region's bounds as six scalars picked by selects (an array indexed
    // at run time would live in scratch memory, whose loads wait behind every
    // outstanding global load of the wave)
    const int lo0 = g.lo[0], lo1 = g.lo[1], lo2 = g.lo[2], len0 = g.len[0], len1 = g.len[1], len2 = g.len[2];
    const int nk = min(n, MASK_BITS);
    for (int w = threadIdx.x; w < nk * 3; w += blockDim.x) {
        const int k = w / 3, a = w - 3 * k;
        const int c = a == 0 ? cx[k] : a == 1 ? cy[k] : cz[k];
        const int lo = a == 0 ? lo0 : a == 1 ? lo1 : lo2, len = a == 0 ? len0 : a == 1 ? len1 : len2;
        for (int dd = -reach; dd <= reach; dd++) {
            int u = wrapi(c + dd, L) - lo;
            if (u < 0) u += L;
            else if (u >= L) u -= L;
            if (u < len) atomicOr(&am.m[a][u], 1u << k);
        }
    }
    __syncthreads();
}
__device__ inline void build_axis_masks(int n, const int *cx, const int *cy, const int *cz, const Region &g, int L,
                                        int reach, AxisMask &am) {
    zero_axis_masks(am);
    __syncthreads();
    fill_axis_masks(n, cx, cy, cz, g, L, reach, am);
}
// run_cells: the colloids' cells at run() time, else their current cells
__device__ inline void build_axis_masks(const TileRun &tc, const Region &g, int L, bool run_cells, int reach,
                                        AxisMask &am) {
    if (run_cells) build_axis_masks(tc.n, tc.qx, tc.qy, tc.qz, g, L, reach, am);
    else build_axis_masks(tc.n, tc.cx, tc.cy, tc.cz, g, L, reach, am);
}
// both mask sets of a tile in one pass: run() cells within reach0 into am0,
// current cells within reach1 into am1 (as two fills would), into masks
// zeroed before the last barrier; ends with a barrier
__device__ inline void fill_axis_masks2(const TileRun &tc, const Region &g, int L, int reach0, AxisMask &am0,
                                        int reach1, AxisMask &am1) {
    const int lo0 = g.lo[0], lo1 = g.lo[1], lo2 = g.lo[2], len0 = g.len[0], len1 = g.len[1], len2 = g.len[2];
    const int nk = min(tc.n, MASK_BITS);
    for (int w = threadIdx.x; w < nk * 6; w += blockDim.x) {
        const bool cur = w >= nk * 3;
        const int v = cur ? w - nk * 3 : w, k = v / 3, a = v - 3 * k;
        const int c = cur ? (a == 0 ? tc.cx[k] : a == 1 ? tc.cy[k] : tc.cz[k])
                          : (a == 0 ? tc.qx[k] : a == 1 ? tc.qy[k] : tc.qz[k]);
        const int lo = a == 0 ? lo0 : a == 1 ? lo1 : lo2, len = a == 0 ? len0 : a == 1 ? len1 : len2;
        const int reach = cur ? reach1 : reach0;
        AxisMask &am = cur ? am1 : am0;
        for (int dd = -reach; dd <= reach; dd++) {
            int u = wrapi(c + dd, L) - lo;
            if (u < 0) u += L;
            else if (u >= L) u -= L;
            if (u < len) atomicOr(&am.m[a][u], 1u << k);
        }
    }
    __syncthreads();
}

// run() rear-shell test of listed colloid k for a particle at binned cell
// (ix,iy,iz) and absolute position (x,y,z) (cpp/run.cpp:17-18)
__device__ __forceinline__ bool shell_test(const Phys &ph, const TileRun &tc, int k, int L, int ix, int iy,
                                           int iz, double x, double y, double z) {
    if (!in_stencil(ix, iy, iz, tc.qx[k], tc.qy[k], tc.qz[k], L)) return false;
    double Ld = (double)L;
    double ax = img1(x - tc.rx[k], Ld), ay = img1(y - tc.ry[k], Ld), az = img1(z - tc.rz[k], Ld);
    double d2 = (ax * ax + ay * ay) + az * az;
    double dot = (ax * tc.vx[k] + ay * tc.vy[k]) + az * tc.vz[k];
    return d2 <= ph.lim_shell && dot <= 0;
}
// candidate colloids of a particle, ascending: mask bits, then the unmasked
// tail k >= MASK_BITS (every use applies the exact stencil test)
template <class F>
__device__ __forceinline__ void for_candidates(const TileCore &tc, unsigned m, F &&f) {
    while (m) {
        int k = __ffs(m) - 1;
        m &= m - 1;
        f(k);
    }
    for (int k = MASK_BITS; k < tc.n; k++) f(k);
}
// every listed colloid, ascending (the order for_candidates gives an all-ones
// mask), unrolled so that the LDS reads of several colloids are in flight together
template <class F>
__device__ __forceinline__ void for_listed(const TileCore &tc, F &&f) {
#pragma unroll 4
    for (int k = 0; k < tc.n; k++) f(k);
}

// ------------------------------------------------------------ colloid work
// Every colloid-only reference entry point, as per-colloid functions shared by
// the one-workgroup-per-replica kernel (individual entry points) and the
// wave-per-colloid kernel of the fused step; same operations in the same order.

// cpp/update.cpp:8-13
__device__ __forceinline__ double powi(double x, int r) {
    double ans = 1;
    for (int i = 1; i <= r; i++) ans *= x;
    return ans;
}
// WCA pair force on the pair (i, j), i the lower index: temp = img(p_i - p_j)
// (cpp/update.cpp:21-28); false beyond r_cutoff
__device__ __forceinline__ bool pair_force(const Phys &ph, double L, const double *pi, const double *pj, double &fx,
                                           double &fy, double &fz) {
    double tx = img1(pi[0] - pj[0], L), ty = img1(pi[1] - pj[1], L), tz = img1(pi[2] - pj[2], L);
    double r = sqrt((tx * tx + ty * ty) + tz * tz);
    if (!(r < ph.r_cutoff)) return false;
    double t1 = ph.sig12 / powi(r, 13), t2 = ph.sig6 / powi(r, 7);
    double mag = 4.0 * ph.eps * (12.0 * t1 - 6.0 * t2) - ph.fc;
    fx = (tx * mag) / r;
    fy = (ty * mag) / r;
    fz = (tz * mag) / r;
    return true;
}
// Partner j's term of f_g in the reference's accumulation (cpp/update.cpp:28):
// f_g -= ff(j,g) for j < g, f_g += ff(g,j) for j > g.  sign carries the
// operation; false beyond r_cutoff.
__device__ __forceinline__ bool force_term(const Phys &ph, const Dev &d, double L, int g, int j, double *c,
                                           bool &sub) {
    sub = j < g;
    return sub ? pair_force(ph, L, d.cpos + 3 * j, d.cpos + 3 * g, c[0], c[1], c[2])
               : pair_force(ph, L, d.cpos + 3 * g, d.cpos + 3 * j, c[0], c[1], c[2]);
}
// The same term as one addend t: f_g + t is f_g - ff(j,g) or f_g + ff(g,j)
// bit for bit (x - y is x + (-y)), and t = +0 beyond r_cutoff leaves f_g
// unchanged (a sum of such terms started at +0 never holds -0).  pi, pj: the
// positions of g and j.
__device__ __forceinline__ void force_addend(const Phys &ph, double L, int g, int j, const double *pi,
                                             const double *pj, double *t) {
    double c[3] = {0, 0, 0};
    const bool sub = j < g;
    const bool has = sub ? pair_force(ph, L, pj, pi, c[0], c[1], c[2]) : pair_force(ph, L, pi, pj, c[0], c[1], c[2]);
    for (int k = 0; k < 3; k++) t[k] = has ? (sub ? -c[k] : c[k]) : 0.0;
}
// cpp/update.cpp:14-32 for one colloid, serially in list order
__device__ inline void colloid_force_serial(const Phys &ph, const Dev &d, double L, int g, double *f) {
    double f0 = 0, f1 = 0, f2 = 0;
    for (int m = 0; m < d.md_cnt[g]; m++) {
        double c[3];
        bool sub;
        if (force_term(ph, d, L, g, d.md_list[g * MD_CAP + m], c, sub)) {
            if (sub) { f0 = f0 - c[0]; f1 = f1 - c[1]; f2 = f2 - c[2]; }
            else { f0 = f0 + c[0]; f1 = f1 + c[1]; f2 = f2 + c[2]; }
        }
    }
    f[0] = f0; f[1] = f1; f[2] = f2;
}
// fixed-point (2^-40) value of a and back; |a| >= 2^22 counts as 0
__device__ __forceinline__ unsigned long long fixq(double a) {
    return fabs(a) < 0x1p22 ? (unsigned long long)__double2ll_rn(a * 0x1p40) : 0ull;
}
__device__ __forceinline__ double unfixq(unsigned long long q) { return (double)(long long)q * 0x1p-40; }

// How the fluid in a surrogate-driven colloid's shell pays the colloid's
// debt of the step: linear  D = n_sub * impulse + mirrored momentum change +
// run() kick (cpp/run.cpp:14), angular  A = n_sub * torque + mirrored angular
// momentum change.  Shell particle i (at r_i from the centre) gets
// dv_i = -u + alpha x r_i, with u and alpha chosen so that the shell's
// momentum changes by -D and its angular momentum about the centre by -A, as
// the explicit collisions conserve both (cpp/fluid_colloid_collision.cpp:46-47):
//   (tr C - C) alpha = -A + (S1 x D) / n,   u = (D + alpha x S1) / n,
// C = S2 - S1 S1^T / n the shell's second moment about its centroid.  lin:
// the shell holds at least SHELL_MIN_LIN particles and no share exceeds
// SHELL_MAX_KICK; ang: besides, at least SHELL_MIN_ANG particles, an inertia
// within a factor ~30 of isotropic and a rotation share below SHELL_MAX_KICK
// at the shell's edge.  Otherwise the part is not paid by the shell (u or
// alpha 0) and the colloid takes the mirrored fluid's momentum or angular
// momentum change back instead of its impulse or torque.
constexpr int SHELL_MIN_LIN = 8, SHELL_MIN_ANG = 16;
constexpr double SHELL_MAX_KICK = 2.0;  // velocity units (the thermal speed is 1)
struct ShellPay {
    double u[3], a[3];
    bool lin, ang;
};
__device__ inline ShellPay shell_pay(const Phys &ph, const Dev &d, int g) {
    ShellPay p;
    const int cnt = d.sshell[g];
    const unsigned long long *q = d.sacc + (size_t)SUR_ACC * g;
    const double ratio = ph.mass_colloid / ph.mass_fl, v0 = d.cv0[g], ns = (double)ph.n_sub;
    double D[3], A[3], s1[3];
    for (int k = 0; k < 3; k++) {
        D[k] = (ns * d.csur[6 * g + k] + unfixq(q[k])) + ratio * (d.cra[3 * g + k] * v0);
        A[k] = ns * d.csur[6 * g + 3 + k] + unfixq(q[3 + k]);
        s1[k] = unfixq(q[6 + k]);
        p.u[k] = 0.0;
        p.a[k] = 0.0;
    }
    p.lin = cnt >= SHELL_MIN_LIN;
    p.ang = false;
    if (!p.lin) return p;
    const double n = (double)cnt;
    const double cxx = unfixq(q[9]) - s1[0] * s1[0] / n, cyy = unfixq(q[10]) - s1[1] * s1[1] / n,
                 czz = unfixq(q[11]) - s1[2] * s1[2] / n, cxy = unfixq(q[12]) - s1[0] * s1[1] / n,
                 cxz = unfixq(q[13]) - s1[0] * s1[2] / n, cyz = unfixq(q[14]) - s1[1] * s1[2] / n;
    const double tr = (cxx + cyy) + czz;
    const double m00 = tr - cxx, m11 = tr - cyy, m22 = tr - czz, m01 = -cxy, m02 = -cxz, m12 = -cyz;
    const double c00 = m11 * m22 - m12 * m12, c01 = m02 * m12 - m01 * m22, c02 = m01 * m12 - m02 * m11;
    const double c11 = m00 * m22 - m02 * m02, c12 = m01 * m02 - m00 * m12, c22 = m00 * m11 - m01 * m01;
    const double det = (m00 * c00 + m01 * c01) + m02 * c02;
    p.ang = cnt >= SHELL_MIN_ANG && tr > 0.0 && det > 0.01 * tr * tr * tr;
    if (p.ang) {
        const double b0 = -A[0] + (s1[1] * D[2] - s1[2] * D[1]) / n;
        const double b1 = -A[1] + (s1[2] * D[0] - s1[0] * D[2]) / n;
        const double b2 = -A[2] + (s1[0] * D[1] - s1[1] * D[0]) / n;
        p.a[0] = ((c00 * b0 + c01 * b1) + c02 * b2) / det;
        p.a[1] = ((c01 * b0 + c11 * b1) + c12 * b2) / det;
        p.a[2] = ((c02 * b0 + c12 * b1) + c22 * b2) / det;
        const double am = sqrt((p.a[0] * p.a[0] + p.a[1] * p.a[1]) + p.a[2] * p.a[2]);
        if (!(am * sqrt(ph.lim_shell) <= SHELL_MAX_KICK)) {
            p.ang = false;
            p.a[0] = 0.0; p.a[1] = 0.0; p.a[2] = 0.0;
        }
    }
    p.u[0] = (D[0] + (p.a[1] * s1[2] - p.a[2] * s1[1])) / n;
    p.u[1] = (D[1] + (p.a[2] * s1[0] - p.a[0] * s1[2])) / n;
    p.u[2] = (D[2] + (p.a[0] * s1[1] - p.a[1] * s1[0])) / n;
    if (!(sqrt((p.u[0] * p.u[0] + p.u[1] * p.u[1]) + p.u[2] * p.u[2]) <= SHELL_MAX_KICK)) {
        p.lin = false;
        p.ang = false;
        for (int k = 0; k < 3; k++) {
            p.u[k] = 0.0;
            p.a[k] = 0.0;
        }
    }
    return p;
}

struct ColIn {  // the inputs of colloid_update, loaded ahead of the force terms
    double v[3], w[3], ra[3], f[3], of[3], p0[3], dv[3], dw[3], sv[6], imp[6], disp[3], pmd[3], v0;
    int cell[3], cmd[3];
    bool sur;
};
__device__ inline void colloid_load(const Phys &ph, const Dev &d, int g, unsigned flags, ColIn &in) {
    for (int k = 0; k < 3; k++) {
        in.v[k] = d.cvel[3 * g + k];
        in.w[k] = d.comega[3 * g + k];
        in.ra[k] = d.cra[3 * g + k];
        in.f[k] = (flags & PH_FORCE) ? 0.0 : d.cf[3 * g + k];
        in.of[k] = d.coldf[3 * g + k];
        in.p0[k] = d.cpos[3 * g + k];
        in.dv[k] = (flags & PH_APPLY) ? d.cdv[3 * g + k] : 0.0;
        in.dw[k] = (flags & PH_APPLY) ? d.cdw[3 * g + k] : 0.0;
        in.cell[k] = (flags & PH_KICK) ? d.ccell[3 * g + k] : 0;
        in.disp[k] = (flags & PH_POS) ? d.cdisp[3 * g + k] : 0.0;
        in.pmd[k] = (flags & PH_POS) ? d.cpos_md[3 * g + k] : 0.0;
        in.cmd[k] = (flags & PH_POS) ? d.ccell_md[3 * g + k] : 0;
    }
    in.sur = (flags & PH_APPLY) && ph.surrogate && d.cmode[g];
    for (int k = 0; k < 6; k++) {
        in.sv[k] = ((flags & PH_APPLY) && ph.surrogate) ? d.csur[6 * g + k] : 0.0;
        in.imp[k] = (flags & PH_APPLY) ? d.cimp[6 * g + k] : 0.0;
    }
    // a surrogate-driven colloid whose shell cannot take its impulse (no
    // fluid) or its torque (no invertible inertia) takes the mirrored fluid's
    // momentum or angular momentum change back instead
    if (in.sur) {
        const ShellPay pay = shell_pay(ph, d, g);
        const unsigned long long *q = d.sacc + (size_t)SUR_ACC * g;
        for (int k = 0; k < 3; k++) {
            if (!pay.lin) in.sv[k] = -unfixq(q[k]) / ph.n_sub;
            if (!pay.ang) in.sv[3 + k] = -unfixq(q[3 + k]) / ph.n_sub;
        }
    }
    in.v0 = (flags & PH_KICK) ? d.cv0[g] : 0.0;
}
// The per-colloid phases of one substep, in PH_* order, by one thread, on the
// inputs loaded by colloid_load: the phases run in registers and the results
// are stored at the end.  fn is the new WCA force when PH_FORCE is set; trig,
// when not null, holds sin and cos of the activity rotation angles (x, y, z)
// as this function would form them.
//  PH_APPLY     cpp/fluid_colloid_collision.cpp:46-47 (the sums dv, dw were
//               formed in the reference's event order); a surrogate-driven
//               colloid receives its sampled impulse instead, and the event
//               sums are accumulated for the surrogate's training pairs
//  PH_ACTIVITY  cpp/update.cpp:34-46
//  PH_VELOCITY  cpp/update.cpp:62-67
//  PH_KICK      cpp/run.cpp:14 (colloid part of run) + the run-time snapshot
//  PH_OLDF      cpp/active_brownian.cpp:28
//  PH_POS       cpp/update.cpp:48-54, into the other position buffer; marks
//               the replica's pair candidates / tile lists stale if the
//               colloid left their margins
__device__ inline void colloid_update(const Phys &ph, const Dev &d, int L, int r, int g, unsigned flags,
                                      const double *fn, const ColIn &in, const double *trig = nullptr) {
    double v[3], w[3], ra[3], f[3], of[3];
    for (int k = 0; k < 3; k++) {
        v[k] = in.v[k];
        w[k] = in.w[k];
        ra[k] = in.ra[k];
        f[k] = (flags & PH_FORCE) ? fn[k] : in.f[k];
        of[k] = in.of[k];
    }
    if (flags & PH_APPLY) {
        for (int k = 0; k < 3; k++) {
            const double iv = in.sur ? in.sv[k] : in.dv[k];
            const double iw = in.sur ? in.sv[3 + k] : in.dw[k];
            v[k] = v[k] + iv * ph.mass_fl / ph.mass_colloid;
            w[k] = w[k] + iw * ph.mass_fl / ph.I_colloid;
        }
    }
    if (flags & PH_ACTIVITY) {
        double sbx, sby, sbz, cbx, cby, cbz;
        if (trig) {
            sbx = trig[0]; sby = trig[1]; sbz = trig[2]; cbx = trig[3]; cby = trig[4]; cbz = trig[5];
        } else {
            double bx = w[0] * ph.dt, by = w[1] * ph.dt, bz = w[2] * ph.dt;
            sbx = sin(bx); sby = sin(by); sbz = sin(bz); cbx = cos(bx); cby = cos(by); cbz = cos(bz);
        }
        double m1x = cby * cbz, m1y = -cby * sbz, m1z = sby;
        double m2x = sbx * sby * cbz + cbx * sbz, m2y = -sbx * sby * sbz + cbx * cbz, m2z = -sbx * cby;
        double m3x = -cbx * sby * cbz + sbx * sbz, m3y = cbx * sby * sbz + sbx * cbz, m3z = cbx * cby;
        double a = ra[0], b = ra[1], c = ra[2];
        ra[0] = (m1x * a + m1y * b) + m1z * c;
        ra[1] = (m2x * a + m2y * b) + m2z * c;
        ra[2] = (m3x * a + m3y * b) + m3z * c;
    }
    if (flags & PH_VELOCITY) {
        double dtb2 = ph.dt / (ph.mass_colloid * 2);
        for (int k = 0; k < 3; k++) v[k] = v[k] + (of[k] + f[k]) * dtb2;
    }
    if (flags & PH_KICK) {
        for (int k = 0; k < 3; k++) {
            v[k] = v[k] + ra[k] * in.v0;
            d.cpos_run[3 * g + k] = in.p0[k];
            d.ccell_run[3 * g + k] = in.cell[k];
        }
        d.ccnt[g] = 0;
        d.sshell[g] = 0;
        for (int k = 0; k < SUR_ACC; k++) d.sacc[(size_t)SUR_ACC * g + k] = 0;
    }
    if (flags & PH_OLDF)
        for (int k = 0; k < 3; k++) of[k] = f[k];
    if (flags & PH_POS) {
        double dt2 = ph.dt * ph.dt, ddt = 0.5 * dt2 / ph.mass_colloid;
        bool far_tl = false;
        double m2 = 0;
        for (int k = 0; k < 3; k++) {
            double dx = v[k] * ph.dt + f[k] * ddt;
            double p = in.p0[k] + dx;
            d.cdisp[3 * g + k] = in.disp[k] + dx;
            p = mod1(p, (double)L);
            d.cpos_nxt[3 * g + k] = p;
            int cc = coord_cell(p, L);
            d.ccell_nxt[3 * g + k] = cc;
            int dm = pdiff(cc, in.cmd[k], L);
            if (dm > 1 || dm < -1) far_tl = true;
            double md = img1(p - in.pmd[k], (double)L);
            m2 += md * md;
        }
        if (m2 > 0.25 * MD_SKIN * MD_SKIN) atomicOr(&d.md_stale[r], 1);
        if (far_tl) atomicOr(&d.tcl_stale[r], 1);
        if ((flags & PH_MARGIN) && m2 > ph.margin * ph.margin) atomicOr(d.status, BW_FLAG_MARGIN);
    }
    for (int k = 0; k < 3; k++) {
        if (flags & PH_APPLY) {
            d.cimp[6 * g + k] = in.imp[k] + in.dv[k];
            d.cimp[6 * g + 3 + k] = in.imp[3 + k] + in.dw[k];
            d.cdv[3 * g + k] = 0;
            d.cdw[3 * g + k] = 0;
        }
        if (flags & (PH_APPLY | PH_VELOCITY | PH_KICK)) d.cvel[3 * g + k] = v[k];
        if (flags & PH_APPLY) d.comega[3 * g + k] = w[k];
        if (flags & PH_ACTIVITY) d.cra[3 * g + k] = ra[k];
        if (flags & PH_FORCE) d.cf[3 * g + k] = f[k];
        if (flags & PH_OLDF) d.coldf[3 * g + k] = of[k];
    }
}
// pair test of cpp/neighbour_list.cpp:33-34 (symmetric: img(-a) = -img(a)) on
// the positions pi of g and pj of j
__device__ __forceinline__ bool md_pair_at(const Phys &ph, double L, int g, int j, const double *pi,
                                           const double *pj) {
    double tx, ty, tz;
    if (j > g) {
        tx = img1(pi[0] - pj[0], L); ty = img1(pi[1] - pj[1], L); tz = img1(pi[2] - pj[2], L);
    } else {
        tx = img1(pj[0] - pi[0], L); ty = img1(pj[1] - pi[1], L); tz = img1(pj[2] - pi[2], L);
    }
    return (tx * tx + ty * ty) + tz * tz < ph.lim_md;
}
__device__ __forceinline__ bool md_pair(const Phys &ph, const Dev &d, double L, int g, int j) {
    return md_pair_at(ph, L, g, j, d.cpos + 3 * g, d.cpos + 3 * j);
}
// the pair list of colloid g (all partners, ascending), by one thread
__device__ inline void colloid_mdlist_serial(const Phys &ph, const Dev &d, const Rep &R, int r, int g) {
    const bool full = d.md_stale[r] != 0;
    int ncand = full ? R.nc : d.mdc_cnt[g], w = 0;
    for (int m = 0; m < ncand; m++) {
        int j = full ? R.first_colloid + m : d.mdc_list[g * MDC_CAP + m];
        if (j == g || !md_pair(ph, d, (double)R.L, g, j)) continue;
        if (w < MD_CAP) d.md_list[g * MD_CAP + w] = j;
        w++;
    }
    if (w > MD_CAP) { atomicOr(d.status, BW_FLAG_LIST_OVERFLOW); w = MD_CAP; }
    d.md_cnt[g] = w;
}

// All colloid work of one replica by one workgroup, phases in PH_* order.
__device__ inline void colloid_phases(const Phys &ph, const Dev &d, const Rep &R, int r, unsigned flags) {
    const int g0 = R.first_colloid, nc = R.nc;
    const double L = (double)R.L;
    for (int i = threadIdx.x; i < nc; i += blockDim.x) {
        int g = g0 + i;
        ColIn in;
        colloid_load(ph, d, g, flags, in);
        double fn[3] = {0, 0, 0};
        if (flags & PH_FORCE) colloid_force_serial(ph, d, L, g, fn);
        colloid_update(ph, d, R.L, r, g, flags, fn, in);
    }
    __syncthreads();
    if (flags & PH_MDLIST)
        for (int i = threadIdx.x; i < nc; i += blockDim.x) colloid_mdlist_serial(ph, d, R, r, g0 + i);
    __syncthreads();
    if (flags & PH_POTENTIAL) {
        // cpp/update.cpp:24, one running sum over (i ascending, j > i ascending)
        if (threadIdx.x == 0) {
            double pot = 0;
            for (int i = 0; i < nc; i++) {
                int g = g0 + i;
                for (int m = 0; m < d.md_cnt[g]; m++) {
                    int j = d.md_list[g * MD_CAP + m];
                    if (j <= g) continue;
                    const double *pi = d.cpos + 3 * g, *pj = d.cpos + 3 * j;
                    double tx = img1(pi[0] - pj[0], L), ty = img1(pi[1] - pj[1], L), tz = img1(pi[2] - pj[2], L);
                    double rr = sqrt((tx * tx + ty * ty) + tz * tz);
                    if (rr < ph.r_cutoff)
                        pot += 4 * ph.eps * (pow(ph.sig_colloid / rr, 12) - pow(ph.sig_colloid / rr, 6)) - ph.ufc +
                               ph.fc * rr;
                }
            }
            d.cpot[r] = pot;
        }
    }
    __syncthreads();
}

}  // namespace bw
