// Shared host/device definitions of libbrownian.
//
// World: n_rep independent replicas (a plain run is one replica).  Fluid
// particles live in tiles -- blocks of Bx x By x Bz cells of the SRD grid as it
// was shifted at the last rotation step -- each tile owning a fixed range of
// `cap` slots in structure-of-arrays buffers.  A particle's position is an
// integer cell, packed as ix | iy << 10 | iz << 20, plus an offset in that cell
// stored in Real (float or double); velocities are Real.  Colloids are few and
// kept in FP64; their positions and cells are double-buffered so that one
// kernel can read every colloid's position (forces) while writing the next.
#pragma once
#include <stdint.h>

#include <algorithm>

namespace bw {

constexpr int EV_CAP = 64;          // reflection events per colloid per substep
constexpr int CHAIN_CAP = 1024;     // serially replayed particles per replica per substep
constexpr int CONT_CAP = 512;       // continuations per replica per substep
constexpr int SERIAL_CAP = 2048;    // serial list incl. entries of overflowed buckets
constexpr int TCL_CAP = 64;         // colloids listed for one tile
constexpr int SUR_ACC = 15;         // fixed-point sums per surrogate-driven colloid and step
constexpr int MASK_BITS = 32;       // listed colloids covered by the LDS stencil masks
constexpr int AXIS_MAX = 64;        // cells per axis of a tile's region (box grown by TILE_REACH)
constexpr int CELL_BITS = 10;       // packed cell coordinates: L <= 1023
constexpr int MD_CAP = 128;         // full MD neighbour list per colloid
constexpr int MDC_CAP = 256;        // MD candidates (cutoff + skin) per colloid
constexpr int MAX_CAP = 4096;       // slots per tile upper bound
constexpr int STREAM_THREADS = 256;
constexpr int NEAR_BLOCKS = 2048;    // near-particle pass: at most this many workgroups of 4 waves
constexpr int DEPART_THREADS = 256;
constexpr int COLLIDE_THREADS = 256; // 4 colloids (one wave each) per workgroup
constexpr int OVF_CAP = 64;         // colloids with an overflowed event bucket, per replica
constexpr int NDIR = 27;
constexpr int TILE_REACH = 3;       // cells a particle may stand outside its tile box
constexpr double MD_SKIN = 2.0;     // candidate-list skin for the colloid pair list
constexpr double STREAM_MARGIN = 0.25;  // colloid displacement bound within a step (blocked streaming)
constexpr int N_FEAT = 144;         // surrogate features per colloid
constexpr int COARSE_MAX = 64;      // coarse cells of one tile
constexpr long long MEAN_CHUNK = 65536;  // particles per workgroup of the start-up's mean
// workgroups per replica of the start-up's mean for replicas of up to max_nf particles
inline int init_mean_chunks(long long max_nf) { return (int)std::max(1LL, (max_nf + MEAN_CHUNK - 1) / MEAN_CHUNK); }

// colloid phase flags (applied in this order)
enum : unsigned {
    PH_APPLY = 1u << 0,     // v += dv*m_f/m_c, w += dw*m_f/I   (cpp/fluid_colloid_collision.cpp:46-47)
    PH_ACTIVITY = 1u << 1,  // cpp/update.cpp:34-46
    PH_FORCE = 1u << 2,     // cpp/update.cpp:14-32 (force only)
    PH_VELOCITY = 1u << 3,  // cpp/update.cpp:62-67
    PH_KICK = 1u << 4,      // cpp/run.cpp:14 (colloid part of run) + run-time snapshot
    PH_OLDF = 1u << 5,      // cpp/active_brownian.cpp:28
    PH_POS = 1u << 6,       // cpp/update.cpp:48-54
    PH_MDLIST = 1u << 7,    // filter candidates into the pair list (cpp/neighbour_list.cpp:27-39)
    PH_POTENTIAL = 1u << 8, // potential in the reference's summation order
    PH_RESET_EVENTS = 1u << 9,
    PH_MARGIN = 1u << 10,   // check the colloid displacement bound of blocked streaming
};

struct Phys {
    double kbt, mass_fl, mass_colloid, sig_colloid, eps, sigma, I_colloid;
    double ndt, dt;
    double theta, ct, st, ict;        // SRD angle and its cos/sin/(1-cos)
    double r_cutoff, fc, ufc, sig12, sig6;
    double lim_coll;                  // pow(sigma,2)*0.25
    double lim_shell;                 // pow(sigma*0.5+0.5,2)
    double lim_md;                    // pow(3*sig_colloid,2)
    double lim_md_skin;               // (3*sig_colloid + skin)^2
    double vscale_thermal;            // sqrt(kbt/mass_fl)
    // blocked streaming (fused step): a particle is "near" when within
    // near_shell2 (squared) of a colloid's run() position, or within
    // near_base + its travel in the step of a colloid's position, or when its
    // travel exceeds vcap_ndt; colloids must stay within `margin` of their
    // position at the step's start
    double margin, near_base, near_shell2, vcap_ndt;
    int near_reach;                   // per-axis cell reach of the near test
    int moments;                      // SRD pass writes coarse-cell moments (surrogate inputs)
    int surrogate;                    // colloids may be driven by the surrogate impulses
    int n_sub;
    int cap;                          // slots per tile
    int srd_threads;                  // SRD workgroup size
    int srd_pt;                       // SRD slots per thread (srd_threads x srd_pt >= cap)
    int near_blocks;                  // grid of the near-particle pass
    int near_wrap;                    // the SRD near test needs the periodic wrap (a small box)
    int n_rep, n_tiles, n_colloid_total;
    int draws;                        // BW_DRAWS_*
};

struct Rep {
    int L, nf, nc;
    int first_colloid;
    int B[3], nt[3];
    int first_tile, n_tiles;
    long long first_cell;   // offset into per-cell arrays
    long long first_fluid;  // offset into id-indexed per-particle arrays
    long long first_coarse; // offset into coarse-cell arrays (2x2x2 cells of the shifted grid)
    uint32_t seed_lo, seed_hi;
};

template <class Real>
struct Fluid {
    int *cell;
    Real *ox, *oy, *oz;
    Real *vx, *vy, *vz;
    int *pid;
};

struct Dev {
    const Rep *rep;
    const int *tile_rep;
    int *count;          // [n_tiles] particles in each tile (current buffer)
    int *count_next;     // [n_tiles] for the buffer being written
    int *near_cnt;       // [n_tiles] particles stored first that the substeps touch
    int *chunk_cnt;      // [1] chunks of near particles this step
    int *chunk_tile;     // [n_tiles * cap/64] chunk -> tile << 6 | chunk index in the tile
    int *dep_off;        // [n_tiles * 28]
    int *dep_idx;        // [n_tiles * cap]
    // colloids (global index g), FP64, 3 per colloid
    double *cpos, *cpos_nxt;  // current / next (written by update_pos_md)
    double *cvel, *comega, *cf, *coldf, *cra, *cdisp, *cpos_run, *cdv, *cdw, *cpos_md;
    double *cv0;
    int *ccell, *ccell_nxt;
    int *crep, *ccell_run, *ccell_md, *ccnt;
    int *md_stale, *tcl_stale;  // per replica: pair candidates / tile lists outdated
    int *tcl_cnt, *tcl_list;  // per tile: colloids that can touch it this step
    int *tsl_cnt, *tsl_list;  // per tile: surrogate-driven colloids near it this step
    int *md_cnt, *md_list, *mdc_cnt, *mdc_list;
    double *md_f;        // [NC * MD_CAP * 3] per pair-list entry: its signed WCA term of the force
    double *cpot;        // [n_rep] potential
    int *ev_cnt, *ev_slot, *chain_cnt, *chain_slot;
    int *cont_cnt, *cont_slot, *cont_from;  // continuations of the colloid loop
    double *cont_dump;
    int *ovf_cnt, *ovf_list, *ev_total;     // per replica: overflowed buckets, events seen
    unsigned *status;
    long long *step;     // MPCD step counter (device generator key)
    double *shift;       // [n_rep*3] shift of the current step
    // injected draws
    const double *inj_shift;   // [n_rep*3]
    const double *inj_axes;    // [cells_total*2]
    const int *inj_off;        // [n_sub*n_colloid_total + 1]
    const int *inj_pid;
    const double *inj_val;     // 6 per tuple
    // maps
    int *map_count, *map_start, *map_fill, *map_list;
    int *neigh_cnt, *neigh_list;
    int neigh_cap;
    int *shifted_cell;         // id-indexed
    double *audit_partial;     // [n_tiles * 8]
    // surrogate closure
    float *cmom;               // [coarse cells * 5]: count, sum v (3), sum |v|^2 after the last SRD pass
    int *mom_valid;            // the moments describe the current fluid
    int *cmode;                // per colloid: 0 explicit collisions, 1 surrogate impulse
    int *cmode_prev;           // per colloid: the mode of the previous step
    double *csur;              // per colloid: surrogate impulse per substep (6, cdv/cdw units)
    double *cimp;              // per colloid: collision impulse accumulated since the features (6)
    // per surrogate-driven colloid and step, formed by the rotation pass: the
    // fluid in its shell (distance <= sigma/2 + 0.5 from the centre), which
    // pays the colloid's sampled impulse and torque, its run() kick and the
    // mirrored fluid's momentum changes; and 15 fixed-point (2^-40) sums in
    // units of the fluid mass: the mirrored fluid's momentum change (3) and
    // angular momentum change about the centre (3), and the shell's first (3)
    // and second (xx yy zz xy xz yz) moments of position about the centre
    int *sshell;
    unsigned long long *sacc;
    float *feat;               // per colloid: surrogate features (144)
    float *sur_out;            // per colloid: impulse mean (6) and log-variance (6)
    const float *sur_lo, *sur_hi;  // feature range of the training data (144 each), then
                                   // impulse range (6 each) at offset N_FEAT
};

// Coarse bins of the colloids for the device start-up's exclusion test: per
// replica a G^3 grid of edge L/G >= sigma/2 (CSR lists of global colloid ids)
struct InitBins {
    const int *G;              // [n_rep]
    const long long *first;    // [n_rep] first bin of the replica
    const int *start;          // [bins + 1]
    const int *list;           // colloid ids
    const double *cpos;        // colloid positions (global)
};

}  // namespace bw
