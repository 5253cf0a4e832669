// PP device structures shared by the cell-list build and the pair kernels.
#pragma once
#include "common.h"

namespace cube2 {

struct PPTile {
    int rcp;           // PP cells per coarse cell (max level of the tile's subtiles)
    int S;             // cells per dim of the tile grid, nt*rcp + 2*rcp
    long long cell_off;  // first cell of the tile in the concatenated grids
};

struct PPRed {
    unsigned long long p_cand, p_acc;  // examined / accepted ordered pairs
    unsigned f2max_bits;               // max |a|^2 (non-negative float bits)
    unsigned vmax_bits[3];             // max |vp| per component
    int err;
    int nbig;    // coarse cells listed for the block-per-cell count and scatter
    int ndense;  // PP cells of >= 64 particles listed for the sub-cell sort
    int pad;
};

// Learned stand-in (learned.hip): a level-2 subtile whose fine cells the
// network evaluates ("job"): its parent tile and its origin in fine cells
// (tile-relative, level 2).
struct LearnedJob {
    int t;
    int o[3];
};

// What the learned kernels need to pick the substituted particles (physical
// particles of level-2 subtiles in PP cells of fewer than `occupancy`) and
// place them in their job's grids.
struct LearnedArgs {
    const LearnedJob* jobs;
    const int* job_of;  // [ntile * nns^3] subtile location -> job, or -1
    float range;        // the level-2 PP range (apm3 of level 2)
    int occupancy;
    float rsf2;         // fine cells per standard cell at level 2
    float sub_std;      // standard cells per subtile side
    int nns;
    float y_scale;      // network output unit (acceleration)
};

struct PPArgs {
    const float4* pos4;        // sorted (x, y, z, range); range 0 = halo source
    const uint32_t* slot;      // sorted -> xp/vp slot
    const uint32_t* cellid;    // sorted -> global cell (tile offset + key)
    const int* cell_start;     // c_tot + 1 exclusive starts
    const PPTile* tiles;
    const int* n_sorted;       // particles in the cell lists (device, from the scan)
    int ntile, cap, ratio_cs;  // cap: capacity of the sorted arrays
    float app;
    const float* scal;         // {mass, a_mid, dt} (device)
    const double* dv64;        // pending PM3 Delta-v (FP64) or nullptr
    float* vp;
    // split pass (non-null): the pair kernel stores each active particle's
    // velocity term a * m^2 * a_mid * dt / 6 (FP32, before the division by
    // pi) at term[3 * slot], and the velocity update (k_pp_apply_rows or
    // k_pp_apply) applies it after PM3 has joined; vp and dv64 are then
    // unused by the pair kernel
    float* term;
    PPRed* red;
    // per parent tile {P_cand, P_acc} [ntile][2], or nullptr (counted only
    // when requested: cube2_pp_count_tiles)
    unsigned long long* tile_pairs;
    // learned kick: chunks of PP cells of fewer than learned_occ particles
    // with range learned_range are left to the network (0: none)
    int learned_occ;
    float learned_range;
    // pairs-once pass (non-null): each pair of two active particles of one
    // range is evaluated once and its force added to both, the per-particle
    // sums kept as signed 64-bit fixed point (value * react_scale) at
    // react[3 * sorted index]; k_pp_newton_terms turns them into term
    unsigned long long* react;
    float react_scale;
    int ntt;  // coarse cells per subtile side
};

int launch_pp_pairs(const PPArgs& a, int mode, hipStream_t s);

// Whether the pair kernel has a pairs-once variant for this PP mode.
bool pp_newton_mode(int mode);

// Whether PP calls count pairs per parent tile (off by default).
bool& pp_count_tiles();

// log2 of the pair kernel's chunk size in PP cells of >= 64 particles: 16
// built in, 32 with pairs once (half the reaction atomics, which share HBM
// with the mesh kernels in the overlapped kick: at 256^3 the kick 112.3 ->
// 111.4 ms, the pair kernel alone 64.5 -> 67.2 ms)
inline int pp_dense_lg(bool newton) {
    const int v = ctx().tune[CUBE2_TUNE_PP_DENSE_LG];
    return v ? v : (newton ? 5 : 4);
}

// j candidates per batch of the pair kernel: 64 (built in) or 128.
inline int pp_jbatch() { return ctx().tune[CUBE2_TUNE_PP_JBATCH] == 2 ? 128 : 64; }

// Chunks of > 2^(v-1) particles sum directly (v = 2 built in).
inline int pp_direct_lg() {
    const int v = ctx().tune[CUBE2_TUNE_PP_DIRECT_LG];
    return v ? v : 2;
}

// Whether split PP passes evaluate each pair of one range once (built in).
inline bool pp_newton_on() { return ctx().tune[CUBE2_TUNE_PP_NEWTON] != 2; }

// log2 of the estimated candidate pairs from which the pair kernel takes a
// chunk ahead of the others (heaviest first), or 0: all in cell order
// (built in).
inline int pp_heavy_lg() {
    const int v = ctx().tune[CUBE2_TUNE_PP_ORDER];
    return v > 1 ? v : 0;
}

}  // namespace cube2
