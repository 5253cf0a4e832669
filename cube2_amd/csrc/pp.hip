// PP: cell lists by counting sort and the softened pair sum, all parent
// tiles of the image in one pass.
//
// Reference semantics (pp_force_kernel.cu:625-958, pp_force_kernel.c:134-391):
//   * a parent tile is gridded with PP cells of 1/rcp coarse cells, rcp the
//     finest level among its subtiles (pp_force_kernel.cu:641-661);
//   * particles of the tile's coarse cells plus one halo cell are binned by
//     key = (i2*S + i1)*S + i0, i_d = floor(rcp*x_d) + rcp, S = nt*rcp + 2*rcp
//     (pp_force_kernel.cu:764-810) and stably sorted (:858-861);
//   * each physical particle i sums (F_ra(r,app) - F_ra(r,range_i)) d/r over
//     the j of the (2R+1)^3 cells around it with 0 < r^2 < range_i^2,
//     R = ceil(range_i*rcp/ratio_cs) (:192-250);
//   * a *= m^2, f2max = max |a|^2, Delta-v = a*a_mid*dt/(6 pi) (:252-301).
//
// MI355X design: the particles of one coarse cell are contiguous in xp (the
// spine layout) and all fall into that cell's rcp^3 PP cells, so one
// wavefront per coarse cell builds the histogram in LDS (no global atomics),
// a device scan turns counts into cell starts, and the same wavefront
// scatters its particles to their sorted slots with ballot-ranked, order-
// preserving offsets (a stable counting sort == the reference's stable radix
// sort).  Sorted particles are one float4 stream (x, y, z, range) with a
// parallel slot/cell-id stream.  The pair kernel walks the 3 contiguous
// x-runs of each (y,z) neighbour row.  The pending PM3 Delta-v (FP64) and the
// PP Delta-v are applied to vp in one FP64 add per component.
#include <algorithm>
#include <chrono>

#include "common.h"
#include "engine.h"
#include "pp_kernels.h"
#include "profile.h"
#include "reduce.h"
#include "scan.h"

namespace cube2 {

int launch_pp_tiled(const PPArgs& a, const int* chunks, const int* nchunks, int grid, int mode,
                    hipStream_t s);
int pp_tiled_grid(int mode, bool overlap, bool newton);

struct PPState {
    DevBuf nzero, tiles, ires_loc, apm3, counts, scan_tmp, pos4, slot, cellid, red, tile_first;
    DevBuf chunk_flags, chunks, big, dense;
    // the scatter's output when the dense cells are then reordered (all
    // modes but EXACT): k_chunk_flags and k_pp_subsort copy it to pos4 / slot
    DevBuf raw_pos4, raw_slot;
    DevBuf order, order_bin, order_hist;  // chunks heaviest first (k_chunk_cost/order)
    DevBuf term;  // split pass: per slot velocity term [n_used][3] (active slots)
    DevBuf rows_len, rows_base;  // split pass: physical rows' slot counts (scanned) and first slots
    DevBuf react;  // pairs-once pass: per sorted particle 64-bit fixed-point sums [cap][3]
    HostBuf h_res;  // PPRed, per-location counts, per-tile pairs, tile_first (pinned)
    std::vector<PPTile> h_tiles;
    std::vector<uint64_t> tile_pairs;  // [ntile][2] of the last call (when counted)
    PPPlan last;                       // plan of the last completed call
    bool count_tiles = false;
    int64_t c_tot = 0;
    int n_tot = 0;
    bool valid = false;
};
static PPState g_pp;

bool& pp_count_tiles() { return g_pp.count_tiles; }

// ---------------------------------------------------------------- kernels

// nzero[t][k][j][i] = idx_b_r[t][k][j] - sum_{q>=i} rhoc[t][k][j][q]
// (pp_force_kernel.cu:744-750): first slot of each coarse cell.
__global__ void k_nzero(const int* __restrict__ rhoc, const int64_t* __restrict__ idx,
                        int64_t* __restrict__ nz, int nte, int64_t nrows) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int* row = rhoc + r * nte;
    int64_t* out = nz + r * nte;
    int64_t s = 0, end = idx[r];
    for (int i = nte - 1; i >= 0; i--) {
        s += row[i];
        out[i] = end - s;
    }
}

struct CellRef {
    int t, ci, cj, ck;  // tile and extended-local coarse coords in [0, nt+2)
    int64_t cell;       // index into rhoc / nzero
};

__device__ inline CellRef coarse_cell(int64_t w, int nt, int ncb, int nte) {
    const int ne = nt + 2;
    const int64_t per = (int64_t)ne * ne * ne;
    CellRef c;
    c.t = (int)(w / per);
    int64_t r = w - (int64_t)c.t * per;
    c.ci = (int)(r % ne);
    c.cj = (int)((r / ne) % ne);
    c.ck = (int)(r / ((int64_t)ne * ne));
    const int i = c.ci + ncb - 1, j = c.cj + ncb - 1, k = c.ck + ncb - 1;
    c.cell = (((int64_t)c.t * nte + k) * nte + j) * nte + i;
    return c;
}

// PP cell coordinate of one axis: floor(rcp * x0) + rcp with
// x0 = (float)((i - ncb) + frac) (pp_force_kernel.cu:764-803).
__device__ inline int pp_index(float x0, int rcp) {
    return (int)floorf(__fmul_rn((float)rcp, x0)) + rcp;
}

__device__ inline float decode_coarse(int rel, int16_t v) {
    return (float)((double)rel + frac16(v));
}

constexpr int kWaves = 4;
constexpr int kMaxSub = 12 * 12 * 12;
// Coarse cells of more than kBigCell particles (C1's clumps put up to 115,201
// in one) are counted and scattered by a block of kBigWaves wavefronts each,
// not by one wavefront, whose serial loop over them outlasted the rest of
// the launch (scatter 3.0 ms, of which 2.3 the largest cells' tail).
constexpr int kBigCell = 2048;
constexpr int kBigWaves = 8;
constexpr int kBigBlocks = 512;  // grid of the big-cell kernels (grid-stride over the list)

// PP sub-cell of a particle of coarse cell c, clamped into the cell's rcp^3
// sub-cells (out: whether it had to be), and its decoded coarse coordinates.
__device__ inline int pp_sub(const CellRef& c, int rcp, const int16_t* p, float& x0, float& x1,
                             float& x2, bool& out) {
    x0 = decode_coarse(c.ci - 1, p[0]);
    x1 = decode_coarse(c.cj - 1, p[1]);
    x2 = decode_coarse(c.ck - 1, p[2]);
    int bx = pp_index(x0, rcp) - c.ci * rcp;
    int by = pp_index(x1, rcp) - c.cj * rcp;
    int bz = pp_index(x2, rcp) - c.ck * rcp;
    out = (unsigned)bx >= (unsigned)rcp || (unsigned)by >= (unsigned)rcp ||
          (unsigned)bz >= (unsigned)rcp;
    bx = min(max(bx, 0), rcp - 1);
    by = min(max(by, 0), rcp - 1);
    bz = min(max(bz, 0), rcp - 1);
    return (bz * rcp + by) * rcp + bx;
}

// Tile-local PP cell key of sub-cell b of coarse cell c.
__device__ inline int64_t pp_key(const CellRef& c, int rcp, int S, int b) {
    const int bx = b % rcp, by = (b / rcp) % rcp, bz = b / (rcp * rcp);
    return ((int64_t)(c.ck * rcp + bz) * S + (c.cj * rcp + by)) * S + (c.ci * rcp + bx);
}

// One wavefront per coarse cell of each tile's extended region: histogram
// of its particles over its rcp^3 PP cells; writes every one of those
// counts (the cells tile the tile's S^3 grid, so no clearing is needed).
__global__ __launch_bounds__(256) void k_pp_count(const int* __restrict__ rhoc,
                                                  const int64_t* __restrict__ nzero,
                                                  const int16_t* __restrict__ xp,
                                                  const PPTile* __restrict__ tiles, int nt,
                                                  int ncb, int nte, int64_t ncells, int64_t n_used,
                                                  int* __restrict__ counts,
                                                  int* __restrict__ err,
                                                  int* __restrict__ big_list,
                                                  int* __restrict__ nbig) {
    __shared__ int bins[kWaves][kMaxSub];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * kWaves + wave;
    const bool ok = w < ncells;
    CellRef c{};
    int rcp = 1, S = 1, nb = 0;
    int64_t off = 0;
    if (ok) {
        c = coarse_cell(w, nt, ncb, nte);
        PPTile T = tiles[c.t];
        rcp = T.rcp;
        S = T.S;
        off = T.cell_off;
        nb = rcp * rcp * rcp;
        for (int b = lane; b < nb; b += 64) bins[wave][b] = 0;
    }
    __syncthreads();
    bool big = false;
    if (ok) {
        int np = rhoc[c.cell];
        const int64_t base = nzero[c.cell];
        if (np < 0 || base < 0 || base + np > n_used) {  // rhoc / idx_b_r inconsistent
            if (lane == 0) atomicOr(err, 2);
            np = 0;
        }
        big = np > kBigCell;
        if (big) {
            if (lane == 0) big_list[atomicAdd(nbig, 1)] = (int)w;
            np = 0;
        }
        for (int l = lane; l < np; l += 64) {
            float x0, x1, x2;
            bool out;
            const int b = pp_sub(c, rcp, xp + 3 * (base + l), x0, x1, x2, out);
            if (out) atomicOr(err, 1);
            atomicAdd(&bins[wave][b], 1);
        }
    }
    __syncthreads();
    if (ok && !big) {
        for (int b = lane; b < nb; b += 64) counts[off + pp_key(c, rcp, S, b)] = bins[wave][b];
    }
}

// The listed coarse cells, one block each: the same histogram.
__global__ __launch_bounds__(1024) void k_pp_count_big(const int* __restrict__ rhoc,
                                                       const int64_t* __restrict__ nzero,
                                                       const int16_t* __restrict__ xp,
                                                       const PPTile* __restrict__ tiles, int nt,
                                                       int ncb, int nte,
                                                       const int* __restrict__ big_list,
                                                       const int* __restrict__ nbig,
                                                       int* __restrict__ counts,
                                                       int* __restrict__ err) {
    __shared__ int bins[kMaxSub];
    const int nbg = *nbig;
    for (int i = blockIdx.x; i < nbg; i += gridDim.x) {
        const CellRef c = coarse_cell(big_list[i], nt, ncb, nte);
        const PPTile T = tiles[c.t];
        const int rcp = T.rcp, nb = rcp * rcp * rcp;
        for (int b = threadIdx.x; b < nb; b += blockDim.x) bins[b] = 0;
        __syncthreads();
        const int np = rhoc[c.cell];  // listed cells passed the consistency check
        const int64_t base = nzero[c.cell];
        for (int l = threadIdx.x; l < np; l += blockDim.x) {
            float x0, x1, x2;
            bool out;
            const int b = pp_sub(c, rcp, xp + 3 * (base + l), x0, x1, x2, out);
            if (out) atomicOr(err, 1);
            atomicAdd(&bins[b], 1);
        }
        __syncthreads();
        for (int b = threadIdx.x; b < nb; b += blockDim.x)
            counts[T.cell_off + pp_key(c, rcp, T.S, b)] = bins[b];
        __syncthreads();  // bins are cleared for the next cell
    }
}

// Range of the physical particles of coarse cell c (their subtile's level),
// 0 for halo copies.
__device__ inline float cell_range(const CellRef& c, const int* __restrict__ ires_loc,
                                   const float* __restrict__ apm3, int nt, int nnt, int nns,
                                   int ntt) {
    const bool phys = c.ci >= 1 && c.ci <= nt && c.cj >= 1 && c.cj <= nt && c.ck >= 1 && c.ck <= nt;
    if (!phys) return 0.0f;
    const int tx = c.t % nnt, ty = (c.t / nnt) % nnt, tz = c.t / (nnt * nnt);
    const int sx = (c.ci - 1) / ntt, sy = (c.cj - 1) / ntt, sz = (c.ck - 1) / ntt;
    const int loc = (((tz * nnt + ty) * nnt + tx) * nns + sz) * nns * nns + sy * nns + sx;
    return apm3[ires_loc[loc] - 1];
}

// Scatters particles l_begin .. l_end-1 of coarse cell c (one wavefront, in
// slot order) to cell_start[key] + cur[sub-cell] + (rank among the earlier
// particles of its PP cell in this run), advancing cur.  Ranks come from
// comparing every lane's sub-cell (scalar broadcasts, no dependency chain);
// the last lane of each sub-cell advances its cursor.
__device__ inline void scatter_run(const CellRef& c, int rcp, int S, int64_t off,
                                   const int16_t* __restrict__ xp, int64_t base, int l_begin,
                                   int l_end, float range, int ratio_cs, int* cur,
                                   const int* __restrict__ cell_start, float4* __restrict__ pos4,
                                   uint32_t* __restrict__ slot, uint32_t* __restrict__ cellid,
                                   int lane) {
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int l0 = l_begin; l0 < l_end; l0 += 64) {
        const int l = l0 + lane;
        const bool valid = l < l_end;
        int b = -1;
        float x0 = 0, x1 = 0, x2 = 0;
        if (valid) {
            bool out;
            b = pp_sub(c, rcp, xp + 3 * (base + l), x0, x1, x2, out);
        }
        uint64_t same = 0;
#pragma unroll
        for (int k = 0; k < 64; k++)
            same |= (uint64_t)(__builtin_amdgcn_readlane(b, k) == b) << k;
        int mypos = 0;
        if (valid) mypos = cur[b] + __popcll(same & lt);
        __builtin_amdgcn_wave_barrier();
        if (valid && !(same & ~lt & ~(1ull << lane))) cur[b] += __popcll(same & lt) + 1;
        __builtin_amdgcn_wave_barrier();
        if (valid) {
            const int64_t key = pp_key(c, rcp, S, b);
            const int dst = cell_start[off + key] + mypos;
            pos4[dst] = make_float4((float)ratio_cs * x0, (float)ratio_cs * x1,
                                    (float)ratio_cs * x2, range);
            slot[dst] = (uint32_t)(base + l);
            cellid[dst] = (uint32_t)(off + key);
        }
    }
}

// Same traversal; scatters each particle to cell_start[key] + (rank among
// the earlier particles of its PP cell), preserving slot order within a
// cell.  Cells of more than kBigCell particles are left to k_pp_scatter_big.
__global__ __launch_bounds__(256) void k_pp_scatter(
    const int* __restrict__ rhoc, const int64_t* __restrict__ nzero, const int16_t* __restrict__ xp,
    const PPTile* __restrict__ tiles, const int* __restrict__ ires_loc,
    const float* __restrict__ apm3, int nt, int ncb, int nte, int nnt, int nns, int ntt,
    int ratio_cs, int64_t ncells, int64_t n_used, const int* __restrict__ cell_start,
    const int* __restrict__ err, float4* __restrict__ pos4, uint32_t* __restrict__ slot,
    uint32_t* __restrict__ cellid) {
    if (*err & 4) return;  // sorted arrays would overflow: nothing is written
    __shared__ int cur[kWaves][kMaxSub];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * kWaves + wave;
    const bool ok = w < ncells;
    CellRef c{};
    int rcp = 1, S = 1, nb = 0;
    int64_t off = 0;
    if (ok) {
        c = coarse_cell(w, nt, ncb, nte);
        PPTile T = tiles[c.t];
        rcp = T.rcp;
        S = T.S;
        off = T.cell_off;
        nb = rcp * rcp * rcp;
        for (int b = lane; b < nb; b += 64) cur[wave][b] = 0;
    }
    __syncthreads();
    if (!ok) return;  // no block-level barrier below
    const int np = rhoc[c.cell];
    const int64_t base = nzero[c.cell];
    if (np <= 0 || np > kBigCell || base < 0 || base + np > n_used) return;
    const float range = cell_range(c, ires_loc, apm3, nt, nnt, nns, ntt);
    scatter_run(c, rcp, S, off, xp, base, 0, np, range, ratio_cs, cur[wave], cell_start, pos4,
                slot, cellid, lane);
}

// The listed coarse cells, one block each: wavefront s takes the s-th
// contiguous segment of the cell's particles; per-segment histograms and
// their exclusive sums over the segments give each wavefront's starting
// cursors, so the order is the one-wavefront order exactly.
__global__ __launch_bounds__(64 * kBigWaves) void k_pp_scatter_big(
    const int* __restrict__ rhoc, const int64_t* __restrict__ nzero, const int16_t* __restrict__ xp,
    const PPTile* __restrict__ tiles, const int* __restrict__ ires_loc,
    const float* __restrict__ apm3, int nt, int ncb, int nte, int nnt, int nns, int ntt,
    int ratio_cs, const int* __restrict__ big_list, const int* __restrict__ nbig,
    const int* __restrict__ cell_start, const int* __restrict__ err, float4* __restrict__ pos4,
    uint32_t* __restrict__ slot, uint32_t* __restrict__ cellid) {
    if (*err & 4) return;
    __shared__ int hist[kBigWaves][kMaxSub];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nbg = *nbig;
    for (int i = blockIdx.x; i < nbg; i += gridDim.x) {
        const CellRef c = coarse_cell(big_list[i], nt, ncb, nte);
        const PPTile T = tiles[c.t];
        const int rcp = T.rcp, nb = rcp * rcp * rcp;
        const int np = rhoc[c.cell];
        const int64_t base = nzero[c.cell];
        const int seg = (np + 64 * kBigWaves - 1) / (64 * kBigWaves) * 64;
        const int l0 = min(np, wave * seg), l1 = min(np, l0 + seg);
        for (int b = lane; b < nb; b += 64) hist[wave][b] = 0;
        __builtin_amdgcn_wave_barrier();
        for (int l = l0 + lane; l < l1; l += 64) {
            float x0, x1, x2;
            bool out;
            atomicAdd(&hist[wave][pp_sub(c, rcp, xp + 3 * (base + l), x0, x1, x2, out)], 1);
        }
        __syncthreads();
        for (int b = threadIdx.x; b < nb; b += blockDim.x) {
            int run = 0;
            for (int s = 0; s < kBigWaves; s++) {
                const int t = hist[s][b];
                hist[s][b] = run;
                run += t;
            }
        }
        __syncthreads();
        const float range = cell_range(c, ires_loc, apm3, nt, nnt, nns, ntt);
        scatter_run(c, rcp, T.S, T.cell_off, xp, base, l0, l1, range, ratio_cs, hist[wave],
                    cell_start, pos4, slot, cellid, lane);
        __syncthreads();  // hist is rewritten for the next cell
    }
}

// Physical particles per subtile location (pp_force_kernel.cu:691-705).
__global__ void k_pp_ptotal(const int* __restrict__ rhoc, int nt, int ncb, int nte, int nnt,
                            int nns, int ntt, unsigned long long* __restrict__ ploc) {
    const int64_t per = (int64_t)nt * nt * nt;
    int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= per * nnt * nnt * nnt) return;
    const int t = (int)(w / per);
    int64_t r = w - (int64_t)t * per;
    const int i = (int)(r % nt), j = (int)((r / nt) % nt), k = (int)(r / ((int64_t)nt * nt));
    const int v = rhoc[(((int64_t)t * nte + k + ncb) * nte + j + ncb) * nte + i + ncb];
    if (!v) return;
    const int tx = t % nnt, ty = (t / nnt) % nnt, tz = t / (nnt * nnt);
    const int loc = (((tz * nnt + ty) * nnt + tx) * nns + k / ntt) * nns * nns + (j / ntt) * nns + i / ntt;
    atomicAdd(&ploc[loc], (unsigned long long)v);
}

// Work chunks of the pair kernel: <= 64 consecutive physical particles of
// one PP cell, <= 2^dense_lg in cells of >= 64 particles.  flags[a] = 1
// where a chunk starts.
// The sorted count n comes from the scan on the device; flags[a] = 0 for
// n <= a < cap, and for every a if the call has failed.
// Reads the scatter's output (rpos, rslot) and copies every particle not in
// an active cell of >= 64 particles to pos4 / slot (k_pp_subsort places the
// others).  Pairs-once pass (small > 0): the particles of active cells of at
// most `small` particles, whose chunks evaluate their pairs one-sidedly, are
// marked by a negative range (w = -range); every reader of w then takes |w|
// as the range and w != 0 as active.
__global__ void k_chunk_flags(const float4* __restrict__ rpos, const uint32_t* __restrict__ rslot,
                              const uint32_t* __restrict__ cellid,
                              const int* __restrict__ cell_start, const int* __restrict__ n_sorted,
                              const int* __restrict__ err, int cap, int dense_lg, int small,
                              int* __restrict__ flags, int* __restrict__ dense_list,
                              int* __restrict__ ndense, float4* __restrict__ pos4,
                              uint32_t* __restrict__ slot) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= cap) return;
    const int n = (*err & 4) ? 0 : *n_sorted;
    int f = 0;
    if (a < n) {
        float4 p = rpos[a];
        bool dense = false;
        if (p.w > 0.0f) {
            const uint32_t c = cellid[a];
            const int c0 = cell_start[c], cn = cell_start[c + 1] - c0;
            dense = cn >= 64;
            const int mask = dense ? (1 << dense_lg) - 1 : 63;
            f = ((a - c0) & mask) == 0;
            if (dense && a == c0) dense_list[atomicAdd(ndense, 1)] = (int)c;
            if (cn <= small) p.w = -p.w;
        }
        if (!dense) {
            pos4[a] = p;
            slot[a] = rslot[a];
        }
    }
    flags[a] = f;
}

__global__ void k_chunk_emit(const int* __restrict__ scanned, int n, int* __restrict__ chunks) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    if (scanned[a + 1] != scanned[a]) chunks[scanned[a]] = a;
}

// Chunk order of the pair kernel.  Its wavefronts take chunks by a static
// stride over the list; in cell order the 6144 chunks in flight at a time
// are neighbours, whose j the L2 caches share, but a chunk in a clump's core
// can cost as much as a wavefront's average total, and where such chunks
// fall late a wavefront ends late (at 256^3 REF wavefronts ended between 32
// and 80 ms).  On request (a tunable; off by default) the heavy chunks, whose
// estimated candidate pairs are at least 2^heavy_lg, go first, heaviest
// first, and the rest follow in cell order.  The estimate is (i slots + 1) x
// (particles of the chunk's PP cell) x (2R+1)^3, as 8 steps per octave; the
// sort is a stable counting sort over 256 bins (k_chunk_cost, a scan,
// k_chunk_order), all light chunks in the last bin, 0.3 ms at 256^3.
// Measured at 256^3 REF (pair kernel alone): cell order 80.9 ms; heavy
// first from 2^14, 2^17, 2^20, 2^22: 86.6, 86.3, 86.3, 88.9 ms; from 2^24
// (few chunks) 80.5 ms; the whole list by cost (the in-flight chunks
// scattered over the volume) 87.2 ms.  The wavefronts that end early leave
// their SIMD's issue slots to the others, so the end-time spread costs less
// than the heavy chunks lose by running together.
constexpr int kOrderBins = 256;
constexpr int kOrderBlocks = 1024;

__device__ inline void order_segment(int n, int& beg, int& end) {
    const int seg = (n + (int)gridDim.x - 1) / (int)gridDim.x;
    beg = min(n, (int)blockIdx.x * seg);
    end = min(n, beg + seg);
}

__global__ __launch_bounds__(kOrderBins) void k_chunk_cost(PPArgs A, const int* __restrict__ chunks,
                                                        const int* __restrict__ nchunks_dev,
                                                        int dense_lg, int heavy_lg,
                                                        unsigned char* __restrict__ bin,
                                                        int* __restrict__ hist) {
    __shared__ int h[kOrderBins];
    int beg, end;
    order_segment(*nchunks_dev, beg, end);
    h[threadIdx.x] = 0;
    __syncthreads();
    for (int ck = beg + (int)threadIdx.x; ck < end; ck += kOrderBins) {
        const int a0 = chunks[ck];
        const uint32_t cid = A.cellid[a0];
        const int cbeg = A.cell_start[cid], cend = A.cell_start[cid + 1];
        const int ni = min(cend - cbeg >= 64 ? 1 << dense_lg : 64, cend - a0);
        int I = 1;
        while (I < ni) I <<= 1;
        int lo = 0, hi = A.ntile - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (A.tiles[mid].cell_off <= (long long)cid) lo = mid;
            else hi = mid - 1;
        }
        const int rcp = A.tiles[lo].rcp;
        const int W = 2 * (int)ceilf(fabsf(A.pos4[a0].w) * (float)rcp / (float)A.ratio_cs) + 1;
        const float cost = (float)(I + 1) * (float)(cend - cbeg) * (float)(W * W * W);
        const int k = min(kOrderBins - 1, (int)(8.0f * log2f(cost + 1.0f)));
        const int b = k >= 8 * heavy_lg ? kOrderBins - 1 - k : kOrderBins - 1;
        bin[ck] = (unsigned char)b;
        // the light chunks share the last bin: one LDS atomic per wavefront
        const uint64_t lm = __ballot(b == kOrderBins - 1);
        if (b != kOrderBins - 1) atomicAdd(&h[b], 1);
        else if ((int)(threadIdx.x & 63) == __ffsll((unsigned long long)lm) - 1)
            atomicAdd(&h[b], __popcll(lm));
    }
    __syncthreads();
    hist[threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];  // bin-major for the scan
}

// Stable scatter of the chunks by bin: within a block's segment, 256 chunks
// at a time in order; a chunk's place is its bin's base, the chunks of its bin
// in the earlier wavefronts of the step, and its rank among its wavefront's
// lanes of the same bin (found with 8 ballots on the bin's bits).
__global__ __launch_bounds__(kOrderBins) void k_chunk_order(const int* __restrict__ chunks,
                                                         const unsigned char* __restrict__ bin,
                                                         const int* __restrict__ nchunks_dev,
                                                         const int* __restrict__ hist,
                                                         int* __restrict__ out) {
    __shared__ int base[kOrderBins];
    __shared__ int wc[4][kOrderBins];
    int beg, end;
    order_segment(*nchunks_dev, beg, end);
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    base[t] = hist[t * gridDim.x + blockIdx.x];
    wc[0][t] = wc[1][t] = wc[2][t] = wc[3][t] = 0;
    __syncthreads();
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int t0 = beg; t0 < end; t0 += kOrderBins) {  // block-uniform trip count
        const int ck = t0 + t;
        const bool valid = ck < end;
        const int b = valid ? bin[ck] : 0;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const uint64_t m = __ballot(valid && ((b >> q) & 1));
            peers &= ((b >> q) & 1) ? m : ~m;
        }
        const int rank = __popcll(peers & lt);
        if (valid && rank == 0) wc[wave][b] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int pos = base[b] + rank;
            for (int w = 0; w < wave; w++) pos += wc[w][b];
            out[pos] = chunks[ck];
        }
        __syncthreads();
        base[t] += wc[0][t] + wc[1][t] + wc[2][t] + wc[3][t];
        wc[0][t] = wc[1][t] = wc[2][t] = wc[3][t] = 0;
        __syncthreads();
    }
}

// Spreads the low 3 bits of v to bits 0, 3, 6 (Morton interleave).
__device__ inline int spread3(int v) { return (v & 1) | ((v & 2) << 2) | ((v & 4) << 4); }

// One wavefront per PP cell of at least 64 particles (listed by
// k_chunk_flags): the cell's particles are copied from the scatter's output
// (rpos, rslot) to pos4 / slot ordered by an 8x8x8 Morton sub-cell key
// (counting sort in LDS; within a sub-cell, the cell order), so that
// consecutive chunks of a dense cell are spatially compact and a chunk is a
// range of sorted indices.  Cells keep their ranges.
__global__ __launch_bounds__(256) void k_pp_subsort(
    const float4* __restrict__ rpos, const uint32_t* __restrict__ rslot,
    const int* __restrict__ cell_start, const PPTile* __restrict__ tiles, int ntile,
    const int* __restrict__ dense_list, const int* __restrict__ ndense_dev, int ratio_cs,
    float4* __restrict__ pos4, uint32_t* __restrict__ slot) {
    __shared__ int bins[kWaves][512];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ndense = *ndense_dev;
    // grid-stride over the listed cells; no block-level barrier in this kernel
    for (int qc = blockIdx.x * kWaves + wave; qc < ndense; qc += gridDim.x * kWaves) {
    const uint32_t cid = (uint32_t)dense_list[qc];
    const int c0 = cell_start[cid], c1 = cell_start[cid + 1];
    int lo = 0, hi = ntile - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tiles[mid].cell_off <= (long long)cid) lo = mid;
        else hi = mid - 1;
    }
    const PPTile T = tiles[lo];
    const long long key = (long long)cid - T.cell_off;
    const int S = T.S;
    const int kx = (int)(key % S), ky = (int)((key / S) % S), kz = (int)(key / ((long long)S * S));
    const float size = (float)ratio_cs / (float)T.rcp;   // cell edge in pos4 units
    const float sx = (float)(kx - T.rcp) * size, sy = (float)(ky - T.rcp) * size,
                sz = (float)(kz - T.rcp) * size;
    const float inv = 8.0f / size;
    int* b = bins[wave];
    for (int q = lane; q < 512; q += 64) b[q] = 0;
    __builtin_amdgcn_wave_barrier();
    auto subkey = [&](int a) {
        const float4 v = rpos[a];
        const int ix = min(max((int)((v.x - sx) * inv), 0), 7);
        const int iy = min(max((int)((v.y - sy) * inv), 0), 7);
        const int iz = min(max((int)((v.z - sz) * inv), 0), 7);
        return spread3(ix) | (spread3(iy) << 1) | (spread3(iz) << 2);
    };
    for (int a = c0 + lane; a < c1; a += 64) atomicAdd(&b[subkey(a)], 1);
    __builtin_amdgcn_wave_barrier();
    // exclusive scan of the 512 bins: 8 consecutive bins per lane
    int loc[8], s = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        loc[q] = b[lane * 8 + q];
        s += loc[q];
    }
    int inc = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    int run = inc - s;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        b[lane * 8 + q] = run;
        run += loc[q];
    }
    __builtin_amdgcn_wave_barrier();
    for (int a = c0 + lane; a < c1; a += 64) {
        const int pos = atomicAdd(&b[subkey(a)], 1);
        pos4[c0 + pos] = rpos[a];
        slot[c0 + pos] = rslot[a];
    }
    __builtin_amdgcn_wave_barrier();  // bins are cleared for the next cell
    }
}

// First sorted index of every tile (for per-tile counts).
__global__ void k_tile_first(const int* __restrict__ cell_start, const PPTile* __restrict__ tiles,
                             int ntile, int64_t c_tot, int* __restrict__ out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntile) return;
    out[t] = cell_start[t < ntile ? tiles[t].cell_off : c_tot];
}

// ------------------------------------------------------------ host side

// The PP reduction block: PPRed, then physical particles per subtile
// location [ns3], then {P_cand, P_acc} per parent tile [ntile][2].
static size_t red_bytes(const Geom& g, int ntile) {
    return sizeof(PPRed) + ((size_t)g.ns3 + 2 * (size_t)ntile) * sizeof(unsigned long long);
}

static int location_maps(const Geom& g, const int* ires, const int* ixyz3, std::vector<int>& iloc,
                         std::vector<int>& owner) {
    iloc.assign(g.ns3, 0);
    owner.assign(g.ns3, -1);
    const int nnt = g.p.nnt, nns = g.p.nns;
    for (int it = 0; it < g.ns3; it++) {
        const int* x = ixyz3 + 6 * it;
        int sx = x[0] - 1, sy = x[1] - 1, sz = x[2] - 1, tx = x[3] - 1, ty = x[4] - 1, tz = x[5] - 1;
        if (sx < 0 || sx >= nns || sy < 0 || sy >= nns || sz < 0 || sz >= nns || tx < 0 ||
            tx >= nnt || ty < 0 || ty >= nnt || tz < 0 || tz >= nnt || ires[it] < 2 ||
            ires[it] > 6) {
            set_error("invalid PP subtile metadata at %d (ires=%d)", it, ires[it]);
            return CUBE2_ERR_ARG;
        }
        int loc = (((tz * nnt + ty) * nnt + tx) * nns + sz) * nns * nns + sy * nns + sx;
        if (owner[loc] >= 0) {
            set_error("subtiles %d and %d have the same location in ixyz3", owner[loc], it);
            return CUBE2_ERR_ARG;
        }
        iloc[loc] = ires[it];
        owner[loc] = it;
    }
    // ns3 distinct in-range locations out of ns3: every location is filled
    return CUBE2_OK;
}

// Cell base slots of the whole image (shared with PM3).
int nzero_reserve() {
    const Geom& g = ctx().g;
    const int64_t rows = (int64_t)g.ntile * g.nte * g.nte;
    return g_pp.nzero.reserve(std::max<int64_t>(rows * g.nte, 1) * sizeof(int64_t));
}

int compute_nzero(const int* d_rhoc, const int64_t* d_idx, const int64_t** out, hipStream_t s) {
    const Geom& g = ctx().g;
    const int64_t rows = (int64_t)g.ntile * g.nte * g.nte;
    C2_TRY(nzero_reserve());
    if (ctx().nzero_shared) {  // computed once for both passes of an overlapped kick
        *out = g_pp.nzero.as<int64_t>();
        return CUBE2_OK;
    }
    if (rows > 0) {
        Stage st("nzero", (double)rows * g.nte * 12 + rows * 8);
        k_nzero<<<ceil_div(rows, 256), 256, 0, s>>>(d_rhoc, d_idx, g_pp.nzero.as<int64_t>(), g.nte, rows);
        C2_HIP(hipGetLastError());
    }
    *out = g_pp.nzero.as<int64_t>();
    return CUBE2_OK;
}

int pp_validate(const int* ires, const int* ixyz3, const int* ratio_sf) {
    ensure_params();
    if (!ires || !ixyz3 || !ratio_sf) {
        set_error("null argument");
        return CUBE2_ERR_ARG;
    }
    std::vector<int> iloc, owner;
    C2_TRY(location_maps(ctx().g, ires, ixyz3, iloc, owner));
    for (int L = 2; L <= 6; L++)  // the levels ires can select (kick.f90:115-127)
        if (ratio_sf[L - 1] <= 0 || ratio_sf[L - 1] > 12) {
            set_error("ratio_sf[%d]=%d outside 1..12", L - 1, ratio_sf[L - 1]);
            return CUBE2_ERR_ARG;
        }
    return CUBE2_OK;
}

// Sorted-array total beyond capacity (idx_b_r rows overlapping): err |= 4,
// which every later kernel of the call checks before touching the arrays.
__global__ void k_pp_check_total(const int* __restrict__ total, int64_t cap, PPRed* red) {
    if (threadIdx.x == 0 && (int64_t)*total > cap) red->err |= 4;
}

int pp_prepare(const int* ires, const int* ixyz3, const float* apm3, const int* ratio_sf,
               int64_t n_used, PPPlan& P, bool split, bool learned) {
    ensure_params();
    Context& C = ctx();
    const Geom& g = C.g;
    if (!ires || !ixyz3 || !apm3 || !ratio_sf) {
        set_error("null argument");
        return CUBE2_ERR_ARG;
    }
    if (n_used < 0 || n_used >= ((int64_t)1 << 31)) {
        set_error("invalid PP particle count %lld", (long long)n_used);
        return CUBE2_ERR_CAPACITY;
    }
    C2_TRY(location_maps(g, ires, ixyz3, P.iloc, P.owner));
    for (int L = 2; L <= 6; L++)  // the levels ires can select (kick.f90:115-127)
        if (ratio_sf[L - 1] <= 0 || ratio_sf[L - 1] > 12) {
            set_error("ratio_sf[%d]=%d outside 1..12", L - 1, ratio_sf[L - 1]);
            return CUBE2_ERR_ARG;
        }
    C2_TRY(ensure_init());
    hipStream_t s = C.stream;
    // per-tile grid (pp_force_kernel.cu:641-661)
    P.ntile = g.ntile;
    P.nsub3 = g.p.nns * g.p.nns * g.p.nns;
    P.n_used = n_used;
    P.mode = C.pp_mode;
    if (split && P.mode == CUBE2_PP_EXACT) {
        set_error("the exact-order PP mode has no split pass");
        return CUBE2_ERR_ARG;
    }
    P.split = split;
    if (learned && (!split || !learned_model_set())) {
        set_error(split ? "learned kick without a model (cube2_learned_set)"
                        : "the learned stand-in needs the split pair pass");
        return split ? CUBE2_ERR_STATE : CUBE2_ERR_ARG;
    }
    P.learned = learned;
    std::copy(ratio_sf, ratio_sf + 7, P.ratio_sf.begin());
    std::copy(apm3, apm3 + 7, P.apm3.begin());
    std::vector<PPTile>& T = g_pp.h_tiles;
    T.assign(P.ntile, PPTile{});
    int64_t off = 0;
    for (int t = 0; t < P.ntile; t++) {
        int rcp = ratio_sf[1];
        for (int q = 0; q < P.nsub3; q++) rcp = std::max(rcp, ratio_sf[P.iloc[t * P.nsub3 + q] - 1]);
        T[t].rcp = rcp;
        T[t].S = g.nt * rcp + 2 * rcp;
        T[t].cell_off = off;
        off += (int64_t)T[t].S * T[t].S * T[t].S;
    }
    P.c_tot = off;
    if (P.c_tot >= ((int64_t)1 << 31) - 1) {
        set_error("PP cell grid too large (%lld cells)", (long long)P.c_tot);
        return CUBE2_ERR_CAPACITY;
    }
    P.ncells_ext = (int64_t)P.ntile * (g.nt + 2) * (g.nt + 2) * (g.nt + 2);
    P.cap = (size_t)std::max<int64_t>(n_used, 1);
    C2_TRY(g_pp.tiles.reserve(P.ntile * sizeof(PPTile)));
    C2_TRY(g_pp.ires_loc.reserve(g.ns3 * sizeof(int)));
    C2_TRY(g_pp.apm3.reserve(8 * sizeof(float)));
    C2_TRY(g_pp.counts.reserve((P.c_tot + 1) * sizeof(int)));
    C2_TRY(g_pp.pos4.reserve(P.cap * sizeof(float4)));
    C2_TRY(g_pp.slot.reserve(P.cap * sizeof(uint32_t)));
    C2_TRY(g_pp.cellid.reserve(P.cap * sizeof(uint32_t)));
    C2_TRY(g_pp.red.reserve(red_bytes(g, P.ntile)));
    C2_TRY(g_pp.tile_first.reserve((P.ntile + 1) * sizeof(int)));
    C2_TRY(g_pp.chunk_flags.reserve((P.cap + 1) * sizeof(int)));
    C2_TRY(g_pp.chunks.reserve((P.cap + 1) * sizeof(int)));
    if (pp_heavy_lg()) {
        C2_TRY(g_pp.order.reserve((P.cap + 1) * sizeof(int)));
        C2_TRY(g_pp.order_bin.reserve(P.cap));
        C2_TRY(g_pp.order_hist.reserve(((size_t)kOrderBins * kOrderBlocks + 1) * sizeof(int)));
    }
    if (P.mode != CUBE2_PP_EXACT) {
        C2_TRY(g_pp.raw_pos4.reserve(P.cap * sizeof(float4)));
        C2_TRY(g_pp.raw_slot.reserve(P.cap * sizeof(uint32_t)));
    }
    C2_TRY(g_pp.big.reserve(std::max<int64_t>(P.ncells_ext, 1) * sizeof(int)));
    C2_TRY(g_pp.dense.reserve((P.cap / 64 + 1) * sizeof(int)));
    // Pairs once: each pair force component is below 10 / app^2 in magnitude
    // (F_ra(r, a) <= 342 / (35 a^2), and 1/r^2 < 1/app^2 beyond app), and a
    // particle has fewer than cap partners, so sums of (value * 2^k) stay
    // below 2^62 for 2^k <= 2^62 / (10 / app^2 * cap); at 256^3 (app 0.06)
    // k = 24, a resolution of 6e-8 against forces of order 1 to 1e3.
    P.newton = false;
    if (pp_newton_on() && pp_newton_mode(P.mode) && !learned && g.p.app > 0.0) {
        const double bound = 10.0 / (g.p.app * g.p.app) * (double)P.cap;
        const int k = std::min(40, (int)std::floor(62.0 - std::log2(bound)));
        if (k >= 12) {
            P.newton = true;
            P.react_scale = std::ldexp(1.0f, k);
        }
    }
    if (split || P.newton) {
        C2_TRY(g_pp.term.reserve(P.cap * 3 * sizeof(float)));
        const int64_t rows = (int64_t)g.ntile * g.nt * g.nt;
        C2_TRY(g_pp.rows_len.reserve((rows + 1) * sizeof(int)));
        C2_TRY(g_pp.rows_base.reserve(std::max<int64_t>(rows, 1) * sizeof(int64_t)));
    }
    if (P.newton) C2_TRY(g_pp.react.reserve(P.cap * 3 * sizeof(unsigned long long)));
    P.all_active = true;
    for (int lv : P.iloc) P.all_active = P.all_active && lv >= 1 && lv <= 7 && apm3[lv - 1] > 0.0f;
    if (learned) {
        C2_TRY(learned_prepare(g, P.iloc, ratio_sf, apm3, P.lp, learned_model_occupancy()));
        C2_TRY(learned_reserve(P.lp));
    }
    C2_TRY(g_pp.scan_tmp.reserve(scan_tmp_ints(std::max<int64_t>(P.c_tot, (int64_t)P.cap)) * sizeof(int)));
    C2_TRY(nzero_reserve());
    C2_TRY(g_pp.h_res.reserve(red_bytes(g, P.ntile) + (P.ntile + 1) * sizeof(int)));
    P.count_tiles = g_pp.count_tiles;
    C2_HIP(hipMemcpyAsync(g_pp.tiles.ptr, T.data(), P.ntile * sizeof(PPTile), hipMemcpyHostToDevice, s));
    C2_HIP(hipMemcpyAsync(g_pp.ires_loc.ptr, P.iloc.data(), g.ns3 * sizeof(int), hipMemcpyHostToDevice, s));
    C2_HIP(hipMemcpyAsync(g_pp.apm3.ptr, apm3, 7 * sizeof(float), hipMemcpyHostToDevice, s));
    C2_HIP(hipStreamSynchronize(s));  // host tables may go out of scope
    buffers_changed();                 // the device tables now describe this call
    P.grid_pairs = pp_tiled_grid(P.mode, split, P.newton);
    g_pp.valid = false;
    return CUBE2_OK;
}

static int pp_results_to_host(const PPPlan& P, hipStream_t s);
static int pp_apply_kernels(const PPPlan& P, float* d_vp, const double* dv64, hipStream_t s);

// Physical rows (t, k, j) of the image: each is one contiguous run of slots,
// from the first slot of its first physical coarse cell to that of the
// first halo cell after its last (nzero there).  len = its slots, base =
// its first slot; a run outside [0, n_used) counts as empty (the cell-list
// kernels flag such a layout as an error).
__global__ void k_phys_rows(const int64_t* __restrict__ nzero, int nt, int ncb, int nte,
                            int64_t nrows, int64_t n_used, int* __restrict__ len,
                            int64_t* __restrict__ base) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int64_t per = (int64_t)nt * nt;
    const int64_t t = r / per, rem = r - t * per;
    const int j = ncb + (int)(rem % nt), k = ncb + (int)(rem / nt);
    const int64_t row0 = ((t * nte + k) * nte + j) * nte;
    const int64_t b = nzero[row0 + ncb], e = nzero[row0 + ncb + nt];
    const bool ok = b >= 0 && e >= b && e <= n_used;
    base[r] = ok ? b : 0;
    len[r] = ok ? (int)(e - b) : 0;
}

// Moves the (x, y, z) rows of slot runs between a full per-slot array and a
// compact one: run r is slots [base[r], base[r] + coff[r+1] - coff[r]) of the
// full array and rows [coff[r], coff[r+1]) of the compact one (runs =
// base[nruns] then coff[nruns + 1]).  One block per run at a time.
__global__ __launch_bounds__(256) void k_runs_move(float* __restrict__ dst,
                                                   const float* __restrict__ src,
                                                   const int64_t* __restrict__ runs, int nruns,
                                                   int to_compact) {
    const int64_t* coff = runs + nruns;
    for (int r = blockIdx.x; r < nruns; r += gridDim.x) {
        const int64_t full = 3 * runs[r], comp = 3 * coff[r], n = 3 * (coff[r + 1] - coff[r]);
        const int64_t so = to_compact ? full : comp, d0 = to_compact ? comp : full;
        for (int64_t q = threadIdx.x; q < n; q += blockDim.x) dst[d0 + q] = src[so + q];
    }
}

int vp_runs_move(float* dst, const float* src, const int64_t* runs, int nruns, bool to_compact,
                 hipStream_t s) {
    if (nruns <= 0) return CUBE2_OK;
    k_runs_move<<<std::min(nruns, 16384), 256, 0, s>>>(dst, src, runs, nruns, to_compact ? 1 : 0);
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

// Pairs-once pass: clears the fixed-point sums of the sorted particles.
__global__ void k_react_zero(unsigned long long* __restrict__ react, const int* __restrict__ n_sorted,
                             int cap, const int* __restrict__ err) {
    const int64_t n = (*err & 4) ? 0 : 3 * (int64_t)min(*n_sorted, cap);
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x)
        react[q] = 0;
}

// Pairs-once pass, after the pair kernel: each active sorted particle's
// acceleration sum (fixed point / scale, rounded once to FP32) becomes its
// velocity term a * m^2 * a_mid * dt / 6 at term[3 * slot], in the unsplit
// kernel's FP32 operation order, and f2max = max |a m^2|^2.
__global__ __launch_bounds__(256) void k_pp_newton_terms(
    const float4* __restrict__ pos4, const uint32_t* __restrict__ slot,
    const unsigned long long* __restrict__ react, const int* __restrict__ n_sorted, int cap,
    const float* __restrict__ scal, double inv_scale, float* __restrict__ term, PPRed* red) {
#pragma clang fp contract(off)
    float f2 = 0.0f;
    const int n = red->err == 0 ? min(*n_sorted, cap) : 0;
    const float mass = scal[0], a_mid = scal[1], dt = scal[2];
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < n; a += gridDim.x * blockDim.x) {
        if (pos4[a].w == 0.0f) continue;  // halo sources
        float acc[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float v = (float)((double)(long long)react[3 * (size_t)a + c] * inv_scale);
            acc[c] = v * mass * mass;
        }
        f2 = fmaxf(f2, acc[0] * acc[0] + acc[1] * acc[1] + acc[2] * acc[2]);
        const size_t ip = (size_t)slot[a] * 3;
#pragma unroll
        for (int c = 0; c < 3; c++) term[ip + c] = acc[c] * a_mid * dt / 6.0f;
    }
    const float mx[1] = {f2};
    block_max_to<1>(&red->f2max_bits, mx);
}

// Velocity update of a split pass in slot order, when every particle of the
// physical rows is active (every subtile has a PP range): the physical slots
// as one list over the scanned rows, kApplySeg at a time per block, each
// thread walking the rows forward from the segment's first; vp = RN_f(vp +
// dv64 + term / pi) in FP64 (the unsplit kernel's operation order), vmax
// over them.  The slots, Delta-v and terms are read in order.
constexpr int kApplySeg = 2048;

__global__ __launch_bounds__(256) void k_pp_apply_rows(const int* __restrict__ prefix,
                                                      const int64_t* __restrict__ base, int nrows,
                                                      const float* __restrict__ term,
                                                      const double* __restrict__ dv64,
                                                      float* __restrict__ vp, PPRed* red) {
    __shared__ int srow;
    float vm[3] = {0.f, 0.f, 0.f};
    const int total = red->err == 0 ? prefix[nrows] : 0;  // a failed call leaves no terms
    for (int seg0 = blockIdx.x * kApplySeg; seg0 < total; seg0 += gridDim.x * kApplySeg) {
        if (threadIdx.x == 0) {  // last row starting at or before seg0
            int lo = 0, hi = nrows - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (prefix[mid] <= seg0) lo = mid;
                else hi = mid - 1;
            }
            srow = lo;
        }
        __syncthreads();
        int row = srow;
        for (int e = seg0 + threadIdx.x; e < min(total, seg0 + kApplySeg); e += blockDim.x) {
            while (prefix[row + 1] <= e) row++;  // e < total = prefix[nrows]
            const size_t ip = (size_t)(base[row] + (e - prefix[row])) * 3;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double t = (double)term[ip + c] / 3.14159265358979323846;
                double v = (double)vp[ip + c];
                if (dv64) v += dv64[ip + c];
                const float nv = (float)(v + t);
                vp[ip + c] = nv;
                vm[c] = fmaxf(vm[c], fabsf(nv));
            }
        }
        __syncthreads();  // srow is rewritten for the next segment
    }
    block_max_to<3>(red->vmax_bits, vm);
}

int pp_enqueue(const PPPlan& P, const int* d_rhoc, const int64_t* d_idx, const int16_t* d_xp,
               float* d_vp, const float* d_scal, const double* dv64, hipStream_t s) {
    const Geom& g = ctx().g;
    PPRed* red = g_pp.red.as<PPRed>();
    unsigned long long* ploc = reinterpret_cast<unsigned long long*>(red + 1);
    int* counts = g_pp.counts.as<int>();
    const int* n_sorted = counts + P.c_tot;
    C2_TRY(fill_zero(red, red_bytes(g, P.ntile), s));
    const int64_t* nzp = nullptr;
    C2_TRY(compute_nzero(d_rhoc, d_idx, &nzp, s));
    const int64_t prow = (int64_t)g.ntile * g.nt * g.nt;
    if ((P.split || P.newton) && P.all_active && prow > 0) {
        Stage st("pp_rows", (double)prow * 28);
        k_phys_rows<<<ceil_div(prow, 256), 256, 0, s>>>(nzp, g.nt, g.ncb, g.nte, prow, P.n_used,
                                                        g_pp.rows_len.as<int>(),
                                                        g_pp.rows_base.as<int64_t>());
        C2_HIP(hipGetLastError());
        C2_TRY(exclusive_scan(g_pp.rows_len.as<int>(), prow, g_pp.scan_tmp, s));
    }
    if (P.ncells_ext > 0) {
        Stage st("pp_count", (double)P.n_used * 6 + P.c_tot * 4.0);
        k_pp_count<<<ceil_div(P.ncells_ext, kWaves), 256, 0, s>>>(
            d_rhoc, nzp, d_xp, g_pp.tiles.as<PPTile>(), g.nt, g.ncb, g.nte, P.ncells_ext,
            P.n_used, counts, &red->err, g_pp.big.as<int>(), &red->nbig);
        C2_HIP(hipGetLastError());
        k_pp_count_big<<<kBigBlocks, 1024, 0, s>>>(d_rhoc, nzp, d_xp, g_pp.tiles.as<PPTile>(), g.nt,
                                                   g.ncb, g.nte, g_pp.big.as<int>(), &red->nbig,
                                                   counts, &red->err);
        C2_HIP(hipGetLastError());
    }
    {
        Stage st("pp_scan", (double)P.c_tot * 8);
        C2_TRY(exclusive_scan(counts, P.c_tot, g_pp.scan_tmp, s));
        k_pp_check_total<<<1, 64, 0, s>>>(n_sorted, (int64_t)P.cap, red);
        C2_HIP(hipGetLastError());
    }
    // the scatter writes pos4 / slot directly in the EXACT mode; otherwise
    // k_chunk_flags and k_pp_subsort copy its output there
    const bool reorder = P.mode != CUBE2_PP_EXACT;
    float4* spos = reorder ? g_pp.raw_pos4.as<float4>() : g_pp.pos4.as<float4>();
    uint32_t* sslot = reorder ? g_pp.raw_slot.as<uint32_t>() : g_pp.slot.as<uint32_t>();
    if (P.ncells_ext > 0) {
        Stage st("pp_scatter", (double)P.n_used * (6 + 24));
        k_pp_scatter<<<ceil_div(P.ncells_ext, kWaves), 256, 0, s>>>(
            d_rhoc, nzp, d_xp, g_pp.tiles.as<PPTile>(), g_pp.ires_loc.as<int>(),
            g_pp.apm3.as<float>(), g.nt, g.ncb, g.nte, g.p.nnt, g.p.nns, g.ntt, g.p.ratio_cs,
            P.ncells_ext, P.n_used, counts, &red->err, spos, sslot, g_pp.cellid.as<uint32_t>());
        C2_HIP(hipGetLastError());
        k_pp_scatter_big<<<kBigBlocks, 64 * kBigWaves, 0, s>>>(
            d_rhoc, nzp, d_xp, g_pp.tiles.as<PPTile>(), g_pp.ires_loc.as<int>(),
            g_pp.apm3.as<float>(), g.nt, g.ncb, g.nte, g.p.nnt, g.p.nns, g.ntt, g.p.ratio_cs,
            g_pp.big.as<int>(), &red->nbig, counts, &red->err, spos, sslot,
            g_pp.cellid.as<uint32_t>());
        C2_HIP(hipGetLastError());
    }
    {
        const int64_t nphys = (int64_t)P.ntile * g.nt * g.nt * g.nt;
        if (nphys > 0)
            k_pp_ptotal<<<ceil_div(nphys, 256), 256, 0, s>>>(d_rhoc, g.nt, g.ncb, g.nte, g.p.nnt,
                                                              g.p.nns, g.ntt, ploc);
        C2_HIP(hipGetLastError());
        k_tile_first<<<ceil_div(P.ntile + 1, 64), 64, 0, s>>>(counts, g_pp.tiles.as<PPTile>(),
                                                              P.ntile, P.c_tot,
                                                              g_pp.tile_first.as<int>());
        C2_HIP(hipGetLastError());
    }
    PPArgs A{};
    A.pos4 = g_pp.pos4.as<float4>();
    A.slot = g_pp.slot.as<uint32_t>();
    A.cellid = g_pp.cellid.as<uint32_t>();
    A.cell_start = counts;
    A.tiles = g_pp.tiles.as<PPTile>();
    A.ntile = P.ntile;
    A.n_sorted = n_sorted;
    A.cap = (int)P.cap;
    A.ratio_cs = g.p.ratio_cs;
    A.app = (float)g.p.app;
    A.scal = d_scal;
    A.dv64 = P.split ? nullptr : dv64;
    A.vp = P.split ? nullptr : d_vp;
    A.term = P.split ? g_pp.term.as<float>() : nullptr;
    A.red = red;
    A.tile_pairs = P.count_tiles ? ploc + g.ns3 : nullptr;
    A.learned_occ = P.learned ? P.lp.args.occupancy : 0;
    A.learned_range = P.learned ? P.lp.args.range : 0.0f;
    A.ntt = g.ntt;
    if (P.newton) {
        A.term = g_pp.term.as<float>();
        A.react = g_pp.react.as<unsigned long long>();
        A.react_scale = P.react_scale;
    }
    if (P.mode == CUBE2_PP_EXACT) {
        C2_TRY(launch_pp_pairs(A, P.mode, s));
    } else {
        int* flags = g_pp.chunk_flags.as<int>();
        {
            Stage st("pp_chunks", (double)P.n_used * 24);
            k_chunk_flags<<<ceil_div(P.cap, 256), 256, 0, s>>>(
                g_pp.raw_pos4.as<float4>(), g_pp.raw_slot.as<uint32_t>(), A.cellid, A.cell_start,
                n_sorted, &red->err, (int)P.cap, pp_dense_lg(P.newton),
                P.newton ? (1 << pp_direct_lg()) / 2 : 0, flags, g_pp.dense.as<int>(),
                &red->ndense, g_pp.pos4.as<float4>(), g_pp.slot.as<uint32_t>());
            C2_HIP(hipGetLastError());
            C2_TRY(exclusive_scan(flags, (int64_t)P.cap, g_pp.scan_tmp, s));
            k_chunk_emit<<<ceil_div(P.cap, 256), 256, 0, s>>>(flags, (int)P.cap, g_pp.chunks.as<int>());
            C2_HIP(hipGetLastError());
        }
        {
            Stage st("pp_subsort", (double)P.n_used * 24);
            // one wavefront per chunk (only first chunks of dense cells work)
            k_pp_subsort<<<(unsigned)std::max(1, P.grid_pairs), 256, 0, s>>>(
                g_pp.raw_pos4.as<float4>(), g_pp.raw_slot.as<uint32_t>(), A.cell_start, A.tiles,
                P.ntile, g_pp.dense.as<int>(), &red->ndense, g.p.ratio_cs, g_pp.pos4.as<float4>(),
                g_pp.slot.as<uint32_t>());
            C2_HIP(hipGetLastError());
        }
        const int* chunks = g_pp.chunks.as<int>();
        if (pp_heavy_lg()) {
            Stage st("pp_order", (double)P.cap * 9);
            int* hist = g_pp.order_hist.as<int>();
            k_chunk_cost<<<kOrderBlocks, kOrderBins, 0, s>>>(A, chunks, flags + P.cap,
                                                             pp_dense_lg(P.newton),
                                                             pp_heavy_lg(),
                                                             g_pp.order_bin.as<unsigned char>(), hist);
            C2_HIP(hipGetLastError());
            C2_TRY(exclusive_scan(hist, (int64_t)kOrderBins * kOrderBlocks, g_pp.scan_tmp, s));
            k_chunk_order<<<kOrderBlocks, kOrderBins, 0, s>>>(chunks, g_pp.order_bin.as<unsigned char>(),
                                                              flags + P.cap, hist, g_pp.order.as<int>());
            C2_HIP(hipGetLastError());
            chunks = g_pp.order.as<int>();
        }
        if (P.newton) {
            Stage st("pp_react_zero", (double)P.n_used * 24);
            k_react_zero<<<2048, 256, 0, s>>>(A.react, n_sorted, (int)P.cap, &red->err);
            C2_HIP(hipGetLastError());
        }
        C2_TRY(launch_pp_tiled(A, chunks, flags + P.cap, P.grid_pairs, P.mode, s));
        if (P.learned) C2_TRY(learned_enqueue(P.lp, A, g_pp.tile_first.as<int>(), s));
        if (P.newton) {
            Stage st("pp_terms", (double)P.n_used * (16 + 24 + 4 + 12));
            k_pp_newton_terms<<<4096, 256, 0, s>>>(A.pos4, A.slot, A.react, n_sorted, (int)P.cap,
                                                    d_scal, 1.0 / (double)P.react_scale, A.term,
                                                    red);
            C2_HIP(hipGetLastError());
        }
    }
    if (P.split) return CUBE2_OK;  // pp_enqueue_apply updates vp and copies the results
    if (P.newton) C2_TRY(pp_apply_kernels(P, d_vp, dv64, s));
    return pp_results_to_host(P, s);
}

// Velocity update of a split pass, once the pair kernel and PM3 are done:
// for every active sorted particle (range > 0), vp = RN_f(vp + dv64 +
// term / pi) in FP64, the unsplit kernel's operation order; vmax over them.
// (Used when some subtile has no PP range; its particles keep their
// velocities, as in the unsplit pass.)
__global__ __launch_bounds__(256) void k_pp_apply(const float4* __restrict__ pos4,
                                                 const uint32_t* __restrict__ slot,
                                                 const float* __restrict__ term,
                                                 const int* __restrict__ n_sorted, int cap,
                                                 const double* __restrict__ dv64,
                                                 float* __restrict__ vp, PPRed* red) {
    float vm[3] = {0.f, 0.f, 0.f};
    const bool ok = red->err == 0;  // a failed call leaves the sorted arrays unread
    const int n = ok ? min(*n_sorted, cap) : 0;
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < n; a += gridDim.x * blockDim.x) {
        if (pos4[a].w == 0.0f) continue;  // halo sources are not updated
        const size_t ip = (size_t)slot[a] * 3;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double t = (double)term[ip + c] / 3.14159265358979323846;
            double v = (double)vp[ip + c];
            if (dv64) v += dv64[ip + c];
            const float nv = (float)(v + t);
            vp[ip + c] = nv;
            vm[c] = fmaxf(vm[c], fabsf(nv));
        }
    }
    block_max_to<3>(red->vmax_bits, vm);
}

static int pp_apply_kernels(const PPPlan& P, float* d_vp, const double* dv64, hipStream_t s) {
    {
        const Geom& g = ctx().g;
        const int64_t prow = (int64_t)g.ntile * g.nt * g.nt;
        if (P.all_active) {
            Stage st("pp_apply", (double)P.n_used * (12 + 24 + 12 + 12));
            if (prow > 0) {
                const unsigned blocks = std::min<unsigned>(ceil_div(P.cap, kApplySeg), 4096);
                k_pp_apply_rows<<<blocks, 256, 0, s>>>(g_pp.rows_len.as<int>(),
                                                       g_pp.rows_base.as<int64_t>(), (int)prow,
                                                       g_pp.term.as<float>(), dv64, d_vp,
                                                       g_pp.red.as<PPRed>());
                C2_HIP(hipGetLastError());
            }
        } else {
            Stage st("pp_apply", (double)P.n_used * (16 + 4 + 12 + 12 + 24 + 12));
            const unsigned blocks = std::min<unsigned>(ceil_div(P.cap, 256), 8192);
            k_pp_apply<<<blocks, 256, 0, s>>>(g_pp.pos4.as<float4>(), g_pp.slot.as<uint32_t>(),
                                              g_pp.term.as<float>(),
                                              g_pp.counts.as<int>() + P.c_tot, (int)P.cap, dv64,
                                              d_vp, g_pp.red.as<PPRed>());
            C2_HIP(hipGetLastError());
        }
    }
    return CUBE2_OK;
}

int pp_enqueue_apply(const PPPlan& P, float* d_vp, const double* dv64, hipStream_t s) {
    if (!P.split) {
        set_error("velocity apply of an unsplit PP pass");
        return CUBE2_ERR_STATE;
    }
    C2_TRY(pp_apply_kernels(P, d_vp, dv64, s));
    return pp_results_to_host(P, s);
}

// PP's reduction block and tile starts into pinned host memory.
static int pp_results_to_host(const PPPlan& P, hipStream_t s) {
    const Geom& g = ctx().g;
    PPRed* red = g_pp.red.as<PPRed>();
    char* h = g_pp.h_res.as<char>();
    C2_HIP(hipMemcpyAsync(h, red, red_bytes(g, P.ntile), hipMemcpyDeviceToHost, s));
    C2_HIP(hipMemcpyAsync(h + red_bytes(g, P.ntile), g_pp.tile_first.ptr,
                          (P.ntile + 1) * sizeof(int), hipMemcpyDeviceToHost, s));
    return CUBE2_OK;
}

int pp_finish(const PPPlan& P, float* f2max, float* vmax, double* ptotal, double* ttotal,
              uint64_t* pair_counts, double wall_s) {
    const Geom& g = ctx().g;
    const char* h = g_pp.h_res.as<const char>();
    PPRed r;
    std::memcpy(&r, h, sizeof(PPRed));
    const unsigned long long* hp = reinterpret_cast<const unsigned long long*>(h + sizeof(PPRed));
    const int* tf = reinterpret_cast<const int*>(h + red_bytes(g, P.ntile));
    if (r.err & 2) {
        set_error("rhoc and idx_b_r describe slots outside [0, n_used=%lld)", (long long)P.n_used);
        return CUBE2_ERR_ARG;
    }
    if (r.err & 4) {
        set_error("cell lists hold more particles than n_used=%lld (overlapping idx_b_r rows)",
                  (long long)P.n_used);
        return CUBE2_ERR_ARG;
    }
    if (r.err & 1) {
        set_error("a particle's PP key fell outside its coarse cell (geometry too large for FP32 keys)");
        return CUBE2_ERR_CAPACITY;
    }
    const int n_tot = tf[P.ntile] - tf[0];
    *f2max = __builtin_bit_cast(float, r.f2max_bits);
    for (int d = 0; d < 3; d++) vmax[d] = __builtin_bit_cast(float, r.vmax_bits[d]);
    for (int it = 0; it < g.ns3; it++) ptotal[it] = 0.0;
    for (int loc = 0; loc < g.ns3; loc++) ptotal[P.owner[loc]] += (double)hp[loc];
    if (pair_counts) {
        pair_counts[0] = r.p_cand;
        pair_counts[1] = r.p_acc;
    }
    if (ttotal) {
        for (int it = 0; it < g.ns3; it++) ttotal[it] = 0.0;
        for (int t = 0; t < P.ntile; t++) {
            const double share = n_tot > 0 ? (double)(tf[t + 1] - tf[t]) / n_tot : 0.0;
            ttotal[P.owner[t * P.nsub3]] = wall_s * share;
        }
    }
    if (P.count_tiles) g_pp.tile_pairs.assign(hp + g.ns3, hp + g.ns3 + 2 * (size_t)P.ntile);
    else g_pp.tile_pairs.clear();
    g_pp.c_tot = P.c_tot;
    g_pp.n_tot = n_tot;
    g_pp.last = P;
    g_pp.valid = true;
    return CUBE2_OK;
}

int pp_last(const PPPlan** P, PPArgs* A, const int** tile_first) {
    if (!g_pp.valid || g_pp.last.mode == CUBE2_PP_EXACT) {
        set_error("no PP cell lists: run a PP call in a tiled mode first");
        return CUBE2_ERR_STATE;
    }
    const PPPlan& L = g_pp.last;
    *P = &L;
    PPArgs a{};
    a.pos4 = g_pp.pos4.as<float4>();
    a.slot = g_pp.slot.as<uint32_t>();
    a.cellid = g_pp.cellid.as<uint32_t>();
    a.cell_start = g_pp.counts.as<int>();
    a.tiles = g_pp.tiles.as<PPTile>();
    a.n_sorted = g_pp.counts.as<int>() + L.c_tot;
    a.ntile = L.ntile;
    a.cap = (int)L.cap;
    a.ratio_cs = ctx().g.p.ratio_cs;
    a.red = g_pp.red.as<PPRed>();
    *A = a;
    *tile_first = g_pp.tile_first.as<int>();
    return CUBE2_OK;
}

int pp_debug_tile_pairs(int ntile, uint64_t* counts) {
    if (!g_pp.valid || g_pp.tile_pairs.empty()) {
        set_error("no per-tile pair counts: enable them (cube2_pp_count_tiles) before a PP call");
        return CUBE2_ERR_STATE;
    }
    if (!counts || (size_t)ntile * 2 != g_pp.tile_pairs.size()) {
        set_error("per-tile pair counts: %zu tiles, not %d", g_pp.tile_pairs.size() / 2, ntile);
        return CUBE2_ERR_ARG;
    }
    std::copy(g_pp.tile_pairs.begin(), g_pp.tile_pairs.end(), counts);
    return CUBE2_OK;
}

int pp_force(const int* ires, const int* ixyz3, const float* apm3, const int* ratio_sf,
             const int* d_rhoc, const int64_t* d_idx, const int16_t* d_xp, float* d_vp,
             int64_t n_used, float mass, float a_mid, float dt, float* f2max, float* vmax,
             double* ptotal, double* ttotal, uint64_t* pair_counts, bool learned,
             const std::function<int(hipStream_t)>& fill_vp) {
    ensure_params();
    Context& C = ctx();
    auto t0 = std::chrono::steady_clock::now();
    if (!f2max || !vmax || !ptotal) {
        set_error("null argument");
        return CUBE2_ERR_ARG;
    }
    if (C.dv_state == 1 && C.dv_n_used != n_used) {
        set_error("PM3/PP particle count changed within a kick");
        return CUBE2_ERR_STATE;
    }
    // fill_vp (the drop-in call: vp comes from host memory): the pair sums
    // run split from the velocity update, and vp is filled on the second
    // stream beside them (not in the exact-order mode, which has no split
    // pass: vp is filled first)
    const bool late = fill_vp && C.pp_mode != CUBE2_PP_EXACT;
    if (fill_vp && !late) C2_TRY(fill_vp(C.stream));
    PPPlan P;
    C2_TRY(pp_prepare(ires, ixyz3, apm3, ratio_sf, n_used, P, learned || late, learned));
    P.grid_pairs = pp_tiled_grid(P.mode, false, P.newton);  // nothing runs beside it here
    const double* dv = (C.dv_state == 1) ? C.dv64.as<double>() : nullptr;
    const float* d_scal = nullptr;
    C2_TRY(scalars_upload(mass, a_mid, dt, C.stream, &d_scal));
    C2_TRY(pp_enqueue(P, d_rhoc, d_idx, d_xp, d_vp, d_scal, dv, C.stream));
    if (late) {
        static hipEvent_t filled = nullptr;
        if (!filled) C2_HIP(hipEventCreateWithFlags(&filled, hipEventDisableTiming));
        C2_TRY(fill_vp(C.stream2));
        C2_HIP(hipEventRecord(filled, C.stream2));
        C2_HIP(hipStreamWaitEvent(C.stream, filled, 0));
    }
    if (P.split) C2_TRY(pp_enqueue_apply(P, d_vp, dv, C.stream));
    C2_TRY(sync_and_flush());
    C.dv_state = 0;
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return pp_finish(P, f2max, vmax, ptotal, ttotal, pair_counts, wall);
}

int pp_debug_celllist(int it, int* rcp, int* S, int* N, uint32_t* keys, uint32_t* slot,
                      int32_t* cell_start) {
    if (!g_pp.valid || it < 0 || it >= (int)g_pp.h_tiles.size()) {
        set_error("no PP cell list for tile %d", it);
        return CUBE2_ERR_STATE;
    }
    const PPTile& T = g_pp.h_tiles[it];
    const int64_t ncell = (int64_t)T.S * T.S * T.S;
    hipStream_t s = ctx().stream;
    int a0 = 0, a1 = 0;
    C2_HIP(hipMemcpyAsync(&a0, g_pp.counts.as<int>() + T.cell_off, sizeof(int), hipMemcpyDeviceToHost, s));
    C2_HIP(hipMemcpyAsync(&a1, g_pp.counts.as<int>() + T.cell_off + ncell, sizeof(int), hipMemcpyDeviceToHost, s));
    C2_HIP(hipStreamSynchronize(s));
    if (rcp) *rcp = T.rcp;
    if (S) *S = T.S;
    if (N) *N = a1 - a0;
    const int n = a1 - a0;
    if (keys && n > 0) {
        std::vector<uint32_t> cid(n);
        C2_HIP(hipMemcpy(cid.data(), g_pp.cellid.as<uint32_t>() + a0, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
        for (int q = 0; q < n; q++) keys[q] = (uint32_t)(cid[q] - T.cell_off);
    }
    if (slot && n > 0) {
        C2_HIP(hipMemcpy(slot, g_pp.slot.as<uint32_t>() + a0, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
        // dense cells are stored in sub-cell order; the reference lists a
        // cell's particles in slot order
        std::vector<uint32_t> cid(n);
        C2_HIP(hipMemcpy(cid.data(), g_pp.cellid.as<uint32_t>() + a0, n * sizeof(uint32_t),
                         hipMemcpyDeviceToHost));
        for (int q = 0; q < n;) {
            int e = q + 1;
            while (e < n && cid[e] == cid[q]) e++;
            std::sort(slot + q, slot + e);
            q = e;
        }
    }
    if (cell_start) {
        C2_HIP(hipMemcpy(cell_start, g_pp.counts.as<int>() + T.cell_off, (ncell + 1) * sizeof(int),
                         hipMemcpyDeviceToHost));
        for (int64_t q = 0; q <= ncell; q++) cell_start[q] -= a0;
    }
    return CUBE2_OK;
}

void pp_release() {
    PPState& p = g_pp;
    for (DevBuf* b : {&p.nzero, &p.tiles, &p.ires_loc, &p.apm3, &p.counts, &p.scan_tmp, &p.pos4,
                      &p.slot, &p.cellid, &p.red, &p.tile_first, &p.chunk_flags, &p.chunks,
                      &p.raw_pos4, &p.raw_slot, &p.big, &p.dense, &p.term, &p.rows_len,
                      &p.rows_base, &p.order, &p.order_bin, &p.order_hist, &p.react})
        b->release();
    p.valid = false;
}

}  // namespace cube2
