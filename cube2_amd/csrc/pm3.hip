// PM3: fine-mesh particle-mesh force of every subtile (kick.f90:138-193,
// pm_thread kick.f90:403-543, potential_to_force :545-568; GPU reference
// pm3_fft_cuda.cu:276-401,402-591).
//
// Per kick the subtiles are grouped by level (one mesh size n per level)
// and each group runs as batches of jobs sized to a memory budget:
//   deposit   one block per 16^3 brick of a job's mesh sums in LDS the TSC
//             weights (FP64 mesh coordinate, FP32 weights, kick.f90:437-445)
//             of the particles of the coarse cells overlapping the brick,
//             then stores the brick (no clearing, no global atomics);
//   FFT       x R2C (scaled by the particle mass), y, z+Green/n^3+z^-1, y^-1,
//             x C2R (fft.hip), so the potential arrives divided by n^3;
//   gradient  4-point difference with the real(4) weights of
//             parameters.f90:49-50, 3 force planes over [0, nfp+1]^3, max
//             |F|^2 over the interior;
//   interp    one wavefront per physical coarse cell, TSC gather of the 27
//             force points, Delta-v = F*a_mid*dt/6/pi accumulated in FP64.
// The Delta-v stays pending in the library's FP64 accumulator until PP
// applies it (pp_force_kernel.cu:496-498).
#include <algorithm>
#include <climits>
#include <map>

#include "common.h"
#include "fft.h"
#include "profile.h"
#include "engine.h"
#include "reduce.h"
#include "scan.h"

namespace cube2 {

struct PM3Job {
    int t;      // parent tile (x-major)
    int u[3];   // subtile index inside the tile
    int it;     // subtile id (ixyz3 row)
};

struct PM3State {
    DevBuf rho, force, jobs, red, ptot, dep_items, dep_scan, dep_list, dep_cnts, int_items,
        int_list;
    HostBuf h_res;  // f2max bits, error flag, per-job particle counts (pinned)
};
static PM3State g_pm3;

struct Green {
    const float* host;
    size_t count;
    float* dev;
};
static std::vector<Green> g_greens;

// ---------------------------------------------------------------- kernels

struct MeshArgs {
    const int* rhoc;
    const int64_t* nzero;
    const int16_t* xp;
    const PM3Job* jobs;
    int njobs, nt, ntt, ncb, nte, nnt, ratio_cs;  // ntt: coarse cells per job edge
    int n, nfp, nfb, rsf;
    int nhalo;  // coarse halo cells around a job whose particles are deposited
    int layer;  // 3: PM3 subtile jobs, 2: PM2 parent-tile jobs
    int64_t n_used;
};

// TSC of one axis (kick.f90:439-445): centre cell Fortran index c+1 where
// c = floor(x); FP32 weights from the FP64 coordinate.
__device__ inline void tsc_axis(double x, int& c, float w[3]) {
#pragma clang fp contract(off)
    const double fl = floor(x);
    c = (int)fl;
    const float l3 = (float)(x - fl);
    w[0] = (1.0f - l3) * (1.0f - l3) / 2.0f;
    w[2] = l3 * l3 / 2.0f;
    w[1] = 1.0f - w[0] - w[2];
}

__device__ inline double mesh_coord(const MeshArgs& A, int c_nte, int16_t v, int u) {
    const double xc = (double)(c_nte - A.ncb) + frac16(v);
    return xc * (double)(A.ratio_cs * A.rsf) - (double)(u * A.nfp);
}

// floor(mesh_coord) exactly, in integer arithmetic: with U the unsigned
// 16-bit position, frac * cf = (2U + 1) cf / 2^17.
__device__ inline int fine_floor(const MeshArgs& A, int c_nte, int16_t v, int u, int cf) {
    const int U = (int)(int16_t)((int)v - 32768) + 32768;
    return (c_nte - A.ncb) * cf - u * A.nfp + (int)(((int64_t)(2 * U + 1) * cf) >> 17);
}

constexpr int kWaves = 4;

// Cell-group deposit.  A work item is (job, group of g^3 coarse cells of
// the job's deposit range, split, z-slab): it visits one share (every
// nsplit-th chunk of 256 in the group's row order) of the group's particles
// once and sums their TSC weights in LDS over the region the group's
// particles reach, (g cf + 2)^2 fine cells x the slab's Zs planes.  A group
// is split so that no item has more than kDepSplit + 256 particles: dense
// clumps spread over many blocks.  LDS sums are 32-bit fixed point: LDS
// integer atomics run several times faster than float ones and the sums do
// not depend on the order of the adds.  An item with at most c < 2^b
// particles scales its weights (each <= 0.75^3) by 2^(33-b), capped at
// 2^30, so that a point's sum stays below 0.75^3 2^33 < 2^32; light items
// keep ~30 fraction bits, dense ones (large sums, where an FP32 sum's own
// rounding is coarser) at least 18.  When a group is not split, its region's interior
// (2 or more cells from its faces) belongs to the item alone and is stored;
// the shell that neighbouring groups overlap and the points of split groups
// are added atomically.  The mesh is cleared before, so zero sums are
// skipped.  Mesh points outside [0, n)^3 are dropped (kick.f90:463: only
// rho_th(m1:m2) is kept).
constexpr int kDepThreads = 256;
constexpr int kDepLdsBytes = 80 * 1024;  // two blocks per CU
constexpr int kDepSplit = 8192;          // particles per work item at most

struct DepositShape {
    int g, cf, E, Zs, ngx, nslab;
    size_t lds;
    int64_t groups;  // ngx^3 per job
};

static DepositShape deposit_shape(const MeshArgs& A) {
    DepositShape S;
    S.cf = A.ratio_cs * A.rsf;
    S.g = std::max(1, std::min(4, 16 / S.cf));
    S.E = S.g * S.cf + 2;
    const int D = A.ntt + 2 * A.nhalo;
    S.ngx = (D + S.g - 1) / S.g;
    const int zmax = std::max(1, kDepLdsBytes / (int)sizeof(unsigned) / (S.E * S.E));
    S.nslab = (S.E + zmax - 1) / zmax;
    S.Zs = (S.E + S.nslab - 1) / S.nslab;
    S.lds = (size_t)S.E * S.E * S.Zs * sizeof(unsigned);
    S.groups = (int64_t)S.ngx * S.ngx * S.ngx;
    return S;
}

// Coarse cells [c0, c1) (nte indices) of group gi of job J.
__device__ inline void group_cells(const MeshArgs& A, const PM3Job& J, int g, int ngx, int q,
                                   int c0[3], int c1[3]) {
    const int gi[3] = {q % ngx, (q / ngx) % ngx, q / (ngx * ngx)};
    const int D = A.ntt + 2 * A.nhalo;  // coarse cells per edge of the deposit range
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int base = J.u[d] * A.ntt + A.ncb - A.nhalo;
        c0[d] = base + gi[d] * g;
        c1[d] = min(c0[d] + g, base + D);
    }
}

// items[jb * G + q] = work items of group q of job jb (splits x slabs);
// cnts[jb * G + q] = its particles.
__global__ void k_dep_count(MeshArgs A, int g, int ngx, int nslab, int* __restrict__ items,
                            int* __restrict__ cnts) {
    const int G = ngx * ngx * ngx;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)A.njobs * G) return;
    const int jb = (int)(t / G), q = (int)(t - (int64_t)jb * G);
    const PM3Job J = A.jobs[jb];
    int c0[3], c1[3];
    group_cells(A, J, g, ngx, q, c0, c1);
    int64_t cnt = 0;
    for (int ck = c0[2]; ck < c1[2]; ck++)
        for (int cj = c0[1]; cj < c1[1]; cj++) {
            const int64_t row = (((int64_t)J.t * A.nte + ck) * A.nte + cj) * A.nte;
            for (int ci = c0[0]; ci < c1[0]; ci++) cnt += max(A.rhoc[row + ci], 0);
        }
    items[t] = (int)((cnt + kDepSplit - 1) / kDepSplit) * nslab;
    cnts[t] = (int)min(cnt, (int64_t)INT_MAX);
}

// list[offs[jq] + r] = jq for the items r of group jq (offs: exclusive scan
// of k_dep_count's items); entries past cap are dropped (the deposit then
// reports an error).
__global__ void k_dep_list(const int* __restrict__ offs, int64_t ngroups, int cap,
                           int* __restrict__ list) {
    const int64_t jq = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (jq >= ngroups) return;
    const int o = offs[jq], c = offs[jq + 1] - o;
    for (int r = 0; r < c && o + r < cap; r++) list[o + r] = (int)jq;
}

// Persistent: blocks take items it = blockIdx.x, + gridDim.x, ... below the
// total offs[ngroups]; list[it] names the item's group.
__global__ __launch_bounds__(kDepThreads) void k_mesh_deposit(MeshArgs A, float* __restrict__ rho,
                                                              size_t rs, int g, int cf, int E,
                                                              int Zs, int ngx, int nslab,
                                                              const int* __restrict__ offs,
                                                              const int* __restrict__ list, int cap,
                                                              const int* __restrict__ cnts,
                                                              int* __restrict__ err) {
    extern __shared__ unsigned box[];  // [Zs][E][E], fixed point
    const int n = A.n, off = A.nfb - 1;  // storage x = Fortran index + nfb - 1
    const int G = ngx * ngx * ngx;
    const int ngroups = A.njobs * G;
    const int total = offs[ngroups];
    if (total > cap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(err, 8);
        return;
    }
    for (int it = blockIdx.x; it < total; it += gridDim.x) {
        const int jq = list[it];  // offs[jq] <= it < offs[jq + 1]
        const int r = it - offs[jq];
        const int nsplit = (offs[jq + 1] - offs[jq]) / nslab;
        const int split = r / nslab, slab = r - split * nslab;
        // fixed-point scale from the item's particle bound (round-robin chunks)
        const int cnt = cnts[jq];
        const int nch = (cnt + (int)blockDim.x - 1) / (int)blockDim.x;  // chunks of the group
        const int bound = max(1, min(cnt, (nch + nsplit - 1) / nsplit * (int)blockDim.x));
        const int fbits = min(30, 33 - (32 - __clz(bound)));
        const float fix = ldexpf(1.0f, fbits), inv_fix = ldexpf(1.0f, -fbits);
        const int jb = jq / G, q = jq - jb * G;
        const PM3Job J = A.jobs[jb];
        int c0[3], c1[3], X0[3], Ee[3];
        group_cells(A, J, g, ngx, q, c0, c1);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            X0[d] = (c0[d] - A.ncb) * cf - J.u[d] * A.nfp + off;  // storage index of local 0
            Ee[d] = (c1[d] - c0[d]) * cf + 2;
        }
        const int zlo = slab * Zs, nz = min(Zs, Ee[2] - zlo);  // region-local planes of the slab
        if (nz <= 0) continue;  // uniform over the block
        __syncthreads();        // the previous item's write-out has read the LDS
        for (int e = threadIdx.x; e < E * E * nz; e += blockDim.x) box[e] = 0u;
        __syncthreads();
        const int gx = c1[0] - c0[0];  // 1..4 cells along x
        const int T = (int)blockDim.x;
        int64_t gbase = 0;  // group-order index of the row's first particle
        for (int ck = c0[2]; ck < c1[2]; ck++) {
            const int zc = (ck - c0[2]) * cf;  // the cell reaches local planes [zc, zc + cf + 2)
            const bool reach = !(zc + cf + 2 <= zlo || zc >= zlo + nz);
            for (int cj = c0[1]; cj < c1[1]; cj++) {
                // the particles of cells c0..c1-1 of a row are contiguous slots
                const int64_t cell0 = (((int64_t)J.t * A.nte + ck) * A.nte + cj) * A.nte + c0[0];
                const int64_t s0 = A.nzero[cell0];
                const int64_t e0 = s0 + A.rhoc[cell0];
                const int64_t e1 = gx > 1 ? e0 + A.rhoc[cell0 + 1] : e0;
                const int64_t e2 = gx > 2 ? e1 + A.rhoc[cell0 + 2] : e1;
                const int64_t e3 = gx > 3 ? e2 + A.rhoc[cell0 + 3] : e2;
                if (e0 == s0 && e1 == e0 && e2 == e1 && e3 == e2) continue;  // no particles
                if (s0 < 0 || e0 < s0 || e1 < e0 || e2 < e1 || e3 < e2 || e3 > A.n_used) {
                    if (threadIdx.x == 0 && split == 0 && slab == 0) atomicOr(err, 2);
                    continue;
                }
                const int64_t rlo = gbase, rhi = gbase + (e3 - s0);
                gbase = rhi;
                if (!reach) continue;
                // chunks of T particles in group order go round-robin to the splits
                int64_t ch = rlo / T;
                ch += ((split - ch % nsplit) + nsplit) % nsplit;
                for (; ch * T < rhi; ch += nsplit) {
                    const int64_t gi = ch * T + threadIdx.x;
                    if (gi < rlo || gi >= rhi) continue;
                    const int64_t l = s0 + (gi - rlo);
                    const int ci = c0[0] + (l >= e0) + (l >= e1) + (l >= e2);
                    const int16_t* p = A.xp + 3 * l;
                    // floor of the mesh coordinate in integer arithmetic (the
                    // value tsc_axis computes); x, y land in [0, Ee - 3]
                    const int Xl = fine_floor(A, ci, p[0], J.u[0], cf) + off - X0[0];
                    const int Yl = fine_floor(A, cj, p[1], J.u[1], cf) + off - X0[1];
                    const int Zl = fine_floor(A, ck, p[2], J.u[2], cf) + off - X0[2] - zlo;
                    if (Zl > nz - 1 || Zl < -2) continue;
                    int cx, cy, cz;
                    float wx[3], wy[3], wz[3];
                    tsc_axis(mesh_coord(A, ci, p[0], J.u[0]), cx, wx);
                    tsc_axis(mesh_coord(A, cj, p[1], J.u[1]), cy, wy);
                    tsc_axis(mesh_coord(A, ck, p[2], J.u[2]), cz, wz);
                    (void)cx, (void)cy, (void)cz;  // equal to the integer floors above
#pragma unroll
                    for (int a3 = 0; a3 < 3; a3++) {
                        const int Z = Zl + a3;
                        if ((unsigned)Z >= (unsigned)nz) continue;
#pragma unroll
                        for (int a2 = 0; a2 < 3; a2++) {
                            unsigned* row = box + (Z * E + Yl + a2) * E + Xl;
#pragma unroll
                            for (int a1 = 0; a1 < 3; a1++) {
                                const float w = wx[a1] * wy[a2] * wz[a3];
                                atomicAdd(row + a1, (unsigned)rintf(w * fix));
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();
        float* R = rho + (size_t)jb * rs;
        const int ex = Ee[0], ey = Ee[1], ez = Ee[2];
        for (int e = threadIdx.x; e < ex * ey * nz; e += blockDim.x) {
            const int x = e % ex, y = (e / ex) % ey, zl = e / (ex * ey), z = zl + zlo;
            const int X = X0[0] + x, Y = X0[1] + y, Z = X0[2] + z;
            if ((unsigned)X >= (unsigned)n || (unsigned)Y >= (unsigned)n ||
                (unsigned)Z >= (unsigned)n)
                continue;
            const unsigned u = box[(zl * E + y) * E + x];
            if (u == 0u) continue;  // the mesh is cleared
            const float v = (float)u * inv_fix;  // one FP32 rounding of the sum
            float* dst = R + ((size_t)Z * n + Y) * (n + 2) + X;
            if (nsplit == 1 && x >= 2 && x < ex - 2 && y >= 2 && y < ey - 2 && z >= 2 && z < ez - 2)
                *dst = v;
            else
                unsafeAtomicAdd(dst, v);
        }
    }
}

// Brick deposit: block (brick, job) owns a kBrick^3 region of the job's
// mesh (storage indices) and sums in LDS every TSC weight that lands in it
// from the particles of the coarse cells that can reach it (64-bit fixed
// point, each weight rounded to 2^-32: integer LDS atomics, so the sums do not
// depend on the order of the adds), then stores the brick.  Every mesh point
// belongs to one brick: the mesh needs no clearing and HBM sees no atomics,
// at the price of visiting a particle once per brick its cell overlaps.
// The block first reads the counts and first slots of all its cells at once
// (at most kBrickCells) and takes their particles as one list, kDepBrickThreads
// at a time, so that neither a chain of per-cell loads nor a cell of a few
// particles keeps the block's other lanes waiting (at 256^3 most bricks of
// the dense levels see a few cells of a few particles each).  A brick's time
// is mostly its chain of dependent loads; with the box's 32 KB of LDS four
// blocks fit a CU, so 8 wavefronts per block rather than 4 keep more of the
// particle loads in flight (at 256^3: 10.3 -> 8.3 ms per kick; 16
// wavefronts: 12.7).
constexpr int kBrick = 16;
constexpr int kBrickCells = 256;
constexpr int kDepBrickThreads = 512;

__device__ inline int floordiv(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// Adds one particle's 27 TSC weights that fall in the brick (64-bit fixed
// point).  kRot: lanes start their 3 x 3 (z, y) rows at different rows
// (rotation by lane), so that the particles of one dense fine cell do not add
// to the same LDS word in the same step; the sums are integers, so the order
// of the adds does not change them.  It pays on the coarsest meshes, whose
// fine cells hold many particles, and costs on the finer ones (at 256^3 it
// is used for 8 fine cells per coarse cell: level 2, 1.9 -> 1.1-1.4 ms).
template <bool kRot>
__device__ inline void brick_add(unsigned long long* box, int X0, int Y0, int Z0,
                                 const float wx[3], const float wy[3], const float wz[3]) {
    const int rz = kRot ? (int)(threadIdx.x % 3u) : 0;
    const int ry = kRot ? (int)((threadIdx.x / 3u) % 3u) : 0;
#pragma unroll
    for (int s3 = 0; s3 < 3; s3++) {
        const int a3 = kRot ? (s3 + rz >= 3 ? s3 + rz - 3 : s3 + rz) : s3;
        const int Z = Z0 + a3;
        if ((unsigned)Z >= (unsigned)kBrick) continue;
        const float wza = kRot ? (a3 == 0 ? wz[0] : a3 == 1 ? wz[1] : wz[2]) : wz[s3];
#pragma unroll
        for (int s2 = 0; s2 < 3; s2++) {
            const int a2 = kRot ? (s2 + ry >= 3 ? s2 + ry - 3 : s2 + ry) : s2;
            const int Y = Y0 + a2;
            if ((unsigned)Y >= (unsigned)kBrick) continue;
            const float wya = kRot ? (a2 == 0 ? wy[0] : a2 == 1 ? wy[1] : wy[2]) : wy[s2];
#pragma unroll
            for (int a1 = 0; a1 < 3; a1++) {
                const int X = X0 + a1;
                if ((unsigned)X >= (unsigned)kBrick) continue;
                const float w = wx[a1] * wya * wza;  // <= 0.75^3
                atomicAdd(&box[(Z * kBrick + Y) * kBrick + X],
                          (unsigned long long)(unsigned)rintf(w * 4294967296.0f));
            }
        }
    }
}

__global__ __launch_bounds__(kDepBrickThreads) void k_brick_deposit(MeshArgs A, float* __restrict__ rho,
                                                       size_t rs, int nbx, int* __restrict__ err) {
    __shared__ unsigned long long box[kBrick * kBrick * kBrick];
    __shared__ int cstart[kBrickCells + 1];  // exclusive prefix of the cells' counts
    __shared__ int64_t cbase[kBrickCells];   // first xp slot of each cell
    __shared__ int cmesh[3][kBrickCells];    // its mesh coordinate origin (ci-ncb)*cf - u*nfp
    __shared__ int wsum[kDepBrickThreads / 64];
    const int n = A.n, off = A.nfb - 1;  // storage x = Fortran index + nfb - 1
    // neighbouring bricks visit the same particles: they run on one XCD
    const unsigned lb = xcd_remap(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int jb = (int)(lb / gridDim.x);
    const int b = (int)(lb % gridDim.x);
    const int bx = b % nbx, by = (b / nbx) % nbx, bz = b / (nbx * nbx);
    const int lo[3] = {bx * kBrick, by * kBrick, bz * kBrick};
    for (int q = threadIdx.x; q < kBrick * kBrick * kBrick; q += blockDim.x) box[q] = 0ull;
    const PM3Job J = A.jobs[jb];
    const int cf = A.ratio_cs * A.rsf;  // fine cells per coarse cell
    // A particle with c = floor(xpos) writes storage c+off .. c+off+2, so it
    // reaches [lo, lo+kBrick) iff c in [lo-off-2, lo+kBrick-off-1]; its
    // coarse cell (nte index) is floor((c + u*nfp)/cf) + ncb.
    int clo[3], ext[3];
    for (int d = 0; d < 3; d++) {
        const int cmin = lo[d] - off - 2 + J.u[d] * A.nfp;
        const int cmax = lo[d] + kBrick - off - 1 + J.u[d] * A.nfp;
        const int a = floordiv(cmin, cf) + A.ncb;
        const int e = floordiv(cmax, cf) + A.ncb;
        const int dlo = J.u[d] * A.ntt + A.ncb - A.nhalo,
                  dhi = (J.u[d] + 1) * A.ntt + A.ncb - 1 + A.nhalo;
        clo[d] = max(a, dlo);
        ext[d] = max(0, min(e, dhi) - clo[d] + 1);
    }
    const int ncell = ext[0] * ext[1] * ext[2];  // <= kBrickCells (checked on the host)
    const bool rot = cf < 16;
    {
        const int q = threadIdx.x, wave = q >> 6, lane = q & 63;
        int np = 0;
        int64_t base = 0;
        if (q < ncell) {
            const int ci = clo[0] + q % ext[0], cj = clo[1] + (q / ext[0]) % ext[1],
                      ck = clo[2] + q / (ext[0] * ext[1]);
            cmesh[0][q] = (ci - A.ncb) * cf - J.u[0] * A.nfp;
            cmesh[1][q] = (cj - A.ncb) * cf - J.u[1] * A.nfp;
            cmesh[2][q] = (ck - A.ncb) * cf - J.u[2] * A.nfp;
            const int64_t cell = (((int64_t)J.t * A.nte + ck) * A.nte + cj) * A.nte + ci;
            np = A.rhoc[cell];
            base = A.nzero[cell];
            if (np <= 0) {
                np = 0;
            } else if (base < 0 || base + np > A.n_used) {
                atomicOr(err, 2);
                np = 0;
            }
        }
        int inc = np;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(inc, o, 64);
            if (lane >= o) inc += t;
        }
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        for (int w = 0; w < wave; w++) inc += wsum[w];
        if (q < kBrickCells) {
            cstart[q + 1] = inc;
            cbase[q] = base;
        }
        if (q == 0) cstart[0] = 0;
    }
    __syncthreads();
    const int M = cstart[ncell];
    int c = 0;
    for (int v = threadIdx.x; v < M; v += blockDim.x) {
        while (cstart[c + 1] <= v) c++;  // v < M: stops at a cell that holds v
        const int16_t* p = A.xp + 3 * (cbase[c] + (v - cstart[c]));
        // The mesh coordinate is the cell's origin plus frac * cf, frac =
        // (U + 1/2) / 2^16 with U the 16-bit code read unsigned; its floor is
        // origin + ((2U + 1) cf >> 17) in integer arithmetic (the value
        // tsc_axis computes), so particles that miss the brick are rejected
        // before the FP64 work.  (2U + 1) cf < 2^24: a 24-bit multiply.
        const int fx = (2 * (int)(uint16_t)p[0] + 1) * cf, fy = (2 * (int)(uint16_t)p[1] + 1) * cf,
                  fz = (2 * (int)(uint16_t)p[2] + 1) * cf;
        const int mx = cmesh[0][c], my = cmesh[1][c], mz = cmesh[2][c];
        const int X0 = mx + (fx >> 17) + off - lo[0];
        const int Y0 = my + (fy >> 17) + off - lo[1];
        const int Z0 = mz + (fz >> 17) + off - lo[2];
        if (X0 > kBrick - 1 || X0 < -2 || Y0 > kBrick - 1 || Y0 < -2 || Z0 > kBrick - 1 ||
            Z0 < -2)
            continue;
        int cx, cy, cz;
        float wx[3], wy[3], wz[3];
        // mesh_coord's value: every term is exact in FP64, so origin +
        // (2U + 1) cf 2^-17 equals ((c - ncb) + frac) cf - u nfp
        tsc_axis((double)mx + (double)fx * 0x1p-17, cx, wx);
        tsc_axis((double)my + (double)fy * 0x1p-17, cy, wy);
        tsc_axis((double)mz + (double)fz * 0x1p-17, cz, wz);
        (void)cx, (void)cy, (void)cz;  // equal to the integer floors above
        if (rot)
            brick_add<true>(box, X0, Y0, Z0, wx, wy, wz);
        else
            brick_add<false>(box, X0, Y0, Z0, wx, wy, wz);
    }
    __syncthreads();
    float* R = rho + (size_t)jb * rs;
    for (int q = threadIdx.x; q < kBrick * kBrick * kBrick; q += blockDim.x) {
        const int X = lo[0] + (q % kBrick), Y = lo[1] + (q / kBrick) % kBrick,
                  Z = lo[2] + q / (kBrick * kBrick);
        if (X < n && Y < n && Z < n)
            R[((size_t)Z * n + Y) * (n + 2) + X] = (float)((double)box[q] * (1.0 / 4294967296.0));
    }
}

// Coarse cells per axis a brick can visit: the c range of 2 + kBrick fine
// cells spans at most (kBrick + 1) / cf + 2 cells.
static bool brick_cells_fit(int cf) {
    const int per = (kBrick + 1) / std::max(1, cf) + 2;
    return per * per * per <= kBrickCells;
}

// Which deposit a mesh gets: bricks on the fine PM3 meshes, cell-group work
// items on PM2's (dense clumps would leave a few bricks with most of the
// work).
static bool use_bricks(const MeshArgs& A) { return A.layer == 3; }

// Blocks of the persistent deposit: what fits the CUs at once.
static int deposit_grid(size_t lds) {
    static std::map<size_t, int> cache;
    auto it = cache.find(lds);
    if (it != cache.end()) return it->second;
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        ncu = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_mesh_deposit, kDepThreads, lds) !=
            hipSuccess ||
        per < 1)
        per = 2;
    // eight times what is resident: blocks retire and are handed out again as
    // the items' uneven costs come in
    cache[lds] = std::max(1, ncu) * per * 8;
    return cache[lds];
}

// Force planes over Fortran indices [0, nfp+1]^3 (kick.f90:545-568) and the
// interior max |F|^2 (kick.f90:494).  A block owns a tile of W x-points x
// kGradH rows and marches through up to Zc output planes: the phi planes it
// needs pass through a 5-plane LDS ring, each read from HBM once (with its
// 2-wide x/y halo), instead of each z-plane block re-reading its 4
// neighbour planes (neighbouring planes' blocks run on other XCDs, each
// with its own L2).
constexpr int kGradH = 8;
constexpr int kGradW = 64;

__global__ __launch_bounds__(256) void k_pm3_grad(const float* __restrict__ phi, size_t rs,
                                                  float* __restrict__ force, int n, int nfp,
                                                  int nfb, int W, int ntx, int Zc,
                                                  unsigned* __restrict__ f2bits) {
    __shared__ float ring[5][kGradH + 4][kGradW + 4];
    const int nf = nfp + 2;
    // neighbouring tiles share halo lines: they run on one XCD
    const unsigned lb = xcd_remap(
        (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x,
        gridDim.x * gridDim.y * gridDim.z);
    const int bxy = (int)(lb % gridDim.x);
    const int tx = bxy % ntx, ty = bxy / ntx;
    const int x0 = tx * W, y0 = ty * kGradH;
    const int k_begin = (int)((lb / gridDim.x) % gridDim.y) * Zc, k_end = min(nf, k_begin + Zc);
    const int jb = (int)(lb / (gridDim.x * gridDim.y));
    const float* P = phi + (size_t)jb * rs;
    const int sr = n + 2;
    const size_t sp = (size_t)n * sr;
    const int ex = min(W, nf - x0), ey = min(kGradH, nf - y0);  // outputs of the tile
    const int lw = ex + 4, lh = ey + 4;
    constexpr int kRW = kGradW + 4;  // ring row pitch
    // Each thread's share of a plane, fixed for the whole march: the ring
    // cells it loads (offset in the phi plane, place in the ring) and the
    // outputs it computes (offset in a force plane, place of the stencil's
    // centre in the ring), so that a plane costs no index arithmetic beyond
    // one base address per array.
    constexpr int kPre = ((kGradH + 4) * kRW + 255) / 256;
    int loff[kPre], lpos[kPre];
#pragma unroll
    for (int i = 0; i < kPre; i++) {
        const int e = threadIdx.x + 256 * i;
        const int lx = e % kRW, ly = e / kRW;  // constant divisor
        const bool v = lx < lw && ly < lh;
        loff[i] = v ? (y0 + ly + nfb - 3) * sr + (x0 + lx + nfb - 3) : -1;
        lpos[i] = ly * kRW + lx;
    }
    constexpr int kOut = kGradW * kGradH / 256;  // outputs per thread
    int ooff[kOut], opos[kOut];
    bool inner[kOut];
#pragma unroll
    for (int q = 0; q < kOut; q++) {
        const int lx = (int)threadIdx.x % kGradW, ly = (int)threadIdx.x / kGradW + q * (256 / kGradW);
        const bool v = lx < ex && ly < ey;
        const int i0 = x0 + lx, j0 = y0 + ly;
        ooff[q] = v ? j0 * nf + i0 : -1;
        opos[q] = (ly + 2) * kRW + lx + 2;
        inner[q] = v && i0 >= 1 && i0 <= nfp && j0 >= 1 && j0 <= nfp;
    }
    float* ring0 = &ring[0][0][0];
    constexpr int kSlot = (kGradH + 4) * kRW;
    // phi plane of output plane k (storage z = k + nfb - 1) into a ring slot,
    // in two halves: the loads into registers are issued before the previous
    // plane's outputs are computed, so their latency hides behind that work;
    // the registers go to the ring after it
    float pre[kPre];
    auto fetch = [&](int k) {
        const float* Pz = P + (size_t)(k + nfb - 1) * sp;
#pragma unroll
        for (int i = 0; i < kPre; i++)
            if (loff[i] >= 0) pre[i] = Pz[loff[i]];
    };
    auto store = [&](int k) {
        float* R = ring0 + ((k + 5) % 5) * kSlot;
#pragma unroll
        for (int i = 0; i < kPre; i++)
            if (loff[i] >= 0) R[lpos[i]] = pre[i];
    };
    // real(4) weights of alpha = 4/3 (parameters.f90:49-50)
    const float al = 4.0f / 3.0f;
    const float w0 = (al - 1.0f) / 4.0f, w1 = -al / 2.0f, w2 = al / 2.0f, w3 = (1.0f - al) / 4.0f;
    const size_t fs = (size_t)nf * nf * nf;
    float* F = force + (size_t)jb * 3 * fs;
    float f2 = 0.0f;
    for (int k = k_begin - 2; k < k_begin + 2; k++) {
        fetch(k);
        store(k);
    }
    if (k_begin < k_end) fetch(k_begin + 2);
    for (int k0 = k_begin; k0 < k_end; k0++) {
        store(k0 + 2);  // its slot held plane k0 - 3, last read before the previous barrier
        __syncthreads();
        if (k0 + 1 < k_end) fetch(k0 + 3);
        const float* Rm2 = ring0 + ((k0 + 3) % 5) * kSlot;
        const float* Rm1 = ring0 + ((k0 + 4) % 5) * kSlot;
        const float* R0 = ring0 + (k0 % 5) * kSlot;
        const float* Rp1 = ring0 + ((k0 + 1) % 5) * kSlot;
        const float* Rp2 = ring0 + ((k0 + 2) % 5) * kSlot;
        float* Fx = F + (size_t)3 * k0 * nf * nf;  // force points interleaved (x, y, z)
        const bool kin = k0 >= 1 && k0 <= nfp;
#pragma unroll
        for (int q = 0; q < kOut; q++) {
            if (ooff[q] < 0) continue;
            const int c = opos[q];
            float fx, fy, fz;
            {
#pragma clang fp contract(off)
                // phi arrives divided by n^3 (folded into the Green's table)
                fx = R0[c - 2] * w0 + R0[c - 1] * w1 + R0[c + 1] * w2 + R0[c + 2] * w3;
                fy = R0[c - 2 * kRW] * w0 + R0[c - kRW] * w1 + R0[c + kRW] * w2 +
                     R0[c + 2 * kRW] * w3;
                fz = Rm2[c] * w0 + Rm1[c] * w1 + Rp1[c] * w2 + Rp2[c] * w3;
            }
            float* Fp = Fx + 3 * (size_t)ooff[q];
            Fp[0] = fx;
            Fp[1] = fy;
            Fp[2] = fz;
            if (inner[q] && kin) f2 = fmaxf(f2, fx * fx + fy * fy + fz * fz);
        }
        __syncthreads();  // the next load overwrites plane k0 - 2
    }
    const float fv[1] = {f2};
    block_max_to<1>(f2bits, fv);
}

// Interpolation work items: one wavefront per (job, physical coarse cell,
// segment of seg64*64 particles), so a dense cell spreads over many
// wavefronts and a light one takes one.  items[jb * per + r] = segments of
// cell r of job jb.
__global__ void k_interp_count(MeshArgs A, int seglen, int* __restrict__ items) {
    const int64_t per = (int64_t)A.ntt * A.ntt * A.ntt;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= per * A.njobs) return;
    const int jb = (int)(t / per);
    const int r = (int)(t - (int64_t)jb * per);
    const PM3Job J = A.jobs[jb];
    const int ci = J.u[0] * A.ntt + A.ncb + r % A.ntt;
    const int cj = J.u[1] * A.ntt + A.ncb + (r / A.ntt) % A.ntt;
    const int ck = J.u[2] * A.ntt + A.ncb + r / (A.ntt * A.ntt);
    const int np = A.rhoc[(((int64_t)J.t * A.nte + ck) * A.nte + cj) * A.nte + ci];
    items[t] = np > 0 ? (np + seglen - 1) / seglen : 0;
}

// TSC gather of the force planes and FP64 Delta-v (kick.f90:497-538), a
// wavefront per work item (list[it] names its cell, offs the cells' first
// items); persistent over the items.
__global__ __launch_bounds__(256) void k_pm3_interp(MeshArgs A, const float* __restrict__ force,
                                                    double* __restrict__ dv64, int accumulate,
                                                    const float* __restrict__ scal, int seglen,
                                                    const int* __restrict__ offs,
                                                    const int* __restrict__ list, int cap,
                                                    unsigned long long* __restrict__ ptot,
                                                    int* __restrict__ err) {
    const float amid = scal[1], dt = scal[2];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t per = (int64_t)A.ntt * A.ntt * A.ntt;
    const int total = offs[per * A.njobs];
    if (total > cap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(err, 8);
        return;
    }
    const int nf = A.nfp + 2;
    const size_t fs = (size_t)nf * nf * nf;
    for (int it = blockIdx.x * kWaves + wave; it < total; it += gridDim.x * kWaves) {
        const int jq = list[it];
        const int seg = it - offs[jq];
        const int jb = (int)(jq / per);
        const int r = (int)(jq - (int64_t)jb * per);
        const PM3Job J = A.jobs[jb];
        const int ci = J.u[0] * A.ntt + A.ncb + r % A.ntt;
        const int cj = J.u[1] * A.ntt + A.ncb + (r / A.ntt) % A.ntt;
        const int ck = J.u[2] * A.ntt + A.ncb + r / (A.ntt * A.ntt);
        const int64_t cell = (((int64_t)J.t * A.nte + ck) * A.nte + cj) * A.nte + ci;
        const int np = A.rhoc[cell];
        const int64_t base = A.nzero[cell];
        if (np <= 0 || base < 0 || base + np > A.n_used) continue;
        if (seg == 0 && lane == 0) atomicAdd(&ptot[jb], (unsigned long long)np);
        const float* F = force + (size_t)jb * 3 * fs;
        const int lend = min(np, (seg + 1) * seglen);
        for (int l = seg * seglen + lane; l < lend; l += 64) {
            const int64_t ip = base + l;
            const int16_t* p = A.xp + 3 * ip;
            int cx, cy, cz;
            float wx[3], wy[3], wz[3];
            tsc_axis(mesh_coord(A, ci, p[0], J.u[0]), cx, wx);
            tsc_axis(mesh_coord(A, cj, p[1], J.u[1]), cy, wy);
            tsc_axis(mesh_coord(A, ck, p[2], J.u[2]), cz, wz);
            float X = 0.0f, Y = 0.0f, Z = 0.0f;
#pragma unroll
            for (int a3 = 0; a3 < 3; a3++)
#pragma unroll
                for (int a2 = 0; a2 < 3; a2++)
#pragma unroll
                    for (int a1 = 0; a1 < 3; a1++) {
#pragma clang fp contract(off)
                        const size_t u = ((size_t)(cz + a3) * nf + (cy + a2)) * nf + (cx + a1);
                        const float* Fp = F + 3 * u;  // one 12-byte point
                        X = X + Fp[0] * wx[a1] * wy[a2] * wz[a3];
                        Y = Y + Fp[1] * wx[a1] * wy[a2] * wz[a3];
                        Z = Z + Fp[2] * wx[a1] * wy[a2] * wz[a3];
                    }
            const float acc[3] = {X, Y, Z};
#pragma unroll
            for (int c = 0; c < 3; c++) {
                double t;
                {
#pragma clang fp contract(off)
                    t = (double)(acc[c] * amid * dt / 6.0f) / 3.14159265358979323846;
                }
                if (accumulate) dv64[3 * ip + c] += t;
                else dv64[3 * ip + c] = t;
            }
        }
    }
}

// Blocks of the interpolation, whose wavefronts take items by a static stride.
static int interp_grid() {
    static int cached = 0;
    if (cached) return cached;
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        ncu = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_pm3_interp, 256, 0) != hipSuccess ||
        per < 1)
        per = 2;
    // eight times what is resident: the hardware hands out blocks as others
    // retire, which evens out the items' uneven costs (256^3: 2.35 -> 2.12 ms)
    cached = std::max(1, ncu) * per * 8;
    return cached;
}

// Pending Delta-v applied without PP; vmax over physical slots.
__global__ __launch_bounds__(256) void k_velocity_apply(const int* __restrict__ rhoc,
                                                        const int64_t* __restrict__ nzero, int nt,
                                                        int ncb, int nte, int ntile, int64_t n_used,
                                                        const double* __restrict__ dv64,
                                                        float* __restrict__ vp,
                                                        unsigned* __restrict__ vbits) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t per = (int64_t)nt * nt * nt;
    const int64_t w = (int64_t)blockIdx.x * kWaves + wave;
    float v0 = 0, v1 = 0, v2 = 0;
    if (w < per * ntile) {
        const int t = (int)(w / per);
        const int r = (int)(w - (int64_t)t * per);
        const int i = ncb + r % nt, j = ncb + (r / nt) % nt, k = ncb + r / (nt * nt);
        const int64_t cell = (((int64_t)t * nte + k) * nte + j) * nte + i;
        const int np = rhoc[cell];
        const int64_t base = nzero[cell];
        if (np > 0 && base >= 0 && base + np <= n_used)
            for (int l = lane; l < np; l += 64) {
                const size_t ip = (size_t)(base + l) * 3;
                float nv[3];
                for (int c = 0; c < 3; c++) {
                    double v = (double)vp[ip + c];
                    if (dv64) v += dv64[ip + c];
                    nv[c] = (float)v;
                    vp[ip + c] = nv[c];
                }
                v0 = fmaxf(v0, fabsf(nv[0]));
                v1 = fmaxf(v1, fabsf(nv[1]));
                v2 = fmaxf(v2, fabsf(nv[2]));
            }
    }
    const float vv[3] = {v0, v1, v2};
    block_max_to<3>(vbits, vv);
}

// Green_3D (Green/Green.f90:1-80) in FP64; kshift = 0, n1 = n/2+1.  The
// Fortran real(4) literals and arguments are kept as float values promoted
// to double: 1./nglobal, 2./15, alpha = 4./3, 1-alpha, gridsize**2.
__global__ __launch_bounds__(256) void k_green(float* __restrict__ G, int n, float apm_soft,
                                               float apm_range, float gridsize) {
    const int nh = n / 2 + 1;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)nh * n * n) return;
    const int k1 = (int)(q % nh), k2 = (int)((q / nh) % n), k3 = (int)(q / ((int64_t)nh * n));
    const double PI = 3.14159265358979323846;
    const double base = 2.0 * PI * (double)(1.0f / (float)n);
    auto fs = [&](int t) { return base * (double)(((t + n / 2) % n) - n / 2); };
    const double kv[3] = {fs(k1), fs(k2), fs(k3)};
    const double c215 = (double)(2.0f / 15.0f);
    const float alf = 4.0f / 3.0f;
    const double alpha = (double)alf, oma = (double)(1.0f - alf);
    double D[3], u2n = 1.0;
    for (int c = 0; c < 3; c++) {
        const double s = sin(kv[c] / 2);
        u2n *= 1 - s * s + c215 * (s * s) * (s * s);
        D[c] = alpha * sin(kv[c]) + oma * sin(2 * kv[c]) / 2;
    }
    double ur[3] = {0, 0, 0};
    for (int i3 = -2; i3 <= 2; i3++)
        for (int i2 = -2; i2 <= 2; i2++)
            for (int i1 = -2; i1 <= 2; i1++) {
                const double kn[3] = {kv[0] + 2 * PI * i1, kv[1] + 2 * PI * i2, kv[2] + 2 * PI * i3};
                const double km = sqrt(kn[0] * kn[0] + kn[1] * kn[1] + kn[2] * kn[2]);
                if (km == 0) continue;  // only at k = 0, zeroed below
                double u = 1.0;
                for (int c = 0; c < 3; c++) u *= kn[c] == 0 ? 1.0 : sin(kn[c] / 2) / (kn[c] / 2);
                const double u2 = u * u, uk2 = u2 * u2 * u2;  // **(2p+2), p = 2 (TSC)
                double y = (double)apm_soft * km / 2;
                const double ss = 12 * pow(y, -4.0) * (2 - 2 * cos(y) - y * sin(y));
                double sr;
                if (apm_range == 0) sr = 0;
                else {
                    y = (double)apm_range * km / 2;
                    sr = 12 * pow(y, -4.0) * (2 - 2 * cos(y) - y * sin(y));
                }
                const double f = (ss * ss - sr * sr) / (km * km);
                for (int c = 0; c < 3; c++) ur[c] += uk2 * kn[c] * f;
            }
    const double num = D[0] * ur[0] + D[1] * ur[1] + D[2] * ur[2];
    const double den = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
    const double gs2 = (double)(gridsize * gridsize);
    double g = 4 * PI * num / den / (u2n * u2n) / gs2;
    const bool corner = (k1 == 0 || k1 == n / 2) && (k2 == 0 || k2 == n / 2) && (k3 == 0 || k3 == n / 2);
    if (corner) g = 0.0;  // Green.f90:71
    G[q] = (float)g;
}

// ------------------------------------------------------------ host side

struct LevelGroup {
    int L;
    std::vector<PM3Job> jobs;
};

static int make_groups(const Geom& g, const int* isort, const int* ires, const int* ixyz3,
                       std::vector<LevelGroup>& groups) {
    groups.clear();
    std::vector<int> seen(g.ns3, 0);
    std::map<int, size_t> idx;
    for (int o = 0; o < g.ns3; o++) {
        const int it = isort ? isort[o] - 1 : o;
        if (it < 0 || it >= g.ns3 || seen[it]) {
            set_error("isort is not a permutation of 1..ns3 (entry %d)", o);
            return CUBE2_ERR_ARG;
        }
        seen[it] = 1;
        const int* x = ixyz3 + 6 * it;
        const int L = ires[it];
        for (int q = 0; q < 3; q++)
            if (x[q] < 1 || x[q] > g.p.nns || x[q + 3] < 1 || x[q + 3] > g.p.nnt || L < 2 || L > 6) {
                set_error("invalid PM3 subtile metadata at %d (ires=%d)", it, L);
                return CUBE2_ERR_ARG;
            }
        PM3Job J;
        J.t = ((x[5] - 1) * g.p.nnt + (x[4] - 1)) * g.p.nnt + (x[3] - 1);
        for (int q = 0; q < 3; q++) J.u[q] = x[q] - 1;
        J.it = it;
        if (!idx.count(L)) {
            idx[L] = groups.size();
            groups.push_back({L, {}});
        }
        groups[idx[L]].jobs.push_back(J);
    }
    return CUBE2_OK;
}

static size_t g_budget = (size_t)16 << 30;  // bytes of mesh + force planes per batch

struct DebugOut {
    float *rho = nullptr, *phi = nullptr, *force = nullptr;
};

// Runs one batch of jobs of one level; f2bits and ptot are device scalars.
// Mesh and force-plane buffers of a batch of nb jobs at mesh size n.
// Work items of a batch at most: a group of c particles has
// ceil(c / kDepSplit) <= c / kDepSplit + 1 splits per slab, and a slot lies
// in the deposit ranges of at most 8 jobs.
static int deposit_cap(const MeshArgs& A, size_t nb) {
    const DepositShape S = deposit_shape(A);
    const int64_t cap = ((int64_t)nb * S.groups + 8 * A.n_used / kDepSplit + 1) * S.nslab;
    return (int)std::min<int64_t>(cap, (int64_t)1 << 30);
}

// Interpolation work items of a batch at most: every cell one, plus one per
// 64 particles (the shortest segment), the jobs' physical cells disjoint.
static int interp_cap(const MeshArgs& A, size_t nb) {
    const int64_t cells = (int64_t)nb * A.ntt * A.ntt * A.ntt;
    return (int)std::min<int64_t>(cells + A.n_used / 64 + 1, (int64_t)1 << 30);
}

static int batch_reserve(const MeshArgs& A, size_t nb) {
    const size_t rs = (size_t)(A.n + 2) * A.n * A.n;
    const size_t fs = (size_t)(A.nfp + 2) * (A.nfp + 2) * (A.nfp + 2);
    C2_TRY(g_pm3.rho.reserve(nb * rs * sizeof(float)));
    C2_TRY(g_pm3.force.reserve(nb * 3 * fs * sizeof(float)));
    const int64_t ng = (int64_t)nb * deposit_shape(A).groups;
    if (ng >= ((int64_t)1 << 30)) {
        set_error("deposit: %lld cell groups in one batch", (long long)ng);
        return CUBE2_ERR_CAPACITY;
    }
    const int64_t nci = (int64_t)nb * A.ntt * A.ntt * A.ntt;  // interpolation cells
    C2_TRY(g_pm3.dep_items.reserve((size_t)(ng + 1) * sizeof(int)));
    C2_TRY(g_pm3.dep_scan.reserve(scan_tmp_ints(std::max(ng, nci)) * sizeof(int)));
    C2_TRY(g_pm3.int_items.reserve((size_t)(nci + 1) * sizeof(int)));
    C2_TRY(g_pm3.int_list.reserve((size_t)interp_cap(A, nb) * sizeof(int)));
    C2_TRY(g_pm3.dep_list.reserve((size_t)deposit_cap(A, nb) * sizeof(int)));
    C2_TRY(g_pm3.dep_cnts.reserve((size_t)std::max<int64_t>(ng, 1) * sizeof(int)));
    return CUBE2_OK;
}

// d_jobs: the batch's nb jobs on the device.  Buffers must be reserved.
static int run_batch(const MeshArgs& base, const PM3Job* d_jobs, size_t nb,
                     const float* green, float mass, const float* d_scal, int accumulate,
                     double* dv64, unsigned* f2bits, unsigned long long* ptot, int* err,
                     DebugOut* dbg, hipStream_t s) {
    const bool l2 = base.layer == 2;  // stage names of the profile
    const int n = base.n, nfp = base.nfp;
    const size_t rs = (size_t)(n + 2) * n * n;
    const size_t fs = (size_t)(nfp + 2) * (nfp + 2) * (nfp + 2);
    if (base.nfb < 3) {  // the 4-point gradient reads 2 cells past the force planes
        set_error("mesh buffer of %d fine cells is below the gradient stencil's 3", base.nfb);
        return CUBE2_ERR_GEOMETRY;
    }
    C2_TRY(batch_reserve(base, nb));
    MeshArgs A = base;
    A.jobs = d_jobs;
    A.njobs = (int)nb;
    float* rho = g_pm3.rho.as<float>();
    float* force = g_pm3.force.as<float>();
    const FFTPlan* P = nullptr;
    C2_TRY(fft_plan(n, &P));
    if (use_bricks(A)) {
        if (!brick_cells_fit(A.ratio_cs * A.rsf)) {
            set_error("PM3 deposit: %d fine cells per coarse cell is below the brick minimum",
                      A.ratio_cs * A.rsf);
            return CUBE2_ERR_GEOMETRY;
        }
        const int nbx = (n + kBrick - 1) / kBrick;
        Stage st(l2 ? "pm2_deposit" : "pm3_deposit", (double)nb * rs * 4);
        dim3 grid((unsigned)(nbx * nbx * nbx), (unsigned)nb);
        k_brick_deposit<<<grid, kDepBrickThreads, 0, s>>>(A, rho, rs, nbx, err);
        C2_HIP(hipGetLastError());
    } else {
        const DepositShape S = deposit_shape(A);
        Stage st(l2 ? "pm2_deposit" : "pm3_deposit", (double)nb * rs * 4 * 2);
        C2_TRY(fill_zero(rho, nb * rs * sizeof(float), s));
        static bool attr = false;
        if (!attr) {
            C2_HIP(hipFuncSetAttribute((const void*)k_mesh_deposit,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, kDepLdsBytes));
            attr = true;
        }
        const int64_t ng = (int64_t)nb * S.groups;
        int* items = g_pm3.dep_items.as<int>();
        int* cnts = g_pm3.dep_cnts.as<int>();
        k_dep_count<<<ceil_div(ng, 256), 256, 0, s>>>(A, S.g, S.ngx, S.nslab, items, cnts);
        C2_HIP(hipGetLastError());
        C2_TRY(exclusive_scan(items, ng, g_pm3.dep_scan, s));  // items -> offsets, total at [ng]
        const int cap = deposit_cap(A, nb);
        int* list = g_pm3.dep_list.as<int>();
        k_dep_list<<<ceil_div(ng, 256), 256, 0, s>>>(items, ng, cap, list);
        C2_HIP(hipGetLastError());
        k_mesh_deposit<<<deposit_grid(S.lds), kDepThreads, S.lds, s>>>(
            A, rho, rs, S.g, S.cf, S.E, S.Zs, S.ngx, S.nslab, items, list, cap, cnts, err);
        C2_HIP(hipGetLastError());
    }
    if (dbg && dbg->rho) {
        std::vector<float> h(rs);
        C2_HIP(hipMemcpyAsync(h.data(), rho, rs * sizeof(float), hipMemcpyDeviceToHost, s));
        C2_HIP(hipStreamSynchronize(s));
        for (int z = 0; z < n; z++)
            for (int y = 0; y < n; y++)
                for (int x = 0; x < n; x++)
                    dbg->rho[((size_t)z * n + y) * n + x] = h[((size_t)z * n + y) * (n + 2) + x] * mass;
    }
    const double mesh_bytes = (double)nb * rs * 4;
    {
        Stage st(l2 ? "pm2_fft_x_r2c" : "pm3_fft_x_r2c", 2 * mesh_bytes);
        C2_TRY(fft_x_r2c(rho, *P, (int)nb, rs, d_scal, s));  // d_scal[0] = mass
    }
    {
        Stage st(l2 ? "pm2_fft_y" : "pm3_fft_y", 2 * mesh_bytes);
        C2_TRY(fft_cols(rho, *P, (int)nb, rs, 1, 0, nullptr, s));
    }
    // The gradient reads phi at storage indices nfb-3 .. nfp+nfb+2 on every
    // axis: the inverse z transform stores only those planes, the inverse y
    // and x transforms run on them and, in them, on those rows (the full
    // potential is kept when a caller inspects it).
    const int wlo = base.nfb - 3, whi = nfp + base.nfb + 3;
    const FFTWin win{wlo, whi - wlo, wlo / 2, (whi + 1) / 2 - wlo / 2};
    const FFTWin* w = (dbg && dbg->phi) ? nullptr : &win;
    {
        Stage st(l2 ? "pm2_fft_z_green" : "pm3_fft_z_green", 2 * mesh_bytes + (double)(n / 2 + 1) * n * n * 4);
        C2_TRY(fft_cols(rho, *P, (int)nb, rs, 2, 2, green, s, w));
    }
    {
        Stage st(l2 ? "pm2_fft_y_inv" : "pm3_fft_y_inv", 2 * mesh_bytes);
        C2_TRY(fft_cols(rho, *P, (int)nb, rs, 1, 1, nullptr, s, w));
    }
    {
        Stage st(l2 ? "pm2_fft_x_c2r" : "pm3_fft_x_c2r", 2 * mesh_bytes);
        C2_TRY(fft_x_c2r(rho, *P, (int)nb, rs, s, w));
    }
    if (dbg && dbg->phi) {
        std::vector<float> h(rs);
        C2_HIP(hipMemcpyAsync(h.data(), rho, rs * sizeof(float), hipMemcpyDeviceToHost, s));
        C2_HIP(hipStreamSynchronize(s));
        for (int z = 0; z < n; z++)
            for (int y = 0; y < n; y++)
                for (int x = 0; x < n; x++)
                    dbg->phi[((size_t)z * n + y) * n + x] = h[((size_t)z * n + y) * (n + 2) + x];
    }
    {
        Stage st(l2 ? "pm2_grad" : "pm3_grad", mesh_bytes + (double)nb * 3 * fs * 4);
        const int nfq = nfp + 2;
        const int ntx = (nfq + kGradW - 1) / kGradW, W = (nfq + ntx - 1) / ntx;
        const int nzc = (nfq + 63) / 64, Zc = (nfq + nzc - 1) / nzc;
        const int nty = (nfq + kGradH - 1) / kGradH;
        dim3 grid((unsigned)(ntx * nty), (unsigned)nzc, (unsigned)nb);
        k_pm3_grad<<<grid, 256, 0, s>>>(rho, rs, force, n, nfp, base.nfb, W, ntx, Zc, f2bits);
        C2_HIP(hipGetLastError());
    }
    if (dbg && dbg->force) {  // the planes [3][nf^3] from the interleaved points
        std::vector<float> h(3 * fs);
        C2_HIP(hipMemcpyAsync(h.data(), force, 3 * fs * sizeof(float), hipMemcpyDeviceToHost, s));
        C2_HIP(hipStreamSynchronize(s));
        for (size_t u = 0; u < fs; u++)
            for (int c = 0; c < 3; c++) dbg->force[c * fs + u] = h[3 * u + c];
    }
    if (dv64) {
        const int seg64 = ctx().tune[CUBE2_TUNE_INTERP_SEG] ? ctx().tune[CUBE2_TUNE_INTERP_SEG] : 4;
        const int seglen = 64 * seg64;
        const int64_t ncell = (int64_t)nb * base.ntt * base.ntt * base.ntt;
        Stage st(l2 ? "pm2_interp" : "pm3_interp", (double)nb * 3 * fs * 4);
        int* items = g_pm3.int_items.as<int>();
        int* list = g_pm3.int_list.as<int>();
        const int cap = interp_cap(base, nb);
        k_interp_count<<<ceil_div(ncell, 256), 256, 0, s>>>(A, seglen, items);
        C2_HIP(hipGetLastError());
        C2_TRY(exclusive_scan(items, ncell, g_pm3.dep_scan, s));
        k_dep_list<<<ceil_div(ncell, 256), 256, 0, s>>>(items, ncell, cap, list);
        C2_HIP(hipGetLastError());
        k_pm3_interp<<<interp_grid(), 256, 0, s>>>(A, force, dv64, accumulate, d_scal, seglen,
                                                   items, list, cap, ptot, err);
        C2_HIP(hipGetLastError());
    }
    return CUBE2_OK;
}

static MeshArgs mesh_args(const Geom& g, const int* d_rhoc, const int64_t* nz, const int16_t* d_xp,
                          int64_t n_used, int rsf) {
    MeshArgs A{};
    A.rhoc = d_rhoc;
    A.nzero = nz;
    A.xp = d_xp;
    A.nt = g.nt;
    A.ntt = g.ntt;
    A.ncb = g.ncb;
    A.nte = g.nte;
    A.nnt = g.p.nnt;
    A.ratio_cs = g.p.ratio_cs;
    A.rsf = rsf;
    A.nfp = g.ntt * g.p.ratio_cs * rsf;
    A.nfb = g.p.ratio_cs * rsf;
    A.n = A.nfp + 2 * A.nfb;
    A.nhalo = 1;  // nloop = 1 (kick.f90:108-130 through pm_thread)
    A.layer = 3;
    A.n_used = n_used;
    return A;
}

// PM2 job geometry (kick.f90:81-87): the whole parent tile (nc1..nc2 =
// 1..nt) on the standard grid (gridsize 1), buffer ngb, all ncb halo cells
// deposited (nloop = ncb).
static MeshArgs mesh_args_pm2(const Geom& g, const int* d_rhoc, const int64_t* nz,
                              const int16_t* d_xp, int64_t n_used) {
    MeshArgs A = mesh_args(g, d_rhoc, nz, d_xp, n_used, 1);
    A.ntt = g.nt;
    A.nfp = g.nt * g.p.ratio_cs;
    A.nfb = g.p.ngb;
    A.n = A.nfp + 2 * A.nfb;
    A.nhalo = g.ncb;
    A.layer = 2;
    return A;
}

int pm3_prepare(const int* isort, const int* ires, const int* ixyz3, const int* ratio_sf,
                int64_t n_used, const float* const* d_green, PM3Plan& P) {
    ensure_params();
    Context& C = ctx();
    const Geom& g = C.g;
    if (!ires || !ixyz3 || !ratio_sf || !d_green) {
        set_error("null argument");
        return CUBE2_ERR_ARG;
    }
    if (n_used < 0 || n_used >= ((int64_t)1 << 31)) {
        set_error("invalid PM3 particle count %lld", (long long)n_used);
        return CUBE2_ERR_CAPACITY;
    }
    std::vector<LevelGroup> groups;
    C2_TRY(make_groups(g, isort, ires, ixyz3, groups));
    for (auto& G : groups)
        if (!d_green[G.L - 2]) {
            set_error("no Green's table for level %d", G.L);
            return CUBE2_ERR_ARG;
        }
    C2_TRY(ensure_init());
    for (int q = 0; q < 5; q++) P.green[q] = nullptr;
    for (auto& G : groups) {  // tile-ordered copies for the z pass
        const int n = (g.ntt * g.p.ratio_cs + 2 * g.p.ratio_cs) * ratio_sf[G.L - 1];
        C2_TRY(green_tiled(d_green[G.L - 2], n, &P.green[G.L - 2], C.stream));
    }
    P.n_used = n_used;
    P.batches.clear();
    P.job_it.clear();
    std::vector<PM3Job> all;
    for (auto& G : groups) {
        const int rsf = ratio_sf[G.L - 1];
        MeshArgs A = mesh_args(g, nullptr, nullptr, nullptr, n_used, rsf);
        const size_t rs = (size_t)(A.n + 2) * A.n * A.n;
        const size_t fsz = (size_t)(A.nfp + 2) * (A.nfp + 2) * (A.nfp + 2);
        const size_t per_job = (rs + 3 * fsz) * sizeof(float);
        const size_t batch = std::max<size_t>(1, std::min(G.jobs.size(), g_budget / per_job));
        const FFTPlan* fp = nullptr;
        C2_TRY(fft_plan(A.n, &fp));
        for (size_t j0 = 0; j0 < G.jobs.size(); j0 += batch) {
            PM3Batch B;
            B.level = G.L;
            B.rsf = rsf;
            B.nb = std::min(batch, G.jobs.size() - j0);
            B.job0 = all.size();
            B.slot0 = P.job_it.size();
            for (size_t q = 0; q < B.nb; q++) {
                all.push_back(G.jobs[j0 + q]);
                P.job_it.push_back(G.jobs[j0 + q].it);
            }
            C2_TRY(batch_reserve(A, B.nb));
            P.batches.push_back(B);
        }
    }
    C2_TRY(nzero_reserve());
    C2_TRY(C.dv64.reserve((size_t)std::max<int64_t>(n_used, 1) * 3 * sizeof(double)));
    C2_TRY(g_pm3.red.reserve(4 * sizeof(unsigned)));
    C2_TRY(g_pm3.ptot.reserve(std::max<size_t>(g.ns3, 1) * sizeof(unsigned long long)));
    C2_TRY(g_pm3.jobs.reserve(std::max<size_t>(all.size(), 1) * sizeof(PM3Job)));
    C2_TRY(g_pm3.h_res.reserve(4 * sizeof(unsigned) + g.ns3 * sizeof(unsigned long long)));
    hipStream_t s = C.stream;
    if (!all.empty())
        C2_HIP(hipMemcpyAsync(g_pm3.jobs.ptr, all.data(), all.size() * sizeof(PM3Job),
                              hipMemcpyHostToDevice, s));
    C2_HIP(hipStreamSynchronize(s));
    buffers_changed();  // the device job table now describes this call
    return CUBE2_OK;
}

int pm3_enqueue(const PM3Plan& P, const int* d_rhoc, const int64_t* d_idx, const int16_t* d_xp,
                const float* d_scal, int accumulate, double* dv64, hipStream_t s) {
    const Geom& g = ctx().g;
    const int64_t* nz = nullptr;
    C2_TRY(compute_nzero(d_rhoc, d_idx, &nz, s));
    unsigned* f2bits = g_pm3.red.as<unsigned>();
    int* err = reinterpret_cast<int*>(f2bits + 1);
    unsigned long long* ptot = g_pm3.ptot.as<unsigned long long>();
    C2_TRY(fill_zero(f2bits, 4 * sizeof(unsigned), s));
    C2_TRY(fill_zero(ptot, g.ns3 * sizeof(unsigned long long), s));
    const PM3Job* d_jobs = g_pm3.jobs.as<PM3Job>();
    for (const PM3Batch& B : P.batches) {
        const MeshArgs A = B.layer == 2 ? mesh_args_pm2(g, d_rhoc, nz, d_xp, P.n_used)
                                        : mesh_args(g, d_rhoc, nz, d_xp, P.n_used, B.rsf);
        const float* green = B.layer == 2 ? P.green2 : P.green[B.level - 2];
        C2_TRY(run_batch(A, d_jobs + B.job0, B.nb, green, 1.0f, d_scal, accumulate, dv64, f2bits,
                         ptot + B.slot0, err, nullptr, s));
    }
    char* h = g_pm3.h_res.as<char>();
    C2_HIP(hipMemcpyAsync(h, f2bits, 4 * sizeof(unsigned), hipMemcpyDeviceToHost, s));
    C2_HIP(hipMemcpyAsync(h + 4 * sizeof(unsigned), ptot, g.ns3 * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, s));
    return CUBE2_OK;
}

int pm3_finish(const PM3Plan& P, float* f2max, double* ptotal) {
    const Geom& g = ctx().g;
    const char* h = g_pm3.h_res.as<const char>();
    unsigned r[4];
    std::memcpy(r, h, sizeof(r));
    const unsigned long long* hp = reinterpret_cast<const unsigned long long*>(h + sizeof(r));
    if (r[1] & 8) {
        set_error("rhoc holds more particles than n_used=%lld slots allow", (long long)P.n_used);
        return CUBE2_ERR_ARG;
    }
    if (r[1]) {
        set_error("rhoc and idx_b_r describe slots outside [0, n_used=%lld)", (long long)P.n_used);
        return CUBE2_ERR_ARG;
    }
    *f2max = __builtin_bit_cast(float, r[0]);
    const int nout = P.nslots_out ? P.nslots_out : g.ns3;
    for (int it = 0; it < nout; it++) ptotal[it] = 0.0;
    for (size_t q = 0; q < P.job_it.size(); q++) ptotal[P.job_it[q]] = (double)hp[q];
    return CUBE2_OK;
}

int pm3_force(const int* isort, const int* ires, const int* ixyz3, const int* ratio_sf,
              const int* d_rhoc, const int64_t* d_idx, const int16_t* d_xp, int64_t n_used,
              const float* const* d_green, float mass, float amid, float dt, float* f2max,
              double* ptotal) {
    ensure_params();
    Context& C = ctx();
    if (!f2max || !ptotal) {
        set_error("null argument");
        return CUBE2_ERR_ARG;
    }
    if (C.dv_state == 1 && C.dv_n_used != n_used) {
        set_error("PM3 particle count changed while a Delta-v is pending");
        return CUBE2_ERR_STATE;
    }
    PM3Plan P;
    C2_TRY(pm3_prepare(isort, ires, ixyz3, ratio_sf, n_used, d_green, P));
    const int accumulate = C.dv_state == 1 ? 1 : 0;
    const float* d_scal = nullptr;
    C2_TRY(scalars_upload(mass, amid, dt, C.stream, &d_scal));
    C2_TRY(pm3_enqueue(P, d_rhoc, d_idx, d_xp, d_scal, accumulate, C.dv64.as<double>(), C.stream));
    C2_TRY(sync_and_flush());
    C2_TRY(pm3_finish(P, f2max, ptotal));
    C.dv_state = 1;
    C.dv_n_used = n_used;
    return CUBE2_OK;
}

int pm2_prepare(int64_t n_used, const float* d_green2, PM3Plan& P) {
    ensure_params();
    Context& C = ctx();
    const Geom& g = C.g;
    if (!d_green2) {
        set_error("no Green's table for PM2");
        return CUBE2_ERR_ARG;
    }
    if (n_used < 0 || n_used >= ((int64_t)1 << 31)) {
        set_error("invalid PM2 particle count %lld", (long long)n_used);
        return CUBE2_ERR_CAPACITY;
    }
    C2_TRY(ensure_init());
    const MeshArgs A = mesh_args_pm2(g, nullptr, nullptr, nullptr, n_used);
    const FFTPlan* fp = nullptr;
    C2_TRY(fft_plan(A.n, &fp));
    C2_TRY(green_tiled(d_green2, A.n, &P.green2, C.stream));
    P.n_used = n_used;
    P.nslots_out = g.ntile;
    P.batches.clear();
    P.job_it.clear();
    std::vector<PM3Job> all(g.ntile);
    for (int t = 0; t < g.ntile; t++) {
        all[t].t = t;
        all[t].u[0] = all[t].u[1] = all[t].u[2] = 0;
        all[t].it = t;
        P.job_it.push_back(t);
    }
    const size_t rs = (size_t)(A.n + 2) * A.n * A.n;
    const size_t fsz = (size_t)(A.nfp + 2) * (A.nfp + 2) * (A.nfp + 2);
    const size_t batch = std::max<size_t>(1, std::min<size_t>(g.ntile, g_budget / ((rs + 3 * fsz) * 4)));
    for (size_t j0 = 0; j0 < (size_t)g.ntile; j0 += batch) {
        PM3Batch B;
        B.layer = 2;
        B.rsf = 1;
        B.nb = std::min(batch, (size_t)g.ntile - j0);
        B.job0 = j0;
        B.slot0 = j0;
        C2_TRY(batch_reserve(A, B.nb));
        P.batches.push_back(B);
    }
    C2_TRY(nzero_reserve());
    C2_TRY(C.dv64.reserve((size_t)std::max<int64_t>(n_used, 1) * 3 * sizeof(double)));
    C2_TRY(g_pm3.red.reserve(4 * sizeof(unsigned)));
    C2_TRY(g_pm3.ptot.reserve(std::max<size_t>(g.ns3, 1) * sizeof(unsigned long long)));
    C2_TRY(g_pm3.jobs.reserve(all.size() * sizeof(PM3Job)));
    C2_TRY(g_pm3.h_res.reserve(4 * sizeof(unsigned) + g.ns3 * sizeof(unsigned long long)));
    C2_HIP(hipMemcpyAsync(g_pm3.jobs.ptr, all.data(), all.size() * sizeof(PM3Job),
                          hipMemcpyHostToDevice, C.stream));
    C2_HIP(hipStreamSynchronize(C.stream));
    buffers_changed();
    return CUBE2_OK;
}

int pm2_force(const int* d_rhoc, const int64_t* d_idx, const int16_t* d_xp, int64_t n_used,
              const float* d_green2, float mass, float amid, float dt, float* f2max,
              double* ptotal) {
    ensure_params();
    Context& C = ctx();
    if (!f2max || !ptotal) {
        set_error("null argument");
        return CUBE2_ERR_ARG;
    }
    if (C.dv_state == 1 && C.dv_n_used != n_used) {
        set_error("PM2 particle count changed while a Delta-v is pending");
        return CUBE2_ERR_STATE;
    }
    PM3Plan P;
    C2_TRY(pm2_prepare(n_used, d_green2, P));
    const int accumulate = C.dv_state == 1 ? 1 : 0;
    const float* d_scal = nullptr;
    C2_TRY(scalars_upload(mass, amid, dt, C.stream, &d_scal));
    C2_TRY(pm3_enqueue(P, d_rhoc, d_idx, d_xp, d_scal, accumulate, C.dv64.as<double>(), C.stream));
    C2_TRY(sync_and_flush());
    C2_TRY(pm3_finish(P, f2max, ptotal));
    C.dv_state = 1;
    C.dv_n_used = n_used;
    return CUBE2_OK;
}

int pm3_debug_job(int it, const int* ires, const int* ixyz3, const int* ratio_sf, const int* d_rhoc,
                  const int64_t* d_idx, const int16_t* d_xp, int64_t n_used, const float* d_green,
                  float mass, float* rho, float* phi, float* force, float* f2max) {
    ensure_params();
    Context& C = ctx();
    const Geom& g = C.g;
    if (it < 0 || it >= g.ns3 || !d_green) {
        set_error("bad subtile %d or no Green's table", it);
        return CUBE2_ERR_ARG;
    }
    std::vector<LevelGroup> groups;
    C2_TRY(make_groups(g, nullptr, ires, ixyz3, groups));
    C2_TRY(ensure_init());
    hipStream_t s = C.stream;
    PM3Job J{};
    for (auto& G : groups)
        for (auto& j : G.jobs)
            if (j.it == it) J = j;
    const int L = ires[it];
    MeshArgs A = mesh_args(g, d_rhoc, nullptr, d_xp, n_used, ratio_sf[L - 1]);
    const int64_t* nz = nullptr;
    C2_TRY(compute_nzero(d_rhoc, d_idx, &nz, s));
    A.nzero = nz;
    C2_TRY(g_pm3.red.reserve(4 * sizeof(unsigned)));
    unsigned* f2bits = g_pm3.red.as<unsigned>();
    C2_TRY(fill_zero(f2bits, 4 * sizeof(unsigned), s));
    C2_TRY(g_pm3.jobs.reserve(sizeof(PM3Job)));
    C2_HIP(hipMemcpyAsync(g_pm3.jobs.ptr, &J, sizeof(PM3Job), hipMemcpyHostToDevice, s));
    buffers_changed();
    DebugOut dbg{rho, phi, force};
    const float* gt = nullptr;
    C2_TRY(green_tiled(d_green, A.n, &gt, s));
    const float* d_scal = nullptr;
    C2_TRY(scalars_upload(mass, 0.0f, 0.0f, s, &d_scal));
    C2_TRY(run_batch(A, g_pm3.jobs.as<PM3Job>(), 1, gt, mass, d_scal, 0, nullptr, f2bits,
                     nullptr, reinterpret_cast<int*>(f2bits + 1), &dbg, s));
    unsigned h[4];
    C2_HIP(hipMemcpyAsync(h, f2bits, 4 * sizeof(unsigned), hipMemcpyDeviceToHost, s));
    C2_TRY(sync_and_flush());
    if (f2max) *f2max = __builtin_bit_cast(float, h[0]);
    return CUBE2_OK;
}

int velocity_apply(const int* d_rhoc, const int64_t* d_idx, float* d_vp, int64_t n_used,
                   float* vmax) {
    C2_TRY(ensure_init());
    Context& C = ctx();
    const Geom& g = C.g;
    hipStream_t s = C.stream;
    if (C.dv_state == 1 && C.dv_n_used != n_used) {
        set_error("particle count changed while a Delta-v is pending");
        return CUBE2_ERR_STATE;
    }
    const int64_t* nz = nullptr;
    C2_TRY(compute_nzero(d_rhoc, d_idx, &nz, s));
    C2_TRY(g_pm3.red.reserve(4 * sizeof(unsigned)));
    unsigned* vb = g_pm3.red.as<unsigned>();
    C2_TRY(fill_zero(vb, 4 * sizeof(unsigned), s));
    const int64_t waves = (int64_t)g.ntile * g.nt * g.nt * g.nt;
    k_velocity_apply<<<ceil_div(waves, kWaves), 256, 0, s>>>(
        d_rhoc, nz, g.nt, g.ncb, g.nte, g.ntile, n_used,
        C.dv_state == 1 ? C.dv64.as<double>() : nullptr, d_vp, vb);
    C2_HIP(hipGetLastError());
    unsigned h[4];
    C2_HIP(hipMemcpyAsync(h, vb, 4 * sizeof(unsigned), hipMemcpyDeviceToHost, s));
    C2_TRY(sync_and_flush());
    for (int d = 0; d < 3; d++) vmax[d] = __builtin_bit_cast(float, h[d]);
    C.dv_state = 0;
    return CUBE2_OK;
}

int green_3d(float* d_G, int n, float apm_soft, float apm_range, float gridsize) {
    if (!d_G || n < 2 || n % 2) {
        set_error("bad Green's table request (n=%d)", n);
        return CUBE2_ERR_ARG;
    }
    C2_TRY(ensure_init());
    hipStream_t s = ctx().stream;
    const int64_t tot = (int64_t)(n / 2 + 1) * n * n;
    {
        Stage st("green_3d", (double)tot * 4);
        k_green<<<ceil_div(tot, 256), 256, 0, s>>>(d_G, n, apm_soft, apm_range, gridsize);
        C2_HIP(hipGetLastError());
    }
    return sync_and_flush();
}

int fft3d(float* d, int n, int nb, int dir) {
    C2_TRY(ensure_init());
    C2_TRY(fft3d_batch(d, n, nb, dir, ctx().stream));
    return sync_and_flush();
}

int green_cache_get(const float* host, size_t count, const float** dev) {
    for (auto& G : g_greens)
        if (G.host == host && G.count == count) {
            *dev = G.dev;
            return CUBE2_OK;
        }
    Green G{host, count, nullptr};
    C2_HIP(hipMalloc(&G.dev, count * sizeof(float)));
    C2_HIP(hipMemcpy(G.dev, host, count * sizeof(float), hipMemcpyHostToDevice));
    g_greens.push_back(G);
    *dev = G.dev;
    return CUBE2_OK;
}

void green_cache_release() {
    buffers_changed();
    for (auto& G : g_greens) (void)hipFree(G.dev);
    g_greens.clear();
}

void pm3_release() {
    g_pm3.rho.release();
    g_pm3.force.release();
    g_pm3.jobs.release();
    g_pm3.red.release();
    g_pm3.ptot.release();
    g_pm3.dep_items.release();
    g_pm3.dep_scan.release();
    g_pm3.dep_list.release();
    g_pm3.dep_cnts.release();
    g_pm3.int_items.release();
    g_pm3.int_list.release();
    g_pm3.h_res.release();
    fft_release();
    green_tiled_release();
}

}  // namespace cube2
