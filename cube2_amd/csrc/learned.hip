// Learned stand-in for PP in sparse cells, inside the kick (no counterpart in
// the reference).
//
// A 3-D CNN of four valid 3^3 convolutions, 4 -> 32 -> 32 -> 32 -> 3
// channels (ReLU after the first three), maps the fine-cell stencil around a
// fine cell of the level-2 mesh (0.5 standard cells) to a PP acceleration at
// that cell; a particle's PP acceleration is the cloud-in-cell interpolation
// of the cell values at its position.  Input channels of a fine cell:
// log(1 + n) of its particle count n and the mean offset of its particles
// from the cell centre (fine cells, per axis; 0 when empty).
//
// In a kick with CUBE2_KICK_LEARNED, the pair kernel skips the chunks of PP
// cells of fewer than `occupancy` particles in level-2 subtiles; for every
// level-2 subtile ("job") this file builds the input grid of the subtile's
// fine cells plus a margin of 5 (4 for the 9^3 receptive field, 1 for the
// interpolation) from PP's sorted particles, runs the four layers over the
// grid (cnn_conv3, activations BF16 in between), and writes each skipped
// particle's velocity term (the pair kernel's a m^2 a_mid dt / 6) from the
// interpolated prediction, so that the split pass's velocity update applies
// it with the rest.
#include <algorithm>
#include <cmath>

#include "common.h"
#include "engine.h"
#include "pp_kernels.h"
#include "profile.h"
#include "reduce.h"

namespace cube2 {

constexpr int kLMargin = 5;
constexpr int kLChannels[5] = {4, 32, 32, 32, 3};
constexpr int kLJobBatch = 64;  // subtiles per pass of the layers (buffer size)

struct LearnedModel {
    bool set = false;
    DevBuf wp[4], bias[4];
    float y_scale = 1.0f;
    int occupancy = 64;
};
static LearnedModel g_lm;

struct LearnedBufs {
    DevBuf x, a1, a2, a3, out, jobs, job_of;
    int last_j0 = 0, last_nj = 0, last_DI = 0, last_PO = 0;  // last pass of the layers
};
static LearnedBufs g_lb;

__device__ inline int find_tile_l(const PPTile* __restrict__ T, int ntile, uint32_t c) {
    int lo = 0, hi = ntile - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (T[mid].cell_off <= (long long)c) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Fine-cell input of the jobs [j0, j0 + nj): count and offset sums, from
// every sorted particle (physical and halo copies) of each job's tile that
// falls inside its grid of DI^3 cells (fine cell o - kLMargin + index).
__global__ __launch_bounds__(256) void k_learned_deposit(const float4* __restrict__ pos4,
                                                         const int* __restrict__ tile_first,
                                                         const LearnedJob* __restrict__ jobs,
                                                         int j0, int DI, float rsf2,
                                                         float* __restrict__ X) {
    const int j = blockIdx.y;
    const LearnedJob J = jobs[j0 + j];
    const int a0 = tile_first[J.t], a1 = tile_first[J.t + 1];
    float* Xj = X + (size_t)j * DI * DI * DI * 4;
    for (int a = a0 + blockIdx.x * blockDim.x + threadIdx.x; a < a1; a += gridDim.x * blockDim.x) {
        const float4 p = pos4[a];
        const float f[3] = {rsf2 * p.x, rsf2 * p.y, rsf2 * p.z};
        int c[3];
        bool in = true;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const float g = floorf(f[d]);
            c[d] = (int)g - (J.o[d] - kLMargin);
            in &= (unsigned)c[d] < (unsigned)DI;
        }
        if (!in) continue;
        float* x = Xj + ((size_t)(c[2] * DI + c[1]) * DI + c[0]) * 4;
        atomicAdd(x, 1.0f);
#pragma unroll
        for (int d = 0; d < 3; d++) atomicAdd(x + 1 + d, f[d] - floorf(f[d]) - 0.5f);
    }
}

// count -> log(1 + count), offset sums -> means.
__global__ void k_learned_finish(float* __restrict__ X, int64_t ncell) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    float4 v = reinterpret_cast<float4*>(X)[c];
    if (v.x > 0.0f) {
        const float r = 1.0f / v.x;
        v.y *= r;
        v.z *= r;
        v.w *= r;
        v.x = log1pf(v.x);
    }
    reinterpret_cast<float4*>(X)[c] = v;
}

// Whether sorted particle a is one the network stands in for, and its job:
// a physical particle of a level-2 subtile (its range is the level-2 range)
// in a PP cell of fewer than occ particles.
__device__ inline int learned_job_of(const PPArgs& A, const LearnedArgs& L, int a, float4& p) {
    p = A.pos4[a];
    if (fabsf(p.w) != L.range) return -1;  // (w < 0: pairs-once small cell)
    const uint32_t cid = A.cellid[a];
    if (A.cell_start[cid + 1] - A.cell_start[cid] >= L.occupancy) return -1;
    const int t = find_tile_l(A.tiles, A.ntile, cid);
    const int sx = min((int)floorf(p.x / L.sub_std), L.nns - 1);
    const int sy = min((int)floorf(p.y / L.sub_std), L.nns - 1);
    const int sz = min((int)floorf(p.z / L.sub_std), L.nns - 1);
    return L.job_of[(size_t)t * L.nns * L.nns * L.nns + (sz * L.nns + sy) * L.nns + sx];
}

// CIC lower corner (output-grid coordinates) and fractions of position p in
// job J's output grid (fine cell o - 1 + index, values at cell centres).
__device__ inline void learned_cic(const LearnedJob& J, float rsf2, float4 p, int q[3], float w[3]) {
    const float f[3] = {rsf2 * p.x, rsf2 * p.y, rsf2 * p.z};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const float u = f[d] - (float)(J.o[d] - 1) - 0.5f;
        const float fl = floorf(u);
        q[d] = (int)fl;
        w[d] = u - fl;
    }
}

// The velocity terms of the substituted particles whose job is in
// [j0, j0 + nj): prediction interpolated from the job's output grid
// (PO^3 x 3), scaled, times mass^2 as the pair kernel's sums; f2max.
__global__ __launch_bounds__(256) void k_learned_apply(PPArgs A, LearnedArgs L, int j0, int nj,
                                                       int PO, const float* __restrict__ O) {
    float f2 = 0.0f;
    const bool ok = A.red->err == 0;
    const int n = ok ? min(*A.n_sorted, A.cap) : 0;
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < n; a += gridDim.x * blockDim.x) {
        float4 p;
        const int j = learned_job_of(A, L, a, p);
        if (j < j0 || j >= j0 + nj) continue;
        const LearnedJob J = L.jobs[j];
        int q[3];
        float w[3];
        learned_cic(J, L.rsf2, p, q, w);
        float acc[3] = {0.f, 0.f, 0.f};
        const float* Oj = O + (size_t)(j - j0) * PO * PO * PO * 3;
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
            const int x = min(max(q[0] + dx, 0), PO - 1), y = min(max(q[1] + dy, 0), PO - 1),
                      z = min(max(q[2] + dz, 0), PO - 1);
            const float wt = (dx ? w[0] : 1.f - w[0]) * (dy ? w[1] : 1.f - w[1]) * (dz ? w[2] : 1.f - w[2]);
            const float* o = Oj + ((size_t)(z * PO + y) * PO + x) * 3;
            acc[0] += wt * o[0];
            acc[1] += wt * o[1];
            acc[2] += wt * o[2];
        }
        const float mass = A.scal[0], a_mid = A.scal[1], dt = A.scal[2];
        float am[3];
#pragma unroll
        for (int d = 0; d < 3; d++) am[d] = acc[d] * L.y_scale * mass * mass;
        f2 = fmaxf(f2, am[0] * am[0] + am[1] * am[1] + am[2] * am[2]);
#pragma unroll
        for (int d = 0; d < 3; d++) A.term[3 * (size_t)A.slot[a] + d] = am[d] * a_mid * dt / 6.0f;
    }
    const float mx[1] = {f2};
    block_max_to<1>(&A.red->f2max_bits, mx);
}

// Training samples: for every substitutable particle, its job, the CIC lower
// corner in the job's output grid (= the origin of its 10^3 input patch in
// the job's input grid), the CIC fractions and its slot.
__global__ __launch_bounds__(256) void k_learned_samples(PPArgs A, LearnedArgs L, int cap,
                                                         int* __restrict__ count,
                                                         int* __restrict__ job,
                                                         int* __restrict__ base,
                                                         float* __restrict__ frac,
                                                         uint32_t* __restrict__ slot) {
    const int n = min(*A.n_sorted, A.cap);
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < n; a += gridDim.x * blockDim.x) {
        float4 p;
        const int j = learned_job_of(A, L, a, p);
        if (j < 0) continue;
        const int k = atomicAdd(count, 1);
        if (k >= cap) continue;
        int q[3];
        float w[3];
        learned_cic(L.jobs[j], L.rsf2, p, q, w);
        job[k] = j;
        for (int d = 0; d < 3; d++) {
            base[3 * k + d] = q[d];
            frac[3 * k + d] = w[d];
        }
        slot[k] = A.slot[a];
    }
}

// out[i][z][y][x][c] = X[job[i]][base+z][base+y][base+x][c] for the 10^3 x 4
// patch of each sample (zeros outside the grid).
__global__ void k_learned_gather(const float* __restrict__ X, int DI, const int* __restrict__ job,
                                 const int* __restrict__ base, int n, float* __restrict__ out) {
    constexpr int PT = 10;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * PT * PT * PT) return;
    const int i = (int)(e / (PT * PT * PT)), r = (int)(e % (PT * PT * PT));
    const int x = base[3 * i] + r % PT, y = base[3 * i + 1] + (r / PT) % PT, z = base[3 * i + 2] + r / (PT * PT);
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if ((unsigned)x < (unsigned)DI && (unsigned)y < (unsigned)DI && (unsigned)z < (unsigned)DI)
        v = reinterpret_cast<const float4*>(X)[(((size_t)job[i] * DI + z) * DI + y) * DI + x];
    reinterpret_cast<float4*>(out)[e] = v;
}

// ------------------------------------------------------------------ host
bool learned_model_set() { return g_lm.set; }
int learned_model_occupancy() { return g_lm.occupancy; }

int learned_model_load(const float* const* W, const float* const* b, float y_scale, int occupancy) {
    if (!W || !b || !(y_scale > 0.0f) || occupancy < 1) {
        set_error("learned model: null weights, y_scale <= 0 or occupancy < 1");
        return CUBE2_ERR_ARG;
    }
    C2_TRY(ensure_init());
    hipStream_t s = ctx().stream;
    DevBuf tmp;
    for (int l = 0; l < 4; l++) {
        const int cin = kLChannels[l], cout = kLChannels[l + 1];
        if (!W[l] || !b[l]) {
            set_error("learned model: null layer %d", l);
            return CUBE2_ERR_ARG;
        }
        const size_t nw = (size_t)27 * cin * cout;
        C2_TRY(tmp.reserve(nw * sizeof(float)));
        C2_TRY(g_lm.wp[l].reserve(cnn_conv_packed_count(cin, cout) * sizeof(uint16_t)));
        C2_TRY(g_lm.bias[l].reserve(cout * sizeof(float)));
        C2_HIP(hipMemcpyAsync(tmp.ptr, W[l], nw * sizeof(float), hipMemcpyHostToDevice, s));
        C2_HIP(hipMemcpyAsync(g_lm.bias[l].ptr, b[l], cout * sizeof(float), hipMemcpyHostToDevice, s));
        C2_TRY(cnn_conv_pack(tmp.as<float>(), g_lm.wp[l].as<uint16_t>(), cin, cout, s));
        C2_HIP(hipStreamSynchronize(s));  // tmp is reused by the next layer
    }
    tmp.release();
    g_lm.y_scale = y_scale;
    g_lm.occupancy = occupancy;
    g_lm.set = true;
    buffers_changed();  // a recorded kick reads the model's buffers
    return CUBE2_OK;
}

void learned_release() {
    for (int l = 0; l < 4; l++) {
        g_lm.wp[l].release();
        g_lm.bias[l].release();
    }
    g_lm.set = false;
    for (DevBuf* b : {&g_lb.x, &g_lb.a1, &g_lb.a2, &g_lb.a3, &g_lb.out, &g_lb.jobs, &g_lb.job_of})
        b->release();
}

// Jobs of the level-2 subtiles (location order) and the location -> job map;
// the geometry of the grids.
int learned_prepare(const Geom& g, const std::vector<int>& iloc, const int* ratio_sf,
                    const float* apm3, LearnedPlan& Lp, int occupancy) {
    Lp.jobs.clear();
    const int nns = g.p.nns, nsub3 = nns * nns * nns;
    std::vector<int> job_of(g.ns3, -1);
    const int rsf2 = ratio_sf[1];
    Lp.nfs = g.ntt * g.p.ratio_cs * rsf2;  // fine cells per subtile side at level 2
    for (int t = 0; t < g.ntile; t++)
        for (int q = 0; q < nsub3; q++) {
            if (iloc[(size_t)t * nsub3 + q] != 2) continue;
            const int sx = q % nns, sy = (q / nns) % nns, sz = q / (nns * nns);
            LearnedJob J;
            J.t = t;
            J.o[0] = sx * Lp.nfs;
            J.o[1] = sy * Lp.nfs;
            J.o[2] = sz * Lp.nfs;
            job_of[(size_t)t * nsub3 + q] = (int)Lp.jobs.size();
            Lp.jobs.push_back(J);
        }
    Lp.DI = Lp.nfs + 2 * kLMargin;
    Lp.PO = Lp.DI - 8;
    Lp.args.range = apm3[1];
    Lp.args.occupancy = occupancy;
    Lp.args.rsf2 = (float)rsf2;
    Lp.args.sub_std = (float)(g.ntt * g.p.ratio_cs);
    Lp.args.nns = nns;
    Lp.args.y_scale = g_lm.y_scale;
    const size_t nj = std::max<size_t>(Lp.jobs.size(), 1);
    C2_TRY(g_lb.jobs.reserve(nj * sizeof(LearnedJob)));
    C2_TRY(g_lb.job_of.reserve(std::max(g.ns3, 1) * sizeof(int)));
    hipStream_t s = ctx().stream;
    if (!Lp.jobs.empty())
        C2_HIP(hipMemcpyAsync(g_lb.jobs.ptr, Lp.jobs.data(), Lp.jobs.size() * sizeof(LearnedJob),
                              hipMemcpyHostToDevice, s));
    C2_HIP(hipMemcpyAsync(g_lb.job_of.ptr, job_of.data(), g.ns3 * sizeof(int), hipMemcpyHostToDevice, s));
    C2_HIP(hipStreamSynchronize(s));  // host tables go out of scope
    Lp.args.jobs = g_lb.jobs.as<LearnedJob>();
    Lp.args.job_of = g_lb.job_of.as<int>();
    Lp.batch = (int)std::max<size_t>(1, std::min<size_t>(Lp.jobs.size(), kLJobBatch));
    buffers_changed();  // the device job tables now describe this plan
    return CUBE2_OK;
}

// Buffers of the layers for batches of Lp.batch jobs (in the kick) or for
// the input grids of all jobs (training).
int learned_reserve(const LearnedPlan& Lp) {
    const size_t nb = std::max(Lp.batch, 1);
    auto cells = [](int d) { return (size_t)d * d * d; };
    C2_TRY(g_lb.x.reserve(nb * cells(Lp.DI) * 4 * sizeof(float)));
    C2_TRY(g_lb.a1.reserve(nb * cells(Lp.DI - 2) * 32 * sizeof(uint16_t)));
    C2_TRY(g_lb.a2.reserve(nb * cells(Lp.DI - 4) * 32 * sizeof(uint16_t)));
    C2_TRY(g_lb.a3.reserve(nb * cells(Lp.DI - 6) * 32 * sizeof(uint16_t)));
    C2_TRY(g_lb.out.reserve(nb * cells(Lp.PO) * 3 * sizeof(float)));
    return CUBE2_OK;
}

static int deposit_inputs(const LearnedPlan& Lp, const PPArgs& A, const int* tile_first, int j0,
                          int nj, float* X, hipStream_t s) {
    const size_t ncell = (size_t)nj * Lp.DI * Lp.DI * Lp.DI;
    C2_TRY(fill_zero(X, ncell * 4 * sizeof(float), s));
    dim3 grid(64, (unsigned)nj);
    k_learned_deposit<<<grid, 256, 0, s>>>(A.pos4, tile_first, Lp.args.jobs, j0, Lp.DI,
                                           Lp.args.rsf2, X);
    C2_HIP(hipGetLastError());
    k_learned_finish<<<ceil_div(ncell, 256), 256, 0, s>>>(X, (int64_t)ncell);
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

// In the kick, after the pair kernel (which skipped the same particles):
// per batch of jobs the input grids, the four layers and the velocity terms.
int learned_enqueue(const LearnedPlan& Lp, const PPArgs& A, const int* tile_first, hipStream_t s) {
    if (!g_lm.set) {
        set_error("learned kick without a model (cube2_learned_set)");
        return CUBE2_ERR_STATE;
    }
    const int njobs = (int)Lp.jobs.size();
    LearnedArgs L = Lp.args;
    for (int j0 = 0; j0 < njobs; j0 += Lp.batch) {
        const int nj = std::min(Lp.batch, njobs - j0);
        const int D = Lp.DI;
        const double cells1 = (double)(D - 2) * (D - 2) * (D - 2), cells2 = (double)(D - 4) * (D - 4) * (D - 4),
                     cells3 = (double)(D - 6) * (D - 6) * (D - 6), cells4 = (double)Lp.PO * Lp.PO * Lp.PO;
        {
            Stage st("learned_input", (double)nj * D * D * D * 16);
            C2_TRY(deposit_inputs(Lp, A, tile_first, j0, nj, g_lb.x.as<float>(), s));
        }
        {
            Stage st("learned_layers", 0.0,
                     2.0 * 27 * nj * (cells1 * 4 * 32 + cells2 * 32 * 32 + cells3 * 32 * 32 + cells4 * 32 * 3));
            C2_TRY(cnn_conv3(g_lb.x.ptr, 0, g_lm.wp[0].as<uint16_t>(), g_lm.bias[0].as<float>(),
                             g_lb.a1.ptr, 1, nj, D, 4, 32, 1, s));
            C2_TRY(cnn_conv3(g_lb.a1.ptr, 1, g_lm.wp[1].as<uint16_t>(), g_lm.bias[1].as<float>(),
                             g_lb.a2.ptr, 1, nj, D - 2, 32, 32, 1, s));
            C2_TRY(cnn_conv3(g_lb.a2.ptr, 1, g_lm.wp[2].as<uint16_t>(), g_lm.bias[2].as<float>(),
                             g_lb.a3.ptr, 1, nj, D - 4, 32, 32, 1, s));
            C2_TRY(cnn_conv3(g_lb.a3.ptr, 1, g_lm.wp[3].as<uint16_t>(), g_lm.bias[3].as<float>(),
                             g_lb.out.ptr, 0, nj, D - 6, 32, 3, 0, s));
        }
        g_lb.last_j0 = j0;
        g_lb.last_nj = nj;
        g_lb.last_DI = Lp.DI;
        g_lb.last_PO = Lp.PO;
        {
            Stage st("learned_apply", (double)A.cap * 24);
            const unsigned blocks = std::min<unsigned>(ceil_div(A.cap, 256), 4096);
            k_learned_apply<<<blocks, 256, 0, s>>>(A, L, j0, nj, Lp.PO, g_lb.out.as<float>());
            C2_HIP(hipGetLastError());
        }
    }
    return CUBE2_OK;
}

// Training support, after a PP call with the cell lists of PP in place: the
// input grids of every level-2 subtile into X (caller's, [njobs][DI^3][4])
// and the substitutable particles as samples.
int learned_training_inputs(const LearnedPlan& Lp, const PPArgs& A, const int* tile_first, float* X,
                            hipStream_t s) {
    const int njobs = (int)Lp.jobs.size();
    for (int j0 = 0; j0 < njobs; j0 += 64) {
        const int nj = std::min(64, njobs - j0);
        C2_TRY(deposit_inputs(Lp, A, tile_first, j0, nj,
                              X + (size_t)j0 * Lp.DI * Lp.DI * Lp.DI * 4, s));
    }
    return CUBE2_OK;
}

int learned_training_samples(const LearnedPlan& Lp, const PPArgs& A, int cap, int* d_count,
                             int* job, int* base, float* frac, uint32_t* slot, hipStream_t s) {
    C2_TRY(fill_zero(d_count, sizeof(int), s));
    const unsigned blocks = std::min<unsigned>(ceil_div(A.cap, 256), 4096);
    k_learned_samples<<<blocks, 256, 0, s>>>(A, Lp.args, cap, d_count, job, base, frac, slot);
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

// The input grids and outputs of the last pass of the layers in a kick
// (for checks): jobs [j0, j0 + nj), X [nj][DI^3][4], O [nj][PO^3][3].
int learned_debug_last(float* X, float* O, int cap_jobs, int* j0, int* nj, int* DI, int* PO,
                       hipStream_t s) {
    *j0 = g_lb.last_j0;
    *nj = g_lb.last_nj;
    *DI = g_lb.last_DI;
    *PO = g_lb.last_PO;
    if (!X && !O) return CUBE2_OK;
    if (cap_jobs < g_lb.last_nj) {
        set_error("learned debug: %d jobs, room for %d", g_lb.last_nj, cap_jobs);
        return CUBE2_ERR_CAPACITY;
    }
    const size_t nx = (size_t)g_lb.last_nj * g_lb.last_DI * g_lb.last_DI * g_lb.last_DI * 4;
    const size_t no = (size_t)g_lb.last_nj * g_lb.last_PO * g_lb.last_PO * g_lb.last_PO * 3;
    if (X) C2_HIP(hipMemcpyAsync(X, g_lb.x.ptr, nx * sizeof(float), hipMemcpyDeviceToDevice, s));
    if (O) C2_HIP(hipMemcpyAsync(O, g_lb.out.ptr, no * sizeof(float), hipMemcpyDeviceToDevice, s));
    C2_HIP(hipStreamSynchronize(s));
    return CUBE2_OK;
}

int learned_gather(const float* X, int DI, const int* job, const int* base, int n, float* out,
                   hipStream_t s) {
    if (n <= 0) return CUBE2_OK;
    const int64_t tot = (int64_t)n * 1000;
    k_learned_gather<<<ceil_div(tot, 256), 256, 0, s>>>(X, DI, job, base, n, out);
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

}  // namespace cube2
