// Short-range kick driver (kick.f90:107-238) over the device-resident
// arrays: level selection from the per-subtile maximum coarse-cell count,
// job order, PM3, PP (which applies PM3's pending Delta-v), and the FP64
// time-step limits.
#include <algorithm>
#include <chrono>
#include <cmath>

#include "common.h"
#include "engine.h"
#include "profile.h"

namespace cube2 {

// Job order of kick.f90:131: indexed heapsort of -npc_max (kick.f90:570-614,
// Numerical Recipes' indexx), 1-based indices, not stable.
void heap_index(int n, const float* arr, int* indx) {
    for (int j = 1; j <= n; j++) indx[j - 1] = j;
    if (n < 2) return;
    auto A = [&](int q) { return arr[q - 1]; };
    auto IX = [&](int q) -> int& { return indx[q - 1]; };
    int l = n / 2 + 1, ir = n;
    for (;;) {
        int it;
        float qv;
        if (l > 1) {
            it = IX(--l);
            qv = A(it);
        } else {
            it = IX(ir);
            qv = A(it);
            IX(ir) = IX(1);
            if (--ir == 1) {
                IX(1) = it;
                return;
            }
        }
        int i = l, j = l + l;
        while (j <= ir) {
            if (j < ir && A(IX(j)) < A(IX(j + 1))) j++;
            if (qv < A(IX(j))) {
                IX(i) = IX(j);
                i = j;
                j += j;
            } else {
                j = ir + 1;
            }
        }
        IX(i) = it;
    }
}

// Level of kick.f90:114-126 from the largest physical coarse-cell count
// of a subtile; npc_max = maxval/64. is exact in float, so the thresholds
// compare exactly.
int level_of(int maxcount) {
    const float p = (float)maxcount / 64.0f;
    return p < 24.0f ? 2 : p < 114.0f ? 3 : p < 320.0f ? 4 : p < 1350.0f ? 5 : 6;
}

int check_ixyz3(const Geom& g, const int* ixyz3) {
    for (int it = 0; it < g.ns3; it++) {
        const int* x = ixyz3 + 6 * it;
        for (int q = 0; q < 3; q++)
            if (x[q] < 1 || x[q] > g.p.nns || x[q + 3] < 1 || x[q + 3] > g.p.nnt) {
                set_error("invalid ixyz3 row %d", it);
                return CUBE2_ERR_ARG;
            }
    }
    return CUBE2_OK;
}

// One block per subtile: maximum of rhoc over its ntt^3 physical coarse
// cells.  sub[it] = {sx, sy, sz, tile} (0-based).
__global__ __launch_bounds__(256) void k_subtile_max(const int* __restrict__ rhoc,
                                                     const int4* __restrict__ sub, int ntt, int ncb,
                                                     int nte, int* __restrict__ out) {
    const int4 s = sub[blockIdx.x];
    const int* rc = rhoc + (size_t)s.w * nte * nte * nte;
    const int ncell = ntt * ntt * ntt;
    int m = 0;
    for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
        const int i = c % ntt, j = (c / ntt) % ntt, k = c / (ntt * ntt);
        const int ci = s.x * ntt + i + ncb, cj = s.y * ntt + j + ncb, ck = s.z * ntt + k + ncb;
        m = max(m, rc[((size_t)ck * nte + cj) * nte + ci]);
    }
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
    __shared__ int wm[4];
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = max(max(wm[0], wm[1]), max(wm[2], wm[3]));
}

struct KickState {
    DevBuf sub, maxc;
};
static KickState g_kick;

int select_levels_dev(const int* ixyz3, const int* d_rhoc, int* ires, int* isort, int* napm) {
    C2_TRY(ensure_init());
    Context& C = ctx();
    const Geom& g = C.g;
    C2_TRY(check_ixyz3(g, ixyz3));
    std::vector<int4> sub(g.ns3);
    for (int it = 0; it < g.ns3; it++) {
        const int* x = ixyz3 + 6 * it;
        const int t = ((x[5] - 1) * g.p.nnt + (x[4] - 1)) * g.p.nnt + (x[3] - 1);
        sub[it] = make_int4(x[0] - 1, x[1] - 1, x[2] - 1, t);
    }
    C2_TRY(g_kick.sub.reserve(sub.size() * sizeof(int4)));
    C2_TRY(g_kick.maxc.reserve(sub.size() * sizeof(int)));
    hipStream_t s = C.stream;
    std::vector<int> mx(g.ns3);
    {
        Stage st("select_levels");
        C2_HIP(hipMemcpyAsync(g_kick.sub.ptr, sub.data(), sub.size() * sizeof(int4),
                              hipMemcpyHostToDevice, s));
        k_subtile_max<<<g.ns3, 256, 0, s>>>(d_rhoc, g_kick.sub.as<int4>(), g.ntt, g.ncb, g.nte,
                                            g_kick.maxc.as<int>());
        C2_HIP(hipGetLastError());
        C2_HIP(hipMemcpyAsync(mx.data(), g_kick.maxc.ptr, mx.size() * sizeof(int),
                              hipMemcpyDeviceToHost, s));
        C2_HIP(hipStreamSynchronize(s));
    }
    std::vector<float> neg(g.ns3);
    int nap[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int it = 0; it < g.ns3; it++) {
        ires[it] = level_of(mx[it]);
        nap[ires[it] - 1]++;
        neg[it] = -((float)mx[it] / 64.0f);
    }
    heap_index(g.ns3, neg.data(), isort);
    if (napm)
        for (int q = 0; q < 7; q++) napm[q] = nap[q];
    return CUBE2_OK;
}

// sim%dt_pm3 / sim%dt_pp of kick.f90:180,227 with G_grid = 1/(6 pi)
// (variables.f90:11), in FP64; +inf for a zero force.
static double dt_force(float f2max, float a_mid) {
    if (!(f2max > 0.0f)) return INFINITY;
    const double g_grid = 1.0 / 6.0 / M_PI;
    return 0.5 * std::sqrt(1.0 / (std::sqrt((double)f2max) * (double)a_mid * g_grid));
}

// PM3 and PP of one kick on two streams.  Nothing couples them until the
// velocity update: PM3 reads the particle arrays and writes the pending FP64
// Delta-v, PP's cell lists and pair sums read the particle arrays and write
// their own buffers.  So PP (a split pass) runs on the second stream beside
// PM3, and the update (vp = RN_f(vp + dv64 + PP term), the order of
// pp_force_kernel.cu:496-498,1043-1047: PM3's Delta-v first, then PP's)
// follows on the main stream once both have finished.  The cell-base table
// both read is computed once, before the fork.
static int enqueue_pm3_pp(const PM3Plan& P3, const PPPlan& P, const int* d_rhoc,
                          const int64_t* d_idx, const int16_t* d_xp, float* d_vp,
                          const float* d_scal, int accumulate, hipStream_t s) {
    Context& C = ctx();
    hipStream_t s2 = C.stream2;
    double* dv64 = C.dv64.as<double>();
    const int64_t* nz = nullptr;
    C2_TRY(compute_nzero(d_rhoc, d_idx, &nz, s));
    struct Shared {
        Shared() { ctx().nzero_shared = true; }
        ~Shared() { ctx().nzero_shared = false; }
    } shared;
    C2_HIP(hipEventRecord(C.ev_fork, s));
    C2_HIP(hipStreamWaitEvent(s2, C.ev_fork, 0));
    {
        StageStream on(s2);
        C2_TRY(pp_enqueue(P, d_rhoc, d_idx, d_xp, d_vp, d_scal, nullptr, s2));
    }
    C2_TRY(pm3_enqueue(P3, d_rhoc, d_idx, d_xp, d_scal, accumulate, dv64, s));
    C2_HIP(hipEventRecord(C.ev_join, s2));
    C2_HIP(hipStreamWaitEvent(s, C.ev_join, 0));
    return pp_enqueue_apply(P, d_vp, dv64, s);
}

// The overlapped PM3 + PP of the eager kick driver (prepare, enqueue, wait,
// finish), in place of pm3_force then pp_force.
static int pm3_pp_force(const int* isort, const int* ires, const int* ixyz3, const float* apm3,
                        const int* ratio_sf, const int* d_rhoc, const int64_t* d_idx,
                        const int16_t* d_xp, float* d_vp, int64_t n_used,
                        const float* const* d_green, float mass, float a_mid, float dt,
                        float* f2max_pm3, double* ptotal_pm3, float* f2max_pp, float* vmax,
                        double* ptotal_pp, double* ttotal_pp, uint64_t* cnt, bool learned) {
    Context& C = ctx();
    const auto t0 = std::chrono::steady_clock::now();
    if (C.dv_state == 1 && C.dv_n_used != n_used) {
        set_error("particle count changed while a Delta-v is pending");
        return CUBE2_ERR_STATE;
    }
    PM3Plan P3;
    C2_TRY(pm3_prepare(isort, ires, ixyz3, ratio_sf, n_used, d_green, P3));
    PPPlan P;
    C2_TRY(pp_prepare(ires, ixyz3, apm3, ratio_sf, n_used, P, true, learned));
    const int accumulate = C.dv_state == 1 ? 1 : 0;
    const float* d_scal = nullptr;
    C2_TRY(scalars_upload(mass, a_mid, dt, C.stream, &d_scal));
    C2_TRY(enqueue_pm3_pp(P3, P, d_rhoc, d_idx, d_xp, d_vp, d_scal, accumulate, C.stream));
    C2_TRY(sync_and_flush());
    C.dv_state = 0;
    C2_TRY(pm3_finish(P3, f2max_pm3, ptotal_pm3));
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return pp_finish(P, f2max_pp, vmax, ptotal_pp, ttotal_pp, cnt, wall);
}

int kick(int flags, const int* ixyz3, const float* apm3, const int* ratio_sf, const int* d_rhoc,
         const int64_t* d_idx, const int16_t* d_xp, float* d_vp, int64_t n_used,
         const float* const* d_green, float mass, float a_mid, float dt, int* ires, int* isort,
         double* ptotal_pm3, double* ptotal_pp, double* ttotal_pp, cube2_kick_result_t* out) {
    if (!ixyz3 || !apm3 || !ratio_sf || !ires || !isort || !out ||
        (flags & ~(CUBE2_KICK_PM3 | CUBE2_KICK_PP | CUBE2_KICK_KEEP_LEVELS | CUBE2_KICK_PM2 |
                   CUBE2_KICK_LEARNED)) ||
        ((flags & CUBE2_KICK_PM2) && (!d_green || !d_green[5]))) {
        set_error("kick: null argument or unknown flag");
        return CUBE2_ERR_ARG;
    }
    ensure_params();
    const Geom& g = ctx().g;
    std::memset(out, 0, sizeof(*out));
    if (flags & CUBE2_KICK_KEEP_LEVELS) {
        C2_TRY(check_ixyz3(g, ixyz3));
        for (int it = 0; it < g.ns3; it++) {
            if (ires[it] < 2 || ires[it] > 6) {
                set_error("kick: ires[%d] = %d outside 2..6", it, ires[it]);
                return CUBE2_ERR_ARG;
            }
            out->napm[ires[it] - 1]++;
        }
        if (flags & CUBE2_KICK_PP) C2_TRY(pp_validate(ires, ixyz3, ratio_sf));
        C2_TRY(ensure_init());
    } else {
        C2_TRY(select_levels_dev(ixyz3, d_rhoc, ires, isort, out->napm));
        if (flags & CUBE2_KICK_PP) C2_TRY(pp_validate(ires, ixyz3, ratio_sf));
    }
    std::vector<double> pt3, ptp, ttp;
    if (!ptotal_pm3) ptotal_pm3 = (pt3.assign(g.ns3, 0.0), pt3.data());
    if (!ptotal_pp) ptotal_pp = (ptp.assign(g.ns3, 0.0), ptp.data());
    if (!ttotal_pp) ttotal_pp = (ttp.assign(g.ns3, 0.0), ttp.data());
    out->dt_pm2 = INFINITY;
    if (flags & CUBE2_KICK_PM2) {  // kick.f90:76-105, before PM3
        std::vector<double> pt2(g.ntile, 0.0);
        C2_TRY(pm2_force(d_rhoc, d_idx, d_xp, n_used, d_green[5], mass, a_mid, dt,
                         &out->f2max_pm2, pt2.data()));
        out->dt_pm2 = dt_force(out->f2max_pm2, a_mid);
    }
    const bool both = (flags & CUBE2_KICK_PM3) && (flags & CUBE2_KICK_PP);
    const bool learned = flags & CUBE2_KICK_LEARNED;
    if (both && kick_overlap() && ctx().pp_mode != CUBE2_PP_EXACT) {
        uint64_t cnt[2] = {0, 0};
        C2_TRY(pm3_pp_force(isort, ires, ixyz3, apm3, ratio_sf, d_rhoc, d_idx, d_xp, d_vp, n_used,
                            d_green, mass, a_mid, dt, &out->f2max_pm3, ptotal_pm3, &out->f2max_pp,
                            out->vmax, ptotal_pp, ttotal_pp, cnt, learned));
        for (int it = 0; it < g.ns3; it++) out->pt_pm3 += ptotal_pm3[it];
        out->dt_pm3 = dt_force(out->f2max_pm3, a_mid);
        out->pair_counts[0] = cnt[0];
        out->pair_counts[1] = cnt[1];
        for (int it = 0; it < g.ns3; it++) out->pt_pp += ptotal_pp[it];
        out->dt_pp = dt_force(out->f2max_pp, a_mid);
    } else {
    if (flags & CUBE2_KICK_PM3) {
        C2_TRY(pm3_force(isort, ires, ixyz3, ratio_sf, d_rhoc, d_idx, d_xp, n_used, d_green, mass,
                         a_mid, dt, &out->f2max_pm3, ptotal_pm3));
        for (int it = 0; it < g.ns3; it++) out->pt_pm3 += ptotal_pm3[it];
        out->dt_pm3 = dt_force(out->f2max_pm3, a_mid);
    }
    if (flags & CUBE2_KICK_PP) {
        uint64_t cnt[2] = {0, 0};
        C2_TRY(pp_force(ires, ixyz3, apm3, ratio_sf, d_rhoc, d_idx, d_xp, d_vp, n_used, mass,
                        a_mid, dt, &out->f2max_pp, out->vmax, ptotal_pp, ttotal_pp, cnt, learned));
        out->pair_counts[0] = cnt[0];
        out->pair_counts[1] = cnt[1];
        for (int it = 0; it < g.ns3; it++) out->pt_pp += ptotal_pp[it];
        out->dt_pp = dt_force(out->f2max_pp, a_mid);
    } else if (flags & (CUBE2_KICK_PM3 | CUBE2_KICK_PM2)) {
        // no PP to apply the pending mesh Delta-v: apply it here
        C2_TRY(velocity_apply(d_rhoc, d_idx, d_vp, n_used, out->vmax));
    }
    }
    if (!(flags & CUBE2_KICK_PM3)) out->dt_pm3 = INFINITY;
    if (!(flags & CUBE2_KICK_PP)) out->dt_pp = INFINITY;
    const float vm = std::max(out->vmax[0], std::max(out->vmax[1], out->vmax[2]));
    out->dt_vmax = vm > 0.0f ? 0.9 * 14.0 / (double)vm : INFINITY;  // kick.f90:236; vbuf = 0.9
    return CUBE2_OK;
}

// ------------------------------------------------------- recorded kick
// The kick with fixed levels, recorded once into a HIP graph and replayed.
// Everything that shapes the launches is part of the key; the scalars
// (mass, a_mid, dt) are read by the kernels from device memory, which the
// graph refills from a pinned block at each launch.
struct GraphKey {
    int flags = 0, mode = 0;
    bool overlap = false;  // PM3 and PP on two streams (recorded as parallel branches)
    bool learned = false;  // the network stands in for PP in sparse level-2 cells
    int64_t n_used = 0;
    const void* arrays[4] = {nullptr, nullptr, nullptr, nullptr};
    const float* green[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    std::vector<int> ires, isort, ixyz3;
    float apm3[7] = {0};
    int ratio_sf[7] = {0};
    cube2_params_t p{};
    bool same(const GraphKey& o) const {
        return flags == o.flags && mode == o.mode && overlap == o.overlap && learned == o.learned &&
               n_used == o.n_used &&
               std::equal(arrays, arrays + 4, o.arrays) && std::equal(green, green + 5, o.green) &&
               ires == o.ires && isort == o.isort && ixyz3 == o.ixyz3 &&
               std::memcmp(apm3, o.apm3, sizeof(apm3)) == 0 &&
               std::memcmp(ratio_sf, o.ratio_sf, sizeof(ratio_sf)) == 0 &&
               std::memcmp(&p, &o.p, sizeof(p)) == 0;
    }
};

struct KickGraph {
    bool valid = false;
    GraphKey key;
    uint64_t generation = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    PM3Plan pm3;
    PPPlan pp;
};
static KickGraph g_graph;

static void graph_destroy() {
    if (g_graph.exec) (void)hipGraphExecDestroy(g_graph.exec);
    if (g_graph.graph) (void)hipGraphDestroy(g_graph.graph);
    g_graph.exec = nullptr;
    g_graph.graph = nullptr;
    g_graph.valid = false;
}

static int graph_enqueue(const GraphKey& K, const int* d_rhoc, const int64_t* d_idx,
                         const int16_t* d_xp, float* d_vp, float mass, float a_mid, float dt,
                         hipStream_t s) {
    Context& C = ctx();
    const float* d_scal = nullptr;
    C2_TRY(scalars_upload(mass, a_mid, dt, s, &d_scal));
    const bool pm3 = K.flags & CUBE2_KICK_PM3;
    if (K.overlap)
        return enqueue_pm3_pp(g_graph.pm3, g_graph.pp, d_rhoc, d_idx, d_xp, d_vp, d_scal, 0, s);
    if (pm3)
        C2_TRY(pm3_enqueue(g_graph.pm3, d_rhoc, d_idx, d_xp, d_scal, 0, C.dv64.as<double>(), s));
    const double* dv = pm3 ? C.dv64.as<double>() : nullptr;
    C2_TRY(pp_enqueue(g_graph.pp, d_rhoc, d_idx, d_xp, d_vp, d_scal, dv, s));
    if (g_graph.pp.split) C2_TRY(pp_enqueue_apply(g_graph.pp, d_vp, dv, s));
    return CUBE2_OK;
}

static int graph_record(const GraphKey& K, const int* ixyz3, const float* apm3,
                        const int* ratio_sf, const int* d_rhoc, const int64_t* d_idx,
                        const int16_t* d_xp, float* d_vp, const float* const* d_green, float mass,
                        float a_mid, float dt) {
    Context& C = ctx();
    hipStream_t s = C.stream;
    graph_destroy();
    if (K.flags & CUBE2_KICK_PM3)
        C2_TRY(pm3_prepare(K.isort.data(), K.ires.data(), ixyz3, ratio_sf, K.n_used, d_green,
                           g_graph.pm3));
    C2_TRY(pp_prepare(K.ires.data(), ixyz3, apm3, ratio_sf, K.n_used, g_graph.pp,
                      K.overlap || K.learned, K.learned));
    if (!K.overlap)
        g_graph.pp.grid_pairs = pp_tiled_grid(g_graph.pp.mode, false, g_graph.pp.newton);
    C2_TRY(C.dv64.reserve((size_t)std::max<int64_t>(K.n_used, 1) * 3 * sizeof(double)));
    const float* d_scal = nullptr;
    C2_TRY(scalars_upload(mass, a_mid, dt, s, &d_scal));  // reserves the scalar blocks
    C2_HIP(hipStreamSynchronize(s));
    const uint64_t gen = buffers_generation();
    C2_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    C.capturing = true;
    const int rc = graph_enqueue(K, d_rhoc, d_idx, d_xp, d_vp, mass, a_mid, dt, s);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &graph);
    C.capturing = false;
    if (rc != CUBE2_OK || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        if (rc != CUBE2_OK) return rc;
        set_error("graph recording failed: %s", hipGetErrorString(e));
        return CUBE2_ERR_HIP;
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(graph);
        set_error("graph instantiation failed: %s", hipGetErrorString(ei));
        return CUBE2_ERR_HIP;
    }
    if (buffers_generation() != gen) {  // nothing may grow while recording
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        set_error("library buffers changed while recording the kick");
        return CUBE2_ERR_STATE;
    }
    g_graph.graph = graph;
    g_graph.exec = exec;
    g_graph.key = K;
    g_graph.generation = gen;
    g_graph.valid = true;
    return CUBE2_OK;
}

int kick_graph(int flags, const int* ixyz3, const float* apm3, const int* ratio_sf,
               const int* d_rhoc, const int64_t* d_idx, const int16_t* d_xp, float* d_vp,
               int64_t n_used, const float* const* d_green, float mass, float a_mid, float dt,
               const int* ires, const int* isort, double* ptotal_pm3, double* ptotal_pp,
               double* ttotal_pp, cube2_kick_result_t* out, int* recorded) {
    if (!ixyz3 || !apm3 || !ratio_sf || !ires || !out || !d_rhoc || !d_idx || !d_xp || !d_vp ||
        ((flags & ~CUBE2_KICK_LEARNED) != (CUBE2_KICK_PM3 | CUBE2_KICK_PP) &&
         (flags & ~CUBE2_KICK_LEARNED) != CUBE2_KICK_PP)) {
        set_error("recorded kick: null argument, or flags other than PM3|PP or PP (with LEARNED)");
        return CUBE2_ERR_ARG;
    }
    if ((flags & CUBE2_KICK_PM3) && (!isort || !d_green)) {
        set_error("recorded kick: PM3 needs isort and Green's tables");
        return CUBE2_ERR_ARG;
    }
    ensure_params();
    Context& C = ctx();
    const Geom& g = C.g;
    C2_TRY(check_ixyz3(g, ixyz3));
    for (int it = 0; it < g.ns3; it++)
        if (ires[it] < 2 || ires[it] > 6) {
            set_error("recorded kick: ires[%d] = %d outside 2..6", it, ires[it]);
            return CUBE2_ERR_ARG;
        }
    C2_TRY(pp_validate(ires, ixyz3, ratio_sf));
    if (C.dv_state == 1) {
        set_error("recorded kick: a PM3 Delta-v is pending from an earlier call");
        return CUBE2_ERR_STATE;
    }
    C2_TRY(ensure_init());
    GraphKey K;
    K.flags = flags;
    K.mode = C.pp_mode;
    K.learned = flags & CUBE2_KICK_LEARNED;
    K.flags = flags & ~CUBE2_KICK_LEARNED;
    K.overlap = (flags & CUBE2_KICK_PM3) && kick_overlap() && C.pp_mode != CUBE2_PP_EXACT;
    K.n_used = n_used;
    K.arrays[0] = d_rhoc;
    K.arrays[1] = d_idx;
    K.arrays[2] = d_xp;
    K.arrays[3] = d_vp;
    if (flags & CUBE2_KICK_PM3) {
        for (int q = 0; q < 5; q++) K.green[q] = d_green[q];
        K.isort.assign(isort, isort + g.ns3);
    }
    K.ires.assign(ires, ires + g.ns3);
    K.ixyz3.assign(ixyz3, ixyz3 + 6 * g.ns3);
    std::memcpy(K.apm3, apm3, sizeof(K.apm3));
    std::memcpy(K.ratio_sf, ratio_sf, sizeof(K.ratio_sf));
    K.p = g.p;
    const bool fresh = !(g_graph.valid && g_graph.key.same(K) &&
                         g_graph.generation == buffers_generation());
    if (fresh)
        C2_TRY(graph_record(K, ixyz3, apm3, ratio_sf, d_rhoc, d_idx, d_xp, d_vp, d_green, mass,
                            a_mid, dt));
    if (recorded) *recorded = fresh ? 1 : 0;
    hipStream_t s = C.stream;
    float* h = C.h_scal.as<float>();  // read by the graph's first node at launch
    h[0] = mass;
    h[1] = a_mid;
    h[2] = dt;
    h[3] = 0.0f;
    const auto t0 = std::chrono::steady_clock::now();
    {
        Stage st("kick_graph");
        C2_HIP(hipGraphLaunch(g_graph.exec, s));
    }
    C2_TRY(sync_and_flush());
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::memset(out, 0, sizeof(*out));
    for (int it = 0; it < g.ns3; it++) out->napm[ires[it] - 1]++;
    std::vector<double> pt3, ptp, ttp;
    if (!ptotal_pm3) ptotal_pm3 = (pt3.assign(g.ns3, 0.0), pt3.data());
    if (!ptotal_pp) ptotal_pp = (ptp.assign(g.ns3, 0.0), ptp.data());
    if (!ttotal_pp) ttotal_pp = (ttp.assign(g.ns3, 0.0), ttp.data());
    if (flags & CUBE2_KICK_PM3) {
        C2_TRY(pm3_finish(g_graph.pm3, &out->f2max_pm3, ptotal_pm3));
        for (int it = 0; it < g.ns3; it++) out->pt_pm3 += ptotal_pm3[it];
        out->dt_pm3 = dt_force(out->f2max_pm3, a_mid);
    } else {
        out->dt_pm3 = INFINITY;
    }
    uint64_t cnt[2] = {0, 0};
    C2_TRY(pp_finish(g_graph.pp, &out->f2max_pp, out->vmax, ptotal_pp, ttotal_pp, cnt, wall));
    out->pair_counts[0] = cnt[0];
    out->pair_counts[1] = cnt[1];
    for (int it = 0; it < g.ns3; it++) out->pt_pp += ptotal_pp[it];
    out->dt_pp = dt_force(out->f2max_pp, a_mid);
    const float vm = std::max(out->vmax[0], std::max(out->vmax[1], out->vmax[2]));
    out->dt_vmax = vm > 0.0f ? 0.9 * 14.0 / (double)vm : INFINITY;
    C.dv_state = 0;
    return CUBE2_OK;
}

void kick_release() {
    graph_destroy();
    g_kick.sub.release();
    g_kick.maxc.release();
}

}  // namespace cube2
