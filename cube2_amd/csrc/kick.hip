// Short-range kick driver (kick.f90:107-238) over the device-resident
// arrays: level selection from the per-subtile maximum coarse-cell count,
// job order, PM3, PP (which applies PM3's pending Delta-v), and the FP64
// time-step limits.
#include <cmath>

#include "common.h"
#include "profile.h"

namespace cube2 {

int pp_validate(const int* ires, const int* ixyz3, const int* ratio_sf);
int pp_force(const int* ires, const int* ixyz3, const float* apm3, const int* ratio_sf,
             const int* d_rhoc, const int64_t* d_idx, const int16_t* d_xp, float* d_vp,
             int64_t n_used, float mass, float a_mid, float dt, float* f2max, float* vmax,
             double* ptotal, double* ttotal, uint64_t* pair_counts);
int pm3_force(const int* isort, const int* ires, const int* ixyz3, const int* ratio_sf,
              const int* d_rhoc, const int64_t* d_idx, const int16_t* d_xp, int64_t n_used,
              const float* const* d_green, float mass, float amid, float dt, float* f2max,
              double* ptotal);
int velocity_apply(const int* d_rhoc, const int64_t* d_idx, float* d_vp, int64_t n_used,
                   float* vmax);

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

int kick(int flags, const int* ixyz3, const float* apm3, const int* ratio_sf, const int* d_rhoc,
         const int64_t* d_idx, const int16_t* d_xp, float* d_vp, int64_t n_used,
         const float* const* d_green, float mass, float a_mid, float dt, int* ires, int* isort,
         double* ptotal_pm3, double* ptotal_pp, double* ttotal_pp, cube2_kick_result_t* out) {
    if (!ixyz3 || !apm3 || !ratio_sf || !ires || !isort || !out ||
        (flags & ~(CUBE2_KICK_PM3 | CUBE2_KICK_PP | CUBE2_KICK_KEEP_LEVELS))) {
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
    if (flags & CUBE2_KICK_PM3) {
        C2_TRY(pm3_force(isort, ires, ixyz3, ratio_sf, d_rhoc, d_idx, d_xp, n_used, d_green, mass,
                         a_mid, dt, &out->f2max_pm3, ptotal_pm3));
        for (int it = 0; it < g.ns3; it++) out->pt_pm3 += ptotal_pm3[it];
        out->dt_pm3 = dt_force(out->f2max_pm3, a_mid);
    }
    if (flags & CUBE2_KICK_PP) {
        uint64_t cnt[2] = {0, 0};
        C2_TRY(pp_force(ires, ixyz3, apm3, ratio_sf, d_rhoc, d_idx, d_xp, d_vp, n_used, mass,
                        a_mid, dt, &out->f2max_pp, out->vmax, ptotal_pp, ttotal_pp, cnt));
        out->pair_counts[0] = cnt[0];
        out->pair_counts[1] = cnt[1];
        for (int it = 0; it < g.ns3; it++) out->pt_pp += ptotal_pp[it];
        out->dt_pp = dt_force(out->f2max_pp, a_mid);
    } else if (flags & CUBE2_KICK_PM3) {
        // no PP to apply the pending PM3 Delta-v: apply it here
        C2_TRY(velocity_apply(d_rhoc, d_idx, d_vp, n_used, out->vmax));
    }
    if (!(flags & CUBE2_KICK_PM3)) out->dt_pm3 = INFINITY;
    if (!(flags & CUBE2_KICK_PP)) out->dt_pp = INFINITY;
    const float vm = std::max(out->vmax[0], std::max(out->vmax[1], out->vmax[2]));
    out->dt_vmax = vm > 0.0f ? 0.9 * 14.0 / (double)vm : INFINITY;  // kick.f90:236; vbuf = 0.9
    return CUBE2_OK;
}

void kick_release() {
    g_kick.sub.release();
    g_kick.maxc.release();
}

}  // namespace cube2
