// PP pair kernels.
//
// exact: one lane per physical particle i, j visited in ascending sorted
//   order over the (2R+1)^2 neighbour rows, each row one contiguous run of
//   2R+1 cells.  Every float operation is the reference C kernel's
//   (pp_force_kernel.c:308-318, F_ra :45-74) with contraction off and IEEE
//   sqrt/divide, and the accumulation order per i is the same, so Delta-v is
//   bit-identical to the reference CPU kernel on the same inputs.
// fast: same traversal, FMA + rsqrt arithmetic, F_ra in Horner form with the
//   12/ep^2 term folded into 1/r^2 (relative RMS vs exact <= 1e-5).
// Both finish with the velocity update of pp_force_kernel.c:332-364: a*=m^2,
// f2max, vp = (float)((double)vp + dv64 + (double)(a*a_mid*dt/6)/pi), vmax.
#include "pp_kernels.h"
#include "profile.h"
#include "reduce.h"

namespace cube2 {

__device__ inline int find_tile(const PPTile* __restrict__ T, int ntile, uint32_t c) {
    int lo = 0, hi = ntile - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (T[mid].cell_off <= (long long)c) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// basic_functions.f08:95-109, term for term (no contraction).
__device__ inline float fra_exact(float r, float apm) {
#pragma clang fp contract(off)
    float ep = 2.0f * r / apm;
    if (apm == 0.0f || ep > 2.0f) return 1.0f / (r * r);
    if (ep > 1.0f) {
        float e2 = ep * ep, e3 = e2 * ep, e4 = e3 * ep, e5 = e4 * ep, e6 = e5 * ep;
        return (1.0f / 35.0f / (apm * apm)) *
               (12.0f / e2 - 224.0f + 896.0f * ep - 840.0f * e2 + 224.0f * e3 + 70.0f * e4 -
                48.0f * e5 + 7.0f * e6);
    }
    if (ep > 0.0f) {
        float e3 = ep * ep * ep, e4 = e3 * ep, e5 = e4 * ep, e6 = e5 * ep;
        return (1.0f / 35.0f / (apm * apm)) *
               (224.0f * ep - 224.0f * e3 + 70.0f * e4 + 48.0f * e5 - 21.0f * e6);
    }
    return 0.0f;
}

// F_ra(r, a) from ep = 2r/a, 1/r^2 and c = 1/(35 a^2), Horner form.
// For 1 < ep <= 2 the 12/ep^2 term equals (3/35) * (1/r^2) * ... / c-scaled:
// c * 12/ep^2 = 12 a^2 / (4 r^2) / (35 a^2) = (3/35) / r^2.
__device__ inline float fra_fast(float ep, float ir2, float c) {
    if (ep > 2.0f) return ir2;
    float hi = fmaf(3.0f / 35.0f, ir2,
                    c * fmaf(ep, fmaf(ep, fmaf(ep, fmaf(ep, fmaf(ep, fmaf(ep, 7.0f, -48.0f), 70.0f),
                                                         224.0f), -840.0f), 896.0f), -224.0f));
    float e2 = ep * ep;
    float lo = c * ep * fmaf(e2, fmaf(ep, fmaf(ep, fmaf(ep, -21.0f, 48.0f), 70.0f), -224.0f), 224.0f);
    return ep > 1.0f ? hi : lo;
}

// a *= m^2; f2; vp update in FP64 (pp_force_kernel.c:350-360).
__device__ inline void pp_update(const PPArgs& A, uint32_t a, float ax, float ay, float az,
                                 float& f2, float& v0, float& v1, float& v2) {
#pragma clang fp contract(off)
    const float mass = A.scal[0], a_mid = A.scal[1], dt = A.scal[2];
    const float a0 = ax * mass * mass;
    const float a1 = ay * mass * mass;
    const float a2 = az * mass * mass;
    f2 = a0 * a0 + a1 * a1 + a2 * a2;
    const size_t ip = (size_t)A.slot[a] * 3;
    const float acc[3] = {a0, a1, a2};
    float nv[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const double term = (double)(acc[q] * a_mid * dt / 6.0f) / 3.14159265358979323846;
        double v = (double)A.vp[ip + q];
        if (A.dv64) v += A.dv64[ip + q];
        nv[q] = (float)(v + term);
        A.vp[ip + q] = nv[q];
    }
    v0 = fabsf(nv[0]);
    v1 = fabsf(nv[1]);
    v2 = fabsf(nv[2]);
}

__device__ inline void pp_reduce(const PPArgs& A, float f2, float v0, float v1, float v2,
                                 unsigned long long pc, unsigned long long pa) {
    const float mx[4] = {f2, v0, v1, v2};
    block_max_to<4>(&A.red->f2max_bits, mx);   // f2max_bits, vmax_bits[3] are contiguous
    const unsigned long long sm[2] = {pc, pa};
    block_sum_to<2>(&A.red->p_cand, sm);        // p_cand, p_acc are contiguous
}

template <bool kExact>
__global__ __launch_bounds__(256) void k_pp_pairs(PPArgs A) {
    if (A.red->err & 4) return;  // failed call: the sorted arrays are not valid
    const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = (uint32_t)*A.n_sorted;
    float f2 = 0.f, v0 = 0.f, v1 = 0.f, v2 = 0.f;
    unsigned long long pc = 0, pa = 0;
    float4 p = make_float4(0, 0, 0, 0);
    if (a < n) p = A.pos4[a];
    if (p.w > 0.0f) {
        const uint32_t cid = A.cellid[a];
        const PPTile T = A.tiles[find_tile(A.tiles, A.ntile, cid)];
        const long long key = (long long)cid - T.cell_off;
        const int S = T.S;
        const int x = (int)(key % S), y = (int)((key / S) % S), z = (int)(key / ((long long)S * S));
        int R;
        {
#pragma clang fp contract(off)
            R = (int)ceilf(p.w * (float)T.rcp / (float)A.ratio_cs);
        }
        const float range = p.w;
        float pp2;
        {
#pragma clang fp contract(off)
            pp2 = range * range;
        }
        // fast-path constants
        const float k2 = 2.0f / range, cr = 1.0f / (35.0f * range * range);
        const float ka = A.app > 0.0f ? 2.0f / A.app : 0.0f;
        const float ca = A.app > 0.0f ? 1.0f / (35.0f * A.app * A.app) : 0.0f;
        float ax = 0.f, ay = 0.f, az = 0.f;
        const int xlo = max(x - R, 0), xhi = min(x + R, S - 1);
        for (int dz = -R; dz <= R; dz++) {
            const int zz = z + dz;
            if (zz < 0 || zz >= S) continue;
            for (int dy = -R; dy <= R; dy++) {
                const int yy = y + dy;
                if (yy < 0 || yy >= S) continue;
                const long long row = T.cell_off + ((long long)zz * S + yy) * S;
                const int j0 = A.cell_start[row + xlo], j1 = A.cell_start[row + xhi + 1];
                pc += (unsigned long long)(j1 - j0);
                for (int j = j0; j < j1; j++) {
                    const float4 q = A.pos4[j];
                    if (kExact) {
#pragma clang fp contract(off)
                        const float dx = q.x - p.x, dy_ = q.y - p.y, dz_ = q.z - p.z;
                        const float r2 = dx * dx + dy_ * dy_ + dz_ * dz_;
                        if (r2 > 0.0f && r2 < pp2) {
                            const float rm = sqrtf(r2);
                            const float fpp = fra_exact(rm, A.app) - fra_exact(rm, range);
                            ax += fpp * dx / rm;
                            ay += fpp * dy_ / rm;
                            az += fpp * dz_ / rm;
                            pa++;
                        }
                    } else {
                        const float dx = q.x - p.x, dy_ = q.y - p.y, dz_ = q.z - p.z;
                        const float r2 = fmaf(dz_, dz_, fmaf(dy_, dy_, dx * dx));
                        if (r2 > 0.0f && r2 < pp2) {
                            const float ir = __frsqrt_rn(r2);
                            const float ir2 = ir * ir;
                            const float r = r2 * ir;
                            float fa = ir2;
                            if (r <= A.app) fa = fra_fast(r * ka, ir2, ca);
                            const float f = (fa - fra_fast(r * k2, ir2, cr)) * ir;
                            ax = fmaf(f, dx, ax);
                            ay = fmaf(f, dy_, ay);
                            az = fmaf(f, dz_, az);
                            pa++;
                        }
                    }
                }
            }
        }
        pc -= 1;  // the particle itself
        pp_update(A, a, ax, ay, az, f2, v0, v1, v2);
        if (A.tile_pairs) {  // per-tile counts (on request only)
            const int ti = find_tile(A.tiles, A.ntile, cid);
            atomicAdd(&A.tile_pairs[2 * ti], pc);
            atomicAdd(&A.tile_pairs[2 * ti + 1], pa);
        }
    }
    pp_reduce(A, f2, v0, v1, v2, pc, pa);
}

int launch_pp_pairs(const PPArgs& a, int mode, hipStream_t s) {
    Stage st(mode == CUBE2_PP_EXACT ? "pp_pairs_exact" : "pp_pairs_fast");
    const unsigned blocks = ceil_div((size_t)a.cap, 256);  // n_sorted <= cap, read on the device
    if (mode == CUBE2_PP_EXACT) k_pp_pairs<true><<<blocks, 256, 0, s>>>(a);
    else k_pp_pairs<false><<<blocks, 256, 0, s>>>(a);
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

}  // namespace cube2
