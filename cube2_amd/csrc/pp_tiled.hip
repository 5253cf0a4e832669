// PP pair kernel.
//
// Wavefronts take chunks of physical particles of one PP cell by a static
// stride (a grid of 64 blocks per CU, far more than are resident, so that
// the hardware balances the chunks' uneven costs): the whole cell if it
// holds fewer than 64, else 2^dense_lg at a time (16 built in, 32 with pairs
// once).  The lanes are split into I i-slots x J j-slots (I the chunk
// size rounded up to a power of two, I*J = 64).  The j of the (2R+1)^2
// neighbour rows are contiguous runs of the sorted particle stream; a
// wavefront concatenates them into one candidate list and fetches it 64 at
// a time (one coalesced load, prefetched a batch ahead, staged in LDS).
//
// Chunks of I >= 2^direct_lg (I >= 4 built in) sum directly: lane (i, s)
// evaluates the force law on the pairs of i with the j of its j-slot s as
// they are accepted, partial sums in registers in the reference's j order,
// combined over the J slots at the end.  Cells of >= 64 particles are
// stored in 8x8x8 Morton sub-cell order (so a chunk is spatially compact), and each fetched batch of j is first reduced to those within
// the range of the chunk's bounding box: the smaller the chunk, the fewer
// lanes idle on a j accepted by only some of its i (REF at 256^3: 112 ms
// with 64-particle chunks, 92 ms with 16).  Smaller chunks append accepted
// pairs to a per-wavefront LDS queue (ballot + prefix count) and evaluate
// the force law 64 pairs at a time.
//
// kRef: the pair force is the reference C kernel's, operation for operation
//   (pp_force_kernel.c:308-318, F_ra :45-74), with each correctly rounded
//   float division a/b evaluated as RN_f(a * y), y an FP64 reciprocal of b
//   accurate to 2^-52.  The quotient of two floats is either exactly
//   representable or at least 2^-49 (relative) away from every rounding
//   midpoint, so this equals IEEE a/b.  The pair forces are bit-identical to
//   the reference; only the order of the per-i sums differs (<= ~4e-7
//   relative RMS), and with pairs once (kNewton, below) a pair of one range
//   is evaluated once and its force summed for both particles in 64-bit
//   fixed point.
// FAST: FMA/rsqrt arithmetic with F_ra in Horner form.
// PLUMMER: Plummer softening r/(r^2 + app^2)^(3/2) in place of F_ra(r, app)
//   (no counterpart in the reference), in PRECISE's arithmetic.
// PRECISE: a cancellation-free force law and FP64 per-particle sums.  For
//   1 < ep <= 2, 35 a^2 ep^2 (F_ra(r,a) - 1/r^2) = (ep-2)^5 (7ep^3 + 22ep^2
//   + 10ep + 4), so 1/r^2 - F_ra = (2-ep)^5 (7ep^3+22ep^2+10ep+4)/(140 r^2)
//   without the term-by-term cancellation near the cutoff; sums of pair
//   terms accumulate in FP64 (full rate on MI355X) instead of Kahan-
//   compensated FP32.
#include <algorithm>
#include <type_traits>

#include "pp_kernels.h"
#include "profile.h"
#include "reduce.h"

namespace cube2 {

constexpr int kQ = 128;  // queue entries per wavefront
// A position farther from every particle than any range: its squared distance
// overflows to +inf, so a pair test with it fails without a bounds check.
constexpr float kFar = 1e30f;

// 1/x from an approximation y0: with e = 1 - x*y0 (one FMA, rounded once),
// 1/x = y0/(1-e) = y0 (1 + e + e^2 + e^3/(1-e)), so y0 + y0 (e + e^2) is
// within |e|^3 + 3*2^-53 (relative) of 1/x.  For |e| <= 2^-18 that is below
// 2^-51.
__device__ inline double recip_refine(double x, double y0) {
    const double e = fma(-x, y0, 1.0);
    return fma(fma(e, e, e), y0, y0);
}

// FP64 reciprocal of a float: v_rcp_f64 and one cubic refinement.  The
// self-test checks |1 - x*y| <= 2^-51 for every float x with exponent in
// [-40, 40].
__device__ inline double recip64(double x) { return recip_refine(x, __builtin_amdgcn_rcp(x)); }

__device__ inline float div_by(float a, double y) { return (float)((double)a * y); }

// RN(sqrt(x)) for normal positive finite x (pair distances: r^2 >= 2^-28,
// the positions being multiples of ratio_cs * 2^-16 apart).  v_sqrt_f32 is
// within one ulp; of s and its neighbours s-, s+ the rounding test keeps s-
// if x <= s- * s and s+ if x > s * s+ (FMA residuals, exact in sign: the
// residuals are integer multiples of ulp^2 and midpoints squared are not).
// No scaling or special-value fix-ups, which only zero, infinities, NaNs and
// tiny inputs need.  The self-test checks it against IEEE sqrtf for every
// float with exponent in [-40, 40].
__device__ inline float sqrt_rn(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float lo = __int_as_float(__float_as_int(s) - 1);
    const float hi = __int_as_float(__float_as_int(s) + 1);
    float r = fmaf(-lo, s, x) <= 0.0f ? lo : s;
    r = fmaf(-hi, s, x) > 0.0f ? hi : r;
    return r;
}

struct FraK {
    float a;        // softening diameter
    double inv_a;   // 1/a in FP64
    double half_a2;  // (a/2)^2 in FP64
    float c;        // (1/35)/(a*a), as basic_functions.f08:103 rounds it
    float r2max;    // a^2 (1 + 2^-17): for r^2 above it, RN(2r/a) > 2 for certain
};

__device__ inline FraK make_frak(float a) {
    FraK K;
    K.a = a;
    K.inv_a = a > 0.0f ? 1.0 / (double)a : 0.0;
    K.half_a2 = 0.25 * (double)a * (double)a;
    K.r2max = (float)((double)a * a * (1.0 + 0x1p-17));
    {
#pragma clang fp contract(off)
        K.c = 1.0f / 35.0f / (a * a);
    }
    return K;
}

// F_ra(r, a) of basic_functions.f08:95-109 with the reference's rounding, for
// ep = RN(2r/a) <= 2 (the caller handles a = 0 and ep > 2, where F_ra is
// 1/(r*r)).  Evaluated without divergence: the powers are shared, both
// polynomials are formed in the reference's operation order, and one is
// selected.  RN(12/e2) is 12 * (FP64 reciprocal), rounded once to float;
// the reciprocal refines y0 = (a/2)^2 (1/r)^2, which is within a few float
// roundings of 1/e2 (ep = RN(2r/a), e2 = RN(ep*ep)).
__device__ inline float fra_ref_poly(float ep, const FraK& K, double inv_r2) {
#pragma clang fp contract(off)
    const float e2 = ep * ep, e3 = e2 * ep, e4 = e3 * ep, e5 = e4 * ep, e6 = e5 * ep;
    const float t12 = (float)(12.0 * recip_refine((double)e2, K.half_a2 * inv_r2));
    const float hi = K.c * (t12 - 224.0f + 896.0f * ep - 840.0f * e2 + 224.0f * e3 + 70.0f * e4 -
                            48.0f * e5 + 7.0f * e6);
    const float lo = K.c * (224.0f * ep - 224.0f * e3 + 70.0f * e4 + 48.0f * e5 - 21.0f * e6);
    // (the reference's F_ra is 0 at ep = 0; pairs here have r^2 >= 2^-28, so
    // ep > 0)
    return ep > 1.0f ? hi : lo;
}

__device__ inline float fra_fast2(float ep, float ir2, float c) {
    if (ep > 2.0f) return ir2;
    const float hi = fmaf(3.0f / 35.0f, ir2,
                          c * fmaf(ep, fmaf(ep, fmaf(ep, fmaf(ep, fmaf(ep, fmaf(ep, 7.0f, -48.0f),
                                                                         70.0f), 224.0f), -840.0f),
                                            896.0f), -224.0f));
    const float e2 = ep * ep;
    const float lo = c * ep * fmaf(e2, fmaf(ep, fmaf(ep, fmaf(ep, -21.0f, 48.0f), 70.0f), -224.0f), 224.0f);
    return ep > 1.0f ? hi : lo;
}

enum { kModeRef = 0, kModeFast = 1, kModePrecise = 2, kModePlummer = 3 };

// 1/r^2 - F_ra(r, a) for r < a (ep = 2r/a in (0, 2]), cancellation-free.
__device__ inline float short_term(float ep, float ir2, float c) {
    const float t = 2.0f - ep, t2 = t * t;
    const float q = fmaf(ep, fmaf(ep, fmaf(ep, 7.0f, 22.0f), 10.0f), 4.0f);
    const float hi = t2 * t2 * t * q * ir2 * (1.0f / 140.0f);
    const float e2 = ep * ep;
    const float lo =
        ir2 - c * ep * fmaf(e2, fmaf(ep, fmaf(ep, fmaf(ep, -21.0f, 48.0f), 70.0f), -224.0f), 224.0f);
    return ep > 1.0f ? hi : lo;
}

// 1/r^2 - r / (r^2 + eps^2)^(3/2) (Plummer softening), cancellation-free:
// with u = eps^2/r^2 and w = (1 + u)^(-1/2) it is (1 - w^3)/r^2, and
// 1 - w = u w^2 / (1 + w), so (1/r^2) u w^2 (1 + w + w^2) / (1 + w).
__device__ inline float plummer_term(float ir2, float eps2) {
    const float u = eps2 * ir2;
    const float w = __frsqrt_rn(1.0f + u);
    const float w2 = w * w;
    return ir2 * u * w2 * (1.0f + w + w2) / (1.0f + w);
}

template <int kMode>
__device__ inline void pair_force(float4 e, const FraK& Ka, const FraK& Kr, float& fx, float& fy,
                                  float& fz) {
    if (kMode == kModeRef) {
#pragma clang fp contract(off)
        const float rm = sqrt_rn(e.w);
        const double inv_rm = recip64((double)rm);
        // RN(1/(r*r)): the reciprocal of RN(r*r) refined from (1/r)^2
        const double inv_rm2 = inv_rm * inv_rm;
        const float frr = (float)recip_refine((double)(rm * rm), inv_rm2);
        // range term: r < range, so ep <= 2 except by rounding at the cutoff
        const float epr = div_by(2.0f * rm, Kr.inv_a);
        const float fr = epr > 2.0f ? frr : fra_ref_poly(epr, Kr, inv_rm2);
        // softening term: 1/(r*r) unless r is within app; the polynomial runs
        // only when some lane of the wavefront needs it
        float fa = frr;
        if (Ka.a != 0.0f && __ballot(e.w <= Ka.r2max)) {
            const float epa = div_by(2.0f * rm, Ka.inv_a);
            if (__ballot(epa <= 2.0f)) {
                const float pa = fra_ref_poly(epa, Ka, inv_rm2);
                if (epa <= 2.0f) fa = pa;
            }
        }
        const float fpp = fa - fr;
        fx = div_by(fpp * e.x, inv_rm);
        fy = div_by(fpp * e.y, inv_rm);
        fz = div_by(fpp * e.z, inv_rm);
    } else if (kMode == kModeFast) {
        const float ir = __frsqrt_rn(e.w);
        const float ir2 = ir * ir;
        const float r = e.w * ir;
        float fa = ir2;
        if (r <= Ka.a) fa = fra_fast2(r * (float)(2.0 * Ka.inv_a), ir2, Ka.c);
        const float f = (fa - fra_fast2(r * (float)(2.0 * Kr.inv_a), ir2, Kr.c)) * ir;
        fx = f * e.x;
        fy = f * e.y;
        fz = f * e.z;
    } else if (kMode == kModePlummer) {
        // pair = (1/r^2 - F_ra(r, range)) - (1/r^2 - r/(r^2 + app^2)^(3/2)),
        // each piece cancellation-free (softening length eps = app)
        const float r = sqrtf(e.w);
        const float ir = 1.0f / r;
        const float ir2 = ir * ir;
        float f = short_term(r * (float)(2.0 * Kr.inv_a), ir2, Kr.c);
        if (Ka.a > 0.0f) f -= plummer_term(ir2, Ka.a * Ka.a);
        f *= ir;
        fx = f * e.x;
        fy = f * e.y;
        fz = f * e.z;
    } else {
        // pair = (1/r^2 - F_ra(r, range)) - (1/r^2 - F_ra(r, app)), the second
        // term zero for r > app; each piece cancellation-free.
        const float r = sqrtf(e.w);
        const float ir = 1.0f / r;
        const float ir2 = ir * ir;
        float f = short_term(r * (float)(2.0 * Kr.inv_a), ir2, Kr.c);
        if (Ka.a > 0.0f && r < Ka.a) f -= short_term(r * (float)(2.0 * Ka.inv_a), ir2, Ka.c);
        f *= ir;
        fx = f * e.x;
        fy = f * e.y;
        fz = f * e.z;
    }
}

// Orders LDS accesses of the lanes of one wavefront.  LDS operations of a
// wavefront are performed in issue order, so a compiler barrier suffices (no
// block barrier: the wavefronts of a block run independent trip counts).
__device__ inline void wave_lds_sync() { __builtin_amdgcn_wave_barrier(); }

template <int kMode, int kJB = 64, bool kNewton = false>
struct WaveLDS {
    using Acc = typename std::conditional<kMode == kModePrecise || kMode == kModePlummer, double,
                                          float>::type;
    float4 q[kQ];           // queued chunks: accepted pairs (dx, dy, dz, r^2)
    unsigned char qi[kQ];   // and their i slot
    Acc acc[64][3];         // per-i force sums
    float4 jb[kJB + 64];    // current batch of j positions, then J far-away ones
    int rs[64], rp[64];     // row group: j = v + rs[k] for rp[k-1] <= v < rp[k]
    // pairs-once pass: per batch slot, the sum over the chunk's i of the
    // pair forces on i from that slot's j (j's reaction, negated; one
    // 16-byte LDS store per slot)
    float4 jr[kNewton ? kJB : 1];
};

// Sums over groups of 2^LG consecutive lanes (LG <= 6), left in each
// group's last lane: DPP row shifts within the 16-lane rows (shift 1, 2, 4,
// 8), then row broadcasts across them (lane 15 to the next row, lane 31 to
// the upper half).  Every lane must be active.
template <int K>
__device__ inline float dpp_add(float x) {
    constexpr int ctrl = K < 4 ? 0x110 + (1 << K) : (K == 4 ? 0x142 : 0x143);
    constexpr int rows = K < 4 ? 0xf : (K == 4 ? 0xa : 0xc);
    return x + __int_as_float(
                   __builtin_amdgcn_update_dpp(0, __float_as_int(x), ctrl, rows, 0xf, K < 4));
}

template <int LG>
__device__ inline void group_sum3(float& x, float& y, float& z) {
    if constexpr (LG > 0) {
        group_sum3<LG - 1>(x, y, z);
        x = dpp_add<LG - 1>(x);
        y = dpp_add<LG - 1>(y);
        z = dpp_add<LG - 1>(z);
    }
}

// Signed fixed point (value * scale, rounded to nearest) of a pairs-once
// partial sum, added to a particle's 64-bit accumulator: integer adds
// commute, so the totals do not depend on the order of the atomics.
__device__ inline void fixed_add(unsigned long long* dst, float v, float scale) {
    if (v != 0.0f) atomicAdd(dst, (unsigned long long)__double2ll_rn((double)v * (double)scale));
}

template <int kMode, int kJB, bool kN>
__device__ inline void flush(WaveLDS<kMode, kJB, kN>& L, int lane, int begin, int count,
                             const FraK& Ka, const FraK& Kr) {
    if (lane < count) {
        const int s = (begin + lane) & (kQ - 1);
        const float4 e = L.q[s];
        const int ii = L.qi[s];
        float fx, fy, fz;
        pair_force<kMode>(e, Ka, Kr, fx, fy, fz);
        using Acc = typename WaveLDS<kMode, kJB, kN>::Acc;
        atomicAdd(&L.acc[ii][0], (Acc)fx);
        atomicAdd(&L.acc[ii][1], (Acc)fy);
        atomicAdd(&L.acc[ii][2], (Acc)fz);
    }
    wave_lds_sync();
}

// Wavefront-uniform copies (scalar registers) of values every lane holds.
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline float uni(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

__device__ inline int find_tile2(const PPTile* __restrict__ T, int ntile, uint32_t c) {
    int lo = 0, hi = ntile - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (T[mid].cell_off <= (long long)c) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Wavefront w of the grid takes chunks w, w + W, w + 2W, ... (W = waves in
// the grid); the chunk count is read on the device.  Chunk-level values are
// kept wavefront-uniform (scalar registers), and the register budget is
// capped for 6 wavefronts per SIMD (REF at 256^3, ordered pairs: 80.7 ms;
// budgets for 5 and 7 wavefronts 89.2 and 86.0 ms; 4 blocks per CU instead
// of 6: 105 ms).  The REF variant spills a few registers outside its inner
// loops.
template <int kMode, int kJB, bool kNewton>
__global__ __launch_bounds__(256, 6) void k_pp_tiled(PPArgs A, const int* __restrict__ chunks,
                                                  const int* __restrict__ nchunks_dev,
                                                  int direct_lg, int dense_lg) {
    static_assert(kJB == 64 || kJB == 128, "j batch of 64 or 128");
    constexpr int kH = kJB / 64;  // fetched float4 per lane per batch
    __shared__ WaveLDS<kMode, kJB, kNewton> lds[4];
    const int wave = uni((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    WaveLDS<kMode, kJB, kNewton>& L = lds[wave];
    const int nchunks = uni(*nchunks_dev);

    float f2 = 0.f, v0 = 0.f, v1 = 0.f, v2 = 0.f;
    unsigned long long pc = 0, pa = 0;
    for (int ck = blockIdx.x * 4 + wave; ck < nchunks; ck += gridDim.x * 4) {
        const int a0 = uni(chunks[ck]);
        const uint32_t cid = (uint32_t)uni((int)A.cellid[a0]);
        const int cend = uni(A.cell_start[cid + 1]);
        const int cbeg = uni(A.cell_start[cid]);
        // cells of >= 64 particles are stored in sub-cell order
        // (k_pp_subsort) and taken 2^dense_lg at a time, so that a chunk is
        // spatially compact
        const bool grouped = cend - cbeg >= 64;
        const int ni = min(grouped ? 1 << dense_lg : 64, cend - a0);
        int lgI = 0;
        while ((1 << lgI) < ni) lgI++;
        const int I = 1 << lgI, J = 64 >> lgI;
        const int islot = lane & (I - 1), jslot = lane >> lgI;
        const bool vi = islot < ni;
        const int ia = a0 + (vi ? islot : 0);
        const float4 p = A.pos4[ia];
        const int ti = uni(find_tile2(A.tiles, A.ntile, cid));
        const PPTile T = A.tiles[ti];
        const unsigned long long pc0 = pc, pa0 = pa;
        const long long key = (long long)cid - T.cell_off;
        const int S = T.S;
        const int x = (int)(key % S), y = (int)((key / S) % S), z = (int)(key / ((long long)S * S));
        // one subtile per PP cell: same range for the chunk (w < 0 marks the
        // pairs-once pass's small cells)
        const float range = fabsf(uni(p.w));
        // learned kick: the network stands in for the sparse level-2 cells
        if (A.learned_occ && range == A.learned_range && cend - cbeg < A.learned_occ) continue;
        int R;
        float pp2;
        {
#pragma clang fp contract(off)
            R = (int)ceilf(range * (float)T.rcp / (float)A.ratio_cs);
            pp2 = range * range;
        }
        const FraK Ka = make_frak(A.app), Kr = make_frak(range);
        for (int q = lane; q < 64 * 3; q += 64) (&L.acc[0][0])[q] = 0;
        wave_lds_sync();
        const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
        int qhead = 0, qtail = 0;
        // pairs once: chunks of cells of more than 2^(direct_lg-1) particles
        // (each such cell's chunks alike, k_chunk_flags marks the others)
        const bool direct = kNewton ? cend - cbeg > (1 << direct_lg) / 2 : lgI >= direct_lg;
        const bool nwc = kNewton && direct;
        using Acc = typename WaveLDS<kMode, kJB, kNewton>::Acc;
        // Pairs once (kNewton): a j of the chunk's own range in a cell of
        // more than 2^(direct_lg-1) particles is an active particle of the
        // same tile and level, an i of a chunk of this kind whose stencil
        // holds this chunk's cell; of the two, the one with the lower sorted
        // index evaluates the pair and adds the force to both (the pair force
        // is exactly antisymmetric: RN(a - b) = -RN(b - a)).  Other j (halo
        // copies, other levels, small cells) act on i alone, and the chunks
        // of small cells evaluate all their pairs themselves.  A chunk whose
        // stencil lies inside its own subtile meets only j of its own range,
        // so it skips the rows and cells before its own.
        bool interior = false;
        if (nwc) {
            const int cps = A.ntt * T.rcp, span = S - 2 * T.rcp;
            auto inside = [&](int c) {
                const int lo = c - R - T.rcp, hi = c + R - T.rcp;
                return lo >= 0 && hi < span && lo / cps == hi / cps;
            };
            interior = inside(x) && inside(y) && inside(z);
        }
        Acc dax = 0, day = 0, daz = 0;
        unsigned npa = 0;
        // direct chunks: bounding box of the chunk's i; a j farther than the
        // range from the box cannot pair with any of them (the 1e-4 relative
        // margin keeps every j the per-pair test could accept)
        float bx0 = p.x, bx1 = p.x, by0 = p.y, by1 = p.y, bz0 = p.z, bz1 = p.z;
        if (direct) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                bx0 = fminf(bx0, __shfl_xor(bx0, o, 64));
                bx1 = fmaxf(bx1, __shfl_xor(bx1, o, 64));
                by0 = fminf(by0, __shfl_xor(by0, o, 64));
                by1 = fmaxf(by1, __shfl_xor(by1, o, 64));
                bz0 = fminf(bz0, __shfl_xor(bz0, o, 64));
                bz1 = fmaxf(bz1, __shfl_xor(bz1, o, 64));
            }
            bx0 = uni(bx0);
            bx1 = uni(bx1);
            by0 = uni(by0);
            by1 = uni(by1);
            bz0 = uni(bz0);
            bz1 = uni(bz1);
        }
        const float pp2m = pp2 * 1.0001f;
        // the pair tests' i position: a far-away one for lanes past the chunk
        const float pix = vi ? p.x : -kFar, piy = vi ? p.y : -kFar, piz = vi ? p.z : -kFar;
        const int xlo = max(x - R, 0), xhi = min(x + R, S - 1);
        const int W = 2 * R + 1, nrows = W * W;
        // Neighbour rows (each a contiguous run of sorted j) in groups of 64:
        // the group's runs form one virtual list of M candidates, fetched 64
        // at a time with one coalesced load, prefetched a batch ahead.
        for (int rg = 0; rg < nrows; rg += 64) {
            int js = 0, cnt = 0, cfull = 0;
            const int r = rg + lane;
            if (r < nrows) {
                const int zz = z + r / W - R, yy = y + r % W - R;
                if (zz >= 0 && zz < S && yy >= 0 && yy < S) {
                    const long long row = T.cell_off + ((long long)zz * S + yy) * S;
                    js = A.cell_start[row + xlo];
                    cnt = A.cell_start[row + xhi + 1] - js;
                    cfull = cnt;
                    if (kNewton && interior && zz * S + yy <= z * S + y) {
                        // rows before the chunk's: none; its own row: from its
                        // first i -- unless a small cell lies there, whose j
                        // act on this chunk's i only from here
                        const bool own = zz * S + yy == z * S + y;
                        const int xe = own ? x : xhi + 1, small = (1 << direct_lg) / 2;
                        bool has_small = false;
                        int c0 = js;
                        for (int xc = xlo; xc < xe; xc++) {
                            const int c1 = A.cell_start[row + xc + 1];
                            has_small |= c1 - c0 > 0 && c1 - c0 <= small;
                            c0 = c1;
                        }
                        if (!has_small) {
                            const int j0 = own ? a0 : js + cnt;
                            cnt -= j0 - js;
                            js = j0;
                        }
                    }
                }
            }
            int inc = cnt;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(inc, o, 64);
                if (lane >= o) inc += t;
            }
            const int M = uni(__shfl(inc, 63, 64));
            L.rs[lane] = js - (inc - cnt);
            L.rp[lane] = inc;
            wave_lds_sync();
            if (kNewton && interior) {  // the candidates counted: the whole stencil
                int sf = cfull;
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) sf += __shfl_xor(sf, o, 64);
                if (vi && jslot == 0) pc += (unsigned long long)sf;
            } else if (vi && jslot == 0) {
                pc += (unsigned long long)M;
            }
            // a lane's candidates v = vb + lane only grow, so its row (the
            // first with rp > v) is found by walking forward from the last
            // one: rp of the rows past the group is M > v, so the walk stops
            // inside the group
            int frow = 0;
            auto fetch = [&](int vb) {
                const int v = vb + lane;
                float4 q4 = make_float4(0.f, 0.f, 0.f, 0.f);
                if (v < M) {
                    while (L.rp[frow] <= v) frow++;
                    const int jx = v + L.rs[frow];
                    q4 = A.pos4[jx];
                    // pairs once: w = the sorted index of a j of the chunk's
                    // range, INT_MAX for a j that acts on i alone
                    if (kNewton) q4.w = __int_as_float(q4.w == range ? jx : INT_MAX);
                }
                return q4;
            };
            // a batch is kJB candidates, kH fetches of 64; the next batch's
            // fetches are issued before this batch's pairs run
            float4 qcur[kH];
#pragma unroll
            for (int h = 0; h < kH; h++) qcur[h] = fetch(64 * h);
            for (int vb = 0; vb < M; vb += kJB) {
                int nb = 0;
                if (direct) {  // keep the j near the chunk's box, compacted
#pragma unroll
                    for (int h = 0; h < kH; h++) {
                        const float4 qh = qcur[h];
                        const float ex = fmaxf(fmaxf(bx0 - qh.x, qh.x - bx1), 0.0f);
                        const float ey = fmaxf(fmaxf(by0 - qh.y, qh.y - by1), 0.0f);
                        const float ez = fmaxf(fmaxf(bz0 - qh.z, qh.z - bz1), 0.0f);
                        // (pairs once: a j of the chunk's range before its first
                        // i evaluates the pair itself)
                        const bool keep = vb + 64 * h + lane < M &&
                                          ex * ex + ey * ey + ez * ez < pp2m &&
                                          (!nwc || __float_as_int(qh.w) >= a0);
                        const uint64_t km = __ballot(keep);
                        if (keep) L.jb[nb + __popcll(km & lt)] = qh;
                        nb += __popcll(km);
                    }
                } else {
#pragma unroll
                    for (int h = 0; h < kH; h++) L.jb[64 * h + lane] = qcur[h];
                    nb = min(kJB, M - vb);
                }
                // the J slots past the batch hold far-away j: a lane's last
                // iteration reads them instead of a clamped slot
                if (lane < J) L.jb[nb + lane] = make_float4(kFar, kFar, kFar, 0.0f);
                wave_lds_sync();
#pragma unroll
                for (int h = 0; h < kH; h++)
                    if (vb + kJB + 64 * h < M) qcur[h] = fetch(vb + kJB + 64 * h);
                // (i, j) of slot t: the difference vector and whether it pairs
                // (branch-free: slots past the batch hold far-away j, and an
                // i slot past the chunk tests from a far-away i)
                auto test = [&](int t, float4& e) {
#pragma clang fp contract(off)
                    const float4 qj = L.jb[t];
                    e.x = qj.x - pix;
                    e.y = qj.y - piy;
                    e.z = qj.z - piz;
                    if (kMode == kModePrecise || kMode == kModePlummer)
                        e.w = fmaf(e.z, e.z, fmaf(e.y, e.y, e.x * e.x));
                    else
                        e.w = e.x * e.x + e.y * e.y + e.z * e.z;
                    return (e.w > 0.0f) & (e.w < pp2);
                };
                if (nwc) {
                    // Lane (i, s) takes slot t + s at step t, as below; the I
                    // lanes of slot t + s then hold the forces on all the
                    // chunk's i from that j, and their sum (a DPP reduction,
                    // no LDS traffic) is j's reaction for the batch, which
                    // the group's last lane stores.  (LDS float atomics cost
                    // ~190 cycles per wavefront instruction on gfx950.)
                    // (one loop per chunk size: the reduction's DPP steps
                    // and the group shape are then constants)
                    auto pairs = [&](auto lg_c) {
                        constexpr int LG = decltype(lg_c)::value;
                        constexpr int IC = 1 << LG, JC = 64 >> LG;
                        for (int t = 0; t < nb; t += JC) {
                            float4 e;
                            const int jw = __float_as_int(L.jb[t + jslot].w);
                            float fx = 0.0f, fy = 0.0f, fz = 0.0f;
                            if (test(t + jslot, e) & (jw > ia)) {
#pragma clang fp contract(off)
                                pair_force<kMode>(e, Ka, Kr, fx, fy, fz);
                                dax += (Acc)fx;
                                day += (Acc)fy;
                                daz += (Acc)fz;
                                npa += jw != INT_MAX ? 2u : 1u;
                            }
                            group_sum3<LG>(fx, fy, fz);
                            if ((lane & (IC - 1)) == IC - 1 && t + jslot < nb)
                                L.jr[t + jslot] = make_float4(fx, fy, fz, 0.0f);
                        }
                    };
                    switch (lgI) {
                        case 0: pairs(std::integral_constant<int, 0>{}); break;
                        case 1: pairs(std::integral_constant<int, 1>{}); break;
                        case 2: pairs(std::integral_constant<int, 2>{}); break;
                        case 3: pairs(std::integral_constant<int, 3>{}); break;
                        case 4: pairs(std::integral_constant<int, 4>{}); break;
                        case 5: pairs(std::integral_constant<int, 5>{}); break;
                        default: pairs(std::integral_constant<int, 6>{}); break;
                    }
                    wave_lds_sync();
                    // the reactions of the batch's j of the chunk's range
#pragma unroll
                    for (int h = 0; h < kH; h++) {
                        const int l = 64 * h + lane;
                        if (l < nb) {
                            const int jw = __float_as_int(L.jb[l].w);
                            if (jw != INT_MAX) {
                                const float4 r = L.jr[l];
                                unsigned long long* d = A.react + 3 * (size_t)jw;
                                fixed_add(d, -r.x, A.react_scale);
                                fixed_add(d + 1, -r.y, A.react_scale);
                                fixed_add(d + 2, -r.z, A.react_scale);
                            }
                        }
                    }
                } else
                // two loops on the chunk-uniform branch, so that the direct
                // loop carries only its accumulators
                if (direct) {
                    for (int t = 0; t < nb; t += J) {
                        float4 e;
                        if (test(t + jslot, e)) {
#pragma clang fp contract(off)
                            float fx, fy, fz;
                            pair_force<kMode>(e, Ka, Kr, fx, fy, fz);
                            dax += (Acc)fx;
                            day += (Acc)fy;
                            daz += (Acc)fz;
                            npa++;
                        }
                    }
                } else {
                    for (int t = 0; t < nb; t += J) {
                        float4 e;
                        const bool ok = test(t + jslot, e);
                        const uint64_t m = __ballot(ok);
                        if (m) {
                            if (ok) {
                                const int s = (qtail + __popcll(m & lt)) & (kQ - 1);
                                L.q[s] = e;
                                L.qi[s] = (unsigned char)islot;
                            }
                            qtail += __popcll(m);
                            wave_lds_sync();
                            if (qtail - qhead >= 64) {
                                flush<kMode>(L, lane, qhead, 64, Ka, Kr);
                                qhead += 64;
                            }
                        }
                    }
                }
                wave_lds_sync();  // jb is rewritten by the next batch
            }
        }
        if (direct) {
            if (vi) {
                atomicAdd(&L.acc[islot][0], dax);
                atomicAdd(&L.acc[islot][1], day);
                atomicAdd(&L.acc[islot][2], daz);
            }
            pa += npa;
            wave_lds_sync();
        }
        if (qtail > qhead) flush<kMode>(L, lane, qhead, qtail - qhead, Ka, Kr);
        if (lane == 0) pa += (unsigned long long)qtail;
        if (kNewton && vi && jslot == 0) {
            pc -= 1;  // the particle itself
            // the i's own sums join its reactions; k_pp_newton_terms forms
            // the velocity term once every chunk has added to it
#pragma unroll
            for (int c = 0; c < 3; c++)
                fixed_add(A.react + 3 * (size_t)ia + c, (float)L.acc[islot][c], A.react_scale);
        } else if (vi && jslot == 0) {
#pragma clang fp contract(off)
            pc -= 1;  // the particle itself
            const int a = ia;
            const float ax = (float)L.acc[islot][0], ay = (float)L.acc[islot][1],
                        az = (float)L.acc[islot][2];
            const float mass = A.scal[0], a_mid = A.scal[1], dt = A.scal[2];
            const float a0m = ax * mass * mass, a1m = ay * mass * mass, a2m = az * mass * mass;
            f2 = fmaxf(f2, a0m * a0m + a1m * a1m + a2m * a2m);
            const float acc[3] = {a0m, a1m, a2m};
            if (A.term) {  // split pass: k_pp_apply finishes the update
#pragma unroll
                for (int c = 0; c < 3; c++)
                    A.term[3 * (size_t)A.slot[a] + c] = acc[c] * a_mid * dt / 6.0f;
            } else {
                const size_t ip = (size_t)A.slot[a] * 3;
                float nv[3];
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const double term = (double)(acc[c] * a_mid * dt / 6.0f) / 3.14159265358979323846;
                    double v = (double)A.vp[ip + c];
                    if (A.dv64) v += A.dv64[ip + c];
                    nv[c] = (float)(v + term);
                    A.vp[ip + c] = nv[c];
                }
                v0 = fmaxf(v0, fabsf(nv[0]));
                v1 = fmaxf(v1, fabsf(nv[1]));
                v2 = fmaxf(v2, fabsf(nv[2]));
            }
        }
        if (A.tile_pairs) {  // per-tile counts (on request only)
            const unsigned long long dc = wave_sum_u64(pc - pc0), da = wave_sum_u64(pa - pa0);
            if (lane == 0) {
                atomicAdd(&A.tile_pairs[2 * ti], dc);
                atomicAdd(&A.tile_pairs[2 * ti + 1], da);
            }
        }
        wave_lds_sync();  // acc and the queue are reused by the next chunk
    }
    const float mx[4] = {f2, v0, v1, v2};
    block_max_to<4>(&A.red->f2max_bits, mx);
    const unsigned long long sm[2] = {pc, pa};
    block_sum_to<2>(&A.red->p_cand, sm);
}

bool pp_newton_mode(int mode) { return mode == CUBE2_PP_REF || mode == CUBE2_PP_FAST; }

// Blocks of the pair kernel: 64 per CU, ten times the resident maximum (6),
// unless tuned.  Its wavefronts take chunks by a static stride over the
// list; with a grid far beyond what is resident, the hardware hands out
// blocks as others retire, which evens out the uneven chunk costs (at 256^3
// pairs once, alone: 67.3 ms at 6 blocks per CU, 63.3 at 24, 56.1 at 64,
// 57.4 at 128) and lets the PM3 mesh kernels on the other stream take the
// slots of retiring blocks throughout (overlapped kick: 111.5 ms at 6, 105.3
// at 24, 103.3 at 64).  overlap: the kernel runs beside PM3
// (CUBE2_TUNE_PP_OVERLAP_BLOCKS per CU if set).
int pp_tiled_grid(int mode, bool overlap, bool newton) {
    (void)mode;
    (void)newton;
    constexpr int kBlocksPerCU = 64;
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        ncu = 256;
    ncu = std::max(1, ncu);
    const int forced = ctx().tune[CUBE2_TUNE_PP_BLOCKS_PER_CU];
    const int over = ctx().tune[CUBE2_TUNE_PP_OVERLAP_BLOCKS];
    if (overlap && over) return ncu * over;
    return ncu * (forced ? forced : kBlocksPerCU);
}

template <int kJB>
static void launch_pp_tiled_jb(const PPArgs& a, const int* chunks, const int* nchunks,
                               unsigned blocks, int mode, int direct_lg, int dense_lg,
                               hipStream_t s) {
    const int dl = direct_lg, gl = dense_lg;
    if (a.react && mode == CUBE2_PP_REF)
        k_pp_tiled<kModeRef, kJB, true><<<blocks, 256, 0, s>>>(a, chunks, nchunks, dl, gl);
    else if (a.react && mode == CUBE2_PP_FAST)
        k_pp_tiled<kModeFast, kJB, true><<<blocks, 256, 0, s>>>(a, chunks, nchunks, dl, gl);
    else if (mode == CUBE2_PP_REF)
        k_pp_tiled<kModeRef, kJB, false><<<blocks, 256, 0, s>>>(a, chunks, nchunks, dl, gl);
    else if (mode == CUBE2_PP_FAST)
        k_pp_tiled<kModeFast, kJB, false><<<blocks, 256, 0, s>>>(a, chunks, nchunks, dl, gl);
    else if (mode == CUBE2_PP_PLUMMER)
        k_pp_tiled<kModePlummer, kJB, false><<<blocks, 256, 0, s>>>(a, chunks, nchunks, dl, gl);
    else
        k_pp_tiled<kModePrecise, kJB, false><<<blocks, 256, 0, s>>>(a, chunks, nchunks, dl, gl);
}

int launch_pp_tiled(const PPArgs& a, const int* chunks, const int* nchunks, int grid, int mode,
                    hipStream_t s) {
    const char* name = mode == CUBE2_PP_REF       ? "pp_pairs_ref"
                       : mode == CUBE2_PP_FAST    ? "pp_pairs_fastq"
                       : mode == CUBE2_PP_PLUMMER ? "pp_pairs_plummer"
                                                  : "pp_pairs_precise";
    Stage st(name);
    const unsigned blocks = (unsigned)std::max(1, grid);
    // chunks of more than 2 particles sum directly unless tuned otherwise
    const int direct_lg = pp_direct_lg();
    const int dense_lg = pp_dense_lg(a.react != nullptr);
    if (pp_jbatch() == 128)
        launch_pp_tiled_jb<128>(a, chunks, nchunks, blocks, mode, direct_lg, dense_lg, s);
    else
        launch_pp_tiled_jb<64>(a, chunks, nchunks, blocks, mode, direct_lg, dense_lg, s);
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

// --- per-pair check: the force of each given pair (dx, dy, dz) with the
// given range, by the pair kernel's own pair_force<mode>, with its
// acceptance test (0 < r^2 < range^2; others give 0).
template <int kMode>
__global__ void k_pair_forces(const float4* __restrict__ in, int n, float app,
                              float* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    float4 e = make_float4(0.f, 0.f, 0.f, 0.f);
    float range = 0.0f;
    if (t < n) {
        const float4 q = in[t];
        e.x = q.x;
        e.y = q.y;
        e.z = q.z;
        range = q.w;
    }
    bool ok;
    {
#pragma clang fp contract(off)
        if (kMode == kModePrecise || kMode == kModePlummer)
            e.w = fmaf(e.z, e.z, fmaf(e.y, e.y, e.x * e.x));
        else
            e.w = e.x * e.x + e.y * e.y + e.z * e.z;
        ok = t < n && e.w > 0.0f && e.w < range * range;
    }
    float fx = 0.f, fy = 0.f, fz = 0.f;
    if (ok) {
        const FraK Ka = make_frak(app), Kr = make_frak(range);
        pair_force<kMode>(e, Ka, Kr, fx, fy, fz);
    }
    if (t < n) {
        out[3 * t] = fx;
        out[3 * t + 1] = fy;
        out[3 * t + 2] = fz;
    }
}

int pair_forces_debug(int mode, const float* pairs, int n, float app, float* forces) {
    C2_TRY(ensure_init());
    if (n < 0 || (n > 0 && (!pairs || !forces))) {
        set_error("pair check: bad arguments");
        return CUBE2_ERR_ARG;
    }
    if (mode != CUBE2_PP_REF && mode != CUBE2_PP_FAST && mode != CUBE2_PP_PRECISE &&
        mode != CUBE2_PP_PLUMMER) {
        set_error("pair check: mode %d has no pair_force", mode);
        return CUBE2_ERR_ARG;
    }
    if (n == 0) return CUBE2_OK;
    hipStream_t s = ctx().stream;
    float4* d_in = nullptr;
    float* d_out = nullptr;
    C2_HIP(hipMalloc(&d_in, (size_t)n * sizeof(float4)));
    if (hipMalloc(&d_out, (size_t)n * 3 * sizeof(float)) != hipSuccess) {
        (void)hipFree(d_in);
        set_error("pair check: out of device memory");
        return CUBE2_ERR_HIP;
    }
    hipError_t e = hipMemcpyAsync(d_in, pairs, (size_t)n * sizeof(float4), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        const unsigned blocks = ceil_div((size_t)n, 256);
        if (mode == CUBE2_PP_REF) k_pair_forces<kModeRef><<<blocks, 256, 0, s>>>(d_in, n, app, d_out);
        else if (mode == CUBE2_PP_FAST) k_pair_forces<kModeFast><<<blocks, 256, 0, s>>>(d_in, n, app, d_out);
        else if (mode == CUBE2_PP_PLUMMER)
            k_pair_forces<kModePlummer><<<blocks, 256, 0, s>>>(d_in, n, app, d_out);
        else k_pair_forces<kModePrecise><<<blocks, 256, 0, s>>>(d_in, n, app, d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(forces, d_out, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    C2_HIP(e);
    return CUBE2_OK;
}

// --- arithmetic self-test: the reciprocal-based quotients of the
// reference-arithmetic kernel against IEEE float division.
// seed != 0: n seeded random operands (a, b with exponents in [-40, 40],
// r and the length l in [-20, 20]); seed == 0: every float with exponent in
// [-40, 40] (n = 81 * 2^23) as b and r, with a = l = 1.  Checked: a/b,
// 12/b, |1 - b*recip64(b)| <= 2^-51, 1/RN(r*r) refined from (1/r)^2, and
// 12/e2 for ep = RN(2r/l), e2 = RN(ep*ep) refined from (l/2)^2 (1/r)^2, and
// sqrt_rn(b) against IEEE sqrtf.
__global__ void k_selftest_div(uint64_t seed, int n, unsigned long long* bad) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint64_t s = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(t + 1);
    auto nxt = [&]() {
        s ^= s >> 12;
        s ^= s << 25;
        s ^= s >> 27;
        return (uint32_t)((s * 0x2545F4914F6CDD1Dull) >> 32);
    };
    auto mk = [](int e, uint32_t m) { return __uint_as_float(((uint32_t)(e + 127) << 23) | (m & 0x7FFFFF)); };
    float a, b, r, l;
    if (seed == 0) {
        b = mk((t >> 23) - 40, (uint32_t)t);
        r = b;
        a = 1.0f;
        l = 1.0f;
    } else {
        a = mk((int)(nxt() % 81) - 40, nxt());
        b = mk((int)(nxt() % 81) - 40, nxt());
        r = mk((int)(nxt() % 41) - 20, nxt());
        l = mk((int)(nxt() % 41) - 20, nxt());
    }
    float ieee, ieee12, ieee_rr, ieee_t12, p, ep, e2, ieee_sqrt;
    {
#pragma clang fp contract(off)
        ieee_sqrt = sqrtf(b);
        ieee = a / b;
        ieee12 = 12.0f / b;
        p = r * r;
        ieee_rr = 1.0f / p;
        ep = 2.0f * r / l;
        e2 = ep * ep;
        ieee_t12 = 12.0f / e2;
    }
    const double yb = recip64((double)b);
    const float mine = div_by(a, yb);
    const float twelve = (float)(12.0 * yb);
    const bool resid = fabs(fma(-(double)b, yb, 1.0)) <= 0x1p-51;
    const double inv_r = recip64((double)r);
    const float rr = (float)recip_refine((double)p, inv_r * inv_r);
    const float ep2 = div_by(2.0f * r, 1.0 / (double)l);
    const float t12 =
        (float)(12.0 * recip_refine((double)e2, 0.25 * (double)l * (double)l * (inv_r * inv_r)));
    if (__float_as_uint(ieee) != __float_as_uint(mine) ||
        __float_as_uint(ieee12) != __float_as_uint(twelve) || !resid ||
        __float_as_uint(ieee_rr) != __float_as_uint(rr) || __float_as_uint(ep) != __float_as_uint(ep2) ||
        __float_as_uint(ieee_t12) != __float_as_uint(t12) ||
        __float_as_uint(ieee_sqrt) != __float_as_uint(sqrt_rn(b)))
        atomicAdd(bad, 1ull);
}

int selftest_div(uint64_t seed, int n, unsigned long long* mismatches) {
    C2_TRY(ensure_init());
    hipStream_t s = ctx().stream;
    unsigned long long* d = nullptr;
    C2_HIP(hipMalloc(&d, sizeof(unsigned long long)));
    C2_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long), s));
    k_selftest_div<<<ceil_div((size_t)n, 256), 256, 0, s>>>(seed, n, d);
    C2_HIP(hipGetLastError());
    C2_HIP(hipMemcpyAsync(mismatches, d, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    C2_HIP(hipStreamSynchronize(s));
    C2_HIP(hipFree(d));
    return CUBE2_OK;
}

}  // namespace cube2
