// Batched PM3 FFTs, hand-written for gfx950.
//
// One pass per axis; every pass stages a block of pencils in LDS and runs a
// mixed-radix (4, 2, 3, 5, 7) self-sorting Stockham transform there, in place
// with register staging, so HBM sees one read and one write of the mesh per
// pass.
//   x, forward : two real rows y, y+1 form one complex line z = a + i b;
//                FFT(z) is split into the two half spectra A, B (R2C).
//   y          : tiles of C consecutive kx columns x n rows.
//   z          : as y; mode 2 transforms forward, multiplies by the Green's
//                table and transforms back while the pencils stay in LDS, so
//                the Green's multiply (kick.f90:479-490) costs no HBM pass.
//   x, inverse : Z = A + i B over the Hermitian extension, one inverse FFT
//                gives both real rows (C2R); imaginary parts of the kx=0 and
//                kx=n/2 terms are ignored, as a C2R transform does.
// Unnormalised in both directions (FFTW/cuFFT convention, kick.f90:552).
#include <cmath>
#include <map>
#include <type_traits>

#include "fft.h"
#include "profile.h"

namespace cube2 {

constexpr int kFFTThreads = 256;
constexpr int kFFTElems = 2048;  // complex elements per LDS tile (measured best: 2048 > 4096 > 1024)
constexpr int kColsMax = 4096;   // largest LDS tile of the y/z passes (wider minimum tiles)

constexpr int lg_floor_c(int v) {
    int l = 0;
    while ((2 << l) <= v) l++;
    return l;
}
// The built-in shapes of a size (the host code applies the same rules):
// log2 of the row pairs per block of the x passes, and of the column-tile
// width of the y/z passes (about kFFTElems LDS elements per tile, no wider
// than the half spectrum, within kColsMax).
constexpr int x_lgl(int n) { return lg_floor_c(kFFTElems / n > 1 ? kFFTElems / n : 1); }
constexpr int cols_lgc_builtin(int n) {
    int lg = x_lgl(n);
    while (lg > 0 && (1 << lg) > 2 * (n / 2 + 1)) lg--;
    while (lg > 0 && (1 << lg) * n > kColsMax) lg--;
    return lg;
}

__device__ inline float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// i*S*z for S = +-1
template <int S>
__device__ inline float2 mul_si(float2 z) {
    return S < 0 ? make_float2(z.y, -z.x) : make_float2(-z.y, z.x);
}

template <int R, int S>
struct DFT;

template <int S>
struct DFT<2, S> {
    __device__ static void run(float2* v) {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <int S>
struct DFT<4, S> {
    __device__ static void run(float2* v) {
        float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        float2 t2 = cadd(v[1], v[3]), t3 = mul_si<S>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

template <int S>
struct DFT<3, S> {
    __device__ static void run(float2* v) {
        const float c = 0.86602540378443864676f;  // sin(2 pi / 3)
        float2 s1 = cadd(v[1], v[2]), d1 = csub(v[1], v[2]);
        float2 m = make_float2(v[0].x - 0.5f * s1.x, v[0].y - 0.5f * s1.y);
        float2 t = mul_si<S>(make_float2(c * d1.x, c * d1.y));
        v[0] = cadd(v[0], s1);
        v[1] = cadd(m, t);
        v[2] = csub(m, t);
    }
};

template <int S>
struct DFT<5, S> {
    __device__ static void run(float2* v) {
        const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
        const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
        float2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
        float2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
        float2 x0 = v[0];
        float2 m1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
        float2 m2 = make_float2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
        float2 n1 = mul_si<S>(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
        float2 n2 = mul_si<S>(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
        v[0] = make_float2(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
        v[1] = cadd(m1, n1);
        v[4] = csub(m1, n1);
        v[2] = cadd(m2, n2);
        v[3] = csub(m2, n2);
    }
};

// Seven points: pairs a_m = x_m + x_{7-m}, b_m = x_m - x_{7-m} (m = 1..3);
// X_k = x_0 + C_k + i S s_k and X_{7-k} = x_0 + C_k - i S s_k with
// C_k = sum_m a_m cos(2 pi m k / 7), s_k = sum_m b_m sin(2 pi m k / 7).
template <int S>
struct DFT<7, S> {
    __device__ static void run(float2* v) {
        const float c1 = 0.62348980185873353053f, c2 = -0.22252093395631440429f,
                    c3 = -0.90096886790241912624f;
        const float s1 = 0.78183148246802980871f, s2 = 0.97492791218182360702f,
                    s3 = 0.43388373911755812048f;
        const float2 x0 = v[0];
        const float2 a1 = cadd(v[1], v[6]), b1 = csub(v[1], v[6]);
        const float2 a2 = cadd(v[2], v[5]), b2 = csub(v[2], v[5]);
        const float2 a3 = cadd(v[3], v[4]), b3 = csub(v[3], v[4]);
        auto lin = [](float p, float2 x, float q, float2 y, float r, float2 z) {
            return make_float2(p * x.x + q * y.x + r * z.x, p * x.y + q * y.y + r * z.y);
        };
        const float2 C1 = cadd(x0, lin(c1, a1, c2, a2, c3, a3));
        const float2 C2 = cadd(x0, lin(c2, a1, c3, a2, c1, a3));
        const float2 C3 = cadd(x0, lin(c3, a1, c1, a2, c2, a3));
        const float2 S1 = mul_si<S>(lin(s1, b1, s2, b2, s3, b3));
        const float2 S2 = mul_si<S>(lin(s2, b1, -s3, b2, -s1, b3));
        const float2 S3 = mul_si<S>(lin(s3, b1, -s1, b2, s2, b3));
        v[0] = make_float2(x0.x + a1.x + a2.x + a3.x, x0.y + a1.y + a2.y + a3.y);
        v[1] = cadd(C1, S1);
        v[6] = csub(C1, S1);
        v[2] = cadd(C2, S2);
        v[5] = csub(C2, S2);
        v[3] = cadd(C3, S3);
        v[4] = csub(C3, S3);
    }
};

// j mod p for 0 <= j < 2^22 via a float reciprocal and one correction.
__device__ inline int mod_p(int j, int p, float invp) {
    int q = (int)((float)j * invp);
    int r = j - q * p;
    r += (r < 0) ? p : 0;
    r -= (r >= p) ? p : 0;
    return r;
}

// One Stockham stage over L = 2^lgL lines held in one LDS buffer (element e
// of line l at l*ls + e*es), in place: every thread first reads the inputs
// of all its butterflies into registers, the block synchronises, then the
// outputs are written.  Consecutive threads take consecutive lines.
// N, PST > 0: the size and the stage's stride p are compile-time constants
// (the kernels specialised for the meshes of the reference geometries), so
// that every index of the stage is shifts and constant multiplies.
template <int R, int S, int CAP, int N = 0, int PST = 0>
__device__ inline void stage(float2* __restrict__ a, const float2* __restrict__ tw, int n_rt,
                             int p_rt, float invp, int lgL, int ls, int es) {
    constexpr int KMAX = (CAP / R + kFFTThreads - 1) / kFFTThreads;
    const int n = N ? N : n_rt, p = PST ? PST : p_rt;
    const int m = n / R, ts = n / (p * R);
    const int L = 1 << lgL;
    const int tasks = L * m;
    float2 v[KMAX][R];
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
        const int task = threadIdx.x + k * kFFTThreads;
        if (task < tasks) {
            const int l = task & (L - 1), j = task >> lgL;
#pragma unroll
            for (int q = 0; q < R; q++) v[k][q] = a[l * ls + (j + q * m) * es];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
        const int task = threadIdx.x + k * kFFTThreads;
        if (task < tasks) {
            const int l = task & (L - 1), j = task >> lgL;
            // p is the product of the earlier radices: a power of two for
            // every stage before the first radix 3, 5 or 7
            const int kk = PST             ? (int)((unsigned)j % (unsigned)PST)
                           : (p & (p - 1)) == 0 ? (j & (p - 1))
                                                : mod_p(j, p, invp);
            // a specialised stage after the first multiplies every task by
            // its twiddles, kk = 0 (w = 1) included: no branch to merge
            if (PST ? PST > 1 : kk != 0) {
#pragma unroll
                for (int q = 1; q < R; q++) {
                    float2 w = tw[q * kk * ts];
                    if (S > 0) w.y = -w.y;
                    v[k][q] = cmul(v[k][q], w);
                }
            }
            DFT<R, S>::run(v[k]);
            const int base = (j - kk) * R + kk;
#pragma unroll
            for (int q = 0; q < R; q++) a[l * ls + (base + q * p) * es] = v[k][q];
        }
    }
    __syncthreads();
}

// Transforms 2^lgL lines (at most CAP elements) held in LDS buffer a in
// place.  Ends with a barrier.
// R7: the radix-7 butterfly is compiled in only for sizes with a factor 7
// (its registers would lower the occupancy of every other size).
template <int S, int CAP = kFFTElems, bool R7 = false>
__device__ void fft_lds(float2* a, const float2* tw, const FFTPlan& P, int lgL, int ls, int es) {
    for (int s = 0; s < P.nst; s++) {
        const int r = P.rad[s], p = P.pst[s];
        const float ip = P.invp[s];
        if (r == 4) stage<4, S, CAP>(a, tw, P.n, p, ip, lgL, ls, es);
        else if (r == 2) stage<2, S, CAP>(a, tw, P.n, p, ip, lgL, ls, es);
        else if (r == 3) stage<3, S, CAP>(a, tw, P.n, p, ip, lgL, ls, es);
        else if (r == 5) stage<5, S, CAP>(a, tw, P.n, p, ip, lgL, ls, es);
        else if constexpr (R7) stage<7, S, CAP>(a, tw, P.n, p, ip, lgL, ls, es);
    }
}

// Compile-time plan of an FFT size: the radices in fft_plan's order (4, 2,
// 3, 5, 7) and each stage's stride.
struct CPlan {
    int nst = 0;
    int rad[16] = {};
    int pst[16] = {};
};
constexpr CPlan make_cplan(int n) {
    CPlan P;
    const int radices[5] = {4, 2, 3, 5, 7};
    int m = n;
    for (int k = 0; k < 5; k++)
        while (m % radices[k] == 0 && P.nst < 16) {
            P.rad[P.nst++] = radices[k];
            m /= radices[k];
        }
    int p = 1;
    for (int st = 0; st < P.nst; st++) {
        P.pst[st] = p;
        p *= P.rad[st];
    }
    return P;
}

// fft_lds for a size known at compile time (every stage's constants folded).
template <int S, int N, int CAP, int ST = 0>
__device__ __forceinline__ void fft_lds_c(float2* a, const float2* tw, int lgL, int ls, int es) {
    constexpr CPlan P = make_cplan(N);
    if constexpr (ST < P.nst) {
        constexpr int r = P.rad[ST], p = P.pst[ST];
        static_assert(r == 2 || r == 3 || r == 4 || r == 5, "no radix 7 in the specialised sizes");
        stage<r, S, CAP, N, p>(a, tw, N, p, 1.0f / (float)p, lgL, ls, es);
        fft_lds_c<S, N, CAP, ST + 1>(a, tw, lgL, ls, es);
    }
}

// The generic transform (N = 0) or the specialised one.
template <int S, int CAP, bool R7, int N>
__device__ __forceinline__ void fft_run(float2* a, const float2* tw, const FFTPlan& P, int lgL, int ls,
                               int es) {
    if constexpr (N == 0)
        fft_lds<S, CAP, R7>(a, tw, P, lgL, ls, es);
    else
        fft_lds_c<S, N, CAP>(a, tw, lgL, ls, es);
}

__device__ inline void load_twiddles(float2* tw, const FFTPlan& P) {
    for (int t = threadIdx.x; t < P.n; t += blockDim.x) tw[t] = P.tw[t];
}

// Row pair g of the batch: first float of row y = 2*(g % half) of plane
// z = (g / half) % n of job g / (n*half).
__device__ inline float* row_pair(float* d, int g, int n, size_t rs) {
    const int half = n >> 1;
    const int per_job = n * half;
    const int job = g / per_job;
    const int rem = g - job * per_job;
    const int z = rem / half, yp = rem - z * half;
    return d + (size_t)job * rs + ((size_t)z * n + 2 * yp) * (n + 2);
}

// Lines are row pairs; 2^lgL of them per block, one wavefront per line at a
// time along the row (coalesced).
template <bool R7, int N = 0>
__global__ __launch_bounds__(kFFTThreads) void k_fft_x_r2c(float* __restrict__ d, FFTPlan P,
                                                           int npairs, size_t rs, int lgL_arg,
                                                           const float* __restrict__ scale_dev) {
    const float scale = scale_dev ? *scale_dev : 1.0f;
    extern __shared__ float2 sm[];
    const int n = N ? N : P.n, nh = n / 2 + 1;
    const int lgL = N ? x_lgl(N) : lgL_arg;
    const int L = 1 << lgL, es = L + 1;  // padded: the lanes of a row hit distinct banks
    float2* tw = sm;
    float2* A = sm + n;
    load_twiddles(tw, P);
    const int g0 = blockIdx.x * L;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll 2
    for (int l = wave; l < L; l += kFFTThreads / 64) {
        if (g0 + l < npairs) {
            const float* ra = row_pair(d, g0 + l, n, rs);
            for (int t = lane; t < n; t += 64)
                A[l + t * es] = make_float2(ra[t] * scale, ra[t + n + 2] * scale);
        }
    }
    __syncthreads();
    fft_run<-1, kFFTElems, R7, N>(A, tw, P, lgL, 1, es);
    for (int l = wave; l < L; l += kFFTThreads / 64) {
        if (g0 + l >= npairs) break;
        float2* ra = reinterpret_cast<float2*>(row_pair(d, g0 + l, n, rs));
        float2* rb = ra + nh;
        for (int k = lane; k < nh; k += 64) {
            const float2 zk = A[l + k * es], zn = A[l + (k == 0 ? 0 : n - k) * es];
            ra[k] = make_float2(0.5f * (zk.x + zn.x), 0.5f * (zk.y - zn.y));
            rb[k] = make_float2(0.5f * (zk.y + zn.y), -0.5f * (zk.x - zn.x));
        }
    }
}

// Row pair g of a window of nz planes from z0 and nyp row pairs from yp0.
__device__ inline float* row_pair_w(float* d, int g, int n, size_t rs, FFTWin w) {
    const int per_job = w.nz * w.nyp;
    const int job = g / per_job;
    const int rem = g - job * per_job;
    const int z = w.z0 + rem / w.nyp, yp = w.yp0 + rem % w.nyp;
    return d + (size_t)job * rs + ((size_t)z * n + 2 * yp) * (n + 2);
}

template <bool R7, int N = 0>
__global__ __launch_bounds__(kFFTThreads) void k_fft_x_c2r(float* __restrict__ d, FFTPlan P,
                                                           int npairs, size_t rs, int lgL_arg,
                                                           FFTWin w) {
    extern __shared__ float2 sm[];
    const int n = N ? N : P.n, nh = n / 2 + 1, half = n / 2;
    const int lgL = N ? x_lgl(N) : lgL_arg;
    const int L = 1 << lgL, es = L + 1;  // padded: the lanes of a row hit distinct banks
    float2* tw = sm;
    float2* A = sm + n;
    load_twiddles(tw, P);
    const int g0 = blockIdx.x * L;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll 2
    for (int l = wave; l < L; l += kFFTThreads / 64) {
        if (g0 + l >= npairs) continue;
        const float2* ra = reinterpret_cast<const float2*>(row_pair_w(d, g0 + l, n, rs, w));
        const float2* rb = ra + nh;
        for (int k = lane; k < n; k += 64) {
            float2 a, b;
            if (k <= half) {
                a = ra[k];
                b = rb[k];
            } else {
                a = ra[n - k];
                b = rb[n - k];
                a.y = -a.y;
                b.y = -b.y;
            }
            if (k == 0 || k == half) {
                a.y = 0.0f;
                b.y = 0.0f;
            }
            A[l + k * es] = make_float2(a.x - b.y, a.y + b.x);
        }
    }
    __syncthreads();
    fft_run<1, kFFTElems, R7, N>(A, tw, P, lgL, 1, es);
    for (int l = wave; l < L; l += kFFTThreads / 64) {
        if (g0 + l >= npairs) break;
        float* ra = row_pair_w(d, g0 + l, n, rs, w);
        for (int t = lane; t < n; t += 64) {
            const float2 v = A[l + t * es];
            ra[t] = v.x;
            ra[t + n + 2] = v.y;
        }
    }
}

// Tiles of C = 2^lgC consecutive kx columns x n rows along y (axis 1) or z
// (axis 2); LDS element (row i, column c) at i*C + c.
template <int MODE, int CAP, bool R7, int N = 0>
__global__ __launch_bounds__(kFFTThreads) void k_fft_cols(float* __restrict__ dreal, FFTPlan P,
                                                          size_t rs, int axis, int lgC_arg,
                                                          const float* __restrict__ green, int o0,
                                                          int ilo, int ihi) {
    extern __shared__ float2 sm[];
    const int n = N ? N : P.n, nh = n / 2 + 1;
    // a specialised size runs with its built-in tile width (the host checks)
    const int lgC = N ? cols_lgc_builtin(N) : lgC_arg;
    const int C = 1 << lgC;
    const int ntiles = (nh + C - 1) >> lgC;
    float2* tw = sm;
    float2* A = sm + n;
    load_twiddles(tw, P);
    // neighbouring tiles of a plane or row share cache lines: they run on one
    // XCD
    const unsigned lb = xcd_remap(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int bx = (int)(lb % gridDim.x);
    const int tile = bx % ntiles;
    const int o = o0 + bx / ntiles;  // plane (axis 1) or row (axis 2)
    const int job = (int)(lb / gridDim.x);
    const int kx0 = tile * C, Cw = min(C, nh - kx0);
    float2* d = reinterpret_cast<float2*>(dreal + (size_t)job * rs);
    const size_t rowstride = axis == 1 ? (size_t)nh : (size_t)n * nh;
    const size_t obase = (axis == 1 ? (size_t)o * n * nh : (size_t)o * nh) + kx0;
    for (int e = threadIdx.x; e < n * C; e += kFFTThreads) {
        const int i = e >> lgC, c = e & (C - 1);
        if (c < Cw) A[e] = d[obase + i * rowstride + c];
    }
    __syncthreads();
    if (MODE == 1) {
        fft_run<1, CAP, R7, N>(A, tw, P, lgC, 1, C);
    } else {
        fft_run<-1, CAP, R7, N>(A, tw, P, lgC, 1, C);
        if (MODE == 2) {
            // Green's table in tile order: [y][tile][z][c], contiguous per tile
            const float* gt = green + ((size_t)o * ntiles + tile) * n * C;
            for (int e = threadIdx.x; e < n * C; e += kFFTThreads) {
                const int c = e & (C - 1);
                if (c < Cw) {
                    const float gk = gt[e];
                    A[e] = make_float2(A[e].x * gk, A[e].y * gk);
                }
            }
            __syncthreads();
            fft_run<1, CAP, R7, N>(A, tw, P, lgC, 1, C);
        }
    }
    // rows ilo .. ihi-1 go back (an inverse z pass stores only the planes
    // the later passes read)
    for (int e = (ilo << lgC) + threadIdx.x; e < (ihi << lgC); e += kFFTThreads) {
        const int i = e >> lgC, c = e & (C - 1);
        if (c < Cw) d[obase + i * rowstride + c] = A[e];
    }
}

// Green's table G[z][y][kx] (z, y < n, kx < n/2+1) to the z pass's tile
// order T[y][tile][z][c], kx = tile*C + c, times 1/n^3 (the normalisation of
// the unnormalised transform pair, kick.f90:552, applied once here instead
// of per mesh point); columns past n/2 are zero.
__global__ void k_green_tile(const float* __restrict__ G, float* __restrict__ T, int n, int lgC,
                             int ntiles) {
    const int nh = n / 2 + 1, C = 1 << lgC;
    const size_t tot = (size_t)n * ntiles * n * C;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= tot) return;
    const int c = (int)(q & (C - 1));
    size_t r = q >> lgC;
    const int z = (int)(r % n);
    r /= n;
    const int tile = (int)(r % ntiles);
    const int y = (int)(r / ntiles);
    const int kx = tile * C + c;
    const double inv_n3 = 1.0 / ((double)n * n * n);
    T[q] = kx < nh ? (float)((double)G[((size_t)z * n + y) * nh + kx] * inv_n3) : 0.0f;
}

// ------------------------------------------------------------- host side

static std::map<int, FFTPlan> g_plans;

// Sizes whose kernels are specialised at compile time (fft_lds_c): the PM3
// meshes of the 256^3 geometry (levels 2..6) and of the 64^3 one; the others
// run the generic kernels.  f is called with std::integral_constant<int, N>,
// N = 0 for the generic kernels.
template <class F>
static void with_size(int n, bool has7, F&& f) {
    if (!has7) switch (n) {
            case 80: return f(std::integral_constant<int, 80>{});
            case 144: return f(std::integral_constant<int, 144>{});
            case 160: return f(std::integral_constant<int, 160>{});
            case 240: return f(std::integral_constant<int, 240>{});
            case 320: return f(std::integral_constant<int, 320>{});
            case 480: return f(std::integral_constant<int, 480>{});
            default: break;
        }
    f(std::integral_constant<int, 0>{});
}

template <int N>
static int set_lds_attributes_n() {
    const int bytes = (kFFTElems + 2048) * (int)sizeof(float2);
    const int bytes_cols = (kColsMax + 1024) * (int)sizeof(float2);
    const hipFuncAttribute a = hipFuncAttributeMaxDynamicSharedMemorySize;
    C2_HIP(hipFuncSetAttribute((const void*)k_fft_x_r2c<false, N>, a, bytes));
    C2_HIP(hipFuncSetAttribute((const void*)k_fft_x_c2r<false, N>, a, bytes));
    C2_HIP(hipFuncSetAttribute((const void*)k_fft_cols<0, kFFTElems, false, N>, a, bytes_cols));
    C2_HIP(hipFuncSetAttribute((const void*)k_fft_cols<1, kFFTElems, false, N>, a, bytes_cols));
    C2_HIP(hipFuncSetAttribute((const void*)k_fft_cols<2, kFFTElems, false, N>, a, bytes_cols));
    return CUBE2_OK;
}
static bool g_attr_set = false;

template <bool R7>
static int set_lds_attributes_r() {
    // x passes: n twiddles + n (L+1) <= n + kFFTElems + n elements
    const int bytes = (kFFTElems + 2048) * (int)sizeof(float2);
    const int bytes_cols = (kColsMax + 1024) * (int)sizeof(float2);
    const hipFuncAttribute a = hipFuncAttributeMaxDynamicSharedMemorySize;
    C2_HIP(hipFuncSetAttribute((const void*)k_fft_x_r2c<R7>, a, bytes));
    C2_HIP(hipFuncSetAttribute((const void*)k_fft_x_c2r<R7>, a, bytes));
    C2_HIP(hipFuncSetAttribute((const void*)k_fft_cols<0, kFFTElems, R7>, a, bytes_cols));
    C2_HIP(hipFuncSetAttribute((const void*)k_fft_cols<1, kFFTElems, R7>, a, bytes_cols));
    C2_HIP(hipFuncSetAttribute((const void*)k_fft_cols<2, kFFTElems, R7>, a, bytes_cols));
    C2_HIP(hipFuncSetAttribute((const void*)k_fft_cols<0, kColsMax, R7>, a, bytes_cols));
    C2_HIP(hipFuncSetAttribute((const void*)k_fft_cols<1, kColsMax, R7>, a, bytes_cols));
    C2_HIP(hipFuncSetAttribute((const void*)k_fft_cols<2, kColsMax, R7>, a, bytes_cols));
    return CUBE2_OK;
}

static int set_lds_attributes() {
    if (g_attr_set) return CUBE2_OK;
    // x passes: n twiddles + n (L+1) <= n + kFFTElems + n elements
    C2_TRY(set_lds_attributes_r<false>());
    C2_TRY(set_lds_attributes_r<true>());
    C2_TRY(set_lds_attributes_n<80>());
    C2_TRY(set_lds_attributes_n<144>());
    C2_TRY(set_lds_attributes_n<160>());
    C2_TRY(set_lds_attributes_n<240>());
    C2_TRY(set_lds_attributes_n<320>());
    C2_TRY(set_lds_attributes_n<480>());
    g_attr_set = true;
    return CUBE2_OK;
}

int fft_plan(int n, const FFTPlan** out) {
    auto it = g_plans.find(n);
    if (it != g_plans.end()) {
        *out = &it->second;
        return CUBE2_OK;
    }
    if (n < 2 || n % 2 || n > 1024) {
        set_error("FFT size %d unsupported (even, <= 1024)", n);
        return CUBE2_ERR_CAPACITY;
    }
    FFTPlan P;
    P.n = n;
    int m = n;
    for (int r : {4, 2, 3, 5, 7})
        while (m % r == 0) {
            if (P.nst == 16) break;
            P.rad[P.nst++] = r;
            P.has7 |= r == 7;
            m /= r;
        }
    if (m != 1) {
        set_error("FFT size %d has a prime factor other than 2, 3, 5, 7", n);
        return CUBE2_ERR_CAPACITY;
    }
    for (int st = 0, p = 1; st < P.nst; st++) {
        P.pst[st] = p;
        P.invp[st] = 1.0f / (float)p;
        p *= P.rad[st];
    }
    std::vector<float2> tw(n);
    for (int t = 0; t < n; t++) {
        const double ang = 2.0 * M_PI * (double)t / (double)n;
        tw[t] = make_float2((float)std::cos(ang), (float)-std::sin(ang));
    }
    C2_HIP(hipMalloc(&P.tw, n * sizeof(float2)));
    C2_HIP(hipMemcpy(P.tw, tw.data(), n * sizeof(float2), hipMemcpyHostToDevice));
    C2_TRY(set_lds_attributes());
    g_plans[n] = P;
    *out = &g_plans[n];
    return CUBE2_OK;
}

void fft_release() {
    buffers_changed();
    for (auto& kv : g_plans) (void)hipFree(kv.second.tw);
    g_plans.clear();
}

static int lg_floor(int v) {
    int l = 0;
    while ((2 << l) <= v) l++;
    return l;
}

int fft_x_r2c(float* d, const FFTPlan& P, int nb, size_t rs, const float* scale, hipStream_t s) {
    const int n = P.n;
    const int lgL = x_lgl(n);  // the specialised kernels assume it
    const int L = 1 << lgL;
    const int64_t npairs = (int64_t)nb * n * (n / 2);
    if (npairs >= (int64_t)1 << 31) {
        set_error("FFT batch too large");
        return CUBE2_ERR_CAPACITY;
    }
    const size_t lds = (size_t)(n + (L + 1) * n) * sizeof(float2);
    const unsigned grid = (unsigned)((npairs + L - 1) / L);
    with_size(n, P.has7, [&](auto c) {
        constexpr int N = decltype(c)::value;
        if (N == 0 && P.has7)
            k_fft_x_r2c<true><<<grid, kFFTThreads, lds, s>>>(d, P, (int)npairs, rs, lgL, scale);
        else
            k_fft_x_r2c<false, N><<<grid, kFFTThreads, lds, s>>>(d, P, (int)npairs, rs, lgL, scale);
    });
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

int fft_x_c2r(float* d, const FFTPlan& P, int nb, size_t rs, hipStream_t s, const FFTWin* win) {
    const int n = P.n;
    const int lgL = x_lgl(n);  // the specialised kernels assume it
    const int L = 1 << lgL;
    const FFTWin w = win ? *win : FFTWin{0, n, 0, n / 2};
    if (w.z0 < 0 || w.nz < 1 || w.z0 + w.nz > n || w.yp0 < 0 || w.nyp < 1 || w.yp0 + w.nyp > n / 2) {
        set_error("FFT window outside the mesh");
        return CUBE2_ERR_ARG;
    }
    const int64_t npairs = (int64_t)nb * w.nz * w.nyp;
    if (npairs >= (int64_t)1 << 31) {
        set_error("FFT batch too large");
        return CUBE2_ERR_CAPACITY;
    }
    const size_t lds = (size_t)(n + (L + 1) * n) * sizeof(float2);
    const unsigned grid = (unsigned)((npairs + L - 1) / L);
    with_size(n, P.has7, [&](auto c) {
        constexpr int N = decltype(c)::value;
        if (N == 0 && P.has7)
            k_fft_x_c2r<true><<<grid, kFFTThreads, lds, s>>>(d, P, (int)npairs, rs, lgL, w);
        else
            k_fft_x_c2r<false, N><<<grid, kFFTThreads, lds, s>>>(d, P, (int)npairs, rs, lgL, w);
    });
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

// Column tile width of the y/z passes: about kFFTElems LDS elements per tile
// (the FFT_MIN_LGC tunable raises the minimum width; 16 columns measured
// slower at n = 320 and 480 for the lower occupancy).
int fft_cols_lgC(int n) {
    const int nh = n / 2 + 1;
    const int min_lg = ctx().tune[CUBE2_TUNE_FFT_MIN_LGC];
    int lgC = std::max(lg_floor(std::max(1, kFFTElems / n)), min_lg);
    while (lgC > 0 && (1 << lgC) > 2 * nh) lgC--;  // no wider than the half spectrum needs
    while (lgC > 0 && (1 << lgC) * n > kColsMax) lgC--;  // the tile must fit the LDS buffer
    return lgC;
}

template <bool R7, int N = 0>
static void launch_cols(bool wide, int mode, dim3 grid, size_t lds, hipStream_t s, float* d,
                        const FFTPlan& P, size_t rs, int axis, int lgC, const float* green,
                        int o0, int ilo, int ihi) {
    if constexpr (N != 0) wide = false;  // never wide (fft_cols dispatches so)
    if (wide) {
        if (mode == 0) k_fft_cols<0, kColsMax, R7, N><<<grid, kFFTThreads, lds, s>>>(d, P, rs, axis, lgC, green, o0, ilo, ihi);
        else if (mode == 1) k_fft_cols<1, kColsMax, R7, N><<<grid, kFFTThreads, lds, s>>>(d, P, rs, axis, lgC, green, o0, ilo, ihi);
        else k_fft_cols<2, kColsMax, R7, N><<<grid, kFFTThreads, lds, s>>>(d, P, rs, axis, lgC, green, o0, ilo, ihi);
    } else {
        if (mode == 0) k_fft_cols<0, kFFTElems, R7, N><<<grid, kFFTThreads, lds, s>>>(d, P, rs, axis, lgC, green, o0, ilo, ihi);
        else if (mode == 1) k_fft_cols<1, kFFTElems, R7, N><<<grid, kFFTThreads, lds, s>>>(d, P, rs, axis, lgC, green, o0, ilo, ihi);
        else k_fft_cols<2, kFFTElems, R7, N><<<grid, kFFTThreads, lds, s>>>(d, P, rs, axis, lgC, green, o0, ilo, ihi);
    }
}

int fft_cols(float* d, const FFTPlan& P, int nb, size_t rs, int axis, int mode,
             const float* green, hipStream_t s, const FFTWin* win) {
    const int n = P.n, nh = n / 2 + 1;
    const int lgC = fft_cols_lgC(n);
    const int C = 1 << lgC;
    const int ntiles = (nh + C - 1) / C;
    // a window restricts the planes of an inverse y pass and the planes an
    // inverse z pass stores
    int o0 = 0, no = n, ilo = 0, ihi = n;
    if (win) {
        if (win->z0 < 0 || win->nz < 1 || win->z0 + win->nz > n) {
            set_error("FFT window outside the mesh");
            return CUBE2_ERR_ARG;
        }
        if (axis == 1 && mode == 1) {
            o0 = win->z0;
            no = win->nz;
        } else if (axis == 2 && mode == 2) {
            ilo = win->z0;
            ihi = win->z0 + win->nz;
        }
    }
    dim3 grid((unsigned)(no * ntiles), (unsigned)nb);
    const size_t lds = (size_t)(n + C * n) * sizeof(float2);
    const bool wide = C * n > kFFTElems;
    // the specialised kernels exist for the built-in tile width only
    const bool builtin = !wide && lgC == cols_lgc_builtin(n);
    with_size(builtin ? n : 0, P.has7, [&](auto c) {
        constexpr int N = decltype(c)::value;
        if (N == 0 && P.has7)
            launch_cols<true>(wide, mode, grid, lds, s, d, P, rs, axis, lgC, green, o0, ilo, ihi);
        else
            launch_cols<false, N>(wide, mode, grid, lds, s, d, P, rs, axis, lgC, green, o0, ilo,
                                  ihi);
    });
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

struct GreenTiled {
    const float* src;
    int n, lgC;
    float* dev;
};
static std::vector<GreenTiled> g_green_tiled;

// The tiled copy is rewritten from G on every call (a plan's preparation):
// a caller's table at a given device address can change (a freed table's
// memory reused for another), so the copy is never trusted across calls;
// only its allocation is kept per (address, n, tile width).
int green_tiled(const float* G, int n, const float** out, hipStream_t s) {
    const int lgC = fft_cols_lgC(n), C = 1 << lgC;
    const int ntiles = (n / 2 + 1 + C - 1) / C;
    const size_t tot = (size_t)n * ntiles * n * C;
    GreenTiled* hit = nullptr;
    for (auto& t : g_green_tiled)
        if (t.src == G && t.n == n && t.lgC == lgC) hit = &t;
    if (!hit) {
        GreenTiled t{G, n, lgC, nullptr};
        buffers_changed();
        C2_HIP(hipMalloc(&t.dev, tot * sizeof(float)));
        g_green_tiled.push_back(t);
        hit = &g_green_tiled.back();
    }
    k_green_tile<<<ceil_div(tot, 256), 256, 0, s>>>(G, hit->dev, n, lgC, ntiles);
    C2_HIP(hipGetLastError());
    C2_HIP(hipStreamSynchronize(s));
    *out = hit->dev;
    return CUBE2_OK;
}

void green_tiled_release() {
    buffers_changed();
    for (auto& t : g_green_tiled) (void)hipFree(t.dev);
    g_green_tiled.clear();
}

int fft3d_batch(float* d, int n, int nb, int dir, hipStream_t s) {
    const FFTPlan* P = nullptr;
    C2_TRY(fft_plan(n, &P));
    const size_t rs = (size_t)(n + 2) * n * n;
    if (dir < 0) {
        C2_TRY(fft_x_r2c(d, *P, nb, rs, nullptr, s));
        C2_TRY(fft_cols(d, *P, nb, rs, 1, 0, nullptr, s));
        C2_TRY(fft_cols(d, *P, nb, rs, 2, 0, nullptr, s));
    } else {
        C2_TRY(fft_cols(d, *P, nb, rs, 2, 1, nullptr, s));
        C2_TRY(fft_cols(d, *P, nb, rs, 1, 1, nullptr, s));
        C2_TRY(fft_x_c2r(d, *P, nb, rs, s));
    }
    return CUBE2_OK;
}

}  // namespace cube2
