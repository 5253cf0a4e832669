// Batched PM3 FFTs, hand-written for gfx950.
//
// One pass per axis; every pass stages a block of pencils in LDS and runs a
// mixed-radix (4, 2, 3, 5) self-sorting Stockham transform there, so HBM sees
// one read and one write of the mesh per pass.
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

#include "fft.h"
#include "profile.h"

namespace cube2 {

constexpr int kFFTThreads = 256;
constexpr int kFFTElems = 4096;  // complex elements per LDS buffer

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

// One Stockham stage over L lines: element e of line l at l*ls + e*es.
template <int R, int S>
__device__ inline void stage(const float2* __restrict__ a, float2* __restrict__ b,
                             const float2* __restrict__ tw, int n, int p, int L, int ls, int es,
                             bool line_fast) {
    const int m = n / R, ts = n / (p * R);
    const int tasks = L * m;
    for (int task = threadIdx.x; task < tasks; task += blockDim.x) {
        int l, j;
        if (line_fast) {
            l = task % L;
            j = task / L;
        } else {
            l = task / m;
            j = task - l * m;
        }
        const int k = j % p;
        float2 v[R];
#pragma unroll
        for (int q = 0; q < R; q++) v[q] = a[l * ls + (j + q * m) * es];
        if (k) {
#pragma unroll
            for (int q = 1; q < R; q++) {
                float2 w = tw[q * k * ts];
                if (S > 0) w.y = -w.y;
                v[q] = cmul(v[q], w);
            }
        }
        DFT<R, S>::run(v);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int q = 0; q < R; q++) b[l * ls + (base + q * p) * es] = v[q];
    }
}

// Transforms L lines held in LDS buffer a (b is the ping-pong partner);
// returns the buffer holding the result.  Ends with a barrier.
template <int S>
__device__ float2* fft_lds(float2* a, float2* b, const float2* tw, const FFTPlan& P, int L, int ls,
                           int es, bool line_fast) {
    int p = 1;
    for (int s = 0; s < P.nst; s++) {
        const int r = P.rad[s];
        if (r == 4) stage<4, S>(a, b, tw, P.n, p, L, ls, es, line_fast);
        else if (r == 2) stage<2, S>(a, b, tw, P.n, p, L, ls, es, line_fast);
        else if (r == 3) stage<3, S>(a, b, tw, P.n, p, L, ls, es, line_fast);
        else stage<5, S>(a, b, tw, P.n, p, L, ls, es, line_fast);
        __syncthreads();
        float2* t = a;
        a = b;
        b = t;
        p *= r;
    }
    return a;
}

__device__ inline void load_twiddles(float2* tw, const FFTPlan& P) {
    for (int t = threadIdx.x; t < P.n; t += blockDim.x) tw[t] = P.tw[t];
}

__global__ __launch_bounds__(kFFTThreads) void k_fft_x_r2c(float* __restrict__ d, FFTPlan P,
                                                           int nb, size_t rs, int PL, float scale) {
    extern __shared__ float2 sm[];
    const int n = P.n, nh = n / 2 + 1, half = n / 2;
    float2* tw = sm;
    float2* A = sm + n;
    float2* B = A + (size_t)PL * n;
    load_twiddles(tw, P);
    const int64_t npairs = (int64_t)nb * n * half;
    const int64_t g0 = (int64_t)blockIdx.x * PL;
    const int L = (int)min((int64_t)PL, npairs - g0);
    for (int e = threadIdx.x; e < L * n; e += blockDim.x) {
        const int l = e / n, t = e - l * n;
        const int64_t g = g0 + l;
        const int64_t job = g / ((int64_t)n * half), rem = g - job * n * half;
        const int z = (int)(rem / half), yp = (int)(rem - (int64_t)z * half);
        const float* ra = d + job * rs + ((size_t)z * n + 2 * yp) * (n + 2);
        A[l * n + t] = make_float2(ra[t] * scale, ra[t + n + 2] * scale);
    }
    __syncthreads();
    const float2* R = fft_lds<-1>(A, B, tw, P, L, n, 1, false);
    for (int e = threadIdx.x; e < L * nh; e += blockDim.x) {
        const int l = e / nh, k = e - l * nh;
        const int64_t g = g0 + l;
        const int64_t job = g / ((int64_t)n * half), rem = g - job * n * half;
        const int z = (int)(rem / half), yp = (int)(rem - (int64_t)z * half);
        float2* ra = reinterpret_cast<float2*>(d + job * rs + ((size_t)z * n + 2 * yp) * (n + 2));
        float2* rb = ra + nh;
        const float2 zk = R[l * n + k], zn = R[l * n + (k == 0 ? 0 : n - k)];
        ra[k] = make_float2(0.5f * (zk.x + zn.x), 0.5f * (zk.y - zn.y));
        rb[k] = make_float2(0.5f * (zk.y + zn.y), -0.5f * (zk.x - zn.x));
    }
}

__global__ __launch_bounds__(kFFTThreads) void k_fft_x_c2r(float* __restrict__ d, FFTPlan P,
                                                           int nb, size_t rs, int PL) {
    extern __shared__ float2 sm[];
    const int n = P.n, nh = n / 2 + 1, half = n / 2;
    float2* tw = sm;
    float2* A = sm + n;
    float2* B = A + (size_t)PL * n;
    load_twiddles(tw, P);
    const int64_t npairs = (int64_t)nb * n * half;
    const int64_t g0 = (int64_t)blockIdx.x * PL;
    const int L = (int)min((int64_t)PL, npairs - g0);
    for (int e = threadIdx.x; e < L * n; e += blockDim.x) {
        const int l = e / n, k = e - l * n;
        const int64_t g = g0 + l;
        const int64_t job = g / ((int64_t)n * half), rem = g - job * n * half;
        const int z = (int)(rem / half), yp = (int)(rem - (int64_t)z * half);
        const float2* ra = reinterpret_cast<const float2*>(d + job * rs + ((size_t)z * n + 2 * yp) * (n + 2));
        const float2* rb = ra + nh;
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
        A[l * n + k] = make_float2(a.x - b.y, a.y + b.x);
    }
    __syncthreads();
    const float2* R = fft_lds<1>(A, B, tw, P, L, n, 1, false);
    for (int e = threadIdx.x; e < L * n; e += blockDim.x) {
        const int l = e / n, t = e - l * n;
        const int64_t g = g0 + l;
        const int64_t job = g / ((int64_t)n * half), rem = g - job * n * half;
        const int z = (int)(rem / half), yp = (int)(rem - (int64_t)z * half);
        float* ra = d + job * rs + ((size_t)z * n + 2 * yp) * (n + 2);
        const float2 v = R[l * n + t];
        ra[t] = v.x;
        ra[t + n + 2] = v.y;
    }
}

template <int MODE>
__global__ __launch_bounds__(kFFTThreads) void k_fft_cols(float* __restrict__ dreal, FFTPlan P,
                                                          int nb, size_t rs, int axis, int C,
                                                          const float* __restrict__ green) {
    extern __shared__ float2 sm[];
    const int n = P.n, nh = n / 2 + 1;
    const int ntiles = (nh + C - 1) / C;
    float2* tw = sm;
    float2* A = sm + n;
    float2* B = A + (size_t)C * n;
    load_twiddles(tw, P);
    const int64_t b = blockIdx.x;
    const int64_t job = b / ((int64_t)n * ntiles);
    const int rem = (int)(b - job * n * ntiles);
    const int o = rem / ntiles, tile = rem - o * ntiles;
    const int kx0 = tile * C, Cw = min(C, nh - kx0);
    float2* d = reinterpret_cast<float2*>(dreal + job * rs);
    const size_t rowstride = axis == 1 ? (size_t)nh : (size_t)n * nh;
    const size_t obase = (axis == 1 ? (size_t)o * n * nh : (size_t)o * nh) + kx0;
    for (int e = threadIdx.x; e < Cw * n; e += blockDim.x) {
        const int i = e / Cw, c = e - i * Cw;
        A[e] = d[obase + i * rowstride + c];
    }
    __syncthreads();
    float2* R;
    if (MODE == 1) {
        R = fft_lds<1>(A, B, tw, P, Cw, 1, Cw, true);
    } else {
        R = fft_lds<-1>(A, B, tw, P, Cw, 1, Cw, true);
        if (MODE == 2) {
            for (int e = threadIdx.x; e < Cw * n; e += blockDim.x) {
                const int i = e / Cw, c = e - i * Cw;
                const float gk = green[obase + i * rowstride + c];
                R[e] = make_float2(R[e].x * gk, R[e].y * gk);
            }
            __syncthreads();
            R = fft_lds<1>(R, R == A ? B : A, tw, P, Cw, 1, Cw, true);
        }
    }
    for (int e = threadIdx.x; e < Cw * n; e += blockDim.x) {
        const int i = e / Cw, c = e - i * Cw;
        d[obase + i * rowstride + c] = R[e];
    }
}

// ------------------------------------------------------------- host side

static std::map<int, FFTPlan> g_plans;
static bool g_attr_set = false;

static int set_lds_attributes() {
    if (g_attr_set) return CUBE2_OK;
    const int bytes = (2 * kFFTElems + 1024) * (int)sizeof(float2);
    C2_HIP(hipFuncSetAttribute((const void*)k_fft_x_r2c, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    C2_HIP(hipFuncSetAttribute((const void*)k_fft_x_c2r, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    C2_HIP(hipFuncSetAttribute((const void*)k_fft_cols<0>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    C2_HIP(hipFuncSetAttribute((const void*)k_fft_cols<1>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    C2_HIP(hipFuncSetAttribute((const void*)k_fft_cols<2>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
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
    for (int r : {4, 2, 3, 5})
        while (m % r == 0) {
            if (P.nst == 16) break;
            P.rad[P.nst++] = r;
            m /= r;
        }
    if (m != 1) {
        set_error("FFT size %d has a prime factor other than 2, 3, 5", n);
        return CUBE2_ERR_CAPACITY;
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
    for (auto& kv : g_plans) (void)hipFree(kv.second.tw);
    g_plans.clear();
}

int fft_x_r2c(float* d, const FFTPlan& P, int nb, size_t rs, float scale, hipStream_t s) {
    const int n = P.n;
    const int PL = std::max(1, kFFTElems / n);
    const int64_t npairs = (int64_t)nb * n * (n / 2);
    const size_t lds = (size_t)(n + 2 * PL * n) * sizeof(float2);
    k_fft_x_r2c<<<(unsigned)((npairs + PL - 1) / PL), kFFTThreads, lds, s>>>(d, P, nb, rs, PL, scale);
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

int fft_x_c2r(float* d, const FFTPlan& P, int nb, size_t rs, hipStream_t s) {
    const int n = P.n;
    const int PL = std::max(1, kFFTElems / n);
    const int64_t npairs = (int64_t)nb * n * (n / 2);
    const size_t lds = (size_t)(n + 2 * PL * n) * sizeof(float2);
    k_fft_x_c2r<<<(unsigned)((npairs + PL - 1) / PL), kFFTThreads, lds, s>>>(d, P, nb, rs, PL);
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

int fft_cols(float* d, const FFTPlan& P, int nb, size_t rs, int axis, int mode,
             const float* green, hipStream_t s) {
    const int n = P.n, nh = n / 2 + 1;
    const int C = std::max(1, std::min(nh, kFFTElems / n));
    const int ntiles = (nh + C - 1) / C;
    const unsigned blocks = (unsigned)((int64_t)nb * n * ntiles);
    const size_t lds = (size_t)(n + 2 * C * n) * sizeof(float2);
    if (mode == 0) k_fft_cols<0><<<blocks, kFFTThreads, lds, s>>>(d, P, nb, rs, axis, C, green);
    else if (mode == 1) k_fft_cols<1><<<blocks, kFFTThreads, lds, s>>>(d, P, nb, rs, axis, C, green);
    else k_fft_cols<2><<<blocks, kFFTThreads, lds, s>>>(d, P, nb, rs, axis, C, green);
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

int fft3d_batch(float* d, int n, int nb, int dir, hipStream_t s) {
    const FFTPlan* P = nullptr;
    C2_TRY(fft_plan(n, &P));
    const size_t rs = (size_t)(n + 2) * n * n;
    if (dir < 0) {
        C2_TRY(fft_x_r2c(d, *P, nb, rs, 1.0f, s));
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
