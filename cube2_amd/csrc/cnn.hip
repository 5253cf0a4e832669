// Kernels of the learned PP stand-in: a 3-D CNN of 3^3 valid convolutions
// in channels-last layout (activation [patch][z][y][x][channel], FP32),
// each layer an im2col (BF16) and a BF16 GEMM on the matrix cores with FP32
// accumulation; backward by the same GEMM on transposed operands.
//
// GEMM: C[m][n] (+)= sum_k A(m,k) B(k,n), A(m,k) = A[m*sam + k*sak],
// B(k,n) = B[k*sbk + n*sbn], C row-major with leading dimension ldc.  One
// wavefront owns a 32 x 32 tile of C as 2 x 2 v_mfma_f32_16x16x16_bf16
// tiles; operand lanes: A row m0 + (lane % 16), K values 4 (lane / 16) +
// 0..3; B column n0 + (lane % 16), the same K values; result lane holds rows
// 4 (lane / 16) + 0..3 of column lane % 16.  Out-of-range operands are
// zeros, so any M, N, K work.
#include "common.h"
#include "engine.h"

namespace cube2 {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (finite inputs).
__device__ inline uint16_t to_bf16(float f) {
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// Split K: block z of gridDim.z takes k in [z*kc, (z+1)*kc) and stores its
// partial product to C + z*zs (the partials are then summed in order).
__global__ __launch_bounds__(256) void k_gemm_bf16(const uint16_t* __restrict__ A, int64_t sam,
                                                   int64_t sak, const uint16_t* __restrict__ B,
                                                   int64_t sbk, int64_t sbn, float* __restrict__ C,
                                                   int64_t ldc, int M, int N, int K, int acc,
                                                   int kc, int64_t zs) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int kbeg = blockIdx.z * kc, kend = min(K, kbeg + kc);
    A += (int64_t)kbeg * sak;
    B += (int64_t)kbeg * sbk;
    K = kend - kbeg;
    C += blockIdx.z * zs;
    const int m0 = blockIdx.y * 64 + (wave >> 1) * 32, n0 = blockIdx.x * 64 + (wave & 1) * 32;
    const int r16 = lane & 15, kq = (lane >> 4) * 4;
    v4f d[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) d[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
    // A rows contiguous in k and 8-byte aligned: one 8-byte load per lane
    const bool avec = sak == 1 && (sam & 3) == 0 && ((uintptr_t)A & 7) == 0;
    for (int k0 = 0; k0 < K; k0 += 16) {
        v4s a[2], b[2];
        const bool kfull = k0 + 16 <= K;
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int m = m0 + i * 16 + r16, n = n0 + i * 16 + r16;
            if (avec && kfull) {
                a[i] = m < M ? *reinterpret_cast<const v4s*>(A + (int64_t)m * sam + k0 + kq)
                             : v4s{0, 0, 0, 0};
            } else {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int k = k0 + kq + q;
                    a[i][q] = (m < M && k < K) ? (short)A[(int64_t)m * sam + (int64_t)k * sak]
                                               : (short)0;
                }
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int k = k0 + kq + q;
                b[i][q] = (n < N && k < K) ? (short)B[(int64_t)k * sbk + (int64_t)n * sbn] : (short)0;
            }
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                d[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], d[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int m = m0 + i * 16 + kq + q, n = n0 + j * 16 + r16;
                if (m < M && n < N) {
                    float* c = C + (int64_t)m * ldc + n;
                    *c = acc ? *c + d[i][j][q] : d[i][j][q];
                }
            }
}

// C[m][n] (+)= sum over the S split-K partials P[s][m][n], in order.
__global__ void k_sum_partials(const float* __restrict__ P, float* __restrict__ C, int64_t ldc,
                               int M, int N, int S, int acc) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)M * N) return;
    const int m = (int)(e / N), n = (int)(e % N);
    float v = 0.f;
    for (int s = 0; s < S; s++) v += P[(int64_t)s * M * N + e];
    float* c = C + (int64_t)m * ldc + n;
    *c = acc ? *c + v : v;
}

// col[(p, z, y, x)][((dz*3 + dy)*3 + dx)*Cin + c] = X[p][z+dz][y+dy][x+dx][c]
// for the (D-2)^3 output positions of a D^3 input.
__global__ void k_im2col3(const float* __restrict__ X, uint16_t* __restrict__ col, int B, int D,
                          int Cin) {
    const int P = D - 2, Kd = 27 * Cin;
    const int64_t tot = (int64_t)B * P * P * P * Kd;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= tot) return;
    const int k = (int)(e % Kd);
    int64_t r = e / Kd;
    const int x = (int)(r % P), y = (int)((r / P) % P), z = (int)((r / ((int64_t)P * P)) % P);
    const int64_t p = r / ((int64_t)P * P * P);
    const int c = k % Cin, o = k / Cin, dx = o % 3, dy = (o / 3) % 3, dz = o / 9;
    col[e] = to_bf16(X[(((p * D + z + dz) * D + y + dy) * D + x + dx) * Cin + c]);
}

// dX[p][iz][iy][ix][c] = sum over the offsets (dz, dy, dx) whose output
// position (iz-dz, iy-dy, ix-dx) exists of dcol[that position][offset, c].
__global__ void k_col2im3(const float* __restrict__ dcol, float* __restrict__ dX, int B, int D,
                          int Cin) {
    const int P = D - 2, Kd = 27 * Cin;
    const int64_t tot = (int64_t)B * D * D * D * Cin;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= tot) return;
    const int c = (int)(e % Cin);
    int64_t r = e / Cin;
    const int ix = (int)(r % D), iy = (int)((r / D) % D), iz = (int)((r / ((int64_t)D * D)) % D);
    const int64_t p = r / ((int64_t)D * D * D);
    float s = 0.f;
    for (int dz = 0; dz < 3; dz++) {
        const int z = iz - dz;
        if (z < 0 || z >= P) continue;
        for (int dy = 0; dy < 3; dy++) {
            const int y = iy - dy;
            if (y < 0 || y >= P) continue;
            for (int dx = 0; dx < 3; dx++) {
                const int x = ix - dx;
                if (x < 0 || x >= P) continue;
                const int64_t row = ((p * P + z) * P + y) * P + x;
                s += dcol[row * Kd + ((dz * 3 + dy) * 3 + dx) * Cin + c];
            }
        }
    }
    dX[e] = s;
}

__global__ void k_bias_act(float* __restrict__ Y, const float* __restrict__ bias, int64_t n, int C,
                           int relu) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    float v = Y[e] + bias[e % C];
    Y[e] = relu ? fmaxf(v, 0.f) : v;
}

// dY *= (Y > 0) for a ReLU layer; dY as BF16 for the GEMMs.
__global__ void k_act_bwd(const float* __restrict__ Y, float* __restrict__ dY,
                          uint16_t* __restrict__ dYb, int64_t n, int relu) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    float g = dY[e];
    if (relu && !(Y[e] > 0.f)) g = 0.f;
    dY[e] = g;
    dYb[e] = to_bf16(g);
}

// dbias[c] = sum over rows of dY[row][c]: one block per channel, fixed order.
__global__ __launch_bounds__(256) void k_colsum(const float* __restrict__ dY, float* __restrict__ db,
                                                int64_t rows, int C) {
    __shared__ float red[256];
    const int c = blockIdx.x;
    float s = 0.f;
    for (int64_t r = threadIdx.x; r < rows; r += blockDim.x) s += dY[r * C + c];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = red[0];
}

__global__ void k_to_bf16(const float* __restrict__ x, uint16_t* __restrict__ y, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) y[e] = to_bf16(x[e]);
}

// Adam on FP32 master weights (bias-corrected step t >= 1); refreshes the
// BF16 copy the GEMMs read when wb is given.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint16_t* __restrict__ wb, int64_t n, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.f - b1) * gi, vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    const float wi = w[e] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[e] = wi;
    if (wb) wb[e] = to_bf16(wi);
}

// Mean squared error over n values and its gradient 2 (out - t) / n.
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ out,
                                             const float* __restrict__ t, float* __restrict__ d,
                                             float* __restrict__ loss, int64_t n) {
    __shared__ float red[256];
    float s = 0.f;
    for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
        const float r = out[e] - t[e];
        d[e] = 2.f * r / (float)n;
        s += r * r;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = red[0] / (float)n;
}

// ------------------------------------------------------ 3^3 convolution
// Y[b][z][y][x][o] = act(bias[o] + sum over taps t = (dz, dy, dx) and input
// channels c of X[b][z+dz][y+dy][x+dx][c] * W[t*CIN + c][o]) for the
// P^3 = (D-2)^3 outputs of each of B inputs of D^3 x CIN cells
// (channels-last), on v_mfma_f32_16x16x32_bf16 (BF16 operands, FP32
// accumulation) with no materialised im2col.
//
// A workgroup (4 wavefronts) is persistent over output blocks of
// 4 (x) x 4 (y) x 8 (z) positions.  For each block it stages the 6 x 6 x 10
// input cells as BF16 in LDS; wavefront w computes the z-planes 2w and 2w+1
// (two 16-row tiles, the 4 x 4 positions of a plane) against every output
// column.  The packed weights Wp[tap][out][in] (BF16) are loaded into LDS
// once per workgroup.  MFMA operands: lane l holds row (position) l % 16 of
// A and column (output channel) l % 16 of B with K indices 8 (l / 16) + 0..7;
// the result lane holds rows 4 (l / 16) + 0..3 of column l % 16.  K runs over
// (tap, channel) with the channel fastest: for CIN = 32 one K step is one
// tap; for CIN = 4 a step is 8 taps (taps 27..31 are zeros).
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef __bf16 v4bf __attribute__((ext_vector_type(4)));

constexpr int kConvBx = 4, kConvBy = 4, kConvBz = 8;
constexpr int kConvIx = kConvBx + 2, kConvIy = kConvBy + 2, kConvIz = kConvBz + 2;
constexpr int kConvCells = kConvIx * kConvIy * kConvIz;  // 360 staged input cells

template <int CIN>
constexpr int conv_taps_padded() { return CIN == 32 ? 27 : 32; }

template <int CIN, int COUTP>
constexpr size_t conv_lds_bytes() {
    return (size_t)conv_taps_padded<CIN>() * COUTP * CIN * 2 + (size_t)kConvCells * CIN * 2;
}

__device__ inline float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

template <class T>
__device__ inline float load_f32(const T* p);
template <>
__device__ inline float load_f32<float>(const float* p) { return *p; }
template <>
__device__ inline float load_f32<uint16_t>(const uint16_t* p) { return bf16_to_f32(*p); }

template <class T>
__device__ inline void store_act(T* p, float v);
template <>
__device__ inline void store_act<float>(float* p, float v) { *p = v; }
template <>
__device__ inline void store_act<uint16_t>(uint16_t* p, float v) { *p = to_bf16(v); }

template <int CIN, int COUTP, class TIn, class TOut>
__global__ __launch_bounds__(256) void k_conv3(const TIn* __restrict__ X,
                                               const uint16_t* __restrict__ Wp,
                                               const float* __restrict__ bias,
                                               TOut* __restrict__ Y, int B, int D, int cout,
                                               int relu) {
    extern __shared__ __align__(16) uint16_t conv_lds[];
    constexpr int TAPS = conv_taps_padded<CIN>();
    uint16_t* wl = conv_lds;                         // [TAPS][COUTP][CIN]
    uint16_t* xl = conv_lds + TAPS * COUTP * CIN;    // [cell][CIN]
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    {
        const uint4* src = reinterpret_cast<const uint4*>(Wp);
        uint4* dst = reinterpret_cast<uint4*>(wl);
        for (int i = tid; i < TAPS * COUTP * CIN / 8; i += 256) dst[i] = src[i];
    }
    const int P = D - 2;
    const int nbx = (P + kConvBx - 1) / kConvBx, nby = (P + kConvBy - 1) / kConvBy,
              nbz = (P + kConvBz - 1) / kConvBz;
    const int64_t per = (int64_t)nbx * nby * nbz, nblk = per * B;
    const int r16 = lane & 15, kq = lane >> 4;
    for (int64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
        const int b = (int)(blk / per);
        const int64_t rr = blk - (int64_t)b * per;
        const int bx = (int)(rr % nbx), by = (int)((rr / nbx) % nby), bz = (int)(rr / ((int64_t)nbx * nby));
        const int x0 = bx * kConvBx, y0 = by * kConvBy, z0 = bz * kConvBz;
        __syncthreads();  // the previous block's reads of xl are done (and wl is loaded)
        // stage the input cells as BF16 (zeros outside the grid), 4 channels per item
        for (int i = tid; i < kConvCells * (CIN / 4); i += 256) {
            const int cell = i / (CIN / 4), q = i % (CIN / 4);
            const int cx = cell % kConvIx, cy = (cell / kConvIx) % kConvIy, cz = cell / (kConvIx * kConvIy);
            const int gx = x0 + cx, gy = y0 + cy, gz = z0 + cz;
            uint16_t h[4] = {0, 0, 0, 0};
            if (gx < D && gy < D && gz < D) {
                const TIn* src = X + ((((int64_t)b * D + gz) * D + gy) * D + gx) * CIN + 4 * q;
#pragma unroll
                for (int j = 0; j < 4; j++) h[j] = to_bf16(load_f32<TIn>(src + j));
            }
            uint2 v;
            v.x = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
            v.y = (uint32_t)h[2] | ((uint32_t)h[3] << 16);
            *reinterpret_cast<uint2*>(xl + cell * CIN + 4 * q) = v;
        }
        __syncthreads();
        v4f acc[2][COUTP / 16];
#pragma unroll
        for (int rt = 0; rt < 2; rt++)
#pragma unroll
            for (int ct = 0; ct < COUTP / 16; ct++) acc[rt][ct] = v4f{0.f, 0.f, 0.f, 0.f};
        const int xr = r16 & 3, yr = r16 >> 2;
        constexpr int STEPS = TAPS * CIN / 32;
        for (int s = 0; s < STEPS; s++) {
            v8bf a[2], bm[COUTP / 16];
            if (CIN == 32) {
                const int t = s, dz = t / 9, dy = (t / 3) % 3, dx = t % 3;
#pragma unroll
                for (int rt = 0; rt < 2; rt++) {
                    const int zl = 2 * wave + rt;
                    const int cell = ((zl + dz) * kConvIy + yr + dy) * kConvIx + xr + dx;
                    a[rt] = *reinterpret_cast<const v8bf*>(xl + cell * CIN + 8 * kq);
                }
#pragma unroll
                for (int ct = 0; ct < COUTP / 16; ct++)
                    bm[ct] = *reinterpret_cast<const v8bf*>(wl + ((t * COUTP + ct * 16 + r16) * CIN + 8 * kq));
            } else {  // CIN == 4: taps 8s + 2kq and 8s + 2kq + 1, four channels each
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int t = 8 * s + 2 * kq + h;
                    const bool tv = t < 27;
                    const int tt = tv ? t : 0;
                    const int dz = tt / 9, dy = (tt / 3) % 3, dx = tt % 3;
#pragma unroll
                    for (int rt = 0; rt < 2; rt++) {
                        const int zl = 2 * wave + rt;
                        const int cell = ((zl + dz) * kConvIy + yr + dy) * kConvIx + xr + dx;
                        v4bf q = *reinterpret_cast<const v4bf*>(xl + cell * CIN);
                        if (!tv) q = v4bf{};
#pragma unroll
                        for (int j = 0; j < 4; j++) a[rt][4 * h + j] = q[j];
                    }
#pragma unroll
                    for (int ct = 0; ct < COUTP / 16; ct++) {
                        const v4bf q = *reinterpret_cast<const v4bf*>(wl + (t * COUTP + ct * 16 + r16) * CIN);
#pragma unroll
                        for (int j = 0; j < 4; j++) bm[ct][4 * h + j] = q[j];
                    }
                }
            }
#pragma unroll
            for (int rt = 0; rt < 2; rt++)
#pragma unroll
                for (int ct = 0; ct < COUTP / 16; ct++)
                    acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[rt], bm[ct], acc[rt][ct], 0, 0, 0);
        }
        // epilogue: bias, activation, store the valid positions and channels
#pragma unroll
        for (int rt = 0; rt < 2; rt++) {
            const int gz = z0 + 2 * wave + rt;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int r = 4 * kq + i;
                const int gx = x0 + (r & 3), gy = y0 + (r >> 2);
                if (gx >= P || gy >= P || gz >= P) continue;
                TOut* out = Y + ((((int64_t)b * P + gz) * P + gy) * P + gx) * cout;
#pragma unroll
                for (int ct = 0; ct < COUTP / 16; ct++) {
                    const int o = ct * 16 + r16;
                    if (o >= cout) continue;
                    float v = acc[rt][ct][i] + bias[o];
                    if (relu) v = fmaxf(v, 0.f);
                    store_act<TOut>(out + o, v);
                }
            }
        }
    }
}

// Wp[t][o][c] = BF16(W[t*cin + c][o]) for t < 27, o < cout, zero elsewhere
// ([taps_padded][coutp][cin]).
__global__ void k_conv_pack(const float* __restrict__ W, uint16_t* __restrict__ Wp, int cin,
                            int cout, int taps, int coutp) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= taps * coutp * cin) return;
    const int c = e % cin, o = (e / cin) % coutp, t = e / (cin * coutp);
    Wp[e] = (t < 27 && o < cout) ? to_bf16(W[(t * cin + c) * cout + o]) : (uint16_t)0;
}

// Cloud-in-cell combination of the 2^3 output cells of each sample with its
// fractions f (pred = sum over corners of prod_d (f_d or 1 - f_d) out), the
// mean squared error against target over n*3 values and its gradient with
// respect to the 2^3 outputs.  One block.
__global__ __launch_bounds__(256) void k_cic_mse(const float* __restrict__ out8,
                                                 const float* __restrict__ frac,
                                                 const float* __restrict__ target, int n,
                                                 float* __restrict__ pred,
                                                 float* __restrict__ dout8,
                                                 float* __restrict__ loss) {
    __shared__ float red[256];
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const float f0 = frac[3 * i], f1 = frac[3 * i + 1], f2 = frac[3 * i + 2];
        float w[8];
#pragma unroll
        for (int c = 0; c < 8; c++)
            w[c] = ((c & 1) ? f0 : 1.f - f0) * ((c & 2) ? f1 : 1.f - f1) * ((c & 4) ? f2 : 1.f - f2);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            float p = 0.f;
#pragma unroll
            for (int c = 0; c < 8; c++) p += w[c] * out8[(8 * i + c) * 3 + d];
            if (pred) pred[3 * i + d] = p;
            const float r = p - target[3 * i + d];
            s += r * r;
            const float g = 2.f * r / (3.f * (float)n);
#pragma unroll
            for (int c = 0; c < 8; c++) dout8[(8 * i + c) * 3 + d] = w[c] * g;
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = red[0] / (3.f * (float)n);
}

static int cnn_check(bool ok, const char* what) {
    if (ok) return CUBE2_OK;
    set_error("cnn: bad arguments to %s", what);
    return CUBE2_ERR_ARG;
}

static unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

// Split-K schedule of a product: S partial products over K ranges of kc
// (a multiple of 16); S = 1 for no split.  Few output tiles and a long K
// (the weight gradients, K = rows) are split over blocks, and the partials
// summed in a fixed order.  S * M * N <= 2^21 for every split shape (tiles
// >= M N / 4096 and S <= 512 / tiles).
static void gemm_split(int M, int N, int K, int* S_out, int* kc_out) {
    const int64_t tiles = (int64_t)((N + 63) / 64) * ((M + 63) / 64);
    int S = 1;
    if (tiles < 256 && K >= 4096) S = (int)std::min<int64_t>(std::min<int64_t>(512 / tiles, K / 2048), 64);
    int kc = K;
    if (S > 1) {
        kc = (int)(((K + S - 1) / S + 15) / 16 * 16);
        S = (K + kc - 1) / kc;
    }
    *S_out = S;
    *kc_out = kc;
}

// The product on stream s, split-K partials (S > 1) in part[S * M * N].
static int gemm_launch(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk,
                       int64_t sbn, float* C, int64_t ldc, int M, int N, int K, int acc, int S,
                       int kc, float* part, hipStream_t s) {
    dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64));
    if (S <= 1) {
        k_gemm_bf16<<<grid, 256, 0, s>>>(A, sam, sak, B, sbk, sbn, C, ldc, M, N, K, acc, K, 0);
        C2_HIP(hipGetLastError());
        return CUBE2_OK;
    }
    grid.z = (unsigned)S;
    k_gemm_bf16<<<grid, 256, 0, s>>>(A, sam, sak, B, sbk, sbn, part, N, M, N, K, 0, kc,
                                     (int64_t)M * N);
    C2_HIP(hipGetLastError());
    k_sum_partials<<<blocks_for((int64_t)M * N), 256, 0, s>>>(part, C, ldc, M, N, S, acc);
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

static DevBuf g_partials;  // split-K partials of cnn_gemm (released by the finalizer)

int cnn_gemm(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk,
             int64_t sbn, float* C, int64_t ldc, int M, int N, int K, int acc) {
    C2_TRY(cnn_check(A && B && C && M >= 0 && N >= 0 && K >= 0 && ldc >= N, "gemm"));
    C2_TRY(ensure_init());
    if (M == 0 || N == 0) return CUBE2_OK;
    int S, kc;
    gemm_split(M, N, K, &S, &kc);
    if (S > 1) C2_TRY(g_partials.reserve((size_t)S * M * N * sizeof(float)));
    return gemm_launch(A, sam, sak, B, sbk, sbn, C, ldc, M, N, K, acc, S, kc, g_partials.as<float>(),
                       ctx().stream);
}

// ------------------------------------------------ allocation-free product
// cube2_gemm_bf16_noalloc_dev keeps its split-K partials in this file-scope
// device array (4 MiB; every GPU the library runs on holds its own copy
// until the process ends).  Host code never passes the array's name to a
// kernel: in host code it is a stand-in, not a device address.  The device
// address comes from hipGetSymbolAddress, once per device with that device
// current, and is kept per device.  Calls that overlap on one GPU share the
// array and race; the partials need no clearing (every call writes all
// S * M * N of them before they are summed).
constexpr size_t kScratchFloats = (size_t)1 << 20;
static __device__ __align__(256) float scratch[kScratchFloats];
constexpr int kMaxDevices = 64;
static float* g_scratch_ptr[kMaxDevices] = {};

static int scratch_address(float** out) {
    int dev = 0;
    C2_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device %d outside the scratch table", dev);
        return CUBE2_ERR_STATE;
    }
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        C2_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[dev];
    return CUBE2_OK;
}

int cnn_gemm_noalloc(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk,
                     int64_t sbn, float* C, int64_t ldc, int M, int N, int K, int acc) {
    C2_TRY(cnn_check(A && B && C && M >= 0 && N >= 0 && K >= 0 && ldc >= N, "gemm_noalloc"));
    C2_TRY(ensure_init());
    float* part = nullptr;
    C2_TRY(scratch_address(&part));
    if (M == 0 || N == 0) return CUBE2_OK;
    int S, kc;
    gemm_split(M, N, K, &S, &kc);
    if (S > 1 && (size_t)S * M * N > kScratchFloats) {
        set_error("gemm_noalloc: M=%d N=%d K=%d needs %zu bytes of split-K scratch; the static "
                  "scratch holds %zu", M, N, K, (size_t)S * M * N * sizeof(float),
                  kScratchFloats * sizeof(float));
        return CUBE2_ERR_CAPACITY;
    }
    return gemm_launch(A, sam, sak, B, sbk, sbn, C, ldc, M, N, K, acc, S, kc, part, ctx().stream);
}

void cnn_release() { g_partials.release(); }

// ---------------------------------------------------------- convolution
int cnn_conv_pack(const float* W, uint16_t* Wp, int cin, int cout, hipStream_t s) {
    C2_TRY(cnn_check(W && Wp && (cin == 4 || cin == 32) && (cout == 32 || (cout == 3 && cin == 32)),
                     "conv_pack"));
    const int taps = cin == 32 ? 27 : 32, coutp = cout == 3 ? 16 : 32;
    const int n = taps * coutp * cin;
    k_conv_pack<<<blocks_for(n), 256, 0, s>>>(W, Wp, cin, cout, taps, coutp);
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

size_t cnn_conv_packed_count(int cin, int cout) {
    return (size_t)(cin == 32 ? 27 : 32) * (cout == 3 ? 16 : 32) * cin;
}

template <int CIN, int COUTP, class TIn, class TOut>
static int conv_launch(const void* X, const uint16_t* Wp, const float* bias, void* Y, int B, int D,
                       int cout, int relu, hipStream_t s) {
    auto kern = k_conv3<CIN, COUTP, TIn, TOut>;
    const size_t lds = conv_lds_bytes<CIN, COUTP>();
    static int per_cu = 0;  // resident workgroups per CU (one value per instantiation)
    if (!per_cu) {
        C2_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
        int per = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, 256, lds) != hipSuccess || per < 1)
            per = 1;
        per_cu = per;
    }
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        ncu = 256;
    const int P = D - 2;
    const int64_t nblk = (int64_t)B * ((P + kConvBx - 1) / kConvBx) * ((P + kConvBy - 1) / kConvBy) *
                         ((P + kConvBz - 1) / kConvBz);
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(nblk, (int64_t)ncu * per_cu));
    kern<<<grid, 256, lds, s>>>(static_cast<const TIn*>(X), Wp, bias, static_cast<TOut*>(Y), B, D,
                                cout, relu);
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

// x_bf16 / y_bf16: activations stored as BF16 (else FP32).
int cnn_conv3(const void* X, int x_bf16, const uint16_t* Wp, const float* bias, void* Y, int y_bf16,
              int B, int D, int cin, int cout, int relu, hipStream_t s) {
    C2_TRY(cnn_check(X && Wp && bias && Y && B >= 0 && D >= 3, "conv3"));
    if (B == 0) return CUBE2_OK;
    if (cin == 4 && cout == 32 && !x_bf16)
        return y_bf16 ? conv_launch<4, 32, float, uint16_t>(X, Wp, bias, Y, B, D, cout, relu, s)
                      : conv_launch<4, 32, float, float>(X, Wp, bias, Y, B, D, cout, relu, s);
    if (cin == 32 && cout == 32)
        return x_bf16 ? (y_bf16 ? conv_launch<32, 32, uint16_t, uint16_t>(X, Wp, bias, Y, B, D, cout, relu, s)
                                : conv_launch<32, 32, uint16_t, float>(X, Wp, bias, Y, B, D, cout, relu, s))
                      : (y_bf16 ? conv_launch<32, 32, float, uint16_t>(X, Wp, bias, Y, B, D, cout, relu, s)
                                : conv_launch<32, 32, float, float>(X, Wp, bias, Y, B, D, cout, relu, s));
    if (cin == 32 && cout == 3 && !y_bf16)
        return x_bf16 ? conv_launch<32, 16, uint16_t, float>(X, Wp, bias, Y, B, D, cout, relu, s)
                      : conv_launch<32, 16, float, float>(X, Wp, bias, Y, B, D, cout, relu, s);
    set_error("conv3: no kernel for %d -> %d channels (x %s, y %s)", cin, cout,
              x_bf16 ? "bf16" : "f32", y_bf16 ? "bf16" : "f32");
    return CUBE2_ERR_ARG;
}

int cnn_im2col(const float* X, uint16_t* col, int B, int D, int Cin) {
    C2_TRY(cnn_check(X && col && B >= 0 && D >= 3 && Cin >= 1, "im2col"));
    C2_TRY(ensure_init());
    const int64_t P = D - 2, tot = (int64_t)B * P * P * P * 27 * Cin;
    if (tot) k_im2col3<<<blocks_for(tot), 256, 0, ctx().stream>>>(X, col, B, D, Cin);
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

int cnn_col2im(const float* dcol, float* dX, int B, int D, int Cin) {
    C2_TRY(cnn_check(dcol && dX && B >= 0 && D >= 3 && Cin >= 1, "col2im"));
    C2_TRY(ensure_init());
    const int64_t tot = (int64_t)B * D * D * D * Cin;
    if (tot) k_col2im3<<<blocks_for(tot), 256, 0, ctx().stream>>>(dcol, dX, B, D, Cin);
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

int cnn_bias_act(float* Y, const float* bias, int64_t rows, int C, int relu) {
    C2_TRY(cnn_check(Y && bias && rows >= 0 && C >= 1, "bias_act"));
    C2_TRY(ensure_init());
    const int64_t n = rows * C;
    if (n) k_bias_act<<<blocks_for(n), 256, 0, ctx().stream>>>(Y, bias, n, C, relu);
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

int cnn_act_bwd(const float* Y, float* dY, uint16_t* dYb, float* dbias, int64_t rows, int C,
                int relu) {
    C2_TRY(cnn_check(Y && dY && dYb && dbias && rows >= 0 && C >= 1, "act_bwd"));
    C2_TRY(ensure_init());
    const int64_t n = rows * C;
    hipStream_t s = ctx().stream;
    if (n) k_act_bwd<<<blocks_for(n), 256, 0, s>>>(Y, dY, dYb, n, relu);
    k_colsum<<<C, 256, 0, s>>>(dY, dbias, rows, C);
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

int cnn_to_bf16(const float* x, uint16_t* y, int64_t n) {
    C2_TRY(cnn_check(x && y && n >= 0, "to_bf16"));
    C2_TRY(ensure_init());
    if (n) k_to_bf16<<<blocks_for(n), 256, 0, ctx().stream>>>(x, y, n);
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

int cnn_adam(float* w, const float* g, float* m, float* v, uint16_t* wb, int64_t n, float lr,
             float b1, float b2, float eps, int t) {
    C2_TRY(cnn_check(w && g && m && v && n >= 0 && t >= 1, "adam"));
    C2_TRY(ensure_init());
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    if (n) k_adam<<<blocks_for(n), 256, 0, ctx().stream>>>(w, g, m, v, wb, n, lr, b1, b2, eps, c1, c2);
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

int cnn_cic_mse(const float* out8, const float* frac, const float* target, int n, float* pred,
                float* dout8, float* loss) {
    C2_TRY(cnn_check(out8 && frac && target && dout8 && loss && n >= 1, "cic_mse"));
    C2_TRY(ensure_init());
    k_cic_mse<<<1, 256, 0, ctx().stream>>>(out8, frac, target, n, pred, dout8, loss);
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

int cnn_mse(const float* out, const float* target, float* dout, float* loss, int64_t n) {
    C2_TRY(cnn_check(out && target && dout && loss && n >= 1, "mse"));
    C2_TRY(ensure_init());
    k_mse<<<1, 256, 0, ctx().stream>>>(out, target, dout, loss, n);
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

}  // namespace cube2
