// Block-level reductions feeding a single global word.  Many wavefronts
// updating one address serialise in L2, so each block combines its values
// first and a max is only attempted when it can raise the stored value.
#pragma once
#include "common.h"

namespace cube2 {

__device__ inline float wave_max_f(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ inline unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block max of non-negative floats of NV independent quantities into
// bits[0..NV); every thread of the block must call it.
template <int NV>
__device__ inline void block_max_to(unsigned* bits, const float (&v)[NV]) {
    __shared__ float part[NV][16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    float w[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) w[q] = wave_max_f(v[q]);
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NV; q++) part[q][wave] = w[q];
    __syncthreads();
    if (threadIdx.x < NV) {
        const int q = threadIdx.x;
        float m = 0.0f;
        for (int i = 0; i < nw; i++) m = fmaxf(m, part[q][i]);
        if (m > 0.0f) {
            const unsigned b = __float_as_uint(m);
            if (b > __hip_atomic_load(bits + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                atomicMax(bits + q, b);
        }
    }
    __syncthreads();
}

template <int NS>
__device__ inline void block_sum_to(unsigned long long* dst, const unsigned long long (&v)[NS]) {
    __shared__ unsigned long long part[NS][16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    unsigned long long w[NS];
#pragma unroll
    for (int q = 0; q < NS; q++) w[q] = wave_sum_u64(v[q]);
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NS; q++) part[q][wave] = w[q];
    __syncthreads();
    if (threadIdx.x < NS) {
        const int q = threadIdx.x;
        unsigned long long s = 0;
        for (int i = 0; i < nw; i++) s += part[q][i];
        if (s) atomicAdd(dst + q, s);
    }
    __syncthreads();
}

}  // namespace cube2
