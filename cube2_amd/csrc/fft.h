// Batched 3-D real FFTs of PM3 fine meshes, in place in the FFTW/cuFFT
// padded layout: nb cubes of n^3 reals, x rows of n+2 floats, which after
// the forward transform hold (n/2+1) complex values [kz][ky][kx].
#pragma once
#include "common.h"

namespace cube2 {

struct FFTPlan {
    int n = 0;
    int nst = 0;
    int rad[16] = {0};
    int pst[16] = {0};      // product of the radices before stage s
    float invp[16] = {0};   // 1/pst[s], for division by a runtime constant
    bool has7 = false;      // a radix-7 stage: the kernels with that butterfly
    float2* tw = nullptr;  // exp(-2 pi i t / n), t < n (device)
};

// Plan for size n (radices 4, 2, 3, 5, 7; cached; freed by fft_release()).
int fft_plan(int n, const FFTPlan** out);
void fft_release();

// R2C along x for every row pair, scaling the real input by *scale (device
// pointer; nullptr: no scaling).
int fft_x_r2c(float* d, const FFTPlan& P, int nb, size_t rs, const float* scale, hipStream_t s);
// C2R along x for every row pair.
// The part of the real result a caller reads: planes z0 .. z0+nz-1 and, in
// them, rows 2*yp0 .. 2*(yp0+nyp)-1.  The inverse passes skip the rest
// (those points are left undefined).
struct FFTWin {
    int z0, nz, yp0, nyp;
};

int fft_x_c2r(float* d, const FFTPlan& P, int nb, size_t rs, hipStream_t s,
              const FFTWin* w = nullptr);
// Complex transforms along y (axis 1) or z (axis 2) of the half spectrum.
//   mode 0: forward, 1: inverse, 2: forward, multiply by the Green's table,
//   inverse; green is the tile-ordered copy from green_tiled().
int fft_cols(float* d, const FFTPlan& P, int nb, size_t rs, int axis, int mode,
             const float* green, hipStream_t s, const FFTWin* w = nullptr);

// Tile-ordered device copy of a Green's table G[kz][ky][kx] / n^3 for the z pass,
// made once per (G, n) and cached (G's contents must not change).
int green_tiled(const float* G, int n, const float** out, hipStream_t s);
void green_tiled_release();

// Full batched transforms (dir -1: R2C, +1: C2R), unnormalised.
int fft3d_batch(float* d, int n, int nb, int dir, hipStream_t s);

}  // namespace cube2
