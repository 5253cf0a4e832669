// The image layout on the device: the kick path's inputs rhoc, idx_b_r, xp,
// vp built from particle positions in HBM, as buffer_grid and spine_image
// produce them for one periodic image (buffer_grid.f90:13-155,
// variables.f90:146-216), halo copies included (the buffer_x copies of
// buffer_particle.f90:23-345 for a single image):
//   encode   global coarse cell and 16-bit in-cell position of every particle
//            (FP64, parameters.f90:87-88);
//   sort     stable radix sort (sort.hip) of (coarse cell, particle index): the order of
//            the particles of one cell is their input order;
//   starts   first sorted position of every coarse cell (binary search);
//   rhoc     particles per coarse cell of every tile's extended region (the
//            periodic neighbour's cell for halo cells);
//   slots    exclusive scan of rhoc: slot ranges per cell, idx_b_r = row ends;
//   fill     one wavefront per extended cell copies its particles' codes and
//            velocities into their slots.
#include <cstring>

#include "common.h"
#include "engine.h"
#include "profile.h"
#include "scan.h"
#include "sort.h"

namespace cube2 {

struct LayoutState {
    DevBuf gid, gid_sorted, idx, idx_sorted, code, cnt, slot, sort_tmp, scan_tmp, red;
};
static LayoutState g_lay;

// Global coarse cell (z-major) and int16 code of each position, in the same
// FP64 operations as the host layout builder: p = pos mod ng (sign of the
// divisor), pc = p / ratio_cs, c = clamp(floor(pc)), u = clamp(floor((pc - c)
// * 65536)) stored as the int16 with the same bits.
__global__ void k_layout_encode(const double* __restrict__ pos, int64_t n, int ng, int ratio_cs,
                                int nc, uint32_t* __restrict__ gid, uint32_t* __restrict__ idx,
                                int16_t* __restrict__ code) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int g[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double p = fmod(pos[3 * i + d], (double)ng);
        if (p != 0.0 && p < 0.0) p += (double)ng;
        const double pc = p / (double)ratio_cs;
        int c = (int)floor(pc);
        c = min(max(c, 0), nc - 1);
        const double frac = pc - (double)c;
        double u = floor(frac * 65536.0);
        u = fmin(fmax(u, 0.0), 65535.0);
        code[3 * i + d] = (int16_t)(uint16_t)u;
        g[d] = c;
    }
    gid[i] = (uint32_t)((g[2] * nc + g[1]) * nc + g[0]);
    idx[i] = (uint32_t)i;
}

// gstart[b] = first sorted position of a cell id >= b (b <= nbin): the
// cell starts straight from the sorted ids, by binary search.
__global__ void k_layout_starts(const uint32_t* __restrict__ gid_sorted, int64_t n, int64_t nbin,
                                int* __restrict__ gstart) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nbin) return;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)gid_sorted[mid] < b) lo = mid + 1;
        else hi = mid;
    }
    gstart[b] = (int)lo;
}

// Global coarse cell of extended cell (i, j, k) of tile t (periodic).
__device__ inline int ext_global(int64_t q, int nt, int ncb, int nte, int nnt, int nc) {
    const int i = (int)(q % nte), j = (int)((q / nte) % nte), k = (int)((q / ((int64_t)nte * nte)) % nte);
    const int t = (int)(q / ((int64_t)nte * nte * nte));
    const int tx = t % nnt, ty = (t / nnt) % nnt, tz = t / (nnt * nnt);
    const int gx = ((tx * nt + i - ncb) % nc + nc) % nc;
    const int gy = ((ty * nt + j - ncb) % nc + nc) % nc;
    const int gz = ((tz * nt + k - ncb) % nc + nc) % nc;
    return (gz * nc + gy) * nc + gx;
}

__global__ void k_layout_rhoc(const int* __restrict__ gstart, int64_t ncells, int nt, int ncb,
                              int nte, int nnt, int nc, int* __restrict__ rhoc,
                              int* __restrict__ slot) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ncells) return;
    const int g = ext_global(q, nt, ncb, nte, nnt, nc);
    const int c = gstart[g + 1] - gstart[g];
    rhoc[q] = c;
    slot[q] = c;  // scanned into slot starts
}

__global__ void k_layout_rows(const int* __restrict__ slot, int64_t nrows, int nte,
                              int64_t* __restrict__ idx_b_r) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrows) idx_b_r[r] = (int64_t)slot[(r + 1) * nte];  // exclusive start of the next row
}

// One wavefront per extended cell.
__global__ __launch_bounds__(256) void k_layout_fill(
    const int* __restrict__ gstart, const int* __restrict__ slot, const uint32_t* __restrict__ src,
    const int16_t* __restrict__ code, const float* __restrict__ vel, int64_t ncells, int nt,
    int ncb, int nte, int nnt, int nc, int16_t* __restrict__ xp, float* __restrict__ vp) {
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (q >= ncells) return;
    const int g = ext_global(q, nt, ncb, nte, nnt, nc);
    const int s0 = gstart[g], c = gstart[g + 1] - s0;
    const int64_t d0 = slot[q];
    for (int l = lane; l < c; l += 64) {
        const uint32_t p = src[s0 + l];
        const int64_t d = d0 + l;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            xp[3 * d + k] = code[3 * (int64_t)p + k];
            vp[3 * d + k] = vel ? vel[3 * (int64_t)p + k] : 0.0f;
        }
    }
}

int layout_build(const double* d_pos, const float* d_vel, int64_t n, int* d_rhoc,
                 int64_t* d_idx_b_r, int16_t* d_xp, float* d_vp, int64_t cap, int64_t* n_used) {
    ensure_params();
    Context& C = ctx();
    const Geom& g = C.g;
    if ((!d_pos && n > 0) || !d_rhoc || !d_idx_b_r || !n_used || n < 0 || n >= ((int64_t)1 << 31) ||
        (int64_t)g.nc * g.nc * g.nc >= ((int64_t)1 << 31)) {
        set_error("layout: null argument or size out of range");
        return CUBE2_ERR_ARG;
    }
    C2_TRY(ensure_init());
    hipStream_t s = C.stream;
    const int64_t nbin = (int64_t)g.nc * g.nc * g.nc;
    const int64_t ncells = (int64_t)g.ntile * g.nte * g.nte * g.nte;
    const int64_t nrows = (int64_t)g.ntile * g.nte * g.nte;
    const size_t nn = (size_t)std::max<int64_t>(n, 1);
    C2_TRY(g_lay.gid.reserve(nn * 4));
    C2_TRY(g_lay.gid_sorted.reserve(nn * 4));
    C2_TRY(g_lay.idx.reserve(nn * 4));
    C2_TRY(g_lay.idx_sorted.reserve(nn * 4));
    C2_TRY(g_lay.code.reserve(nn * 6));
    C2_TRY(g_lay.cnt.reserve((nbin + 1) * sizeof(int)));
    C2_TRY(g_lay.slot.reserve((ncells + 1) * sizeof(int)));
    C2_TRY(g_lay.scan_tmp.reserve(scan_tmp_ints(std::max(nbin, ncells)) * sizeof(int)));
    int end_bit = 1;
    while (((int64_t)1 << end_bit) < nbin) end_bit++;
    {
        Stage st("layout_encode", (double)n * (24 + 14));
        if (n) {
            k_layout_encode<<<ceil_div(n, 256), 256, 0, s>>>(d_pos, n, g.p.ng, g.p.ratio_cs, g.nc,
                                                            g_lay.gid.as<uint32_t>(),
                                                            g_lay.idx.as<uint32_t>(),
                                                            g_lay.code.as<int16_t>());
            C2_HIP(hipGetLastError());
        }
    }
    {
        Stage st("layout_sort", (double)n * 16);
        C2_TRY(radix_sort_pairs_u32(g_lay.gid.as<uint32_t>(), g_lay.gid_sorted.as<uint32_t>(),
                                    g_lay.idx.as<uint32_t>(), g_lay.idx_sorted.as<uint32_t>(), n,
                                    end_bit, g_lay.sort_tmp, g_lay.scan_tmp, s));
    }
    int* cnt = g_lay.cnt.as<int>();
    int* slot = g_lay.slot.as<int>();
    {
        Stage st("layout_cells", (double)n * 8 + (double)ncells * 16);
        k_layout_starts<<<ceil_div(nbin + 1, 256), 256, 0, s>>>(g_lay.gid_sorted.as<uint32_t>(), n,
                                                                nbin, cnt);  // total at [nbin]
        C2_HIP(hipGetLastError());
        k_layout_rhoc<<<ceil_div(ncells, 256), 256, 0, s>>>(cnt, ncells, g.nt, g.ncb, g.nte,
                                                            g.p.nnt, g.nc, d_rhoc, slot);
        C2_HIP(hipGetLastError());
        C2_TRY(exclusive_scan(slot, ncells, g_lay.scan_tmp, s));
        k_layout_rows<<<ceil_div(nrows, 256), 256, 0, s>>>(slot, nrows, g.nte, d_idx_b_r);
        C2_HIP(hipGetLastError());
    }
    int total = 0;
    C2_HIP(hipMemcpyAsync(&total, slot + ncells, sizeof(int), hipMemcpyDeviceToHost, s));
    C2_HIP(hipStreamSynchronize(s));
    *n_used = total;
    if (total < 0) {
        set_error("layout: more than 2^31 particle slots");
        return CUBE2_ERR_CAPACITY;
    }
    if (total > cap || !d_xp || !d_vp) {
        set_error("layout: %d slots needed, capacity %lld", total, (long long)cap);
        return CUBE2_ERR_CAPACITY;
    }
    {
        Stage st("layout_fill", (double)total * (6 + 12) * 2);
        k_layout_fill<<<ceil_div(ncells, 4), 256, 0, s>>>(
            cnt, slot, g_lay.idx_sorted.as<uint32_t>(), g_lay.code.as<int16_t>(), d_vel, ncells,
            g.nt, g.ncb, g.nte, g.p.nnt, g.nc, d_xp, d_vp);
        C2_HIP(hipGetLastError());
    }
    return sync_and_flush();
}

// buffer_v for one image (buffer_particle.f90:11-21 through buffer_xvp,
// :23-345): the velocities of every halo cell's slots from the periodic
// source cell's slots, which hold the same particles in the same order.  A
// wavefront per extended cell; physical cells return at once.
__global__ __launch_bounds__(256) void k_halo_vp(const int* __restrict__ rhoc,
                                                 const int64_t* __restrict__ idx_b_r,
                                                 float* __restrict__ vp, int64_t n_used, int nt,
                                                 int ncb, int nte, int nnt, int* __restrict__ err) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t q = (int64_t)blockIdx.x * 4 + wave;
    const int64_t te3 = (int64_t)nte * nte * nte;
    if (q >= te3 * nnt * nnt * nnt) return;
    const int t = (int)(q / te3);
    const int r = (int)(q - (int64_t)t * te3);
    const int c[3] = {r % nte, (r / nte) % nte, r / (nte * nte)};
    bool halo = false;
    for (int d = 0; d < 3; d++) halo |= c[d] < ncb || c[d] >= ncb + nt;
    if (!halo) return;
    const int np = rhoc[q];
    if (np <= 0) return;
    // source: the physical cell of the periodic image holding this cell
    const int nc = nt * nnt;
    const int tc[3] = {t % nnt, (t / nnt) % nnt, t / (nnt * nnt)};
    int st[3], sc[3];
    for (int d = 0; d < 3; d++) {
        const int gcell = ((tc[d] * nt + c[d] - ncb) % nc + nc) % nc;
        st[d] = gcell / nt;
        sc[d] = gcell % nt + ncb;
    }
    const int64_t ts = ((int64_t)st[2] * nnt + st[1]) * nnt + st[0];
    const int64_t qs = ts * te3 + ((int64_t)sc[2] * nte + sc[1]) * nte + sc[0];
    auto base = [&](int64_t cell, int64_t tile, int k, int j, int i) {
        int64_t suf = 0;
        for (int a = i; a < nte; a++) suf += rhoc[cell - i + a];
        return idx_b_r[tile * nte * nte + (int64_t)k * nte + j] - suf;
    };
    const int64_t b0 = base(q, t, c[2], c[1], c[0]);
    const int64_t b1 = base(qs, ts, sc[2], sc[1], sc[0]);
    if (rhoc[qs] != np || b0 < 0 || b1 < 0 || b0 + np > n_used || b1 + np > n_used) {
        if (lane == 0) atomicOr(err, 1);
        return;
    }
    for (int e = lane; e < 3 * np; e += 64) vp[3 * b0 + e] = vp[3 * b1 + e];
}

int halo_velocities(const int* d_rhoc, const int64_t* d_idx_b_r, float* d_vp, int64_t n_used) {
    ensure_params();
    Context& C = ctx();
    const Geom& g = C.g;
    if (!d_rhoc || !d_idx_b_r || n_used < 0 || (n_used > 0 && !d_vp)) {
        set_error("halo velocities: null argument");
        return CUBE2_ERR_ARG;
    }
    C2_TRY(ensure_init());
    hipStream_t s = C.stream;
    C2_TRY(g_lay.red.reserve(sizeof(int)));
    int* err = g_lay.red.as<int>();
    C2_TRY(fill_zero(err, sizeof(int), s));
    const int64_t ncells = (int64_t)g.ntile * g.nte * g.nte * g.nte;
    {
        Stage st("halo_vp", (double)n_used * 24);
        k_halo_vp<<<ceil_div(ncells, 4), 256, 0, s>>>(d_rhoc, d_idx_b_r, d_vp, n_used, g.nt, g.ncb,
                                                      g.nte, g.p.nnt, err);
        C2_HIP(hipGetLastError());
    }
    int h = 0;
    C2_HIP(hipMemcpyAsync(&h, err, sizeof(int), hipMemcpyDeviceToHost, s));
    C2_TRY(sync_and_flush());
    if (h) {
        set_error("halo velocities: a halo cell and its source cell differ (not one image's layout)");
        return CUBE2_ERR_ARG;
    }
    return CUBE2_OK;
}

void layout_release() {
    LayoutState& L = g_lay;
    for (DevBuf* b : {&L.gid, &L.gid_sorted, &L.idx, &L.idx_sorted, &L.code, &L.cnt, &L.slot,
                      &L.sort_tmp, &L.scan_tmp, &L.red})
        b->release();
}

}  // namespace cube2
