// Drift of one image on the device (drift.f90:1-158, spine_image of
// variables.f90:146-216): from the image with halo copies to the new
// physical-only image.
//   keys     a wavefront per extended cell of every tile moves each slot by
//            dt_mid*vp (the reference's FP32/FP64 mix): target cell
//            ceiling(xq + deltax), new int16 code xp + nint(dt_mid*vp /
//            (x_resolution*ratio_cs)) (wrapping).  A slot whose target is a
//            physical cell of the same tile gets the key (tile, cell, layer
//            of its source z), the others a key past every valid one;
//   sort     stable radix sort (sort.hip) of (key, slot): within a target cell the
//            particles come by source layer, then in slot order (source z,
//            y, x, slot), the arrival order of the layered loop
//            drift.f90:36-62, 90-122 (one source z per layer can reach a
//            cell when |dz| <= (nlayer-2)/2, which the kernel checks);
//   counts   run boundaries of the sorted cells give the new rhoc (halo
//            cells 0), a row scan gives idx_b_r;
//   gather   xp, vp of the kept slots in sorted order.
#include <cstring>

#include "common.h"
#include "engine.h"
#include "profile.h"
#include "scan.h"
#include "sort.h"

namespace cube2 {

struct DriftState {
    DevBuf key, key_sorted, val, val_sorted, xpn, sort_tmp, bounds, rows, scan_tmp, red;
    HostBuf h_red;
};
static DriftState g_dr;

struct DriftArgs {
    const int* rhoc;
    const int64_t* idx_b_r;
    const int16_t* xp;
    const float* vp;
    int64_t n_used;
    int nt, ncb, nte, nnt, ratio_cs, nlayer;
    float dt_mid;
    uint32_t key_none;  // key of a slot that leaves the tile
};

// red[0]: kept slots; red[1]: error bits (1: slot range, 2: moved beyond ncb
// cells, 4: |dz| beyond the layer spacing).
__global__ __launch_bounds__(256) void k_drift_keys(DriftArgs A, uint32_t* __restrict__ key,
                                                    uint32_t* __restrict__ val,
                                                    int16_t* __restrict__ xpn,
                                                    unsigned long long* __restrict__ red) {
#pragma clang fp contract(off)
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t q = (int64_t)blockIdx.x * 4 + wave;  // extended cell of the image
    const int64_t te3 = (int64_t)A.nte * A.nte * A.nte;
    if (q >= te3 * A.nnt * A.nnt * A.nnt) return;
    const int t = (int)(q / te3);
    const int r = (int)(q - (int64_t)t * te3);
    const int i = r % A.nte, j = (r / A.nte) % A.nte, k = r / (A.nte * A.nte);
    const int np = A.rhoc[q];
    if (np <= 0) return;
    // nzero = row end - sum(rhoc(i:)) of the row
    const int64_t row = (int64_t)t * A.nte * A.nte + (int64_t)k * A.nte + j;
    const int* rrow = A.rhoc + (q - i);
    int64_t suf = 0;
    for (int a = i; a < A.nte; a++) suf += rrow[a];
    const int64_t nzero = A.idx_b_r[row] - suf;
    if (nzero < 0 || nzero + np > A.n_used) {
        if (lane == 0) atomicOr(&red[1], 1ull);
        return;
    }
    const int src[3] = {i, j, k};
    const int dmax = (A.nlayer - 2) / 2;
    unsigned long long bad = 0;
    int kept = 0;
    for (int l = lane; l < np; l += 64) {
        const int64_t ip = nzero + l;
        int ig[3];
        int16_t code[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const int16_t v = A.xp[3 * ip + d];
            const double xq = (double)(src[d] - A.ncb) + frac16(v);  // (cell - 1) + frac
            const float f = A.dt_mid * A.vp[3 * ip + d];              // FP32 product
            const double deltax = (double)(f / (float)A.ratio_cs);    // FP32 quotient
            const int c = (int)ceil(xq + deltax);                     // 1-based target cell
            ig[d] = c - 1;                                            // 0-based physical index
            const int moved = ig[d] - (src[d] - A.ncb);
            if (moved > A.ncb || moved < -A.ncb) bad |= 2ull;
            if (d == 2 && (moved > dmax || moved < -dmax)) bad |= 4ull;
            const double sh = (double)f / (1.0 / 65536.0 * (double)A.ratio_cs);
            const long long m = llround(sh);
            code[d] = (int16_t)(uint16_t)(uint32_t)((long long)v + m);
        }
        const bool keep = ig[0] >= 0 && ig[0] < A.nt && ig[1] >= 0 && ig[1] < A.nt && ig[2] >= 0 &&
                          ig[2] < A.nt;
        uint32_t kk = A.key_none;
        if (keep) {
            const uint32_t cell = (uint32_t)(((int64_t)t * A.nt + ig[2]) * A.nt + ig[1]) * A.nt + ig[0];
            kk = cell * (uint32_t)A.nlayer + (uint32_t)(k % A.nlayer);
            kept++;
        }
        key[ip] = kk;
        val[ip] = (uint32_t)ip;
#pragma unroll
        for (int d = 0; d < 3; d++) xpn[3 * ip + d] = code[d];
    }
    // wavefront sums: one atomic per wavefront
    for (int o = 32; o > 0; o >>= 1) kept += __shfl_xor(kept, o, 64);
    bad |= __shfl_xor(bad, 32, 64);
    for (int o = 16; o > 0; o >>= 1) bad |= __shfl_xor(bad, o, 64);
    if (lane == 0) {
        if (kept) atomicAdd(&red[0], (unsigned long long)kept);
        if (bad) atomicOr(&red[1], bad);
    }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Positions (standard cells, FP64, wrapped into [0, ng)) of every slot of
// the image: a wavefront per extended cell.
__global__ __launch_bounds__(256) void k_slot_positions(const int* __restrict__ rhoc,
                                                        const int64_t* __restrict__ idx_b_r,
                                                        const int16_t* __restrict__ xp,
                                                        int64_t n_used, int nt, int ncb, int nte,
                                                        int nnt, int ratio_cs,
                                                        double* __restrict__ pos,
                                                        unsigned long long* __restrict__ err) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t q = (int64_t)blockIdx.x * 4 + wave;
    const int64_t te3 = (int64_t)nte * nte * nte;
    if (q >= te3 * nnt * nnt * nnt) return;
    const int np = rhoc[q];
    if (np <= 0) return;
    const int t = (int)(q / te3);
    const int r = (int)(q - (int64_t)t * te3);
    const int i = r % nte, j = (r / nte) % nte, k = r / (nte * nte);
    const int64_t row = (int64_t)t * nte * nte + (int64_t)k * nte + j;
    int64_t suf = 0;
    for (int a = i; a < nte; a++) suf += rhoc[q - i + a];
    const int64_t nzero = idx_b_r[row] - suf;
    if (nzero < 0 || nzero + np > n_used) {
        if (lane == 0) atomicOr(err, 1ull);
        return;
    }
    const int nc = nt * nnt;
    const int tc[3] = {t % nnt, (t / nnt) % nnt, t / (nnt * nnt)};
    const int cc[3] = {i, j, k};
    int gc[3];
#pragma unroll
    for (int d = 0; d < 3; d++) gc[d] = ((tc[d] * nt + cc[d] - ncb) % nc + nc) % nc;
    for (int l = lane; l < np; l += 64) {
        const int64_t ip = nzero + l;
#pragma unroll
        for (int d = 0; d < 3; d++)
            pos[3 * ip + d] = ((double)gc[d] + frac16(xp[3 * ip + d])) * (double)ratio_cs;
    }
}

// bounds[2c] / bounds[2c+1]: first / past-last sorted position of cell c.
__global__ void k_drift_bounds(const uint32_t* __restrict__ ks, int64_t n_kept, int nlayer,
                               int* __restrict__ bounds) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_kept) return;
    const uint32_t c = ks[p] / (uint32_t)nlayer;
    if (p == 0 || ks[p - 1] / (uint32_t)nlayer != c) bounds[2 * (int64_t)c] = (int)p;
    if (p == n_kept - 1 || ks[p + 1] / (uint32_t)nlayer != c) bounds[2 * (int64_t)c + 1] = (int)(p + 1);
}

// New rhoc over every extended cell (halo cells 0) and its row sums.
__global__ void k_drift_rhoc(const int* __restrict__ bounds, int nt, int ncb, int nte, int ntile,
                             int* __restrict__ rhoc_out, int* __restrict__ rows) {
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // (t, k, j)
    const int64_t nrows = (int64_t)ntile * nte * nte;
    if (row >= nrows) return;
    const int t = (int)(row / ((int64_t)nte * nte));
    const int rr = (int)(row - (int64_t)t * nte * nte);
    const int k = rr / nte - ncb, j = rr % nte - ncb;  // physical indices
    int* out = rhoc_out + row * nte;
    int sum = 0;
    const bool phys_row = k >= 0 && k < nt && j >= 0 && j < nt;
    for (int i = 0; i < nte; i++) {
        int c = 0;
        const int pi = i - ncb;
        if (phys_row && pi >= 0 && pi < nt) {
            const int64_t cell = (((int64_t)t * nt + k) * nt + j) * nt + pi;
            c = bounds[2 * cell + 1] - bounds[2 * cell];
        }
        out[i] = c;
        sum += c;
    }
    rows[row] = sum;
}

__global__ void k_drift_rows(const int* __restrict__ rows, int64_t nrows, int64_t* __restrict__ idx_out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrows) idx_out[r] = (int64_t)rows[r + 1];  // running total through row r
}

__global__ void k_drift_gather(const uint32_t* __restrict__ vs, int64_t n_kept,
                               const int16_t* __restrict__ xpn, const float* __restrict__ vp,
                               int16_t* __restrict__ xp_out, float* __restrict__ vp_out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_kept) return;
    const int64_t s = vs[p];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        xp_out[3 * p + d] = xpn[3 * s + d];
        vp_out[3 * p + d] = vp[3 * s + d];
    }
}

int drift(const int* d_rhoc, const int64_t* d_idx_b_r, const int16_t* d_xp, const float* d_vp,
          int64_t n_used, float dt_mid, float vz_max, int* d_rhoc_out, int64_t* d_idx_out,
          int16_t* d_xp_out, float* d_vp_out, int64_t cap, int64_t* n_out) {
    ensure_params();
    Context& C = ctx();
    const Geom& g = C.g;
    if (!d_rhoc || !d_idx_b_r || !d_rhoc_out || !d_idx_out || !n_out || n_used < 0 ||
        n_used >= ((int64_t)1 << 31) || (n_used > 0 && (!d_xp || !d_vp)) || !(dt_mid >= 0.0f) ||
        !(vz_max >= 0.0f)) {
        set_error("drift: null argument or value out of range");
        return CUBE2_ERR_ARG;
    }
    // nlayer = 2*ceiling(dt_mid*vz_max/ratio_cs)+2 (drift.f90:26), FP32
    volatile float qv = dt_mid * vz_max;
    const float qf = qv / (float)g.p.ratio_cs;
    if (!(qf <= (float)g.ncb + 1.0f)) {
        set_error("drift: dt_mid*vz_max/ratio_cs = %g exceeds the halo depth %d", (double)qf, g.ncb);
        return CUBE2_ERR_ARG;
    }
    const int nlayer = 2 * (int)std::ceil(qf) + 2;
    const int64_t ncell_phys = (int64_t)g.ntile * g.nt * g.nt * g.nt;
    if (ncell_phys * nlayer + 1 >= ((int64_t)1 << 32) || ncell_phys >= ((int64_t)1 << 30)) {
        set_error("drift: image too large for 32-bit keys");
        return CUBE2_ERR_CAPACITY;
    }
    C2_TRY(ensure_init());
    hipStream_t s = C.stream;
    const int64_t te3 = (int64_t)g.nte * g.nte * g.nte;
    const int64_t ncells = te3 * g.ntile;
    const int64_t nrows = (int64_t)g.ntile * g.nte * g.nte;
    const size_t nn = (size_t)std::max<int64_t>(n_used, 1);
    C2_TRY(g_dr.key.reserve(nn * 4));
    C2_TRY(g_dr.key_sorted.reserve(nn * 4));
    C2_TRY(g_dr.val.reserve(nn * 4));
    C2_TRY(g_dr.val_sorted.reserve(nn * 4));
    C2_TRY(g_dr.xpn.reserve(nn * 6));
    C2_TRY(g_dr.bounds.reserve((size_t)(2 * ncell_phys) * sizeof(int)));
    C2_TRY(g_dr.rows.reserve((size_t)(nrows + 1) * sizeof(int)));
    C2_TRY(g_dr.scan_tmp.reserve(scan_tmp_ints(nrows) * sizeof(int)));
    C2_TRY(g_dr.red.reserve(2 * sizeof(unsigned long long)));
    C2_TRY(g_dr.h_red.reserve(2 * sizeof(unsigned long long)));
    const uint32_t key_none = (uint32_t)(ncell_phys * nlayer);
    unsigned end_bit = 1;
    while (end_bit < 32 && ((uint64_t)1 << end_bit) <= key_none) end_bit++;
    unsigned long long* red = g_dr.red.as<unsigned long long>();
    C2_TRY(fill_zero(red, 2 * sizeof(unsigned long long), s));
    if (n_used) {  // slots no cell names (an inconsistent layout) sort last
        k_fill_u32<<<ceil_div(n_used, 256), 256, 0, s>>>(g_dr.key.as<uint32_t>(), n_used, 0xffffffffu);
        C2_HIP(hipGetLastError());
    }
    {
        Stage st("drift_keys", (double)n_used * (6 + 12 + 4 + 4 + 6));
        DriftArgs A;
        A.rhoc = d_rhoc;
        A.idx_b_r = d_idx_b_r;
        A.xp = d_xp;
        A.vp = d_vp;
        A.n_used = n_used;
        A.nt = g.nt;
        A.ncb = g.ncb;
        A.nte = g.nte;
        A.nnt = g.p.nnt;
        A.ratio_cs = g.p.ratio_cs;
        A.nlayer = nlayer;
        A.dt_mid = dt_mid;
        A.key_none = key_none;
        k_drift_keys<<<ceil_div(ncells, 4), 256, 0, s>>>(A, g_dr.key.as<uint32_t>(),
                                                         g_dr.val.as<uint32_t>(),
                                                         g_dr.xpn.as<int16_t>(), red);
        C2_HIP(hipGetLastError());
    }
    unsigned long long* h = g_dr.h_red.as<unsigned long long>();
    C2_HIP(hipMemcpyAsync(h, red, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    C2_HIP(hipStreamSynchronize(s));
    if (h[1]) {
        set_error(h[1] & 1 ? "drift: rhoc and idx_b_r describe slots outside [0, n_used)"
                  : h[1] & 2 ? "drift: a particle moved more than ncb coarse cells"
                             : "drift: a particle moved further in z than vz_max allows");
        return CUBE2_ERR_ARG;
    }
    const int64_t n_kept = (int64_t)h[0];
    *n_out = n_kept;
    if (n_kept > cap || (n_kept > 0 && (!d_xp_out || !d_vp_out))) {
        set_error("drift: %lld particles kept, capacity %lld", (long long)n_kept, (long long)cap);
        return CUBE2_ERR_CAPACITY;
    }
    {
        Stage st("drift_sort", (double)n_used * 16);
        C2_TRY(radix_sort_pairs_u32(g_dr.key.as<uint32_t>(), g_dr.key_sorted.as<uint32_t>(),
                                    g_dr.val.as<uint32_t>(), g_dr.val_sorted.as<uint32_t>(),
                                    n_used, (int)end_bit, g_dr.sort_tmp, g_dr.scan_tmp, s));
    }
    {
        Stage st("drift_cells", (double)n_kept * 4 + (double)ncells * 8);
        int* bounds = g_dr.bounds.as<int>();
        C2_TRY(fill_zero(bounds, (size_t)(2 * ncell_phys) * sizeof(int), s));
        if (n_kept) {
            k_drift_bounds<<<ceil_div(n_kept, 256), 256, 0, s>>>(g_dr.key_sorted.as<uint32_t>(),
                                                                 n_kept, nlayer, bounds);
            C2_HIP(hipGetLastError());
        }
        int* rows = g_dr.rows.as<int>();
        k_drift_rhoc<<<ceil_div(nrows, 256), 256, 0, s>>>(bounds, g.nt, g.ncb, g.nte, g.ntile,
                                                          d_rhoc_out, rows);
        C2_HIP(hipGetLastError());
        C2_TRY(exclusive_scan(rows, nrows, g_dr.scan_tmp, s));
        k_drift_rows<<<ceil_div(nrows, 256), 256, 0, s>>>(rows, nrows, d_idx_out);
        C2_HIP(hipGetLastError());
    }
    if (n_kept) {
        Stage st("drift_gather", (double)n_kept * (4 + 2 * (6 + 12)));
        k_drift_gather<<<ceil_div(n_kept, 256), 256, 0, s>>>(g_dr.val_sorted.as<uint32_t>(),
                                                             n_kept, g_dr.xpn.as<int16_t>(), d_vp,
                                                             d_xp_out, d_vp_out);
        C2_HIP(hipGetLastError());
    }
    return sync_and_flush();
}

int slot_positions(const int* d_rhoc, const int64_t* d_idx_b_r, const int16_t* d_xp,
                   int64_t n_used, double* d_pos) {
    ensure_params();
    Context& C = ctx();
    const Geom& g = C.g;
    if (!d_rhoc || !d_idx_b_r || n_used < 0 || (n_used > 0 && (!d_xp || !d_pos))) {
        set_error("positions: null argument");
        return CUBE2_ERR_ARG;
    }
    C2_TRY(ensure_init());
    hipStream_t s = C.stream;
    C2_TRY(g_dr.red.reserve(2 * sizeof(unsigned long long)));
    C2_TRY(g_dr.h_red.reserve(2 * sizeof(unsigned long long)));
    unsigned long long* red = g_dr.red.as<unsigned long long>();
    C2_TRY(fill_zero(red, 2 * sizeof(unsigned long long), s));
    const int64_t ncells = (int64_t)g.nte * g.nte * g.nte * g.ntile;
    k_slot_positions<<<ceil_div(ncells, 4), 256, 0, s>>>(d_rhoc, d_idx_b_r, d_xp, n_used, g.nt,
                                                         g.ncb, g.nte, g.p.nnt, g.p.ratio_cs, d_pos,
                                                         red);
    C2_HIP(hipGetLastError());
    unsigned long long* h = g_dr.h_red.as<unsigned long long>();
    C2_HIP(hipMemcpyAsync(h, red, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    C2_TRY(sync_and_flush());
    if (h[0]) {
        set_error("positions: rhoc and idx_b_r describe slots outside [0, n_used)");
        return CUBE2_ERR_ARG;
    }
    return CUBE2_OK;
}

void drift_release() {
    DriftState& D = g_dr;
    for (DevBuf* b : {&D.key, &D.key_sorted, &D.val, &D.val_sorted, &D.xpn, &D.sort_tmp,
                      &D.bounds, &D.rows, &D.scan_tmp, &D.red})
        b->release();
    D.h_red.release();
}

}  // namespace cube2
