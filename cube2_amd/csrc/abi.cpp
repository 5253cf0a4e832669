// C-ABI entry points: the reference's four bind(C) symbols (host pointers,
// fatal errors) and the cube2_* device-resident API (status codes).
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "common.h"
#include "engine.h"
#include "pp_kernels.h"
#include "profile.h"

using namespace cube2;

static std::recursive_mutex& g_mx = api_mutex();  // shared with context.cpp

static int64_t n_used_of(const Geom& g, const int64_t* idx_b_r) {
    return idx_b_r[(size_t)g.ntile * g.nte * g.nte - 1];
}

// The drop-in entry points take pageable host arrays.  They cross PCIe
// through two pinned staging buffers: host threads copy chunk k into one
// while the DMA engine moves chunk k-1 from the other (a copy from pageable
// memory otherwise stages through one host thread).
namespace {
constexpr size_t kStageBytes = (size_t)32 << 20;
// Slot runs of the physical coarse cells' particles (see phys_runs).
struct Runs {
    std::vector<int64_t> dev;    // first slots, then the compact prefix (nruns + 1), in rows
    std::vector<size_t> boff;    // compact prefix in bytes of 3 floats per row (nruns + 1)
    std::vector<float*> host;    // each run's first row in the caller's vp
    int nruns = 0;
    int64_t rows = 0;
};
struct Staging {
    HostBuf buf[2];
    hipEvent_t ev[2] = {nullptr, nullptr};
    DevBuf runs, compact;  // the run table and the compact rows on the device
    Runs r;
};
Staging g_stage;

int staging_ready() {
    for (int b = 0; b < 2; b++) {
        if (g_stage.buf[b].cap < kStageBytes) C2_TRY(g_stage.buf[b].reserve(kStageBytes));
        if (!g_stage.ev[b]) C2_HIP(hipEventCreateWithFlags(&g_stage.ev[b], hipEventDisableTiming));
    }
    return CUBE2_OK;
}

// f(lo, hi) over [0, n), split among up to 8 host threads (one below 4 MB).
template <class F>
void par_bytes(size_t n, const F& f) {
    const unsigned hw = std::thread::hardware_concurrency();
    const unsigned T = std::min(8u, std::max(1u, hw / 2));
    if (T == 1 || n < ((size_t)4 << 20)) {
        f((size_t)0, n);
        return;
    }
    const size_t per = ((n + T - 1) / T + 63) & ~(size_t)63;
    std::thread th[8];
    for (unsigned t = 1; t < T; t++) {
        const size_t o = std::min(n, per * t), e = std::min(n, o + per);
        th[t] = std::thread([&f, o, e] { f(o, e); });
    }
    f((size_t)0, std::min(n, per));
    for (unsigned t = 1; t < T; t++) th[t].join();
}

// Bytes [off, off + len) of the runs' compact stream to or from a staging
// buffer that holds exactly that range.
void runs_copy(const Runs& R, char* pinned, size_t off, size_t len, bool to_pinned) {
    par_bytes(len, [&](size_t lo, size_t hi) {
        size_t a = off + lo;
        const size_t e = off + hi;
        size_t r = (size_t)(std::upper_bound(R.boff.begin(), R.boff.end(), a) - R.boff.begin()) - 1;
        while (a < e) {
            const size_t re = std::min(e, R.boff[r + 1]);
            char* h = reinterpret_cast<char*>(R.host[r]) + (a - R.boff[r]);
            char* p = pinned + (a - off);
            if (to_pinned)
                std::memcpy(p, h, re - a);
            else
                std::memcpy(h, p, re - a);
            a = re;
            r++;
        }
    });
}

// Host to device, ordered on stream s; fill(pinned, off, len) writes bytes
// [off, off + len) of the source into the staging buffer.  The source may
// change on return.
template <class Fill>
int stage_up(void* dst, size_t bytes, hipStream_t s, const Fill& fill) {
    if (!bytes) return CUBE2_OK;
    C2_TRY(staging_ready());
    for (size_t off = 0, k = 0; off < bytes; off += kStageBytes, k++) {
        const int b = (int)(k & 1);
        const size_t len = std::min(kStageBytes, bytes - off);
        C2_HIP(hipEventSynchronize(g_stage.ev[b]));  // the buffer's last DMA is done
        fill(static_cast<char*>(g_stage.buf[b].ptr), off, len);
        C2_HIP(hipMemcpyAsync(static_cast<char*>(dst) + off, g_stage.buf[b].ptr, len,
                              hipMemcpyHostToDevice, s));
        C2_HIP(hipEventRecord(g_stage.ev[b], s));
    }
    return CUBE2_OK;
}

// Device to host after the work queued on stream s; drain(pinned, off, len)
// takes bytes [off, off + len).  Complete on return.
template <class Drain>
int stage_down(const void* src, size_t bytes, hipStream_t s, const Drain& drain) {
    if (!bytes) return CUBE2_OK;
    C2_TRY(staging_ready());
    const size_t nch = (bytes + kStageBytes - 1) / kStageBytes;
    auto issue = [&](size_t k) -> int {
        const int b = (int)(k & 1);
        const size_t off = k * kStageBytes, len = std::min(kStageBytes, bytes - off);
        C2_HIP(hipMemcpyAsync(g_stage.buf[b].ptr, static_cast<const char*>(src) + off, len,
                              hipMemcpyDeviceToHost, s));
        C2_HIP(hipEventRecord(g_stage.ev[b], s));
        return CUBE2_OK;
    };
    C2_TRY(issue(0));
    for (size_t k = 0; k < nch; k++) {
        if (k + 1 < nch) C2_TRY(issue(k + 1));  // into the other buffer, already drained
        const int b = (int)(k & 1);
        const size_t off = k * kStageBytes, len = std::min(kStageBytes, bytes - off);
        C2_HIP(hipEventSynchronize(g_stage.ev[b]));
        drain(static_cast<const char*>(g_stage.buf[b].ptr), off, len);
    }
    return CUBE2_OK;
}

int copy_h2d(void* dst, const void* src, size_t bytes, hipStream_t s) {
    const char* h = static_cast<const char*>(src);
    return stage_up(dst, bytes, s, [h](char* p, size_t off, size_t len) {
        par_bytes(len, [=](size_t lo, size_t hi) { std::memcpy(p + lo, h + off + lo, hi - lo); });
    });
}

int copy_d2h(void* dst, const void* src, size_t bytes, hipStream_t s) {
    char* h = static_cast<char*>(dst);
    return stage_down(src, bytes, s, [h](const char* p, size_t off, size_t len) {
        par_bytes(len, [=](size_t lo, size_t hi) { std::memcpy(h + off + lo, p + lo, hi - lo); });
    });
}

// The slot runs of the physical coarse cells' particles, one per physical
// row (t, k, j), as k_phys_rows finds them (the first slot of a cell is
// idx_b_r of its row minus the particles of the row's cells from it on),
// adjacent runs merged.  Only these particles' velocities change in a PP
// call (halo copies are sources only), so the drop-in call moves just their
// rows of vp across PCIe.  false when a run falls outside [0, n): the call
// then moves the whole array, and the cell-list kernels report the layout.
bool phys_runs(const Geom& g, const int* rhoc, const int64_t* idx_b_r, int64_t n, float* vp,
               Runs& R) {
    R.dev.clear();
    R.host.clear();
    std::vector<int64_t> base, len;
    for (int t = 0; t < g.ntile; t++)
        for (int k = g.ncb; k < g.ncb + g.nt; k++)
            for (int j = g.ncb; j < g.ncb + g.nt; j++) {
                const int64_t row = ((int64_t)t * g.nte + k) * g.nte + j;
                const int* rc = rhoc + row * g.nte;
                int64_t tail = 0, phys = 0;
                for (int i = g.ncb + g.nt; i < g.nte; i++) tail += rc[i];
                for (int i = g.ncb; i < g.ncb + g.nt; i++) phys += rc[i];
                const int64_t e = idx_b_r[row] - tail, b = e - phys;
                if (phys < 0 || tail < 0 || b < 0 || e > n) return false;
                if (phys == 0) continue;
                if (!base.empty() && base.back() + len.back() == b)
                    len.back() += phys;
                else {
                    base.push_back(b);
                    len.push_back(phys);
                }
            }
    if (base.size() >= (size_t)INT32_MAX) return false;
    R.nruns = (int)base.size();
    R.dev = base;
    R.boff.assign(1, 0);
    int64_t acc = 0;
    R.dev.push_back(0);
    for (int r = 0; r < R.nruns; r++) {
        R.host.push_back(vp + 3 * base[r]);
        acc += len[r];
        R.dev.push_back(acc);
        R.boff.push_back((size_t)acc * 3 * sizeof(float));
    }
    R.rows = acc;
    return true;
}

// vp's physical rows up (on stream s, into d_vp) and back (after the work on
// C.stream): the run table and the compact rows on the device, the rows
// gathered from and scattered to the caller's array by host threads.
int runs_reserve() {
    const Runs& R = g_stage.r;
    C2_TRY(g_stage.runs.reserve(R.dev.size() * sizeof(int64_t)));
    return g_stage.compact.reserve(std::max<size_t>((size_t)R.rows * 3 * sizeof(float), 4));
}

int runs_h2d(float* d_vp, hipStream_t s) {  // after runs_reserve
    const Runs& R = g_stage.r;
    const size_t bytes = (size_t)R.rows * 3 * sizeof(float);
    C2_TRY(copy_h2d(g_stage.runs.ptr, R.dev.data(), R.dev.size() * sizeof(int64_t), s));
    C2_TRY(stage_up(g_stage.compact.ptr, bytes, s, [&R](char* p, size_t off, size_t len) {
        runs_copy(R, p, off, len, true);
    }));
    return vp_runs_move(d_vp, g_stage.compact.as<float>(), g_stage.runs.as<int64_t>(), R.nruns,
                        false, s);
}

int runs_d2h(const float* d_vp, hipStream_t s) {
    const Runs& R = g_stage.r;
    C2_TRY(vp_runs_move(g_stage.compact.as<float>(), d_vp, g_stage.runs.as<int64_t>(), R.nruns,
                        true, s));
    return stage_down(g_stage.compact.ptr, (size_t)R.rows * 3 * sizeof(float), s,
                      [&R](const char* p, size_t off, size_t len) {
                          runs_copy(R, const_cast<char*>(p), off, len, false);
                      });
}
}  // namespace

// Stage host arrays for the drop-in entry points.
static int stage_inputs(const int* rhoc, const int64_t* idx_b_r, const int16_t* xp, bool reuse_xp,
                        int64_t* n_used) {
    Context& C = ctx();
    const Geom& g = C.g;
    hipStream_t s = C.stream;
    const int64_t n = n_used_of(g, idx_b_r);
    if (n < 0 || n >= ((int64_t)1 << 31)) {
        set_error("invalid particle count %lld", (long long)n);
        return CUBE2_ERR_CAPACITY;
    }
    const size_t nr = (size_t)g.ntile * g.nte * g.nte * g.nte, ni = (size_t)g.ntile * g.nte * g.nte;
    C2_TRY(C.d_rhoc.reserve(nr * sizeof(int)));
    C2_TRY(C.d_idx.reserve(ni * sizeof(int64_t)));
    C2_TRY(copy_h2d(C.d_rhoc.ptr, rhoc, nr * sizeof(int), s));
    C2_TRY(copy_h2d(C.d_idx.ptr, idx_b_r, ni * sizeof(int64_t), s));
    if (!(reuse_xp && C.staged_xp_host == xp && C.staged_xp_n == n)) {
        C2_TRY(C.d_xp.reserve((size_t)std::max<int64_t>(n, 1) * 3 * sizeof(int16_t)));
        C2_TRY(copy_h2d(C.d_xp.ptr, xp, (size_t)n * 3 * sizeof(int16_t), s));
        C.staged_xp_host = xp;
        C.staged_xp_n = n;
    }
    *n_used = n;
    return CUBE2_OK;
}

[[noreturn]] static void die(const char* who, bool abort_) {
    fprintf(stderr, "%s: %s\n", who, last_error());
    if (abort_) std::abort();
    std::exit(EXIT_FAILURE);
}

// Learned PP stand-in: each operation, then a wait for the library stream.
template <class F>
static int cnn_call(F f) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    C2_TRY(f());
    C2_HIP(hipStreamSynchronize(ctx().stream));
    return CUBE2_OK;
}

extern "C" {

void c_pp_force_kernel(int* isort, int* ires, int* ixyz3, float* apm3, int* ratio_sf, int* rhoc,
                       int64_t* idx_b_r, int16_t* xp, float* vp, float mass_p_cdm, float a_mid,
                       float dt, float* f2max, float* vmax, double* ptotal, double* ttotal) {
    (void)isort;  // PP updates are disjoint per subtile; order is irrelevant
    std::lock_guard<std::recursive_mutex> l(g_mx);
    const auto t0 = std::chrono::steady_clock::now();
    // CUBE_PP_TIMING (pp_force_kernel.c:35-43): 1 a summary line, 2 also one
    // line per parent tile
    const char* tenv = std::getenv("CUBE_PP_TIMING");
    const int timing = tenv ? std::atoi(tenv) : 0;
    const bool prof = timing >= 1 && call_profile_begin("CUBE_PP_TIMING");
    if (pp_validate(ires, ixyz3, ratio_sf) != CUBE2_OK) die("[PP]", false);
    if (ensure_init() != CUBE2_OK) die("[PP]", false);
    Context& C = ctx();
    int64_t n = 0;
    if (stage_inputs(rhoc, idx_b_r, xp, C.dv_state == 1, &n) != CUBE2_OK) die("[PP]", false);
    if (C.d_vp.reserve((size_t)std::max<int64_t>(n, 1) * 3 * sizeof(float)) != CUBE2_OK) die("[PP]", false);
    // vp crosses PCIe while the pair kernel runs (it is read only by the
    // velocity update after it); with the split pass (not the exact-order
    // mode, whose pair kernel updates vp itself) only the physical particles'
    // rows cross, the halo copies' being neither read nor changed
    const bool runs = C.pp_mode != CUBE2_PP_EXACT &&
                      phys_runs(C.g, rhoc, idx_b_r, n, vp, g_stage.r) && g_stage.r.rows < n;
    if (runs && runs_reserve() != CUBE2_OK) die("[PP]", false);
    auto fill_vp = [&](hipStream_t s) {
        if (runs) return runs_h2d(C.d_vp.as<float>(), s);
        return copy_h2d(C.d_vp.ptr, vp, (size_t)n * 3 * sizeof(float), s);
    };
    if (pp_force(ires, ixyz3, apm3, ratio_sf, C.d_rhoc.as<int>(), C.d_idx.as<int64_t>(),
                 C.d_xp.as<int16_t>(), C.d_vp.as<float>(), n, mass_p_cdm, a_mid, dt, f2max, vmax,
                 ptotal, ttotal, nullptr, false, fill_vp) != CUBE2_OK)
        die("[PP]", false);
    if ((runs ? runs_d2h(C.d_vp.as<float>(), C.stream)
              : copy_d2h(vp, C.d_vp.ptr, (size_t)n * 3 * sizeof(float), C.stream)) != CUBE2_OK)
        die("[PP]", false);
    C.staged_xp_host = nullptr;  // a new kick re-uploads positions
    if (prof) {
        const Geom& g = C.g;
        const std::string st = call_profile_end();
        const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        double tsum = 0.0, psum = 0.0;
        for (int it = 0; it < g.ns3; it++) {
            tsum += ttotal[it];
            psum += ptotal[it];
        }
        if (timing >= 2) {
            const int nsub3 = g.p.nns * g.p.nns * g.p.nns;
            for (int t = 0; t < g.ntile; t++) {
                double pt = 0.0, tt = 0.0;
                for (int it = 0; it < g.ns3; it++) {
                    const int* x = ixyz3 + 6 * it;
                    if (((x[5] - 1) * g.p.nnt + x[4] - 1) * g.p.nnt + x[3] - 1 != t) continue;
                    pt += ptotal[it];
                    tt += ttotal[it];
                }
                std::fprintf(stderr, "[PP-GPU-TILE] itile=%d subtiles=%d ncount=%.0f total=%.6fs\n", t,
                             nsub3, pt, tt);
            }
        }
        std::fprintf(stderr,
                     "[PP-GPU] total=%.6fs tiles_sum=%.6fs pcount_sum=%.0f f2max=%.6e "
                     "vmax=(%.6e, %.6e, %.6e) stages:%s\n",
                     wall, tsum, psum, (double)*f2max, (double)vmax[0], (double)vmax[1],
                     (double)vmax[2], st.c_str());
    }
}

void c_pm3_force_kernel(int* isort, int* ires, int* ixyz3, int* ratio_sf, int* rhoc,
                        int64_t* idx_b_r, int16_t* xp, float* g2, float* g4, float* g6, float* g8,
                        float* g12, float* vp, float mass, float amid, float dt, float* f2max,
                        float* vmax, double* ptotal) {
    (void)vp;  // the Delta-v stays pending until c_pp_force_kernel (pm3_fft_cuda.cu:519)
    std::lock_guard<std::recursive_mutex> l(g_mx);
    const auto t0 = std::chrono::steady_clock::now();
    // CUBE_PM3_PROFILE (pm3_fft_cuda.cu:522-523,592-619): one line of stage times
    const bool prof = call_profile_begin("CUBE_PM3_PROFILE");
    if (ensure_init() != CUBE2_OK) die("PM3", true);
    Context& C = ctx();
    const Geom& g = C.g;
    int64_t n = 0;
    if (prof) {
        Stage st("xp_h2d", 0.0);
        if (stage_inputs(rhoc, idx_b_r, xp, false, &n) != CUBE2_OK) die("PM3", true);
    } else if (stage_inputs(rhoc, idx_b_r, xp, false, &n) != CUBE2_OK) {
        die("PM3", true);
    }
    const float* host_g[5] = {g2, g4, g6, g8, g12};
    const float* dev_g[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    bool need[5] = {false, false, false, false, false};
    for (int it = 0; it < g.ns3; it++)
        if (ires[it] >= 2 && ires[it] <= 6) need[ires[it] - 2] = true;
    for (int L = 2; L <= 6; L++) {
        if (!need[L - 2]) continue;
        const int nn = (g.ntt * g.p.ratio_cs + 2 * g.p.ratio_cs) * ratio_sf[L - 1];
        const size_t count = (size_t)(nn / 2 + 1) * nn * nn;
        if (!host_g[L - 2] || green_cache_get(host_g[L - 2], count, &dev_g[L - 2]) != CUBE2_OK)
            die("PM3", true);
    }
    if (pm3_force(isort, ires, ixyz3, ratio_sf, C.d_rhoc.as<int>(), C.d_idx.as<int64_t>(),
                  C.d_xp.as<int16_t>(), n, dev_g, mass, amid, dt, f2max, ptotal) != CUBE2_OK)
        die("PM3", true);
    vmax[0] = vmax[1] = vmax[2] = 0.0f;  // pm3_fft_cuda.cu:552
    if (prof) {
        const std::string st = call_profile_end();
        int napm[7] = {0, 0, 0, 0, 0, 0, 0};
        double pt = 0.0;
        for (int it = 0; it < g.ns3; it++) {
            if (ires[it] >= 1 && ires[it] <= 7) napm[ires[it] - 1]++;
            pt += ptotal[it];
        }
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        std::fprintf(stderr,
                     "[PM3 profile] subtiles L2..L6=%d/%d/%d/%d/%d particles=%.0f slots=%lld "
                     "total=%.2fms f2max=%.6e stages:%s\n",
                     napm[1], napm[2], napm[3], napm[4], napm[5], pt, (long long)n, ms,
                     (double)*f2max, st.c_str());
    }
}

void c_pm3_fft_finalize(void) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    pm3_release();
    green_cache_release();
    Context& C = ctx();
    C.d_xp.release();
    C.staged_xp_host = nullptr;
}

void c_gpu_velocity_finalize(void) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    pp_release();
    kick_release();
    layout_release();
    drift_release();
    cnn_release();
    learned_release();
    Context& C = ctx();
    C.dv64.release();
    C.d_vp.release();
    C.d_rhoc.release();
    C.d_idx.release();
    g_stage.buf[0].release();
    g_stage.buf[1].release();
    g_stage.runs.release();
    g_stage.compact.release();
    C.dv_state = 0;
    C.dv_n_used = 0;
}

int cube2_pm3_force_dev(const int* isort, const int* ires, const int* ixyz3, const int* ratio_sf,
                        const int* d_rhoc, const int64_t* d_idx_b_r, const int16_t* d_xp,
                        int64_t n_used, const float* const* d_green, float mass, float amid,
                        float dt, float* f2max, double* ptotal) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    return pm3_force(isort, ires, ixyz3, ratio_sf, d_rhoc, d_idx_b_r, d_xp, n_used, d_green, mass,
                     amid, dt, f2max, ptotal);
}

int cube2_pp_force_dev(const int* ires, const int* ixyz3, const float* apm3, const int* ratio_sf,
                       const int* d_rhoc, const int64_t* d_idx_b_r, const int16_t* d_xp,
                       float* d_vp, int64_t n_used, float mass, float a_mid, float dt,
                       float* f2max, float* vmax, double* ptotal, double* ttotal,
                       uint64_t* pair_counts) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    return pp_force(ires, ixyz3, apm3, ratio_sf, d_rhoc, d_idx_b_r, d_xp, d_vp, n_used, mass,
                    a_mid, dt, f2max, vmax, ptotal, ttotal, pair_counts);
}

int cube2_pm2_force_dev(const int* d_rhoc, const int64_t* d_idx_b_r, const int16_t* d_xp,
                        int64_t n_used, const float* d_green2, float mass, float amid, float dt,
                        float* f2max, double* ptotal) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    return pm2_force(d_rhoc, d_idx_b_r, d_xp, n_used, d_green2, mass, amid, dt, f2max, ptotal);
}

int cube2_velocity_apply_dev(const int* d_rhoc, const int64_t* d_idx_b_r, float* d_vp,
                             int64_t n_used, float* vmax) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    return velocity_apply(d_rhoc, d_idx_b_r, d_vp, n_used, vmax);
}

int cube2_green_3d_dev(float* d_G, int n, float apm_soft, float apm_range, float gridsize) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    return green_3d(d_G, n, apm_soft, apm_range, gridsize);
}

int cube2_green_file_write_dev(const float* d_G, int n, const char* path) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    return green_file_write(d_G, n, path);
}

int cube2_green_file_read_dev(float* d_G, int n, const char* path) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    return green_file_read(d_G, n, path);
}

int cube2_gemm_bf16_dev(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B,
                        int64_t sbk, int64_t sbn, float* C, int64_t ldc, int M, int N, int K,
                        int acc) {
    return cnn_call([&] { return cnn_gemm(A, sam, sak, B, sbk, sbn, C, ldc, M, N, K, acc); });
}
int cube2_gemm_bf16_noalloc_dev(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B,
                                int64_t sbk, int64_t sbn, float* C, int64_t ldc, int M, int N,
                                int K, int acc) {
    return cnn_call([&] { return cnn_gemm_noalloc(A, sam, sak, B, sbk, sbn, C, ldc, M, N, K, acc); });
}
int cube2_cnn_conv_pack_dev(const float* W, uint16_t* Wp, int cin, int cout) {
    return cnn_call([&] { return ensure_init() != CUBE2_OK ? CUBE2_ERR_HIP : cnn_conv_pack(W, Wp, cin, cout, ctx().stream); });
}
int cube2_cnn_conv3_dev(const void* X, int x_bf16, const uint16_t* Wp, const float* bias, void* Y,
                        int y_bf16, int B, int D, int cin, int cout, int relu) {
    return cnn_call([&] {
        C2_TRY(ensure_init());
        return cnn_conv3(X, x_bf16, Wp, bias, Y, y_bf16, B, D, cin, cout, relu, ctx().stream);
    });
}
int cube2_learned_set(const float* const* W, const float* const* b, float y_scale, int occupancy) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    return learned_model_load(W, b, y_scale, occupancy);
}

int cube2_learned_clear(void) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    learned_release();
    buffers_changed();
    return CUBE2_OK;
}

int cube2_learned_inputs_dev(float* d_X, int cap_jobs, int* njobs, int* DI) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    const PPPlan* P = nullptr;
    PPArgs A{};
    const int* tf = nullptr;
    C2_TRY(pp_last(&P, &A, &tf));
    LearnedPlan Lp;
    C2_TRY(learned_prepare(ctx().g, P->iloc, P->ratio_sf.data(), P->apm3.data(), Lp, 1));
    if (njobs) *njobs = (int)Lp.jobs.size();
    if (DI) *DI = Lp.DI;
    if (!d_X) return CUBE2_OK;  // a size query
    if (cap_jobs < (int)Lp.jobs.size()) {
        set_error("learned inputs: %zu level-2 subtiles, room for %d", Lp.jobs.size(), cap_jobs);
        return CUBE2_ERR_CAPACITY;
    }
    C2_TRY(learned_training_inputs(Lp, A, tf, d_X, ctx().stream));
    C2_HIP(hipStreamSynchronize(ctx().stream));
    return CUBE2_OK;
}

int cube2_learned_samples_dev(int occupancy, int cap, int* d_job, int* d_base, float* d_frac,
                              uint32_t* d_slot, int* n) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    if (occupancy < 1 || cap < 0 || !n || (cap > 0 && (!d_job || !d_base || !d_frac || !d_slot))) {
        set_error("learned samples: bad arguments");
        return CUBE2_ERR_ARG;
    }
    const PPPlan* P = nullptr;
    PPArgs A{};
    const int* tf = nullptr;
    C2_TRY(pp_last(&P, &A, &tf));
    LearnedPlan Lp;
    C2_TRY(learned_prepare(ctx().g, P->iloc, P->ratio_sf.data(), P->apm3.data(), Lp, occupancy));
    DevBuf cnt;
    C2_TRY(cnt.reserve(sizeof(int)));
    C2_TRY(learned_training_samples(Lp, A, cap, cnt.as<int>(), d_job, d_base, d_frac, d_slot,
                                    ctx().stream));
    C2_HIP(hipMemcpyAsync(n, cnt.ptr, sizeof(int), hipMemcpyDeviceToHost, ctx().stream));
    C2_HIP(hipStreamSynchronize(ctx().stream));
    cnt.release();
    if (*n > cap) {
        set_error("learned samples: %d substitutable particles, room for %d", *n, cap);
        return CUBE2_ERR_CAPACITY;
    }
    return CUBE2_OK;
}

int cube2_learned_debug_last(float* d_X, float* d_O, int cap_jobs, int* j0, int* nj, int* DI,
                             int* PO) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    if (!j0 || !nj || !DI || !PO) {
        set_error("learned debug: null argument");
        return CUBE2_ERR_ARG;
    }
    C2_TRY(ensure_init());
    return learned_debug_last(d_X, d_O, cap_jobs, j0, nj, DI, PO, ctx().stream);
}

int cube2_learned_gather_dev(const float* d_X, int DI, const int* d_job, const int* d_base, int n,
                             float* d_out) {
    return cnn_call([&] {
        C2_TRY(ensure_init());
        if (!d_X || DI < 10 || n < 0 || (n > 0 && (!d_job || !d_base || !d_out))) {
            set_error("learned gather: bad arguments");
            return (int)CUBE2_ERR_ARG;
        }
        return learned_gather(d_X, DI, d_job, d_base, n, d_out, ctx().stream);
    });
}

int cube2_cnn_im2col_dev(const float* X, uint16_t* col, int B, int D, int Cin) {
    return cnn_call([&] { return cnn_im2col(X, col, B, D, Cin); });
}
int cube2_cnn_col2im_dev(const float* dcol, float* dX, int B, int D, int Cin) {
    return cnn_call([&] { return cnn_col2im(dcol, dX, B, D, Cin); });
}
int cube2_cnn_bias_act_dev(float* Y, const float* bias, int64_t rows, int C, int relu) {
    return cnn_call([&] { return cnn_bias_act(Y, bias, rows, C, relu); });
}
int cube2_cnn_act_bwd_dev(const float* Y, float* dY, uint16_t* dYb, float* dbias, int64_t rows,
                          int C, int relu) {
    return cnn_call([&] { return cnn_act_bwd(Y, dY, dYb, dbias, rows, C, relu); });
}
int cube2_to_bf16_dev(const float* x, uint16_t* y, int64_t n) {
    return cnn_call([&] { return cnn_to_bf16(x, y, n); });
}
int cube2_cnn_adam_dev(float* w, const float* g, float* m, float* v, uint16_t* wb, int64_t n,
                       float lr, float b1, float b2, float eps, int t) {
    return cnn_call([&] { return cnn_adam(w, g, m, v, wb, n, lr, b1, b2, eps, t); });
}
int cube2_cnn_mse_dev(const float* out, const float* target, float* dout, float* loss, int64_t n) {
    return cnn_call([&] { return cnn_mse(out, target, dout, loss, n); });
}
int cube2_cnn_cic_mse_dev(const float* out8, const float* frac, const float* target, int n,
                          float* pred, float* dout8, float* loss) {
    return cnn_call([&] { return cnn_cic_mse(out8, frac, target, n, pred, dout8, loss); });
}

int cube2_fft3d_dev(float* d_data, int n, int nb, int dir) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    return fft3d(d_data, n, nb, dir);
}

int cube2_select_levels(const int* rhoc, const int* ixyz3, int* ires, int* isort, int* napm) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    Context& C = ctx();
    if (!C.params_set) {
        cube2_params_t p;
        cube2_get_params(&p);
    }
    const Geom& g = C.g;
    if (!rhoc || !ixyz3 || !ires || !isort) {
        set_error("null argument");
        return CUBE2_ERR_ARG;
    }
    C2_TRY(check_ixyz3(g, ixyz3));
    std::vector<float> neg(g.ns3);
    int nap[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int it = 0; it < g.ns3; it++) {
        const int* x = ixyz3 + 6 * it;
        const int t = ((x[5] - 1) * g.p.nnt + (x[4] - 1)) * g.p.nnt + (x[3] - 1);
        const int* rc = rhoc + (size_t)t * g.nte * g.nte * g.nte;
        int mx = 0;
        for (int k = (x[2] - 1) * g.ntt; k < x[2] * g.ntt; k++)
            for (int j = (x[1] - 1) * g.ntt; j < x[1] * g.ntt; j++)
                for (int i = (x[0] - 1) * g.ntt; i < x[0] * g.ntt; i++)
                    mx = std::max(mx, rc[((size_t)(k + g.ncb) * g.nte + (j + g.ncb)) * g.nte + (i + g.ncb)]);
        const int L = level_of(mx);
        ires[it] = L;
        nap[L - 1]++;
        neg[it] = -((float)mx / 64.0f);
    }
    heap_index(g.ns3, neg.data(), isort);
    if (napm)
        for (int q = 0; q < 7; q++) napm[q] = nap[q];
    return CUBE2_OK;
}

int cube2_kick_dev(int flags, const int* ixyz3, const float* apm3, const int* ratio_sf,
                   const int* d_rhoc, const int64_t* d_idx_b_r, const int16_t* d_xp, float* d_vp,
                   int64_t n_used, const float* const* d_green, float mass, float a_mid, float dt,
                   int* ires, int* isort, double* ptotal_pm3, double* ptotal_pp,
                   double* ttotal_pp, cube2_kick_result_t* out) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    return kick(flags, ixyz3, apm3, ratio_sf, d_rhoc, d_idx_b_r, d_xp, d_vp, n_used, d_green, mass,
                a_mid, dt, ires, isort, ptotal_pm3, ptotal_pp, ttotal_pp, out);
}

int cube2_kick_graph_dev(int flags, const int* ixyz3, const float* apm3, const int* ratio_sf,
                         const int* d_rhoc, const int64_t* d_idx_b_r, const int16_t* d_xp,
                         float* d_vp, int64_t n_used, const float* const* d_green, float mass,
                         float a_mid, float dt, const int* ires, const int* isort,
                         double* ptotal_pm3, double* ptotal_pp, double* ttotal_pp,
                         cube2_kick_result_t* out, int* recorded) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    return kick_graph(flags, ixyz3, apm3, ratio_sf, d_rhoc, d_idx_b_r, d_xp, d_vp, n_used,
                      d_green, mass, a_mid, dt, ires, isort, ptotal_pm3, ptotal_pp, ttotal_pp,
                      out, recorded);
}

int cube2_layout_build_dev(const double* d_pos, const float* d_vel, int64_t n, int* d_rhoc,
                           int64_t* d_idx_b_r, int16_t* d_xp, float* d_vp, int64_t cap,
                           int64_t* n_used) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    return layout_build(d_pos, d_vel, n, d_rhoc, d_idx_b_r, d_xp, d_vp, cap, n_used);
}

int cube2_halo_velocities_dev(const int* d_rhoc, const int64_t* d_idx_b_r, float* d_vp,
                              int64_t n_used) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    return halo_velocities(d_rhoc, d_idx_b_r, d_vp, n_used);
}

int cube2_drift_dev(const int* d_rhoc, const int64_t* d_idx_b_r, const int16_t* d_xp,
                    const float* d_vp, int64_t n_used, float dt_mid, float vz_max,
                    int* d_rhoc_out, int64_t* d_idx_b_r_out, int16_t* d_xp_out, float* d_vp_out,
                    int64_t cap, int64_t* n_out) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    return drift(d_rhoc, d_idx_b_r, d_xp, d_vp, n_used, dt_mid, vz_max, d_rhoc_out, d_idx_b_r_out,
                 d_xp_out, d_vp_out, cap, n_out);
}

int cube2_slot_positions_dev(const int* d_rhoc, const int64_t* d_idx_b_r, const int16_t* d_xp,
                             int64_t n_used, double* d_pos) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    return slot_positions(d_rhoc, d_idx_b_r, d_xp, n_used, d_pos);
}

int cube2_selftest_div(uint64_t seed, int n, unsigned long long* mismatches) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    if (n < 0 || !mismatches) {
        set_error("bad self-test arguments");
        return CUBE2_ERR_ARG;
    }
    return selftest_div(seed, n, mismatches);
}

int cube2_pp_count_tiles(int on) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    if (pp_count_tiles() != (on != 0)) {
        pp_count_tiles() = on != 0;
        buffers_changed();  // a recorded kick records again with the new kernel arguments
    }
    return CUBE2_OK;
}

int cube2_pp_debug_tile_pairs(int ntile, uint64_t* counts) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    return pp_debug_tile_pairs(ntile, counts);
}

int cube2_pp_debug_pair_forces(int mode, const float* pairs, int n, float app, float* forces) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    return pair_forces_debug(mode, pairs, n, app, forces);
}

int cube2_pp_debug_celllist(int it, int* rcp, int* S, int* N, uint32_t* keys_sorted,
                            uint32_t* slot, int32_t* cell_start) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    return pp_debug_celllist(it, rcp, S, N, keys_sorted, slot, cell_start);
}

int cube2_pm3_debug_job(int it, const int* ires, const int* ixyz3, const int* ratio_sf,
                        const int* d_rhoc, const int64_t* d_idx_b_r, const int16_t* d_xp,
                        int64_t n_used, const float* d_green, float mass, float* rho, float* phi,
                        float* force, float* f2max) {
    std::lock_guard<std::recursive_mutex> l(g_mx);
    return pm3_debug_job(it, ires, ixyz3, ratio_sf, d_rhoc, d_idx_b_r, d_xp, n_used, d_green, mass,
                         rho, phi, force, f2max);
}

}  // extern "C"
