// Shared definitions of the MI355X short-range force library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cube2_short.h"

namespace cube2 {

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
const char* last_error();

#define C2_HIP(call)                                                                 \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            ::cube2::set_error("HIP error %s at %s:%d (%s)", hipGetErrorString(e_),  \
                               __FILE__, __LINE__, #call);                           \
            return CUBE2_ERR_HIP;                                                    \
        }                                                                            \
    } while (0)

#define C2_TRY(expr)                  \
    do {                              \
        int rc_ = (expr);             \
        if (rc_ != CUBE2_OK) return rc_; \
    } while (0)

// ------------------------------------------------------------- geometry
struct Geom {
    cube2_params_t p;
    int nc, nt, ntt, ncb, nte, nsub, ns3, ntile;
};
Geom make_geom(const cube2_params_t& p);

// --------------------------------------------------- device memory pool
// Grow-only device buffers owned by the library (freed by the finalizers).
struct DevBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    int reserve(size_t bytes);
    void release();
    template <class T>
    T* as() const { return static_cast<T*>(ptr); }
};

// Generation of the library's device/host buffers and uploaded tables:
// bumped by every (re)allocation, release and table upload, so that a
// recorded graph can tell that the memory or tables it refers to changed.
uint64_t buffers_generation();
void buffers_changed();

// Grow-only pinned host buffer (hipHostMalloc): D2H targets of results, so
// that result copies can be recorded into a HIP graph.
struct HostBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    int reserve(size_t bytes);
    void release();
    template <class T>
    T* as() const { return static_cast<T*>(ptr); }
};

// ------------------------------------------------------------ profiling
// Per-stage GPU time, recorded with events on the library stream when
// profiling is on (cube2_profile_enable).
struct StageTimer {
    hipEvent_t a = nullptr, b = nullptr;
    const char* name = nullptr;
    bool on = false;
    void begin(const char* stage, hipStream_t s);
    void end(hipStream_t s);
};
void profile_add(const char* stage, double ms, double bytes, double flops);
bool profiling();

// --------------------------------------------------------- context
struct Context {
    bool params_set = false;
    Geom g{};
    int device = -1;
    hipStream_t stream = nullptr;
    // second library stream: PP's pass beside PM3 in an overlapped kick
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // stream the stage profiler records on (nullptr: stream)
    hipStream_t stage_stream = nullptr;
    // set while one kick's PM3 and PP share a cell-base table computed up front
    bool nzero_shared = false;
    int pp_mode = CUBE2_PP_REF;
    // velocity accumulator (dv64, FP64, 3 per slot) shared by PM3 and PP
    DevBuf dv64;
    int64_t dv_n_used = 0;
    int dv_state = 0;  // 0 none pending, 1 PM3 has written a pending Delta-v
    // host-pointer drop-in staging
    DevBuf d_xp, d_vp, d_rhoc, d_idx;
    const int16_t* staged_xp_host = nullptr;
    int64_t staged_xp_n = -1;
    DevBuf d_scal;
    HostBuf h_scal;
    int tune[CUBE2_TUNE_COUNT] = {};  // 0: built-in choice
    // true while work is being recorded into a HIP graph: no host syncs,
    // allocations, memsets or profiler events may be issued
    bool capturing = false;
};
Context& ctx();
int ensure_init();
// Serialises every C-ABI entry point (recursive: entry points call each other).
std::recursive_mutex& api_mutex();

// Per-call scalars that kernels read from device memory, so that a recorded
// graph replays with new values: {mass, a_mid, dt, 0}.  The upload copies
// from a pinned host block, which a replay reads at launch time.
int scalars_upload(float mass, float a_mid, float dt, hipStream_t s, const float** dev);   // params + HIP device + library stream
void ensure_params();  // params only (no HIP call)


// Whether the kick driver overlaps PP with PM3 (CUBE2_TUNE_KICK_OVERLAP).
inline bool kick_overlap() { return ctx().tune[CUBE2_TUNE_KICK_OVERLAP] != 2; }

// ------------------------------------------------------------- helpers
inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// Workgroups of a launch go to the 8 XCDs of the device round-robin (id b
// to XCD b % 8), and each XCD has its own L2.  This bijection of [0, W)
// gives XCD x the contiguous logical range [x*q + min(x, r), ...) (q = W/8,
// r = W%8), taken in order, so that neighbouring logical blocks (which share
// cache lines) run on one XCD at about the same time.
__device__ inline unsigned xcd_remap(unsigned b, unsigned W) {
    const unsigned q = W >> 3, r = W & 7, x = b & 7, k = b >> 3;
    return x * q + min(x, r) + k;
}

// Quantised position of a particle inside its coarse cell, as the
// reference decodes it: ((int16)(xp + ishift) + rshift) * x_resolution
// with ishift = -2^15, rshift = 0.5 - ishift, x_resolution = 2^-16
// (parameters.f90:87-88,118; pp_force_kernel.cu:764-769).
__host__ __device__ inline double frac16(int16_t v) {
    return ((double)(int16_t)((int64_t)v - 32768) + 32768.5) * (1.0 / 65536.0);
}

}  // namespace cube2
