// Library state: geometry, stream, errors, grow-only device buffers and
// the per-stage event profiler.
#include <cstdarg>
#include <cstdlib>
#include <map>
#include <mutex>

#include "common.h"
#include "profile.h"

namespace cube2 {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}
const char* last_error() { return g_err.c_str(); }

Geom make_geom(const cube2_params_t& p) {
    Geom g;
    g.p = p;
    g.nc = p.ng / p.ratio_cs;
    g.nt = g.nc / p.nnt;
    g.ntt = g.nt / p.nns;
    g.ncb = p.ngb / p.ratio_cs;
    g.nte = g.nt + 2 * g.ncb;
    g.nsub = p.nnt * p.nns;
    g.ns3 = g.nsub * g.nsub * g.nsub;
    g.ntile = p.nnt * p.nnt * p.nnt;
    return g;
}

static uint64_t g_generation = 1;
uint64_t buffers_generation() { return g_generation; }
void buffers_changed() { ++g_generation; }

int DevBuf::reserve(size_t bytes) {
    if (bytes <= cap) return CUBE2_OK;
    if (ctx().capturing) {
        set_error("device buffer growth (%zu bytes) while recording a graph", bytes);
        return CUBE2_ERR_STATE;
    }
    if (ptr) {
        C2_HIP(hipFree(ptr));
        ptr = nullptr;
        cap = 0;
    }
    buffers_changed();
    size_t want = bytes + bytes / 8 + 256;
    C2_HIP(hipMalloc(&ptr, want));
    cap = want;
    return CUBE2_OK;
}

void DevBuf::release() {
    buffers_changed();
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
}

int HostBuf::reserve(size_t bytes) {
    if (bytes <= cap) return CUBE2_OK;
    if (ctx().capturing) {
        set_error("host buffer growth (%zu bytes) while recording a graph", bytes);
        return CUBE2_ERR_STATE;
    }
    if (ptr) {
        C2_HIP(hipHostFree(ptr));
        ptr = nullptr;
        cap = 0;
    }
    buffers_changed();
    size_t want = bytes + bytes / 8 + 256;
    C2_HIP(hipHostMalloc(&ptr, want, hipHostMallocDefault));
    cap = want;
    return CUBE2_OK;
}

void HostBuf::release() {
    buffers_changed();
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    cap = 0;
}

Context& ctx() {
    static Context c;
    return c;
}

static cube2_params_t default_params() {
    // The reference's compiled defaults (cuda_parameters_generated.h).
    cube2_params_t p;
    p.ng = 512;
    p.nnt = 4;
    p.nns = 4;
    p.ratio_cs = 4;
    p.np_nc = 4;
    p.ngb = 16;
    p.ncore = 32;
    p.body_centered_cubic = 0;
    p.image_buffer = 2.0;
    p.tile_buffer = 3.0;
    p.app = 0.06;
    return p;
}

void ensure_params() {
    Context& c = ctx();
    if (!c.params_set) {
        c.g = make_geom(default_params());
        c.params_set = true;
    }
}

int scalars_upload(float mass, float a_mid, float dt, hipStream_t s, const float** dev) {
    Context& c = ctx();
    C2_TRY(c.d_scal.reserve(4 * sizeof(float)));
    C2_TRY(c.h_scal.reserve(4 * sizeof(float)));
    float* h = c.h_scal.as<float>();
    h[0] = mass;
    h[1] = a_mid;
    h[2] = dt;
    h[3] = 0.0f;
    C2_HIP(hipMemcpyAsync(c.d_scal.ptr, h, 4 * sizeof(float), hipMemcpyHostToDevice, s));
    *dev = c.d_scal.as<float>();
    return CUBE2_OK;
}

int ensure_init() {
    Context& c = ctx();
    ensure_params();
    if (!c.stream) {
        int n = 0;
        C2_HIP(hipGetDeviceCount(&n));
        if (n < 1) {
            set_error("no HIP device is visible");
            return CUBE2_ERR_HIP;
        }
        C2_HIP(hipGetDevice(&c.device));
        // Blocking streams (not hipStreamNonBlocking): work on them is ordered
        // after work a caller queued on the default (null) stream, as a
        // PyTorch caller's tensor writes are, and default-stream work after
        // theirs; the two library streams still run beside each other.
        C2_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamDefault));
        C2_HIP(hipStreamCreateWithFlags(&c.stream2, hipStreamDefault));
        C2_HIP(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
        C2_HIP(hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming));
    }
    return CUBE2_OK;
}

// ------------------------------------------------------------ profiling
struct Stat {
    double ms = 0, bytes = 0, flops = 0;
    int64_t n = 0;
};
static std::map<std::string, Stat> g_stats;
static bool g_prof = false;
struct Pending {
    std::string name;
    hipEvent_t a, b;
    double bytes, flops;
};
static std::vector<Pending> g_pending;
static std::vector<hipEvent_t> g_pool;

bool profiling() { return g_prof; }

static hipEvent_t get_event() {
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

Stage::Stage(const char* name, double bytes, double flops) {
    if (!g_prof || ctx().capturing) return;
    on_ = true;
    name_ = name;
    bytes_ = bytes;
    flops_ = flops;
    a_ = get_event();
    s_ = ctx().stage_stream ? ctx().stage_stream : ctx().stream;
    (void)hipEventRecord(a_, s_);
}

void Stage::set_work(double bytes, double flops) {
    bytes_ = bytes;
    flops_ = flops;
}

Stage::~Stage() {
    if (!on_) return;
    hipEvent_t b = get_event();
    (void)hipEventRecord(b, s_);
    g_pending.push_back({name_, a_, b, bytes_, flops_});
}

void profile_flush() {
    for (auto& p : g_pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            Stat& s = g_stats[p.name];
            s.ms += ms;
            s.n += 1;
            s.bytes += p.bytes;
            s.flops += p.flops;
        }
        g_pool.push_back(p.a);
        g_pool.push_back(p.b);
    }
    g_pending.clear();
}

void profile_add(const char* stage, double ms, double bytes, double flops) {
    Stat& s = g_stats[stage];
    s.ms += ms;
    s.n += 1;
    s.bytes += bytes;
    s.flops += flops;
}

// Stage profile of one drop-in call, switched on by an environment variable
// as the reference's CUBE_PM3_PROFILE / CUBE_PP_TIMING: the caller's
// accumulated stage statistics are set aside for the call and restored.
static std::map<std::string, Stat> g_saved_stats;
static bool g_saved_prof = false;

bool call_profile_begin(const char* env) {
    if (!std::getenv(env)) return false;
    g_saved_prof = g_prof;
    g_saved_stats.swap(g_stats);
    g_stats.clear();
    g_prof = true;
    return true;
}

std::string call_profile_end() {
    profile_flush();
    std::string out;
    char buf[96];
    for (const auto& kv : g_stats) {
        std::snprintf(buf, sizeof(buf), " %s=%.3fms", kv.first.c_str(), kv.second.ms);
        out += buf;
    }
    g_stats.swap(g_saved_stats);
    g_saved_stats.clear();
    g_prof = g_saved_prof;
    return out;
}

int sync_and_flush() {
    C2_HIP(hipStreamSynchronize(ctx().stream));
    profile_flush();
    return CUBE2_OK;
}

}  // namespace cube2

using namespace cube2;

namespace cube2 {
std::recursive_mutex& api_mutex() {
    static std::recursive_mutex m;
    return m;
}
}  // namespace cube2

extern "C" {

int cube2_set_params(const cube2_params_t* p) {
    std::lock_guard<std::recursive_mutex> l(api_mutex());
    if (!p || p->ng <= 0 || p->nnt <= 0 || p->nns <= 0 || p->ratio_cs <= 0 || p->ngb <= 0 ||
        p->ng % p->ratio_cs || (p->ng / p->ratio_cs) % p->nnt ||
        (p->ng / p->ratio_cs / p->nnt) % p->nns || p->ngb % p->ratio_cs || p->app < 0) {
        set_error("inconsistent geometry");
        return CUBE2_ERR_GEOMETRY;
    }
    Context& c = ctx();
    const cube2_params_t& q = c.g.p;
    const bool same = c.params_set && q.ng == p->ng && q.nnt == p->nnt && q.nns == p->nns &&
                      q.ratio_cs == p->ratio_cs && q.np_nc == p->np_nc && q.ngb == p->ngb &&
                      q.ncore == p->ncore && q.body_centered_cubic == p->body_centered_cubic &&
                      q.image_buffer == p->image_buffer && q.tile_buffer == p->tile_buffer &&
                      q.app == p->app;
    if (same) return CUBE2_OK;  // a pending PM3 Delta-v stays pending
    c.g = make_geom(*p);
    c.params_set = true;
    c.dv_state = 0;  // a new geometry drops a pending Delta-v
    c.staged_xp_host = nullptr;
    return CUBE2_OK;
}

int cube2_set_tuning(int key, int value) {
    static const int lo[CUBE2_TUNE_COUNT] = {1, 1, 1, 1, 3, 1, 1, 1, 1, 1},
                     hi[CUBE2_TUNE_COUNT] = {256, 7, 6, 32, 6, 2, 256, 30, 2, 2};
    if (key < 0 || key >= CUBE2_TUNE_COUNT || (value != 0 && (value < lo[key] || value > hi[key]))) {
        set_error("tuning key %d value %d out of range", key, value);
        return CUBE2_ERR_ARG;
    }
    std::lock_guard<std::recursive_mutex> l(api_mutex());
    Context& c = ctx();
    if (c.tune[key] != value) {
        c.tune[key] = value;
        buffers_changed();  // recorded launches depend on it
    }
    return CUBE2_OK;
}

int cube2_get_tuning(int key, int* value) {
    if (key < 0 || key >= CUBE2_TUNE_COUNT || !value) {
        set_error("bad tuning key %d", key);
        return CUBE2_ERR_ARG;
    }
    static const int builtin[CUBE2_TUNE_COUNT] = {0, 2, 0, 4, 4, 1, 0, 1, 1, 1};
    const int v = ctx().tune[key];
    *value = v ? v : builtin[key];
    if (key == CUBE2_TUNE_PP_BLOCKS_PER_CU && !v) *value = 0;  // occupancy-derived at launch
    return CUBE2_OK;
}

int cube2_get_params(cube2_params_t* p) {
    std::lock_guard<std::recursive_mutex> l(api_mutex());
    Context& c = ctx();
    if (!c.params_set) {
        c.g = make_geom(default_params());
        c.params_set = true;
    }
    *p = c.g.p;
    return CUBE2_OK;
}

int cube2_set_pp_mode(int mode) {
    if (mode < CUBE2_PP_EXACT || mode > CUBE2_PP_PLUMMER) {
        set_error("unknown PP mode %d", mode);
        return CUBE2_ERR_ARG;
    }
    std::lock_guard<std::recursive_mutex> l(api_mutex());
    ctx().pp_mode = mode;
    return CUBE2_OK;
}

int cube2_profile_enable(int on) {
    g_prof = on != 0;
    return CUBE2_OK;
}

int cube2_profile_reset(void) {
    g_stats.clear();
    return CUBE2_OK;
}

int cube2_profile_get(const char* stage, double* ms, int64_t* launches, double* bytes,
                      double* flops) {
    auto it = g_stats.find(stage ? stage : "");
    if (it == g_stats.end()) {
        set_error("no profile for stage %s", stage ? stage : "(null)");
        return CUBE2_ERR_ARG;
    }
    if (ms) *ms = it->second.ms;
    if (launches) *launches = it->second.n;
    if (bytes) *bytes = it->second.bytes;
    if (flops) *flops = it->second.flops;
    return CUBE2_OK;
}

int cube2_profile_stages(char* buf, int len) {
    std::string s;
    for (auto& kv : g_stats) {
        if (!s.empty()) s += ",";
        s += kv.first;
    }
    if (!buf || len <= 0) return CUBE2_ERR_ARG;
    snprintf(buf, (size_t)len, "%s", s.c_str());
    return CUBE2_OK;
}

int cube2_synchronize(void) {
    C2_TRY(ensure_init());
    return sync_and_flush();
}

const char* cube2_last_error(void) { return last_error(); }

}  // extern "C"
