// Internal entry points shared by the C ABI, the kick driver and the graph
// recorder.  Each phase split:
//   *_prepare: validation, host tables, buffer growth and table uploads
//              (may allocate and synchronise; never inside graph recording);
//   *_enqueue: kernels and result copies into pinned host memory only
//              (no host synchronisation: recordable into a HIP graph);
//   *_finish:  after the stream has completed, host-side results and errors.
#pragma once
#include <array>
#include <functional>

#include "common.h"
#include "pp_kernels.h"

namespace cube2 {

// ----------------------------------------------------------------- PP
// Host plan of the learned stand-in for one PP call (learned.hip): the
// level-2 subtiles ("jobs"), the grid sizes (input DI^3, output PO^3 fine
// cells per job) and the jobs per pass of the layers.
struct LearnedPlan {
    std::vector<LearnedJob> jobs;
    LearnedArgs args{};
    int nfs = 0, DI = 0, PO = 0, batch = 1;
};
struct PPPlan {
    std::vector<int> iloc, owner;  // subtile location -> level / -> subtile index
    int ntile = 0, nsub3 = 0;
    int64_t c_tot = 0, ncells_ext = 0, n_used = 0;
    size_t cap = 1;    // capacity of the sorted arrays (n_used)
    int mode = 0;      // CUBE2_PP_*
    int grid_pairs = 0;  // blocks of the persistent pair kernel
    bool count_tiles = false;  // count pairs per parent tile
    bool split = false;        // pair sums apart from the velocity update (pp_enqueue_apply)
    bool learned = false;      // the network stands in for PP in sparse level-2 cells
    bool all_active = false;   // every subtile has a PP range (apm3 > 0 at its level)
    bool newton = false;       // pairs of one range evaluated once (split pass)
    float react_scale = 0.0f;  // its fixed-point scale, a power of two
    LearnedPlan lp;
    std::array<int, 7> ratio_sf{};
    std::array<float, 7> apm3{};
};
int pp_validate(const int* ires, const int* ixyz3, const int* ratio_sf);
int pp_prepare(const int* ires, const int* ixyz3, const float* apm3, const int* ratio_sf,
               int64_t n_used, PPPlan& P, bool split = false, bool learned = false);
int pp_enqueue(const PPPlan& P, const int* d_rhoc, const int64_t* d_idx, const int16_t* d_xp,
               float* d_vp, const float* d_scal, const double* dv64, hipStream_t s);
// After a split pp_enqueue and whatever else writes dv64 (PM3), on stream s:
// the velocity update and vmax, then the result copies.
int pp_enqueue_apply(const PPPlan& P, float* d_vp, const double* dv64, hipStream_t s);
int pp_finish(const PPPlan& P, float* f2max, float* vmax, double* ptotal, double* ttotal,
              uint64_t* pair_counts, double wall_s);
int pp_force(const int* ires, const int* ixyz3, const float* apm3, const int* ratio_sf,
             const int* d_rhoc, const int64_t* d_idx, const int16_t* d_xp, float* d_vp,
             int64_t n_used, float mass, float a_mid, float dt, float* f2max, float* vmax,
             double* ptotal, double* ttotal, uint64_t* pair_counts, bool learned = false,
             const std::function<int(hipStream_t)>& fill_vp = nullptr);
// Copies the (x, y, z) float rows of nruns slot runs, on stream s, between a
// full per-slot array and a compact one (runs: the runs' first slots, then
// their compact offsets as an exclusive prefix of nruns + 1 entries, in
// rows); to_compact: full src -> compact dst, else compact src -> full dst.
int vp_runs_move(float* dst, const float* src, const int64_t* runs, int nruns, bool to_compact,
                 hipStream_t s);
// Blocks of the persistent pair kernel (overlap: beside PM3).
int pp_tiled_grid(int mode, bool overlap, bool newton = false);
// The last PP call's sorted particles and cell lists (for the learned
// stand-in's training inputs); CUBE2_ERR_STATE if there is none.
int pp_last(const PPPlan** P, PPArgs* A, const int** tile_first);

// Learned stand-in (learned.hip).
bool learned_model_set();
int learned_model_load(const float* const* W, const float* const* b, float y_scale, int occupancy);
int learned_model_occupancy();
void learned_release();
int learned_prepare(const Geom& g, const std::vector<int>& iloc, const int* ratio_sf,
                    const float* apm3, LearnedPlan& Lp, int occupancy);
int learned_reserve(const LearnedPlan& Lp);
int learned_enqueue(const LearnedPlan& Lp, const PPArgs& A, const int* tile_first, hipStream_t s);
int learned_training_inputs(const LearnedPlan& Lp, const PPArgs& A, const int* tile_first, float* X,
                            hipStream_t s);
int learned_training_samples(const LearnedPlan& Lp, const PPArgs& A, int cap, int* d_count,
                             int* job, int* base, float* frac, uint32_t* slot, hipStream_t s);
int learned_gather(const float* X, int DI, const int* job, const int* base, int n, float* out,
                   hipStream_t s);
int learned_debug_last(float* X, float* O, int cap_jobs, int* j0, int* nj, int* DI, int* PO,
                       hipStream_t s);
int pp_debug_celllist(int it, int* rcp, int* S, int* N, uint32_t* keys, uint32_t* slot,
                      int32_t* cell_start);
void pp_release();
int compute_nzero(const int* d_rhoc, const int64_t* d_idx, const int64_t** out, hipStream_t s);
int nzero_reserve();

// ---------------------------------------------------------------- PM3
struct PM3Batch {
    int layer = 3;            // 3: PM3 subtile jobs; 2: PM2 parent-tile jobs
    int level = 0, rsf = 0;
    size_t job0 = 0, nb = 0;  // first job in the device job table, jobs in the batch
    size_t slot0 = 0;         // first ptotal slot
};
struct PM3Plan {
    std::vector<PM3Batch> batches;
    std::vector<int> job_it;  // subtile of each ptotal slot
    const float* green[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    const float* green2 = nullptr;  // PM2's table (tile order)
    int nslots_out = 0;             // entries of the caller's ptotal (ns3, or ntile for PM2)
    int64_t n_used = 0;
};
int pm3_prepare(const int* isort, const int* ires, const int* ixyz3, const int* ratio_sf,
                int64_t n_used, const float* const* d_green, PM3Plan& P);
// d_scal: {mass, a_mid, dt} on the device (scalars_upload)
int pm3_enqueue(const PM3Plan& P, const int* d_rhoc, const int64_t* d_idx, const int16_t* d_xp,
                const float* d_scal, int accumulate, double* dv64, hipStream_t s);
int pm3_finish(const PM3Plan& P, float* f2max, double* ptotal);
int pm3_force(const int* isort, const int* ires, const int* ixyz3, const int* ratio_sf,
              const int* d_rhoc, const int64_t* d_idx, const int16_t* d_xp, int64_t n_used,
              const float* const* d_green, float mass, float amid, float dt, float* f2max,
              double* ptotal);
// PM2 (kick.f90:76-105): one job per parent tile on the standard grid with
// the Green's table Gk2 (initialize.f90:135); adds to the pending Delta-v.
int pm2_prepare(int64_t n_used, const float* d_green2, PM3Plan& P);
int pm2_force(const int* d_rhoc, const int64_t* d_idx, const int16_t* d_xp, int64_t n_used,
              const float* d_green2, float mass, float amid, float dt, float* f2max,
              double* ptotal);
int pm3_debug_job(int it, const int* ires, const int* ixyz3, const int* ratio_sf, const int* d_rhoc,
                  const int64_t* d_idx, const int16_t* d_xp, int64_t n_used, const float* d_green,
                  float mass, float* rho, float* phi, float* force, float* f2max);
int velocity_apply(const int* d_rhoc, const int64_t* d_idx, float* d_vp, int64_t n_used,
                   float* vmax);
int green_3d(float* d_G, int n, float apm_soft, float apm_range, float gridsize);
int green_file_write(const float* d_G, int n, const char* path);
int green_file_read(float* d_G, int n, const char* path);
int cnn_gemm(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk,
             int64_t sbn, float* C, int64_t ldc, int M, int N, int K, int acc);
int cnn_gemm_noalloc(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk,
                     int64_t sbn, float* C, int64_t ldc, int M, int N, int K, int acc);
void cnn_release();
// 3^3 valid convolution on the matrix cores (cin 4 or 32; cout 32, or 3
// from 32); Wp packed by cnn_conv_pack (cnn_conv_packed_count entries).
int cnn_conv_pack(const float* W, uint16_t* Wp, int cin, int cout, hipStream_t s);
size_t cnn_conv_packed_count(int cin, int cout);
int cnn_conv3(const void* X, int x_bf16, const uint16_t* Wp, const float* bias, void* Y, int y_bf16,
              int B, int D, int cin, int cout, int relu, hipStream_t s);
int cnn_im2col(const float* X, uint16_t* col, int B, int D, int Cin);
int cnn_col2im(const float* dcol, float* dX, int B, int D, int Cin);
int cnn_bias_act(float* Y, const float* bias, int64_t rows, int C, int relu);
int cnn_act_bwd(const float* Y, float* dY, uint16_t* dYb, float* dbias, int64_t rows, int C,
                int relu);
int cnn_to_bf16(const float* x, uint16_t* y, int64_t n);
int cnn_adam(float* w, const float* g, float* m, float* v, uint16_t* wb, int64_t n, float lr,
             float b1, float b2, float eps, int t);
int cnn_mse(const float* out, const float* target, float* dout, float* loss, int64_t n);
int cnn_cic_mse(const float* out8, const float* frac, const float* target, int n, float* pred,
                float* dout8, float* loss);
int fft3d(float* d, int n, int nb, int dir);
void pm3_release();
int green_cache_get(const float* host, size_t count, const float** dev);
void green_cache_release();

// --------------------------------------------------------------- kick
void heap_index(int n, const float* arr, int* indx);
int level_of(int maxcount);
int check_ixyz3(const Geom& g, const int* ixyz3);
int select_levels_dev(const int* ixyz3, const int* d_rhoc, int* ires, int* isort, int* napm);
int kick(int flags, const int* ixyz3, const float* apm3, const int* ratio_sf, const int* d_rhoc,
         const int64_t* d_idx, const int16_t* d_xp, float* d_vp, int64_t n_used,
         const float* const* d_green, float mass, float a_mid, float dt, int* ires, int* isort,
         double* ptotal_pm3, double* ptotal_pp, double* ttotal_pp, cube2_kick_result_t* out);
int kick_graph(int flags, const int* ixyz3, const float* apm3, const int* ratio_sf,
               const int* d_rhoc, const int64_t* d_idx, const int16_t* d_xp, float* d_vp,
               int64_t n_used, const float* const* d_green, float mass, float a_mid, float dt,
               const int* ires, const int* isort, double* ptotal_pm3, double* ptotal_pp,
               double* ttotal_pp, cube2_kick_result_t* out, int* recorded);
void kick_release();

// ------------------------------------------------------------- layout
int layout_build(const double* d_pos, const float* d_vel, int64_t n, int* d_rhoc,
                 int64_t* d_idx_b_r, int16_t* d_xp, float* d_vp, int64_t cap, int64_t* n_used);
int halo_velocities(const int* d_rhoc, const int64_t* d_idx_b_r, float* d_vp, int64_t n_used);
void layout_release();

// Drift of one image to its physical-only successor (drift.hip).
int drift(const int* d_rhoc, const int64_t* d_idx_b_r, const int16_t* d_xp, const float* d_vp,
          int64_t n_used, float dt_mid, float vz_max, int* d_rhoc_out, int64_t* d_idx_out,
          int16_t* d_xp_out, float* d_vp_out, int64_t cap, int64_t* n_out);
int slot_positions(const int* d_rhoc, const int64_t* d_idx_b_r, const int16_t* d_xp,
                   int64_t n_used, double* d_pos);
void drift_release();

// ---------------------------------------------------------------- misc
int fill_zero(void* d, size_t bytes, hipStream_t s);  // a kernel, not a memset
int selftest_div(uint64_t seed, int n, unsigned long long* mismatches);
int pair_forces_debug(int mode, const float* pairs, int n, float app, float* forces);
int pp_debug_tile_pairs(int ntile, uint64_t* counts);

}  // namespace cube2
