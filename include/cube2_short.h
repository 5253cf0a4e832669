/*
 * libcube2_short -- MI355X (gfx950) short-range force path of CUBE2:
 * fine-mesh PM3 and direct-summation PP of the kick (kick.f90:107-236).
 *
 * C ABI.  Plain pointers and sizes only.  The first four entry points keep
 * the names, argument order and meaning of the reference's bind(C)
 * interfaces so the Fortran host links against this library unchanged.
 * Errors in those four print to stderr and terminate the process, as the
 * reference does (pp_force_kernel.cu:68-76 exit(EXIT_FAILURE);
 * pm3_fft_cuda.cu:16-31 abort()).  Every cube2_* entry point instead returns
 * a status code (CUBE2_OK == 0) and leaves a message in cube2_last_error().
 *
 * Array layouts (Fortran column-major seen from C):
 *   rhoc     int32  [nnt^3][nte][nte][nte]   (rhoc(i,j,k,tx,ty,tz))
 *   idx_b_r  int64  [nnt^3][nte][nte]        (idx_b_r(j,k,tx,ty,tz))
 *   xp       int16  [N_used][3]              N_used = idx_b_r[last]
 *   vp       float  [N_used][3]
 *   ixyz3    int32  [ns3][6] = {sx,sy,sz,tx,ty,tz}, 1-based
 *   ires     int32  [ns3] in 2..6;  isort int32 [ns3], 1-based subtile ids
 *   Gk3_L    float  [n][n][n/2+1] (Gk3(n/2+1,n,n)), n = nfp+2*nfb at level L
 * with nte = nt + 2*ncb, nt = ng/ratio_cs/nnt, ncb = ngb/ratio_cs, ns3 = (nnt*nns)^3.
 */
#ifndef CUBE2_SHORT_H
#define CUBE2_SHORT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CUBE2_OK = 0,
    CUBE2_ERR_ARG = 1,      /* invalid argument or metadata */
    CUBE2_ERR_GEOMETRY = 2, /* geometry not set or inconsistent */
    CUBE2_ERR_CAPACITY = 3, /* size beyond a supported limit */
    CUBE2_ERR_HIP = 4,      /* HIP runtime failure */
    CUBE2_ERR_STATE = 5     /* call order (velocity accumulator) violated */
};

enum {
    CUBE2_PP_EXACT = 0,
    CUBE2_PP_FAST = 1,
    CUBE2_PP_REF = 2,
    CUBE2_PP_PRECISE = 3,
    CUBE2_PP_PLUMMER = 4
};

/* ---------------------------------------------------------------------
 * Drop-in entry points (host pointers; the Fortran caller owns every array)
 * ------------------------------------------------------------------- */

/* Replaces pp_force_kernel.cu:960-963 (interface kick.f90:13-30, call site
 * kick.f90:202-205).  PP over every parent tile; applies the pending PM3
 * Delta-v together with the PP Delta-v to vp; f2max = max |a*m^2|^2,
 * vmax[d] = max |vp[d]| over physical particles; ptotal[it] = physical
 * particles of subtile it; ttotal[first subtile of each parent tile] =
 * that tile's share of the PP wall time.  isort is ignored, as in the
 * reference (pp_force_kernel.cu:1049). */
void c_pp_force_kernel(int* isort, int* ires, int* ixyz3, float* apm3, int* ratio_sf, int* rhoc,
                       int64_t* idx_b_r, int16_t* xp, float* vp, float mass_p_cdm, float a_mid,
                       float dt, float* f2max, float* vmax, double* ptotal, double* ttotal);

/* Replaces pm3_fft_cuda.cu:516-520 (interface kick.f90:31-45, call site
 * kick.f90:144-146).  PM3 over every subtile; leaves the Delta-v pending
 * for the next c_pp_force_kernel (vp is not touched); f2max = max over
 * subtile interiors of |F|^2; vmax = 0; ptotal[it] = physical particles of
 * subtile it (pm3_fft_cuda.cu:436-437).  g2..g12 are the Gk3 tables of
 * levels 2..6, cached on the device by (pointer, size) as the reference
 * does (pm3_fft_cuda.cu:153-159): their contents must not change at a
 * given address. */
void c_pm3_force_kernel(int* isort, int* ires, int* ixyz3, int* ratio_sf, int* rhoc,
                        int64_t* idx_b_r, int16_t* xp, float* g2, float* g4, float* g6, float* g8,
                        float* g12, float* vp, float mass, float amid, float dt, float* f2max,
                        float* vmax, double* ptotal);

/* Replaces pm3_fft_cuda.cu:621-628 (finalize.f90:19-20): frees PM3 device
 * workspaces and the cached Green's tables. */
void c_pm3_fft_finalize(void);

/* Replaces pp_force_kernel.cu:539-547 (finalize.f90:19): frees PP device
 * workspaces and the velocity accumulator. */
void c_gpu_velocity_finalize(void);

/* ---------------------------------------------------------------------
 * Runtime geometry.  Replaces the compile-time cuda_parameters_generated.h
 * (its eleven names, tools/gen_cuda_parameters.py:9-21).  Until it is
 * called the library uses the reference's default values.
 * ------------------------------------------------------------------- */
typedef struct {
    int ng, nnt, nns, ratio_cs, np_nc, ngb, ncore, body_centered_cubic;
    double image_buffer, tile_buffer, app;
} cube2_params_t;

/* Setting the current values again is a no-op; a different geometry drops
 * a pending PM3 Delta-v. */
int cube2_set_params(const cube2_params_t* p);
int cube2_get_params(cube2_params_t* p);

/* PP pair kernel:
 *   CUBE2_PP_REF (default): the reference C kernel's pair arithmetic
 *     operation for operation (bit-identical pair forces), work-compacted;
 *     per-particle sums in another order (relative RMS ~4e-7 vs reference;
 *     with PP_NEWTON built in, each pair of one range is evaluated once for
 *     both particles and the sums are 64-bit fixed point);
 *   CUBE2_PP_EXACT: one lane per particle, the reference's summation order
 *     too: Delta-v bit-identical to pp_force_kernel.c (slow, for checking);
 *   CUBE2_PP_FAST: FMA/rsqrt arithmetic, work-compacted (its own rounding;
 *     accuracy stated against an FP64 direct sum);
 *   CUBE2_PP_PRECISE: cancellation-free force law near the cutoff and FP64
 *     per-particle sums (accuracy stated against an FP64 direct sum);
 *   CUBE2_PP_PLUMMER: Plummer softening, no counterpart in the reference:
 *     the pair force r / (r^2 + app^2)^(3/2) - F_ra(r, range_i) (app is the
 *     softening length eps), in PRECISE's arithmetic (accuracy stated
 *     against an FP64 direct sum of the same law). */
int cube2_set_pp_mode(int mode);

/* Launch-shape tunables (no counterpart in the reference; set by the
 * autotuner in cube2_amd/autotune.py).  value 0 restores the built-in
 * choice.  Valid values:
 *   PP_BLOCKS_PER_CU  1..256, blocks per CU of the pair kernel (built-in:
 *                     64, beyond the resident maximum, so that the hardware
 *                     balances the chunks' uneven costs);
 *   PP_DIRECT_LG      1..7, chunks of > 2^(v-1) particles (lanes for >= 2^v
 *                     i slots) sum directly, smaller ones queue pairs (2);
 *   FFT_MIN_LGC       1..6, minimum log2 column-tile width of the y/z FFT
 *                     passes (built-in: ~2048 LDS elements per tile);
 *   INTERP_SEG        1..32, particles per wavefront work item of the mesh
 *                     interpolation, in units of 64 (4);
 *   PP_DENSE_LG       3..6, log2 of the pair kernel's chunk size in PP
 *                     cells of >= 64 particles (4; 5 when PP_NEWTON
 *                     evaluates pairs once);
 *   KICK_OVERLAP      1: the kick driver runs PP's cell lists and pair sums
 *                     on a second stream beside PM3, the velocity update
 *                     after both (built-in); 2: PM3, then PP;
 *   PP_OVERLAP_BLOCKS 1..256, blocks per CU of the pair kernel while PM3
 *                     runs beside it (built-in: as PP_BLOCKS_PER_CU);
 *   PP_ORDER          1: the pair kernel takes its chunks in cell order
 *                     (built-in); 2..30: chunks of an estimated 2^v
 *                     candidate pairs or more first, heaviest first, then
 *                     the rest in cell order;
 *   PP_JBATCH         1: the pair kernel fetches and tests j candidates 64
 *                     at a time (built-in); 2: 128 at a time;
 *   PP_NEWTON         1: REF and FAST passes (not the learned kick)
 *                     evaluate each pair of two active particles of one
 *                     range once and add its force to both, summing in
 *                     64-bit fixed point at 2^k, k = min(40, floor(62 -
 *                     log2(10 cap / app^2))), cap the particle capacity
 *                     (built-in; used when app > 0 and k >= 12, else as 2);
 *                     2: every ordered pair is evaluated by its i.
 * A change makes the next recorded kick record again. */
enum {
    CUBE2_TUNE_PP_BLOCKS_PER_CU = 0,
    CUBE2_TUNE_PP_DIRECT_LG = 1,
    CUBE2_TUNE_FFT_MIN_LGC = 2,
    CUBE2_TUNE_INTERP_SEG = 3,
    CUBE2_TUNE_PP_DENSE_LG = 4,
    CUBE2_TUNE_KICK_OVERLAP = 5,
    CUBE2_TUNE_PP_OVERLAP_BLOCKS = 6,
    CUBE2_TUNE_PP_ORDER = 7,
    CUBE2_TUNE_PP_JBATCH = 8,
    CUBE2_TUNE_PP_NEWTON = 9,
    CUBE2_TUNE_COUNT = 10
};
int cube2_set_tuning(int key, int value);
int cube2_get_tuning(int key, int* value); /* the value in use (built-in choice resolved) */

/* ---------------------------------------------------------------------
 * Device-resident variants (no counterpart in the reference; they are the
 * same path with particle and cell arrays already in HBM).  rhoc, idx_b_r,
 * xp, vp are device pointers; isort/ires/ixyz3/apm3/ratio_sf and outputs
 * f2max/vmax/ptotal/ttotal are host pointers.  n_used = idx_b_r[last].
 * Work runs on the library stream; calls return after it completes.  The
 * library's streams are blocking streams: their work starts after work the
 * caller queued earlier on the default (null) stream, such as a PyTorch
 * caller's tensor writes.
 * ------------------------------------------------------------------- */
int cube2_pm3_force_dev(const int* isort, const int* ires, const int* ixyz3, const int* ratio_sf,
                        const int* d_rhoc, const int64_t* d_idx_b_r, const int16_t* d_xp,
                        int64_t n_used, const float* const* d_green /* [5]: levels 2..6 */,
                        float mass, float amid, float dt, float* f2max, double* ptotal);

int cube2_pp_force_dev(const int* ires, const int* ixyz3, const float* apm3, const int* ratio_sf,
                       const int* d_rhoc, const int64_t* d_idx_b_r, const int16_t* d_xp,
                       float* d_vp, int64_t n_used, float mass, float a_mid, float dt,
                       float* f2max, float* vmax, double* ptotal, double* ttotal,
                       uint64_t* pair_counts /* [2] P_cand, P_acc; may be NULL */);

/* PM2 (kick.f90:76-105; the reference runs it on the CPU, pm_thread with
 * FFTW): one job per parent tile on the standard grid, mesh n = ngp + 2 ngb
 * (ngp = ng / nnt), the particles of all ncb halo cells deposited, the
 * Green's table d_green2 = Gk2 [n][n][n/2+1] (initialize.f90:135:
 * Green_3D(n, apm2 = 3.5, apm1 = 3.5 ratio_cs, gridsize 1)).  Its Delta-v
 * joins the pending FP64 accumulator like PM3's (applied by PP or
 * cube2_velocity_apply_dev).  ptotal [nnt^3]: physical particles per tile. */
int cube2_pm2_force_dev(const int* d_rhoc, const int64_t* d_idx_b_r, const int16_t* d_xp,
                        int64_t n_used, const float* d_green2, float mass, float amid, float dt,
                        float* f2max, double* ptotal);

/* Applies a pending PM3 Delta-v without PP (the reference leaves it
 * pending until PP, pp_force_kernel.cu:496-498); vmax over physical slots. */
int cube2_velocity_apply_dev(const int* d_rhoc, const int64_t* d_idx_b_r, float* d_vp,
                             int64_t n_used, float* vmax);

/* Green_3D (Green/Green.f90:1-80) in FP64 on the device, stored FP32:
 * d_G[n][n][n/2+1] for apm_soft, apm_range, gridsize as in
 * initialize.f90:142-168. */
int cube2_green_3d_dev(float* d_G, int n, float apm_soft, float apm_range, float gridsize);

/* Green's table cache files (initialize.f90:102-175, the read_Gks switch):
 * Fortran access='stream' files of real(4) Gk(n/2+1,n,n), column-major, no
 * record markers, i.e. the bytes of d_G[n][n][n/2+1]; the reference names
 * them output_dir()//'Gk2' or 'Gk3_<ratio_sf>' //'_<image>.bin'
 * (basic_functions.f08:76-81).  write: d_G to path, replacing the file.
 * read: the first n*n*(n/2+1) floats of path into d_G; CUBE2_ERR_ARG if the
 * file is missing or shorter (the Fortran read fails at end of file). */
int cube2_green_file_write_dev(const float* d_G, int n, const char* path);
int cube2_green_file_read_dev(float* d_G, int n, const char* path);

/* ---------------------------------------------------------------------
 * Learned PP stand-in (no counterpart in the reference): the operations of
 * a 3-D CNN of 3^3 valid convolutions, channels-last FP32 activations
 * [patch][z][y][x][c], BF16 operands (uint16 bit patterns) on the matrix
 * cores with FP32 accumulation.  Device pointers; calls return after the
 * work completes.  cube2_amd/learned.py builds the network on them.
 * ------------------------------------------------------------------- */
/* C[m*ldc + n] (+= if acc) sum_k A[m*sam + k*sak] * B[k*sbk + n*sbn]. */
int cube2_gemm_bf16_dev(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B,
                        int64_t sbk, int64_t sbn, float* C, int64_t ldc, int M, int N, int K,
                        int acc);
/* The same product, allocating nothing: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Split-K partials live in
 * a static 4 MiB device array per GPU; a split shape needing more returns
 * CUBE2_ERR_CAPACITY (S*M*N*4 bytes, S the split count).  Results equal
 * cube2_gemm_bf16_dev's bit for bit.  Calls that overlap on one GPU race. */
int cube2_gemm_bf16_noalloc_dev(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B,
                                int64_t sbk, int64_t sbn, float* C, int64_t ldc, int M, int N,
                                int K, int acc);
/* Layer weights W[27*cin][cout] (k = tap*cin + c, tap = (dz*3+dy)*3+dx; FP32)
 * packed as BF16 for cube2_cnn_conv3_dev: Wp holds T*Cp*cin entries, T = 27
 * (cin 32) or 32 (cin 4), Cp = 32 (cout 32) or 16 (cout 3). */
int cube2_cnn_conv_pack_dev(const float* W, uint16_t* Wp, int cin, int cout);
/* Y[b][z][y][x][o] = act(bias[o] + sum_{tap, c} X[b][z+dz][y+dy][x+dx][c] W[tap*cin+c][o])
 * for the (D-2)^3 outputs of each of B grids of D^3 x cin cells
 * (channels-last), on the matrix cores (BF16 operands, FP32 accumulation),
 * with no im2col.  act = ReLU if relu.  x_bf16 / y_bf16: X / Y hold BF16
 * bit patterns (else FP32).  Kernels: cin 4 -> cout 32 (X FP32),
 * 32 -> 32, 32 -> 3 (Y FP32). */
int cube2_cnn_conv3_dev(const void* X, int x_bf16, const uint16_t* Wp, const float* bias, void* Y,
                        int y_bf16, int B, int D, int cin, int cout, int relu);
/* Learned stand-in in the kick (CUBE2_KICK_LEARNED).  The network: four
 * valid 3^3 convolutions 4 -> 32 -> 32 -> 32 -> 3 channels, ReLU after the
 * first three, over fine cells of the level-2 mesh; input channels
 * log(1 + count) and the mean offset of the cell's particles from its centre
 * (fine cells, x y z); a particle's acceleration is y_scale times the
 * cloud-in-cell interpolation of the output at its position.  It stands in
 * for PP for the physical particles of level-2 subtiles in PP cells of fewer
 * than `occupancy` particles.  W[l]: 27*cin*cout FP32 (host, k = tap*cin +
 * c), b[l]: cout FP32 (host). */
int cube2_learned_set(const float* const* W /* [4] */, const float* const* b /* [4] */,
                      float y_scale, int occupancy);
int cube2_learned_clear(void);
/* Training data after a PP call (tiled modes): the input grids of every
 * level-2 subtile, d_X [njobs][DI][DI][DI][4] FP32 (d_X NULL: sizes only),
 * and the particles the network would stand in for: their subtile (job), the
 * origin of their 10^3 input patch in the job's grid (= the CIC lower corner
 * in the job's output grid), the CIC fractions, and their slot. */
int cube2_learned_inputs_dev(float* d_X, int cap_jobs, int* njobs, int* DI);
int cube2_learned_samples_dev(int occupancy, int cap, int* d_job, int* d_base /* [cap][3] */,
                              float* d_frac /* [cap][3] */, uint32_t* d_slot, int* n);
/* The last learned kick's last pass of the layers (for checks): jobs
 * [*j0, *j0 + *nj), their input grids d_X [nj][DI^3][4] and outputs d_O
 * [nj][PO^3][3] (either may be NULL; both NULL: sizes only). */
int cube2_learned_debug_last(float* d_X, float* d_O, int cap_jobs, int* j0, int* nj, int* DI,
                             int* PO);
/* d_out[i] = the 10^3 x 4 patch of d_X[d_job[i]] at d_base[i] (zeros outside). */
int cube2_learned_gather_dev(const float* d_X, int DI, const int* d_job, const int* d_base, int n,
                             float* d_out);
/* col[(p,z,y,x)][((dz*3+dy)*3+dx)*Cin + c] = X[p][z+dz][y+dy][x+dx][c] (BF16),
 * for the (D-2)^3 outputs of each of B inputs of D^3 x Cin. */
int cube2_cnn_im2col_dev(const float* X, uint16_t* col, int B, int D, int Cin);
/* The adjoint: dX (B x D^3 x Cin) = sum of the dcol entries that read each element. */
int cube2_cnn_col2im_dev(const float* dcol, float* dX, int B, int D, int Cin);
/* Y[r][c] += bias[c], then max(Y, 0) if relu. */
int cube2_cnn_bias_act_dev(float* Y, const float* bias, int64_t rows, int C, int relu);
/* dY *= (Y > 0) if relu; dYb = BF16(dY); dbias[c] = sum_r dY[r][c]. */
int cube2_cnn_act_bwd_dev(const float* Y, float* dY, uint16_t* dYb, float* dbias, int64_t rows,
                          int C, int relu);
int cube2_to_bf16_dev(const float* x, uint16_t* y, int64_t n);
/* One Adam step (t >= 1, bias-corrected) on FP32 weights; wb (may be NULL)
 * receives their BF16 copy. */
int cube2_cnn_adam_dev(float* w, const float* g, float* m, float* v, uint16_t* wb, int64_t n,
                       float lr, float b1, float b2, float eps, int t);
/* *loss = mean (out - target)^2 over n values, dout = 2 (out - target) / n. */
int cube2_cnn_mse_dev(const float* out, const float* target, float* dout, float* loss, int64_t n);
/* pred[i][d] = sum over the 2^3 corners c (x fastest) of the CIC weight of
 * frac[i] times out8[i][c][d]; loss = mean over n*3 of (pred - target)^2;
 * dout8 = its gradient with respect to out8.  pred may be NULL. */
int cube2_cnn_cic_mse_dev(const float* out8, const float* frac, const float* target, int n,
                          float* pred, float* dout8, float* loss);

/* Level selection and job order of kick.f90:107-131 (host arrays). */
int cube2_select_levels(const int* rhoc, const int* ixyz3, int* ires, int* isort, int* napm);

/* ---------------------------------------------------------------------
 * Image layout on the device (no counterpart in the reference's CUDA
 * files; it is the device form of buffer_grid / spine_image and the
 * single-image halo copies, buffer_grid.f90:13-155, variables.f90:146-216,
 * buffer_particle.f90:23-345).  From n positions d_pos[n][3] (FP64,
 * standard cells, periodic in [0, ng)) and optional velocities d_vel[n][3]
 * (NULL: zero) it writes d_rhoc [ntile][nte^3], d_idx_b_r [ntile][nte^2]
 * and, when cap >= the slot count, d_xp [n_used][3] and d_vp [n_used][3].
 * *n_used is always written; CUBE2_ERR_CAPACITY when cap is too small (or
 * d_xp/d_vp NULL), so a caller can size and call again.  The particles of
 * one cell keep their input order (a stable sort).
 * ------------------------------------------------------------------- */
int cube2_layout_build_dev(const double* d_pos, const float* d_vel, int64_t n, int* d_rhoc,
                           int64_t* d_idx_b_r, int16_t* d_xp, float* d_vp, int64_t cap,
                           int64_t* n_used);

/* buffer_v for one periodic image (buffer_particle.f90:11-21, through
 * buffer_xvp :23-345): copies the velocities of every halo cell's slots
 * from its periodic source cell.  A kick updates the physical slots only;
 * call this between the kick and cube2_drift_dev, as the reference's step
 * runs kick then buffer_v (main.f90:26-31). */
int cube2_halo_velocities_dev(const int* d_rhoc, const int64_t* d_idx_b_r, float* d_vp,
                              int64_t n_used);

/* ---------------------------------------------------------------------
 * Drift (drift.f90:1-158 for one image, then spine_image of
 * variables.f90:146-216).  The reference runs it on the host, per tile
 * with OpenMP.  From the image with halo copies (the kick's d_rhoc,
 * d_idx_b_r, d_xp, d_vp, n_used) every slot moves by dt_mid*vp; the
 * particles landing in a physical cell of their slot's tile form the new
 * physical-only image: d_rhoc_out [ntile][nte^3] (halo cells 0),
 * d_idx_b_r_out [ntile][nte^2], d_xp_out / d_vp_out [*n_out][3] (cap
 * entries; n_used always suffices).  Within a cell the particles keep the
 * arrival order of drift.f90's layered loop.  vz_max bounds |vp_z| as
 * sim%vz_max does (kick.f90:239; drift.f90:26 sets the layer spacing from
 * it); CUBE2_ERR_ARG if a particle moves further in z than vz_max allows
 * or more than ncb coarse cells in any direction.  The halo copies for the
 * next kick come from cube2_slot_positions_dev + cube2_layout_build_dev.
 * ------------------------------------------------------------------- */
int cube2_drift_dev(const int* d_rhoc, const int64_t* d_idx_b_r, const int16_t* d_xp,
                    const float* d_vp, int64_t n_used, float dt_mid, float vz_max,
                    int* d_rhoc_out, int64_t* d_idx_b_r_out, int16_t* d_xp_out, float* d_vp_out,
                    int64_t cap, int64_t* n_out);

/* Positions d_pos[n_used][3] (FP64 standard cells, wrapped into [0, ng))
 * of every slot of an image, halo copies included: (global coarse cell +
 * the decoded in-cell offset, parameters.f90:87-88,118) * ratio_cs. */
int cube2_slot_positions_dev(const int* d_rhoc, const int64_t* d_idx_b_r, const int16_t* d_xp,
                             int64_t n_used, double* d_pos);

/* ---------------------------------------------------------------------
 * Kick driver: the short-range part of kick.f90:107-238 in one call over
 * device-resident arrays (the Fortran host's level selection, the PM3 and
 * PP calls of kick.f90:144-146,202-205, and the dt limits of
 * kick.f90:180,227,236).
 * ------------------------------------------------------------------- */
typedef struct {
    float f2max_pm2;                /* kick.f90:96 (CUBE2_KICK_PM2) */
    float f2max_pm3, f2max_pp;      /* kick.f90:146,205 */
    float vmax[3];                  /* max |vp| over physical slots after the kick */
    double dt_pm2, dt_pm3, dt_pp, dt_vmax;  /* FP64; +inf for a zero force / velocity or a skipped part */
    double pt_pm3, pt_pp;           /* sum(ptotal) of PM3 and of PP */
    uint64_t pair_counts[2];        /* PP P_cand, P_acc */
    int napm[7];                    /* subtiles per level (kick.f90:129) */
} cube2_kick_result_t;

enum {
    CUBE2_KICK_PM3 = 1,          /* run PM3 */
    CUBE2_KICK_PP = 2,           /* run PP (applies PM3's pending Delta-v) */
    CUBE2_KICK_KEEP_LEVELS = 4,  /* use the caller's ires/isort instead of selecting them */
    CUBE2_KICK_PM2 = 8,          /* run PM2 first; d_green[5] is its table (Gk2) */
    CUBE2_KICK_LEARNED = 16      /* the learned stand-in replaces PP in sparse level-2 cells
                                    (cube2_learned_set); PP modes other than EXACT */
};

/* flags: CUBE2_KICK_*.  ires/isort [ns3] are written (level selection of
 * kick.f90:107-131 from the device rhoc) unless KEEP_LEVELS.  PM3 without
 * PP applies its Delta-v before returning.  ptotal_pm3, ptotal_pp,
 * ttotal_pp [ns3] may be NULL. */
int cube2_kick_dev(int flags, const int* ixyz3, const float* apm3, const int* ratio_sf,
                   const int* d_rhoc, const int64_t* d_idx_b_r, const int16_t* d_xp, float* d_vp,
                   int64_t n_used, const float* const* d_green /* [5]: levels 2..6 */,
                   float mass, float a_mid, float dt, int* ires, int* isort, double* ptotal_pm3,
                   double* ptotal_pp, double* ttotal_pp, cube2_kick_result_t* out);

/* The same kick with the caller's levels (flags CUBE2_KICK_PM3|CUBE2_KICK_PP
 * or CUBE2_KICK_PP), recorded into a HIP graph on the first call and
 * replayed while the levels, job order, arrays, sizes, parameters and PP
 * mode stay the same (mass, a_mid and dt may change between calls).  Any
 * other library call that grows buffers or uploads tables makes the next
 * call record again.  *recorded (may be NULL) = 1 when this call recorded.
 * Fails with CUBE2_ERR_STATE while a PM3 Delta-v from cube2_pm3_force_dev
 * is pending. */
int cube2_kick_graph_dev(int flags, const int* ixyz3, const float* apm3, const int* ratio_sf,
                         const int* d_rhoc, const int64_t* d_idx_b_r, const int16_t* d_xp,
                         float* d_vp, int64_t n_used, const float* const* d_green, float mass,
                         float a_mid, float dt, const int* ires, const int* isort,
                         double* ptotal_pm3, double* ptotal_pp, double* ttotal_pp,
                         cube2_kick_result_t* out, int* recorded);

/* ---------------------------------------------------------------------
 * Inspection (tests and benchmarks)
 * ------------------------------------------------------------------- */
/* Cell list of parent tile `it` from the last PP call: sizes first
 * (pointers may be NULL), then copies.  keys_sorted[N] (key in the tile's
 * S^3 grid), slot[N] (xp index), cell_start[S^3+1] (tile-relative). */
int cube2_pp_debug_celllist(int it, int* rcp, int* S, int* N, uint32_t* keys_sorted,
                            uint32_t* slot, int32_t* cell_start);

/* Per-parent-tile pair counts.  After cube2_pp_count_tiles(1), each PP call
 * also counts P_cand and P_acc per parent tile (the tile totals of
 * pp_force_kernel.c's per-tile runs, :281-325); cube2_pp_debug_tile_pairs
 * copies them from the last PP call: counts[ntile][2] = {P_cand, P_acc}.
 * Off by default (it adds atomics to the pair kernel). */
int cube2_pp_count_tiles(int on);
int cube2_pp_debug_tile_pairs(int ntile, uint64_t* counts);

/* The pair kernel's pair force for given pairs (host arrays):
 * pairs[n][4] = {dx, dy, dz, range}, forces[n][3] = the force of that pair
 * in mode CUBE2_PP_REF / FAST / PRECISE / PLUMMER with softening app, evaluated by the
 * same device function and acceptance test (0 < r^2 < range^2, else 0) as
 * the pair kernel.  In REF mode it equals (F_ra(r,app) - F_ra(r,range))*d/r
 * of pp_force_kernel.c:308-318 bit for bit. */
int cube2_pp_debug_pair_forces(int mode, const float* pairs, int n, float app, float* forces);

/* One PM3 subtile job's stage outputs from a fresh computation:
 * rho[n^3] (density incl. mass), phi[n^3] (potential / n^3),
 * force[3][(nfp+2)^3]; host pointers, any may be NULL. */
int cube2_pm3_debug_job(int it, const int* ires, const int* ixyz3, const int* ratio_sf,
                        const int* d_rhoc, const int64_t* d_idx_b_r, const int16_t* d_xp,
                        int64_t n_used, const float* d_green, float mass, float* rho, float* phi,
                        float* force, float* f2max);

/* Batched 3-D FFT of nb real cubes of size n (in place, padded x rows of
 * n+2 floats, FFTW/cuFFT in-place layout): dir=-1 R2C, +1 C2R, unnormalised. */
int cube2_fft3d_dev(float* d_data, int n, int nb, int dir);

/* Checks the FP64-reciprocal quotients used by CUBE2_PP_REF (a/b, 12/b,
 * 1/(r*r), 12/ep^2) against IEEE float division on n seeded random operand
 * sets; seed 0 instead takes every float with exponent in [-40, 40]
 * (n = 81 * 2^23).  *mismatches = count. */
int cube2_selftest_div(uint64_t seed, int n, unsigned long long* mismatches);

int cube2_profile_enable(int on);
int cube2_profile_reset(void);
/* Accumulated GPU time (ms), launch count and algorithmic bytes / flops of
 * a stage; returns CUBE2_ERR_ARG for an unknown stage. */
int cube2_profile_get(const char* stage, double* ms, int64_t* launches, double* bytes,
                      double* flops);
/* Names of the recorded stages, comma separated, into buf. */
int cube2_profile_stages(char* buf, int len);

int cube2_synchronize(void);
const char* cube2_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
