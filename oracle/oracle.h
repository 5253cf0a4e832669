/*
 * CPU oracle for the CUBE2 short-range force path (PM3 + PP).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never
 * links, calls or falls back to it.
 *
 * It restates, in plain C with OpenMP, the reference algorithms:
 *   - F_ra                    basic_functions.f08:95-109, pp_force_kernel.c:45-74
 *   - PP gather/keys/sort/sum pp_force_kernel.c:134-391 (arithmetic) with the
 *                             per-particle range / active semantics of
 *                             pp_force_kernel.cu:625-958
 *   - level select + isort    kick.f90:107-131, 570-614
 *   - PM3 pm_thread chain     kick.f90:403-568 (deposit, FFT, Green multiply,
 *                             potential_to_force, interpolation)
 *   - Green_3D                Green/Green.f90:1-80 (FP64)
 *   - drift                   drift.f90:1-158, variables.f90:101-216
 *
 * Parity pins (see tests/): golden vectors from the compiled reference
 * pp_force_kernel.c (oracle/_ref), numpy's pocketfft for the FFT, closed
 * forms and invariants for TSC, gradient and Green's function.
 */
#ifndef CUBE2_ORACLE_H
#define CUBE2_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Runtime form of the reference's compile-time geometry
 * (parameters.f90:19-28,42,57; cuda_parameters_generated.h). */
typedef struct {
    int ng;        /* standard cells per image per dim */
    int nnt;       /* tiles per image per dim */
    int nns;       /* subtiles per tile per dim */
    int ratio_cs;  /* standard cells per coarse cell */
    int ngb;       /* tile buffer in standard cells */
    int nteam;     /* OpenMP tile threads (outer) */
    int nnest;     /* OpenMP cell threads (inner) */
    float app;     /* PP softening diameter, standard cells */
} or_geom;

/* ---- PP (Q1-Q7) ---- */
float or_F_ra(float r, float apm);
/* Pair forces of the reference's pair loop for difference vectors d[n][3]. */
void or_pair_forces(const float* d, const float* range, float app, int n, float* f);

/* Stable counting sort of u32 keys (values permuted alongside). */
void or_stable_sort_u32_pair(uint32_t* keys, uint32_t* vals, int n, uint32_t key_bound);

/* Cell list of one parent tile (it, x-major numbering as in
 * pp_force_kernel.cu:1022).  Outputs (caller-allocated, sizes from
 * or_pp_tile_sizes): keys_sorted[N], perm[N] (sorted -> gather order),
 * slot[N] (xp index of each sorted particle), cell_start[S^3+1]. */
void or_pp_tile_sizes(const or_geom* g, const int* ires, const int* ixyz3, const int* ratio_sf,
                      const int* rhoc, int it, int* rcp, int* S, int* N);
void or_pp_tile_celllist(const or_geom* g, const int* ires, const int* ixyz3, const float* apm3,
                         const int* ratio_sf, const int* rhoc, const int64_t* idx_b_r,
                         const int16_t* xp, int it, uint32_t* keys_sorted, uint32_t* perm,
                         uint32_t* slot, int32_t* cell_start, float* xf_sorted /*[3][N]*/,
                         float* range_sorted);

/* Full PP kick over parent tiles [tile_first, tile_first+tile_count).
 * vp is updated in place: vp = (float)((double)vp + dv64 + pp_term), the
 * double promotion of pp_force_kernel.c:355-357.  dv64 may be NULL.
 * counts (may be NULL): [0]=P_cand (examined ordered pairs, i active,
 * j != i), [1]=P_acc (accepted ordered pairs). */
int or_pp_force(const or_geom* g, const int* ires, const int* ixyz3, const float* apm3,
                const int* ratio_sf, const int* rhoc, const int64_t* idx_b_r, const int16_t* xp,
                float* vp, const double* dv64, float mass, float a_mid, float dt, float* f2max,
                float* vmax, double* ptotal, double* ttotal, uint64_t* counts, int tile_first,
                int tile_count);

/* ---- level selection (P1) ---- */
void or_select_levels(const or_geom* g, const int* rhoc, const int* ixyz3, int* ires, int* isort,
                      int* napm /*[7]*/, float* npc_max);
void or_indexed_sort(int n, const float* arrin, int* indx);

/* ---- PM3 (P4-P12) ---- */
void or_fft3d_r2c(const float* in /* n^3 real, x fastest */, double* out /* (n/2+1)*n*n*2 */, int n);
void or_fft3d_c2r(const double* in /* (n/2+1)*n*n*2 */, double* out /* n^3 */, int n);
void or_fft1d(double* re_im /* 2n interleaved */, int n, int sign);
void or_green_3d(float* Gk /* (n/2+1)*n*n */, int n, float apm_soft, float apm_range,
                 float gridsize);

/* potential_to_force (kick.f90:545-568) on an n^3 potential phi (x fastest):
 * F[3][nfp+2]^3 at phi storage offset nfb-1, f2max over the interior. */
void or_potential_to_force(const double* phi, int n, int nfb, int nfp, double* F, double* f2max);
/* One subtile job, stage outputs for stage-wise checks (any may be NULL):
 * rho[n^3] (float, fine-mesh density incl. mass), phi[n^3] (double,
 * unnormalised C2R(G*R2C(rho)) / n^3), force[3][(nfp+2)^3] (double). */
int or_pm3_job(const or_geom* g, int it, const int* ires, const int* ixyz3, const int* ratio_sf,
               const int* rhoc, const int64_t* idx_b_r, const int16_t* xp, const float* green,
               float mass, float* rho, double* phi, double* force, double* f2max_job);

/* Full PM3 over all subtiles: accumulates Delta-v into dv64[3*N_used],
 * f2max (max over subtile interiors of |F|^2), ptotal[ns3] (counts of
 * particles in each subtile's physical cells, pm3_fft_cuda.cu:436-437).
 * greens[L] for L=2..6 (others NULL). */
int or_pm3_force(const or_geom* g, const int* isort, const int* ires, const int* ixyz3,
                 const int* ratio_sf, const int* rhoc, const int64_t* idx_b_r, const int16_t* xp,
                 const float* const* greens, float mass, float amid, float dt, double* dv64,
                 float* f2max, double* ptotal);

/* PM2 (kick.f90:76-105) over tiles tile0 .. tile0+ntiles-1 (ntiles <= 0:
 * all): accumulates Delta-v into dv64[3*N_used], f2max (max over tile
 * interiors of |F|^2), ptotal[nnt^3] (particles of each tile's physical
 * cells).  green2: Gk2 [n][n][n/2+1], n = nt ratio_cs + 2 ngb. */
int or_pm2_force(const or_geom* g, const int* rhoc, const int64_t* idx_b_r, const int16_t* xp,
                 const float* green2, float mass, float amid, float dt, double* dv64,
                 float* f2max, double* ptotal, int tile0, int ntiles);

/* Drift of one image (drift.f90:1-158, spine_image): from the image with
 * halo copies (rhoc/idx_b_r/xp/vp as the kick takes them) to the new
 * physical-only image: rhoc_out [nnt^3][nte^3] (halo cells 0), idx_out
 * [nnt^3][nte^2] (running row totals), xp_out/vp_out [*nplocal][3] (cap
 * entries at most).  Returns 0, -2 if a particle moves beyond 2 ncb, -3 if
 * cap is too small. */
int or_drift(const or_geom* g, const int* rhoc, const int64_t* idx_b_r, const int16_t* xp,
             const float* vp, float dt_mid, float vz_max, int* rhoc_out, int64_t* idx_out,
             int16_t* xp_out, float* vp_out, int64_t cap, int64_t* nplocal);

/* Apply a pending Delta-v to vp for slots in physical cells, return vmax. */
void or_velocity_apply(const or_geom* g, const int* rhoc, const int64_t* idx_b_r, float* vp,
                       const double* dv64, float* vmax);

/* FP64 direct sum of the PP kernel over all pairs within range_i, with the
 * periodic minimum image, for a list of target particles.  pos in
 * standard cells [0, box), range per target. */
/* law 0: the S2 F_ra(r, app) of basic_functions.f08:95-109; law 1: Plummer
 * r / (r^2 + app^2)^(3/2) (an addition with no counterpart in the
 * reference); minus F_ra(r, range_i) in both. */
void or_direct_pp_f64_cells(const double* pos, int n, double box, const int* targets, int nt,
                            const double* range_t, double app, int law, double* acc);
void or_direct_pp_f64(const double* pos /*[n][3]*/, int n, double box, const int* targets,
                      int nt, const double* range_t, double app, int law, double* acc /*[nt][3]*/);

#ifdef __cplusplus
}
#endif
#endif
