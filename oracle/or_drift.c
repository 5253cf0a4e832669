/*
 * Drift of one periodic image: CPU restatement (TEST INFRASTRUCTURE ONLY;
 * see oracle.h).
 *
 * drift.f90:1-158 per tile, single image: every slot of the tile's extended
 * region (cells 1-ncb..nt+ncb) moves by dt_mid*vp; the particles that land
 * in the tile's physical cells 1..nt form the tile's new part of the image,
 * ordered by cell (z, y, x) and, within a cell, by arrival in the layered
 * loop of drift.f90:36-62 (layer, then source j, i, slot).  rhoc keeps the
 * physical counts only (halo cells zero, drift.f90:143-144); idx_b_r is
 * spine_image's running row total (variables.f90:146-216).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static const int64_t kIshiftD = -((int64_t)1 << 15); /* parameters.f90:87 */
static const double kRshiftD = 0.5 + 32768.0;        /* parameters.f90:88 */
static const double kXresD = 1.0 / 65536.0;          /* parameters.f90:118 */

static inline double frac16d(int16_t v) {
    return ((double)(int16_t)((int64_t)v + kIshiftD) + kRshiftD) * kXresD;
}

int or_drift(const or_geom* g, const int* rhoc, const int64_t* idx_b_r, const int16_t* xp,
             const float* vp, float dt_mid, float vz_max, int* rhoc_out, int64_t* idx_out,
             int16_t* xp_out, float* vp_out, int64_t cap, int64_t* nplocal) {
    const int nc = g->ng / g->ratio_cs, nt = nc / g->nnt, ncb = g->ngb / g->ratio_cs;
    const int nte = nt + 2 * ncb, ne2 = nt + 4 * ncb, nnt = g->nnt;
    const int ntile = nnt * nnt * nnt;
    const size_t te3 = (size_t)nte * nte * nte, e23 = (size_t)ne2 * ne2 * ne2;
    /* nlayer = 2*ceiling(dt_mid*vz_max/ratio_cs)+2 (drift.f90:26), FP32 */
    volatile float q = dt_mid * vz_max;
    q = q / (float)g->ratio_cs;
    const int nlayer = 2 * (int)ceilf(q) + 2;
    int* rhoce = (int*)malloc(sizeof(int) * e23);
    int* rholocal = (int*)malloc(sizeof(int) * e23);
    int64_t* idx_ex_r = (int64_t*)malloc(sizeof(int64_t) * (size_t)ne2 * ne2);
    int64_t np_prev = 0;
    int rc = 0;
    memset(rhoc_out, 0, sizeof(int) * te3 * ntile);
    for (int t = 0; t < ntile && !rc; t++) {
        const int* rct = rhoc + te3 * t;
        const int64_t* ibt = idx_b_r + (size_t)nte * nte * t;
        /* cumulative ends of the tile's rows: nzero of cell i = row end - sum(i:) */
        memset(rhoce, 0, sizeof(int) * e23);
        memset(rholocal, 0, sizeof(int) * e23);
        int64_t tile_total = 0;
        for (size_t c = 0; c < te3; c++) tile_total += rct[c];
        int16_t* xp_new = (int16_t*)malloc(sizeof(int16_t) * 3 * (size_t)(tile_total + 1));
        float* vp_new = (float*)malloc(sizeof(float) * 3 * (size_t)(tile_total + 1));
        for (int pass = 0; pass < 2 && !rc; pass++) {
            for (int il = 0; il < nlayer && !rc; il++)
                for (int k = il; k < nte; k += nlayer) /* k = 1-ncb+ilayer.. as nte index */
                    for (int j = 0; j < nte; j++)
                        for (int i = 0; i < nte; i++) {
                            const int np = rct[((size_t)k * nte + j) * nte + i];
                            int64_t suf = 0;
                            for (int a = i; a < nte; a++) suf += rct[((size_t)k * nte + j) * nte + a];
                            const int64_t nzero = ibt[(size_t)k * nte + j] - suf;
                            for (int l = 0; l < np; l++) {
                                const int64_t ip = nzero + l;
                                const int src[3] = {i, j, k};
                                int ig[3];
                                for (int d = 0; d < 3; d++) {
                                    /* xq = (cell-1) + frac, cell = nte index + 1 - ncb */
                                    const double xq = (double)(src[d] - ncb) + frac16d(xp[3 * ip + d]);
                                    volatile float dv = dt_mid * vp[3 * ip + d];
                                    dv = dv / (float)g->ratio_cs;
                                    const double deltax = (double)dv;
                                    /* 1-based cell ceiling(xq+deltax) -> e2 index */
                                    ig[d] = (int)ceil(xq + deltax) + 2 * ncb - 1;
                                    if (ig[d] < 0 || ig[d] >= ne2) rc = -2;
                                }
                                if (rc) break;
                                const size_t e = ((size_t)ig[2] * ne2 + ig[1]) * ne2 + ig[0];
                                if (pass == 0) {
                                    rhoce[e]++;
                                } else {
                                    rholocal[e]++;
                                    int64_t s2 = 0;
                                    for (int a = ig[0]; a < ne2; a++)
                                        s2 += rhoce[((size_t)ig[2] * ne2 + ig[1]) * ne2 + a];
                                    const int64_t idx =
                                        idx_ex_r[(size_t)ig[2] * ne2 + ig[1]] - s2 + rholocal[e];
                                    for (int d = 0; d < 3; d++) {
                                        /* nint(dt_mid*vp/(x_resolution*ratio_cs)), int16 wrap */
                                        volatile float f = dt_mid * vp[3 * ip + d];
                                        const double sh = (double)f / (kXresD * g->ratio_cs);
                                        const long m = lround(sh);
                                        xp_new[3 * (idx - 1) + d] =
                                            (int16_t)(uint16_t)((int32_t)xp[3 * ip + d] + (int32_t)m);
                                        vp_new[3 * (idx - 1) + d] = vp[3 * ip + d];
                                    }
                                }
                            }
                        }
            if (pass == 0) { /* spine_tile, variables.f90:101-144 */
                int64_t nsum = 0;
                for (int z = 0; z < ne2; z++)
                    for (int y = 0; y < ne2; y++) {
                        for (int x = 0; x < ne2; x++) nsum += rhoce[((size_t)z * ne2 + y) * ne2 + x];
                        idx_ex_r[(size_t)z * ne2 + y] = nsum;
                    }
            }
        }
        if (!rc) {
            /* copy the physical rows (drift.f90:129-140) */
            for (int k = 0; k < nt; k++)
                for (int j = 0; j < nt; j++) {
                    const int z = k + 2 * ncb, y = j + 2 * ncb;
                    int64_t right = 0, phys = 0;
                    for (int x = nt + 2 * ncb; x < ne2; x++) right += rhoce[((size_t)z * ne2 + y) * ne2 + x];
                    for (int x = 2 * ncb; x < nt + 2 * ncb; x++) phys += rhoce[((size_t)z * ne2 + y) * ne2 + x];
                    const int64_t ppe_r = idx_ex_r[(size_t)z * ne2 + y] - right;
                    const int64_t ppe_l = ppe_r - phys + 1;
                    if (np_prev + phys > cap) {
                        rc = -3;
                        break;
                    }
                    memcpy(xp_out + 3 * np_prev, xp_new + 3 * (ppe_l - 1), sizeof(int16_t) * 3 * phys);
                    memcpy(vp_out + 3 * np_prev, vp_new + 3 * (ppe_l - 1), sizeof(float) * 3 * phys);
                    np_prev += phys;
                    for (int i = 0; i < nt; i++)
                        rhoc_out[te3 * t + ((size_t)(k + ncb) * nte + j + ncb) * nte + i + ncb] =
                            rhoce[((size_t)z * ne2 + y) * ne2 + i + 2 * ncb];
                }
        }
        free(xp_new);
        free(vp_new);
    }
    /* spine_image: running total of every extended row, tiles z-major */
    int64_t run = 0;
    for (int t = 0; t < ntile; t++)
        for (int k = 0; k < nte; k++)
            for (int j = 0; j < nte; j++) {
                for (int i = 0; i < nte; i++) run += rhoc_out[te3 * t + ((size_t)k * nte + j) * nte + i];
                idx_out[(size_t)nte * nte * t + (size_t)k * nte + j] = run;
            }
    *nplocal = np_prev;
    free(rhoce);
    free(rholocal);
    free(idx_ex_r);
    return rc;
}
