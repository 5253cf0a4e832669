/*
 * PP oracle (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 * Restates the reference PP path for one image:
 *   gather + keys      pp_force_kernel.c:199-240 / pp_force_kernel.cu:736-815
 *   stable sort        pp_force_kernel.c:76-132 (radix, stable) == counting sort
 *   cell ranges        pp_force_kernel.c:257-274 / pp_force_kernel.cu:878-911
 *   pair sum           pp_force_kernel.c:281-325 (arithmetic order), with the
 *                      per-particle range / cell radius / active flag of
 *                      pp_force_kernel.cu:192-250
 *   velocity update    pp_force_kernel.c:332-364 (double promotion of /pi)
 *   per-tile rcp       pp_force_kernel.cu:641-661
 *   level select       kick.f90:107-131, heapsort kick.f90:570-614
 * Compiled with -ffp-contract=off so every float operation rounds as written.
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static const double kPi = 3.14159265358979323846;
static const int64_t kIshift = -((int64_t)1 << 15);  /* parameters.f90:87 */
static const double kRshift = 0.5 + 32768.0;         /* parameters.f90:88 */
static const double kXres = 1.0 / 65536.0;           /* parameters.f90:118 */

typedef struct {
    int nc, nt, ntt, ncb, nte, nsub, ns3;
} derived;

static derived derive(const or_geom* g) {
    derived d;
    d.nc = g->ng / g->ratio_cs;
    d.nt = d.nc / g->nnt;
    d.ntt = d.nt / g->nns;
    d.ncb = g->ngb / g->ratio_cs;
    d.nte = d.nt + 2 * d.ncb;
    d.nsub = g->nnt * g->nns;
    d.ns3 = d.nsub * d.nsub * d.nsub;
    return d;
}

static double now_sec(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* basic_functions.f08:95-109: the S2-sphere force law; the expression order
 * is kept term for term so that FP32 results are bit-identical. */
float or_F_ra(float r, float apm) {
    float ep = 2.0f * r / apm;
    if (apm == 0.0f || ep > 2.0f) return 1.0f / (r * r);
    if (ep > 1.0f) {
        float e2 = ep * ep, e3 = e2 * ep, e4 = e3 * ep, e5 = e4 * ep, e6 = e5 * ep;
        return (1.0f / 35.0f / (apm * apm)) *
               (12.0f / e2 - 224.0f + 896.0f * ep - 840.0f * e2 + 224.0f * e3 + 70.0f * e4 -
                48.0f * e5 + 7.0f * e6);
    }
    if (ep > 0.0f) {
        float e3 = ep * ep * ep, e4 = e3 * ep, e5 = e4 * ep, e6 = e5 * ep;
        return (1.0f / 35.0f / (apm * apm)) *
               (224.0f * ep - 224.0f * e3 + 70.0f * e4 + 48.0f * e5 - 21.0f * e6);
    }
    return 0.0f;
}

/* The pair force of pp_force_kernel.c:308-318 for given difference vectors:
 * f = (F_ra(r,app) - F_ra(r,range)) * d / r for 0 < r^2 < range^2, else 0. */
void or_pair_forces(const float* d, const float* range, float app, int n, float* f) {
    for (int q = 0; q < n; q++) {
        float ddx = d[3 * q], ddy = d[3 * q + 1], ddz = d[3 * q + 2];
        float r2 = ddx * ddx + ddy * ddy + ddz * ddz;
        float pp2 = range[q] * range[q];
        f[3 * q] = f[3 * q + 1] = f[3 * q + 2] = 0.0f;
        if (r2 > 0.0f && r2 < pp2) {
            float rm = sqrtf(r2);
            float fpp = or_F_ra(rm, app) - or_F_ra(rm, range[q]);
            f[3 * q] = fpp * ddx / rm;
            f[3 * q + 1] = fpp * ddy / rm;
            f[3 * q + 2] = fpp * ddz / rm;
        }
    }
}

void or_stable_sort_u32_pair(uint32_t* keys, uint32_t* vals, int n, uint32_t key_bound) {
    uint32_t* cnt = (uint32_t*)calloc((size_t)key_bound + 1, sizeof(uint32_t));
    uint32_t* k2 = (uint32_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(uint32_t));
    uint32_t* v2 = (uint32_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(uint32_t));
    for (int i = 0; i < n; i++) cnt[keys[i] + 1]++;
    for (uint32_t b = 0; b < key_bound; b++) cnt[b + 1] += cnt[b];
    for (int i = 0; i < n; i++) {
        uint32_t p = cnt[keys[i]]++;
        k2[p] = keys[i];
        v2[p] = vals[i];
    }
    memcpy(keys, k2, (size_t)n * sizeof(uint32_t));
    memcpy(vals, v2, (size_t)n * sizeof(uint32_t));
    free(cnt);
    free(k2);
    free(v2);
}

/* ixyz3 -> per-location ires and owner (pp_force_kernel.cu:972-991). */
static int location_maps(const or_geom* g, const derived* d, const int* ires, const int* ixyz3,
                         int* ires_loc, int* owner) {
    for (int it = 0; it < d->ns3; it++) {
        const int* x = ixyz3 + 6 * it;
        int sx = x[0] - 1, sy = x[1] - 1, sz = x[2] - 1, tx = x[3] - 1, ty = x[4] - 1,
            tz = x[5] - 1;
        if (sx < 0 || sx >= g->nns || sy < 0 || sy >= g->nns || sz < 0 || sz >= g->nns ||
            tx < 0 || tx >= g->nnt || ty < 0 || ty >= g->nnt || tz < 0 || tz >= g->nnt ||
            ires[it] < 2 || ires[it] > 6)
            return -1;
        int loc = (((tz * g->nnt + ty) * g->nnt + tx) * g->nns + sz) * g->nns * g->nns +
                  sy * g->nns + sx;
        ires_loc[loc] = ires[it];
        owner[loc] = it;
    }
    return 0;
}

static inline float decode(int c_rel, int16_t v) {
    return (float)((double)c_rel +
                   ((double)(int16_t)((int64_t)v + kIshift) + kRshift) * kXres);
}

static int tile_rcp(const or_geom* g, const int* ires_loc, const int* ratio_sf, int it) {
    int rcp = ratio_sf[1];
    int base = it * g->nns * g->nns * g->nns;
    for (int s = 0; s < g->nns * g->nns * g->nns; s++) {
        int r = ratio_sf[ires_loc[base + s] - 1];
        if (r > rcp) rcp = r;
    }
    return rcp;
}

typedef struct {
    int rcp, S, N;
    uint32_t* keys;   /* sorted keys */
    uint32_t* perm;   /* sorted -> gather index */
    uint32_t* slot;   /* sorted -> xp index */
    int32_t* start;   /* S^3 + 1 */
    float* xf;        /* [3][N] sorted */
    float* range;     /* sorted; 0 for halo */
} tile_list;

static void free_tile(tile_list* t) {
    free(t->keys);
    free(t->perm);
    free(t->slot);
    free(t->start);
    free(t->xf);
    free(t->range);
}

static int64_t nzero_of(const derived* d, const int* rc, const int64_t* ib, int k, int j, int i) {
    const int* row = rc + ((size_t)k * d->nte + j) * d->nte;
    int64_t s = 0;
    for (int q = i; q < d->nte; q++) s += row[q];
    return ib[(size_t)k * d->nte + j] - s;
}

/* Gather + keys + stable sort for parent tile it (x-major numbering). */
static void build_tile(const or_geom* g, const derived* d, int it, const int* ires_loc,
                       const float* apm3, const int* ratio_sf, const int* rhoc,
                       const int64_t* idx_b_r, const int16_t* xp, tile_list* tl) {
    const int nte = d->nte, ncb = d->ncb, nt = d->nt;
    const int* rc = rhoc + (size_t)it * nte * nte * nte;
    const int64_t* ib = idx_b_r + (size_t)it * nte * nte;
    const int tx = it % g->nnt, ty = (it / g->nnt) % g->nnt, tz = it / (g->nnt * g->nnt);
    int rcp = tile_rcp(g, ires_loc, ratio_sf, it);
    int S = nt * rcp + 2 * rcp;
    int N = 0;
    for (int k = ncb - 1; k <= ncb + nt; k++)
        for (int j = ncb - 1; j <= ncb + nt; j++)
            for (int i = ncb - 1; i <= ncb + nt; i++) N += rc[((size_t)k * nte + j) * nte + i];
    uint32_t* keys = (uint32_t*)malloc((size_t)(N + 1) * sizeof(uint32_t));
    uint32_t* vals = (uint32_t*)malloc((size_t)(N + 1) * sizeof(uint32_t));
    uint32_t* slot_g = (uint32_t*)malloc((size_t)(N + 1) * sizeof(uint32_t));
    float* xf_g = (float*)malloc((size_t)(3 * N + 1) * sizeof(float));
    float* rg_g = (float*)malloc((size_t)(N + 1) * sizeof(float));
    int ip1 = 0;
    for (int k = ncb - 1; k <= ncb + nt; k++)
        for (int j = ncb - 1; j <= ncb + nt; j++)
            for (int i = ncb - 1; i <= ncb + nt; i++) {
                int np = rc[((size_t)k * nte + j) * nte + i];
                int64_t nz = nzero_of(d, rc, ib, k, j, i);
                int phys = i >= ncb && i < ncb + nt && j >= ncb && j < ncb + nt && k >= ncb &&
                           k < ncb + nt;
                float rng = 0.0f;
                if (phys) {
                    int sx = (i - ncb) / d->ntt, sy = (j - ncb) / d->ntt, sz = (k - ncb) / d->ntt;
                    int loc = (((tz * g->nnt + ty) * g->nnt + tx) * g->nns + sz) * g->nns * g->nns +
                              sy * g->nns + sx;
                    rng = apm3[ires_loc[loc] - 1];
                }
                for (int l = 0; l < np; l++) {
                    int64_t ip = nz + l;
                    float x0 = decode(i - ncb, xp[3 * ip + 0]);
                    float x1 = decode(j - ncb, xp[3 * ip + 1]);
                    float x2 = decode(k - ncb, xp[3 * ip + 2]);
                    xf_g[3 * ip1 + 0] = (float)g->ratio_cs * x0;
                    xf_g[3 * ip1 + 1] = (float)g->ratio_cs * x1;
                    xf_g[3 * ip1 + 2] = (float)g->ratio_cs * x2;
                    int i0 = (int)floorf((float)rcp * x0) + rcp;
                    int i1 = (int)floorf((float)rcp * x1) + rcp;
                    int i2 = (int)floorf((float)rcp * x2) + rcp;
                    keys[ip1] = (uint32_t)i2 * (uint32_t)S * (uint32_t)S +
                                (uint32_t)i1 * (uint32_t)S + (uint32_t)i0;
                    vals[ip1] = (uint32_t)ip1;
                    slot_g[ip1] = (uint32_t)ip;
                    rg_g[ip1] = rng;
                    ip1++;
                }
            }
    uint32_t bound = (uint32_t)S * (uint32_t)S * (uint32_t)S;
    or_stable_sort_u32_pair(keys, vals, N, bound);
    tl->rcp = rcp;
    tl->S = S;
    tl->N = N;
    tl->keys = keys;
    tl->perm = vals;
    tl->slot = (uint32_t*)malloc((size_t)(N + 1) * sizeof(uint32_t));
    tl->xf = (float*)malloc((size_t)(3 * N + 1) * sizeof(float));
    tl->range = (float*)malloc((size_t)(N + 1) * sizeof(float));
    for (int a = 0; a < N; a++) {
        uint32_t o = vals[a];
        tl->slot[a] = slot_g[o];
        tl->xf[a] = xf_g[3 * o + 0];
        tl->xf[N + a] = xf_g[3 * o + 1];
        tl->xf[2 * N + a] = xf_g[3 * o + 2];
        tl->range[a] = rg_g[o];
    }
    tl->start = (int32_t*)malloc(((size_t)bound + 1) * sizeof(int32_t));
    {
        size_t c = 0;
        for (int a = 0; a <= N; a++) {
            uint32_t kk = a < N ? keys[a] : bound;
            while (c <= kk && c <= bound) tl->start[c++] = a;
        }
    }
    free(slot_g);
    free(xf_g);
    free(rg_g);
}

void or_pp_tile_sizes(const or_geom* g, const int* ires, const int* ixyz3, const int* ratio_sf,
                      const int* rhoc, int it, int* rcp, int* S, int* N) {
    derived d = derive(g);
    int* il = (int*)malloc(sizeof(int) * d.ns3);
    int* ow = (int*)malloc(sizeof(int) * d.ns3);
    location_maps(g, &d, ires, ixyz3, il, ow);
    *rcp = tile_rcp(g, il, ratio_sf, it);
    *S = d.nt * *rcp + 2 * *rcp;
    const int* rc = rhoc + (size_t)it * d.nte * d.nte * d.nte;
    int n = 0;
    for (int k = d.ncb - 1; k <= d.ncb + d.nt; k++)
        for (int j = d.ncb - 1; j <= d.ncb + d.nt; j++)
            for (int i = d.ncb - 1; i <= d.ncb + d.nt; i++)
                n += rc[((size_t)k * d.nte + j) * d.nte + i];
    *N = n;
    free(il);
    free(ow);
}

void or_pp_tile_celllist(const or_geom* g, const int* ires, const int* ixyz3, const float* apm3,
                         const int* ratio_sf, const int* rhoc, const int64_t* idx_b_r,
                         const int16_t* xp, int it, uint32_t* keys_sorted, uint32_t* perm,
                         uint32_t* slot, int32_t* cell_start, float* xf_sorted,
                         float* range_sorted) {
    derived d = derive(g);
    int* il = (int*)malloc(sizeof(int) * d.ns3);
    int* ow = (int*)malloc(sizeof(int) * d.ns3);
    location_maps(g, &d, ires, ixyz3, il, ow);
    tile_list t;
    build_tile(g, &d, it, il, apm3, ratio_sf, rhoc, idx_b_r, xp, &t);
    size_t nc = (size_t)t.S * t.S * t.S;
    if (keys_sorted) memcpy(keys_sorted, t.keys, (size_t)t.N * 4);
    if (perm) memcpy(perm, t.perm, (size_t)t.N * 4);
    if (slot) memcpy(slot, t.slot, (size_t)t.N * 4);
    if (cell_start) memcpy(cell_start, t.start, (nc + 1) * 4);
    if (xf_sorted) memcpy(xf_sorted, t.xf, (size_t)3 * t.N * 4);
    if (range_sorted) memcpy(range_sorted, t.range, (size_t)t.N * 4);
    free_tile(&t);
    free(il);
    free(ow);
}

static void pp_tile(const or_geom* g, const derived* d, int it, const int* ires_loc,
                    const int* owner, const float* apm3, const int* ratio_sf, const int* rhoc,
                    const int64_t* idx_b_r, const int16_t* xp, float* vp, const double* dv64,
                    float mass, float a_mid, float dt, float* f2max_t, float* vmax_t,
                    double* ptotal, double* ttotal, uint64_t* cnt) {
    double t0 = now_sec();
    tile_list t;
    build_tile(g, d, it, ires_loc, apm3, ratio_sf, rhoc, idx_b_r, xp, &t);
    const int N = t.N, S = t.S, rcp = t.rcp;
    const float app = g->app;
    const float* X = t.xf;
    const float* Y = t.xf + N;
    const float* Z = t.xf + 2 * N;
    float* af = (float*)calloc((size_t)3 * (N + 1), sizeof(float));
    uint64_t pc = 0, pa = 0;
#pragma omp parallel for num_threads(g->nnest) schedule(dynamic, 64) reduction(+ : pc, pa)
    for (int a = 0; a < N; a++) {
        float range = t.range[a];
        if (range == 0.0f) continue; /* halo particles are sources only */
        uint32_t key = t.keys[a];
        int x = (int)(key % (uint32_t)S), y = (int)((key / (uint32_t)S) % (uint32_t)S),
            z = (int)(key / ((uint32_t)S * (uint32_t)S));
        int R = (int)ceilf(range * (float)rcp / (float)g->ratio_cs);
        float xo = X[a], yo = Y[a], zo = Z[a];
        float pp2 = range * range;
        float ax = 0.f, ay = 0.f, az = 0.f;
        for (int dz = -R; dz <= R; dz++) {
            int zz = z + dz;
            if (zz < 0 || zz >= S) continue;
            for (int dy = -R; dy <= R; dy++) {
                int yy = y + dy;
                if (yy < 0 || yy >= S) continue;
                for (int dx = -R; dx <= R; dx++) {
                    int xx = x + dx;
                    if (xx < 0 || xx >= S) continue;
                    size_t c = ((size_t)zz * S + yy) * S + xx;
                    int j0 = t.start[c], j1 = t.start[c + 1];
                    pc += (uint64_t)(j1 - j0);
                    for (int j = j0; j < j1; j++) {
                        float ddx = X[j] - xo, ddy = Y[j] - yo, ddz = Z[j] - zo;
                        float r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                        if (r2 > 0.0f && r2 < pp2) {
                            float rm = sqrtf(r2);
                            float fpp = or_F_ra(rm, app) - or_F_ra(rm, range);
                            ax += fpp * ddx / rm;
                            ay += fpp * ddy / rm;
                            az += fpp * ddz / rm;
                            pa++;
                        }
                    }
                }
            }
        }
        pc -= 1; /* the particle itself */
        af[3 * a + 0] = ax;
        af[3 * a + 1] = ay;
        af[3 * a + 2] = az;
    }
    float f2 = 0.0f, v0 = 0.0f, v1 = 0.0f, v2 = 0.0f;
    const int tx = it % g->nnt, ty = (it / g->nnt) % g->nnt, tz = it / (g->nnt * g->nnt);
    for (int a = 0; a < N; a++) {
        if (t.range[a] == 0.0f) continue;
        float a0 = af[3 * a + 0] * mass * mass;
        float a1 = af[3 * a + 1] * mass * mass;
        float a2 = af[3 * a + 2] * mass * mass;
        f2 = fmaxf(f2, a0 * a0 + a1 * a1 + a2 * a2);
        size_t ip = t.slot[a];
        float acc[3] = {a0, a1, a2};
        float vnew[3];
        for (int q = 0; q < 3; q++) {
            double term = (double)(acc[q] * a_mid * dt / 6.0f) / kPi;
            double v = (double)vp[3 * ip + q];
            if (dv64) v += dv64[3 * ip + q];
            vnew[q] = (float)(v + term);
            vp[3 * ip + q] = vnew[q];
        }
        v0 = fmaxf(v0, fabsf(vnew[0]));
        v1 = fmaxf(v1, fabsf(vnew[1]));
        v2 = fmaxf(v2, fabsf(vnew[2]));
        /* owner subtile of this physical particle, from its PP cell */
        uint32_t key = t.keys[a];
        int x = (int)(key % (uint32_t)S) - rcp, y = (int)((key / (uint32_t)S) % (uint32_t)S) - rcp,
            z = (int)(key / ((uint32_t)S * (uint32_t)S)) - rcp;
        int cs = rcp * d->ntt;
        int loc = (((tz * g->nnt + ty) * g->nnt + tx) * g->nns + z / cs) * g->nns * g->nns +
                  (y / cs) * g->nns + (x / cs);
#pragma omp atomic
        ptotal[owner[loc]] += 1.0;
    }
    *f2max_t = f2;
    vmax_t[0] = v0;
    vmax_t[1] = v1;
    vmax_t[2] = v2;
    if (cnt) {
        cnt[0] = pc;
        cnt[1] = pa;
    }
    free(af);
    free_tile(&t);
    if (ttotal) {
        int first = it * g->nns * g->nns * g->nns;
        ttotal[owner[first]] = now_sec() - t0;
    }
}

int or_pp_force(const or_geom* g, const int* ires, const int* ixyz3, const float* apm3,
                const int* ratio_sf, const int* rhoc, const int64_t* idx_b_r, const int16_t* xp,
                float* vp, const double* dv64, float mass, float a_mid, float dt, float* f2max,
                float* vmax, double* ptotal, double* ttotal, uint64_t* counts, int tile_first,
                int tile_count) {
    derived d = derive(g);
    int* il = (int*)malloc(sizeof(int) * d.ns3);
    int* ow = (int*)malloc(sizeof(int) * d.ns3);
    if (location_maps(g, &d, ires, ixyz3, il, ow) != 0) {
        free(il);
        free(ow);
        return -1;
    }
    int ntile = g->nnt * g->nnt * g->nnt;
    if (tile_count < 0) tile_count = ntile - tile_first;
    for (int s = 0; s < d.ns3; s++) {
        ptotal[s] = 0.0;
        if (ttotal) ttotal[s] = 0.0;
    }
    float f2 = 0.0f, v0 = 0.0f, v1 = 0.0f, v2 = 0.0f;
    uint64_t pc = 0, pa = 0;
    omp_set_max_active_levels(2);
#pragma omp parallel for num_threads(g->nteam) schedule(dynamic, 1) \
    reduction(max : f2, v0, v1, v2) reduction(+ : pc, pa)
    for (int q = tile_first; q < tile_first + tile_count; q++) {
        float tf2, tv[3];
        uint64_t c[2] = {0, 0};
        pp_tile(g, &d, q, il, ow, apm3, ratio_sf, rhoc, idx_b_r, xp, vp, dv64, mass, a_mid, dt,
                &tf2, tv, ptotal, ttotal, c);
        f2 = fmaxf(f2, tf2);
        v0 = fmaxf(v0, tv[0]);
        v1 = fmaxf(v1, tv[1]);
        v2 = fmaxf(v2, tv[2]);
        pc += c[0];
        pa += c[1];
    }
    *f2max = f2;
    vmax[0] = v0;
    vmax[1] = v1;
    vmax[2] = v2;
    if (counts) {
        counts[0] = pc;
        counts[1] = pa;
    }
    free(il);
    free(ow);
    return 0;
}

/* kick.f90:570-614 (Numerical Recipes heapsort of an index array, 1-based). */
void or_indexed_sort(int n, const float* arrin, int* indx) {
    for (int j = 1; j <= n; j++) indx[j - 1] = j;
    if (n < 2) return;
#define A(q) arrin[(q)-1]
#define IX(q) indx[(q)-1]
    int l = n / 2 + 1, ir = n;
    for (;;) {
        int indxt;
        float qv;
        if (l > 1) {
            l = l - 1;
            indxt = IX(l);
            qv = A(indxt);
        } else {
            indxt = IX(ir);
            qv = A(indxt);
            IX(ir) = IX(1);
            ir = ir - 1;
            if (ir == 1) {
                IX(1) = indxt;
                return;
            }
        }
        int i = l, j = l + l;
        while (j <= ir) {
            if (j < ir && A(IX(j)) < A(IX(j + 1))) j = j + 1;
            if (qv < A(IX(j))) {
                IX(i) = IX(j);
                i = j;
                j = j + j;
            } else {
                j = ir + 1;
            }
        }
        IX(i) = indxt;
    }
#undef A
#undef IX
}

/* kick.f90:107-131: per-subtile refinement level from the densest coarse cell. */
void or_select_levels(const or_geom* g, const int* rhoc, const int* ixyz3, int* ires, int* isort,
                      int* napm, float* npc_max_out) {
    derived d = derive(g);
    float* npc = (float*)malloc(sizeof(float) * d.ns3);
    float* neg = (float*)malloc(sizeof(float) * d.ns3);
    for (int q = 0; q < 7; q++) napm[q] = 0;
    for (int it = 0; it < d.ns3; it++) {
        const int* x = ixyz3 + 6 * it;
        int t = ((x[5] - 1) * g->nnt + (x[4] - 1)) * g->nnt + (x[3] - 1);
        const int* rc = rhoc + (size_t)t * d.nte * d.nte * d.nte;
        int mx = 0;
        for (int k = (x[2] - 1) * d.ntt; k < x[2] * d.ntt; k++)
            for (int j = (x[1] - 1) * d.ntt; j < x[1] * d.ntt; j++)
                for (int i = (x[0] - 1) * d.ntt; i < x[0] * d.ntt; i++) {
                    int v = rc[((size_t)(k + d.ncb) * d.nte + (j + d.ncb)) * d.nte + (i + d.ncb)];
                    if (v > mx) mx = v;
                }
        float p = (float)mx / 64.0f;
        npc[it] = p;
        neg[it] = -p;
        int L;
        if (p < 24.0f) L = 2;
        else if (p < 114.0f) L = 3;
        else if (p < 320.0f) L = 4;
        else if (p < 1350.0f) L = 5;
        else L = 6;
        ires[it] = L;
        napm[L - 1]++;
    }
    or_indexed_sort(d.ns3, neg, isort);
    if (npc_max_out) memcpy(npc_max_out, npc, sizeof(float) * d.ns3);
    free(npc);
    free(neg);
}

/* Apply a pending Delta-v (PM3-only kick) and return vmax over physical slots. */
void or_velocity_apply(const or_geom* g, const int* rhoc, const int64_t* idx_b_r, float* vp,
                       const double* dv64, float* vmax) {
    derived d = derive(g);
    int ntile = g->nnt * g->nnt * g->nnt;
    float v[3] = {0.f, 0.f, 0.f};
    for (int t = 0; t < ntile; t++) {
        const int* rc = rhoc + (size_t)t * d.nte * d.nte * d.nte;
        const int64_t* ib = idx_b_r + (size_t)t * d.nte * d.nte;
        for (int k = d.ncb; k < d.ncb + d.nt; k++)
            for (int j = d.ncb; j < d.ncb + d.nt; j++)
                for (int i = d.ncb; i < d.ncb + d.nt; i++) {
                    int np = rc[((size_t)k * d.nte + j) * d.nte + i];
                    int64_t nz = nzero_of(&d, rc, ib, k, j, i);
                    for (int l = 0; l < np; l++) {
                        size_t ip = (size_t)(nz + l);
                        for (int q = 0; q < 3; q++) {
                            float nv = (float)((double)vp[3 * ip + q] + dv64[3 * ip + q]);
                            vp[3 * ip + q] = nv;
                            v[q] = fmaxf(v[q], fabsf(nv));
                        }
                    }
                }
    }
    vmax[0] = v[0];
    vmax[1] = v[1];
    vmax[2] = v[2];
}

/* FP64 direct summation of the PP kernel F_ra(r,app) - F_ra(r,range_i) over
 * every particle within range_i (minimum image), for accuracy checks. */
static double F_ra64(double r, double apm) {
    double ep = 2.0 * r / apm;
    if (apm == 0.0 || ep > 2.0) return 1.0 / (r * r);
    if (ep > 1.0)
        return (1.0 / 35.0 / (apm * apm)) *
               (12.0 / (ep * ep) - 224.0 + 896.0 * ep - 840.0 * ep * ep + 224.0 * pow(ep, 3) +
                70.0 * pow(ep, 4) - 48.0 * pow(ep, 5) + 7.0 * pow(ep, 6));
    if (ep > 0.0)
        return (1.0 / 35.0 / (apm * apm)) * (224.0 * ep - 224.0 * pow(ep, 3) + 70.0 * pow(ep, 4) +
                                             48.0 * pow(ep, 5) - 21.0 * pow(ep, 6));
    return 0.0;
}

/* Plummer softening (no counterpart in the reference): the force of a unit
 * mass at distance r is r / (r^2 + eps^2)^(3/2), minus the same F_ra(r,
 * range) the mesh carries; eps = app. */
static double F_plummer64(double r, double eps) {
    const double q = r * r + eps * eps;
    return r / (q * sqrt(q));
}

/* The softened pair force of the given law (0: S2 F_ra, 1: Plummer). */
static double F_soft64(double r, double app, int law) {
    return law == 1 ? F_plummer64(r, app) : F_ra64(r, app);
}

void or_direct_pp_f64(const double* pos, int n, double box, const int* targets, int nt,
                      const double* range_t, double app, int law, double* acc) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int q = 0; q < nt; q++) {
        int i = targets[q];
        double rg = range_t[q], a[3] = {0, 0, 0};
        for (int j = 0; j < n; j++) {
            double dd[3], r2 = 0;
            for (int c = 0; c < 3; c++) {
                double v = pos[3 * j + c] - pos[3 * i + c];
                v -= box * floor(v / box + 0.5);
                dd[c] = v;
                r2 += v * v;
            }
            if (r2 > 0 && r2 < rg * rg) {
                double r = sqrt(r2);
                double f = F_soft64(r, app, law) - F_ra64(r, rg);
                for (int c = 0; c < 3; c++) a[c] += f * dd[c] / r;
            }
        }
        acc[3 * q + 0] = a[0];
        acc[3 * q + 1] = a[1];
        acc[3 * q + 2] = a[2];
    }
}

/* The same FP64 direct sum with the particles binned into cubic cells of
 * about the smallest range (at most 256 per side), each target scanning the
 * cells within its own range (falls back to all pairs when fewer than 3
 * cells fit the box). */
void or_direct_pp_f64_cells(const double* pos, int n, double box, const int* targets, int nt,
                            const double* range_t, double app, int law, double* acc) {
    double rmin = 0.0;
    for (int q = 0; q < nt; q++)
        if (range_t[q] > 0.0 && (rmin == 0.0 || range_t[q] < rmin)) rmin = range_t[q];
    int nc = rmin > 0.0 ? (int)floor(box / rmin) : 1;
    if (nc > 256) nc = 256;
    if (nc < 3) {
        or_direct_pp_f64(pos, n, box, targets, nt, range_t, app, law, acc);
        return;
    }
    const double cs = box / nc;
    const size_t ncell = (size_t)nc * nc * nc;
    int* cnt = (int*)calloc(ncell + 1, sizeof(int));
    int* cell_of = (int*)malloc(sizeof(int) * (size_t)n);
    int* order = (int*)malloc(sizeof(int) * (size_t)n);
    for (int j = 0; j < n; j++) {
        int c[3];
        for (int d = 0; d < 3; d++) {
            double x = pos[3 * j + d] - box * floor(pos[3 * j + d] / box);
            int k = (int)(x / cs);
            c[d] = k >= nc ? nc - 1 : (k < 0 ? 0 : k);
        }
        cell_of[j] = (c[2] * nc + c[1]) * nc + c[0];
        cnt[cell_of[j] + 1]++;
    }
    for (size_t c = 0; c < ncell; c++) cnt[c + 1] += cnt[c];
    int* fill = (int*)malloc(sizeof(int) * ncell);
    memcpy(fill, cnt, sizeof(int) * ncell);
    for (int j = 0; j < n; j++) order[fill[cell_of[j]]++] = j;
    free(fill);
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < nt; q++) {
        const int i = targets[q];
        const double rg = range_t[q];
        double a[3] = {0, 0, 0};
        const int ci = cell_of[i];
        const int cx = ci % nc, cy = (ci / nc) % nc, cz = ci / (nc * nc);
        int m = (int)ceil(rg / cs);
        if (2 * m + 1 > nc) m = (nc - 1) / 2;  /* the whole box, each cell once */
        for (int dz = -m; dz <= m; dz++)
            for (int dy = -m; dy <= m; dy++)
                for (int dx = -m; dx <= m; dx++) {
                    const int c = ((((cz + dz) % nc + nc) % nc) * nc + ((cy + dy) % nc + nc) % nc) * nc +
                                  ((cx + dx) % nc + nc) % nc;
                    for (int s = cnt[c]; s < cnt[c + 1]; s++) {
                        const int j = order[s];
                        double dd[3], r2 = 0;
                        for (int d = 0; d < 3; d++) {
                            double v = pos[3 * j + d] - pos[3 * i + d];
                            v -= box * floor(v / box + 0.5);
                            dd[d] = v;
                            r2 += v * v;
                        }
                        if (r2 > 0 && r2 < rg * rg) {
                            const double r = sqrt(r2);
                            const double f = F_soft64(r, app, law) - F_ra64(r, rg);
                            for (int d = 0; d < 3; d++) a[d] += f * dd[d] / r;
                        }
                    }
                }
        acc[3 * q + 0] = a[0];
        acc[3 * q + 1] = a[1];
        acc[3 * q + 2] = a[2];
    }
    free(cnt);
    free(cell_of);
    free(order);
}
