/*
 * PM3 oracle (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 * Restates the fine-mesh PM3 chain of pm_thread (kick.f90:403-543) and
 * potential_to_force (kick.f90:545-568) for one subtile job, with the GPU
 * drop-in semantics of pm3_fft_cuda.cu where they differ from the CPU path:
 *   - mesh coordinate scale ratio_cs*rsf applied exactly
 *     (pm3_fft_cuda.cu:293-295); kick.f90:438 divides by a real(4) 1/rsf;
 *   - ptotal counts physical particles (pm3_fft_cuda.cu:436-437);
 *   - Delta-v accumulates into a pending buffer (pm3_fft_cuda.cu:386-389).
 * Positions are decoded in FP64 (kick.f90:437-438); TSC weights are FP32
 * (kick.f90:415,442-445); the FFT, Green multiply, gradient and
 * interpolation run in FP64 so that this is a truth model for the FP32 GPU
 * path.  The FFT is an independent recursive mixed-radix DFT, pinned
 * against numpy.fft in tests.  Green_3D follows Green/Green.f90:1-80 in
 * FP64, including its real(4) constants.
 */
#include <complex.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static const double kPi = 3.14159265358979323846;
static const int64_t kIshift = -((int64_t)1 << 15);
static const double kRshift = 0.5 + 32768.0;
static const double kXres = 1.0 / 65536.0;

/* ------------------------------------------------------------------ FFT */

static void dft_rec(const double complex* in, double complex* out, int n, int is,
                    const double complex* w, int wstep, int ntot) {
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    int p = 2;
    while (n % p) p++;
    int m = n / p;
    for (int q = 0; q < p; q++) dft_rec(in + (size_t)q * is, out + (size_t)q * m, m, is * p, w, wstep * p, ntot);
    double complex tmp[16];
    double complex* t = p <= 16 ? tmp : (double complex*)malloc(sizeof(double complex) * p);
    for (int k = 0; k < m; k++) {
        for (int s = 0; s < p; s++) {
            double complex acc = 0;
            for (int q = 0; q < p; q++) {
                long e = ((long)q * (k + (long)s * m)) % n;
                acc += out[(size_t)q * m + k] * w[(size_t)e * wstep % ntot];
            }
            t[s] = acc;
        }
        for (int s = 0; s < p; s++) out[k + (size_t)s * m] = t[s];
    }
    if (t != tmp) free(t);
}

/* In-place complex DFT of length n on interleaved re/im, sign -1 forward. */
void or_fft1d(double* re_im, int n, int sign) {
    double complex* w = (double complex*)malloc(sizeof(double complex) * n);
    double complex* a = (double complex*)malloc(sizeof(double complex) * n);
    double complex* b = (double complex*)malloc(sizeof(double complex) * n);
    for (int t = 0; t < n; t++) {
        double ang = 2.0 * kPi * (double)t / (double)n;
        w[t] = cos(ang) + (double)sign * sin(ang) * I;
        a[t] = re_im[2 * t] + re_im[2 * t + 1] * I;
    }
    dft_rec(a, b, n, 1, w, 1, n);
    for (int t = 0; t < n; t++) {
        re_im[2 * t] = creal(b[t]);
        re_im[2 * t + 1] = cimag(b[t]);
    }
    free(w);
    free(a);
    free(b);
}

typedef struct {
    int n;
    double complex* wf; /* forward twiddles */
    double complex* wb; /* backward twiddles */
} fftplan;

static fftplan mkplan(int n) {
    fftplan p;
    p.n = n;
    p.wf = (double complex*)malloc(sizeof(double complex) * n);
    p.wb = (double complex*)malloc(sizeof(double complex) * n);
    for (int t = 0; t < n; t++) {
        double ang = 2.0 * kPi * (double)t / (double)n;
        p.wf[t] = cos(ang) - sin(ang) * I;
        p.wb[t] = cos(ang) + sin(ang) * I;
    }
    return p;
}
static void rmplan(fftplan* p) {
    free(p->wf);
    free(p->wb);
}

/* 3-D transforms; complex layout [kz][ky][kx] with kx in [0, n/2]. */
static void r2c3(const float* in, double complex* out, int n) {
    const int nh = n / 2 + 1;
    fftplan pl = mkplan(n);
#pragma omp parallel
    {
        double complex* a = (double complex*)malloc(sizeof(double complex) * n);
        double complex* b = (double complex*)malloc(sizeof(double complex) * n);
#pragma omp for schedule(static)
        for (int zy = 0; zy < n * n; zy++) {
            for (int x = 0; x < n; x++) a[x] = in[(size_t)zy * n + x];
            dft_rec(a, b, n, 1, pl.wf, 1, n);
            for (int x = 0; x < nh; x++) out[(size_t)zy * nh + x] = b[x];
        }
#pragma omp for schedule(static)
        for (int zx = 0; zx < n * nh; zx++) {
            int z = zx / nh, x = zx % nh;
            for (int y = 0; y < n; y++) a[y] = out[((size_t)z * n + y) * nh + x];
            dft_rec(a, b, n, 1, pl.wf, 1, n);
            for (int y = 0; y < n; y++) out[((size_t)z * n + y) * nh + x] = b[y];
        }
#pragma omp for schedule(static)
        for (int yx = 0; yx < n * nh; yx++) {
            int y = yx / nh, x = yx % nh;
            for (int z = 0; z < n; z++) a[z] = out[((size_t)z * n + y) * nh + x];
            dft_rec(a, b, n, 1, pl.wf, 1, n);
            for (int z = 0; z < n; z++) out[((size_t)z * n + y) * nh + x] = b[z];
        }
        free(a);
        free(b);
    }
    rmplan(&pl);
}

/* Unnormalised inverse; the imaginary parts of the kx=0 and kx=n/2 terms
 * are ignored, as a Hermitian C2R transform does. Destroys `c`. */
static void c2r3(double complex* c, double* out, int n) {
    const int nh = n / 2 + 1;
    fftplan pl = mkplan(n);
#pragma omp parallel
    {
        double complex* a = (double complex*)malloc(sizeof(double complex) * n);
        double complex* b = (double complex*)malloc(sizeof(double complex) * n);
#pragma omp for schedule(static)
        for (int yx = 0; yx < n * nh; yx++) {
            int y = yx / nh, x = yx % nh;
            for (int z = 0; z < n; z++) a[z] = c[((size_t)z * n + y) * nh + x];
            dft_rec(a, b, n, 1, pl.wb, 1, n);
            for (int z = 0; z < n; z++) c[((size_t)z * n + y) * nh + x] = b[z];
        }
#pragma omp for schedule(static)
        for (int zx = 0; zx < n * nh; zx++) {
            int z = zx / nh, x = zx % nh;
            for (int y = 0; y < n; y++) a[y] = c[((size_t)z * n + y) * nh + x];
            dft_rec(a, b, n, 1, pl.wb, 1, n);
            for (int y = 0; y < n; y++) c[((size_t)z * n + y) * nh + x] = b[y];
        }
#pragma omp for schedule(static)
        for (int zy = 0; zy < n * n; zy++) {
            const double complex* h = c + (size_t)zy * nh;
            for (int x = 0; x < nh; x++) a[x] = h[x];
            a[0] = creal(a[0]);
            if (n % 2 == 0) a[n / 2] = creal(a[n / 2]);
            for (int x = nh; x < n; x++) a[x] = conj(h[n - x]);
            dft_rec(a, b, n, 1, pl.wb, 1, n);
            for (int x = 0; x < n; x++) out[(size_t)zy * n + x] = creal(b[x]);
        }
        free(a);
        free(b);
    }
    rmplan(&pl);
}

void or_fft3d_r2c(const float* in, double* out, int n) { r2c3(in, (double complex*)out, n); }

void or_fft3d_c2r(const double* in, double* out, int n) {
    size_t nk = (size_t)(n / 2 + 1) * n * n;
    double complex* c = (double complex*)malloc(sizeof(double complex) * nk);
    memcpy(c, in, sizeof(double complex) * nk);
    c2r3(c, out, n);
    free(c);
}

/* ------------------------------------------------------------ Green_3D */

/* Green/Green.f90:1-80 with kshift=0, nglobal=n, (n1,n2,n3)=(n/2+1,n,n).
 * The real(4) literals and arguments of the Fortran are reproduced as
 * float values promoted to double. */
void or_green_3d(float* Gk, int n, float apm_soft, float apm_range, float gridsize) {
    const int nh = n / 2 + 1;
    const double twopi_over_n = 2.0 * kPi * (double)(1.0f / (float)n);
    const double c215 = (double)(2.0f / 15.0f);
    const float alpha_f = 4.0f / 3.0f;
    const double alpha = (double)alpha_f;
    const double one_m_alpha = (double)(1.0f - alpha_f);
    const double gs2 = (double)(gridsize * gridsize);
    double* fs = (double*)malloc(sizeof(double) * n);
    for (int t = 0; t < n; t++) fs[t] = twopi_over_n * (double)(((t + n / 2) % n) - n / 2);
#pragma omp parallel for schedule(dynamic, 1)
    for (int k3 = 0; k3 < n; k3++)
        for (int k2 = 0; k2 < n; k2++)
            for (int k1 = 0; k1 < nh; k1++) {
                double kv[3] = {fs[k1], fs[k2], fs[k3]};
                double D[3], ur[3] = {0, 0, 0}, u2n = 1.0;
                for (int c = 0; c < 3; c++) {
                    double s = sin(kv[c] / 2);
                    u2n *= 1 - s * s + c215 * (s * s) * (s * s);
                    D[c] = alpha * sin(kv[c]) + one_m_alpha * sin(2 * kv[c]) / 2;
                }
                for (int i3 = -2; i3 <= 2; i3++)
                    for (int i2 = -2; i2 <= 2; i2++)
                        for (int i1 = -2; i1 <= 2; i1++) {
                            double kn[3] = {kv[0] + 2 * kPi * i1, kv[1] + 2 * kPi * i2,
                                            kv[2] + 2 * kPi * i3};
                            double km = sqrt(kn[0] * kn[0] + kn[1] * kn[1] + kn[2] * kn[2]);
                            double u = 1.0;
                            for (int c = 0; c < 3; c++)
                                u *= kn[c] == 0 ? 1.0 : sin(kn[c] / 2) / (kn[c] / 2);
                            double u2 = u * u, uk2 = u2 * u2 * u2; /* **(2p+2), p=2 */
                            double ss, sr, y;
                            if (km == 0) ss = 1;
                            else {
                                y = (double)apm_soft * km / 2;
                                ss = 12 * pow(y, -4) * (2 - 2 * cos(y) - y * sin(y));
                            }
                            if (apm_range == 0) sr = 0;
                            else if (km == 0) sr = 1;
                            else {
                                y = (double)apm_range * km / 2;
                                sr = 12 * pow(y, -4) * (2 - 2 * cos(y) - y * sin(y));
                            }
                            if (km == 0) continue; /* only at k=0, zeroed below */
                            double f = (ss * ss - sr * sr) / (km * km);
                            for (int c = 0; c < 3; c++) ur[c] += uk2 * kn[c] * f;
                        }
                double num = D[0] * ur[0] + D[1] * ur[1] + D[2] * ur[2];
                double den = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
                double g = 4 * kPi * num / den / (u2n * u2n) / gs2;
                Gk[((size_t)k3 * n + k2) * nh + k1] = (float)g;
            }
    /* Gk(::n/2, ::n/2, ::n/2) = 0 (Green.f90:71) */
    for (int a = 0; a <= 1; a++)
        for (int b = 0; b <= 1; b++)
            for (int c = 0; c <= 1; c++)
                Gk[((size_t)(c * (n / 2)) * n + b * (n / 2)) * nh + a * (n / 2)] = 0.0f;
    free(fs);
}

/* ----------------------------------------------------------- PM3 chain */

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

static int64_t nzero_of(const derived* d, const int* rc, const int64_t* ib, int k, int j, int i) {
    const int* row = rc + ((size_t)k * d->nte + j) * d->nte;
    int64_t s = 0;
    for (int q = i; q < d->nte; q++) s += row[q];
    return ib[(size_t)k * d->nte + j] - s;
}

static inline double frac16(int16_t v) {
    return ((double)(int16_t)((int64_t)v + kIshift) + kRshift) * kXres;
}

/* TSC weights of kick.f90:439-445 (FP32 weights from an FP64 coordinate). */
static inline void tsc_w(double xpos, int* c, float w[3]) {
    double fl = floor(xpos);
    *c = (int)fl; /* 0-based index of the centre cell */
    float l3 = (float)(xpos - fl);
    w[0] = (1.0f - l3) * (1.0f - l3) / 2.0f;
    w[2] = l3 * l3 / 2.0f;
    w[1] = 1.0f - w[0] - w[2];
}

/* One mesh job: ncj coarse cells per edge starting at coarse cell u*ncj of
 * the tile (PM3: a subtile, ncj = ntt; PM2: the whole tile, ncj = nt), the
 * particles of nhalo further coarse cells on every side deposited
 * (pm%nloop: 1 for PM3, ncb for PM2). */
typedef struct {
    int L, rsf, nfp, nfb, n, tile, u[3], ncj, nhalo;
} job_t;

static int make_job(const or_geom* g, const derived* d, int it, const int* ires, const int* ixyz3,
                    const int* ratio_sf, job_t* j) {
    const int* x = ixyz3 + 6 * it;
    j->L = ires[it];
    if (j->L < 2 || j->L > 6) return -1;
    j->rsf = ratio_sf[j->L - 1];
    j->nfp = d->ntt * g->ratio_cs * j->rsf;
    j->nfb = g->ratio_cs * j->rsf;
    j->n = j->nfp + 2 * j->nfb;
    j->tile = ((x[5] - 1) * g->nnt + (x[4] - 1)) * g->nnt + (x[3] - 1);
    for (int c = 0; c < 3; c++) j->u[c] = x[c] - 1;
    j->ncj = d->ntt;
    j->nhalo = 1;
    return 0;
}

/* PM2 job of tile t (kick.f90:80-87): standard grid (gridsize 1), mesh
 * ngp + 2 ngb with ngp = nt ratio_cs, all ncb halo cells (nloop = ncb). */
static void make_job_pm2(const or_geom* g, const derived* d, int t, job_t* j) {
    j->L = 0;
    j->rsf = 1;
    j->nfp = d->nt * g->ratio_cs;
    j->nfb = g->ngb;
    j->n = j->nfp + 2 * j->nfb;
    j->tile = t;
    for (int c = 0; c < 3; c++) j->u[c] = 0;
    j->ncj = d->nt;
    j->nhalo = d->ncb;
}

/* Mesh coordinate (fine cells, subtile origin) of particle slot ip in
 * coarse cell (ci,cj,ck) (0-based nte indices) along one axis. */
static inline double mesh_coord(const or_geom* g, const derived* d, const job_t* j, int c_idx,
                                int16_t v, int axis) {
    double xc = (double)(c_idx - d->ncb) + frac16(v);
    return xc * (double)(g->ratio_cs * j->rsf) - (double)(j->u[axis] * j->nfp);
}

/* pm_thread's deposit, FFT solve and potential_to_force for one job
 * (kick.f90:419-494, 545-568). */
static void run_job(const or_geom* g, const derived* dp, job_t jb, const int* rhoc,
                    const int64_t* idx_b_r, const int16_t* xp, const float* green, float mass,
                    float* rho_out, double* phi_out, double* force_out, double* f2max_job) {
    const derived d = *dp;
    const int n = jb.n, nfb = jb.nfb, nfp = jb.nfp, ne = n + 2;
    const int* rc = rhoc + (size_t)jb.tile * d.nte * d.nte * d.nte;
    const int64_t* ib = idx_b_r + (size_t)jb.tile * d.nte * d.nte;
    /* rho_th over Fortran indices [-nfb, nfp+nfb+1] (nex = nfb+1) */
    float* rt = (float*)calloc((size_t)ne * ne * ne, sizeof(float));
    int lo[3], hi[3]; /* nc1-nloop .. nc2+nloop as 0-based nte indices */
    for (int c = 0; c < 3; c++) {
        lo[c] = jb.u[c] * jb.ncj + d.ncb - jb.nhalo;
        hi[c] = (jb.u[c] + 1) * jb.ncj + d.ncb - 1 + jb.nhalo;
    }
    for (int layer = 0; layer < 3; layer++)
        for (int k = lo[2] + layer; k <= hi[2]; k += 3)
            for (int jj = lo[1]; jj <= hi[1]; jj++)
                for (int i = lo[0]; i <= hi[0]; i++) {
                    int np = rc[((size_t)k * d.nte + jj) * d.nte + i];
                    int64_t nz = nzero_of(&d, rc, ib, k, jj, i);
                    for (int l = 0; l < np; l++) {
                        const int16_t* p = xp + 3 * (nz + l);
                        int cx, cy, cz;
                        float wx[3], wy[3], wz[3];
                        tsc_w(mesh_coord(g, &d, &jb, i, p[0], 0), &cx, wx);
                        tsc_w(mesh_coord(g, &d, &jb, jj, p[1], 1), &cy, wy);
                        tsc_w(mesh_coord(g, &d, &jb, k, p[2], 2), &cz, wz);
                        /* Fortran index of the centre cell is cx+1; storage
                         * offset = fortran + nfb (origin at -nfb). */
                        for (int a3 = 0; a3 < 3; a3++)
                            for (int a2 = 0; a2 < 3; a2++)
                                for (int a1 = 0; a1 < 3; a1++) {
                                    int X = cx + a1 + nfb, Y = cy + a2 + nfb, Z = cz + a3 + nfb;
                                    rt[((size_t)Z * ne + Y) * ne + X] += wx[a1] * wy[a2] * wz[a3];
                                }
                    }
                }
    /* density over Fortran [1-nfb, nfp+nfb] == storage [1, n] */
    float* rho = (float*)malloc(sizeof(float) * (size_t)n * n * n);
    for (int z = 0; z < n; z++)
        for (int y = 0; y < n; y++)
            for (int x = 0; x < n; x++)
                rho[((size_t)z * n + y) * n + x] = rt[((size_t)(z + 1) * ne + y + 1) * ne + x + 1] * mass;
    free(rt);
    if (rho_out) memcpy(rho_out, rho, sizeof(float) * (size_t)n * n * n);
    const int nh = n / 2 + 1;
    size_t nk = (size_t)nh * n * n;
    double complex* ck = (double complex*)malloc(sizeof(double complex) * nk);
    r2c3(rho, ck, n);
    free(rho);
    for (size_t q = 0; q < nk; q++) ck[q] *= (double)green[q];
    double* phi = (double*)malloc(sizeof(double) * (size_t)n * n * n);
    c2r3(ck, phi, n);
    free(ck);
    const double inv = 1.0 / ((double)n * n * n);
    for (size_t q = 0; q < (size_t)n * n * n; q++) phi[q] *= inv;
    if (phi_out) memcpy(phi_out, phi, sizeof(double) * (size_t)n * n * n);
    const int nf = nfp + 2;
    double* F = force_out ? force_out : (double*)malloc(sizeof(double) * 3 * (size_t)nf * nf * nf);
    double f2 = 0.0;
    or_potential_to_force(phi, n, nfb, nfp, F, &f2);
    free(phi);
    if (f2max_job) *f2max_job = f2;
    if (!force_out) free(F);
}

/* potential_to_force (kick.f90:545-568) with the weights of
 * parameters.f90:49-50: F_d = sum_m w_m phi(x + o_m e_d), o = (-2,-1,1,2),
 * w = ((alpha-1)/4, -alpha/2, alpha/2, (1-alpha)/4) in real(4), alpha = 4/3,
 * over force indices f in [0, nfp+1]^3 <-> phi storage f + nfb - 1 of the
 * n^3 potential; *f2max = max |F|^2 over the interior [1, nfp]^3. */
void or_potential_to_force(const double* phi, int n, int nfb, int nfp, double* F, double* f2max) {
    const float af = 4.0f / 3.0f;
    const double w[4] = {(double)((af - 1.0f) / 4.0f), (double)(-af / 2.0f), (double)(af / 2.0f),
                         (double)((1.0f - af) / 4.0f)};
    const int off[4] = {-2, -1, 1, 2};
    const int nf = nfp + 2;
    double f2 = 0.0;
#pragma omp parallel for reduction(max : f2) schedule(static)
    for (int k0 = 0; k0 < nf; k0++)
        for (int j0 = 0; j0 < nf; j0++)
            for (int i0 = 0; i0 < nf; i0++) {
                int pi = i0 + nfb - 1, pj = j0 + nfb - 1, pk = k0 + nfb - 1;
                double fx = 0, fy = 0, fz = 0;
                for (int m = 0; m < 4; m++) {
                    fx += phi[((size_t)pk * n + pj) * n + pi + off[m]] * w[m];
                    fy += phi[((size_t)pk * n + pj + off[m]) * n + pi] * w[m];
                    fz += phi[((size_t)(pk + off[m]) * n + pj) * n + pi] * w[m];
                }
                size_t u = ((size_t)k0 * nf + j0) * nf + i0;
                F[u] = fx;
                F[(size_t)nf * nf * nf + u] = fy;
                F[2 * (size_t)nf * nf * nf + u] = fz;
                if (i0 >= 1 && i0 <= nfp && j0 >= 1 && j0 <= nfp && k0 >= 1 && k0 <= nfp) {
                    double s = fx * fx + fy * fy + fz * fz;
                    if (s > f2) f2 = s;
                }
            }
    *f2max = f2;
}

int or_pm3_job(const or_geom* g, int it, const int* ires, const int* ixyz3, const int* ratio_sf,
               const int* rhoc, const int64_t* idx_b_r, const int16_t* xp, const float* green,
               float mass, float* rho_out, double* phi_out, double* force_out,
               double* f2max_job) {
    derived d = derive(g);
    job_t jb;
    if (make_job(g, &d, it, ires, ixyz3, ratio_sf, &jb)) return -1;
    run_job(g, &d, jb, rhoc, idx_b_r, xp, green, mass, rho_out, phi_out, force_out, f2max_job);
    return 0;
}

/* pm_thread's velocity update (kick.f90:497-538) of the job's physical
 * cells into the FP64 Delta-v; returns their particle count. */
static long interp_job(const or_geom* g, const derived* dp, const job_t* jbp, const int* rhoc,
                       const int64_t* idx_b_r, const int16_t* xp, const double* F, double kick,
                       double* dv64) {
    const derived d = *dp;
    const job_t jb = *jbp;
    const int nf = jb.nfp + 2;
    const int* rc = rhoc + (size_t)jb.tile * d.nte * d.nte * d.nte;
    const int64_t* ib = idx_b_r + (size_t)jb.tile * d.nte * d.nte;
    const size_t fs = (size_t)nf * nf * nf;
    long cnt = 0;
    for (int k = jb.u[2] * jb.ncj + d.ncb; k < (jb.u[2] + 1) * jb.ncj + d.ncb; k++)
        for (int jj = jb.u[1] * jb.ncj + d.ncb; jj < (jb.u[1] + 1) * jb.ncj + d.ncb; jj++)
            for (int i = jb.u[0] * jb.ncj + d.ncb; i < (jb.u[0] + 1) * jb.ncj + d.ncb; i++) {
                int np = rc[((size_t)k * d.nte + jj) * d.nte + i];
                int64_t nz = nzero_of(&d, rc, ib, k, jj, i);
                cnt += np;
                for (int l = 0; l < np; l++) {
                    size_t ip = (size_t)(nz + l);
                    const int16_t* p = xp + 3 * ip;
                    int cx, cy, cz;
                    float wx[3], wy[3], wz[3];
                    tsc_w(mesh_coord(g, &d, &jb, i, p[0], 0), &cx, wx);
                    tsc_w(mesh_coord(g, &d, &jb, jj, p[1], 1), &cy, wy);
                    tsc_w(mesh_coord(g, &d, &jb, k, p[2], 2), &cz, wz);
                    double dv[3] = {0, 0, 0};
                    /* force index of the centre cell = cx+1 (Fortran) */
                    for (int a3 = 0; a3 < 3; a3++)
                        for (int a2 = 0; a2 < 3; a2++)
                            for (int a1 = 0; a1 < 3; a1++) {
                                size_t u = ((size_t)(cz + a3) * nf + (cy + a2)) * nf + (cx + a1);
                                double wgt = (double)wx[a1] * (double)wy[a2] * (double)wz[a3];
                                dv[0] += F[u] * wgt;
                                dv[1] += F[fs + u] * wgt;
                                dv[2] += F[2 * fs + u] * wgt;
                            }
                    for (int c = 0; c < 3; c++) dv64[3 * ip + c] += dv[c] * kick;
                }
            }
    return cnt;
}

int or_pm3_force(const or_geom* g, const int* isort, const int* ires, const int* ixyz3,
                 const int* ratio_sf, const int* rhoc, const int64_t* idx_b_r, const int16_t* xp,
                 const float* const* greens, float mass, float amid, float dt, double* dv64,
                 float* f2max, double* ptotal) {
    derived d = derive(g);
    double f2 = 0.0;
    for (int s = 0; s < d.ns3; s++) ptotal[s] = 0.0;
    const double kick = (double)amid * (double)dt / (6.0 * kPi);
    for (int o = 0; o < d.ns3; o++) {
        int it = isort[o] - 1;
        job_t jb;
        if (make_job(g, &d, it, ires, ixyz3, ratio_sf, &jb)) return -1;
        const int nf = jb.nfp + 2;
        double* F = (double*)malloc(sizeof(double) * 3 * (size_t)nf * nf * nf);
        double fj = 0;
        run_job(g, &d, jb, rhoc, idx_b_r, xp, greens[jb.L], mass, NULL, NULL, F, &fj);
        if (fj > f2) f2 = fj;
        const long cnt = interp_job(g, &d, &jb, rhoc, idx_b_r, xp, F, kick, dv64);
        ptotal[it] = (double)cnt;
        free(F);
    }
    *f2max = (float)f2;
    return 0;
}

/* PM2 (kick.f90:76-105): pm_thread with pm_layer 2 on tiles tile0 ..
 * tile0+ntiles-1 (ntiles <= 0: all), Green's table Gk2 [n][n][n/2+1],
 * n = ngp + 2 ngb.  Adds to dv64; ptotal[t]: the particles of tile t's
 * physical cells. */
int or_pm2_force(const or_geom* g, const int* rhoc, const int64_t* idx_b_r, const int16_t* xp,
                 const float* green2, float mass, float amid, float dt, double* dv64,
                 float* f2max, double* ptotal, int tile0, int ntiles) {
    derived d = derive(g);
    double f2 = 0.0;
    const int ntile = g->nnt * g->nnt * g->nnt;
    const int t_end = ntiles > 0 ? tile0 + ntiles : ntile;
    if (tile0 < 0 || t_end > ntile) return -1;
    const double kick = (double)amid * (double)dt / (6.0 * kPi);
    for (int t = tile0; t < t_end; t++) {
        job_t jb;
        make_job_pm2(g, &d, t, &jb);
        const int nf = jb.nfp + 2;
        double* F = (double*)malloc(sizeof(double) * 3 * (size_t)nf * nf * nf);
        double fj = 0;
        run_job(g, &d, jb, rhoc, idx_b_r, xp, green2, mass, NULL, NULL, F, &fj);
        if (fj > f2) f2 = fj;
        ptotal[t] = (double)interp_job(g, &d, &jb, rhoc, idx_b_r, xp, F, kick, dv64);
        free(F);
    }
    *f2max = (float)f2;
    return 0;
}
