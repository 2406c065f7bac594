/*
 * C restatement of the GPU-ISAC multi-reference polar alignment (oracle).
 *
 * TEST INFRASTRUCTURE ONLY: built into oracle/_build/liboracle64.so (double)
 * and liboracle32.so (float, -DORACLE_FLOAT).  Only the tests, the smoke
 * check and the CPU-baseline leg of bench.py load these libraries, and only
 * to check or time against; the shipped library never links them.
 *
 * It restates, per (particle, reference) pair, the same semantics as
 * oracle/isac_oracle.py (the numpy restatement), which in turn cites the
 * reference CUDA sources:
 *   template  cuda/gpu_aln_common.cu:39-62      shifts  cuda/gpu_aln_common.cu:64-86
 *   sampling  cuda/gpu_aln_noref.cu:1257-1288   normalisation :1294-1351
 *   ring R2C  :1885, :2121-2125                 CCF  :1483-1537
 *   psi C2R   :2393, :2459                      decode  :1667-1794
 *   output    :517-540
 * The psi C2R of both mirrors is one complex inverse FFT of
 * Z[k] = C0[k] + i C1[k] (Hermitian-extended), so c0 = Re z, c1 = Im z.
 * Parallel over particles with OpenMP.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#ifdef ORACLE_FLOAT
typedef float real;
#define RCOS cosf
#define RSIN sinf
#define RSQRT sqrtf
#define RFLOOR floorf
#define RFABS fabsf
#else
typedef double real;
#define RCOS cos
#define RSIN sin
#define RSQRT sqrt
#define RFLOOR floor
#define RFABS fabs
#endif

typedef struct { real re, im; } cplx;

/* ------------------------------------------------------------- radix-2 FFT */
typedef struct { int n, logn; cplx *tw; int *rev; } fftplan;

static void plan_init(fftplan *p, int n) {
    p->n = n; p->logn = 0;
    while ((1 << p->logn) < n) p->logn++;
    p->tw = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    p->rev = (int *)malloc(sizeof(int) * (size_t)n);
    for (int k = 0; k < n; k++) {
        double a = -2.0 * M_PI * (double)k / (double)n;
        p->tw[k].re = (real)cos(a);
        p->tw[k].im = (real)sin(a);
        int r = 0;
        for (int b = 0; b < p->logn; b++) if (k & (1 << b)) r |= 1 << (p->logn - 1 - b);
        p->rev[k] = r;
    }
}
static void plan_free(fftplan *p) { free(p->tw); free(p->rev); }

/* in-place complex FFT; sign = -1 forward (e^{-i}), +1 inverse (unnormalised) */
static void fft(const fftplan *p, cplx *x, int sign) {
    int n = p->n;
    for (int i = 0; i < n; i++) {
        int j = p->rev[i];
        if (j > i) { cplx t = x[i]; x[i] = x[j]; x[j] = t; }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < half; k++) {
                cplx w = p->tw[k * step];
                if (sign > 0) w.im = -w.im;
                cplx a = x[i + k], b = x[i + k + half];
                cplx bw = { b.re * w.re - b.im * w.im, b.re * w.im + b.im * w.re };
                x[i + k].re = a.re + bw.re; x[i + k].im = a.im + bw.im;
                x[i + k + half].re = a.re - bw.re; x[i + k + half].im = a.im - bw.im;
            }
        }
    }
}

/* --------------------------------------------------------------- sampling */
static real bilinear(const float *img, int D, real x, real y) {
    real x0 = RFLOOR(x), y0 = RFLOOR(y);
    real fx = x - x0, fy = y - y0;
    int ix = (int)x0, iy = (int)y0;
    int xa = ix < 0 ? 0 : (ix > D - 1 ? D - 1 : ix);
    int xb = ix + 1 < 0 ? 0 : (ix + 1 > D - 1 ? D - 1 : ix + 1);
    int ya = iy < 0 ? 0 : (iy > D - 1 ? D - 1 : iy);
    int yb = iy + 1 < 0 ? 0 : (iy + 1 > D - 1 ? D - 1 : iy + 1);
    real v00 = img[ya * D + xa], v01 = img[ya * D + xb];
    real v10 = img[yb * D + xa], v11 = img[yb * D + xb];
    return (1 - fy) * ((1 - fx) * v00 + fx * v01) + fy * ((1 - fx) * v10 + fx * v11);
}

/* normalised ring spectra of one image at centre (cx, cy): out[R][K] */
static void spectra_one(const float *img, int D, real cx, real cy, int R, int L,
                        const real *ox, const real *oy, const fftplan *pl,
                        real *pol, cplx *buf, cplx *out, int weighted) {
    int K = L / 2 + 1;
    double sA = 0, sQ = 0, nn = 0;
    for (int i = 0; i < R; i++) {
        double A = 0, Q = 0;
        for (int j = 0; j < L; j++) {
            real v = bilinear(img, D, cx + ox[i * L + j], cy + oy[i * L + j]);
            pol[i * L + j] = v;
            A += v; Q += (double)v * v;
        }
        sA += (i + 1) * A; sQ += (i + 1) * Q; nn += (i + 1);
    }
    nn *= L;
    double mu = sA / nn;
    double sd = sqrt(fabs(sQ / nn - mu * mu));
    real rmu = (real)mu, rsd = (real)(sd + 1e-7);
    for (int i = 0; i < R; i++) {
        for (int j = 0; j < L; j++) {
            buf[j].re = (pol[i * L + j] - rmu) / rsd;
            buf[j].im = 0;
        }
        fft(pl, buf, -1);
        real w = weighted ? (real)(i + 1) : (real)1;
        for (int k = 0; k < K; k++) {
            out[i * K + k].re = buf[k].re * w;
            out[i * K + k].im = buf[k].im * w;
        }
    }
}

/* 7-point parabola (cuda/gpu_aln_noref.cu:1781-1787) */
static double parabola(const double *x) {
    double c2 = 49. * x[0] + 6. * x[1] - 21. * x[2] - 32. * x[3] - 27. * x[4] - 6. * x[5] + 31. * x[6];
    double c3 = 5. * x[0] - 3. * x[2] - 4. * x[3] - 3. * x[4] + 5. * x[6];
    return (c3 != 0.0) ? c2 / (2.0 * c3) - 4.0 : 0.0;
}

/* Shift grid in the reference's float loop order; returns count. */
int oracle_shift_grid(float rx, float ry, float step, float *out, int cap) {
    int n = 0;
    for (float sx = -rx; sx <= rx; sx += step)
        for (float sy = -ry; sy <= ry; sy += step) {
            if (n < cap) { out[2 * n] = sx; out[2 * n + 1] = sy; }
            n++;
        }
    return n;
}

/*
 * Multi-reference alignment of P particles against N references.
 * sbj [P][D][D], ref [N][D][D], acc_xy [P][2] (accumulated shifts).
 * Outputs (all [N][P] unless noted): d (raw max), peak [N][4P] = {angle, sx, sy, mirror},
 * shift_idx, t_idx, delta (parabola offset), runner (best distinct alignment).
 * Returns 0 on success, negative on bad arguments.
 */
int oracle_mref(const float *sbj, const float *ref, const float *acc_xy,
                int P, int N, int D, int R, int L,
                float rng_x, float rng_y, float step, int nthreads,
                float *d, float *peak, int *shift_idx, int *t_idx,
                double *delta, double *runner) {
    if (P <= 0 || N <= 0 || D <= 0 || R <= 0 || L < 8 || (L & (L - 1))) return -1;
    int K = L / 2 + 1;
    float shifts[2 * 4096];
    int S = oracle_shift_grid(rng_x, rng_y, step, shifts, 4096);
    if (S > 4096) return -2;
    real *ox = (real *)malloc(sizeof(real) * (size_t)R * L);
    real *oy = (real *)malloc(sizeof(real) * (size_t)R * L);
    for (int i = 0; i < R; i++)
        for (int j = 0; j < L; j++) {
            double a = 2.0 * M_PI * (double)j / (double)L;
            ox[i * L + j] = (real)((i + 1) * cos(a));
            oy[i * L + j] = (real)((i + 1) * sin(a));
        }
    fftplan pl;
    plan_init(&pl, L);
    int c = D / 2;
    /* reference spectra, ring-weighted */
    cplx *rsp = (cplx *)malloc(sizeof(cplx) * (size_t)N * R * K);
    {
        real *pol = (real *)malloc(sizeof(real) * (size_t)R * L);
        cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)L);
        for (int n = 0; n < N; n++)
            spectra_one(ref + (size_t)n * D * D, D, (real)c, (real)c, R, L, ox, oy, &pl,
                        pol, buf, rsp + (size_t)n * R * K, 1);
        free(pol); free(buf);
    }
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
        real *pol = (real *)malloc(sizeof(real) * (size_t)R * L);
        cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)L);
        cplx *ssp = (cplx *)malloc(sizeof(cplx) * (size_t)S * R * K);
        real *tab = (real *)malloc(sizeof(real) * (size_t)2 * S * L);
        cplx *z = (cplx *)malloc(sizeof(cplx) * (size_t)L);
#pragma omp for schedule(dynamic, 1)
        for (int p = 0; p < P; p++) {
            const float *img = sbj + (size_t)p * D * D;
            for (int s = 0; s < S; s++) {
                int isx = (int)shifts[2 * s], isy = (int)shifts[2 * s + 1];
                real cx = (real)(c + isx) + (real)acc_xy[2 * p];
                real cy = (real)(c + isy) + (real)acc_xy[2 * p + 1];
                spectra_one(img, D, cx, cy, R, L, ox, oy, &pl, pol, buf,
                            ssp + (size_t)s * R * K, 0);
            }
            for (int n = 0; n < N; n++) {
                const cplx *rw = rsp + (size_t)n * R * K;
                for (int s = 0; s < S; s++) {
                    const cplx *sp = ssp + (size_t)s * R * K;
                    for (int k = 0; k < L; k++) { z[k].re = 0; z[k].im = 0; }
                    for (int k = 0; k < K; k++) {
                        real p1 = 0, p2 = 0, p3 = 0, p4 = 0;
                        for (int i = 0; i < R; i++) {
                            cplx a = rw[i * K + k], b = sp[i * K + k];
                            p1 += a.re * b.re; p2 += a.im * b.im;
                            p3 += a.re * b.im; p4 += a.im * b.re;
                        }
                        /* C0 = R conj(S) = (p1+p2, p4-p3); C1 = conj(R S) = (p1-p2, -(p3+p4)) */
                        real c0r = p1 + p2, c0i = p4 - p3, c1r = p1 - p2, c1i = -(p3 + p4);
                        if (k == 0 || k == L / 2) { c0i = 0; c1i = 0; }
                        /* Z[k] = C0 + i C1 ; Z[L-k] = conj(C0) + i conj(C1) */
                        z[k].re = c0r - c1i; z[k].im = c0i + c1r;
                        if (k != 0 && k != L / 2) {
                            z[L - k].re = c0r + c1i; z[L - k].im = -c0i + c1r;
                        }
                    }
                    fft(&pl, z, +1);
                    for (int t = 0; t < L; t++) {
                        tab[(size_t)(0 * S + s) * L + t] = z[t].re;
                        tab[(size_t)(1 * S + s) * L + t] = z[t].im;
                    }
                }
                /* first max in (m, s, t) order, strict '>' from -FLT_MAX */
                real best = (real)-3.40282e+38;
                long bi = 0;
                for (long q = 0; q < 2L * S * L; q++)
                    if (tab[q] > best) { best = tab[q]; bi = q; }
                int m = (int)(bi / ((long)S * L));
                int s = (int)((bi / L) % S);
                int t = (int)(bi % L);
                const real *win = tab + (size_t)(m * S + s) * L;
                double x7[7];
                for (int q = -3; q <= 3; q++) x7[q + 3] = (double)win[(t + q + L) % L];
                double dl = parabola(x7);
                double ang = (360.0 / L) * t + (360.0 / L) * dl;
                ang = 360.0 - ang;
                if (m) ang += 180.0;
                if (ang >= 360.0) ang -= 360.0;
                /* runner-up over distinct alignments */
                double run = -INFINITY;
                for (int sl = 0; sl < 2 * S; sl++) {
                    const real *row = tab + (size_t)sl * L;
                    for (int tt = 0; tt < L; tt++) {
                        if (sl == m * S + s) {
                            int dd = ((tt - t) % L + L) % L;
                            if (dd > L / 2) dd = L - dd;
                            if (dd <= 3) continue;
                        }
                        if (row[tt] > run) run = row[tt];
                    }
                }
                size_t o = (size_t)n * P + p;
                d[o] = (float)best;
                peak[(size_t)n * 4 * P + 4 * p + 0] = (float)ang;
                peak[(size_t)n * 4 * P + 4 * p + 1] = acc_xy[2 * p] + shifts[2 * s];
                peak[(size_t)n * 4 * P + 4 * p + 2] = acc_xy[2 * p + 1] + shifts[2 * s + 1];
                peak[(size_t)n * 4 * P + 4 * p + 3] = (float)m;
                if (shift_idx) shift_idx[o] = s;
                if (t_idx) t_idx[o] = t;
                if (delta) delta[o] = dl;
                if (runner) runner[o] = run;
            }
        }
        free(pol); free(buf); free(ssp); free(tab); free(z);
    }
    free(rsp); free(ox); free(oy);
    plan_free(&pl);
    return 0;
}
