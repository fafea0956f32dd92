/*
 * CPU oracle for the GenPK power-spectrum path.  TEST INFRASTRUCTURE ONLY:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker (or the timed CPU baseline),
 * never as part of the GPU product path.
 *
 * It restates, in plain FP64 C with OpenMP, the reference algorithm of
 *   - fieldize()       fieldize.cpp:46-114   (CIC deposit, padded layout, +=)
 *   - onedinvwindow()  fieldize.cpp:117-121  (1-D inverse CIC window, FP64)
 *   - invwindow()      fieldize.cpp:125-133  (FP32 product of 3 factors, squared)
 *   - powerspectrum()  powerspectrum.c:35-110 (log-binned, CIC-deconvolved Re(F1 F2*))
 * plus an FP64 3-D real-to-complex FFT standing in for the FFTW call the
 * reference driver makes between fieldize() and powerspectrum()
 * (gen-pk.cpp:193,233: fftw_plan_dft_r2c_3d, in place, unnormalised, sign -1).
 * FFTW itself is not part of the reference tree; the FFT here follows the
 * DFT definition and is pinned in the tests against a brute-force DFT and
 * numpy's pocketfft.
 *
 * Parity pins (tests/test_oracle.py): test.cpp:31-86 known answers, brute
 * force on tiny inputs, closed forms and invariants (mass conservation,
 * sum of counts = d^3-1, translation = roll, linearity).
 */
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.1415926535897932384626433832795
#endif

/* Particles handled per buffered chunk before the serialised scatter
 * (fieldize.cpp:24 uses the same chunk length). */
#define ORC_CHUNK 64

/* --------------------------------------------------------------------- */
/* CIC deposit: follows fieldize.cpp:46-114.                              */
/* Row length of the last axis is 2*(dims/2+extra) (fieldize.cpp:48).    */
/* Positions are promoted to double and scaled by dims/boxsize            */
/* (fieldize.cpp:52,66); the "next" cell index is built from the          */
/* unwrapped floor plus one and wrapped, then the floor itself is wrapped */
/* (fieldize.cpp:67-75); corner c has bit0 = x, bit1 = y, bit2 = z        */
/* (fieldize.cpp:77-92).  The scatter of each 64-particle chunk is        */
/* serialised, as the reference does with its critical section            */
/* (fieldize.cpp:98-111).                                                 */
/* --------------------------------------------------------------------- */
int orc_fieldize(double boxsize, int dims, double *out, int64_t n,
                 const float *pos, const float *masses, double mass, int extra)
{
    const int64_t rowlen = 2 * (dims / 2 + extra);
    const int64_t plane = rowlen * (int64_t)dims;
    const double units = dims / boxsize;
    const int64_t nchunks = (n + ORC_CHUNK - 1) / ORC_CHUNK;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nchunks; c++) {
        double w[ORC_CHUNK][8];
        int64_t at[ORC_CHUNK][8];
        const int64_t first = c * ORC_CHUNK;
        const int64_t cnt = (first + ORC_CHUNK < n) ? ORC_CHUNK : n - first;
        for (int64_t p = 0; p < cnt; p++) {
            const int64_t j = first + p;
            const double m = masses ? (double)masses[j] : mass;
            double lo[3], hi[3];
            int64_t i0[3], i1[3];
            for (int a = 0; a < 3; a++) {
                const double u = (double)pos[3 * j + a] * units;
                /* fieldize.cpp:67 stores floor(x) into an int; on x86-64
                 * that conversion (cvttsd2si) gives INT_MIN for a NaN, an
                 * infinity or a value outside the int range, spelled out
                 * here so the C code has no undefined conversion */
                const double fd = floor(u);
                int fl = (fd >= -2147483648.0 && fd < 2147483648.0) ? (int)fd : INT_MIN;
                const double f = u - fl;
                hi[a] = f;
                lo[a] = 1.0 - f;
                int nx = (fl + 1) % dims;
                if (nx < 0) nx += dims;
                fl %= dims;
                if (fl < 0) fl += dims;
                i0[a] = fl;
                i1[a] = nx;
            }
            for (int corner = 0; corner < 8; corner++) {
                const int bx = corner & 1, by = (corner >> 1) & 1, bz = (corner >> 2) & 1;
                w[p][corner] = m * (bx ? hi[0] : lo[0]) * (by ? hi[1] : lo[1]) * (bz ? hi[2] : lo[2]);
                at[p][corner] = plane * (bx ? i1[0] : i0[0]) + rowlen * (by ? i1[1] : i0[1]) + (bz ? i1[2] : i0[2]);
            }
        }
#pragma omp critical(orc_fieldize_scatter)
        {
            for (int64_t p = 0; p < cnt; p++)
                for (int corner = 0; corner < 8; corner++)
                    out[at[p][corner]] += w[p][corner];
        }
    }
    return 0;
}

/* 1-D inverse CIC window: fieldize.cpp:117-121.  pi*k/(n*sin(pi*k/n)) in
 * double, 1 for k = 0.  The argument of sin divides by (float)n as the
 * reference does. */
double orc_onedinvwindow(int64_t k, int64_t n)
{
    return k ? M_PI * k / (n * sin(M_PI * k / (float)n)) : 1.0;
}

/* invwindow: fieldize.cpp:125-133.  Each factor is rounded to float, the
 * product of the three floats is formed in float, and its square is
 * returned in double; 0 for n = 0. */
double orc_invwindow(int64_t kx, int64_t ky, int64_t kz, int64_t n)
{
    if (n == 0) return 0;
    const float wx = (float)orc_onedinvwindow(kx, n);
    const float wy = (float)orc_onedinvwindow(ky, n);
    const float wz = (float)orc_onedinvwindow(kz, n);
    const float prod = (wx * wy) * wz;
    return (double)prod * (double)prod;
}

/* --------------------------------------------------------------------- */
/* FP64 FFT.  Iterative Stockham autosort on a batch of B interleaved     */
/* transforms (element n of transform b at x[n*B+b]); radix-4 and        */
/* radix-2 butterflies are written out, radices 3, 5 and any other prime */
/* use a direct DFT of that length.  Forward sign -1, unnormalised (the   */
/* FFTW convention the reference relies on).                              */
/* --------------------------------------------------------------------- */
typedef struct { double re, im; } cplx;

typedef struct {
    int n;
    int nfac;
    int fac[64];
    int sign;
    cplx *tw;   /* tw[k] = exp(sign * 2 pi i k / n), k < n */
} fftplan;

static void plan_init(fftplan *p, int n, int sign)
{
    p->n = n;
    p->sign = sign;
    p->nfac = 0;
    int m = n;
    while (m % 4 == 0) { p->fac[p->nfac++] = 4; m /= 4; }
    while (m % 2 == 0) { p->fac[p->nfac++] = 2; m /= 2; }
    while (m % 3 == 0) { p->fac[p->nfac++] = 3; m /= 3; }
    while (m % 5 == 0) { p->fac[p->nfac++] = 5; m /= 5; }
    for (int f = 7; m > 1; f += 2)
        while (m % f == 0) { p->fac[p->nfac++] = f; m /= f; }
    p->tw = (cplx *)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
    for (int k = 0; k < n; k++) {
        /* angle from the exact fraction k/n in long double, rounded once */
        long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
        p->tw[k].re = (double)cosl(a);
        p->tw[k].im = (double)(sign * sinl(a));
    }
}

static void plan_free(fftplan *p) { free(p->tw); p->tw = NULL; }

static inline cplx cmul(cplx a, cplx b)
{
    cplx r = { a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re };
    return r;
}

/* Stockham passes over B interleaved transforms; x, y: work buffers of
 * n*B elements; returns the buffer that holds the result. */
static cplx *fft_run_batch(const fftplan *p, cplx *x, cplx *y, int B)
{
    const int n = p->n;
    int ns = 1;
    for (int s = 0; s < p->nfac; s++) {
        const int r = p->fac[s];
        const int nb = n / r;
        const int twstep = n / (ns * r);
        for (int j = 0; j < nb; j++) {
            const int k = j % ns;
            const int base = (j / ns) * ns * r + k;
            if (r == 4) {
                const cplx t1 = p->tw[k * twstep], t2 = p->tw[2 * k * twstep], t3 = p->tw[3 * k * twstep];
                for (int b = 0; b < B; b++) {
                    cplx a0 = x[(j) * B + b];
                    cplx a1 = cmul(x[(j + nb) * B + b], t1);
                    cplx a2 = cmul(x[(j + 2 * nb) * B + b], t2);
                    cplx a3 = cmul(x[(j + 3 * nb) * B + b], t3);
                    cplx s02 = { a0.re + a2.re, a0.im + a2.im }, d02 = { a0.re - a2.re, a0.im - a2.im };
                    cplx s13 = { a1.re + a3.re, a1.im + a3.im }, d13 = { a1.re - a3.re, a1.im - a3.im };
                    /* multiply d13 by sign*i */
                    cplx rd13 = { -p->sign * d13.im, p->sign * d13.re };
                    y[(base) * B + b] = (cplx){ s02.re + s13.re, s02.im + s13.im };
                    y[(base + ns) * B + b] = (cplx){ d02.re + rd13.re, d02.im + rd13.im };
                    y[(base + 2 * ns) * B + b] = (cplx){ s02.re - s13.re, s02.im - s13.im };
                    y[(base + 3 * ns) * B + b] = (cplx){ d02.re - rd13.re, d02.im - rd13.im };
                }
            } else if (r == 2) {
                const cplx t1 = p->tw[k * twstep];
                for (int b = 0; b < B; b++) {
                    cplx a0 = x[(j) * B + b];
                    cplx a1 = cmul(x[(j + nb) * B + b], t1);
                    y[(base) * B + b] = (cplx){ a0.re + a1.re, a0.im + a1.im };
                    y[(base + ns) * B + b] = (cplx){ a0.re - a1.re, a0.im - a1.im };
                }
            } else {
                const int rstep = n / r;
                for (int b = 0; b < B; b++) {
                    cplx v[64];
                    for (int q = 0; q < r; q++)
                        v[q] = cmul(x[(j + q * nb) * B + b], p->tw[(q * k * twstep) % n]);
                    for (int m2 = 0; m2 < r; m2++) {
                        double sr = 0, si = 0;
                        for (int q = 0; q < r; q++) {
                            const cplx t = p->tw[((q * m2) % r) * rstep];
                            sr += v[q].re * t.re - v[q].im * t.im;
                            si += v[q].re * t.im + v[q].im * t.re;
                        }
                        y[(base + m2 * ns) * B + b] = (cplx){ sr, si };
                    }
                }
            }
        }
        cplx *t = x; x = y; y = t;
        ns *= r;
    }
    return x;
}

#define ORC_FFT_BATCH 8

/* In-place 3-D r2c on the padded layout [d][d][rowlen], rowlen = 2(d/2+1):
 * the same memory contract as the reference's in-place FFTW plan
 * (gen-pk.cpp:176-193).  Output: complex [d][d][d/2+1]. */
int orc_fft_r2c_3d(int d, double *field)
{
    if (d <= 0) return 1;
    const int64_t hz = d / 2 + 1;
    const int64_t rowlen = 2 * hz;
    const int BB = ORC_FFT_BATCH;
    fftplan pz, py;
    plan_init(&pz, d, -1);
    plan_init(&py, d, -1);
#pragma omp parallel
    {
        cplx *a = (cplx *)malloc(sizeof(cplx) * d * BB);
        cplx *b = (cplx *)malloc(sizeof(cplx) * d * BB);
        /* z: real -> half complex, BB lines at a time */
#pragma omp for schedule(static)
        for (int64_t l0 = 0; l0 < (int64_t)d * d; l0 += BB) {
            const int nl = (l0 + BB <= (int64_t)d * d) ? BB : (int)((int64_t)d * d - l0);
            for (int l = 0; l < nl; l++) {
                const double *row = field + (l0 + l) * rowlen;
                for (int z = 0; z < d; z++) { a[z * BB + l].re = row[z]; a[z * BB + l].im = 0; }
            }
            for (int l = nl; l < BB; l++)
                for (int z = 0; z < d; z++) a[z * BB + l] = (cplx){ 0, 0 };
            cplx *r = fft_run_batch(&pz, a, b, BB);
            for (int l = 0; l < nl; l++) {
                double *row = field + (l0 + l) * rowlen;
                for (int64_t k = 0; k < hz; k++) { row[2 * k] = r[k * BB + l].re; row[2 * k + 1] = r[k * BB + l].im; }
            }
        }
        /* y: for each x plane, BB consecutive kz columns at a time */
        const int64_t ncb = (hz + BB - 1) / BB;
#pragma omp for schedule(static)
        for (int64_t t = 0; t < (int64_t)d * ncb; t++) {
            const int64_t x = t / ncb, k0 = (t % ncb) * BB;
            const int nc = (k0 + BB <= hz) ? BB : (int)(hz - k0);
            cplx *base = (cplx *)field + x * d * hz + k0;
            for (int y = 0; y < d; y++)
                for (int c = 0; c < BB; c++) a[y * BB + c] = c < nc ? base[y * hz + c] : (cplx){ 0, 0 };
            cplx *r = fft_run_batch(&py, a, b, BB);
            for (int y = 0; y < d; y++)
                for (int c = 0; c < nc; c++) base[y * hz + c] = r[y * BB + c];
        }
        /* x: for each (ky, kz block) */
#pragma omp for schedule(static)
        for (int64_t t = 0; t < (int64_t)d * ncb; t++) {
            const int64_t yy = t / ncb, k0 = (t % ncb) * BB;
            const int nc = (k0 + BB <= hz) ? BB : (int)(hz - k0);
            cplx *base = (cplx *)field + yy * hz + k0;
            for (int x = 0; x < d; x++)
                for (int c = 0; c < BB; c++) a[x * BB + c] = c < nc ? base[(int64_t)x * d * hz + c] : (cplx){ 0, 0 };
            cplx *r = fft_run_batch(&py, a, b, BB);
            for (int x = 0; x < d; x++)
                for (int c = 0; c < nc; c++) base[(int64_t)x * d * hz + c] = r[x * BB + c];
        }
        free(a);
        free(b);
    }
    plan_free(&pz);
    plan_free(&py);
    return 0;
}

/* Unnormalised inverse of orc_fft_r2c_3d (sign +1, Hermitian half input),
 * in place on the padded layout.  Used to pin the GPU c2r pass. */
int orc_fft_c2r_3d(int d, double *field)
{
    if (d <= 0 || (d & 1)) return 1;
    const int64_t hz = d / 2 + 1;
    const int64_t rowlen = 2 * hz;
    fftplan p;
    plan_init(&p, d, +1);
#pragma omp parallel
    {
        cplx *a = (cplx *)malloc(sizeof(cplx) * d);
        cplx *b = (cplx *)malloc(sizeof(cplx) * d);
#pragma omp for schedule(static)
        for (int64_t col = 0; col < (int64_t)d * hz; col++) {
            cplx *base = (cplx *)field + col;
            for (int x = 0; x < d; x++) a[x] = base[(int64_t)x * d * hz];
            cplx *r = fft_run_batch(&p, a, b, 1);
            for (int x = 0; x < d; x++) base[(int64_t)x * d * hz] = r[x];
        }
#pragma omp for schedule(static)
        for (int64_t col = 0; col < (int64_t)d * hz; col++) {
            const int64_t x = col / hz, kz = col % hz;
            cplx *base = (cplx *)field + x * d * hz + kz;
            for (int y = 0; y < d; y++) a[y] = base[y * hz];
            cplx *r = fft_run_batch(&p, a, b, 1);
            for (int y = 0; y < d; y++) base[y * hz] = r[y];
        }
#pragma omp for schedule(static)
        for (int64_t line = 0; line < (int64_t)d * d; line++) {
            double *row = field + line * rowlen;
            /* rebuild the full Hermitian line, then a complex inverse DFT */
            for (int64_t k = 0; k < hz; k++) { a[k].re = row[2 * k]; a[k].im = row[2 * k + 1]; }
            for (int64_t k = hz; k < d; k++) { a[k].re = row[2 * (d - k)]; a[k].im = -row[2 * (d - k) + 1]; }
            /* the imaginary parts of the self-conjugate modes do not enter a c2r */
            a[0].im = 0;
            a[d / 2].im = 0;
            cplx *r = fft_run_batch(&p, a, b, 1);
            for (int z = 0; z < d; z++) row[z] = r[z].re;
            row[d] = 0;
            row[d + 1] = 0;
        }
        free(a);
        free(b);
    }
    plan_free(&p);
    return 0;
}

/* --------------------------------------------------------------------- */
/* powerspectrum: powerspectrum.c:35-110.                                 */
/* KVAL(n) = n <= d/2 ? n : n - d (powerspectrum.c:33).  Bin index        */
/* floor(binsperunit * log(kk)), binsperunit = (nrbins-1)/log(sqrt(3)d/2) */
/* (powerspectrum.c:38,66,75,84).  Planes kz = 0 (DC skipped) and         */
/* kz = d/2 carry weight 1, the rest weight 2 (powerspectrum.c:61-89).    */
/* Per-thread bins merged at the end (powerspectrum.c:93-100), then the   */
/* normalisation of powerspectrum.c:102-108.                              */
/* The complex fields are passed as interleaved doubles.                  */
/* --------------------------------------------------------------------- */
static inline int64_t kval(int64_t n, int64_t d) { return n <= d / 2 ? n : n - d; }

int orc_powerspectrum(int64_t dims, const double *f1, const double *f2, int nrbins,
                      double *power, int *count, double *keffs, double m1, double m2)
{
    const double binsperunit = (nrbins - 1) / log(sqrt(3) * dims / 2.0);
    const int64_t hz = dims / 2 + 1;
    memset(power, 0, sizeof(double) * nrbins);
    memset(count, 0, sizeof(int) * nrbins);
    memset(keffs, 0, sizeof(double) * nrbins);
    int bad = 0;
#pragma omp parallel
    {
        double *pp = (double *)calloc(nrbins, sizeof(double));
        double *kp = (double *)calloc(nrbins, sizeof(double));
        int *cp = (int *)calloc(nrbins, sizeof(int));
#pragma omp for schedule(static) nowait
        for (int64_t i = 0; i < dims; i++) {
            const double ki = (double)kval(i, dims);
            for (int64_t j = 0; j < dims; j++) {
                const double kj = (double)kval(j, dims);
                const int64_t base = (i * dims + j) * hz;
                for (int64_t k = 0; k <= dims / 2; k++) {
                    /* kz = 0 plane first (DC skipped), then kz = d/2, then the rest,
                     * exactly the three cases of the reference */
                    const double kk_ = kval(k, dims);
                    const double kk = sqrt(ki * ki + kj * kj + kk_ * kk_);
                    double h;
                    if (k == 0) {
                        if (!(kk > 0)) continue;
                        h = 1;
                    } else if (k == dims / 2) {
                        h = 1;
                    } else {
                        h = 2;
                    }
                    const int64_t b = (int64_t)floor(binsperunit * log(kk));
                    if (b < 0 || b >= nrbins) { bad = 1; continue; }
                    const int64_t at = 2 * (base + k);
                    const double re = f1[at] * f2[at] + f1[at + 1] * f2[at + 1];
                    const double w = orc_invwindow(kval(i, dims), kval(j, dims), kval(k, dims), dims);
                    pp[b] += h * re * (w * w);
                    kp[b] += h * kk;
                    cp[b] += (int)h;
                }
            }
        }
#pragma omp critical(orc_ps_merge)
        {
            for (int b = 0; b < nrbins; b++) {
                power[b] += pp[b];
                count[b] += cp[b];
                keffs[b] += kp[b];
            }
        }
        free(pp);
        free(kp);
        free(cp);
    }
    for (int b = 0; b < nrbins; b++) {
        if (count[b]) {
            power[b] /= m1 * m2;
            power[b] /= count[b];
            keffs[b] /= count[b];
        }
    }
    return bad;
}

/* Bin index of the half-spectrum mode with integer |k|^2 = q, computed by
 * the same FP64 formula as orc_powerspectrum (powerspectrum.c:38,66). */
int64_t orc_bin_of_q(int64_t dims, int nrbins, int64_t q)
{
    const double binsperunit = (nrbins - 1) / log(sqrt(3) * dims / 2.0);
    return (int64_t)floor(binsperunit * log(sqrt((double)q)));
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
