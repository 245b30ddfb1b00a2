/*
 * fft.c -- FP64 oracle for the FFT layer of ct2 (src/fft/).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Conventions pinned by the reference's own test (fft_test.cpp):
 *   - 1-D inputs are bit-reversed first (scramble, fft_test.cpp:88,110,139);
 *   - forward sign and normalisation are FFTW_FORWARD, unnormalised
 *     (fft_test.cpp:66);
 *   - real output is half-complex: X[k] = Re F_k, X[N-k] = Im F_k for
 *     0<k<N/2, X[0], X[N/2] real (fft_test.cpp:114-126, 143-155);
 *   - lanes are interleaved: complex<simd_f64> is {re[4], im[4]} per element
 *     (fft_test.cpp:82-97), simd_f64 real data is [n][4].
 * dir = -1 is the conjugate transform (the reference's inverse is wrong,
 * fft_complex_vector.cpp:57-58; the intent is restated).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int ilog2i(int n) {
    int l = 0;
    while ((1 << l) < n) l++;
    return l;
}

static int bitrev(int j, int bits) {
    int r = 0;
    for (int b = 0; b < bits; b++) r |= ((j >> b) & 1) << (bits - 1 - b);
    return r;
}

/* scramble.cpp:26-43: within each of N1 groups, permute N2 blocks of N3
 * contiguous values by bit-reversal of the block index. */
void orc_scramble(double* X, int N1, int N2, int N3) {
    const int bits = ilog2i(N2);
    double* tmp = (double*)malloc(sizeof(double) * (size_t)N3);
    for (int i = 0; i < N1; i++) {
        for (int j = 0; j < N2; j++) {
            const int l = bitrev(j, bits);
            if (j < l) {
                double* a = X + (size_t)N3 * ((size_t)N2 * i + j);
                double* b = X + (size_t)N3 * ((size_t)N2 * i + l);
                memcpy(tmp, a, sizeof(double) * (size_t)N3);
                memcpy(a, b, sizeof(double) * (size_t)N3);
                memcpy(b, tmp, sizeof(double) * (size_t)N3);
            }
        }
    }
    free(tmp);
}

/* twiddles.cpp:6-54: cos(-2 pi n/N), sin(-2 pi n/N), sin(+2 pi n/N). */
void orc_twiddles(int N, double* cosw, double* sinw, double* invsinw) {
    const double twopi = 2.0 * M_PI;
    for (int n = 0; n < N; n++) {
        if (cosw) cosw[n] = cos(-twopi * n / N);
        if (sinw) sinw[n] = sin(-twopi * n / N);
        if (invsinw) invsinw[n] = sin(twopi * n / N);
    }
}

typedef struct {
    double re, im;
} cplx;

static inline cplx cmul(cplx a, cplx b) {
    cplx c = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return c;
}

/* Recursive decimation-in-time on bit-reversed input (fft_complex_vector.cpp:
 * 9-65): radix-2 when log2 N is odd and N <= 32, radix-4 otherwise.  In
 * bit-reversed storage the four quarter blocks hold the sub-transforms of
 * the residues 0,2,1,3 (mod 4); the halves those of 0,1 (mod 2). */
static void fft_rec(cplx* X, int N, const cplx* W, int wstride, int dir) {
    if (N <= 1) return;
    if ((ilog2i(N) % 2 == 1) && (N <= 32)) {
        const int h = N / 2;
        fft_rec(X, h, W, wstride * 2, dir);
        fft_rec(X + h, h, W, wstride * 2, dir);
        for (int k = 0; k < h; k++) {
            const cplx e = X[k], o = cmul(W[(size_t)k * wstride], X[k + h]);
            X[k].re = e.re + o.re;
            X[k].im = e.im + o.im;
            X[k + h].re = e.re - o.re;
            X[k + h].im = e.im - o.im;
        }
    } else {
        const int q = N / 4;
        for (int b = 0; b < 4; b++) fft_rec(X + b * q, q, W, wstride * 4, dir);
        for (int k = 0; k < q; k++) {
            /* residue r lives in block rev2(r): 0->0, 1->2, 2->1, 3->3 */
            const cplx t0 = X[k];
            const cplx t1 = cmul(W[(size_t)k * wstride], X[k + 2 * q]);
            const cplx t2 = cmul(W[(size_t)2 * k * wstride], X[k + q]);
            const cplx t3 = cmul(W[(size_t)3 * k * wstride], X[k + 3 * q]);
            const cplx a = {t0.re + t2.re, t0.im + t2.im};
            const cplx b = {t0.re - t2.re, t0.im - t2.im};
            const cplx c = {t1.re + t3.re, t1.im + t3.im};
            const cplx d = {t1.re - t3.re, t1.im - t3.im};
            /* omega_4 = -i*dir: omega*d */
            const cplx od = {dir * d.im, -dir * d.re};
            X[k].re = a.re + c.re;
            X[k].im = a.im + c.im;
            X[k + q].re = b.re + od.re;
            X[k + q].im = b.im + od.im;
            X[k + 2 * q].re = a.re - c.re;
            X[k + 2 * q].im = a.im - c.im;
            X[k + 3 * q].re = b.re - od.re;
            X[k + 3 * q].im = b.im - od.im;
        }
    }
}

/* twiddle table w^n = e^{-2 pi i dir n/N} from the reference's tables */
static cplx* make_w(int N, int dir) {
    cplx* W = (cplx*)malloc(sizeof(cplx) * (size_t)N);
    double* c = (double*)malloc(sizeof(double) * (size_t)N);
    double* s = (double*)malloc(sizeof(double) * (size_t)N);
    double* is = (double*)malloc(sizeof(double) * (size_t)N);
    orc_twiddles(N, c, s, is);
    for (int n = 0; n < N; n++) {
        W[n].re = c[n];
        W[n].im = dir > 0 ? s[n] : is[n];
    }
    free(c);
    free(s);
    free(is);
    return W;
}

/* fft_complex(complex<fft_simd>*, N, dir), fft_complex_vector.cpp:9-15.
 * Element n lane l: re at X[2*lanes*n + l], im at X[2*lanes*n + lanes + l]. */
void orc_fft_complex(double* X, int N, int dir, int lanes) {
    cplx* W = make_w(N, dir);
    cplx* buf = (cplx*)malloc(sizeof(cplx) * (size_t)N);
    for (int l = 0; l < lanes; l++) {
        for (int n = 0; n < N; n++) {
            buf[n].re = X[(size_t)2 * lanes * n + l];
            buf[n].im = X[(size_t)2 * lanes * n + lanes + l];
        }
        fft_rec(buf, N, W, 1, dir);
        for (int n = 0; n < N; n++) {
            X[(size_t)2 * lanes * n + l] = buf[n].re;
            X[(size_t)2 * lanes * n + lanes + l] = buf[n].im;
        }
    }
    free(buf);
    free(W);
}

/* fft_real(fft_float*, N, dir) (fft_real_scalar.cpp:9-15) and
 * fft_real(fft_simd*, N, dir) (fft_real_vector.cpp:9-15): real bit-reversed
 * input -> half-complex output.  Restated through the complex transform of
 * the same (still bit-reversed) sequence. */
void orc_fft_real(double* X, int N, int dir, int lanes) {
    cplx* W = make_w(N, dir);
    cplx* buf = (cplx*)malloc(sizeof(cplx) * (size_t)N);
    for (int l = 0; l < lanes; l++) {
        for (int n = 0; n < N; n++) {
            buf[n].re = X[(size_t)lanes * n + l];
            buf[n].im = 0.0;
        }
        fft_rec(buf, N, W, 1, dir);
        X[l] = buf[0].re;
        if (N > 1) X[(size_t)lanes * (N / 2) + l] = buf[N / 2].re;
        for (int k = 1; k < N / 2; k++) {
            X[(size_t)lanes * k + l] = buf[k].re;
            X[(size_t)lanes * (N - k) + l] = buf[k].im;
        }
    }
    free(buf);
    free(W);
}

/* 3-D complex transform, natural order in and out, interleaved complex,
 * index (i*n + j)*n + k.  Separable: a 1-D transform along each axis, the
 * plan the 3-D stubs intend (fft_r2c_3d.cpp:5-30). */
void orc_fft3d(double* X, int n, int dir) {
    cplx* W = make_w(n, dir);
    cplx* line = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    cplx* D = (cplx*)X;
    const int bits = ilog2i(n);
    const size_t s[3] = {(size_t)n * n, (size_t)n, 1};
    for (int ax = 0; ax < 3; ax++) {
        const size_t sa = s[ax], sb = s[(ax + 1) % 3], sc = s[(ax + 2) % 3];
        for (int b = 0; b < n; b++) {
            for (int c = 0; c < n; c++) {
                const size_t base = b * sb + c * sc;
                for (int t = 0; t < n; t++) line[bitrev(t, bits)] = D[base + t * sa];
                fft_rec(line, n, W, 1, dir);
                for (int t = 0; t < n; t++) D[base + t * sa] = line[t];
            }
        }
    }
    free(line);
    free(W);
}

/* fft_r2c_3d intent (fft_r2c_3d.cpp:5-30 with the commented-out test at
 * fft_test.cpp:161-197): real N^3 in, FFTW r2c half spectrum out in planar
 * form: out[(N/2+1)(N i + j) + k] = Re, out[dimag + same] = Im,
 * dimag = N^2 (N/2+1), k in [0, N/2]. */
void orc_fft_r2c_3d(const double* in, double* out, int N) {
    const size_t n3 = (size_t)N * N * N;
    const int h = N / 2 + 1;
    const size_t dimag = (size_t)N * N * h;
    double* c = (double*)calloc(2 * n3, sizeof(double));
    for (size_t q = 0; q < n3; q++) c[2 * q] = in[q];
    orc_fft3d(c, N, 1);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            for (int k = 0; k < h; k++) {
                const size_t src = ((size_t)i * N + j) * N + k;
                const size_t dst = ((size_t)i * N + j) * h + k;
                out[dst] = c[2 * src];
                out[dimag + dst] = c[2 * src + 1];
            }
    free(c);
}

/* fft_c2r_3d intent (fft_c2r_3d.cpp:4-40): the planar half spectrum above
 * back to N^3 reals, unnormalised (FFTW c2r convention, which the Green's
 * path assumes by scaling with 1/N^3 once, particle.cpp:152,170). */
void orc_fft_c2r_3d(const double* in, double* out, int N) {
    const size_t n3 = (size_t)N * N * N;
    const int h = N / 2 + 1;
    const size_t dimag = (size_t)N * N * h;
    double* c = (double*)calloc(2 * n3, sizeof(double));
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            for (int k = 0; k < N; k++) {
                double re, im;
                if (k < h) {
                    const size_t q = ((size_t)i * N + j) * h + k;
                    re = in[q];
                    im = in[dimag + q];
                } else { /* Hermitian partner (-i,-j,-k) */
                    const int i2 = (N - i) % N, j2 = (N - j) % N, k2 = N - k;
                    const size_t q = ((size_t)i2 * N + j2) * h + k2;
                    re = in[q];
                    im = -in[dimag + q];
                }
                const size_t d = ((size_t)i * N + j) * N + k;
                c[2 * d] = re;
                c[2 * d + 1] = im;
            }
    orc_fft3d(c, N, -1);
    for (size_t q = 0; q < n3; q++) out[q] = c[2 * q];
    free(c);
}

/* transpose.cpp:7-31 intent: swap the y and z axes of a real N^3 cube in
 * place (the reference swaps every pair twice, i.e. does nothing). */
void orc_transpose_yz(double* X, int N) {
    for (int x = 0; x < N; x++)
        for (int y = 0; y < N; y++)
            for (int z = y + 1; z < N; z++) {
                const size_t i = ((size_t)x * N + y) * N + z;
                const size_t j = ((size_t)x * N + z) * N + y;
                const double t = X[i];
                X[i] = X[j];
                X[j] = t;
            }
}

/* transpose.cpp:33-73: half-complex rows along y of in[x][y][z] (real N^3)
 * -> planar complex out[x][z][ky], ky in [0,N/2], imaginary plane at
 * dimag = N^2 (N/2+1). */
void orc_transpose_r2c(const double* in, double* out, int N) {
    const int h = N / 2 + 1, No2 = N / 2;
    const size_t dimag = (size_t)N * N * h;
    for (int ix = 0; ix < N; ix++)
        for (int iy = 0; iy < h; iy++)
            for (int iz = 0; iz < N; iz++) {
                const size_t ir = ((size_t)ix * N + iy) * N + iz;
                const size_t jr = ((size_t)ix * N + iz) * h + iy;
                out[jr] = in[ir];
                if (iy == 0 || iy == No2) {
                    out[jr + dimag] = 0.0;
                } else {
                    const size_t ii = ((size_t)ix * N + (N - iy)) * N + iz;
                    out[jr + dimag] = in[ii];
                }
            }
}

/* transpose.cpp:75-116 intent: the inverse of orc_transpose_r2c (the
 * reference's loop bounds and output stride are wrong). */
void orc_transpose_c2r(const double* in, double* out, int N) {
    const int h = N / 2 + 1, No2 = N / 2;
    const size_t dimag = (size_t)N * N * h;
    for (int ix = 0; ix < N; ix++)
        for (int iy = 0; iy < h; iy++)
            for (int iz = 0; iz < N; iz++) {
                const size_t ir = ((size_t)ix * N + iy) * N + iz;
                const size_t jr = ((size_t)ix * N + iz) * h + iy;
                out[ir] = in[jr];
                if (iy != 0 && iy != No2) {
                    const size_t ii = ((size_t)ix * N + (N - iy)) * N + iz;
                    out[ii] = in[jr + dimag];
                }
            }
}
