/*
 * mm.c -- FP64 restatement of ct2's multipole-mesh far field
 * (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 * ct2's kick() takes the force on the particles of a cell from direct pairs
 * with the cells of its neighbour list and from a local expansion about the
 * cell centre (kick.cpp:11-29, 105-125).  The expansions come from a
 * convolution over the cell grid done with FFTs: multipoles of each cell
 * about its centre (particle_compute_multipoles, particle.cpp:180-209), one
 * transform per component, the product with the transformed table of
 * derivatives of the periodic Green's function (particle_greens,
 * particle.cpp:116-178: greens + greens_ewald, zeroed inside |offset|^2 <= 10
 * cells for the direct part), and inverse transforms (particle_compute_
 * expansions, particle.cpp:211-271).  The reference's operators P2M, MG2L and
 * L2P come from sfmm (spherical harmonics of order PORDER = 8,
 * cosmictiger.hpp:16), which is not in the tree; this file restates the same
 * pipeline with Cartesian Taylor expansions of total order P:
 *
 *   units      lengths in cells (h = 1/nc), unit masses; the box-unit force
 *              is GM nc^2 times the cell-unit one, the potential GM nc times;
 *   P2M        M_l(c) = sum_{j in c} (-b_j)^l / l!,  b_j = x_j - centre(c);
 *   table      D_n(rho) = d^n G_far(rho) at every cell offset rho, where
 *              G_far is the periodic Green's function (Ewald, neutralising
 *              background, the convention of orc_ewald_kick) minus 1/|rho|
 *              for offsets of the near set |rho|^2 <= near2 (those pairs are
 *              summed directly);
 *   MG2L       L_k(d) = sum_c sum_{|k|+|l| <= P} M_l(c) D_{k+l}(d - c), a cyclic
 *              convolution: per wave vector  L^_k = sum_l M^_l D^_{k+l};
 *   L2P        pot = sum_{|k|<=P} a^k/k! L_k,  F_x = sum_{|k|<=P-1} a^k/k! L_{k+e_x},
 *              a = x - centre(d)  (F = grad pot, toward the sources);
 *   near       the softened direct sum over the cells of the near set.
 *
 * The reference's loops test |offset|^2 on unwrapped indices (particle.cpp:
 * 136 uses j*j+k*k+l*l of 0..Ncell-1, particle.cpp:305 tests the wrapped
 * index i1 and pushes the sink cell itself); the restatement uses the evident
 * intent, the minimum-image cell offset, for both.
 *
 * The derivative table: G = G_short + G_long with Ewald parameter beta (per
 * cell).  G_short = sum_m erfc(beta r_m)/r_m, r_m = |rho + nc m|, is expanded
 * in real space: for a radial function f(s), s = |rho + h|^2, the Taylor
 * coefficients in h are sum_j f^(j)(s0)/j! [h^n] (2 rho.h + h.h)^j, with
 * f^(j)(s) = (-1/2)^j B_j(r) and the recurrence B_0 = erfc(beta r)/r,
 * B_j = ((2j-1) B_{j-1} + (2 beta^2)^j e^{-beta^2 r^2} / (beta sqrt(pi))) / r^2
 * (Smith, "Point multipoles in the Ewald summation", 1998); the nearest
 * image of a near offset uses -erf(beta r)/r from its power series in s.
 * G_long has Fourier coefficients e^{-pi^2 kappa^2/beta^2}/(pi kappa^2) per
 * cycle-per-cell wave vector kappa (and -pi/beta^2 at 0), so its sampled
 * table is transformed analytically, aliases included.  G_short's samples are
 * transformed with orc_fft3d.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define MM_PMAX 8
#define MM_NMAX 165 /* (PMAX+1)(PMAX+2)(PMAX+3)/6 */

typedef struct {
    int P, ncomp;
    int e[MM_NMAX][3];                     /* exponents of component c      */
    int id[MM_PMAX + 1][MM_PMAX + 1][MM_PMAX + 1]; /* component of (a,b,c), -1 */
    double fact[MM_NMAX];                  /* a! b! c!                      */
    int npair;                             /* products within degree P       */
    short (*pair)[3];                      /* (i, j, i+j)                   */
} MMIndex;

int orc_mm_ncomp(int P) { return (P + 1) * (P + 2) * (P + 3) / 6; }

/* Components by total degree, then by descending x exponent, then y. */
void orc_mm_index(int P, int* idx) {
    int c = 0;
    for (int d = 0; d <= P; d++)
        for (int a = d; a >= 0; a--)
            for (int b = d - a; b >= 0; b--) {
                idx[3 * c] = a;
                idx[3 * c + 1] = b;
                idx[3 * c + 2] = d - a - b;
                c++;
            }
}

static double factorial(int k) {
    double f = 1.0;
    for (int i = 2; i <= k; i++) f *= i;
    return f;
}

static int mm_index_init(MMIndex* I, int P) {
    if (P < 1 || P > MM_PMAX) return -1;
    I->P = P;
    I->ncomp = orc_mm_ncomp(P);
    int tmp[3 * MM_NMAX];
    orc_mm_index(P, tmp);
    memset(I->id, -1, sizeof I->id);
    for (int c = 0; c < I->ncomp; c++) {
        for (int q = 0; q < 3; q++) I->e[c][q] = tmp[3 * c + q];
        I->id[tmp[3 * c]][tmp[3 * c + 1]][tmp[3 * c + 2]] = c;
        I->fact[c] = factorial(tmp[3 * c]) * factorial(tmp[3 * c + 1]) * factorial(tmp[3 * c + 2]);
    }
    I->pair = malloc(sizeof(*I->pair) * I->ncomp * I->ncomp);
    I->npair = 0;
    for (int i = 0; i < I->ncomp; i++)
        for (int j = 0; j < I->ncomp; j++) {
            const int a = I->e[i][0] + I->e[j][0], b = I->e[i][1] + I->e[j][1], c = I->e[i][2] + I->e[j][2];
            if (a + b + c > P) continue;
            I->pair[I->npair][0] = (short)i;
            I->pair[I->npair][1] = (short)j;
            I->pair[I->npair][2] = (short)I->id[a][b][c];
            I->npair++;
        }
    return 0;
}

static void mm_index_free(MMIndex* I) { free(I->pair); }

static int degree(const MMIndex* I, int c) { return I->e[c][0] + I->e[c][1] + I->e[c][2]; }

/* C = A * B truncated at degree P (C may not alias A or B) */
static void jet_mul(const MMIndex* I, const double* A, const double* B, double* C) {
    memset(C, 0, sizeof(double) * I->ncomp);
    for (int p = 0; p < I->npair; p++) C[I->pair[p][2]] += A[I->pair[p][0]] * B[I->pair[p][1]];
}

/* Adds the Taylor coefficients T_n = d^n f / n! at rho of f(|rho + h|^2),
 * given fs[j] = f^(j)(s0), s0 = |rho|^2, j = 0..P. */
static void radial_taylor_add(const MMIndex* I, const double rho[3], const double* fs, double* T) {
    double delta[MM_NMAX], pw[MM_NMAX], nx[MM_NMAX];
    memset(delta, 0, sizeof delta);
    for (int q = 0; q < 3; q++) {
        int e1[3] = {0, 0, 0}, e2[3] = {0, 0, 0};
        e1[q] = 1;
        e2[q] = 2;
        delta[I->id[e1[0]][e1[1]][e1[2]]] = 2.0 * rho[q];
        if (I->P >= 2) delta[I->id[e2[0]][e2[1]][e2[2]]] = 1.0;
    }
    memset(pw, 0, sizeof pw);
    pw[0] = 1.0;
    T[0] += fs[0];
    double jf = 1.0;
    for (int j = 1; j <= I->P; j++) {
        jet_mul(I, pw, delta, nx);
        memcpy(pw, nx, sizeof(double) * I->ncomp);
        jf *= j;
        const double c = fs[j] / jf;
        for (int k = 0; k < I->ncomp; k++) T[k] += c * pw[k];
    }
}

/* s-derivatives of erfc(beta sqrt(s)) / sqrt(s) at s = r^2 */
static void erfc_sderiv(int P, double beta, double r, double* fs) {
    const double r2 = r * r, ex = exp(-beta * beta * r2) / (beta * sqrt(M_PI));
    double B = erfc(beta * r) / r, tb = 1.0;
    fs[0] = B;
    double sgn = 1.0;
    for (int j = 1; j <= P; j++) {
        tb *= 2.0 * beta * beta;
        B = ((2 * j - 1) * B + tb * ex) / r2;
        sgn *= -0.5;
        fs[j] = sgn * B;
    }
}

/* s-derivatives of -erf(beta sqrt(s)) / sqrt(s) from its power series:
 * -(2 beta / sqrt(pi)) sum_i (-beta^2 s)^i / (i! (2i+1)) */
static void erf_sderiv(int P, double beta, double s, double* fs) {
    const double b2 = beta * beta, c0 = -2.0 * beta / sqrt(M_PI);
    for (int j = 0; j <= P; j++) {
        /* d^j/ds^j: sum_{i>=j} (-b2)^i s^(i-j) / ((i-j)! (2i+1)) */
        double sum = 0.0, term = pow(-b2, j); /* (-b2)^i s^(i-j)/(i-j)! at i = j */
        for (int i = j; i < j + 400; i++) {
            const double t = term / (2 * i + 1);
            sum += t;
            if (fabs(t) < 1e-19 * fabs(sum) && i > j + 4) break;
            term *= -b2 * s / (double)(i - j + 1);
        }
        fs[j] = c0 * sum;
    }
}

/* Cutoff of the real-space part, in cells: the highest derivative of
 * erfc(beta r)/r is below 1e-16 of its size at r = 8/beta. */
static double mm_rcut(double beta) { return 8.0 / beta; }

/* Taylor coefficients of G_short(rho + h) minus, for near, the nearest
 * image's 1/|rho + h|; rho in cells. */
static void green_short_taylor(const MMIndex* I, int nc, double beta, int near, const double rho[3], double* T) {
    const double rc = mm_rcut(beta);
    const int mr = (int)ceil(rc / nc) + 1;
    double fs[MM_PMAX + 1];
    for (int m0 = -mr; m0 <= mr; m0++)
        for (int m1 = -mr; m1 <= mr; m1++)
            for (int m2 = -mr; m2 <= mr; m2++) {
                const double r[3] = {rho[0] + nc * m0, rho[1] + nc * m1, rho[2] + nc * m2};
                const double s = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                const int own = !m0 && !m1 && !m2;
                if (own && near) {
                    erf_sderiv(I->P, beta, s, fs);
                } else {
                    if (s >= rc * rc) continue;
                    erfc_sderiv(I->P, beta, sqrt(s), fs);
                }
                radial_taylor_add(I, r, fs, T);
            }
}

/* D_n(rho) at an arbitrary point (cells): real-space part plus the k-space
 * sum done directly (for checking the table). */
void orc_mm_green_point(int P, int nc, double beta, int near, const double* rho, double* D) {
    MMIndex I;
    if (mm_index_init(&I, P)) return;
    double T[MM_NMAX];
    memset(T, 0, sizeof T);
    green_short_taylor(&I, nc, beta, near, rho, T);
    for (int c = 0; c < I.ncomp; c++) D[c] = T[c] * I.fact[c];
    /* G_long = (1/(pi nc)) sum_{k != 0} e^{-pi^2 k^2 / (beta nc)^2} / k^2 cos(2 pi k.rho / nc) - pi / (beta^2 nc^3) */
    const int kmax = (int)ceil(2.6 * beta * nc) + 1;
    const double bn2 = beta * beta * (double)nc * nc, w = 2.0 * M_PI / nc;
    for (int k0 = -kmax; k0 <= kmax; k0++)
        for (int k1 = -kmax; k1 <= kmax; k1++)
            for (int k2 = -kmax; k2 <= kmax; k2++) {
                if (!k0 && !k1 && !k2) continue;
                const double kk = (double)(k0 * k0 + k1 * k1 + k2 * k2);
                const double g = exp(-M_PI * M_PI * kk / bn2) / (M_PI * nc * kk);
                if (g < 1e-300) continue;
                const double th = w * (k0 * rho[0] + k1 * rho[1] + k2 * rho[2]);
                const double cs = cos(th), sn = sin(th);
                const double kw[3] = {w * k0, w * k1, w * k2};
                for (int c = 0; c < I.ncomp; c++) {
                    const int a = I.e[c][0], b = I.e[c][1], d = I.e[c][2], deg = a + b + d;
                    const double mono = pow(kw[0], a) * pow(kw[1], b) * pow(kw[2], d);
                    /* d^deg/dth^deg cos(th) = cos(th + deg pi/2) */
                    double tr;
                    switch (deg & 3) {
                        case 0: tr = cs; break;
                        case 1: tr = -sn; break;
                        case 2: tr = -cs; break;
                        default: tr = sn; break;
                    }
                    D[c] += g * mono * tr;
                }
            }
    D[0] -= M_PI / (beta * beta * (double)nc * nc * nc);
    mm_index_free(&I);
}

static int wrap_offset(int i, int nc) { return i < nc / 2 ? i : i - nc; } /* [-nc/2, nc/2) */

/*
 * The transformed derivative table: Dt[c * nc^3 + q] for component c and
 * wave-vector index q = (q0 nc + q1) nc + q2 (FFT order).  D^_c is real for
 * even degree and imaginary for odd degree (D_c(-rho) = (-1)^deg D_c(rho));
 * Dt holds Re D^_c / nc^3 (even) or Im D^_c / nc^3 (odd), so that an
 * unnormalised inverse transform of sum_l M^_l D^_{k+l} gives L_k.
 */
void orc_mm_green_table(int P, int nc, double beta, int near2, double* Dt) {
    MMIndex I;
    if (mm_index_init(&I, P)) return;
    const size_t n3 = (size_t)nc * nc * nc;
    double* S = calloc(n3 * I.ncomp, sizeof(double)); /* short-range samples, by component */
#pragma omp parallel for schedule(dynamic, 8)
    for (ptrdiff_t cell = 0; cell < (ptrdiff_t)n3; cell++) {
        const int i0 = (int)(cell / ((size_t)nc * nc)), i1 = (int)(cell / nc % nc), i2 = (int)(cell % nc);
        const int o[3] = {wrap_offset(i0, nc), wrap_offset(i1, nc), wrap_offset(i2, nc)};
        const double rho[3] = {o[0], o[1], o[2]};
        const int near = o[0] * o[0] + o[1] * o[1] + o[2] * o[2] <= near2;
        double T[MM_NMAX];
        memset(T, 0, sizeof T);
        green_short_taylor(&I, nc, beta, near, rho, T);
        for (int c = 0; c < I.ncomp; c++) S[(size_t)c * n3 + cell] = T[c] * I.fact[c];
    }
    double* Z = malloc(sizeof(double) * 2 * n3);
    const double inv = 1.0 / (double)n3, b2 = beta * beta;
    for (int c = 0; c < I.ncomp; c++) {
        for (size_t q = 0; q < n3; q++) {
            Z[2 * q] = S[(size_t)c * n3 + q];
            Z[2 * q + 1] = 0.0;
        }
        orc_fft3d(Z, nc, 1);
        const int a = I.e[c][0], b = I.e[c][1], d = I.e[c][2], deg = a + b + d;
        for (size_t q = 0; q < n3; q++) {
            const int q0 = wrap_offset((int)(q / ((size_t)nc * nc)), nc), q1 = wrap_offset((int)(q / nc % nc), nc),
                      q2 = wrap_offset((int)(q % nc), nc);
            /* long range: sum over aliases of (2 pi kappa)^n e^{-pi^2 kappa^2/beta^2} / (pi kappa^2) (times i^deg) */
            double lr = 0.0;
            for (int m0 = -1; m0 <= 1; m0++)
                for (int m1 = -1; m1 <= 1; m1++)
                    for (int m2 = -1; m2 <= 1; m2++) {
                        const double k0 = (double)q0 / nc + m0, k1 = (double)q1 / nc + m1, k2 = (double)q2 / nc + m2;
                        const double kk = k0 * k0 + k1 * k1 + k2 * k2;
                        if (kk == 0.0) continue;
                        const double g = exp(-M_PI * M_PI * kk / b2) / (M_PI * kk);
                        lr += g * pow(2 * M_PI * k0, a) * pow(2 * M_PI * k1, b) * pow(2 * M_PI * k2, d);
                    }
            /* i^deg = (-1)^(deg/2) for even deg, i (-1)^((deg-1)/2) for odd */
            const double s = ((deg >> 1) & 1) ? -1.0 : 1.0;
            double v;
            if (deg & 1)
                v = Z[2 * q + 1] + s * lr;
            else
                v = Z[2 * q] + s * lr;
            if (deg == 0 && q == 0) v -= M_PI / b2;
            Dt[(size_t)c * n3 + q] = v * inv;
        }
    }
    free(Z);
    free(S);
    mm_index_free(&I);
}

/*
 * Force and potential of the multipole mesh at every particle (box units,
 * toward the sources, potential positive).  parts: 1 the direct pairs of the
 * near set only, 2 the far field only, 3 both.  nc >= 8 (a power of two), so
 * that the per-axis minimum image of a near pair is the image its cell
 * offset implies.
 */
void orc_mm_forces(size_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, int nc, int P, double beta,
                   int near2, double gm, double soft, int kernel, int parts, double* fx, double* fy, double* fz,
                   double* pot) {
    MMIndex I;
    if (mm_index_init(&I, P)) return;
    int lg = 0;
    while ((1 << lg) < nc) lg++;
    const size_t n3 = (size_t)nc * nc * nc;
    const int sh = 32 - lg;
    /* cells: counting sort */
    size_t* start = calloc(n3 + 1, sizeof(size_t));
    size_t* order = malloc(sizeof(size_t) * (n ? n : 1));
    size_t* cellof = malloc(sizeof(size_t) * (n ? n : 1));
    for (size_t i = 0; i < n; i++) {
        cellof[i] = ((size_t)(x[i] >> sh) * nc + (y[i] >> sh)) * nc + (z[i] >> sh);
        start[cellof[i] + 1]++;
    }
    for (size_t c = 0; c < n3; c++) start[c + 1] += start[c];
    {
        size_t* fill = malloc(sizeof(size_t) * n3);
        memcpy(fill, start, sizeof(size_t) * n3);
        for (size_t i = 0; i < n; i++) order[fill[cellof[i]]++] = i;
        free(fill);
    }
    /* in-cell offsets (cells) about the centre: x nc - (cell + 1/2) */
    const double cs = ldexp(1.0, -sh);
    for (size_t i = 0; i < n; i++) fx[i] = fy[i] = fz[i] = pot[i] = 0.0;

    if (parts & 2) {
        const int K = I.ncomp;
        double* Mre = calloc(n3 * K, sizeof(double));
        for (size_t i = 0; i < n; i++) {
            const uint32_t m = (uint32_t)((1u << sh) - 1u);
            const double b[3] = {(double)(x[i] & m) * cs - 0.5, (double)(y[i] & m) * cs - 0.5,
                                 (double)(z[i] & m) * cs - 0.5};
            for (int c = 0; c < K; c++)
                Mre[(size_t)c * n3 + cellof[i]] +=
                    pow(-b[0], I.e[c][0]) * pow(-b[1], I.e[c][1]) * pow(-b[2], I.e[c][2]) / I.fact[c];
        }
        double* Mh = malloc(sizeof(double) * 2 * n3 * K);
        for (int c = 0; c < K; c++) {
            double* Z = Mh + (size_t)c * 2 * n3;
            for (size_t q = 0; q < n3; q++) {
                Z[2 * q] = Mre[(size_t)c * n3 + q];
                Z[2 * q + 1] = 0.0;
            }
            orc_fft3d(Z, nc, 1);
        }
        free(Mre);
        double* Dt = malloc(sizeof(double) * n3 * K);
        orc_mm_green_table(P, nc, beta, near2, Dt);
        double* Lh = calloc(2 * n3 * K, sizeof(double));
        for (int p = 0; p < I.npair; p++) {
            const int k = I.pair[p][0], l = I.pair[p][1], kl = I.pair[p][2];
            const int odd = degree(&I, kl) & 1;
            const double* M = Mh + (size_t)l * 2 * n3;
            const double* D = Dt + (size_t)kl * n3;
            double* L = Lh + (size_t)k * 2 * n3;
            for (size_t q = 0; q < n3; q++) {
                const double mr = M[2 * q], mi = M[2 * q + 1], t = D[q];
                if (odd) { /* M * (i t) */
                    L[2 * q] -= mi * t;
                    L[2 * q + 1] += mr * t;
                } else {
                    L[2 * q] += mr * t;
                    L[2 * q + 1] += mi * t;
                }
            }
        }
        free(Mh);
        free(Dt);
        for (int c = 0; c < K; c++) orc_fft3d(Lh + (size_t)c * 2 * n3, nc, -1);
        for (size_t i = 0; i < n; i++) {
            const uint32_t m = (uint32_t)((1u << sh) - 1u);
            const double a[3] = {(double)(x[i] & m) * cs - 0.5, (double)(y[i] & m) * cs - 0.5,
                                 (double)(z[i] & m) * cs - 0.5};
            double ph = 0.0, f[3] = {0, 0, 0};
            for (int c = 0; c < K; c++) {
                const double mono = pow(a[0], I.e[c][0]) * pow(a[1], I.e[c][1]) * pow(a[2], I.e[c][2]) / I.fact[c];
                ph += mono * Lh[(size_t)c * 2 * n3 + 2 * cellof[i]];
                if (degree(&I, c) < P)
                    for (int q = 0; q < 3; q++) {
                        int e[3] = {I.e[c][0], I.e[c][1], I.e[c][2]};
                        e[q]++;
                        f[q] += mono * Lh[(size_t)I.id[e[0]][e[1]][e[2]] * 2 * n3 + 2 * cellof[i]];
                    }
            }
            fx[i] += gm * nc * (double)nc * f[0];
            fy[i] += gm * nc * (double)nc * f[1];
            fz[i] += gm * nc * (double)nc * f[2];
            pot[i] += gm * nc * ph;
        }
        free(Lh);
    }

    if (parts & 1) {
        const double e2 = soft * soft;
        const int br = (int)floor(sqrt((double)near2));
#pragma omp parallel for schedule(dynamic, 16)
        for (ptrdiff_t ii = 0; ii < (ptrdiff_t)n; ii++) {
            const size_t i = (size_t)ii;
            const int c0 = (int)(x[i] >> sh), c1 = (int)(y[i] >> sh), c2 = (int)(z[i] >> sh);
            double f[3] = {0, 0, 0}, p = 0.0;
            for (int d0 = -br; d0 <= br; d0++)
                for (int d1 = -br; d1 <= br; d1++)
                    for (int d2 = -br; d2 <= br; d2++) {
                        if (d0 * d0 + d1 * d1 + d2 * d2 > near2) continue;
                        const size_t cc = ((size_t)((c0 + d0 + nc) % nc) * nc + (size_t)((c1 + d1 + nc) % nc)) * nc +
                                          (size_t)((c2 + d2 + nc) % nc);
                        for (size_t t = start[cc]; t < start[cc + 1]; t++) {
                            const size_t j = order[t];
                            const double d[3] = {(double)(int32_t)(x[j] - x[i]) * 0x1p-32,
                                                 (double)(int32_t)(y[j] - y[i]) * 0x1p-32,
                                                 (double)(int32_t)(z[j] - z[i]) * 0x1p-32};
                            const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                            double k, ps;
                            if (kernel == ORC_KERNEL_PLUMMER) {
                                const double w = 1.0 / sqrt(r2 + e2);
                                k = w * w * w;
                                ps = w;
                            } else if (r2 < e2) {
                                k = 1.0 / (soft * e2);
                                ps = 1.5 / soft - 0.5 * r2 / (soft * e2);
                            } else {
                                const double w = 1.0 / sqrt(r2);
                                k = w * w * w;
                                ps = w;
                            }
                            for (int q = 0; q < 3; q++) f[q] += d[q] * k;
                            p += ps;
                        }
                    }
            fx[i] += gm * f[0];
            fy[i] += gm * f[1];
            fz[i] += gm * f[2];
            pot[i] += gm * p;
        }
    }
    free(start);
    free(order);
    free(cellof);
    mm_index_free(&I);
}
