/*
 * pm.c -- FP64 oracle for the particle-mesh force and the periodic (Ewald)
 * direct sum.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference's PM path (particle.cpp:116-271) convolves per-cell
 * multipoles with an FP64 periodic Green's function (near field excluded,
 * Ewald images added, particle.cpp:136-144) through 3-D FFTs.  The build
 * replaces its order-8 multipoles with cloud-in-cell mass assignment
 * (order 0) and keeps the same pipeline shape: deposit -> FFT -> multiply by
 * an FP64 Green's table -> inverse FFT -> interpolate at the particles.
 *
 * Mass-assignment definitions (shared bit for bit with the GPU), grid
 * ng = 2^g, node c at c/ng, s = 32 - g, 21-bit weights per axis summing to 2^21:
 *   CIC (order 1): cell i0 = x >> s, fraction t = x & (2^s - 1),
 *     a1 = t >> (s - 21) for node i0 + 1, a0 = 2^21 - a1 for node i0;
 *   TSC (order 2): nearest node i0 = (x + 2^(s-1)) >> s, signed offset
 *     D = ((x + 2^(s-1)) & (2^s - 1)) - 2^(s-1), shifted right (arithmetic) by
 *     s - 21 into [-2^20, 2^20); node i0 - 1 gets round((2^20 - D)^2 / 2^22),
 *     node i0 + 1 gets round((2^20 + D)^2 / 2^22) (the quadratic spline
 *     (1/2)(1/2 -+ d)^2), node i0 the rest of 2^21.
 *   A node gets W = (wx * wy * wz) >> 31, accumulated in uint64 (units of
 *   2^-32 particles).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define CIC_BITS 21

/* Per-axis weights of `order` (1 CIC, 2 TSC): nodes lo, lo+1[, lo+2]. */
static int assign_weights(uint32_t x, int g, int order, uint32_t* lo, uint64_t w[3]) {
    const int s = 32 - g;
    if (order == 1) {
        *lo = x >> s;
        const uint32_t t = x & (uint32_t)((1ull << s) - 1ull);
        w[1] = (uint64_t)(t >> (s - CIC_BITS));
        w[0] = (1ull << CIC_BITS) - w[1];
        return 2;
    }
    const uint32_t half = 1u << (s - 1);
    const uint32_t xs = x + half;
    const int32_t d = (int32_t)(xs & (uint32_t)((1ull << s) - 1ull)) - (int32_t)half;
    const int64_t D = (int64_t)(d >> (s - CIC_BITS));
    const uint64_t a = (uint64_t)((1 << 20) - D), b = (uint64_t)((1 << 20) + D);
    w[0] = (a * a + (1ull << 21)) >> 22;
    w[2] = (b * b + (1ull << 21)) >> 22;
    w[1] = (1ull << CIC_BITS) - w[0] - w[2];
    *lo = (xs >> s) - 1u;
    return 3;
}

void orc_mesh_deposit(size_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, int g, int order,
                      uint64_t* grid) {
    const uint32_t ng = 1u << g, mask = ng - 1;
    memset(grid, 0, sizeof(uint64_t) * (size_t)ng * ng * ng);
    for (size_t p = 0; p < n; p++) {
        uint32_t i0, j0, k0;
        uint64_t ax[3], ay[3], az[3];
        const int nw = assign_weights(x[p], g, order, &i0, ax);
        assign_weights(y[p], g, order, &j0, ay);
        assign_weights(z[p], g, order, &k0, az);
        for (int dx = 0; dx < nw; dx++)
            for (int dy = 0; dy < nw; dy++)
                for (int dz = 0; dz < nw; dz++) {
                    const size_t c = ((size_t)((i0 + dx) & mask) * ng + ((j0 + dy) & mask)) * ng +
                                     ((k0 + dz) & mask);
                    grid[c] += (ax[dx] * ay[dy] * az[dz]) >> 31;
                }
    }
}

void orc_cic_deposit(size_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, int g,
                     uint64_t* grid) {
    orc_mesh_deposit(n, x, y, z, g, 1, grid);
}

/* Green's function of the PM force, FP64: for k = 2 pi m != 0,
 *   G(m) = -4 pi GM exp(-k^2 rs^2) / (k^2 W(m)^2),  W = prod_d sinc^(order+1)(pi m_d / ng),
 * rs = rs_cells / ng.  W is the CIC (TSC) window, deconvolved once for the deposit
 * and once for the interpolation; the Gaussian keeps the long-range part of
 * the force (the split of TreePM/P3M codes) so that the grid-scale modes that
 * the deconvolution amplifies are cut off.  G(0) = 0 (mean density removed:
 * the periodic, Ewald-consistent solution).
 * Tabulated on the octant |m_d| in [0, ng/2]: table[(a*h + b)*h + c], h = ng/2+1. */
void orc_pm_green_table(int ng, double gm, double rs_cells, int order, double* table) {
    const double rs = rs_cells / ng;
    const int h = ng / 2 + 1;
    for (int a = 0; a < h; a++)
        for (int b = 0; b < h; b++)
            for (int c = 0; c < h; c++) {
                double v = 0.0;
                if (a || b || c) {
                    const int m[3] = {a, b, c};
                    double w = 1.0, k2 = 0.0;
                    for (int d = 0; d < 3; d++) {
                        const double u = M_PI * m[d] / ng;
                        const double sc = m[d] ? sin(u) / u : 1.0;
                        w *= order == 1 ? sc * sc : sc * sc * sc;
                        k2 += (2.0 * M_PI * m[d]) * (2.0 * M_PI * m[d]);
                    }
                    v = -4.0 * M_PI * gm * exp(-k2 * rs * rs) / (k2 * w * w);
                }
                table[((size_t)a * h + b) * h + c] = v;
            }
}

/* Full FP64 PM force and potential at the particles:
 * deposit -> forward FFT -> phi = G rho -> F = -i k phi (Nyquist component of
 * the gradient zeroed), pot = -phi -> inverse FFTs -> interpolation with
 * the same weights as the deposit (order 1 CIC, 2 TSC). */
void orc_pm_forces(size_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, int g,
                   double gm, double rs_cells, int order, double* fx, double* fy, double* fz, double* pot) {
    const int ng = 1 << g, h = ng / 2 + 1;
    const size_t n3 = (size_t)ng * ng * ng;
    uint64_t* W = (uint64_t*)malloc(sizeof(uint64_t) * n3);
    orc_mesh_deposit(n, x, y, z, g, order, W);
    double* rho = (double*)malloc(sizeof(double) * 2 * n3);
    for (size_t c = 0; c < n3; c++) {
        rho[2 * c] = (double)W[c] * 0x1p-32;
        rho[2 * c + 1] = 0.0;
    }
    free(W);
    orc_fft3d(rho, ng, 1);
    double* G = (double*)malloc(sizeof(double) * (size_t)h * h * h);
    orc_pm_green_table(ng, gm, rs_cells, order, G);
    double* out[4];
    for (int q = 0; q < 4; q++) out[q] = (double*)malloc(sizeof(double) * 2 * n3);
    for (int i = 0; i < ng; i++)
        for (int j = 0; j < ng; j++)
            for (int l = 0; l < ng; l++) {
                const size_t c = ((size_t)i * ng + j) * ng + l;
                const int mi = i < ng / 2 ? i : i - ng, mj = j < ng / 2 ? j : j - ng, ml = l < ng / 2 ? l : l - ng;
                const int ai = abs(mi), aj = abs(mj), al = abs(ml);
                const double gv = G[((size_t)ai * h + aj) * h + al];
                const double pr = gv * rho[2 * c], pi = gv * rho[2 * c + 1]; /* phi_hat */
                const double k[3] = {i == ng / 2 ? 0.0 : 2.0 * M_PI * mi, j == ng / 2 ? 0.0 : 2.0 * M_PI * mj,
                                     l == ng / 2 ? 0.0 : 2.0 * M_PI * ml};
                for (int d = 0; d < 3; d++) { /* F_hat = -i k phi_hat */
                    out[d][2 * c] = k[d] * pi;
                    out[d][2 * c + 1] = -k[d] * pr;
                }
                out[3][2 * c] = -pr; /* pot = -phi */
                out[3][2 * c + 1] = -pi;
            }
    free(rho);
    free(G);
    for (int q = 0; q < 4; q++) orc_fft3d(out[q], ng, -1);
    const uint32_t mask = (uint32_t)ng - 1;
    double* dst[4] = {fx, fy, fz, pot};
    for (size_t p = 0; p < n; p++) {
        uint32_t i0, j0, k0;
        uint64_t ax[3], ay[3], az[3];
        const int nw = assign_weights(x[p], g, order, &i0, ax);
        assign_weights(y[p], g, order, &j0, ay);
        assign_weights(z[p], g, order, &k0, az);
        double acc[4] = {0, 0, 0, 0};
        for (int dx = 0; dx < nw; dx++)
            for (int dy = 0; dy < nw; dy++)
                for (int dz = 0; dz < nw; dz++) {
                    const size_t c = ((size_t)((i0 + dx) & mask) * ng + ((j0 + dy) & mask)) * ng +
                                     ((k0 + dz) & mask);
                    const double w = (double)ax[dx] * (double)ay[dy] * (double)az[dz] * 0x1p-63;
                    for (int q = 0; q < 4; q++) acc[q] += w * out[q][2 * c];
                }
        for (int q = 0; q < 4; q++) dst[q][p] = acc[q];
    }
    for (int q = 0; q < 4; q++) free(out[q]);
}

/*
 * Periodic direct sum by Ewald summation, FP64, with explicit images
 * (Hernquist, Bouchet & Suto 1991, eqs. 2.14-2.16), unit box, G = 1,
 * uniform neutralising background:
 *   K(d) = sum_n (d+n)/s^3 [erfc(a s) + 2 a s/sqrt(pi) exp(-a^2 s^2)],  s = |d+n|
 *        + 2 sum_{k != 0} k/k^2 exp(-pi^2 k^2/a^2) sin(2 pi k.d)
 *   psi(d) = sum_n erfc(a s)/s + (1/pi) sum_{k != 0} exp(-pi^2 k^2/a^2)/k^2 cos(2 pi k.d) - pi/a^2
 * d = x_j - x_i is the minimum image; its own n = 0 term uses Plummer
 * softening: the Newtonian d/s^3 and 1/s are replaced by d (s^2+eps^2)^-3/2 and
 * (s^2+eps^2)^-1/2.  The self pair contributes psi's regular part at 0 plus
 * 1/eps to the potential.  Real-space images |n_d| <= nimg, k-space |k_d| <= kmax.
 * This is the periodic force ct2's Green's function intends
 * (particle.cpp:141-144: greens + greens_ewald).
 */
void orc_ewald_kick(size_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, size_t m,
                    const int64_t* sinks, double gm, double soft, double alpha, int nimg, int kmax,
                    double* ax, double* ay, double* az, double* pot) {
    const double e2 = soft * soft;
    const double a2 = alpha * alpha, tsp = 2.0 * alpha / sqrt(M_PI);
#pragma omp parallel for schedule(dynamic, 1)
    for (ptrdiff_t s = 0; s < (ptrdiff_t)m; s++) {
        const size_t i = sinks ? (size_t)sinks[s] : (size_t)s;
        double f[3] = {0, 0, 0}, p = 0.0;
        for (size_t j = 0; j < n; j++) {
            const double d[3] = {(double)(int32_t)(x[j] - x[i]) * 0x1p-32, (double)(int32_t)(y[j] - y[i]) * 0x1p-32,
                                 (double)(int32_t)(z[j] - z[i]) * 0x1p-32};
            /* real space */
            for (int n0 = -nimg; n0 <= nimg; n0++)
                for (int n1 = -nimg; n1 <= nimg; n1++)
                    for (int n2 = -nimg; n2 <= nimg; n2++) {
                        const double r[3] = {d[0] + n0, d[1] + n1, d[2] + n2};
                        const double s2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                        const int own = (n0 == 0 && n1 == 0 && n2 == 0);
                        if (own && j == i) {
                            p += 1.0 / soft - tsp; /* softened self term + regular limit of -erf(a s)/s */
                            continue;
                        }
                        const double sr = sqrt(s2);
                        const double ec = erfc(alpha * sr), ex = exp(-a2 * s2);
                        double kf = (ec + tsp * sr * ex) / (s2 * sr);
                        double kp = ec / sr;
                        if (own) { /* swap the Newtonian part for Plummer */
                            const double w = 1.0 / sqrt(s2 + e2);
                            kf += w * w * w - 1.0 / (s2 * sr);
                            kp += w - 1.0 / sr;
                        }
                        for (int q = 0; q < 3; q++) f[q] += r[q] * kf;
                        p += kp;
                    }
            /* reciprocal space */
            for (int k0 = -kmax; k0 <= kmax; k0++)
                for (int k1 = -kmax; k1 <= kmax; k1++)
                    for (int k2 = -kmax; k2 <= kmax; k2++) {
                        if (!k0 && !k1 && !k2) continue;
                        const double kk = (double)(k0 * k0 + k1 * k1 + k2 * k2);
                        const double gk = exp(-M_PI * M_PI * kk / a2) / kk;
                        const double ph = 2.0 * M_PI * (k0 * d[0] + k1 * d[1] + k2 * d[2]);
                        const double sn = sin(ph), cs = cos(ph);
                        f[0] += 2.0 * gk * k0 * sn;
                        f[1] += 2.0 * gk * k1 * sn;
                        f[2] += 2.0 * gk * k2 * sn;
                        p += gk * cs / M_PI;
                    }
            p -= M_PI / a2;
        }
        ax[s] = gm * f[0];
        ay[s] = gm * f[1];
        az[s] = gm * f[2];
        pot[s] = gm * p;
    }
}

/*
 * Near-field complement of the PM force, FP64 brute force.
 *
 * ct2's kick() sums the near field directly over each cell's neighbour list
 * and takes the far field from the mesh (kick.cpp:105-125: neighbour list at
 * :117, expansion at :118, direct pairs in kick_cell :36-67).  With the
 * Gaussian-filtered PM force above, the mesh already holds, for every pair, the
 * force and potential of the filtered Newtonian kernel
 *     phi_long(r) = erf(u) / r,   u = r / (2 rs),
 *     F_long(r)   = dx [erf(u) - (2u/sqrt(pi)) exp(-u^2)] / r^3,
 * so the near field adds, for each minimum-image pair with r < rcut,
 * the softened kernel (Plummer, or ct2's kick.cpp:52-60) minus that:
 *     F_i += GM dx (k_soft(r) - [erf(u) - (2u/sqrt(pi)) e^{-u^2}] / r^3)
 *     pot_i += GM (p_soft(r) - erf(u) / r)
 * The self pair adds GM (p_soft(0) - 1/(rs sqrt(pi))) to the potential.  The
 * PM + near-field potential then differs from the periodic (Ewald) one by the
 * integral of erfc(u)/r over the box, GM N 4 pi rs^2 per particle, which is
 * subtracted here, so that PM + near field converges to orc_ewald_kick as
 * rcut grows and the mesh is refined.  rs and rcut are in box units.
 */
void orc_pm_near(size_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, size_t m,
                 const int64_t* sinks, double gm, double soft, int kernel, double rs, double rcut,
                 double* ax, double* ay, double* az, double* pot) {
    const double e2 = soft * soft, rc2 = rcut * rcut;
    const double c2 = 2.0 / sqrt(M_PI), inv2rs = 0.5 / rs;
    const double background = 4.0 * M_PI * rs * rs * (double)n;
#pragma omp parallel for schedule(dynamic, 4)
    for (ptrdiff_t s = 0; s < (ptrdiff_t)m; s++) {
        const size_t i = sinks ? (size_t)sinks[s] : (size_t)s;
        double f[3] = {0.0, 0.0, 0.0}, p = 0.0;
        for (size_t j = 0; j < n; j++) {
            const double d[3] = {(double)(int32_t)(x[j] - x[i]) * 0x1p-32, (double)(int32_t)(y[j] - y[i]) * 0x1p-32,
                                 (double)(int32_t)(z[j] - z[i]) * 0x1p-32};
            const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (r2 >= rc2) continue;
            double ks, ps;
            if (kernel == ORC_KERNEL_PLUMMER) {
                const double w = 1.0 / sqrt(r2 + e2);
                ks = w * w * w;
                ps = w;
            } else if (r2 < e2) {
                ks = 1.0 / (soft * e2);
                ps = 1.5 / soft - 0.5 * r2 / (soft * e2);
            } else {
                const double w = 1.0 / sqrt(r2);
                ks = w * w * w;
                ps = w;
            }
            const double r = sqrt(r2), u = r * inv2rs;
            double kl, pl;
            if (u < 1e-3) { /* leading terms of the series; r -> 0 limits */
                kl = c2 * (2.0 / 3.0 - 0.4 * u * u) * inv2rs * inv2rs * inv2rs;
                pl = c2 * (1.0 - u * u / 3.0) * inv2rs;
            } else {
                kl = (erf(u) - c2 * u * exp(-u * u)) / (r2 * r);
                pl = erf(u) / r;
            }
            for (int q = 0; q < 3; q++) f[q] += d[q] * (ks - kl);
            p += ps - pl;
        }
        ax[s] = gm * f[0];
        ay[s] = gm * f[1];
        az[s] = gm * f[2];
        pot[s] = gm * (p - background);
    }
}
