/*
 * oracle.h -- CPU FP64 restatement of CosmicTiger 2's force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path
 * in ct2_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links or calls it.
 *
 * Every function names the reference file:line whose definition it
 * restates.  Where the reference is buggy, the restatement follows the
 * evident intent and says so.
 *
 * Conventions shared with the product:
 *   positions  uint32 fixed point, [0,1) <-> [0,2^32), periodic by wrap;
 *   distance   dx = (int32)(x_j - x_i) * 2^-32 (minimum image);
 *   force      F_i = GM * sum_j dx_ij * k(r) (points toward j);
 *   potential  stored positive (= -phi), self pair included.
 */
#ifndef CT2_ORACLE_H
#define CT2_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* softening kernels */
#define ORC_KERNEL_PLUMMER 0 /* (r^2+eps^2)^-1/2 potential            */
#define ORC_KERNEL_CT2 1     /* uniform-density sphere, kick.cpp:52-60 */

/* kick_params, kick.hpp:10-20 */
typedef struct {
    int ascending, descending, top, first;
    int rung;
    float t0, eta, max_dt, scale;
} orc_kick_params;

/* ---- kick (kick.cpp:36-74) ---- */
void orc_kick(size_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, size_t m,
              const int64_t* sinks, double gm, double soft, int kernel, double* ax, double* ay,
              double* az, double* pot);
/* long-double brute force of the same sum, for tiny n */
void orc_kick_ld(size_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, size_t m,
                 const int64_t* sinks, double gm, double soft, int kernel, double* ax,
                 double* ay, double* az, double* pot);

/* ---- kick epilogue (kick.cpp:75-99), FP32 like the reference ---- */
int orc_kick_epilogue(size_t n, const float* fx, const float* fy, const float* fz, float* vx,
                      float* vy, float* vz, int8_t* rung, const orc_kick_params* p, float hsoft);

/* ---- drift (drift.cpp:4-17) ---- */
uint32_t orc_fixed32_from_float(float s);
void orc_set_threads(int t); /* OpenMP threads of the oracle's loops */
int orc_max_threads(void);
void orc_drift(size_t n, uint32_t* x, uint32_t* y, uint32_t* z, const float* vx,
               const float* vy, const float* vz, float dt, float scale);

/* ---- FFT support (scramble.cpp, twiddles.cpp, transpose.cpp) ---- */
void orc_scramble(double* X, int N1, int N2, int N3);
void orc_twiddles(int N, double* cosw, double* sinw, double* invsinw);
void orc_transpose_yz(double* X, int N);
void orc_transpose_r2c(const double* in, double* out, int N);
void orc_transpose_c2r(const double* in, double* out, int N);

/* ---- 1-D FFTs (fft_complex_vector.cpp, fft_real_scalar.cpp, fft_real_vector.cpp) ----
 * input bit-reversed, output natural order; dir=+1 is e^{-2 pi i nk/N},
 * dir=-1 its conjugate; unnormalised.  lanes = 1 or 4 interleaved lanes. */
void orc_fft_complex(double* X, int N, int dir, int lanes);
void orc_fft_real(double* X, int N, int dir, int lanes);

/* ---- 3-D FFTs ---- */
void orc_fft3d(double* X, int n, int dir); /* interleaved complex n^3, natural order */
void orc_fft_r2c_3d(const double* in, double* out, int N); /* fft_r2c_3d.cpp intent  */
void orc_fft_c2r_3d(const double* in, double* out, int N); /* fft_c2r_3d.cpp intent  */

/* ---- particle mesh (particle.cpp:116-271 shape, CIC assignment) ---- */
void orc_cic_deposit(size_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, int g,
                     uint64_t* grid);
/* order 1 = cloud in cell, 2 = triangular-shaped cloud */
void orc_mesh_deposit(size_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, int g, int order,
                      uint64_t* grid);
void orc_pm_green_table(int ng, double gm, double rs_cells, int order, double* table);
void orc_pm_forces(size_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, int g,
                   double gm, double rs_cells, int order, double* fx, double* fy, double* fz, double* pot);
/* near-field complement of the Gaussian-filtered PM force (kick.cpp:105-125
 * structure: direct pairs within rcut, mesh beyond); rs, rcut in box units */
void orc_pm_near(size_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, size_t m,
                 const int64_t* sinks, double gm, double soft, int kernel, double rs, double rcut,
                 double* ax, double* ay, double* az, double* pot);
/* periodic (Ewald) direct sum with Plummer softening of the nearest image */
void orc_ewald_kick(size_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, size_t m,
                    const int64_t* sinks, double gm, double soft, double alpha, int nimg, int kmax,
                    double* ax, double* ay, double* az, double* pot);

/* ---- multipole mesh (particle.cpp:116-271, kick.cpp:11-29), Cartesian order P <= 8 ---- */
int orc_mm_ncomp(int P);
void orc_mm_index(int P, int* idx); /* [ncomp][3] exponents */
/* D_n(rho) = d^n of the periodic Green's function (cell units), minus the
 * nearest image 1/|rho| if near; rho anywhere, k-space summed directly */
void orc_mm_green_point(int P, int nc, double beta, int near, const double* rho, double* D);
/* the transformed table Dt[c][q] (Re or Im of D^_c by degree parity, / nc^3) */
void orc_mm_green_table(int P, int nc, double beta, int near2, double* Dt);
/* parts: 1 near-set direct pairs, 2 far field, 3 both */
void orc_mm_forces(size_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, int nc, int P, double beta,
                   int near2, double gm, double soft, int kernel, int parts, double* fx, double* fy, double* fz,
                   double* pot);

/* ---- learned short-range correction (no ct2 counterpart) ---- */
size_t orc_mlp_num_params(void);
double orc_bf16_round(double v);
void orc_gelu(size_t n, const double* z, double* g, double* dg);
void orc_mlp_forward(const double* w, size_t m, const double* X, int bf16, double* Y);
double orc_mlp_grad(const double* w, size_t m, const double* X, const double* T, int bf16, double* grad);

/* ---- cosmology (cosmology.cpp:11-102), explicit parameters ---- */
typedef struct {
    double H;         /* Hubble constant in code time units */
    double omega_m, omega_r, omega_lam;
} orc_cosmo;
void orc_cosmos_update(const orc_cosmo* c, double* adotdot, double* adot, double* a, double dt0);
double orc_cosmos_dadt(const orc_cosmo* c, double a);
double orc_cosmos_conformal_time(const orc_cosmo* c, double a0, double a1);
double orc_cosmos_time(const orc_cosmo* c, double a0, double a1);

/* ---- derived options and code units (options.cpp:88-95) ----
 * out: GM, code_to_g, code_to_cm, code_to_s, omega_r, omega_lam, H.  GM and
 * the omegas are rounded to FP32 as the reference's options block stores
 * them; the code units are kept in FP64 (in FP32 M_sun mass_res overflows at
 * the reference's default mass_res); H = H0 code_to_s hubble
 * (cosmology.cpp:15). */
void orc_options_derive(uint64_t ngrid, float mass_res, float hubble, float omega_m, float theta, double* out);

#ifdef __cplusplus
}
#endif
#endif
