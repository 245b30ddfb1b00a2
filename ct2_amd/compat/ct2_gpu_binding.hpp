// ct2_gpu_binding.hpp -- the reference-side binding: ct2's own free-function
// signatures (kick.hpp:10-24, drift.hpp:12, particle.hpp:22-58,
// particle.cpp:273) implemented over the C ABI of libct2hip.so.
//
// A ct2 build links this file instead of src/kick.cpp and src/drift.cpp.
// The particle globals stay host arrays, as ct2's accessors expect; each call
// uploads the arrays it reads and downloads the ones it writes.  Code that
// keeps particles resident on the device calls the ct2_* entry points directly.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <utility>
#include <vector>

#include "ct2hip.h"

// sfmm::fixed32 is a 32-bit unsigned fixed-point position (cosmictiger.hpp:22).
using ct2_pos_fixed = uint32_t;

struct kick_params {  // kick.hpp:10-20
    bool ascending;
    bool descending;
    bool top;
    bool first;
    int rung;
    float t0;
    float eta;
    float max_dt;
    float scale;
};

// particle.hpp:22-28
extern ct2_pos_fixed* particle_x_ptr;
extern ct2_pos_fixed* particle_y_ptr;
extern ct2_pos_fixed* particle_z_ptr;
extern float* particle_vx_ptr;
extern float* particle_vy_ptr;
extern float* particle_vz_ptr;
extern char* particle_rung_ptr;

void particle_allocate(size_t count);  // particle.cpp:273
size_t particle_count();               // particle.cpp:27
void particles_init();                 // particle.cpp:364-374 (seeded uniform positions, v = 0, rung 0)
int kick(const kick_params& params);   // kick.hpp:24
void drift(float dt, float scale);     // drift.hpp:12

// Rung bookkeeping of particle.hpp:57-66 (the primitives ct2's kick_step,
// main.cpp:8-92, calls); the sort permutes the particle arrays.
void particle_sort_by_rung(int minrung);
std::vector<size_t> particle_rung_counts();
std::pair<size_t, size_t> particle_current_range();
// particle_to_grid (particle.cpp:349-362) sorts the current range by cell on
// the device (ct2_particles_to_grid) and permutes the host arrays the same
// way; Ncell = cbrt(count / ppcell) + 4 as in particle.cpp:355.
void particle_to_grid();
void particle_pop_rungs();
void particle_push_rungs();
void particle_set_minrung(int minrung);
// The cell accessors of particle.cpp:290-318 over the last particle_to_grid.
// Neighbours are the cells (i + i0, j + j0, k + k0) mod Ncell with
// i0^2 + j0^2 + k0^2 <= bwidth^2 (the intent of particle.cpp:302-318, whose
// test uses absolute indices and pushes the sink cell).
std::vector<std::pair<size_t, size_t>> particle_cell_neighbors(int i, int j, int k);
int particle_cell_dim();
std::pair<size_t, size_t> particle_get_cell(int i, int j, int k);

// Options the force path reads from get_options() (options.cpp:53-61,155,190).
// The pair kernel defaults to ct2's own uniform-density sphere of radius hsoft
// (kick.cpp:52-60, CT2_KERNEL_CT2); CT2_KERNEL_PLUMMER selects Plummer softening.
void ct2_gpu_set_options(float hsoft, float GM, float eta = 0.15f, int kernel = CT2_KERNEL_CT2);
// ppcell and bwidth of particle_to_grid and particle_cell_neighbors (options.cpp:54,61).
void ct2_gpu_set_grid(int ppcell = 256, int bwidth = 3);
// kick() takes the direct sum by default; this switches it to PM + near field
// (ct2_pm_assignment, ct2_pm_init, ct2_pm_near); ng = 0 switches back.
void ct2_gpu_use_mesh(int ng, double rs_cells, double rcut_rs, int order);
// ... or to PM + the learned correction with the given weights (n =
// ct2_mlp_num_params(); trained for that split, assignment and, when
// particles_per_node > 0, that number of particles per grid node).  kick()
// then falls back to the direct sum whenever the mesh is not the one the
// weights were trained for (ct2_kick_pm_mlp).
void ct2_gpu_use_learned(const float* weights, size_t n, int ng, double rs_cells, int order,
                         double particles_per_node = 0.0);
// ... or to ct2's own structure (kick.cpp:105-125, particle.cpp:116-271):
// direct pairs over the cells with |offset|^2 <= 10 and the multipole mesh of
// the given order (<= 8, ct2's PORDER) on ncell^3 cells, ncell a power of two
// (about 8 particles per cell is fastest; ncell < 0 picks that grid for the
// allocated particles, ct2_mm_cells_for); ncell = 0 switches back to the
// direct sum.
void ct2_gpu_use_multipoles(int ncell, int order);

// kick_step (kick.hpp:23, main.cpp:8-92) over the functions above, for
// callers that do not link ct2's main.cpp (which runs unchanged on them).
// Returns its max_rung argument, as the reference does.
int ct2_gpu_kick_step(int& minrung, int max_rung, double scale, double adot, double tau, double t0, int minrung0,
                      bool nocrop);
// Optional trace of kick_step's calls (one line per kick / drift / rung call).
void ct2_gpu_trace(std::FILE* f);

// ---- the FFT API of fft.hpp:18-29 (fft_float is double, cosmictiger.hpp:18-23),
// with fft.hpp's parameter types, so that the mangled names are the ones ct2's
// callers reference (the inputs are read only).
using fft_float = double;
void fft_real(fft_float* X, int N, int dir = 1);                    // fft_real_scalar.cpp:9
void fft_r2c_3d(fft_float* xin, fft_float* xout, int N);            // fft_r2c_3d.cpp:4
void fft_c2r_3d(fft_float* xin, fft_float* xout, int N);            // fft_c2r_3d.cpp:4
void scramble(fft_float* X, int N1, int N2, int N3);                // scramble.cpp:26
void transpose_yz(fft_float* X, int N);                             // transpose.cpp:7
void transpose_r2c(fft_float* xin, fft_float* xout, int N);         // transpose.cpp:33
void transpose_c2r(fft_float* xin, fft_float* xout, int N);         // transpose.cpp:75
const std::vector<fft_float>& cos_twiddles(int N);                  // twiddles.cpp:6
const std::vector<fft_float>& sin_twiddles(int N);                  // twiddles.cpp:24
const std::vector<fft_float>& inv_sin_twiddles(int N);              // twiddles.cpp:42
// The 4-lane overloads fft_real(simd_f64*, N, dir) and
// fft_complex(complex<simd_f64>*, N, dir) take sfmm's SIMD types, which this
// file cannot name; a ct2 build forwards them to these (the layouts are the
// same: [n][lane] reals, {re[4], im[4]} complex elements):
void ct2_gpu_fft_real_lanes4(double* X, int N, int dir);
void ct2_gpu_fft_complex_lanes4(double* X, int N, int dir);
