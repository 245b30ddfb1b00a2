/*
 * ct2hip.h -- C ABI of libct2hip.so, the MI355X (gfx950) force path of
 * CosmicTiger 2: direct-sum kick, drift, 3-D FFT and the particle-mesh
 * force, on ct2's structure-of-arrays particle container.
 *
 * Conventions
 *   - Every call returns an int status (CT2_OK == 0); ct2_last_error()
 *     describes the last failure.  No C++ exception crosses this boundary.
 *   - The library owns the device buffers it allocates (particles, forces,
 *     grids); the caller owns every host pointer it passes.
 *   - One device and one HIP stream per process (ct2_init).  Calls are
 *     thread-compatible, not re-entrant: serialise them.
 *   - Positions are ct2 fixed32 (sfmm::fixed32): [0,1) <-> [0,2^32), periodic
 *     by integer wrap; distances are the wrapped difference read as int32.
 *
 * Each declaration names the reference interface it replaces (paths are
 * relative to the ct2 source tree).
 */
#ifndef CT2HIP_H
#define CT2HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status */
#define CT2_OK 0
#define CT2_ERR_INVALID 1  /* bad argument: size, null pointer, not a power of two */
#define CT2_ERR_STATE 2    /* wrong call order: not initialised, no particles */
#define CT2_ERR_NOMEM 3    /* device allocation failed */
#define CT2_ERR_HIP 4      /* HIP runtime error */
#define CT2_ERR_NODEVICE 5 /* no usable gfx950 device */

/* Text of the last error (static storage, overwritten by the next error). */
const char* ct2_last_error(void);

/* ------------------------------------------------------------ context */
/* Replaces the process set-up of src/main.cpp:94-102 (hpx_main): binds the
 * device, creates the library stream.  Idempotent for the same device. */
int ct2_init(int device);
/* Frees every device buffer and the stream. */
int ct2_finalize(void);
/* The library's HIP stream (hipStream_t) for callers that time or order work. */
int ct2_get_stream(void** stream);
int ct2_synchronize(void);

/* ------------------------------------------------------------ options */
#define CT2_KERNEL_PLUMMER 0 /* F = GM dx (r^2+h^2)^-3/2 (default) */
#define CT2_KERNEL_CT2 1     /* uniform-density sphere of radius h, kick.cpp:52-60 */

/* The options the force path reads from get_options() (options.hpp:14-35):
 * hsoft (options.cpp:155), GM = 1/Ngrid^3 (options.cpp:190). */
typedef struct {
    float hsoft;
    float GM;
    int kernel;
} ct2_force_options;
int ct2_set_force_options(const ct2_force_options* opts);

/* ------------------------------------------------------------ particles */
/* Replaces particle_allocate(size_t) (particle.cpp:273-288): device SoA
 * arrays x,y,z (uint32 fixed32), vx,vy,vz (float), rung (int8).  Contents
 * are zeroed.  Re-allocating discards the previous set. */
int ct2_particles_alloc(size_t n);
/* Replaces particle_count() (particle.hpp:60, particle.cpp:27-29). */
size_t ct2_particle_count(void);
/* Host <-> device copies of the container (the accessors of
 * particle.hpp:30-56 operate on host copies).  A NULL pointer skips that
 * array. */
int ct2_particles_upload(const uint32_t* x, const uint32_t* y, const uint32_t* z, const float* vx,
                         const float* vy, const float* vz, const int8_t* rung);
int ct2_particles_download(uint32_t* x, uint32_t* y, uint32_t* z, float* vx, float* vy,
                           float* vz, int8_t* rung);

/* Page-locked host memory for the host side of those copies (no counterpart
 * in the reference, whose arrays are ordinary heap blocks): copies from and
 * into it run at the host link's rate, without the driver's staging of
 * pageable memory.  ct2_host_alloc sets *out (NULL for 0 bytes) and returns
 * CT2_ERR_NOMEM when the allocation fails; ct2_host_free takes NULL. */
int ct2_host_alloc(size_t bytes, void** out);
int ct2_host_free(void* p);

/* Device pointers of the container (particle_x_ptr ... particle_rung_ptr,
 * particle.hpp:22-28) for callers with their own HIP code. */
typedef struct {
    uint32_t *x, *y, *z;
    float *vx, *vy, *vz;
    int8_t* rung;
    size_t n;
} ct2_particle_ptrs;
int ct2_particles_device_ptrs(ct2_particle_ptrs* out);

/* ------------------------------------------------------------ kick */
/* kick_params, kick.hpp:10-20 (bools as int). */
typedef struct {
    int ascending;
    int descending;
    int top;
    int first;
    int rung;
    float t0;
    float eta;
    float max_dt;
    float scale;
} ct2_kick_params;

#define CT2_KICK_NAIVE 0x1 /* plain FP32 accumulation instead of compensated */
#define CT2_KICK_STORE 0x2 /* keep per-particle force and potential for ct2_forces_download */
#define CT2_KICK_ORDERED 0x4 /* evaluate every ordered pair (no use of Newton's third law) */
#define CT2_KICK_SYMMETRIC 0x8 /* use Newton's third law on any core of >= 1024 sinks (default: >= 2^16) */

/* Replaces int kick(const kick_params&) (kick.hpp:24, kick.cpp:105-125):
 * all-pairs FP32 direct sum over the whole container with compensated
 * accumulation, then the half-kick and rung epilogue of kick.cpp:75-99.
 * *max_rung receives the return value of kick().  The sinks are the current
 * rung range [i0, n).  The last m = floor((n - i0) / 1024) * 1024 sinks form
 * a core whose unordered pairs are each evaluated once and applied to both
 * particles (Newton's third law), when m >= 2^16 (64 superblocks, below
 * which the ordered kernel is faster) or CT2_KICK_SYMMETRIC is set, and neither CT2_KICK_NAIVE
 * nor CT2_KICK_ORDERED is; the core's pairs with the particles before it and
 * the remaining sinks are summed pair by pair.  Without a core every sink
 * sums over every source.  All paths are deterministic. */
int ct2_kick(const ct2_kick_params* params, unsigned flags, int* max_rung);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same kick as ct2_kick with CT2_KICK_ORDERED (every
 * sink sums over every source, same kernels, the same epilogue and return
 * value) whose partial sums and max-rung counter live in a static 8 MiB
 * device array instead of allocated buffers; it calls no allocator.  At most
 * CT2_NOALLOC_MAX_SINKS active sinks (the rung range [i0, n); sources are all
 * n particles), else CT2_ERR_INVALID and nothing runs.  Flags: CT2_KICK_NAIVE,
 * CT2_KICK_STORE, CT2_KICK_ORDERED.  Each GPU holds one copy of the array for
 * the life of the process and every call on that GPU shares it: calls that
 * overlap on one GPU (other threads, other streams) race.  Results equal
 * ct2_kick(CT2_KICK_ORDERED) bit for bit wherever ct2_kick's partial sums of
 * the same sinks fit in 8 MiB, and agree to FP32 rounding otherwise. */
#define CT2_NOALLOC_MAX_SINKS (1 << 18)
int ct2_kick_noalloc(const ct2_kick_params* params, unsigned flags, int* max_rung);
/* The force (GM-scaled, toward the sources) and positive potential of the
 * last ct2_kick made with CT2_KICK_STORE.  NULL skips an array. */
int ct2_forces_download(float* fx, float* fy, float* fz, float* pot);

/* FP64 direct sum for m sampled sinks (indices into the container), the
 * judge of every FP32 force: same kernel and options as ct2_kick.
 * periodic != 0 sums every periodic image of every source with a uniform
 * neutralising background (FP64 Ewald summation, Plummer softening of the
 * nearest image), the force of particle.cpp:141-144 (greens + greens_ewald). */
int ct2_kick_sample_fp64(const int64_t* idx, size_t m, int periodic, double* ax, double* ay,
                         double* az, double* pot);

/* ------------------------------------------------------------ particle mesh */
/* Particle-mesh far field, replacing the multipole-mesh path of
 * particle_compute_multipoles / particle_compute_expansions / particle_greens
 * (particle.cpp:116-271) and its evaluation by L2P (kick.cpp:29).
 * ng^3 grid (power of two in [8, 1024]); cloud-in-cell assignment with
 * fixed-point weights (deterministic); FP32 FFTs; FP64 Green's table
 *   G(k) = -4 pi GM exp(-k^2 rs^2) / (k^2 W_cic(k)^2),  rs = rs_cells/ng,
 * i.e. the long-range part of the periodic force for rs_cells > 0. */
int ct2_pm_init(int ng, double rs_cells);
/* PM force and potential of every particle into the stored-force arrays
 * (ct2_forces_download). */
int ct2_pm_forces(void);
/* Near field of the PM force.  ct2's kick() sums the near field directly over
 * each cell's neighbour list and takes the far field from the mesh
 * (kick.cpp:105-125: neighbour list :117, expansion :118, kick_cell :36-67).
 * With rcut_rs > 0 every PM force evaluation here (ct2_pm_forces,
 * ct2_kick_pm, ct2_pm_mlp_forces, ct2_kick_pm_mlp, ct2_mlp_data_append) adds,
 * for each minimum-image pair closer than rcut = rcut_rs * rs (rs =
 * rs_cells / ng of ct2_pm_init), the softened kernel minus the mesh's
 * Gaussian-filtered Newtonian kernel, and removes the potential's background
 * term GM N 4 pi rs^2, so that the result tends to the periodic direct sum.
 * Particles are binned into cells of side >= rcut (at most 64 per side) and
 * ordered by (cell, index): the sum is deterministic.  rcut_rs = 0 (the
 * default) switches it off; the setting persists across ct2_pm_init.  rcut
 * may not exceed a third of the box; rs must be > 0. */
int ct2_pm_near(double rcut_rs);
/* Cells per side of the last near-field evaluation (0 before the first). */
int ct2_pm_near_cells(void);
/* The near-field kernel's filtered (long-range) part as the GPU evaluates it,
 * for a split radius rs_u and cutoff rcut_u in 2^-32 position units: nt
 * cubic pieces in r^2 over [0, rcut_u^2) of
 *   k_l = [erf(u) - (2u/sqrt(pi)) e^{-u^2}] / (u^3 (2 rs)^3),
 *   p_l = erf(u) / (u (2 rs)),   u = r / (2 rs),
 * piece i evaluated at f = r^2 nt / rcut_u^2 - i as c0 + f (c1 + f (c2 + f c3)),
 * coefficients (c0..c3) of piece i at k4[4i..4i+3] and p4[4i..4i+3].  The
 * pair kernel subtracts these from the softened Newtonian kernel.  Host only,
 * no device state.  No counterpart in the reference (its near field is the
 * plain softened kernel with the mesh's far field left out). */
int ct2_near_kernel_table(double rs_u, double rcut_u, int nt, float* k4, float* p4);
/* Mass assignment of the PM force: 1 = cloud in cell (default), 2 =
 * triangular-shaped cloud (27 nodes, quadratic spline; the Green's table
 * deconvolves sinc^6 instead of sinc^4).  TSC lowers the mesh's
 * discretisation error about tenfold at a given split radius.  Kept across
 * ct2_pm_init.  (ct2's own mesh assigns order-8 multipoles, particle.cpp:180-209.) */
int ct2_pm_assignment(int order);
/* kick() with the PM force: forces as ct2_pm_forces, then the epilogue of
 * kick.cpp:75-99. */
int ct2_kick_pm(const ct2_kick_params* params, unsigned flags, int* max_rung);
/* Grids of the last PM evaluation: the CIC density (uint64, 2^-32 particle
 * units, ng^3) and the real-space fields (which = 0: Fx + i Fy, 1: Fz + i pot;
 * interleaved float pairs, ng^3). */
int ct2_pm_density_download(uint64_t* host);
int ct2_pm_grid_download(int which, float* host);

/* ------------------------------------------------------------ learned correction */
/* No ct2 counterpart (the reference's far field is L2P, kick.cpp:29).  A
 * network 111 -> 256 -> 256 -> 3 (GELU, BF16 weights, FP32 accumulation on
 * MFMA) maps local PM-grid features (27 nodes around the particle's nearest
 * node x (density/mean, PM force/F_unit) + 3 offsets, F_unit = GM ng^2) to the
 * force the PM misses.  Flat FP32 weight layout (padded):
 *   W1[128][256] b1[256] W2[256][256] b2[256] W3[256][32] b3[32]. */
size_t ct2_mlp_num_params(void);
int ct2_mlp_init(uint64_t seed);
int ct2_mlp_get_weights(float* w, size_t n);
int ct2_mlp_set_weights(const float* w, size_t n);
/* GPU forward of m given feature rows (x[m][128] -> y[m][3]) and the loss
 * L = sum_r t[r][3] |y_r - t[r][0..2]|^2 / m with its gradient (t is m x 4,
 * m a multiple of 128): the kernels the training and the inference use,
 * exposed for verification.  Training sets weight each sample by
 * 1/|F|^2 (relative force error). */
int ct2_mlp_forward(const float* x, size_t m, float* y);
int ct2_mlp_grad(const float* x, const float* t, size_t m, float* grad, float* loss);
/* Features of sampled particles from the last PM evaluation ([m][128]). */
int ct2_mlp_features(const int64_t* idx, size_t m, float* out);
/* Training set on the device: ct2_mlp_data_append runs the PM force and the
 * periodic FP64 judge for the sampled particles and stores their features
 * and targets (judge - PM) / F_unit. */
int ct2_mlp_data_reset(size_t capacity);
size_t ct2_mlp_data_size(void);
int ct2_mlp_data_append(const int64_t* idx, size_t m);
/* `steps` Adam steps on random batches; losses[steps] optional. */
int ct2_mlp_train(int steps, int batch, float lr, uint64_t seed, float* losses);
/* The mesh the current weights were trained to correct: split radius in
 * cells, assignment order (1 CIC, 2 TSC), near-field cutoff in split radii
 * (0: none) and particles per grid node (0: not checked).  The grid side
 * itself may differ (every feature is in grid units).  ct2_mlp_train
 * records the mesh of its training set itself; ct2_mlp_set_weights forgets
 * it until this call. */
int ct2_mlp_set_mesh(double rs_cells, int order, double rcut_rs, double particles_per_node);
/* PM force plus the learned correction into the stored-force arrays, and the
 * kick() built on it (epilogue of kick.cpp:75-99).  ct2_kick_pm_mlp falls
 * back to the direct-sum kick (ct2_kick with the same parameters and flags)
 * when the correction is not usable: no trained weights (only
 * ct2_mlp_init), no PM mesh, or a mesh other than the one recorded for the
 * weights; ct2_last_force_path says which ran. */
int ct2_pm_mlp_forces(void);
int ct2_kick_pm_mlp(const ct2_kick_params* params, unsigned flags, int* max_rung);

/* The force path the last kick took. */
#define CT2_PATH_NONE 0
#define CT2_PATH_DIRECT 1  /* all-pairs direct sum (ct2_kick) */
#define CT2_PATH_PM 2      /* PM alone */
#define CT2_PATH_PM_NEAR 3 /* PM + near field */
#define CT2_PATH_PM_MLP 4  /* PM + learned correction */
#define CT2_PATH_MM 5      /* multipole mesh + near-set direct pairs */
int ct2_last_force_path(void);

/* ------------------------------------------------------------ rungs */
/* ct2's rung bookkeeping (particle.cpp:20-58, 320-347), the primitives of
 * kick_step (main.cpp:8-92).  Every kick (ct2_kick, ct2_kick_pm,
 * ct2_kick_pm_mlp) acts on the sinks of the current range [first, n) --
 * kick() grids only that range (particle.cpp:349-362) -- with every particle
 * as a source; the PM kicks compute the mesh force for all particles and
 * apply the epilogue to the range.  After ct2_particles_alloc the range is
 * [0, n).
 *   ct2_particle_sort_by_rung  particle_sort_by_rung (particle.cpp:320-340):
 *       partitions [first, n) so that rungs < minrung come first and moves
 *       first to the first particle with rung >= minrung; minrung 0 resets
 *       first to 0.  The partition is stable (the reference's swap loop is
 *       not); it permutes the container, and the stored forces are invalid
 *       until the next kick.
 *   ct2_particle_push_rungs / ct2_particle_pop_rungs  particle.cpp:31-37.
 *   ct2_particle_current_range  particle_current_range (particle.cpp:53-58).
 *   ct2_particle_rung_counts  particle_rung_counts (particle.cpp:39-51):
 *       counts[r] for r < *nrungs = 1 + the highest rung present.
 *   ct2_particle_set_minrung  particle_set_minrung (particle.cpp:342-347). */
int ct2_particle_sort_by_rung(int minrung);
int ct2_particle_push_rungs(void);
int ct2_particle_pop_rungs(void);
int ct2_particle_current_range(size_t* first, size_t* last);
int ct2_particle_rung_counts(uint64_t* counts, int max_rungs, int* nrungs);
int ct2_particle_set_minrung(int minrung);

/* Replaces particle_to_grid() (particle.cpp:59-114, 349-362) and the cell
 * accessors it fills (particle_get_cell, particle_cell_dim,
 * particle.cpp:290-296): sorts the current rung range [first, n) by cell,
 * cell (ix, iy, iz) with ix = floor(x ncell / 2^32), in the order
 * (ix ncell + iy) ncell + iz, permuting every container array.  The sort is
 * stable (ct2's recursive partition is not).  starts (host, ncell^3 + 1
 * entries, may be NULL) receives each cell's first container index: cell c
 * holds [starts[c], starts[c+1]).  The stored forces are invalid until the
 * next kick. */
#define CT2_GRID_MAX_NCELL 1024
int ct2_particles_to_grid(int ncell, uint64_t* starts);

/* Test hook for failure handling (no reference counterpart: ct2 aborts on a
 * failed allocation, memory.cpp:7-10): after `after` more device
 * allocations succeed, the next `count` the library makes fail with
 * CT2_ERR_NOMEM without calling hipMalloc (count 0 turns the hook off). */
int ct2_debug_fail_allocations(int after, int count);

/* ------------------------------------------------------------ drift */
/* Replaces void drift(float dt, float scale) (drift.hpp:12, drift.cpp:4-18):
 * x += fixed32(v * (dt/scale)) per axis, with fixed32(s) defined as
 * RNE(s * 2^32) mod 2^32 (0 for non-finite s).  Bit-exact. */
int ct2_drift(float dt, float scale);

/* ------------------------------------------------------------ multipole mesh */
/* ct2's own far field (particle.cpp:116-271, kick.cpp:11-29, 105-125) with
 * Cartesian Taylor expansions of total order `order` (1..8; ct2 uses 8) on
 * nc^3 cells (nc a power of two in [8, 256]): multipoles of each cell
 * (particle_compute_multipoles), one FP32 3-D FFT per pair of components, the
 * product with the FP64 table of derivatives of the periodic Green's function
 * (particle_greens: greens + greens_ewald, less the nearest image on the cells
 * with |offset|^2 <= 10, which are summed directly), inverse FFTs
 * (particle_compute_expansions), and at each particle the softened direct
 * pairs with the near set's cells (kick_cell's neighbour loop, kick.cpp:36-67)
 * plus the local expansion (L2P, kick.cpp:29).  ct2_mm_init builds the table
 * (once per nc and order; FP64, K nc^3 values for K = (order+1)(order+2)
 * (order+3)/6 components, at most 8 GiB). */
int ct2_mm_init(int nc, int order);
/* No counterpart in the reference (its cell grid follows ppcell in
 * particle_to_grid, particle.cpp:59-114): *nc = the power-of-two cell grid in
 * [8, 256] with at least `particles_per_cell` (>= 1) particles per cell on
 * average for n particles, the largest such; 8 per cell balances the near-set
 * pairs against the far field on this GPU. */
int ct2_mm_cells_for(size_t n, int particles_per_cell, int* nc);
/* Force and potential of every particle into the stored force arrays. */
int ct2_mm_forces(void);
/* kick() with the multipole-mesh force: ct2_mm_forces then the epilogue of
 * kick.cpp:75-99 on the current rung range.  flags: CT2_KICK_STORE. */
int ct2_kick_mm(const ct2_kick_params* params, unsigned flags, int* max_rung);
/* The transformed derivative table, [K][nc^3] doubles: Re (even degree) or
 * Im (odd degree) of the DFT of D_n over the cell offsets, divided by nc^3. */
int ct2_mm_table_download(double* host);

/* ------------------------------------------------------------ FFT */
/* In-place 3-D complex transform of an n^3 cube in device memory, natural
 * order, index (i*n + j)*n + k, interleaved (re, im).  dir = +1 is
 * e^{-2 pi i k.x/n} (FFTW_FORWARD), dir = -1 its conjugate; unnormalised.
 * n is a power of two in [8, 1024] (FP32) or [8, 512] (FP64).  Replaces the
 * 3-D transforms the PM path calls (fft.hpp:28-29, particle.cpp:230,261). */
int ct2_fft3d_c2c(void* data, int n, int dir);
int ct2_fft3d_c2c_f64(void* data, int n, int dir);

/* ct2's own FFT API (fft.hpp:18-29) on device memory, in ct2's conventions
 * (FP64; synchronise with ct2_synchronize before reading results on the host):
 *   ct2_fft_complex_1d  replaces fft_complex(complex<simd_f64>*, N, dir)
 *       (fft_complex_vector.cpp:9-65): `batch` groups of `lanes` (1 or 4)
 *       interleaved transforms, element n lane l at re X[2 lanes n + l],
 *       im X[2 lanes n + lanes + l], group stride 2 lanes N doubles.
 *       Input bit-reversed, output natural; dir = +1 is FFTW_FORWARD,
 *       dir = -1 the conjugate transform; unnormalised.  N in [2, 4096].
 *   ct2_fft_real_1d  replaces fft_real(double*, N, dir) and
 *       fft_real(simd_f64*, N, dir) (fft_real_scalar.cpp:9-180,
 *       fft_real_vector.cpp:9-183): real bit-reversed input [n][lane],
 *       half-complex output X[k] = Re F_k, X[N-k] = Im F_k.
 *   ct2_scramble  replaces scramble(X, N1, N2, N3) (scramble.cpp:8-43).
 *   ct2_transpose_yz / _r2c / _c2r  replace transpose.cpp:7-128 (intended
 *       index maps; planar out[x][z][ky], imaginary plane at N^2 (N/2+1)).
 *   ct2_fft_r2c_3d / ct2_fft_c2r_3d  replace fft_r2c_3d / fft_c2r_3d
 *       (fft_r2c_3d.cpp:4-31, fft_c2r_3d.cpp:4-40, stubs in ct2): real N^3
 *       <-> planar half spectrum out[(N/2+1)(N i + j) + k] = Re,
 *       out[N^2 (N/2+1) + same] = Im; unnormalised.  N in [8, 512].
 *   ct2_twiddles  replaces cos_twiddles/sin_twiddles/inv_sin_twiddles
 *       (twiddles.cpp:6-54): host arrays cos(-2 pi n/N), sin(-2 pi n/N),
 *       sin(2 pi n/N). */
int ct2_fft_complex_1d(double* X, int N, int dir, int lanes, size_t batch);
int ct2_fft_real_1d(double* X, int N, int dir, int lanes, size_t batch);
int ct2_scramble(double* X, int N1, int N2, int N3);
int ct2_transpose_yz(double* X, int N);
int ct2_transpose_r2c(const double* in, double* out, int N);
int ct2_transpose_c2r(const double* in, double* out, int N);
int ct2_fft_r2c_3d(const double* in, double* out, int N);
int ct2_fft_c2r_3d(const double* in, double* out, int N);
int ct2_twiddles(int N, double* cosw, double* sinw, double* invsinw);

/* Device memory for callers without HIP code of their own. */
int ct2_dev_alloc(size_t bytes, void** ptr);
int ct2_dev_free(void* ptr);
int ct2_memcpy_h2d(void* dst, const void* src, size_t bytes);
int ct2_memcpy_d2h(void* dst, const void* src, size_t bytes);

/* ------------------------------------------------------------ initial conditions */
/* Seeded synthetic particle sets (host side; replace particles_init,
 * particle.cpp:364-374, whose libc rand() is not reproducible).
 * Lattice (i+1/2)/n per axis plus a Gaussian displacement of sigma_cells
 * lattice cells per axis, wrapped, rounded RNE to fixed32.  n^3 particles,
 * ordered (i*n + j)*n + k. */
int ct2_ic_lattice(int n, double sigma_cells, uint64_t seed, uint32_t* x, uint32_t* y,
                   uint32_t* z);
/* n particles: the first n_sph drawn from `nspheres` Plummer spheres of
 * scale radius `a` truncated at rmax (centres uniform), the rest uniform. */
int ct2_ic_plummer_mix(size_t n, size_t n_sph, int nspheres, double a, double rmax,
                       uint64_t seed, uint32_t* x, uint32_t* y, uint32_t* z);

/* Lattice of n^3 particles (n a power of two <= 512) displaced by a Gaussian
 * random field in the Zel'dovich approximation, written straight into the
 * device container (which must hold n^3 particles): density power
 * P(k) ~ k^spectral_index up to the Nyquist sphere, displacement rms per
 * component rms_cells lattice cells, velocities vfac * displacement.  FP64
 * FFTs; seeded and deterministic. */
int ct2_ic_grf(int n, double rms_cells, double spectral_index, uint64_t seed, double vfac);

/* ------------------------------------------------------------ cosmology */
/* FP64 scale-factor and time integrals of cosmology.hpp:13-16
 * (cosmology.cpp:11-102) with explicit parameters instead of get_options():
 * H is the Hubble constant in code time units. */
typedef struct {
    double H;
    double omega_m;
    double omega_r;
    double omega_lam;
} ct2_cosmology;
int ct2_cosmos_update(const ct2_cosmology* c, double* adotdot, double* adot, double* a, double dt0);
double ct2_cosmos_dadt(const ct2_cosmology* c, double a);
double ct2_cosmos_conformal_time(const ct2_cosmology* c, double a0, double a1);
double ct2_cosmos_time(const ct2_cosmology* c, double a0, double a1);
/* Leapfrog factors between two scale factors: int dt/a (kick of the
 * canonical momentum) and int dt/a^2 (drift), to 1e-12 relative. */
double ct2_cosmos_kick_factor(const ct2_cosmology* c, double a0, double a1);
double ct2_cosmos_drift_factor(const ct2_cosmology* c, double a0, double a1);

/* ------------------------------------------------------ options, units */
/* The options block (options.hpp:14-32) with the reference's field types.
 * ct2_options_defaults: the command-line defaults (options.cpp:51-61).
 * ct2_options_derive: GM = 1/Ngrid^3, nparts = Ngrid^3, code units in cgs
 *   (code_to_g = M_sun mass_res, code_to_cm the side of a box holding nparts
 *   such masses at omega_m times the critical density, code_to_s =
 *   code_to_cm / c), omega_r from the CMB temperature 2.73 Theta K with
 *   N_eff = 3.086, and omega_lam = 1 - omega_r - omega_m (options.cpp:88-95,
 *   constants of constants.hpp:14-39).  CT2_ERR_INVALID if Ngrid, mass_res,
 *   hubble or omega_m is not positive.
 * ct2_options_cosmology: the block the cosmology functions read, with
 *   H = H0 code_to_s hubble in code time units (cosmology.cpp:15,47).
 * The three code units are FP64 here, FP32 in the reference: M_sun mass_res
 * exceeds the FP32 range for mass_res above 1.7e5 solar masses, so the
 * reference's code_to_g, code_to_cm, code_to_s and H are infinite at its own
 * default mass_res of 1e9.  The other fields keep the FP32 rounding. */
typedef struct {
    float bwidth;
    double code_to_cm;
    double code_to_g;
    double code_to_s;
    float eta;
    float GM;
    float hsoft;
    float hubble;
    float omega_lam;
    float omega_m;
    float omega_r;
    float mass_res;
    uint64_t Ngrid;
    uint64_t nparts;
    int ppcell;
    float Theta;
} ct2_options;
int ct2_options_defaults(ct2_options* o);
int ct2_options_derive(ct2_options* o);
int ct2_options_cosmology(const ct2_options* o, ct2_cosmology* c);

#ifdef __cplusplus
}
#endif
#endif
