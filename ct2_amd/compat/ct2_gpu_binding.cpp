// ct2_gpu_binding.cpp -- see ct2_gpu_binding.hpp.
#include "ct2_gpu_binding.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "ct2hip.h"

ct2_pos_fixed* particle_x_ptr = nullptr;
ct2_pos_fixed* particle_y_ptr = nullptr;
ct2_pos_fixed* particle_z_ptr = nullptr;
float* particle_vx_ptr = nullptr;
float* particle_vy_ptr = nullptr;
float* particle_vz_ptr = nullptr;
char* particle_rung_ptr = nullptr;

namespace {
size_t g_count = 0;
char* g_store = nullptr;  // the host arrays, one page-locked block (ct2_host_alloc)
float g_eta = 0.15f;
int g_ppcell = 256, g_bwidth = 3;  // options.cpp:61,54
int g_ncell = 0;                   // the last particle_to_grid
std::vector<uint64_t> g_starts;    // its cell starts, g_ncell^3 + 1
int g_mesh_ng = 0;   // 0: direct-sum kick
bool g_learned = false;  // with g_mesh_ng > 0: PM + learned correction instead of PM + near field
int g_mm_cells = 0;      // > 0: the multipole mesh (takes precedence over the PM modes)
std::FILE* g_trace = nullptr;

// ct2 aborts on allocation failure (memory.cpp:7-10); the binding does the
// same on any library error, with the library's message.
void check(int rc, const char* what) {
    if (rc != CT2_OK) {
        std::fprintf(stderr, "ct2 GPU path: %s failed: %s\n", what, ct2_last_error());
        std::abort();
    }
}

void ensure_init() {
    static bool done = false;
    if (!done) {
        check(ct2_init(0), "ct2_init");
        done = true;
    }
}
}  // namespace

void ct2_gpu_set_options(float hsoft, float GM, float eta, int kernel) {
    ensure_init();
    ct2_force_options o{hsoft, GM, kernel};
    check(ct2_set_force_options(&o), "ct2_set_force_options");
    g_eta = eta;
}

void ct2_gpu_set_grid(int ppcell, int bwidth) {
    if (ppcell <= 0 || bwidth < 0) {
        std::fprintf(stderr, "ct2 GPU path: bad grid options ppcell %d bwidth %d\n", ppcell, bwidth);
        std::abort();
    }
    g_ppcell = ppcell;
    g_bwidth = bwidth;
}

void ct2_gpu_use_mesh(int ng, double rs_cells, double rcut_rs, int order) {
    ensure_init();
    g_mm_cells = 0;
    g_mesh_ng = ng;
    g_learned = false;
    if (ng > 0) {
        check(ct2_pm_assignment(order), "ct2_pm_assignment");
        check(ct2_pm_init(ng, rs_cells), "ct2_pm_init");
        check(ct2_pm_near(rcut_rs), "ct2_pm_near");
    }
}

void ct2_gpu_use_learned(const float* weights, size_t n, int ng, double rs_cells, int order,
                         double particles_per_node) {
    ensure_init();
    check(ct2_pm_assignment(order), "ct2_pm_assignment");
    check(ct2_pm_init(ng, rs_cells), "ct2_pm_init");
    check(ct2_pm_near(0.0), "ct2_pm_near");
    check(ct2_mlp_init(1), "ct2_mlp_init");
    check(ct2_mlp_set_weights(weights, n), "ct2_mlp_set_weights");
    check(ct2_mlp_set_mesh(rs_cells, order, 0.0, particles_per_node), "ct2_mlp_set_mesh");
    g_mm_cells = 0;
    g_mesh_ng = ng;
    g_learned = true;
}

void ct2_gpu_use_multipoles(int ncell, int order) {
    ensure_init();
    if (ncell < 0) check(ct2_mm_cells_for(g_count, 8, &ncell), "ct2_mm_cells_for");
    if (ncell > 0) check(ct2_mm_init(ncell, order), "ct2_mm_init");
    g_mm_cells = ncell;
}

void ct2_gpu_trace(std::FILE* f) { g_trace = f; }

void particle_allocate(size_t count) {
    ensure_init();
    g_count = count;
    // one block, sizeof(T) * count per array (particle.cpp:275-283 intent),
    // page-locked so that the per-call copies run at the host link's rate
    const size_t bytes = count * (3 * sizeof(ct2_pos_fixed) + 3 * sizeof(float) + 1);
    check(ct2_host_free(g_store), "ct2_host_free");
    g_store = nullptr;
    void* block = nullptr;
    check(ct2_host_alloc(bytes, &block), "ct2_host_alloc");
    g_store = (char*)block;
    if (bytes) std::memset(g_store, 0, bytes);
    char* p = g_store;
    particle_x_ptr = (ct2_pos_fixed*)p;
    particle_y_ptr = particle_x_ptr + count;
    particle_z_ptr = particle_y_ptr + count;
    particle_vx_ptr = (float*)(particle_z_ptr + count);
    particle_vy_ptr = particle_vx_ptr + count;
    particle_vz_ptr = particle_vy_ptr + count;
    particle_rung_ptr = (char*)(particle_vz_ptr + count);
    check(ct2_particles_alloc(count), "ct2_particles_alloc");
}

size_t particle_count() { return g_count; }

void particles_init() {
    // ct2 draws libc rand() uniforms (particle.cpp:364-374, util.cpp:3-5); the
    // library's seeded generator replaces them, so a run is reproducible
    ensure_init();
    check(ct2_ic_plummer_mix(g_count, 0, 0, 0.0, 0.0, 1, particle_x_ptr, particle_y_ptr, particle_z_ptr),
          "ct2_ic_plummer_mix");
    for (size_t i = 0; i < g_count; ++i) {
        particle_vx_ptr[i] = particle_vy_ptr[i] = particle_vz_ptr[i] = 0.0f;
        particle_rung_ptr[i] = 0;
    }
}

int kick(const kick_params& p) {
    check(ct2_particles_upload(particle_x_ptr, particle_y_ptr, particle_z_ptr, particle_vx_ptr, particle_vy_ptr,
                               particle_vz_ptr, (const int8_t*)particle_rung_ptr),
          "upload");
    const ct2_kick_params q{p.ascending, p.descending, p.top, p.first, p.rung, p.t0, p.eta, p.max_dt, p.scale};
    int max_rung = 0;
    if (g_mm_cells > 0)
        check(ct2_kick_mm(&q, CT2_KICK_STORE, &max_rung), "ct2_kick_mm");
    else if (g_mesh_ng > 0 && g_learned)
        check(ct2_kick_pm_mlp(&q, CT2_KICK_STORE, &max_rung), "ct2_kick_pm_mlp");
    else if (g_mesh_ng > 0)
        check(ct2_kick_pm(&q, CT2_KICK_STORE, &max_rung), "ct2_kick_pm");
    else
        check(ct2_kick(&q, CT2_KICK_STORE, &max_rung), "ct2_kick");
    if (g_trace) {
        size_t first = 0, last = 0;
        check(ct2_particle_current_range(&first, &last), "ct2_particle_current_range");
        std::fprintf(g_trace, "kick rung %d asc %d desc %d top %d first %d range %zu -> max_rung %d\n", p.rung,
                     (int)p.ascending, (int)p.descending, (int)p.top, (int)p.first, first, max_rung);
    }
    check(ct2_particles_download(nullptr, nullptr, nullptr, particle_vx_ptr, particle_vy_ptr, particle_vz_ptr,
                                 (int8_t*)particle_rung_ptr),
          "download");
    return max_rung;
}

void drift(float dt, float scale) {
    if (g_trace) std::fprintf(g_trace, "drift %.9e %.9e\n", dt, scale);
    check(ct2_particles_upload(particle_x_ptr, particle_y_ptr, particle_z_ptr, particle_vx_ptr, particle_vy_ptr,
                               particle_vz_ptr, nullptr),
          "upload");
    check(ct2_drift(dt, scale), "ct2_drift");
    check(ct2_particles_download(particle_x_ptr, particle_y_ptr, particle_z_ptr, nullptr, nullptr, nullptr, nullptr),
          "download");
}

// ---- rung bookkeeping: the host arrays stay authoritative, as in ct2
namespace {
void upload_all() {
    check(ct2_particles_upload(particle_x_ptr, particle_y_ptr, particle_z_ptr, particle_vx_ptr, particle_vy_ptr,
                               particle_vz_ptr, (const int8_t*)particle_rung_ptr),
          "upload");
}
}  // namespace

void particle_sort_by_rung(int minrung) {
    upload_all();
    check(ct2_particle_sort_by_rung(minrung), "ct2_particle_sort_by_rung");
    check(ct2_particles_download(particle_x_ptr, particle_y_ptr, particle_z_ptr, particle_vx_ptr, particle_vy_ptr,
                                 particle_vz_ptr, (int8_t*)particle_rung_ptr),
          "download");
    if (g_trace) std::fprintf(g_trace, "sort_by_rung %d\n", minrung);
}

std::vector<size_t> particle_rung_counts() {
    check(ct2_particles_upload(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                               (const int8_t*)particle_rung_ptr),
          "upload");
    uint64_t c[128];
    int nr = 0;
    check(ct2_particle_rung_counts(c, 128, &nr), "ct2_particle_rung_counts");
    if (g_trace) {
        std::fprintf(g_trace, "rung_counts");
        for (int r = 0; r < nr; ++r) std::fprintf(g_trace, " %llu", (unsigned long long)c[r]);
        std::fprintf(g_trace, "\n");
    }
    return std::vector<size_t>(c, c + nr);
}

std::pair<size_t, size_t> particle_current_range() {
    size_t first = 0, last = 0;
    check(ct2_particle_current_range(&first, &last), "ct2_particle_current_range");
    return {first, last};
}

void particle_to_grid() {
    const std::pair<size_t, size_t> rng = particle_current_range();
    const size_t count = rng.second - rng.first;
    const int ncell = (int)(std::pow((double)(count / (size_t)g_ppcell), 1.0 / 3.0) + 4);
    upload_all();
    g_starts.assign((size_t)ncell * ncell * ncell + 1, 0);
    check(ct2_particles_to_grid(ncell, g_starts.data()), "ct2_particles_to_grid");
    check(ct2_particles_download(particle_x_ptr, particle_y_ptr, particle_z_ptr, particle_vx_ptr, particle_vy_ptr,
                                 particle_vz_ptr, (int8_t*)particle_rung_ptr),
          "download");
    g_ncell = ncell;
    if (g_trace) std::fprintf(g_trace, "to_grid %d\n", ncell);
}

int particle_cell_dim() { return g_ncell; }

std::pair<size_t, size_t> particle_get_cell(int i, int j, int k) {
    const int n = g_ncell;
    if (n <= 0 || i < 0 || j < 0 || k < 0 || i >= n || j >= n || k >= n) {
        std::fprintf(stderr, "ct2 GPU path: cell (%d, %d, %d) outside the %d^3 grid\n", i, j, k, n);
        std::abort();
    }
    const size_t c = ((size_t)i * n + j) * n + k;
    return {(size_t)g_starts[c], (size_t)g_starts[c + 1]};
}

std::vector<std::pair<size_t, size_t>> particle_cell_neighbors(int i, int j, int k) {
    const int n = g_ncell, b = g_bwidth;
    std::vector<std::pair<size_t, size_t>> out;
    for (int i0 = -b; i0 <= b; ++i0)
        for (int j0 = -b; j0 <= b; ++j0)
            for (int k0 = -b; k0 <= b; ++k0)
                if (i0 * i0 + j0 * j0 + k0 * k0 <= b * b)
                    out.push_back(particle_get_cell(((i + i0) % n + n) % n, ((j + j0) % n + n) % n,
                                                    ((k + k0) % n + n) % n));
    return out;
}

void particle_pop_rungs() {
    check(ct2_particle_pop_rungs(), "ct2_particle_pop_rungs");
    if (g_trace) std::fprintf(g_trace, "pop\n");
}

void particle_push_rungs() {
    check(ct2_particle_push_rungs(), "ct2_particle_push_rungs");
    if (g_trace) std::fprintf(g_trace, "push\n");
}

void particle_set_minrung(int minrung) {
    check(ct2_particles_upload(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                               (const int8_t*)particle_rung_ptr),
          "upload");
    check(ct2_particle_set_minrung(minrung), "ct2_particle_set_minrung");
    check(ct2_particles_download(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, (int8_t*)particle_rung_ptr),
          "download");
    if (g_trace) std::fprintf(g_trace, "set_minrung %d\n", minrung);
}

// The level schedule of kick_step (main.cpp:8-92): levels max_rung .. minrung+1
// kick the ascending half only; minrung is the top level (both halves, sort by
// rung); a kick that reports faster particles queues the next finer level;
// drifts follow each descending kick with that level's dt; when the rungs
// above minrung0 hold more than a third of the slow particles, the top level is
// clipped (nocrop off) and minrung rises by one.
int ct2_gpu_kick_step(int& minrung, int max_rung, double scale, double adot, double tau, double t0, int minrung0,
                      bool nocrop) {
    std::vector<int> levels;
    for (int r = max_rung; r > minrung; --r) levels.push_back(r);
    levels.push_back(minrung);
    bool ascending = true, clip_top = false;
    for (size_t li = 0; li < levels.size(); ++li) {
        const int level = levels[li];
        const bool top = level == minrung;
        if (top) ascending = false;
        if (!ascending && !top) particle_push_rungs();
        if (!ascending) particle_sort_by_rung(level);
        if (top && minrung == minrung0) {
            const std::vector<size_t> counts = particle_rung_counts();
            if (counts.size() > (size_t)minrung0 + 1) {
                size_t fast = 0;
                for (size_t r = minrung0 + 1; r < counts.size(); ++r) fast += counts[r];
                if (3 * fast > counts[minrung0]) clip_top = !nocrop;
            }
        }
        particle_to_grid();
        const bool clipped_top = clip_top && top;
        kick_params kp;
        kp.top = (clip_top && level == minrung0 + 1) || top;
        kp.ascending = ascending || top;
        kp.descending = clipped_top ? false : (!ascending || top);
        kp.eta = g_eta;
        kp.scale = (float)scale;
        kp.first = tau == 0.0;
        kp.rung = level;
        kp.t0 = (float)t0;
        kp.max_dt = (float)(0.01f / adot);
        const int reached = kick(kp);
        if (clipped_top) {
            particle_set_minrung(minrung0 + 1);
            levels.push_back(minrung0 + 1);
        } else if ((!ascending || top) && reached > level) {
            levels.push_back(level + 1);
        }
        if ((!ascending || top) && !clipped_top) drift((float)(t0 / (double)(1 << level)), (float)scale);
        if (ascending && !top) particle_pop_rungs();
    }
    if (clip_top) ++minrung;
    return max_rung;
}

// ---- FFT API: host arrays in and out, the transform on the device.
namespace {
struct DeviceCopy {  // a device copy of a host array of n doubles
    void* p = nullptr;
    size_t bytes = 0;
    DeviceCopy(const double* host, size_t n) : bytes(n * sizeof(double)) {
        check(ct2_dev_alloc(bytes, &p), "ct2_dev_alloc");
        if (host) check(ct2_memcpy_h2d(p, host, bytes), "ct2_memcpy_h2d");
    }
    void to_host(double* host) const { check(ct2_memcpy_d2h(host, p, bytes), "ct2_memcpy_d2h"); }
    double* ptr() const { return (double*)p; }
    ~DeviceCopy() { ct2_dev_free(p); }
};

void real_1d(double* X, int N, int dir, int lanes) {
    ensure_init();
    DeviceCopy d(X, (size_t)N * lanes);
    check(ct2_fft_real_1d(d.ptr(), N, dir, lanes, 1), "ct2_fft_real_1d");
    check(ct2_synchronize(), "ct2_synchronize");
    d.to_host(X);
}
}  // namespace

void fft_real(fft_float* X, int N, int dir) { real_1d(X, N, dir, 1); }
void ct2_gpu_fft_real_lanes4(double* X, int N, int dir) { real_1d(X, N, dir, 4); }

void ct2_gpu_fft_complex_lanes4(double* X, int N, int dir) {
    ensure_init();
    DeviceCopy d(X, (size_t)N * 8);
    check(ct2_fft_complex_1d(d.ptr(), N, dir, 4, 1), "ct2_fft_complex_1d");
    check(ct2_synchronize(), "ct2_synchronize");
    d.to_host(X);
}

void fft_r2c_3d(fft_float* in, fft_float* out, int N) {
    ensure_init();
    const size_t n3 = (size_t)N * N * N, half = (size_t)N * N * (N / 2 + 1);
    DeviceCopy a(in, n3), b(nullptr, 2 * half);
    check(ct2_fft_r2c_3d(a.ptr(), b.ptr(), N), "ct2_fft_r2c_3d");
    check(ct2_synchronize(), "ct2_synchronize");
    b.to_host(out);
}

void fft_c2r_3d(fft_float* in, fft_float* out, int N) {
    ensure_init();
    const size_t n3 = (size_t)N * N * N, half = (size_t)N * N * (N / 2 + 1);
    DeviceCopy a(in, 2 * half), b(nullptr, n3);
    check(ct2_fft_c2r_3d(a.ptr(), b.ptr(), N), "ct2_fft_c2r_3d");
    check(ct2_synchronize(), "ct2_synchronize");
    b.to_host(out);
}

void scramble(fft_float* X, int N1, int N2, int N3) {
    ensure_init();
    DeviceCopy d(X, (size_t)N1 * N2 * N3);
    check(ct2_scramble(d.ptr(), N1, N2, N3), "ct2_scramble");
    check(ct2_synchronize(), "ct2_synchronize");
    d.to_host(X);
}

void transpose_yz(fft_float* X, int N) {
    ensure_init();
    DeviceCopy d(X, (size_t)N * N * N);
    check(ct2_transpose_yz(d.ptr(), N), "ct2_transpose_yz");
    check(ct2_synchronize(), "ct2_synchronize");
    d.to_host(X);
}

void transpose_r2c(fft_float* in, fft_float* out, int N) {
    ensure_init();
    DeviceCopy a(in, (size_t)N * N * N), b(nullptr, 2 * (size_t)N * N * (N / 2 + 1));
    check(ct2_transpose_r2c(a.ptr(), b.ptr(), N), "ct2_transpose_r2c");
    check(ct2_synchronize(), "ct2_synchronize");
    b.to_host(out);
}

void transpose_c2r(fft_float* in, fft_float* out, int N) {
    ensure_init();
    DeviceCopy a(in, 2 * (size_t)N * N * (N / 2 + 1)), b(out, (size_t)N * N * N);
    check(ct2_transpose_c2r(a.ptr(), b.ptr(), N), "ct2_transpose_c2r");
    check(ct2_synchronize(), "ct2_synchronize");
    b.to_host(out);
}

namespace {
// twiddles.cpp keeps one table per N; so does this (not thread-safe either)
const std::vector<double>& twiddle_table(int N, int which) {
    static std::map<int, std::vector<double>> tables[3];
    auto it = tables[which].find(N);
    if (it == tables[which].end()) {
        std::vector<double> c(N), s(N), i(N);
        check(ct2_twiddles(N, c.data(), s.data(), i.data()), "ct2_twiddles");
        tables[0][N] = c;
        tables[1][N] = s;
        tables[2][N] = i;
        it = tables[which].find(N);
    }
    return it->second;
}
}  // namespace

const std::vector<fft_float>& cos_twiddles(int N) { return twiddle_table(N, 0); }
const std::vector<fft_float>& sin_twiddles(int N) { return twiddle_table(N, 1); }
const std::vector<fft_float>& inv_sin_twiddles(int N) { return twiddle_table(N, 2); }
