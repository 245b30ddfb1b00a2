// ct2's own declarations of the functions the GPU binding replaces, written as
// ct2's headers write them (kick.hpp:10-24, drift.hpp:12, fft.hpp:18-29
// without the two sfmm-typed overloads, particle.hpp:58-70 without the
// sfmm-typed particle_get_expansion), and one reference to each.  This file
// does not include the binding's header: linking it against
// ct2_gpu_binding.o proves that the binding defines every symbol, with the
// exact parameter types, that ct2's callers reference.  It is linked, not run.
#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

using fft_float = double;  // cosmictiger.hpp:18

// kick.hpp:10-24
struct kick_params {
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
int kick_step(int& minrung, int max_rung, double scale, double adot, double tau, double t0, int minrung0, bool nocrop);
int kick(const kick_params& params);

// drift.hpp:12
void drift(float dt, float scale);

// fft.hpp:18-29
void transpose_yz(fft_float* X, int N);
void transpose_r2c(fft_float* xin, fft_float* xout, int N);
void transpose_c2r(fft_float* xin, fft_float* xout, int N);
void scramble(fft_float* X, int N1, int N2, int N3);
const std::vector<fft_float>& cos_twiddles(int N);
const std::vector<fft_float>& sin_twiddles(int N);
const std::vector<fft_float>& inv_sin_twiddles(int N);
void fft_real(fft_float* X, int N, int dir = 1);
void fft_r2c_3d(fft_float* xin, fft_float* xout, int N);
void fft_c2r_3d(fft_float* xin, fft_float* xout, int N);

// particle.hpp:58-70
size_t particle_count();
void particles_init();
void particle_sort_by_rung(int minrung);
std::vector<size_t> particle_rung_counts();
void particle_to_grid();
std::pair<size_t, size_t> particle_current_range();
void particle_to_grid();
void particle_pop_rungs();
void particle_push_rungs();
void particle_set_minrung(int minrung);
std::vector<std::pair<size_t, size_t>> particle_cell_neighbors(int i, int j, int k);
int particle_cell_dim();
std::pair<size_t, size_t> particle_get_cell(int i, int j, int k);

// particle_allocate is defined in particle.cpp:273 but declared in no header;
// main.cpp's caller would declare it as the definition reads.
void particle_allocate(size_t count);

// kick_step is defined by ct2's main.cpp:8-92, which a ct2 build keeps, so it
// is declared above and not referenced here.
int main(int argc, char**) {
    if (argc > 0) return 0;  // never runs the calls below; they only have to link
    double buf[8] = {0};
    kick_params kp{};
    int r = kick(kp);
    drift(0.f, 1.f);
    transpose_yz(buf, 2);
    transpose_r2c(buf, buf, 2);
    transpose_c2r(buf, buf, 2);
    scramble(buf, 1, 2, 1);
    r += (int)cos_twiddles(2).size() + (int)sin_twiddles(2).size() + (int)inv_sin_twiddles(2).size();
    fft_real(buf, 2);
    fft_r2c_3d(buf, buf, 2);
    fft_c2r_3d(buf, buf, 2);
    particle_allocate(1);
    particles_init();
    r += (int)particle_count();
    particle_sort_by_rung(0);
    r += (int)particle_rung_counts().size();
    particle_to_grid();
    r += (int)particle_current_range().first;
    particle_push_rungs();
    particle_pop_rungs();
    particle_set_minrung(0);
    r += (int)particle_cell_neighbors(0, 0, 0).size();
    r += particle_cell_dim();
    r += (int)particle_get_cell(0, 0, 0).first;
    return r;
}
