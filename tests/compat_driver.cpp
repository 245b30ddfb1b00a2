// Drives the ct2-signature binding (kick(), drift(), particle globals) the way
// ct2's kick_step would (main.cpp:62-81) and prints the resulting state, for
// tests/test_gpu_compat.py to compare with the CPU oracle.  With the argument
// "fft" it drives the FFT API the way fft_test.cpp does (scramble, then
// fft_real; fft_r2c_3d, then fft_c2r_3d) and prints the results.  With
// "learned <weights file>" it switches kick() to PM + the learned correction
// (raw FP32 weights from the file) and prints the velocities of one kick.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ct2_gpu_binding.hpp"
#include "ct2hip.h"

static double lcg_unit(unsigned& s) {
    s = s * 1664525u + 1013904223u;
    return (double)s / 4294967296.0 - 0.5;
}

static int fft_mode() {
    unsigned s = 777u;
    const int N = 1024;
    std::vector<double> x(N);
    for (double& v : x) v = lcg_unit(s);
    scramble(x.data(), 1, N, 1);
    fft_real(x.data(), N, 1);
    for (double v : x) std::printf("%.17g\n", v);
    const int M = 16, h = M / 2 + 1;
    std::vector<double> cube(M * M * M), spec(2 * M * M * h), back(M * M * M);
    for (double& v : cube) v = lcg_unit(s);
    fft_r2c_3d(cube.data(), spec.data(), M);
    fft_c2r_3d(spec.data(), back.data(), M);
    for (double v : spec) std::printf("%.17g\n", v);
    for (double v : back) std::printf("%.17g\n", v);
    const std::vector<double>& c = cos_twiddles(8);
    for (double v : c) std::printf("%.17g\n", v);
    return 0;
}

// Random particles plus tight pairs (fast rungs); kick_step five times with
// the kick_step trace on stdout, then the final rungs.
static int rungs_mode() {
    const int n = 1024;
    particle_allocate(n);
    ct2_gpu_set_options(2e-3f, 1.0f / n, 0.15f);
    unsigned s = 4242u;
    for (int i = 0; i < n; i++) {
        s = s * 1664525u + 1013904223u;
        particle_x_ptr[i] = s;
        s = s * 1664525u + 1013904223u;
        particle_y_ptr[i] = s;
        s = s * 1664525u + 1013904223u;
        particle_z_ptr[i] = s;
        particle_vx_ptr[i] = particle_vy_ptr[i] = particle_vz_ptr[i] = 0.0f;
        particle_rung_ptr[i] = 0;
    }
    for (int k = 0; k < 48; k++) {  // partner 2k+1 sits 2^-12 .. 2^-8 box from 2k
        const unsigned d = 1u << (20 + k % 5);
        particle_x_ptr[2 * k + 1] = particle_x_ptr[2 * k] + d;
        particle_y_ptr[2 * k + 1] = particle_y_ptr[2 * k];
        particle_z_ptr[2 * k + 1] = particle_z_ptr[2 * k];
    }
    ct2_gpu_trace(stdout);
    int minrung = 0, max_rung = 0;
    double tau = 0.0;
    const double t0 = 0.01;
    for (int step = 0; step < 5; step++) {
        std::printf("step %d minrung %d max_rung %d\n", step, minrung, max_rung);
        ct2_gpu_kick_step(minrung, max_rung, 1.0, 1.0, tau, t0, 0, false);
        tau += t0;
        ct2_gpu_trace(nullptr);
        max_rung = (int)particle_rung_counts().size() - 1;
        ct2_gpu_trace(stdout);
    }
    std::printf("end minrung %d max_rung %d\n", minrung, max_rung);
    return 0;
}

static int learned_mode(const char* path) {
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return 2;
    std::vector<float> w(ct2_mlp_num_params());
    const size_t got = std::fread(w.data(), sizeof(float), w.size(), f);
    std::fclose(f);
    if (got != w.size()) return 3;
    const int n = 4096;
    particle_allocate(n);
    ct2_gpu_set_options(0.01f, 1.0f / n);
    unsigned s = 777u;
    for (int i = 0; i < n; i++) {
        s = s * 1664525u + 1013904223u;
        particle_x_ptr[i] = s;
        s = s * 1664525u + 1013904223u;
        particle_y_ptr[i] = s;
        s = s * 1664525u + 1013904223u;
        particle_z_ptr[i] = s;
        particle_vx_ptr[i] = particle_vy_ptr[i] = particle_vz_ptr[i] = 0.0f;
        particle_rung_ptr[i] = 0;
    }
    ct2_gpu_use_learned(w.data(), w.size(), 32, 1.0, 2);
    kick_params kp{true, true, true, false, 0, 0.01f, 0.15f, 0.01f, 1.0f};
    const int max_rung = kick(kp);
    std::printf("max_rung %d\n", max_rung);
    for (int i = 0; i < n; i++)
        std::printf("%.9e %.9e %.9e %d\n", particle_vx_ptr[i], particle_vy_ptr[i], particle_vz_ptr[i],
                    (int)particle_rung_ptr[i]);
    return 0;
}

// kick() switched to the multipole mesh (16^3 cells, order 6) on 4,096 LCG
// particles with the binding's default kernel; prints max_rung, then vx vy vz
// rung per particle.
static int mm_mode() {
    const int n = 4096;
    particle_allocate(n);
    ct2_gpu_set_options(0.01f, 1.0f / n);
    unsigned s = 4242u;
    for (int i = 0; i < n; i++) {
        s = s * 1664525u + 1013904223u;
        particle_x_ptr[i] = s;
        s = s * 1664525u + 1013904223u;
        particle_y_ptr[i] = s;
        s = s * 1664525u + 1013904223u;
        particle_z_ptr[i] = s;
        particle_vx_ptr[i] = particle_vy_ptr[i] = particle_vz_ptr[i] = 0.0f;
        particle_rung_ptr[i] = 0;
    }
    ct2_gpu_use_multipoles(16, 6);
    kick_params kp{true, true, true, false, 0, 0.01f, 0.15f, 0.01f, 1.0f};
    const int max_rung = kick(kp);
    std::printf("max_rung %d\n", max_rung);
    for (int i = 0; i < n; i++)
        std::printf("%.9e %.9e %.9e %d\n", particle_vx_ptr[i], particle_vy_ptr[i], particle_vz_ptr[i],
                    (int)particle_rung_ptr[i]);
    return 0;
}

// particles_init, then particle_to_grid over the rung range [first, n) left by
// particle_sort_by_rung(1) (rungs 0/1 alternate in blocks of 7); prints the
// range, the grid's dimension, each cell's range, the neighbour count of a
// corner cell, and the particles in container order.
static int grid_mode() {
    const int n = 20000;
    particle_allocate(n);
    particles_init();
    for (int i = 0; i < n; i++) {
        particle_vx_ptr[i] = (float)i;  // a tag that follows each particle
        particle_rung_ptr[i] = (char)((i / 7) % 2);
    }
    ct2_gpu_set_grid(40, 2);
    particle_sort_by_rung(1);
    const std::pair<size_t, size_t> rng = particle_current_range();
    particle_to_grid();
    const int nc = particle_cell_dim();
    std::printf("range %zu %zu ncell %d neighbours %zu\n", rng.first, rng.second, nc,
                particle_cell_neighbors(0, 0, 0).size());
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++)
            for (int k = 0; k < nc; k++) {
                const std::pair<size_t, size_t> c = particle_get_cell(i, j, k);
                std::printf("%zu %zu\n", c.first, c.second);
            }
    for (int i = 0; i < n; i++)
        std::printf("%u %u %u %.1f %d\n", particle_x_ptr[i], particle_y_ptr[i], particle_z_ptr[i], particle_vx_ptr[i],
                    (int)particle_rung_ptr[i]);
    return 0;
}

// Default mode: `compat_driver n [kernel]`; without a kernel argument the
// binding's default (ct2's uniform-density sphere, kick.cpp:52-60) is used.
int main(int argc, char** argv) {
    if (argc > 2 && std::strcmp(argv[1], "learned") == 0) return learned_mode(argv[2]);
    if (argc > 1 && std::strcmp(argv[1], "fft") == 0) return fft_mode();
    if (argc > 1 && std::strcmp(argv[1], "rungs") == 0) return rungs_mode();
    if (argc > 1 && std::strcmp(argv[1], "grid") == 0) return grid_mode();
    if (argc > 1 && std::strcmp(argv[1], "mm") == 0) return mm_mode();
    const int n = argc > 1 ? std::atoi(argv[1]) : 512;
    particle_allocate(n);
    if (argc > 2)
        ct2_gpu_set_options(0.01f, 1.0f / n, 0.15f, std::atoi(argv[2]));
    else
        ct2_gpu_set_options(0.01f, 1.0f / n);
    unsigned s = 12345u;
    for (int i = 0; i < n; i++) {
        s = s * 1664525u + 1013904223u;
        particle_x_ptr[i] = s;
        s = s * 1664525u + 1013904223u;
        particle_y_ptr[i] = s;
        s = s * 1664525u + 1013904223u;
        particle_z_ptr[i] = s;
        particle_vx_ptr[i] = particle_vy_ptr[i] = particle_vz_ptr[i] = 0.0f;
        particle_rung_ptr[i] = 0;
    }
    kick_params kp{true, true, true, false, 0, 0.01f, 0.15f, 0.01f, 1.0f};
    const int max_rung = kick(kp);
    drift(0.01f, 1.0f);
    std::printf("max_rung %d\n", max_rung);
    for (int i = 0; i < n; i++)
        std::printf("%u %u %u %.9e %.9e %.9e %d\n", particle_x_ptr[i], particle_y_ptr[i], particle_z_ptr[i],
                    particle_vx_ptr[i], particle_vy_ptr[i], particle_vz_ptr[i], (int)particle_rung_ptr[i]);
    return 0;
}
