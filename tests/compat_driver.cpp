// Drives the ct2-signature binding (kick(), drift(), particle globals) the way
// ct2's kick_step would (main.cpp:62-81) and prints the resulting state, for
// tests/test_gpu_compat.py to compare with the CPU oracle.  With the argument
// "fft" it drives the FFT API the way fft_test.cpp does (scramble, then
// fft_real; fft_r2c_3d, then fft_c2r_3d) and prints the results.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ct2_gpu_binding.hpp"

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

int main(int argc, char** argv) {
    if (argc > 1 && std::strcmp(argv[1], "fft") == 0) return fft_mode();
    const int n = argc > 1 ? std::atoi(argv[1]) : 512;
    particle_allocate(n);
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
