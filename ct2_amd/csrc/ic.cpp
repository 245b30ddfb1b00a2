// ic.cpp -- seeded synthetic particle sets (host code).
//
// Replaces particles_init (particle.cpp:364-374), whose libc rand() stream
// is not reproducible across platforms.  Every random number is a pure
// function of (seed, particle, draw) through SplitMix64, so the sets are
// identical on any host and independent of the thread count.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "ct2hip.h"

namespace ct2 {
int set_error(int code, const char* fmt, ...);

namespace {

inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Counter-based stream: draw k of particle p under seed s.
struct Stream {
    uint64_t key;
    uint64_t k = 0;
    Stream(uint64_t seed, uint64_t p) : key(splitmix64(seed * 0xD1B54A32D192ED03ull ^ splitmix64(p))) {}
    uint64_t next() { return splitmix64(key + 0x632BE59BD9B4E019ull * (++k)); }
    // uniform on (0,1), 53 bits
    double uniform() { return ((double)(next() >> 11) + 0.5) * 0x1p-53; }
    // Box-Muller, one value per call (the partner is discarded for simplicity)
    double normal() {
        const double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

// [0,1) coordinate -> fixed32, RNE, periodic wrap.
inline uint32_t to_fixed(double x) {
    x -= std::floor(x);
    const double v = std::nearbyint(x * 4294967296.0);
    return (uint32_t)(uint64_t)(int64_t)v;  // 2^32 wraps to 0
}

template <class F>
void parallel_for(size_t n, F f) {
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const size_t chunk = (n + hw - 1) / hw;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < hw; ++t) {
        const size_t a = t * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        th.emplace_back([=] {
            for (size_t i = a; i < b; ++i) f(i);
        });
    }
    for (auto& t : th) t.join();
}

}  // namespace
}  // namespace ct2

using namespace ct2;

extern "C" int ct2_ic_lattice(int n, double sigma_cells, uint64_t seed, uint32_t* x, uint32_t* y, uint32_t* z) {
    if (n <= 0 || n > 1290) return set_error(CT2_ERR_INVALID, "lattice side %d out of range", n);
    if (!x || !y || !z) return set_error(CT2_ERR_INVALID, "null output");
    if (!(sigma_cells >= 0.0)) return set_error(CT2_ERR_INVALID, "negative sigma");
    const size_t N = (size_t)n * n * n;
    const double h = 1.0 / n;
    parallel_for(N, [&](size_t p) {
        const size_t i = p / ((size_t)n * n), j = (p / n) % n, k = p % n;
        Stream s(seed, p);
        const double gx = sigma_cells > 0 ? s.normal() : 0.0;
        const double gy = sigma_cells > 0 ? s.normal() : 0.0;
        const double gz = sigma_cells > 0 ? s.normal() : 0.0;
        x[p] = to_fixed((i + 0.5) * h + sigma_cells * h * gx);
        y[p] = to_fixed((j + 0.5) * h + sigma_cells * h * gy);
        z[p] = to_fixed((k + 0.5) * h + sigma_cells * h * gz);
    });
    return CT2_OK;
}

extern "C" int ct2_ic_plummer_mix(size_t n, size_t n_sph, int nspheres, double a, double rmax, uint64_t seed,
                                  uint32_t* x, uint32_t* y, uint32_t* z) {
    if (n == 0 || n_sph > n) return set_error(CT2_ERR_INVALID, "bad particle counts");
    if (n_sph > 0 && (nspheres <= 0 || !(a > 0.0) || !(rmax > 0.0)))
        return set_error(CT2_ERR_INVALID, "bad Plummer sphere parameters");
    if (!x || !y || !z) return set_error(CT2_ERR_INVALID, "null output");
    // sphere centres: uniform, drawn from the stream of particle index 2^63 + s
    std::vector<double> cen(3 * (size_t)std::max(nspheres, 1));
    for (int s = 0; s < nspheres; ++s) {
        Stream st(seed, (1ull << 63) + (uint64_t)s);
        for (int d = 0; d < 3; ++d) cen[3 * s + d] = st.uniform();
    }
    parallel_for(n, [&](size_t p) {
        Stream st(seed, p);
        double q[3];
        if (p < n_sph) {
            const int s = (int)(p % (size_t)nspheres);
            // Plummer radius by inversion of M(<r) = r^3/(r^2+a^2)^{3/2}, redrawn beyond rmax
            double r;
            do {
                const double u = st.uniform();
                r = a / std::sqrt(std::pow(u, -2.0 / 3.0) - 1.0);
            } while (r > rmax);
            const double cz = 2.0 * st.uniform() - 1.0;
            const double ph = 2.0 * M_PI * st.uniform();
            const double sz = std::sqrt(std::max(0.0, 1.0 - cz * cz));
            q[0] = cen[3 * s + 0] + r * sz * std::cos(ph);
            q[1] = cen[3 * s + 1] + r * sz * std::sin(ph);
            q[2] = cen[3 * s + 2] + r * cz;
        } else {
            for (int d = 0; d < 3; ++d) q[d] = st.uniform();
        }
        x[p] = to_fixed(q[0]);
        y[p] = to_fixed(q[1]);
        z[p] = to_fixed(q[2]);
    });
    return CT2_OK;
}
