// cosmology.cpp -- FP64 scale-factor and time integrals (host code).
//
// The functions of cosmology.hpp:13-16 (cosmology.cpp:11-102) with the
// cosmological parameters passed explicitly instead of read from
// get_options().  H is the Hubble constant in code time units (the reference
// forms it as H0 * code_to_s * hubble, cosmology.cpp:14).  Plus the two
// leapfrog factors the stepping loop needs: int dt/a (kick) and int dt/a^2
// (drift) between two scale factors.
#include <cmath>

#include "ct2hip.h"

namespace ct2 {
int set_error(int code, const char* fmt, ...);

namespace {

inline double sq(double a) { return a * a; }

double ddot_a(const ct2_cosmology* c, double a) {
    return a * sq(c->H) * (-c->omega_r / sq(sq(a)) - 0.5 * c->omega_m / (sq(a) * a) + c->omega_lam);
}

// cosmology.cpp:11-36: one three-stage step in the reference's time variable
// (da = adot * a * dt).
void update0(const ct2_cosmology* c, double& adotdot, double& adot, double& a, double dt) {
    adotdot = ddot_a(c, a);
    const double da1 = adot * a * dt;
    const double dadot1 = adotdot * a * dt;
    a += da1;
    adot += dadot1;

    adotdot = ddot_a(c, a);
    const double da2 = adot * a * dt;
    const double dadot2 = adotdot * a * dt;
    a += -da1 + 0.25 * (da1 + da2);
    adot += -dadot1 + 0.25 * (dadot1 + dadot2);

    adotdot = ddot_a(c, a);
    const double da3 = adot * a * dt;
    const double dadot3 = adotdot * a * dt;
    a += -0.25 * (da1 + da2) + (1.0 / 6.0) * (da1 + da2 + 4.0 * da3);
    adot += -0.25 * (dadot1 + dadot2) + (1.0 / 6.0) * (dadot1 + dadot2 + 4.0 * dadot3);
}

// int f(a) dt from a0 to a1 as int f(a) a / adot d(ln a), trapezoid rule,
// doubling the panel count until two estimates agree to `tol` (the
// convergence rule of cosmology.cpp:58-81).
template <class F>
double integrate_lna(const ct2_cosmology* c, double a0, double a1, F f, double tol) {
    int N = 2;
    double t = 0.0, tlast;
    do {
        tlast = t;
        t = 0.0;
        const double l0 = std::log(a0), l1 = std::log(a1), dl = (l1 - l0) / N;
        double a = std::exp(l0);
        double g = f(a) * a / ct2_cosmos_dadt(c, a);
        for (int i = 0; i < N; ++i) {
            t += 0.5 * g * dl;
            a = std::exp(l0 + dl * (i + 1));
            g = f(a) * a / ct2_cosmos_dadt(c, a);
            t += 0.5 * g * dl;
        }
        N *= 2;
    } while (std::fabs(tlast - t) > tol * std::fabs(t) && N <= (1 << 24));
    return t;
}

}  // namespace
}  // namespace ct2

using namespace ct2;

extern "C" int ct2_cosmos_update(const ct2_cosmology* c, double* adotdot, double* adot, double* a, double dt0) {
    if (!c || !adotdot || !adot || !a) return set_error(CT2_ERR_INVALID, "null argument");
    const int N = 128;  // cosmology.cpp:39
    const double dt = dt0 / N;
    for (int i = 0; i < N; ++i) update0(c, *adotdot, *adot, *a, dt);
    return CT2_OK;
}

extern "C" double ct2_cosmos_dadt(const ct2_cosmology* c, double a) {
    return c->H * a * std::sqrt(c->omega_r / sq(sq(a)) + c->omega_m / (a * sq(a)) + c->omega_lam);
}

extern "C" double ct2_cosmos_conformal_time(const ct2_cosmology* c, double a0, double a1) {
    return integrate_lna(c, a0, a1, [](double a) { return 1.0 / a; }, 1e-9);
}

extern "C" double ct2_cosmos_time(const ct2_cosmology* c, double a0, double a1) {
    // cosmology.cpp:84-102: a fixed 2^20-panel trapezoid in ln a
    const int N = 1024 * 1024;
    const double l0 = std::log(a0), l1 = std::log(a1), dl = (l1 - l0) / N;
    double t = 0.0;
    double a = std::exp(l0);
    double g = a / ct2_cosmos_dadt(c, a);
    for (int i = 0; i < N; ++i) {
        t += 0.5 * g * dl;
        a = std::exp(l0 + dl * (i + 1));
        g = a / ct2_cosmos_dadt(c, a);
        t += 0.5 * g * dl;
    }
    return t;
}

extern "C" double ct2_cosmos_drift_factor(const ct2_cosmology* c, double a0, double a1) {
    return integrate_lna(c, a0, a1, [](double a) { return 1.0 / (a * a); }, 1e-12);
}

extern "C" double ct2_cosmos_kick_factor(const ct2_cosmology* c, double a0, double a1) {
    return integrate_lna(c, a0, a1, [](double a) { return 1.0 / a; }, 1e-12);
}
