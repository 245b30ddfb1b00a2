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

// ------------------------------------------------------------ options, units
// options.cpp:51-61 (defaults) and :88-95 (derived fields).  Each derived
// field is formed in FP64; the dimensionless ones are rounded to FP32 as the
// reference's float members are, the code units stay FP64 (in FP32 the
// mass unit overflows at the default mass_res).
namespace {
// cgs constants of constants.hpp:14-39
constexpr double kC = 2.99792458e10;      // speed of light, cm/s
constexpr double kG = 6.67259e-8;         // gravitational constant
constexpr double kSigma = 5.67051e-5;     // Stefan-Boltzmann
constexpr double kH0 = 1e7 / 3.086e24;    // 100 km/s/Mpc in 1/s
constexpr double kMsun = 1.989e33;        // solar mass, g
constexpr double kTcmb = 2.73;            // K
constexpr double kNeff = 3.086;           // effective neutrino species
constexpr double kPi = 3.14159265358979323846;
}  // namespace

extern "C" int ct2_options_defaults(ct2_options* o) {
    if (!o) return ct2::set_error(CT2_ERR_INVALID, "ct2_options_defaults: null options");
    *o = ct2_options{};
    o->hsoft = (float)(1.0 / 50.0);
    o->bwidth = 3.0f;
    o->omega_m = 0.25f;
    o->mass_res = 1e9f;
    o->hubble = 0.7f;
    o->eta = 0.15f;
    o->Theta = 1.0f;
    o->Ngrid = 32;
    o->ppcell = 256;
    return CT2_OK;
}

extern "C" int ct2_options_derive(ct2_options* o) {
    if (!o) return ct2::set_error(CT2_ERR_INVALID, "ct2_options_derive: null options");
    if (o->Ngrid == 0 || !(o->mass_res > 0.f) || !(o->hubble > 0.f) || !(o->omega_m > 0.f))
        return ct2::set_error(CT2_ERR_INVALID, "ct2_options_derive: Ngrid, mass_res, hubble and omega_m must be positive");
    const uint64_t n = o->Ngrid * o->Ngrid * o->Ngrid;
    o->GM = (float)(1.0 / (double)n);
    o->nparts = n;
    o->code_to_g = kMsun * (double)o->mass_res;
    // mean density nparts * code_to_g / code_to_cm^3 = omega_m * 3 H^2 / (8 pi G)
    const double h = kH0 * (double)o->hubble;
    o->code_to_cm = std::cbrt(o->code_to_g * (8.0 * kPi) * (double)n * kG / (3.0 * (double)o->omega_m * h * h));
    o->code_to_s = o->code_to_cm / kC;
    // photons (energy density 4 sigma T^4 / c) plus N_eff neutrino species at
    // 7/8 (4/11)^(4/3) of the photon density each, over 3 H^2 c^2 / (8 pi G)
    const double T = kTcmb * (double)o->Theta;
    const double nu = 1.0 + kNeff * (7.0 / 8.0) * std::pow(4.0 / 11.0, 4.0 / 3.0);
    o->omega_r = (float)(32.0 * kPi / 3.0 * kG * kSigma * nu * (T * T) * (T * T) / (kC * kC * kC * h * h));
    o->omega_lam = (float)(1.0 - (double)o->omega_r - (double)o->omega_m);
    return CT2_OK;
}

extern "C" int ct2_options_cosmology(const ct2_options* o, ct2_cosmology* c) {
    if (!o || !c) return ct2::set_error(CT2_ERR_INVALID, "ct2_options_cosmology: null argument");
    c->H = kH0 * o->code_to_s * (double)o->hubble;
    c->omega_m = o->omega_m;
    c->omega_r = o->omega_r;
    c->omega_lam = o->omega_lam;
    return CT2_OK;
}
