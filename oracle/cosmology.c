/*
 * cosmology.c -- FP64 oracle for the scale-factor and time integrals,
 * restating src/cosmology.cpp:11-102 with the options read from an explicit
 * parameter block instead of get_options().
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <math.h>

#include "oracle.h"

static inline double sq(double a) { return a * a; }

/* a * H^2 * (-Or/a^4 - Om/(2 a^3) + OL), cosmology.cpp:18,24,30 */
static double addot(const orc_cosmo* c, double a) {
    return a * sq(c->H) * (-c->omega_r / sq(sq(a)) - 0.5 * c->omega_m / (sq(a) * a) + c->omega_lam);
}

/* cosmology.cpp:11-36: one three-stage step of (a, adot) in the time
 * variable of the reference (d a = adot * a * dt). */
static void update0(const orc_cosmo* c, double* adotdot, double* adot, double* a, double dt) {
    *adotdot = addot(c, *a);
    const double da1 = *adot * *a * dt;
    const double dadot1 = *adotdot * *a * dt;
    *a += da1;
    *adot += dadot1;

    *adotdot = addot(c, *a);
    const double da2 = *adot * *a * dt;
    const double dadot2 = *adotdot * *a * dt;
    *a += -da1 + 0.25 * (da1 + da2);
    *adot += -dadot1 + 0.25 * (dadot1 + dadot2);

    *adotdot = addot(c, *a);
    const double da3 = *adot * *a * dt;
    const double dadot3 = *adotdot * *a * dt;
    *a += -0.25 * (da1 + da2) + (1.0 / 6.0) * (da1 + da2 + 4.0 * da3);
    *adot += -0.25 * (dadot1 + dadot2) + (1.0 / 6.0) * (dadot1 + dadot2 + 4.0 * dadot3);
}

/* cosmology.cpp:38-44: 128 sub-steps */
void orc_cosmos_update(const orc_cosmo* c, double* adotdot, double* adot, double* a, double dt0) {
    const int N = 128;
    const double dt = dt0 / N;
    for (int i = 0; i < N; i++) update0(c, adotdot, adot, a, dt);
}

/* cosmology.cpp:46-52: da/dt = H a sqrt(Or/a^4 + Om/a^3 + OL) */
double orc_cosmos_dadt(const orc_cosmo* c, double a) {
    return c->H * a * sqrt(c->omega_r / sq(sq(a)) + c->omega_m / (a * sq(a)) + c->omega_lam);
}

/* cosmology.cpp:58-81: int dt/a from a0 to a1, trapezoid rule in ln a,
 * doubling the count until successive sums agree to 1e-9. */
double orc_cosmos_conformal_time(const orc_cosmo* c, double a0, double a1) {
    int N = 2;
    double t = 0.0, tlast;
    do {
        tlast = t;
        t = 0.0;
        const double loga0 = log(a0), loga1 = log(a1);
        const double dloga = (loga1 - loga0) / N;
        double a = exp(loga0);
        double dadtau = a * orc_cosmos_dadt(c, a);
        for (int i = 0; i < N; i++) {
            t += 0.5 * a / dadtau * dloga;
            a = exp(loga0 + dloga * (i + 1));
            dadtau = a * orc_cosmos_dadt(c, a);
            t += 0.5 * a / dadtau * dloga;
        }
        N *= 2;
    } while (fabs(tlast - t) > 1e-9 * t);
    return t;
}

/* cosmology.cpp:84-102: int dt from a0 to a1, 2^20 trapezoids in ln a. */
double orc_cosmos_time(const orc_cosmo* c, double a0, double a1) {
    const int N = 1024 * 1024;
    const double loga0 = log(a0), loga1 = log(a1);
    const double dloga = (loga1 - loga0) / N;
    double t = 0.0;
    double a = exp(loga0);
    double dadt = orc_cosmos_dadt(c, a);
    for (int i = 0; i < N; i++) {
        t += 0.5 * a / dadt * dloga;
        a = exp(loga0 + dloga * (i + 1));
        dadt = orc_cosmos_dadt(c, a);
        t += 0.5 * a / dadt * dloga;
    }
    return t;
}

/* options.cpp:88-95 in the reference's order of evaluation, each right-hand
 * side in FP64 (the float operands promoted); GM and the omegas stored to
 * FP32 as there, the code units kept in FP64.  Constants of
 * constants.hpp:14-39 (cgs). */
void orc_options_derive(uint64_t ngrid, float mass_res, float hubble, float omega_m, float theta, double* out) {
    const double c = 2.99792458e10, G = 6.67259e-8, sigma = 5.67051e-5, H0 = 1e7 / 3.086e24, M0 = 1.989e33;
    const double Neff = 3.086;
    const uint64_t nparts = ngrid * ngrid * ngrid;
    const float GM = (float)(1.0 / (double)nparts);
    const double code_to_g = M0 * mass_res;
    const double code_to_cm =
        pow(code_to_g * (8.0 * M_PI) * (double)nparts * G / (3.0 * omega_m * sq(H0 * hubble)), 1.0 / 3.0);
    const double code_to_s = code_to_cm / c;
    const float omega_r = (float)(32.0 * M_PI / 3.0 * G * sigma * (1 + Neff * (7. / 8.0) * pow(4. / 11., 4. / 3.)) *
                                  pow(H0, -2) * pow(c, -3) * pow(2.73 * theta, 4) * pow(hubble, -2));
    const float omega_lam = (float)(1.0 - omega_r - omega_m);
    out[0] = GM;
    out[1] = code_to_g;
    out[2] = code_to_cm;
    out[3] = code_to_s;
    out[4] = omega_r;
    out[5] = omega_lam;
    out[6] = H0 * code_to_s * hubble;
}
