// host_tools.cpp -- host-side arithmetic of the event pipeline and step control.
// Built with g++ -ffp-contract=off (see host_tools.hpp for why not with the HIP compiler).
#include "host_tools.hpp"

#include <cmath>

namespace red {
namespace tools {


// tools.cpp:238-243
double calculate_radius(double m, double density)
{
    static const double four_pi_over_three = 4.1887902047863909846168578443727;
    return std::pow(1.0 / four_pi_over_three * m / density, 1.0 / 3.0);
}

// tools.cpp:263-279 (positions and velocities use the same weighting)
void mass_weighted_mean(double m1, double m2, const double* a, const double* b, double* out)
{
    const double M = m1 + m2;
    out[0] = (m1 * a[0] + m2 * b[0]) / M;
    out[1] = (m1 * a[1] + m2 * b[1]) / M;
    out[2] = (m1 * a[2] + m2 * b[2]) / M;
}

// tools.cpp:281-290: p = {mass, radius, density, cd}
void calc_physical_properties(const double* p1, const double* p2, double* p)
{
    const double volume = 4.188790204786391 * ((p1[1] * p1[1] * p1[1]) + (p2[1] * p2[1] * p2[1]));
    p[0] = p1[0] + p2[0];
    p[2] = p[0] / volume;
    p[1] = calculate_radius(p[0], p[2]);
    p[3] = p1[3];
}

// tools.cpp:473-484
void shift_into_range(double lower, double upper, double& value)
{
    const double range = upper - lower;
    while (value >= upper) value -= range;
    while (value < lower) value += range;
}

// tools.cpp:486-499 (the log form the reference defines and calls)
double log_asinh(double value)
{
    if (value > 0) return std::log(value + std::sqrt(value * value + 1));
    return -std::log(-value + std::sqrt(value * value + 1));
}

// tools.cpp:361-471; oe = {sma, ecc, inc, peri, node, mean}
void calculate_orbital_element(double mu, const double* r, const double* v, double* oe)
{
    const double sq2 = 1.0e-14, sq3 = 1.0e-14;
    const double v2 = (v[0] * v[0]) + (v[1] * v[1]) + (v[2] * v[2]);
    const double r2 = (r[0] * r[0]) + (r[1] * r[1]) + (r[2] * r[2]);
    const double rr = std::sqrt(r2);
    const double h = v2 / 2.0 - (mu / std::sqrt(r2));
    const double c3[3] = {r[1] * v[2] - r[2] * v[1], r[2] * v[0] - r[0] * v[2], r[0] * v[1] - r[1] * v[0]};
    const double l3[3] = {(-mu / rr) * r[0] + v[1] * c3[2] - v[2] * c3[1],
                          (-mu / rr) * r[1] + v[2] * c3[0] - v[0] * c3[2],
                          (-mu / rr) * r[2] + v[0] * c3[1] - v[1] * c3[0]};
    const double rv = r[0] * v[0] + r[1] * v[1] + r[2] * v[2];
    const double c = std::sqrt((c3[0] * c3[0]) + (c3[1] * c3[1]) + (c3[2] * c3[2]));
    const double l = std::sqrt((l3[0] * l3[0]) + (l3[1] * l3[1]) + (l3[2] * l3[2]));
    double e2 = 1.0 + 2.0 * (c * c) * h / (mu * mu);
    if (std::fabs(e2) < sq3) e2 = 0.0;
    const double e = std::sqrt(e2);
    const double a = -mu / (2.0 * h);
    const double cosi = c3[2] / c;
    const double sini = std::sqrt((c3[0] * c3[0]) + (c3[1] * c3[1])) / c;
    double incl = std::acos(cosi);
    if (incl < sq2) incl = 0.0;
    else if (incl > (PI_ - sq2)) incl = PI_;
    double node = 0.0;
    if (incl != 0.0 && incl != PI_) {
        const double tx = -c3[1] / (c * sini);
        const double ty = c3[0] / (c * sini);
        node = std::atan2(ty, tx);
        shift_into_range(0.0, 2.0 * PI_, node);
    }
    double E = 0.0, H = 0.0, shH = 0.0, peri = 0.0;
    if (e2 != 0.0) {
        double tx = (l3[0] * std::cos(node) + l3[1] * std::sin(node)) / l;
        double ty = (-l3[0] * std::sin(node) + l3[1] * std::cos(node)) / (l * cosi);
        peri = std::atan2(ty, tx);
        shift_into_range(0.0, 2.0 * PI_, peri);
        if (h >= 0.0) {
            shH = rv / (((a * mu) * (a * mu)) / ((c) * (c) * (c)) * e * std::pow(e2 - 1, 1.5));
            H = log_asinh(shH);
        } else {
            tx = 1.0 / e * (1.0 - rr / a);
            ty = rv / (std::sqrt(mu * a) * e);
            E = std::atan2(ty, tx);
            shift_into_range(0.0, 2.0 * PI_, E);
        }
    } else {
        peri = 0.0;
        E = std::atan2(r[1], r[0]);
        shift_into_range(0, 2.0 * PI_, E);
    }
    double M = 0.0;
    if (h >= 0.0) M = e * shH - H;
    else { M = E - e * std::sin(E); shift_into_range(0, 2.0 * PI_, M); }
    oe[0] = a; oe[1] = e; oe[2] = incl; oe[3] = peri; oe[4] = node; oe[5] = M;
}

// tools.cpp:292-315
int kepler_equation_solver(double ecc, double mean, double eps, double* E)
{
    if (ecc == 0.0 || mean == 0.0 || mean == PI_) { *E = mean; return 0; }
    *E = mean + ecc * (std::sin(mean)) / (1.0 - std::sin(mean + ecc) + std::sin(mean));
    double E1 = 0.0, error;
    int step = 0;
    do {
        E1 = *E - (*E - ecc * std::sin(*E) - mean) / (1.0 - ecc * std::cos(*E));
        error = std::fabs(E1 - *E);
        *E = E1;
    } while (error > eps && step++ <= 15);
    return step > 15 ? 1 : 0;
}

// tools.cpp:317-359; oe = {sma, ecc, inc, peri, node, mean}
int calculate_phase(double mu, const double* oe, double* r, double* v)
{
    const double ecc = oe[1];
    double E = 0.0;
    if (kepler_equation_solver(ecc, oe[5], 1.0e-14, &E) == 1) return 1;
    const double nu = 2.0 * std::atan(std::sqrt((1.0 + ecc) / (1.0 - ecc)) * std::tan(E / 2.0));
    const double p = oe[0] * (1.0 - (ecc * ecc));
    const double rr = p / (1.0 + ecc * std::cos(nu));
    const double kszi = rr * std::cos(nu), eta = rr * std::sin(nu);
    const double vKszi = -std::sqrt(mu / p) * std::sin(nu);
    const double vEta = std::sqrt(mu / p) * (ecc + std::cos(nu));
    const double cw = std::cos(oe[3]), sw = std::sin(oe[3]), cO = std::cos(oe[4]), sO = std::sin(oe[4]);
    const double ci = std::cos(oe[2]), si = std::sin(oe[2]);
    const double P[3] = {cw * cO - sw * sO * ci, cw * sO + sw * cO * ci, sw * si};
    const double Q[3] = {-sw * cO - cw * sO * ci, -sw * sO + cw * cO * ci, cw * si};
    for (int k = 0; k < 3; k++) {
        r[k] = kszi * P[k] + eta * Q[k];
        v[k] = vKszi * P[k] + vEta * Q[k];
    }
    return 0;
}

double controller_update(double* dt, double lambda, double errmax, double tol, double exponent)
{
    const double max_err = std::fabs(*dt * lambda * errmax);
    *dt *= 0.9 * std::pow(tol / max_err, exponent);
    return max_err;
}

double tolerance_step(double dt, double err, double tol, double exponent)
{
    if (!(err > 0.0)) return HUGE_VAL;
    return dt * std::pow(tol / err, exponent);
}

}  // namespace tools
}  // namespace red
