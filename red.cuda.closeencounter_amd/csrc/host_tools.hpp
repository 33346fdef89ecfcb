// host_tools.hpp -- host-side arithmetic of the event pipeline and step control.
//
// Merge physics and orbital elements as red.cuda/src/red.cuda.utility/tools.cpp defines
// them, and the step-size update of the controllers.  The definitions live in
// host_tools.cpp, which is compiled by the host C++ compiler (g++ -ffp-contract=off), not
// by the HIP compiler: its host pass evaluates some of these libm expressions
// differently in the last bit, and these results must equal the CPU oracle's.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

namespace red {
namespace tools {

constexpr double GAUSS = 1.720209895e-2;           // red_constants.h:21
constexpr double GAUSS2 = 2.959122082855911025e-4; // red_constants.h:22
constexpr double PI_ = 3.1415926535897932384626;   // red_macro.h:13

double calculate_radius(double m, double density);                                    // tools.cpp:238-243
void mass_weighted_mean(double m1, double m2, const double* a, const double* b, double* out);  // tools.cpp:263-279
void calc_physical_properties(const double* p1, const double* p2, double* p);          // tools.cpp:281-290
void calculate_orbital_element(double mu, const double* r, const double* v, double* oe); // tools.cpp:361-471
int kepler_equation_solver(double ecc, double mean, double eps, double* E);             // tools.cpp:292-315
int calculate_phase(double mu, const double* oe, double* r, double* v);                 // tools.cpp:317-359
// Controller update of the adaptive schemes (integrator.cu:255, int_rungekutta8.cu:776-777):
// returns |dt * lambda * errmax| and multiplies dt by 0.9 (tol / err)^exponent.
double controller_update(double* dt, double lambda, double errmax, double tol, double exponent);
// Tolerance-matching step of an attempt: dt (tol / err)^exponent (+inf when err <= 0).
double tolerance_step(double dt, double err, double tol, double exponent);

}  // namespace tools
}  // namespace red
