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
#include <string>
#include <vector>

#include "../../include/rednative.h"

namespace red {

// Scatter of one merged body's new state into the device arrays (k_apply_updates).
struct BodyUpdate {
    int idx;
    int id;
    double r[4];
    double v[4];
    double mass, radius;
};

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
// returns |dt * lambda * errmax| and multiplies dt by 0.9 (tol / err)^exponent, the power
// correctly rounded (step_pow.hpp, the ensemble's device controller computes the same bits).
double controller_update(double* dt, double lambda, double errmax, double tol, double exponent);
// Tolerance-matching step of an attempt: dt (tol / err)^exponent (+inf when err <= 0).
double tolerance_step(double dt, double err, double tol, double exponent);

}  // namespace tools

// ---- the host side of the event pipeline (pp_disk.cu:683-1202), shared by single disks and
// the members of an ensemble, so both handle events with the same code ------------------
namespace events {

enum { CNT_TOTAL = 0, CNT_LAST_CLEAR = 1, CNT_LAST_STEP = 2 };

// Event counters of one disk (pp_disk.h:129-133): total, since the last clear / rebuild,
// in the last step.
struct Counters {
    int hit_centrum[3] = {0, 0, 0}, ejection[3] = {0, 0, 0}, collision[3] = {0, 0, 0};
    int close_encounter[3] = {0, 0, 0}, event[3] = {0, 0, 0};
    void increment(int* cnt);            // pp_disk.cu:1059-1067
    void set(int field, int value);      // pp_disk.cu:1069-1076
};

// Host mirror of one disk's per-body parameters; indices are the disk's own (body 0 first).
struct BodyView {
    double* mass;
    double* radius;
    double* density;
    double* cd;
    int* id;
    const int* type;
    int n_massive;
};

// Mass of the first star among the massive bodies, -1 if there is none.
double mass_of_star(const BodyView& B);
// Canonical order of pair records: (attempt, stage, sink, source).
bool event_key_less(const red_event& a, const red_event& b);
// Puts records into canonical order, stably (the order of equal keys kept); sorts small keys
// and moves each record once.
void canonical_sort(std::vector<red_event>& ev);
// The same order for a list of records held elsewhere (records stay where they are).
void canonical_sort(std::vector<const red_event*>& ev);
// Records the device kept one per pair (ensembles), the stage of each pair's first record in
// the attempt field's upper 16 bits: the order create_sp_events gives them.
void kept_order(std::vector<const red_event*>& ev);
// pp_disk.cu:1111-1149: one record per (id1, id2) in first-occurrence order, the earliest t
// kept (the first of equal times).
void create_sp_events(const std::vector<red_event>& raw, std::vector<red_event>& sp);
// The same for a list of records held elsewhere; the sp list is appended to out.
void create_sp_events(const std::vector<const red_event*>& raw, std::vector<red_event>& out);
// pp_disk.cu:1151-1202 for the record's pair (idx1 survives): the merge on the host mirror
// and the two device updates it needs.  kernel_negated: the merger's id was already made
// negative by the detection pass.  Returns RED_OK or RED_ERR_UNSUPPORTED (with msg).
int handle_collision_pair(BodyView& B, red_event* e, bool kernel_negated, std::vector<BodyUpdate>& ups,
                          std::string& msg);
// Astro-centric elements of a close encounter from the recorded stage states, and the
// bodies' parameters (pp_disk.cu:494-566).
void close_encounter_elements(const BodyView& B, red_event* e);
// The ejection / hit-centrum pass on the records of the scan (pp_disk.cu:683-715, 1091-1109):
// ev is put into body order and becomes the pass's sp list.  The scan has negated the ids on
// the device; the host mirror follows here.
int ejection_pass(BodyView& B, std::vector<red_event>& ev, Counters& C, std::vector<BodyUpdate>& ups,
                  std::string& msg);
// The collision pass on the canonical raw list (pp_disk.cu:717-738, 1078-1149): collisions
// merged, close-encounter records of the Hill test reported.
int collision_pass(BodyView& B, const std::vector<red_event>& raw, std::vector<red_event>& sp, Counters& C,
                   std::vector<BodyUpdate>& ups, std::string& msg);
// The same on the pass's sp list (n records) already made by create_sp_events.
int collision_sp_pass(BodyView& B, red_event* sp, size_t n, Counters& C, std::vector<BodyUpdate>& ups,
                      std::string& msg);
// Close-encounter mode (pp_disk.cu:740-770, inner_steps): every raw record reported.
void close_encounter_pass(const BodyView& B, const std::vector<red_event>& raw, std::vector<red_event>& sp,
                          Counters& C);
// The same on sp (n records) = the raw records.
void close_encounter_sp_pass(const BodyView& B, red_event* sp, size_t n, Counters& C);
// The ejection / hit-centrum predicate of pp_disk.cu:72-142 on the host (the same IEEE
// operations as the device's, so the same answer), and the scan of one disk's state:
// records for every active non-star body that meets it, ids negated, as the device scan.
int ejection_name(const double* q, double ej, double hc);
void ejection_scan(int n, const double* r4, const double* v4, int* id, const int* type, double t, double ej,
                   double hc, std::vector<red_event>& out);

}  // namespace events
}  // namespace red
