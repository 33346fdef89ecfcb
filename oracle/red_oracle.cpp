// red_oracle.cpp -- FP64 CPU restatement of red.cuda's integration path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke
// entry point and the CPU-baseline leg of the benchmark may load it; the
// product library never links, loads or calls it.
//
// It restates, single-threaded and in the same floating-point operation order,
// the CPU code path of the reference (compile with -ffp-contract=off so that no
// multiply-add is fused, matching nvcc -fmad=false of the reference build):
//
//   force + collision events   red.cuda/src/red.cuda/pp_disk.cu:404-607
//   interaction bounds         red.cuda/src/red.cuda/number_of_bodies.cpp:106-258
//   ejection / hit centrum     red.cuda/src/red.cuda/pp_disk.cu:909-976
//   event pipeline             red.cuda/src/red.cuda/pp_disk.cu:1059-1202
//   compaction                 red.cuda/src/red.cuda/pp_disk.cu:1324-1409,
//                              number_of_bodies.cpp:27-104
//   error norm, counters       red.cuda/src/red.cuda/integrator.cu:211-263
//   Euler                      red.cuda/src/red.cuda/int_euler.cu:36-141
//   RK2                        red.cuda/src/red.cuda/int_rungekutta2.cu:13-152
//   RK4 (+embedded estimate)   red.cuda/src/red.cuda/int_rungekutta4.cu:27-273
//   RKF7(8)                    red.cuda/src/red.cuda/int_rungekutta8.cu:27-47,336-797
//   Cash-Karp RK5(4)           red.cuda/src/red.cuda/int_rungekutta5.cu:24-93,202-282
//                              (GPU-only and unwired in the reference; restated with
//                              the documented fixes listed at rk5_step below)
//   merge physics, transforms  red.cuda/src/red.cuda.utility/tools.cpp:109-290
//   orbital elements           red.cuda/src/red.cuda.utility/tools.cpp:292-500
//
// Two flags select between the reference's CPU and GPU behaviour where they differ:
//   stage_time_events : 1 = every launch stamps events with the stage time (GPU kernel,
//                       pp_disk.cu:265-279); 0 = NSI/NI launches stamp the step-start
//                       time (CPU quirk, pp_disk.cu:574-582).
//   hc_reactivate     : 1 = a hit-centrum merger's id is negated twice (kernel + merge),
//                       which re-activates it as a massless body (CPU path,
//                       pp_disk.cu:959,1178); 0 = it stays inactive (GPU path, where the
//                       merge negates the stale host copy).
// padded = 1 restates the padded-storage kernel (-ups, pp_disk.cu:144-215): no id test on
// sinks, sources or pairs.  Its padding particles (massless, id 0, pp_disk.cu:1550-1571)
// add exact zeros to every sum and never meet a pair test, so none are created here.
// The step controllers' factor pow(tol / err, 1/(order+1)) is taken correctly rounded
// (step_factor below, through libquadmath's powq) instead of from the platform's libm, whose
// last bit differs between C runtimes (the reference is built with MSVC's and nvcc's).
// Additions the reference does not have (marked NEW): total energy and angular momentum,
// an OpenMP sink loop whose results are bit-identical to the serial one, and an optional
// mutual-Hill-radius close-encounter criterion.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <quadmath.h>

namespace {

// x^y rounded to nearest: powq carries 113 bits, so rounding its result to a double is the
// correct rounding unless x^y lies within about 2^-110 of a midpoint between two doubles
// (int_rungekutta8.cu:777, int_rungekutta4.cu:253 call the C runtime's pow here).
double step_factor(double x, double y) { return (double)powq((__float128)x, (__float128)y); }

const double Gauss  = 1.720209895e-2;            // red_constants.h:21
const double Gauss2 = 2.959122082855911025e-4;   // red_constants.h:22
const double PI_    = 3.1415926535897932384626;  // red_macro.h:13

enum { BT_STAR, BT_GIANT, BT_ROCKY, BT_PROTO, BT_SUPERPL, BT_PL, BT_TP, BT_PAD };
enum { EV_NONE, EV_HIT_CENTRUM, EV_EJECTION, EV_CLOSE_ENCOUNTER, EV_COLLISION };
enum { TH_HC, TH_EJ, TH_RADII, TH_HC2, TH_EJ2 };
enum { INT_EULER, INT_RK2, INT_RK4, INT_RK5, INT_RKF78 };
enum { FC_ASTRO, FC_BARY };
enum { CNT_TOTAL, CNT_LAST_CLEAR, CNT_LAST_STEP };

struct vec_t   { double x, y, z, w; };
struct param_t { double mass, radius, density, cd; };
struct md_t    { int id, body_type, mig_type; double mig_stop_at; };
struct oe_t    { double sma, ecc, inc, peri, node, mean; };

}  // namespace

extern "C" {
// Flat event record; field layout is shared with the product's red_event.
typedef struct orc_event {
    int    name;
    int    id1, idx1, id2, idx2;
    int    attempt, stage;      // attempt number within the step, stage index r of calc_dydx
    int    sink, source;        // (i, j) of the detecting pair before the mass ordering
    int    pad;
    double t, d;
    double p1[4], p2[4], ps[4]; // mass radius density cd
    double r1[3], v1[3], r2[3], v2[3], rs[3], vs[3];
    double oe1[6], oe2[6], oes[6];
} orc_event;

typedef struct orc_params {
    int    integrator;
    int    adaptive;
    double tolerance;
    int    error_check_for_tp;
    int    close_encounter;
    int    inner_steps;
    int    frame_center;
    int    stage_time_events;
    int    hc_reactivate;
    int    n_threads;
    double hit_centrum, ejection, radii_enhance_factor;
    double encounter_hill_factor;   // NEW: 0 = off
    double dt_day;                  // initial step in days (reference hard-codes 0.001, main.cu:141)
    int    padded;                  // the padded-storage kernel's semantics (pp_disk.cu:144-215)
} orc_params;
}

namespace {

static void set3(double* d, const vec_t& v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; }
static void set4(double* d, const param_t& p) { d[0] = p.mass; d[1] = p.radius; d[2] = p.density; d[3] = p.cd; }
static vec_t get3(const double* d) { vec_t v = {d[0], d[1], d[2], 0.0}; return v; }
static param_t get4(const double* d) { param_t p = {d[0], d[1], d[2], d[3]}; return p; }
static void set6(double* d, const oe_t& o) { d[0] = o.sma; d[1] = o.ecc; d[2] = o.inc; d[3] = o.peri; d[4] = o.node; d[5] = o.mean; }

// ---- tools.cpp restatements ---------------------------------------------------------
static void shift_into_range(double lower, double upper, double& value)  // tools.cpp:473-484
{
    double range = upper - lower;
    while (value >= upper) value -= range;
    while (value < lower) value += range;
}

static double tools_asinh(double value)  // tools.cpp:486-499 (log form, not libm asinh)
{
    if (value > 0) return log(value + sqrt(value * value + 1));
    return -log(-value + sqrt(value * value + 1));
}

// NEW (Hill criterion): cube root by bit-seeded Newton iterations using only +, -, *, /,
// so that every platform rounds it identically.
static double cbrt_ieee(double x)
{
    if (!(x > 0.0)) return 0.0;
    uint64_t b;
    memcpy(&b, &x, 8);
    int64_t eb = (int64_t)((b >> 52) & 0x7ff) - 1023;
    int64_t q = eb >= 0 ? eb / 3 : -((-eb + 2) / 3);
    uint64_t yb = (uint64_t)(q + 1023) << 52;
    double y;
    memcpy(&y, &yb, 8);
    for (int k = 0; k < 60; k++) {
        double yn = y - (y * y * y - x) / (3.0 * y * y);
        if (yn == y) break;
        y = yn;
    }
    return y;
}

static double calculate_radius(double m, double density)  // tools.cpp:238-243
{
    static double four_pi_over_three = 4.1887902047863909846168578443727;
    return pow(1.0 / four_pi_over_three * m / density, 1.0 / 3.0);
}

static void calc_position_after_collision(double m1, double m2, const double* r1, const double* r2, double* r)  // tools.cpp:263-270
{
    const double M = m1 + m2;
    r[0] = (m1 * r1[0] + m2 * r2[0]) / M;
    r[1] = (m1 * r1[1] + m2 * r2[1]) / M;
    r[2] = (m1 * r1[2] + m2 * r2[2]) / M;
}

static void calc_physical_properties(const param_t& p1, const param_t& p2, param_t& p)  // tools.cpp:281-290
{
    double volume = 4.188790204786391 * ((p1.radius * p1.radius * p1.radius) + (p2.radius * p2.radius * p2.radius));
    p.mass = p1.mass + p2.mass;
    p.density = p.mass / volume;
    p.radius = calculate_radius(p.mass, p.density);
    p.cd = p1.cd;
}

static int kepler_equation_solver(double ecc, double mean, double eps, double* E)  // tools.cpp:292-315
{
    if (ecc == 0.0 || mean == 0.0 || mean == PI_) { *E = mean; return 0; }
    *E = mean + ecc * (sin(mean)) / (1.0 - sin(mean + ecc) + sin(mean));
    double E1 = 0.0, error;
    int step = 0;
    do {
        E1 = *E - (*E - ecc * sin(*E) - mean) / (1.0 - ecc * cos(*E));
        error = fabs(E1 - *E);
        *E = E1;
    } while (error > eps && step++ <= 15);
    return step > 15 ? 1 : 0;
}

static int calculate_phase(double mu, const oe_t* oe, vec_t* rVec, vec_t* vVec)  // tools.cpp:317-359
{
    double ecc = oe->ecc, E = 0.0;
    if (kepler_equation_solver(ecc, oe->mean, 1.0e-14, &E) == 1) return 1;
    double v = 2.0 * atan(sqrt((1.0 + ecc) / (1.0 - ecc)) * tan(E / 2.0));
    double p = oe->sma * (1.0 - (ecc * ecc));
    double r = p / (1.0 + ecc * cos(v));
    double kszi = r * cos(v), eta = r * sin(v);
    double vKszi = -sqrt(mu / p) * sin(v);
    double vEta = sqrt(mu / p) * (ecc + cos(v));
    double cw = cos(oe->peri), sw = sin(oe->peri), cO = cos(oe->node), sO = sin(oe->node);
    double ci = cos(oe->inc), si = sin(oe->inc);
    vec_t P, Q;
    P.x = cw * cO - sw * sO * ci;  P.y = cw * sO + sw * cO * ci;  P.z = sw * si;
    Q.x = -sw * cO - cw * sO * ci; Q.y = -sw * sO + cw * cO * ci; Q.z = cw * si;
    rVec->x = kszi * P.x + eta * Q.x; rVec->y = kszi * P.y + eta * Q.y; rVec->z = kszi * P.z + eta * Q.z;
    vVec->x = vKszi * P.x + vEta * Q.x; vVec->y = vKszi * P.y + vEta * Q.y; vVec->z = vKszi * P.z + vEta * Q.z;
    rVec->w = vVec->w = 0.0;
    return 0;
}

static int calculate_orbital_element(const double mu, oe_t* oe, const vec_t rVec, const vec_t vVec)  // tools.cpp:361-471
{
    const double sq2 = 1.0e-14, sq3 = 1.0e-14;
    double v2 = (vVec.x * vVec.x) + (vVec.y * vVec.y) + (vVec.z * vVec.z);
    double r2 = (rVec.x * rVec.x) + (rVec.y * rVec.y) + (rVec.z * rVec.z);
    double r = sqrt(r2);
    double h = v2 / 2.0 - (mu / sqrt(r2));
    vec_t cVec = {rVec.y * vVec.z - rVec.z * vVec.y, rVec.z * vVec.x - rVec.x * vVec.z, rVec.x * vVec.y - rVec.y * vVec.x, 0.0};
    vec_t lVec = {(-mu / r) * rVec.x + vVec.y * cVec.z - vVec.z * cVec.y,
                  (-mu / r) * rVec.y + vVec.z * cVec.x - vVec.x * cVec.z,
                  (-mu / r) * rVec.z + vVec.x * cVec.y - vVec.y * cVec.x, 0.0};
    double rv = rVec.x * vVec.x + rVec.y * vVec.y + rVec.z * vVec.z;
    double c = sqrt((cVec.x * cVec.x) + (cVec.y * cVec.y) + (cVec.z * cVec.z));
    double l = sqrt((lVec.x * lVec.x) + (lVec.y * lVec.y) + (lVec.z * lVec.z));
    double e2 = 1.0 + 2.0 * (c * c) * h / (mu * mu);
    if (fabs(e2) < sq3) e2 = 0.0;
    double e = sqrt(e2);
    double a = -mu / (2.0 * h);
    double cosi = cVec.z / c;
    double sini = sqrt((cVec.x * cVec.x) + (cVec.y * cVec.y)) / c;
    double incl = acos(cosi);
    if (incl < sq2) incl = 0.0;
    else if (incl > (PI_ - sq2)) incl = PI_;
    double node = 0.0;
    if (incl != 0.0 && incl != PI_) {
        double tmpx = -cVec.y / (c * sini);
        double tmpy = cVec.x / (c * sini);
        node = atan2(tmpy, tmpx);
        shift_into_range(0.0, 2.0 * PI_, node);
    }
    double E = 0.0, H = 0.0, shH = 0.0, peri = 0.0;
    if (e2 != 0.0) {
        double tmpx = (lVec.x * cos(node) + lVec.y * sin(node)) / l;
        double tmpy = (-lVec.x * sin(node) + lVec.y * cos(node)) / (l * cosi);
        peri = atan2(tmpy, tmpx);
        shift_into_range(0.0, 2.0 * PI_, peri);
        if (h >= 0.0) {
            shH = rv / (((a * mu) * (a * mu)) / ((c) * (c) * (c)) * e * pow(e2 - 1, 1.5));
            H = tools_asinh(shH);
        } else {
            tmpx = 1.0 / e * (1.0 - r / a);
            tmpy = rv / (sqrt(mu * a) * e);
            E = atan2(tmpy, tmpx);
            shift_into_range(0.0, 2.0 * PI_, E);
        }
    } else {
        peri = 0.0;
        E = atan2(rVec.y, rVec.x);
        shift_into_range(0, 2.0 * PI_, E);
    }
    double M = 0.0;
    if (h >= 0.0) M = e * shH - H;
    else { M = E - e * sin(E); shift_into_range(0, 2.0 * PI_, M); }
    oe->sma = a; oe->ecc = e; oe->inc = incl; oe->peri = peri; oe->node = node; oe->mean = M;
    return 0;
}

struct Bound { int sink0, sink1, src0, src1; };

// ---- pp_disk (CPU path) ---------------------------------------------------------------
// pp_disk.cu:1111-1149: one record per (id1, id2) pair in first-occurrence order, the
// record with the earliest t kept (the first of equal times).
static void create_sp_events_list(const std::vector<orc_event>& events, std::vector<orc_event>& sp_events)
{
    int ec = (int)events.size();
    sp_events.resize(ec);
    std::vector<bool> processed(ec, false);
    int n = 0;
    for (int k = 0; k < ec; k++) {
        if (!processed[k]) { processed[k] = true; sp_events[n] = events[k]; }
        else continue;
        for (int i = k + 1; i < ec; i++) {
            if (sp_events[n].id1 == events[i].id1 && sp_events[n].id2 == events[i].id2) {
                processed[i] = true;
                if (sp_events[n].t > events[i].t) sp_events[n] = events[i];
            }
        }
        n++;
    }
    sp_events.resize(n);
}

// integrator.cu:211-256, CPU branch: the largest signed component of each error vector,
// starting from 0 (a NaN component never compares larger and is passed over), scaled by
// |dt * lambda|.
static double max_error(const double* err_r, const double* err_v, int n_var, double dt, double lambda)
{
    double max_err_r = 0.0, max_err_v = 0.0;
    for (int i = 0; i < n_var; i++) {
        if (max_err_r < err_r[i]) max_err_r = err_r[i];
        if (max_err_v < err_v[i]) max_err_v = err_v[i];
    }
    return fabs(dt * lambda * std::max(max_err_r, max_err_v));
}

struct Disk {
    int n_s, n_gp, n_rp, n_pp, n_spl, n_pl, n_tp;
    int n_i[7];
    std::vector<vec_t> y[2], yout[2];
    std::vector<param_t> p;
    std::vector<md_t> md;
    std::vector<double> epoch;
    double t = 0.0;
    double thr[5];
    bool close_encounter = false, inner_steps = true;
    int frame_center = FC_BARY;
    bool stage_time_events = true, hc_reactivate = false;
    bool padded = false;
    double hill_factor = 0.0;
    int n_threads = 1;
    std::vector<orc_event> events;      // raw events since the last clear
    std::vector<orc_event> sp_events;   // events of the last handling pass
    int n_hit_centrum[3] = {0, 0, 0}, n_ejection[3] = {0, 0, 0}, n_collision[3] = {0, 0, 0};
    int n_close_encounter[3] = {0, 0, 0}, n_event[3] = {0, 0, 0};
    int cur_attempt = 0, cur_stage = 0;

    int n_SI() const { return n_s + n_gp + n_rp + n_pp; }
    int n_NSI() const { return n_spl + n_pl; }
    int n_NI() const { return n_tp; }
    int n_massive() const { return n_SI() + n_NSI(); }
    int n_total() const { return n_SI() + n_NSI() + n_NI(); }

    // number_of_bodies.cpp:196-258 (no padding)
    Bound bound_SI() const {
        if (frame_center == FC_BARY) return Bound{0, n_SI(), 0, n_massive()};
        return Bound{1, n_SI(), 1, n_massive()};
    }
    Bound bound_NSI() const { return Bound{n_SI(), n_SI() + n_NSI(), 0, n_SI()}; }
    Bound bound_NI() const { return Bound{n_massive(), n_massive() + n_NI(), 0, n_massive()}; }

    double mass_of_star() const {  // pp_disk.cu:1483-1496
        for (int j = 0; j < n_massive(); j++)
            if (md[j].body_type == BT_STAR) return p[j].mass;
        throw std::string("No star is included!");
    }

    void increment_event_counter(int* c) {  // pp_disk.cu:1059-1067
        for (int i = 0; i < 3; i++) { c[i]++; n_event[i]++; }
    }
    void set_event_counter(int field, int value) {  // pp_disk.cu:1069-1076
        n_hit_centrum[field] = n_ejection[field] = n_collision[field] = n_close_encounter[field] = n_event[field] = value;
    }

    // astro-centric copy of a state (tools.cpp:180-236 incl. its body-0 star test)
    void to_ac(const vec_t* r, const vec_t* v, std::vector<vec_t>& rac, std::vector<vec_t>& vac) const {
        int n = n_total();
        rac.assign(r, r + n); vac.assign(v, v + n);
        vec_t R0 = {0, 0, 0, 0}, V0 = {0, 0, 0, 0};
        bool isstar = false;
        for (int j = 0; j < n; j++) {
            if (BT_STAR == md[0].body_type) {
                R0.x = r[j].x; R0.y = r[j].y; R0.z = r[j].z;
                V0.x = v[j].x; V0.y = v[j].y; V0.z = v[j].z;
                isstar = true;
                break;
            }
        }
        if (!isstar) throw std::string("No star is included!");
        for (int j = n - 1; j >= 0; j--) {
            rac[j].x -= R0.x; rac[j].y -= R0.y; rac[j].z -= R0.z;
            vac[j].x -= V0.x; vac[j].y -= V0.y; vac[j].z -= V0.z;
        }
    }

    // One sink's force sum and its pair tests (pp_disk.cu:404-572).
    void sink_force(double ev_t, const Bound& b, int i, const vec_t* r, const vec_t* v, vec_t* a,
                    std::vector<orc_event>& out, double mstar) const
    {
        if (frame_center == FC_BARY) {
            a[i].x = a[i].y = a[i].z = a[i].w = 0.0;
        } else {
            const double mu = mstar + p[i].mass;
            const double r2 = (r[i].x * r[i].x) + (r[i].y * r[i].y) + (r[i].z * r[i].z);
            a[i].x = -mu / (r2 * sqrt(r2)) * r[i].x;
            a[i].y = -mu / (r2 * sqrt(r2)) * r[i].y;
            a[i].z = -mu / (r2 * sqrt(r2)) * r[i].z;
            a[i].w = 0.0;
        }
        // pp_disk.cu:231-241; the padded kernel (pp_disk.cu:158-167) tests no id
        if (!padded && !(0 < md[i].id)) return;
        for (int j = b.src0; j < b.src1; j++) {
            if (i == j || (!padded && 0 > md[j].id)) continue;
            double dx = r[j].x - r[i].x;
            double dy = r[j].y - r[i].y;
            double dz = r[j].z - r[i].z;
            double r2 = (dx * dx) + (dy * dy) + (dz * dz);
            double d = sqrt(r2);
            double s = p[j].mass / (d * r2);
            a[i].x += s * dx;
            a[i].y += s * dy;
            a[i].z += s * dz;
            if (frame_center == FC_ASTRO) {
                const double rj2 = (r[j].x * r[j].x) + (r[j].y * r[j].y) + (r[j].z * r[j].z);
                a[i].x -= p[j].mass / (rj2 * sqrt(rj2)) * r[j].x;
                a[i].y -= p[j].mass / (rj2 * sqrt(rj2)) * r[j].y;
                a[i].z -= p[j].mass / (rj2 * sqrt(rj2)) * r[j].z;
            }
            if (!inner_steps || !(i > 0 && i > j)) continue;
            bool collide = d < thr[TH_RADII] * (p[i].radius + p[j].radius);
            bool hill = false;
            if (!collide && hill_factor > 0.0 && j >= n_s) {
                // NEW: mutual Hill radius R_H = ((m_i+m_j)/(3 M_star))^(1/3) * (|r_i|+|r_j|)/2
                // between two bodies orbiting the star (stars lead the list; a body is always
                // within a few such radii of its star, so star pairs are not tested)
                double ri = sqrt((r[i].x * r[i].x) + (r[i].y * r[i].y) + (r[i].z * r[i].z));
                double rj = sqrt((r[j].x * r[j].x) + (r[j].y * r[j].y) + (r[j].z * r[j].z));
                double rh = cbrt_ieee((p[i].mass + p[j].mass) / (3.0 * mstar)) * ((ri + rj) * 0.5);
                hill = d < hill_factor * rh;
            }
            if (!collide && !hill) continue;
            orc_event e;
            memset(&e, 0, sizeof(e));
            int surv = i, merg = j;
            e.d = d;
            e.t = ev_t;
            e.attempt = cur_attempt;
            e.stage = cur_stage;
            e.sink = i;
            e.source = j;
            if (collide && !close_encounter) {
                if (p[merg].mass > p[surv].mass) std::swap(surv, merg);
                e.name = EV_COLLISION;
            } else {
                e.name = EV_CLOSE_ENCOUNTER;
                set4(e.p1, p[surv]);
                set4(e.p2, p[merg]);
            }
            e.id1 = md[surv].id; e.id2 = md[merg].id;
            e.idx1 = surv; e.idx2 = merg;
            set3(e.r1, r[surv]); set3(e.v1, v[surv]);
            set3(e.r2, r[merg]); set3(e.v2, v[merg]);
            if (e.name == EV_CLOSE_ENCOUNTER) {
                const double mu1 = Gauss2 * (mstar + p[surv].mass);
                const double mu2 = Gauss2 * (mstar + p[merg].mass);
                vec_t rs, rm, vs_, vm;
                if (frame_center == FC_BARY) {
                    std::vector<vec_t> rac, vac;
                    to_ac(r, v, rac, vac);
                    rs = rac[surv]; rm = rac[merg]; vs_ = vac[surv]; vm = vac[merg];
                } else {
                    rs = r[surv]; rm = r[merg]; vs_ = v[surv]; vm = v[merg];
                }
                const vec_t vts = {vs_.x * Gauss, vs_.y * Gauss, vs_.z * Gauss, 0.0};
                const vec_t vtm = {vm.x * Gauss, vm.y * Gauss, vm.z * Gauss, 0.0};
                oe_t o1, o2;
                calculate_orbital_element(mu1, &o1, rs, vts);
                calculate_orbital_element(mu2, &o2, rm, vtm);
                set6(e.oe1, o1);
                set6(e.oe2, o2);
            }
            out.push_back(e);
        }
    }

    void accel_bound(double ev_t, const Bound& b, const vec_t* r, const vec_t* v, vec_t* a) {
        double mstar = (frame_center == FC_ASTRO || hill_factor > 0.0 || close_encounter) ? mass_of_star() : 0.0;
        if (n_threads <= 1) {
            for (int i = b.sink0; i < b.sink1; i++) sink_force(ev_t, b, i, r, v, a, events, mstar);
            return;
        }
        // NEW: same per-sink j order on several threads; events merged in sink order.
        int n = b.sink1 - b.sink0;
        std::vector<std::vector<orc_event>> per(n_threads);
#ifdef _OPENMP
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 64)
#endif
        for (int k = 0; k < n; k++) {
            int tid = 0;
#ifdef _OPENMP
            tid = omp_get_thread_num();
#endif
            sink_force(ev_t, b, b.sink0 + k, r, v, a, per[tid], mstar);
        }
        std::vector<orc_event> all;
        for (auto& pv : per) all.insert(all.end(), pv.begin(), pv.end());
        std::stable_sort(all.begin(), all.end(), [](const orc_event& x, const orc_event& y) {
            return x.sink != y.sink ? x.sink < y.sink : x.source < y.source;
        });
        events.insert(events.end(), all.begin(), all.end());
    }

    // pp_disk.cu:585-607
    void cpu_calc_grav_accel(double curr_t, const vec_t* r, const vec_t* v, vec_t* dy) {
        double nsi_t = stage_time_events ? curr_t : t;
        if (0 < n_SI()) accel_bound(curr_t, bound_SI(), r, v, dy);
        if (0 < n_NSI()) accel_bound(nsi_t, bound_NSI(), r, v, dy);
        if (0 < n_NI()) accel_bound(nsi_t, bound_NI(), r, v, dy);
    }

    // pp_disk.cu:998-1041
    void calc_dydx(int i, int rr, double curr_t, const vec_t* r, const vec_t* v, vec_t* dy) {
        int n = n_total();
        if (i == 0) {
            memcpy(dy, v, n * sizeof(vec_t));
        } else {
            cur_stage = rr;
            cpu_calc_grav_accel(curr_t, r, v, dy);
        }
    }

    void swap_y() { for (int i = 0; i < 2; i++) std::swap(yout[i], y[i]); }  // pp_disk.cu:1043-1057

    // pp_disk.cu:909-932 + 934-976
    int check_ejection_hit_centrum_raw() {
        const vec_t* r = y[0].data();
        for (int i = 0; i < n_total(); i++) {
            if (0 < md[i].id && BT_STAR != md[i].body_type) {
                double r2 = (r[i].x * r[i].x) + (r[i].y * r[i].y) + (r[i].z * r[i].z);
                int name = EV_NONE;
                if (0.0 < thr[TH_EJ] && thr[TH_EJ2] < r2) name = EV_EJECTION;
                else if (0.0 < thr[TH_HC] && thr[TH_HC2] > r2) name = EV_HIT_CENTRUM;
                if (name == EV_NONE) continue;
                orc_event e;
                memset(&e, 0, sizeof(e));
                e.name = name;
                e.d = sqrt(r2);
                e.t = t;
                e.id1 = md[0].id; e.id2 = md[i].id;
                e.idx1 = 0; e.idx2 = i;
                e.attempt = -1; e.stage = -1; e.sink = i; e.source = 0;
                set3(e.r1, y[0][0]); set3(e.v1, y[1][0]);
                set3(e.r2, y[0][i]); set3(e.v2, y[1][i]);
                if (name == EV_EJECTION) {
                    set4(e.p1, p[0]); set4(e.p2, p[i]);
                    memcpy(e.rs, e.r1, sizeof(e.rs)); memcpy(e.vs, e.v1, sizeof(e.vs)); memcpy(e.ps, e.p1, sizeof(e.ps));
                }
                md[i].id *= -1;
                events.push_back(e);
            }
        }
        return (int)events.size();
    }

    // pp_disk.cu:1151-1202
    void handle_collision_pair(orc_event* c, bool kernel_negated = false) {
        int surv = c->idx1, merg = c->idx2;
        if (BT_SUPERPL == md[merg].body_type)
            throw std::string("Collision between a massive body and a super-planetesimal is not yet implemented.");
        set4(c->p1, p[surv]);
        set4(c->p2, p[merg]);
        calc_position_after_collision(c->p1[0], c->p2[0], c->r1, c->r2, c->rs);
        calc_position_after_collision(c->p1[0], c->p2[0], c->v1, c->v2, c->vs);
        y[0][surv].x = c->rs[0]; y[0][surv].y = c->rs[1]; y[0][surv].z = c->rs[2]; y[0][surv].w = 0.0;
        y[1][surv].x = c->vs[0]; y[1][surv].y = c->vs[1]; y[1][surv].z = c->vs[2]; y[1][surv].w = 0.0;
        param_t ps;
        calc_physical_properties(get4(c->p1), get4(c->p2), ps);
        set4(c->ps, ps);
        p[surv] = ps;
        // CPU path: the id negated by the detection pass is negated again (re-activation).
        // GPU path: the merge negates the host copy, which still holds the positive id.
        if (kernel_negated && !hc_reactivate) md[merg].id = -std::abs(md[merg].id);
        else md[merg].id *= -1;
        p[merg].mass = 0.0; p[merg].density = 0.0; p[merg].radius = 0.0;
        y[0][merg].x = 1.0e20; y[0][merg].y = 0.0; y[0][merg].z = 0.0;
        y[1][merg].x = 0.0; y[1][merg].y = 0.0; y[1][merg].z = 0.0;
    }

    // pp_disk.cu:1111-1149
    void create_sp_events() { create_sp_events_list(events, sp_events); }

    // pp_disk.cu:683-715 + 1091-1109
    bool check_for_ejection_hit_centrum() {
        int n_ev = check_ejection_hit_centrum_raw();
        if (0 < n_ev) {
            sp_events = events;
            for (size_t i = 0; i < sp_events.size(); i++) {
                if (sp_events[i].name == EV_EJECTION) increment_event_counter(n_ejection);
                else { handle_collision_pair(&sp_events[i], true); increment_event_counter(n_hit_centrum); }
            }
            n_ejection[CNT_LAST_STEP] = 0;
            n_hit_centrum[CNT_LAST_STEP] = 0;
            return true;
        }
        return false;
    }

    // pp_disk.cu:717-738 + 1078-1089.  Close-encounter records (NEW Hill criterion) are
    // reported and counted, not merged.
    bool check_for_collision() {
        if (0 < (int)events.size()) {
            create_sp_events();
            for (size_t i = 0; i < sp_events.size(); i++) {
                if (sp_events[i].name == EV_COLLISION) {
                    handle_collision_pair(&sp_events[i]);
                    increment_event_counter(n_collision);
                } else {
                    increment_event_counter(n_close_encounter);
                }
            }
            n_collision[CNT_LAST_STEP] = 0;
            n_close_encounter[CNT_LAST_STEP] = 0;
            return true;
        }
        return false;
    }

    // pp_disk.cu:740-770 (inner_steps == true branch)
    bool check_for_close_encounter_inner() {
        if (0 < (int)events.size()) {
            sp_events = events;
            for (size_t i = 0; i < events.size(); i++) increment_event_counter(n_close_encounter);
            n_close_encounter[CNT_LAST_STEP] = 0;
            return true;
        }
        return false;
    }

    // number_of_bodies.cpp:27-104
    void update_numbers() {
        int c[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int i = 0; i < n_total(); i++) {
            if (0 < md[i].id && BT_PAD > md[i].body_type) continue;
            if (md[i].body_type < 0 || md[i].body_type > BT_PAD) throw std::string("Undefined body type!");
            if (md[i].body_type == BT_PAD) continue;
            c[md[i].body_type]++;
            n_i[md[i].body_type]++;
        }
        n_s -= c[0]; n_gp -= c[1]; n_rp -= c[2]; n_pp -= c[3]; n_spl -= c[4]; n_pl -= c[5]; n_tp -= c[6];
    }

    // pp_disk.cu:1324-1409 (stable compaction; the class order is preserved by construction)
    void remove_inactive_bodies() {
        int old_n = n_total();
        update_numbers();
        int k = 0;
        for (int i = 0; i < old_n; i++) {
            if (0 < md[i].id && BT_PAD > md[i].body_type) {
                y[0][k] = y[0][i]; y[1][k] = y[1][i]; p[k] = p[i]; md[k] = md[i]; epoch[k] = epoch[i];
                k++;
            }
        }
        if (n_total() != k) throw std::string("Error: number of copied bodies does not equal to the number of active bodies.");
        for (int i = 0; i < 2; i++) { y[i].resize(k); yout[i].resize(k); }
        p.resize(k); md.resize(k); epoch.resize(k);
    }

    // pp_disk.cu:893-907
    bool check_for_rebuild_vectors(int n) {
        if (n_event[CNT_LAST_CLEAR] >= n) {
            remove_inactive_bodies();
            set_event_counter(CNT_LAST_CLEAR, 0);
            return true;
        }
        return false;
    }

    // NEW: E = sum 1/2 m v^2 - sum_{interacting i<j} m_i m_j / r_ij ; L = sum m r x v (G = 1 units)
    // Compensated (Neumaier) sums: millions of tiny pair terms added to a potential
    // dominated by the star would otherwise lose their low bits.
    static void nsum(double& s, double& c, double x) {
        const double t = s + x;
        if (std::fabs(s) >= std::fabs(x)) c += (s - t) + x;
        else c += (x - t) + s;
        s = t;
    }
    void energy(double* E, double* L) const {
        double T = 0.0, U = 0.0, cT = 0.0, cU = 0.0, Lx = 0.0, Ly = 0.0, Lz = 0.0;
        int nm = n_massive(), nsi = n_SI();
        for (int i = 0; i < nm; i++) {
            if (md[i].id <= 0) continue;
            const vec_t& r = y[0][i]; const vec_t& v = y[1][i];
            double m = p[i].mass;
            nsum(T, cT, 0.5 * m * ((v.x * v.x) + (v.y * v.y) + (v.z * v.z)));
            Lx += m * (r.y * v.z - r.z * v.y);
            Ly += m * (r.z * v.x - r.x * v.z);
            Lz += m * (r.x * v.y - r.y * v.x);
            int jend = i < nsi ? nm : nsi;   // NSI-NSI pairs do not interact
            for (int j = i + 1; j < jend; j++) {
                if (md[j].id <= 0) continue;
                double dx = y[0][j].x - r.x, dy = y[0][j].y - r.y, dz = y[0][j].z - r.z;
                nsum(U, cU, m * p[j].mass / sqrt((dx * dx) + (dy * dy) + (dz * dz)));
            }
        }
        // NSI sinks interact with SI sources; count each SI-NSI pair once (done above for i < nsi)
        *E = (T + cT) - (U + cU);
        L[0] = Lx; L[1] = Ly; L[2] = Lz;
    }
};

// ---- integrators ---------------------------------------------------------------------
struct Integrator {
    Disk* ppd;
    double dt_try, dt_did = 0.0;
    bool adaptive;
    double tolerance;
    int r_max, order = 1;
    bool error_check_for_tp = false;
    int n_failed_step = 0, n_passed_step = 0, n_tried_step = 0;
    std::vector<std::vector<vec_t>> dydx[2];
    std::vector<vec_t> ytemp[2];
    std::vector<double> err[2];
    double last_err = 0.0;

    // integrator.cu:27-52
    Integrator(Disk* d, double dt, bool adapt, double tol, int rmax)
        : ppd(d), dt_try(dt * Gauss), adaptive(adapt), tolerance(tol), r_max(rmax) {}
    virtual ~Integrator() {}
    void alloc() {
        int n = ppd->n_total();
        for (int i = 0; i < 2; i++) {
            ytemp[i].assign(n, vec_t{0, 0, 0, 0});
            dydx[i].resize(r_max);
            for (int r = 0; r < r_max; r++) dydx[i][r].assign(n, vec_t{0, 0, 0, 0});
            err[i].assign(4 * (size_t)n, 0.0);
        }
    }
    // integrator.cu:211-256 (CPU branch)
    double get_max_error(int n_var, double lambda) { return max_error(err[0].data(), err[1].data(), n_var, dt_try, lambda); }
    void update_counters(int iter) { n_tried_step += iter; n_failed_step += (iter - 1); n_passed_step++; }
    double* Y(int i) { return (double*)ppd->y[i].data(); }
    double* YO(int i) { return (double*)ppd->yout[i].data(); }
    double* F(int i, int r) { return (double*)dydx[i][r].data(); }
    double* YT(int i) { return (double*)ytemp[i].data(); }
    void eval(int r, double ttemp, const vec_t* coor, const vec_t* velo) {
        for (int i = 0; i < 2; i++) ppd->calc_dydx(i, r, ttemp, coor, velo, dydx[i][r].data());
    }
    int n_var_err() const {
        return 4 * (error_check_for_tp ? ppd->n_total() : ppd->n_massive());
    }
    virtual double step() = 0;
};

// result = a + b_factor * b  (int_euler.cu:36-42, int_rungekutta2.cu:34-40, int_rungekutta4.cu:77-83)
static void sum_vector(int n, const double* a, const double* b, double b_factor, double* result)
{
    for (int i = 0; i < n; i++) result[i] = a[i] + b_factor * b[i];
}

struct Euler : Integrator {  // int_euler.cu:44-141
    Euler(Disk* d, double dt) : Integrator(d, dt, false, 0.0, 1) { order = 1; alloc(); }
    double step() override {
        int nv = 4 * ppd->n_total();
        double t = ppd->t;
        ppd->cur_attempt = 0;
        eval(0, t, ppd->y[0].data(), ppd->y[1].data());
        for (int i = 0; i < 2; i++) sum_vector(nv, Y(i), F(i, 0), dt_try, YO(i));
        dt_did = dt_try;
        update_counters(1);
        ppd->t += dt_did;
        ppd->swap_y();
        return dt_did;
    }
};

struct RK2 : Integrator {  // int_rungekutta2.cu:13-152
    double a[2] = {0.0, 1.0 / 2.0}, b[2] = {0.0, 1.0}, c[2] = {0.0, 1.0 / 2.0};
    RK2(Disk* d, double dt) : Integrator(d, dt, false, 0.0, 2) { order = 2; alloc(); }
    double step() override {
        int nv = 4 * ppd->n_total();
        ppd->cur_attempt = 0;
        int r = 0;
        double ttemp = ppd->t + c[r] * dt_try;
        eval(r, ttemp, ppd->y[0].data(), ppd->y[1].data());
        r = 1;
        ttemp = ppd->t + c[r] * dt_try;
        for (int i = 0; i < 2; i++) sum_vector(nv, Y(i), F(i, r - 1), a[r] * dt_try, YT(i));
        eval(r, ttemp, ytemp[0].data(), ytemp[1].data());
        for (int i = 0; i < 2; i++) sum_vector(nv, Y(i), F(i, 1), b[1] * dt_try, YO(i));
        dt_did = dt_try;
        update_counters(1);
        ppd->t += dt_did;
        ppd->swap_y();
        return dt_did;
    }
};

struct RK4 : Integrator {  // int_rungekutta4.cu:27-273
    static constexpr double LAMBDA = 1.0 / 10.0;
    double c[5] = {0.0, 1.0 / 2.0, 1.0 / 2.0, 1.0, 1.0};
    double a[8] = {0.0, 1.0 / 2.0, 1.0 / 2.0, 1.0, 1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    double b[5] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0 - 1.0 / 10.0, 1.0 / 10.0};
    RK4(Disk* d, double dt, bool adapt, double tol) : Integrator(d, dt, adapt, tol, adapt ? 5 : 4) { order = 4; alloc(); }
    double step() override {
        int nv = 4 * ppd->n_total();
        int r = 0;
        double ttemp = ppd->t + c[r] * dt_try;
        ppd->cur_attempt = 0;
        eval(r, ttemp, ppd->y[0].data(), ppd->y[1].data());
        double max_err = 0.0;
        int iter = 0;
        do {
            ppd->cur_attempt = iter;
            dt_did = dt_try;
            for (r = 1; r < order; r++) {
                ttemp = ppd->t + c[r] * dt_try;
                for (int i = 0; i < 2; i++) sum_vector(nv, Y(i), F(i, r - 1), a[r] * dt_try, YT(i));
                eval(r, ttemp, ytemp[0].data(), ytemp[1].data());
            }
            for (int i = 0; i < 2; i++) {
                double b0 = b[0] * dt_try, b1 = b[1] * dt_try;
                const double *yn = Y(i), *f1 = F(i, 0), *f2 = F(i, 1), *f3 = F(i, 2), *f4 = F(i, 3);
                double* ynp1 = YO(i);
                for (int k = 0; k < nv; k++) ynp1[k] = yn[k] + b0 * (f1[k] + f4[k]) + b1 * (f2[k] + f3[k]);
            }
            if (adaptive) {
                r = 4;
                ttemp = ppd->t + c[r] * dt_try;
                eval(r, ttemp, ppd->yout[0].data(), ppd->yout[1].data());
                int n_var = n_var_err();
                for (int i = 0; i < 2; i++) {
                    const double *f4 = F(i, 3), *f5 = F(i, 4);
                    for (int k = 0; k < n_var; k++) err[i][k] = f4[k] - f5[k];
                }
                max_err = get_max_error(n_var, LAMBDA);
                last_err = max_err;
                dt_try *= 0.9 * step_factor(tolerance / max_err, 1.0 / 4.0);
                if ((int)ppd->events.size() > 0) {
                    // the reference leaves the loop before iter++ (int_rungekutta4.cu:255-264)
                    if (dt_try < dt_did) dt_try = dt_did;
                    break;
                }
            }
            iter++;
        } while (adaptive && max_err > tolerance);
        update_counters(iter);
        ppd->t += dt_did;
        ppd->swap_y();
        return dt_did;
    }
};

// Cash-Karp RK5(4).  The reference (int_rungekutta5.cu) is GPU-only and not wired into
// the integrator factory (options.cpp:252-255); as written it forms y_(n+1) from a stale
// f5 and reads beyond its stage table in fixed-step mode.  Restated here with:
//  - every stage evaluated before y_(n+1) is formed;
//  - y_(n+1) propagated with the 5th-order weights (bh, local extrapolation);
//  - err = sum_i (bh_i - b_i) f_i, max-norm as integrator.cu:211-256 with lambda = 1
//    (int_rungekutta5.cu:284-305), dt exponent 1/5 (:259), event break as :261-268;
//  - the stage-sum expression shape of the reference kernels (:41-77): y, then
//    y += (dt * a_ri) * f_i for every non-zero a_ri in ascending i.
struct RK5 : Integrator {
    double c[6] = {0.0, 1.0 / 5.0, 3.0 / 10.0, 3.0 / 5.0, 1.0, 7.0 / 8.0};
    double a[16] = {0.0,
                    1.0 / 5.0,
                    3.0 / 40.0, 9.0 / 40.0,
                    3.0 / 10.0, -9.0 / 10.0, 6.0 / 5.0,
                    -11.0 / 54.0, 5.0 / 2.0, -70.0 / 27.0, 35.0 / 27.0,
                    1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0};
    double bh[6] = {37.0 / 378.0, 0.0, 250.0 / 621.0, 125.0 / 594.0, 0.0, 512.0 / 1771.0};
    double b[6] = {2825.0 / 27648.0, 0.0, 18575.0 / 48384.0, 13525.0 / 55296.0, 277.0 / 14336.0, 1.0 / 4.0};
    int idx_array[6] = {0, 1, 2, 4, 7, 11};
    RK5(Disk* d, double dt, bool adapt, double tol) : Integrator(d, dt, adapt, tol, 6) { order = 5; alloc(); }
    void ysum(int nv, const double* yn, const double* coef, int nc, int i, double* out) {
        for (int k = 0; k < nv; k++) {
            double acc = yn[k];
            for (int s = 0; s < nc; s++) {
                if (0.0 == coef[s]) continue;
                acc += dt_try * coef[s] * F(i, s)[k];
            }
            out[k] = acc;
        }
    }
    double step() override {
        int nv = 4 * ppd->n_total();
        int r = 0;
        double ttemp = ppd->t + c[r] * dt_try;
        ppd->cur_attempt = 0;
        eval(r, ttemp, ppd->y[0].data(), ppd->y[1].data());
        double max_err = 0.0;
        int iter = 0;
        do {
            ppd->cur_attempt = iter;
            dt_did = dt_try;
            for (r = 1; r < 6; r++) {
                ttemp = ppd->t + c[r] * dt_try;
                for (int i = 0; i < 2; i++) ysum(nv, Y(i), &a[idx_array[r]], r, i, YT(i));
                eval(r, ttemp, ytemp[0].data(), ytemp[1].data());
            }
            for (int i = 0; i < 2; i++) ysum(nv, Y(i), bh, 6, i, YO(i));
            if (adaptive) {
                int n_var = n_var_err();
                for (int i = 0; i < 2; i++) {
                    for (int k = 0; k < n_var; k++) {
                        double e = (bh[0] - b[0]) * F(i, 0)[k];
                        for (int s = 1; s < 6; s++) e += (bh[s] - b[s]) * F(i, s)[k];
                        err[i][k] = e;
                    }
                }
                max_err = get_max_error(n_var, 1.0);
                last_err = max_err;
                dt_try *= 0.9 * step_factor(tolerance / max_err, 1.0 / 5.0);
                if ((int)ppd->events.size() > 0) {
                    // the reference leaves the loop before iter++ (int_rungekutta4.cu:255-264)
                    if (dt_try < dt_did) dt_try = dt_did;
                    break;
                }
            }
            iter++;
        } while (adaptive && max_err > tolerance);
        update_counters(iter);
        ppd->t += dt_did;
        ppd->swap_y();
        return dt_did;
    }
};

struct RKF78 : Integrator {  // int_rungekutta8.cu:27-47,463-797
    static constexpr double LAMBDA = 41.0 / 840.0;
    double c[13] = {0.0, 2.0 / 27.0, 1.0 / 9.0, 1.0 / 6.0, 5.0 / 12.0, 1.0 / 2.0, 5.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0, 1.0 / 3.0, 1.0, 0.0, 1.0};
    double a[79] = {0.0,
                    2.0 / 27.0,
                    1.0 / 36.0, 1.0 / 12.0,
                    1.0 / 24.0, 0.0, 1.0 / 8.0,
                    5.0 / 12.0, 0.0, -25.0 / 16.0, 25.0 / 16.0,
                    1.0 / 20.0, 0.0, 0.0, 1.0 / 4.0, 1.0 / 5.0,
                    -25.0 / 108.0, 0.0, 0.0, 125.0 / 108.0, -65.0 / 27.0, 125.0 / 54.0,
                    31.0 / 300.0, 0.0, 0.0, 0.0, 61.0 / 225.0, -2.0 / 9.0, 13.0 / 900.0,
                    2.0, 0.0, 0.0, -53.0 / 6.0, 704.0 / 45.0, -107.0 / 9.0, 67.0 / 90.0, 3.0,
                    -91.0 / 108.0, 0.0, 0.0, 23.0 / 108.0, -976.0 / 135.0, 311.0 / 54.0, -19.0 / 60.0, 17.0 / 6.0, -1.0 / 12.0,
                    2383.0 / 4100.0, 0.0, 0.0, -341.0 / 164.0, 4496.0 / 1025.0, -301.0 / 82.0, 2133.0 / 4100.0, 45.0 / 82.0, 45.0 / 164.0, 18.0 / 41.0,
                    3.0 / 205.0, 0.0, 0.0, 0.0, 0.0, -6.0 / 41.0, -3.0 / 205.0, -3.0 / 41.0, 3.0 / 41.0, 6.0 / 41.0, 0.0,
                    -1777.0 / 4100.0, 0.0, 0.0, -341.0 / 164.0, 4496.0 / 1025.0, -289.0 / 82.0, 2193.0 / 4100.0, 51.0 / 82.0, 33.0 / 164.0, 12.0 / 41.0, 0.0, 1.0};
    double b[11] = {41.0 / 840.0, 0.0, 0.0, 0.0, 0.0, 34.0 / 105.0, 9.0 / 35.0, 9.0 / 35.0, 9.0 / 280.0, 9.0 / 280.0, 41.0 / 840.0};
    RKF78(Disk* d, double dt, bool adapt, double tol) : Integrator(d, dt, adapt, tol, adapt ? 13 : 11) { order = 7; alloc(); }

    // int_rungekutta8.cu:337-442: ytemp = yn + dt*(sum over the non-zero a_rk f_k in ascending k)
    void calc_ytemp_for_fr(int nv, int r) {
        static const int idx_of[13] = {0, 1, 2, 4, 7, 11, 16, 22, 29, 37, 46, 56, 67};
        static const int terms[13][10] = {
            {0}, {0}, {0, 1}, {0, 2}, {0, 2, 3}, {0, 3, 4}, {0, 3, 4, 5}, {0, 4, 5, 6},
            {0, 3, 4, 5, 6, 7}, {0, 3, 4, 5, 6, 7, 8}, {0, 3, 4, 5, 6, 7, 8, 9},
            {0, 5, 6, 7, 8, 9}, {0, 3, 4, 5, 6, 7, 8, 9, 11}};
        static const int nterms[13] = {0, 1, 2, 2, 3, 3, 4, 4, 6, 7, 8, 6, 9};
        int idx = idx_of[r];
        for (int i = 0; i < 2; i++) {
            const double* yn = Y(i);
            double* yt = YT(i);
            const double* f[10];
            double co[10];
            for (int s = 0; s < nterms[r]; s++) { f[s] = F(i, terms[r][s]); co[s] = a[idx + terms[r][s]]; }
            for (int k = 0; k < nv; k++) {
                double sum = co[0] * f[0][k];
                for (int s = 1; s < nterms[r]; s++) sum = sum + co[s] * f[s][k];
                yt[k] = yn[k] + dt_try * sum;
            }
        }
    }
    double step() override {
        int nv = 4 * ppd->n_total();
        int r = 0;
        double ttemp = ppd->t + c[r] * dt_try;
        ppd->cur_attempt = 0;
        eval(r, ttemp, ppd->y[0].data(), ppd->y[1].data());
        double max_err = 0.0;
        int iter = 0;
        do {
            if (0 < (int)ppd->events.size()) ppd->events.clear();
            ppd->cur_attempt = iter;
            dt_did = dt_try;
            for (r = 1; r <= 10; r++) {
                ttemp = ppd->t + c[r] * dt_try;
                calc_ytemp_for_fr(nv, r);
                eval(r, ttemp, ytemp[0].data(), ytemp[1].data());
            }
            for (int i = 0; i < 2; i++) {  // int_rungekutta8.cu:454-460
                const double *yn = Y(i), *f0 = F(i, 0), *f5 = F(i, 5), *f6 = F(i, 6), *f7 = F(i, 7), *f8 = F(i, 8), *f9 = F(i, 9), *f10 = F(i, 10);
                double* o = YO(i);
                for (int k = 0; k < nv; k++)
                    o[k] = yn[k] + dt_try * (b[0] * f0[k] + b[5] * f5[k] + b[6] * (f6[k] + f7[k]) + b[8] * (f8[k] + f9[k]) + b[10] * f10[k]);
            }
            if (adaptive) {
                for (r = 11; r < r_max; r++) {
                    ttemp = ppd->t + c[r] * dt_try;
                    calc_ytemp_for_fr(nv, r);
                    eval(r, ttemp, ytemp[0].data(), ytemp[1].data());
                }
                int n_var = n_var_err();
                for (int i = 0; i < 2; i++) {  // int_rungekutta8.cu:445-451
                    const double *f0 = F(i, 0), *f10 = F(i, 10), *f11 = F(i, 11), *f12 = F(i, 12);
                    for (int k = 0; k < n_var; k++) err[i][k] = (f0[k] + f10[k] - f11[k] - f12[k]);
                }
                max_err = get_max_error(n_var, LAMBDA);
                last_err = max_err;
                dt_try *= 0.9 * step_factor(tolerance / max_err, 1.0 / (order + 1));
            }
            iter++;
        } while (adaptive && max_err > tolerance);
        update_counters(iter);
        ppd->t += dt_did;
        ppd->swap_y();
        return dt_did;
    }
};

struct Oracle {
    Disk disk;
    Integrator* intgr = nullptr;
    std::string last_error;
    ~Oracle() { delete intgr; }
};

}  // namespace

extern "C" {

// counts[7] = ns ngp nrp npp nspl npl ntp; arrays in file order; p is N*4 (mass radius density cd);
// r, v are N*3.  The state is taken as given (internal units are the caller's business).
void* orc_create(const int* counts, const int* id, const int* type, const int* mig_type, const double* mig_stop_at,
                 const double* epoch, const double* p, const double* r, const double* v, const orc_params* prm)
{
    Oracle* o = new Oracle();
    Disk& d = o->disk;
    d.n_s = counts[0]; d.n_gp = counts[1]; d.n_rp = counts[2]; d.n_pp = counts[3];
    d.n_spl = counts[4]; d.n_pl = counts[5]; d.n_tp = counts[6];
    for (int i = 0; i < 7; i++) d.n_i[i] = 0;
    int n = d.n_total();
    for (int i = 0; i < 2; i++) { d.y[i].assign(n, vec_t{0, 0, 0, 0}); d.yout[i].assign(n, vec_t{0, 0, 0, 0}); }
    d.p.resize(n); d.md.resize(n); d.epoch.resize(n);
    for (int k = 0; k < n; k++) {
        d.md[k].id = id[k]; d.md[k].body_type = type[k];
        d.md[k].mig_type = mig_type ? mig_type[k] : 0;
        d.md[k].mig_stop_at = mig_stop_at ? mig_stop_at[k] : 0.0;
        d.epoch[k] = epoch ? epoch[k] : 0.0;
        d.p[k] = get4(p + 4 * k);
        d.y[0][k] = get3(r + 3 * k);
        d.y[1][k] = get3(v + 3 * k);
    }
    d.thr[TH_HC] = prm->hit_centrum;
    d.thr[TH_EJ] = prm->ejection;
    d.thr[TH_RADII] = prm->radii_enhance_factor;
    d.thr[TH_HC2] = prm->hit_centrum * prm->hit_centrum;
    d.thr[TH_EJ2] = prm->ejection * prm->ejection;
    d.close_encounter = prm->close_encounter != 0;
    d.inner_steps = prm->inner_steps != 0;
    d.frame_center = prm->frame_center;
    d.stage_time_events = prm->stage_time_events != 0;
    d.hc_reactivate = prm->hc_reactivate != 0;
    d.padded = prm->padded != 0;
    d.hill_factor = prm->encounter_hill_factor;
    d.n_threads = prm->n_threads > 0 ? prm->n_threads : 1;
    double dt = prm->dt_day;
    switch (prm->integrator) {
    case INT_EULER: o->intgr = new Euler(&d, dt); break;
    case INT_RK2: o->intgr = new RK2(&d, dt); break;
    case INT_RK4: o->intgr = new RK4(&d, dt, prm->adaptive != 0, prm->tolerance); break;
    case INT_RK5: o->intgr = new RK5(&d, dt, prm->adaptive != 0, prm->tolerance); break;
    case INT_RKF78: o->intgr = new RKF78(&d, dt, prm->adaptive != 0, prm->tolerance); break;
    default: delete o; return nullptr;
    }
    o->intgr->error_check_for_tp = prm->error_check_for_tp != 0;
    return o;
}

void orc_destroy(void* h) { delete (Oracle*)h; }
int orc_n_total(void* h) { return ((Oracle*)h)->disk.n_total(); }
void orc_get_counts(void* h, int* c)
{
    Disk& d = ((Oracle*)h)->disk;
    c[0] = d.n_s; c[1] = d.n_gp; c[2] = d.n_rp; c[3] = d.n_pp; c[4] = d.n_spl; c[5] = d.n_pl; c[6] = d.n_tp;
}
const char* orc_last_error(void* h) { return ((Oracle*)h)->last_error.c_str(); }

// tools.cpp:109-178 (sums run j = n-1 .. 0)
void orc_transform_to_bc(void* h)
{
    Disk& d = ((Oracle*)h)->disk;
    int n = d.n_total();
    vec_t R0 = {0, 0, 0, 0}, V0 = {0, 0, 0, 0};
    for (int j = n - 1; j >= 0; j--) {
        double m = d.p[j].mass;
        R0.x += m * d.y[0][j].x; R0.y += m * d.y[0][j].y; R0.z += m * d.y[0][j].z;
        V0.x += m * d.y[1][j].x; V0.y += m * d.y[1][j].y; V0.z += m * d.y[1][j].z;
    }
    double M0 = 0.0;
    for (int j = n - 1; j >= 0; j--) M0 += d.p[j].mass;
    R0.x /= M0; R0.y /= M0; R0.z /= M0;
    V0.x /= M0; V0.y /= M0; V0.z /= M0;
    for (int j = n - 1; j >= 0; j--) {
        d.y[0][j].x -= R0.x; d.y[0][j].y -= R0.y; d.y[0][j].z -= R0.z;
        d.y[1][j].x -= V0.x; d.y[1][j].y -= V0.y; d.y[1][j].z -= V0.z;
    }
}

// tools.cpp:209-236
void orc_transform_to_ac(void* h)
{
    Disk& d = ((Oracle*)h)->disk;
    std::vector<vec_t> rac, vac;
    d.to_ac(d.y[0].data(), d.y[1].data(), rac, vac);
    d.y[0] = rac; d.y[1] = vac;
}

// pp_disk.cu:1512-1551
void orc_transform_time_velocity(void* h)
{
    Disk& d = ((Oracle*)h)->disk;
    for (int j = 0; j < d.n_total(); j++) d.epoch[j] *= Gauss;
    for (int j = 0; j < d.n_total(); j++) { d.y[1][j].x /= Gauss; d.y[1][j].y /= Gauss; d.y[1][j].z /= Gauss; }
}

void orc_set_t(void* h, double t) { ((Oracle*)h)->disk.t = t; }
double orc_get_t(void* h) { return ((Oracle*)h)->disk.t; }
double orc_dt_try(void* h) { return ((Oracle*)h)->intgr->dt_try; }
void orc_set_dt_try(void* h, double dt) { ((Oracle*)h)->intgr->dt_try = dt; }
double orc_last_err(void* h) { return ((Oracle*)h)->intgr->last_err; }

// Returns dt_did, or NaN on an exception (message via orc_last_error).
double orc_step(void* h)
{
    Oracle* o = (Oracle*)h;
    try { return o->intgr->step(); }
    catch (const std::string& s) { o->last_error = s; return NAN; }
}

void orc_get_state(void* h, double* r, double* v, double* p, int* id, int* type)
{
    Disk& d = ((Oracle*)h)->disk;
    for (int k = 0; k < d.n_total(); k++) {
        if (r) set3(r + 3 * k, d.y[0][k]);
        if (v) set3(v + 3 * k, d.y[1][k]);
        if (p) set4(p + 4 * k, d.p[k]);
        if (id) id[k] = d.md[k].id;
        if (type) type[k] = d.md[k].body_type;
    }
}

void orc_set_state(void* h, const double* r, const double* v)
{
    Disk& d = ((Oracle*)h)->disk;
    for (int k = 0; k < d.n_total(); k++) {
        if (r) d.y[0][k] = get3(r + 3 * k);
        if (v) d.y[1][k] = get3(v + 3 * k);
    }
}

// Returns 1 if the corresponding handler ran, 0 if there was nothing to do, -1 on exception.
int orc_check_ejection_hit_centrum(void* h)
{
    Oracle* o = (Oracle*)h;
    try { return o->disk.check_for_ejection_hit_centrum() ? 1 : 0; }
    catch (const std::string& s) { o->last_error = s; return -1; }
}
int orc_check_collision(void* h)
{
    Oracle* o = (Oracle*)h;
    try { return o->disk.check_for_collision() ? 1 : 0; }
    catch (const std::string& s) { o->last_error = s; return -1; }
}
int orc_check_close_encounter(void* h) { return ((Oracle*)h)->disk.check_for_close_encounter_inner() ? 1 : 0; }
int orc_check_rebuild(void* h, int n)
{
    Oracle* o = (Oracle*)h;
    try { return o->disk.check_for_rebuild_vectors(n) ? 1 : 0; }
    catch (const std::string& s) { o->last_error = s; return -1; }
}
void orc_clear_events(void* h) { ((Oracle*)h)->disk.events.clear(); }
int orc_n_raw_events(void* h) { return (int)((Oracle*)h)->disk.events.size(); }

int orc_get_raw_events(void* h, orc_event* out, int cap)
{
    Disk& d = ((Oracle*)h)->disk;
    int n = (int)d.events.size();
    for (int k = 0; k < n && k < cap; k++) out[k] = d.events[k];
    return n;
}
int orc_get_sp_events(void* h, orc_event* out, int cap)
{
    Disk& d = ((Oracle*)h)->disk;
    int n = (int)d.sp_events.size();
    for (int k = 0; k < n && k < cap; k++) out[k] = d.sp_events[k];
    return n;
}

// Direct force evaluation on an arbitrary state (r, v: N*3) at stage time t; a: N*3.
// Appends detected pair events to the raw event list (stage index 'stage').
void orc_accel(void* h, double t, int stage, const double* r3, const double* v3, double* a3)
{
    Disk& d = ((Oracle*)h)->disk;
    int n = d.n_total();
    std::vector<vec_t> r(n), v(n), a(n, vec_t{0, 0, 0, 0});
    for (int k = 0; k < n; k++) { r[k] = get3(r3 + 3 * k); v[k] = get3(v3 + 3 * k); }
    d.cur_stage = stage;
    d.cpu_calc_grav_accel(t, r.data(), v.data(), a.data());
    for (int k = 0; k < n; k++) set3(a3 + 3 * k, a[k]);
}

// out: tried failed passed, n_collision n_ejection n_hit_centrum n_close_encounter (totals), n_event[LAST_CLEAR]
void orc_counters(void* h, int* out)
{
    Oracle* o = (Oracle*)h;
    Disk& d = o->disk;
    out[0] = o->intgr->n_tried_step; out[1] = o->intgr->n_failed_step; out[2] = o->intgr->n_passed_step;
    out[3] = d.n_collision[CNT_TOTAL]; out[4] = d.n_ejection[CNT_TOTAL]; out[5] = d.n_hit_centrum[CNT_TOTAL];
    out[6] = d.n_close_encounter[CNT_TOTAL]; out[7] = d.n_event[CNT_LAST_CLEAR];
}

void orc_energy(void* h, double* E, double* L) { ((Oracle*)h)->disk.energy(E, L); }

// tools.cpp:238-290, 292-471 (exposed for the known-answer tests)
double orc_calculate_radius(double m, double density) { return calculate_radius(m, density); }
void orc_calc_physical_properties(const double* p1, const double* p2, double* p)
{
    param_t q;
    calc_physical_properties(get4(p1), get4(p2), q);
    set4(p, q);
}
void orc_calc_position_after_collision(double m1, double m2, const double* r1, const double* r2, double* r)
{
    calc_position_after_collision(m1, m2, r1, r2, r);
}
int orc_kepler_equation_solver(double ecc, double mean, double eps, double* E) { return kepler_equation_solver(ecc, mean, eps, E); }
int orc_calculate_phase(double mu, const double* oe6, double* r3, double* v3)
{
    oe_t o = {oe6[0], oe6[1], oe6[2], oe6[3], oe6[4], oe6[5]};
    vec_t r = {0, 0, 0, 0}, v = {0, 0, 0, 0};
    int rc = calculate_phase(mu, &o, &r, &v);
    set3(r3, r); set3(v3, v);
    return rc;
}
int orc_calculate_orbital_element(double mu, const double* r3, const double* v3, double* oe6)
{
    oe_t o;
    int rc = calculate_orbital_element(mu, &o, get3(r3), get3(v3));
    set6(oe6, o);
    return rc;
}

// pp_disk.cu:1111-1149 on a caller-built raw list (exposed for the hand-built tests).
int orc_create_sp_events(const orc_event* in, int n, orc_event* out)
{
    std::vector<orc_event> ev(in, in + n), sp;
    create_sp_events_list(ev, sp);
    for (size_t k = 0; k < sp.size(); k++) out[k] = sp[k];
    return (int)sp.size();
}

// integrator.cu:211-256 on caller-given error vectors (n_var components each).
double orc_step_factor(double x, double y) { return step_factor(x, y); }

double orc_max_error(const double* err_r, const double* err_v, int n_var, double dt, double lambda)
{
    return max_error(err_r, err_v, n_var, dt, lambda);
}

// pp_disk.cu:1151-1202: merges the pair of a caller-built collision record into the disk
// (idx1 survives, idx2 is removed; r1, v1, r2, v2 are the recorded stage states).  The
// record receives p1, p2, ps, rs and vs.  Returns 0, or -1 on an exception.
int orc_handle_collision_pair(void* h, orc_event* e, int kernel_negated)
{
    Oracle* o = (Oracle*)h;
    try { o->disk.handle_collision_pair(e, kernel_negated != 0); return 0; }
    catch (const std::string& s) { o->last_error = s; return -1; }
}

}  // extern "C"
