// host_tools.cpp -- host-side arithmetic of the event pipeline and step control.
// Built with g++ -ffp-contract=off (see host_tools.hpp for why not with the HIP compiler).
#include "host_tools.hpp"
#include "step_pow.hpp"

#include <algorithm>
#include <cmath>
#include <utility>

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
    *dt *= 0.9 * step_pow(tol / max_err, exponent);
    return max_err;
}

double tolerance_step(double dt, double err, double tol, double exponent)
{
    if (!(err > 0.0)) return HUGE_VAL;
    return dt * step_pow(tol / err, exponent);
}

}  // namespace tools
}  // namespace red

// ---- event pipeline -------------------------------------------------------------------------
namespace red {
namespace events {

namespace {
enum { BT_STAR = 0, BT_SUPERPL = 4 };
}

void Counters::increment(int* cnt)
{
    for (int i = 0; i < 3; i++) {
        cnt[i]++;
        event[i]++;
    }
}

void Counters::set(int field, int value)
{
    hit_centrum[field] = ejection[field] = collision[field] = close_encounter[field] = event[field] = value;
}

double mass_of_star(const BodyView& B)
{
    for (int j = 0; j < B.n_massive; j++)
        if (B.type[j] == BT_STAR) return B.mass[j];
    return -1.0;
}

bool event_key_less(const red_event& a, const red_event& b)
{
    if (a.attempt != b.attempt) return a.attempt < b.attempt;
    if (a.stage != b.stage) return a.stage < b.stage;
    if (a.sink != b.sink) return a.sink < b.sink;
    return a.source < b.source;
}

namespace {

// A record's canonical key packed into two words that compare as (attempt, stage) and
// (sink, source) do as signed ints, and its position, which orders equal keys.
struct PackedKey {
    uint64_t hi, lo;
    uint32_t idx;
};

inline uint64_t biased(int x) { return (uint64_t)((uint32_t)x ^ 0x80000000u); }

template <class Get>
std::vector<PackedKey> canonical_keys(size_t n, Get get)
{
    std::vector<PackedKey> k(n);
    for (size_t i = 0; i < n; i++) {
        const red_event& e = get(i);
        k[i] = PackedKey{(biased(e.attempt) << 32) | biased(e.stage), (biased(e.sink) << 32) | biased(e.source),
                         (uint32_t)i};
    }
    std::sort(k.begin(), k.end(), [](const PackedKey& a, const PackedKey& b) {
        if (a.hi != b.hi) return a.hi < b.hi;
        if (a.lo != b.lo) return a.lo < b.lo;
        return a.idx < b.idx;
    });
    return k;
}

}  // namespace

void canonical_sort(std::vector<const red_event*>& ev)
{
    const size_t n = ev.size();
    if (n < 2) return;
    const std::vector<PackedKey> k = canonical_keys(n, [&](size_t i) -> const red_event& { return *ev[i]; });
    std::vector<const red_event*> out(n);
    for (size_t i = 0; i < n; i++) out[i] = ev[k[i].idx];
    ev.swap(out);
}

void canonical_sort(std::vector<red_event>& ev)
{
    const size_t n = ev.size();
    if (n < 2) return;
    const std::vector<PackedKey> k = canonical_keys(n, [&](size_t i) -> const red_event& { return ev[i]; });
    std::vector<red_event> out(n);
    for (size_t i = 0; i < n; i++) out[i] = ev[k[i].idx];
    ev.swap(out);
}

// Records kept one per pair by the device, the stage where their pair first occurs in the
// attempt field's upper half: ordered by (that stage, sink, source), which is the order of
// the pairs' first records in the canonical order of one attempt.
void kept_order(std::vector<const red_event*>& ev)
{
    const size_t n = ev.size();
    if (n < 2) return;
    std::vector<PackedKey> k(n);
    for (size_t i = 0; i < n; i++) {
        const red_event& e = *ev[i];
        k[i] = PackedKey{(uint64_t)((uint32_t)e.attempt >> 16), (biased(e.sink) << 32) | biased(e.source), (uint32_t)i};
    }
    std::sort(k.begin(), k.end(), [](const PackedKey& a, const PackedKey& b) {
        if (a.hi != b.hi) return a.hi < b.hi;
        if (a.lo != b.lo) return a.lo < b.lo;
        return a.idx < b.idx;
    });
    std::vector<const red_event*> out(n);
    for (size_t i = 0; i < n; i++) out[i] = ev[k[i].idx];
    ev.swap(out);
}

// The reference compares every record with every later one (O(E^2)); its result is, for each
// (id1, id2) in the order of its first record, the record of smallest t, the first of equal
// times (a later record replaces the kept one only when its t is smaller).  Sorting the
// records by (pair, position) gives the same list in O(E log E).
void create_sp_events(const std::vector<const red_event*>& ev, std::vector<red_event>& sp)
{
    // (pair, position) sorted: each pair's records in list order, so its first record and
    // the kept one (replaced only by a strictly smaller t) are found in one sweep
    struct PairPos {
        uint64_t pair;
        uint32_t pos;
    };
    const size_t n = ev.size();
    std::vector<PairPos> v(n);
    for (size_t i = 0; i < n; i++)
        v[i] = PairPos{((uint64_t)(uint32_t)ev[i]->id1 << 32) | (uint32_t)ev[i]->id2, (uint32_t)i};
    std::sort(v.begin(), v.end(), [](const PairPos& a, const PairPos& b) {
        return a.pair != b.pair ? a.pair < b.pair : a.pos < b.pos;
    });
    std::vector<std::pair<uint32_t, uint32_t>> groups;   // (first position, kept position)
    for (size_t a = 0; a < n;) {
        size_t b = a;
        uint32_t kept = v[a].pos;
        for (; b < n && v[b].pair == v[a].pair; b++)
            if (ev[v[b].pos]->t < ev[kept]->t) kept = v[b].pos;
        groups.emplace_back(v[a].pos, kept);
        a = b;
    }
    std::sort(groups.begin(), groups.end());
    const size_t base = sp.size();
    sp.resize(base + groups.size());
    for (size_t k = 0; k < groups.size(); k++) sp[base + k] = *ev[groups[k].second];
}

void create_sp_events(const std::vector<red_event>& ev, std::vector<red_event>& sp)
{
    std::vector<const red_event*> p(ev.size());
    for (size_t k = 0; k < ev.size(); k++) p[k] = &ev[k];
    sp.clear();
    create_sp_events(p, sp);
}

int handle_collision_pair(BodyView& B, red_event* e, bool kernel_negated, std::vector<BodyUpdate>& ups,
                          std::string& msg)
{
    const int surv = e->idx1, merg = e->idx2;
    if (B.type[merg] == BT_SUPERPL) {
        msg = "Collision between a massive body and a super-planetesimal is not yet implemented.";
        return RED_ERR_UNSUPPORTED;
    }
    const double p1[4] = {B.mass[surv], B.radius[surv], B.density[surv], B.cd[surv]};
    const double p2[4] = {B.mass[merg], B.radius[merg], B.density[merg], B.cd[merg]};
    std::memcpy(e->p1, p1, sizeof(p1));
    std::memcpy(e->p2, p2, sizeof(p2));
    tools::mass_weighted_mean(p1[0], p2[0], e->r1, e->r2, e->rs);
    tools::mass_weighted_mean(p1[0], p2[0], e->v1, e->v2, e->vs);
    tools::calc_physical_properties(p1, p2, e->ps);
    B.mass[surv] = e->ps[0]; B.radius[surv] = e->ps[1]; B.density[surv] = e->ps[2]; B.cd[surv] = e->ps[3];
    // merger: inactive, massless, far away (pp_disk.cu:1177-1189)
    if (kernel_negated) B.id[merg] = -std::abs(B.id[merg]);
    else B.id[merg] *= -1;
    B.mass[merg] = 0.0; B.density[merg] = 0.0; B.radius[merg] = 0.0;
    BodyUpdate us;
    us.idx = surv; us.id = B.id[surv];
    us.r[0] = e->rs[0]; us.r[1] = e->rs[1]; us.r[2] = e->rs[2]; us.r[3] = B.mass[surv];
    us.v[0] = e->vs[0]; us.v[1] = e->vs[1]; us.v[2] = e->vs[2]; us.v[3] = 0.0;
    us.mass = B.mass[surv]; us.radius = B.radius[surv];
    BodyUpdate um;
    um.idx = merg; um.id = B.id[merg];
    um.r[0] = 1.0e20; um.r[1] = 0.0; um.r[2] = 0.0; um.r[3] = 0.0;
    um.v[0] = um.v[1] = um.v[2] = um.v[3] = 0.0;
    um.mass = 0.0; um.radius = 0.0;
    ups.push_back(us);
    ups.push_back(um);
    return RED_OK;
}

void close_encounter_elements(const BodyView& B, red_event* e)
{
    const double ms = mass_of_star(B);
    const double mu1 = tools::GAUSS2 * (ms + B.mass[e->idx1]);
    const double mu2 = tools::GAUSS2 * (ms + B.mass[e->idx2]);
    double r1[3], v1[3], r2[3], v2[3];
    for (int k = 0; k < 3; k++) {
        r1[k] = e->r1[k] - e->rs[k];
        r2[k] = e->r2[k] - e->rs[k];
        v1[k] = (e->v1[k] - e->vs[k]) * tools::GAUSS;
        v2[k] = (e->v2[k] - e->vs[k]) * tools::GAUSS;
    }
    tools::calculate_orbital_element(mu1, r1, v1, e->oe1);
    tools::calculate_orbital_element(mu2, r2, v2, e->oe2);
    const double p1[4] = {B.mass[e->idx1], B.radius[e->idx1], B.density[e->idx1], B.cd[e->idx1]};
    const double p2[4] = {B.mass[e->idx2], B.radius[e->idx2], B.density[e->idx2], B.cd[e->idx2]};
    std::memcpy(e->p1, p1, sizeof(p1));
    std::memcpy(e->p2, p2, sizeof(p2));
    for (int k = 0; k < 3; k++) { e->rs[k] = 0.0; e->vs[k] = 0.0; }
}

int ejection_pass(BodyView& B, std::vector<red_event>& ev, Counters& C, std::vector<BodyUpdate>& ups,
                  std::string& msg)
{
    std::stable_sort(ev.begin(), ev.end(), [](const red_event& a, const red_event& b) { return a.sink < b.sink; });
    for (auto& e : ev) B.id[e.idx2] = -B.id[e.idx2];   // mirror the scan's negation
    // ejection records carry the parameters at detection (pp_disk.cu:58-66), before any
    // hit-centrum merge of this pass changes the star
    for (auto& e : ev) {
        if (e.name != RED_EV_EJECTION) continue;
        const int i = e.idx2;
        const double p0[4] = {B.mass[0], B.radius[0], B.density[0], B.cd[0]};
        const double pi[4] = {B.mass[i], B.radius[i], B.density[i], B.cd[i]};
        std::memcpy(e.p1, p0, sizeof(p0));
        std::memcpy(e.p2, pi, sizeof(pi));
        std::memcpy(e.rs, e.r1, sizeof(e.rs));
        std::memcpy(e.vs, e.v1, sizeof(e.vs));
        std::memcpy(e.ps, e.p1, sizeof(e.ps));
    }
    for (auto& e : ev) {
        if (e.name == RED_EV_EJECTION) {
            C.increment(C.ejection);
        } else {
            const int rc = handle_collision_pair(B, &e, true, ups, msg);
            if (rc) return rc;
            C.increment(C.hit_centrum);
        }
    }
    C.ejection[CNT_LAST_STEP] = 0;
    C.hit_centrum[CNT_LAST_STEP] = 0;
    return RED_OK;
}

int collision_pass(BodyView& B, const std::vector<red_event>& raw, std::vector<red_event>& sp, Counters& C,
                   std::vector<BodyUpdate>& ups, std::string& msg)
{
    create_sp_events(raw, sp);
    return collision_sp_pass(B, sp.data(), sp.size(), C, ups, msg);
}

int collision_sp_pass(BodyView& B, red_event* sp, size_t n, Counters& C, std::vector<BodyUpdate>& ups,
                      std::string& msg)
{
    // close-encounter records take the parameters the bodies had at detection, i.e. before
    // any merge of this pass
    for (size_t k = 0; k < n; k++)
        if (sp[k].name != RED_EV_COLLISION) close_encounter_elements(B, &sp[k]);
    for (size_t k = 0; k < n; k++) {
        red_event& e = sp[k];
        if (e.name == RED_EV_COLLISION) {
            const int rc = handle_collision_pair(B, &e, false, ups, msg);
            if (rc) return rc;
            C.increment(C.collision);
        } else {
            C.increment(C.close_encounter);
        }
    }
    C.collision[CNT_LAST_STEP] = 0;
    C.close_encounter[CNT_LAST_STEP] = 0;
    return RED_OK;
}

void close_encounter_pass(const BodyView& B, const std::vector<red_event>& raw, std::vector<red_event>& sp,
                          Counters& C)
{
    sp = raw;
    close_encounter_sp_pass(B, sp.data(), sp.size(), C);
}

void close_encounter_sp_pass(const BodyView& B, red_event* sp, size_t n, Counters& C)
{
    for (size_t k = 0; k < n; k++) {
        close_encounter_elements(B, &sp[k]);
        C.increment(C.close_encounter);
    }
    C.close_encounter[CNT_LAST_STEP] = 0;
}

int ejection_name(const double* q, double ej, double hc)
{
    const double r2 = (q[0] * q[0]) + (q[1] * q[1]) + (q[2] * q[2]);
    if (0.0 < ej && ej * ej < r2) return RED_EV_EJECTION;
    if (0.0 < hc && hc * hc > r2) return RED_EV_HIT_CENTRUM;
    return RED_EV_NONE;
}

void ejection_scan(int n, const double* r4, const double* v4, int* id, const int* type, double t, double ej,
                   double hc, std::vector<red_event>& out)
{
    out.clear();
    for (int i = 0; i < n; i++) {
        if (!(id[i] > 0 && type[i] != BT_STAR)) continue;
        const double* q = r4 + 4 * i;
        const int name = ejection_name(q, ej, hc);
        if (name == RED_EV_NONE) continue;
        const double r2 = (q[0] * q[0]) + (q[1] * q[1]) + (q[2] * q[2]);
        red_event e;
        std::memset(&e, 0, sizeof(e));
        e.name = name;
        e.id1 = id[0]; e.idx1 = 0;
        e.id2 = id[i]; e.idx2 = i;
        e.attempt = -1; e.stage = -1; e.sink = i; e.source = 0; e.pad = 0;
        e.t = t;
        e.d = std::sqrt(r2);
        for (int c = 0; c < 3; c++) {
            e.r1[c] = r4[c]; e.v1[c] = v4[c];
            e.r2[c] = q[c]; e.v2[c] = v4[4 * i + c];
        }
        id[i] = -id[i];
        out.push_back(e);
    }
}

}  // namespace events
}  // namespace red
