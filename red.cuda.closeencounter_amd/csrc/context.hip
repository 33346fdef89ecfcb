// context.hip -- the C ABI: context life cycle, integrator step control, event pipeline.
//
// Host-side counterpart of pp_disk + integrator (red.cuda/src/red.cuda/pp_disk.cu,
// integrator.cu, int_*.cu).  The per-attempt work runs on one HIP stream, optionally
// replayed from a captured HIP graph; the host reads back one small control block per
// attempt (error maximum, NaN flag, event count), which is what the reference reads with
// its two blocking copies in integrator::get_max_error (integrator.cu:211-256).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "host_tools.hpp"
#include "red_internal.hpp"
#include "scheme.hpp"

using namespace red;

namespace {

const double GAUSS = tools::GAUSS;

enum { BT_STAR = 0, BT_GIANT = 1, BT_ROCKY = 2, BT_PROTO = 3, BT_SUPERPL = 4, BT_PL = 5, BT_TP = 6, BT_PAD = 7 };

__global__ void k_set_ctl(Ctl* ctl, double t, double dt, int attempt, int reset_err, int clear_events, int clear_hc)
{
    ctl->t = t;
    ctl->dt = dt;
    ctl->attempt = attempt;
    if (reset_err) { ctl->err_bits = 0ull; ctl->nan_flag = 0; }
    if (clear_events) ctl->n_events = 0;
    if (clear_hc) ctl->n_events_hc = 0;
}

template <typename T>
hipError_t dmalloc(T** p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc((void**)p, count * sizeof(T));
}

}  // namespace

struct red_ctx {
    red_params prm{};
    hipStream_t stream = nullptr;
    std::string err;
    int cnt[7] = {0};
    int n_inactive[7] = {0};
    int n = 0, cap = 0;
    // host mirror of the per-body parameters (pp_disk::h_p, h_body_md)
    std::vector<double> h_mass, h_radius, h_density, h_cd, h_epoch, h_mig_stop;
    std::vector<int> h_id, h_type, h_mig_type;
    // device storage
    d4* d_r[2] = {nullptr, nullptr};
    d4* d_v[2] = {nullptr, nullptr};
    int cur = 0;
    std::vector<d4*> d_sr, d_sv, d_a;
    double* d_mass = nullptr;
    double* d_radius = nullptr;
    int* d_id = nullptr;
    int* d_type = nullptr;
    int* d_ones = nullptr;                // padded storage: every body gates as active
    bool padded = false;
    Ctl* d_ctl = nullptr;
    Ctl* h_ctl = nullptr;                 // pinned
    red_event* d_events = nullptr;
    int ev_cap = 0;
    red_event* d_events_hc = nullptr;
    void* d_scratch = nullptr;
    size_t scratch_bytes = 0;
    double* d_seg = nullptr;              // segment sums of free-order split sweeps
    size_t seg_elems = 0;
    // the SI sweep runs on a second stream beside the NSI and NI sweeps (fork / join events)
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // the proposer's pair features run on their own stream beside the step's f0
    hipStream_t learn_side = nullptr;
    hipEvent_t ev_learn_fork = nullptr, ev_learn_join = nullptr;
    bool learn_on_side = false;           // this step's pair features went to learn_side
    LearnPlan learn_plan_cur{};
    SrcMixed* d_src_mixed = nullptr;
    d4* d_src_fp64 = nullptr;             // FP64 one-per-sink sweeps: packed sources
    TpsChunk* d_tps_chunk = nullptr;      // and their chunks
    d4* d_partial = nullptr;
    size_t partial_elems = 0;
    std::vector<void*> user_buffers;      // red_device_alloc
    BodyUpdate* d_updates = nullptr;
    int updates_cap = 0;
    double* d_energy = nullptr;
    double* d_energy_part = nullptr;      // potential partials, n_si x chunks
    StageArgs* d_fused_rows = nullptr;    // fused fixed-step loop: stage arguments, both parities
    FusedOut* d_fused_out = nullptr;
    size_t energy_part_cap = 0;
    // integrator state (integrator.h:359-383)
    Scheme S;
    double t = 0.0, dt_try = 0.0, dt_did = 0.0, last_err = 0.0;
    int n_tried = 0, n_failed = 0, n_passed = 0;
    int attempts_cap = 1000;
    // event counters (pp_disk.h:129-133)
    events::Counters evc;
    std::vector<red_event> raw;           // pair events, canonical order
    bool raw_valid = true;
    std::vector<red_event> sp;
    std::vector<int> rebuild_map;
    // launch configuration
    GravConfig gcfg{};
    PairTest ptest{};
    // optional kernel timing: HIP events around every launch group, resolved on demand
    struct ProfRec { int cls; hipEvent_t a, b; };
    bool profiling = false;
    std::vector<ProfRec> prof;
    double prof_ms[RED_PROF_CLASSES] = {0.0};
    int prof_n[RED_PROF_CLASSES] = {0};
    // captured graphs, per state-buffer parity
    hipGraphExec_t g_attempt[2] = {nullptr, nullptr};
    hipGraphExec_t g_f0[2] = {nullptr, nullptr};
    // learned step-size proposer and step records
    red_mlp* prop_net = nullptr;
    double prop_safety = 0.9, prop_max_log = 3.0;
    int prop_count = 0, prop_first_ok = 0;
    int hist_cap = 0;
    std::vector<float> hist;              // RED_HIST_FIELDS floats per record
    float* d_learn_part = nullptr;        // per-block feature extrema
    unsigned* d_learn_counter = nullptr;
    float* h_prop_out = nullptr;          // pinned, written by the step-start kernel
    float* h_prop_out_dev = nullptr;      // its device-side address
    // accepted-step history: eta[0] = latest tolerance-matching step
    double eta_hist[4] = {0.0, 0.0, 0.0, 0.0};
    int n_eta = 0;
    double last_dt = 0.0, last_err_ratio = 0.0;
    int last_attempts = 0, last_events = 0;
    float prev_log_min_sep = 0.0f, prev_log_max_pert = 0.0f;
    bool have_prev_feat = false;

    int nSI() const { return cnt[0] + cnt[1] + cnt[2] + cnt[3]; }
    // the host mirror of the per-body parameters as the event pipeline sees it
    events::BodyView view()
    {
        return events::BodyView{h_mass.data(), h_radius.data(), h_density.data(), h_cd.data(), h_id.data(),
                                h_type.data(), nMassive()};
    }
    int nNSI() const { return cnt[4] + cnt[5]; }
    int nNI() const { return cnt[6]; }
    int nMassive() const { return nSI() + nNSI(); }

    int fail(int code, const std::string& msg) { err = msg; return code; }
    int hipfail(hipError_t e, const char* where)
    {
        err = std::string(where) + ": " + hipGetErrorString(e);
        return RED_ERR_HIP;
    }
};

#define CTX_HIP(ctx, expr)                                       \
    do {                                                         \
        hipError_t _e = (expr);                                  \
        if (_e != hipSuccess) return (ctx)->hipfail(_e, #expr);  \
    } while (0)

namespace red {
std::recursive_mutex& setup_mutex()
{
    static std::recursive_mutex m;
    return m;
}
}  // namespace red

namespace {

void destroy_graphs(red_ctx* c)
{
    SetupLock lk;
    for (int k = 0; k < 2; k++) {
        if (c->g_attempt[k]) (void)hipGraphExecDestroy(c->g_attempt[k]);
        if (c->g_f0[k]) (void)hipGraphExecDestroy(c->g_f0[k]);
        c->g_attempt[k] = c->g_f0[k] = nullptr;
    }
}

double mass_of_star(const red_ctx* c)
{
    for (int j = 0; j < c->nMassive(); j++)
        if (c->h_type[j] == BT_STAR) return c->h_mass[j];
    return -1.0;
}

void refresh_pair_test(red_ctx* c)
{
    PairTest& P = c->ptest;
    P.radii_f = c->prm.radii_enhance_factor;
    P.hill_f = c->prm.encounter_hill_factor;
    const double ms = mass_of_star(c);
    P.three_mstar = 3.0 * (ms > 0 ? ms : 1.0);
    // the mixed sweep screens only sources past the FP64 head (star, giants)
    double rmax = 0.0, mmax = 0.0;
    for (int i = std::max(0, c->cnt[0] + c->cnt[1]); i < c->n; i++) {
        rmax = std::max(rmax, c->h_radius[i]);
        mmax = std::max(mmax, c->h_mass[i]);
    }
    P.rmax = rmax;
    P.mmax = mmax;
    double rall = 0.0, mns = 0.0;
    for (int i = 0; i < c->n; i++) {
        rall = std::max(rall, c->h_radius[i]);
        if (i >= c->cnt[0]) mns = std::max(mns, c->h_mass[i]);
    }
    P.rmax_all = rall;
    P.mmax_ns = mns;
    // ids only turn negative on the device between refreshes, so the host's active set
    // holds the device's and the flag stays conservative
    P.mass_core = 1;
    for (int i = 0; i < c->nMassive(); i++) {
        if (c->h_id[i] < 0 && !c->padded) continue;   // padded: inactive sources are summed too
        uint64_t bits;
        memcpy(&bits, &c->h_mass[i], sizeof(bits));
        if ((uint32_t)((uint32_t)(bits >> 32) - 0x2D300000u) >= 0x25800000u) P.mass_core = 0;
    }
    P.close_encounter = c->prm.close_encounter;
    P.n_star = c->cnt[0];
    P.enabled = c->prm.inner_steps && (P.radii_f > 0.0 || P.hill_f > 0.0);
    // the astro-centric start term uses the star's current mass (it grows by hit-centrum merges)
    if (c->gcfg.astro) c->gcfg.astro_mstar = c->h_mass[0];
}

// Stage state records: stage 0 = current state, stage k >= 1 = S_k, RK4's stage 4 = y_(n+1).
const d4* stage_r(red_ctx* c, int k)
{
    if (k == 0) return c->d_r[c->cur];
    if (c->S.extra_on_yout && k == 4) return c->d_r[1 - c->cur];
    return c->d_sr[k];
}
const d4* stage_v(red_ctx* c, int k)
{
    if (k == 0) return c->d_v[c->cur];
    if (c->S.extra_on_yout && k == 4) return c->d_v[1 - c->cur];
    return c->d_sv[k];
}

// Records HIP events around fn() on the context stream when profiling is on.
// The SI, NSI and NI sweeps (number_of_bodies.cpp:196-258); the astro-centric SI sweep
// leaves the star out (number_of_bodies.cpp:218-221).
void sweeps_of(const red_ctx* c, Sweep sw[3])
{
    const int nsi = c->nSI(), nm = c->nMassive(), n = c->n;
    const int s0 = c->gcfg.astro ? 1 : 0;
    const int* gate = c->padded ? c->d_ones : c->d_id;
    sw[0] = Sweep{s0, nsi, s0, nm, gate};
    sw[1] = Sweep{nsi, nm, 0, nsi, gate};
    sw[2] = Sweep{nm, n, 0, nm, gate};
}

template <typename F>
hipError_t timed(red_ctx* c, int cls, F fn, hipStream_t s = nullptr)
{
    if (!c->profiling) return fn();
    if (!s) s = c->stream;
    red_ctx::ProfRec rec{cls, nullptr, nullptr};
    RED_HIP_TRY(hipEventCreate(&rec.a));
    RED_HIP_TRY(hipEventCreate(&rec.b));
    RED_HIP_TRY(hipEventRecord(rec.a, s));
    hipError_t e = fn();
    RED_HIP_TRY(hipEventRecord(rec.b, s));
    c->prof.push_back(rec);
    return e;
}

// Force evaluation of stage k on state (r, v) into a: the SI, NSI and NI sweeps
// (pp_disk.cu:609-681).
hipError_t eval_forces(red_ctx* c, int stage, double c_stage, const d4* r, const d4* v, d4* a, bool own_state = true)
{
    // caller-owned positions carry masses the context has not seen
    PairTest P = c->ptest;
    if (!own_state) P.mass_core = 0;
    GravBuffers B;
    B.r = r; B.v = v; B.a = a;
    B.mass = c->d_mass; B.radius = c->d_radius; B.id = c->d_id; B.ctl = c->d_ctl;
    B.scratch = c->d_scratch; B.scratch_bytes = c->scratch_bytes;
    B.src_mixed = c->d_src_mixed;
    B.src_fp64 = c->d_src_fp64;
    B.tps_chunk = c->d_tps_chunk;
    B.partial = c->d_partial;
    B.partial_elems = c->partial_elems;
    B.seg = c->d_seg;
    B.seg_elems = c->seg_elems;
    EventSink E;
    E.buf = c->d_events; E.cap = c->ev_cap; E.stage = stage; E.c_stage = c_stage;
    if (c->gcfg.precision == RED_PREC_MIXED)
        RED_HIP_TRY(timed(c, RED_PROF_PACK, [&] { return launch_pack_mixed(r, c->padded ? c->d_ones : c->d_id, c->d_src_mixed, c->nMassive(), c->stream); }));
    Sweep sw[3];
    sweeps_of(c, sw);
    if (sweep_takes_tps(sw[1], c->gcfg, B) || sweep_takes_tps(sw[2], c->gcfg, B))
        RED_HIP_TRY(timed(c, RED_PROF_PACK, [&] {
            return launch_pack_fp64(r, sw[2].gate, c->d_radius, c->d_src_fp64, c->d_tps_chunk, c->nMassive(),
                                    c->stream);
        }));
    // The three sweeps write disjoint rows of a and append events atomically (the list is
    // sorted afterwards), so the SI sweep -- few sinks, long serial sums -- can run on the
    // side stream while the NSI and NI sweeps fill the chip, unless two of them would share
    // the split scratch.
    // The fork costs two event records and two waits per evaluation (on the host without
    // graph replay, as graph edges with it), which only pays off when the other sweeps are
    // large enough to hide the SI sweep under them.
    const double main_pairs = (double)(sw[1].sink1 - sw[1].sink0) * (sw[1].src1 - sw[1].src0) +
                              (double)(sw[2].sink1 - sw[2].sink0) * (sw[2].src1 - sw[2].src0);
    const bool fork = c->side && sw[0].sink1 > sw[0].sink0 &&
                      main_pairs >= (c->prm.use_graph ? 4194304.0 : 16777216.0) &&
                      !(sweep_wants_split(sw[1], c->gcfg) || sweep_wants_split(sw[2], c->gcfg));
    if (fork) {
        RED_HIP_TRY(hipEventRecord(c->ev_fork, c->stream));
        RED_HIP_TRY(hipStreamWaitEvent(c->side, c->ev_fork, 0));
    }
    for (int k = 0; k < 3; k++) {
        if (sw[k].sink1 <= sw[k].sink0) continue;
        hipStream_t st = (fork && k == 0) ? c->side : c->stream;
        RED_HIP_TRY(timed(c, RED_PROF_SI + k, [&] { return launch_sweep(sw[k], B, c->gcfg, P, E, st); }, st));
    }
    if (fork) {
        RED_HIP_TRY(hipEventRecord(c->ev_join, c->side));
        RED_HIP_TRY(hipStreamWaitEvent(c->stream, c->ev_join, 0));
    }
    return hipSuccess;
}

StageArgs make_args(red_ctx* c, const Row& row, const d4* yr, const d4* yv, d4* outr, d4* outv, int n)
{
    StageArgs A;
    memset(&A, 0, sizeof(A));
    A.yr = yr; A.yv = yv; A.outr = outr; A.outv = outv; A.n = n;
    A.nterms = row.nterms; A.c1 = row.c1; A.c2 = row.c2; A.ctl = c->d_ctl;
    for (int k = 0; k < row.nterms; k++) {
        A.t[k].coef = row.coef[k];
        A.t[k].fv = stage_v(c, row.stage[k]);
        A.t[k].fa = c->d_a[row.stage[k]];
    }
    return A;
}

int n_err_bodies(red_ctx* c) { return c->prm.error_check_for_tp ? c->n : c->nMassive(); }

// The work of one attempt after f0 (int_rungekutta8.cu:735-777 and its siblings).
hipError_t enqueue_attempt(red_ctx* c)
{
    const Scheme& S = c->S;
    const d4* yr = c->d_r[c->cur];
    const d4* yv = c->d_v[c->cur];
    d4* outr = c->d_r[1 - c->cur];
    d4* outv = c->d_v[1 - c->cur];
    for (int r = 1; r < S.n_main; r++) {
        StageArgs A = make_args(c, S.rows[r], yr, yv, c->d_sr[r], c->d_sv[r], c->n);
        RED_HIP_TRY(timed(c, RED_PROF_STAGE, [&] { return launch_stage(S.rows[r].form, A, c->stream); }));
        RED_HIP_TRY(eval_forces(c, r, S.c[r], c->d_sr[r], c->d_sv[r], c->d_a[r]));
    }
    StageArgs A = make_args(c, S.np1, yr, yv, outr, outv, c->n);
    RED_HIP_TRY(timed(c, RED_PROF_STAGE, [&] { return launch_stage(S.np1.form, A, c->stream); }));
    if (c->prm.adaptive && S.err_form >= 0) {
        for (int r = S.n_main; r < S.n_stage; r++) {
            if (S.extra_on_yout) {
                RED_HIP_TRY(eval_forces(c, r, S.c[r], outr, outv, c->d_a[r]));
            } else {
                StageArgs B = make_args(c, S.rows[r], yr, yv, c->d_sr[r], c->d_sv[r], c->n);
                RED_HIP_TRY(timed(c, RED_PROF_STAGE, [&] { return launch_stage(S.rows[r].form, B, c->stream); }));
                RED_HIP_TRY(eval_forces(c, r, S.c[r], c->d_sr[r], c->d_sv[r], c->d_a[r]));
            }
        }
        StageArgs E = make_args(c, S.err, yr, yv, nullptr, nullptr, n_err_bodies(c));
        RED_HIP_TRY(timed(c, RED_PROF_ERR, [&] { return launch_err(S.err_form, E, c->stream); }));
    }
    return hipSuccess;
}

hipError_t enqueue_f0(red_ctx* c)
{
    return eval_forces(c, 0, c->S.c[0], c->d_r[c->cur], c->d_v[c->cur], c->d_a[0]);
}

// Captures fn into a graph for the current buffer parity on first use, then replays it.
template <typename F>
hipError_t run_maybe_graph(red_ctx* c, hipGraphExec_t* slot, F fn)
{
    if (!c->prm.use_graph) return fn(c);
    if (!*slot) {
        SetupLock lk;
        hipGraph_t g = nullptr;
        RED_HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = fn(c);
        hipError_t e2 = hipStreamEndCapture(c->stream, &g);
        if (e != hipSuccess) { if (g) (void)hipGraphDestroy(g); return e; }
        RED_HIP_TRY(e2);
        e = hipGraphInstantiate(slot, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        RED_HIP_TRY(e);
    }
    return hipGraphLaunch(*slot, c->stream);
}

int read_ctl(red_ctx* c)
{
    CTX_HIP(c, hipMemcpyAsync(c->h_ctl, c->d_ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
    CTX_HIP(c, hipStreamSynchronize(c->stream));
    return RED_OK;
}

double bits_to_double(unsigned long long b)
{
    double d;
    memcpy(&d, &b, 8);
    return d;
}

// Copies the pair events of the device buffer to the host in canonical order.
int fetch_raw_events(red_ctx* c)
{
    if (c->raw_valid) return RED_OK;
    int rc = read_ctl(c);
    if (rc) return rc;
    const int n = c->h_ctl->n_events;
    if (n > c->ev_cap)
        return c->fail(RED_ERR_EVENT_OVERFLOW, "pair events exceed the event capacity (" + std::to_string(n) + " > " +
                                                   std::to_string(c->ev_cap) + ")");
    c->raw.resize(n);
    if (n > 0) {
        CTX_HIP(c, stream_copy(c->stream, c->raw.data(), c->d_events, sizeof(red_event) * n, hipMemcpyDeviceToHost));
        events::canonical_sort(c->raw);
    }
    c->raw_valid = true;
    return RED_OK;
}

int apply_updates(red_ctx* c, const std::vector<BodyUpdate>& ups)
{
    if (ups.empty()) return RED_OK;
    if ((int)ups.size() > c->updates_cap) {
        SetupLock lk;
        if (c->d_updates) (void)hipFree(c->d_updates);
        c->updates_cap = (int)ups.size() * 2;
        CTX_HIP(c, dmalloc(&c->d_updates, c->updates_cap));
    }
    CTX_HIP(c, hipMemcpyAsync(c->d_updates, ups.data(), sizeof(BodyUpdate) * ups.size(), hipMemcpyHostToDevice, c->stream));
    CTX_HIP(c, launch_apply_updates(c->d_updates, (int)ups.size(), c->d_r[c->cur], c->d_v[c->cur], c->d_mass,
                                    c->d_radius, c->d_id, c->stream));
    CTX_HIP(c, hipStreamSynchronize(c->stream));
    refresh_pair_test(c);
    destroy_graphs(c);   // captured launches hold the pair-test parameters by value
    return RED_OK;
}

int upload_all(red_ctx* c)
{
    std::vector<d4> r(c->n), v(c->n);
    // keep the caller's values: only the parameters and ids are re-uploaded here
    (void)r; (void)v;
    CTX_HIP(c, stream_copy(c->stream, c->d_mass, c->h_mass.data(), sizeof(double) * c->n, hipMemcpyHostToDevice));
    CTX_HIP(c, stream_copy(c->stream, c->d_radius, c->h_radius.data(), sizeof(double) * c->n, hipMemcpyHostToDevice));
    CTX_HIP(c, stream_copy(c->stream, c->d_id, c->h_id.data(), sizeof(int) * c->n, hipMemcpyHostToDevice));
    CTX_HIP(c, stream_copy(c->stream, c->d_type, c->h_type.data(), sizeof(int) * c->n, hipMemcpyHostToDevice));
    return RED_OK;
}

int download_state(red_ctx* c, std::vector<d4>& r, std::vector<d4>& v)
{
    r.resize(c->n); v.resize(c->n);
    if (c->n == 0) return RED_OK;
    CTX_HIP(c, hipStreamSynchronize(c->stream));
    CTX_HIP(c, stream_copy(c->stream, r.data(), c->d_r[c->cur], sizeof(d4) * c->n, hipMemcpyDeviceToHost));
    CTX_HIP(c, stream_copy(c->stream, v.data(), c->d_v[c->cur], sizeof(d4) * c->n, hipMemcpyDeviceToHost));
    return RED_OK;
}

}  // namespace

extern "C" {

const char* red_version(void) { return "rednative 0.1 (gfx950)"; }

const char* red_last_error(const red_ctx* ctx)
{
    if (!ctx) return "no context";
    return ctx->err.c_str();
}

int red_ctx_destroy(red_ctx* c)
{
    if (!c) return RED_OK;
    SetupLock lk;
    destroy_graphs(c);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);               // work forked off and
    if (c->learn_side) (void)hipStreamSynchronize(c->learn_side);   // not yet joined
    for (int k = 0; k < 2; k++) { (void)hipFree(c->d_r[k]); (void)hipFree(c->d_v[k]); }
    for (auto p : c->d_sr) (void)hipFree(p);
    for (auto p : c->d_sv) (void)hipFree(p);
    for (auto p : c->d_a) (void)hipFree(p);
    (void)hipFree(c->d_mass); (void)hipFree(c->d_radius); (void)hipFree(c->d_id); (void)hipFree(c->d_type);
    (void)hipFree(c->d_ones);
    (void)hipFree(c->d_ctl); (void)hipFree(c->d_events); (void)hipFree(c->d_events_hc); (void)hipFree(c->d_scratch);
    (void)hipFree(c->d_src_mixed); (void)hipFree(c->d_src_fp64); (void)hipFree(c->d_tps_chunk); (void)hipFree(c->d_updates); (void)hipFree(c->d_energy); (void)hipFree(c->d_energy_part);
    (void)hipFree(c->d_partial); (void)hipFree(c->d_seg);
    (void)hipFree(c->d_learn_part); (void)hipFree(c->d_learn_counter);
    (void)hipFree(c->d_fused_rows); (void)hipFree(c->d_fused_out);
    for (void* p : c->user_buffers) (void)hipFree(p);
    if (c->h_ctl) (void)hipHostFree(c->h_ctl);
    if (c->h_prop_out) (void)hipHostFree(c->h_prop_out);
    if (c->ev_learn_fork) (void)hipEventDestroy(c->ev_learn_fork);
    if (c->ev_learn_join) (void)hipEventDestroy(c->ev_learn_join);
    if (c->learn_side) (void)hipStreamDestroy(c->learn_side);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return RED_OK;
}

int red_ctx_create(const red_bodies* b, const red_params* p, red_ctx** out)
{
    if (!b || !p || !out) return RED_ERR_ARG;
    SetupLock lk;
    *out = nullptr;
    red_ctx* c = new red_ctx();
    c->prm = *p;
    int n = 0;
    for (int k = 0; k < 7; k++) {
        if (b->counts[k] < 0) { delete c; return RED_ERR_ARG; }
        c->cnt[k] = b->counts[k];
        n += b->counts[k];
    }
    if (p->integrator < RED_EULER || p->integrator > RED_RKF78) { delete c; return RED_ERR_UNSUPPORTED; }
    if (p->frame_center != RED_FRAME_BARY && p->frame_center != RED_FRAME_ASTRO) { delete c; return RED_ERR_ARG; }
    if (p->precision != RED_PREC_FP64 && p->precision != RED_PREC_MIXED) { delete c; return RED_ERR_ARG; }
    if (p->strategy < 0 || p->strategy > 4) { delete c; return RED_ERR_ARG; }
    if (!p->inner_steps && (p->radii_enhance_factor > 0.0 || p->encounter_hill_factor > 0.0)) { delete c; return RED_ERR_UNSUPPORTED; }
    if (n > 0 && (!b->id || !b->type || !b->mass || !b->radius || !b->density || !b->cd || !b->r || !b->v)) {
        delete c;
        return RED_ERR_ARG;
    }
    c->n = n;
    c->cap = std::max(64, (n + 63) / 64 * 64);
    c->h_mass.assign(b->mass, b->mass + n);
    c->h_radius.assign(b->radius, b->radius + n);
    c->h_density.assign(b->density, b->density + n);
    c->h_cd.assign(b->cd, b->cd + n);
    c->h_id.assign(b->id, b->id + n);
    c->h_type.assign(b->type, b->type + n);
    c->h_mig_type.assign(n, 0);
    c->h_mig_stop.assign(n, 0.0);
    c->h_epoch.assign(n, 0.0);
    if (b->mig_type) c->h_mig_type.assign(b->mig_type, b->mig_type + n);
    if (b->mig_stop_at) c->h_mig_stop.assign(b->mig_stop_at, b->mig_stop_at + n);
    if (b->epoch) c->h_epoch.assign(b->epoch, b->epoch + n);
    if (n > 0 && mass_of_star(c) < 0.0 && (p->encounter_hill_factor > 0.0 || p->close_encounter)) {
        delete c;
        return RED_ERR_NO_STAR;
    }
    c->S = make_scheme(p->integrator, p->adaptive != 0);
    c->dt_try = p->dt_day * GAUSS;   // integrator.cu:29
    c->t = 0.0;
    // launch configuration
    c->gcfg.precision = p->precision;
    c->gcfg.n_fp64_src = c->cnt[0] + c->cnt[1];
    c->gcfg.block = p->block > 0 ? std::min(256, std::max(64, p->block / 64 * 64)) : 256;
    c->gcfg.block_seq = p->block >= 256 ? (p->block >= 1024 ? 1024 : (p->block >= 512 ? 512 : 256)) : 0;
    c->gcfg.seq_sinks = p->sinks_per_thread >= 4 ? p->sinks_per_thread : 0;
    c->gcfg.spt = p->sinks_per_thread > 0 ? p->sinks_per_thread : (p->precision == RED_PREC_MIXED ? 2 : 1);
    // FP64 with the reference's sequential sums: producer/consumer sweeps by default
    c->gcfg.strategy = p->strategy > 0 ? p->strategy : (p->precision == RED_PREC_FP64 && !p->fp64_split ? 3 : 1);
    c->gcfg.kahan_pairs = p->kahan_pairs;
    c->gcfg.split_max_sinks = 2048;
    c->gcfg.fp64_split = p->fp64_split;
    c->padded = p->padded_storage != 0;
    if (c->padded && p->frame_center == RED_FRAME_ASTRO) {
        // the padded kernel (pp_disk.cu:144-215) is the reference's GPU path, which has no
        // astro-centric indirect term; the astro frame here follows its CPU path
        int rc = c->fail(RED_ERR_UNSUPPORTED, "padded storage runs in the barycentric frame only");
        *out = c;
        return rc;
    }
    if (p->frame_center == RED_FRAME_ASTRO) {
        // The reference's NSI and NI sweeps run the SI code with bounds that keep the star as
        // source 0 at the origin (number_of_bodies.cpp:227-257, pp_disk.cu:574-583): its
        // indirect term is 0/0 there, so only self-interacting disks are defined.
        if (c->nNSI() > 0 || c->nNI() > 0 || p->precision != RED_PREC_FP64) {
            int rc = c->fail(RED_ERR_UNSUPPORTED, "frame_center = astro needs FP64 and self-interacting bodies only");
            *out = c;
            return rc;
        }
        // the star is body 0 and rests at the origin (tools.cpp:180-199 tests body 0's type)
        if (n == 0 || c->h_type[0] != BT_STAR) { delete c; return RED_ERR_NO_STAR; }
        c->gcfg.astro = 1;
    }
    c->gcfg.source_parts = p->source_parts;
    refresh_pair_test(c);

    hipError_t e = hipSetDevice(p->device);
    if (e != hipSuccess) { int rc = c->hipfail(e, "hipSetDevice"); *out = c; return rc; }
    *out = c;
    CTX_HIP(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    const int cap = c->cap;
    for (int k = 0; k < 2; k++) {
        CTX_HIP(c, dmalloc(&c->d_r[k], cap));
        CTX_HIP(c, dmalloc(&c->d_v[k], cap));
        CTX_HIP(c, hipMemsetAsync(c->d_r[k], 0, sizeof(d4) * cap, c->stream));
        CTX_HIP(c, hipMemsetAsync(c->d_v[k], 0, sizeof(d4) * cap, c->stream));
    }
    const int ns = c->S.n_stage;
    c->d_sr.assign(ns, nullptr); c->d_sv.assign(ns, nullptr); c->d_a.assign(ns, nullptr);
    for (int k = 0; k < ns; k++) {
        CTX_HIP(c, dmalloc(&c->d_a[k], cap));
        CTX_HIP(c, hipMemsetAsync(c->d_a[k], 0, sizeof(d4) * cap, c->stream));
        if (k >= 1) {
            CTX_HIP(c, dmalloc(&c->d_sr[k], cap));
            CTX_HIP(c, dmalloc(&c->d_sv[k], cap));
        }
    }
    CTX_HIP(c, dmalloc(&c->d_mass, cap));
    CTX_HIP(c, dmalloc(&c->d_radius, cap));
    CTX_HIP(c, dmalloc(&c->d_id, cap));
    CTX_HIP(c, dmalloc(&c->d_type, cap));
    if (c->padded) {
        std::vector<int> ones(std::max(cap, 1), 1);
        CTX_HIP(c, dmalloc(&c->d_ones, cap));
        CTX_HIP(c, stream_copy(c->stream, c->d_ones, ones.data(), sizeof(int) * ones.size(), hipMemcpyHostToDevice));
    }
    CTX_HIP(c, dmalloc(&c->d_ctl, 1));
    CTX_HIP(c, hipMemsetAsync(c->d_ctl, 0, sizeof(Ctl), c->stream));
    CTX_HIP(c, hipHostMalloc((void**)&c->h_ctl, sizeof(Ctl), hipHostMallocDefault));
    memset(c->h_ctl, 0, sizeof(Ctl));
    c->ev_cap = p->event_capacity > 0 ? p->event_capacity : std::max(4096, n);
    CTX_HIP(c, dmalloc(&c->d_events, c->ev_cap));
    CTX_HIP(c, dmalloc(&c->d_events_hc, std::max(1, n)));
    if (p->precision == RED_PREC_MIXED) CTX_HIP(c, dmalloc(&c->d_src_mixed, cap));
    if (p->precision == RED_PREC_FP64 && !p->fp64_split && c->gcfg.strategy >= 3) {
        CTX_HIP(c, dmalloc(&c->d_src_fp64, cap));
        CTX_HIP(c, dmalloc(&c->d_tps_chunk, (cap + TPS_CHUNK - 1) / TPS_CHUNK));
    }
    if (p->precision == RED_PREC_MIXED || p->fp64_split) {
        // per-part sums: up to 64 parts of the largest sweep, capped at 1 GiB
        c->partial_elems = std::min((size_t)64 * cap, ((size_t)1 << 30) / sizeof(d4));
        CTX_HIP(c, dmalloc(&c->d_partial, c->partial_elems));
    }
    CTX_HIP(c, dmalloc(&c->d_energy, 6 * (size_t)cap));
    // split-sweep scratch: the largest sweep that forms its terms there; segment sums for
    // the free-order ones
    {
        size_t need = 0, rows = 0;
        Sweep sws[3];
        sweeps_of(c, sws);
        for (const Sweep& s : sws) {
            if (!sweep_wants_split(s, c->gcfg)) continue;
            need = std::max(need, split_scratch_bytes(s));
            rows = std::max(rows, (size_t)(s.sink1 - s.sink0) * 3);
        }
        const size_t limit = (size_t)4 << 30;
        if (need > limit) need = 0;   // too large: those sweeps use the broadcast path
        if (need > 0) CTX_HIP(c, hipMalloc(&c->d_scratch, need));
        c->scratch_bytes = need;
        if (need > 0 && !(p->precision == RED_PREC_FP64 && !p->fp64_split && c->gcfg.strategy >= 3)) {
            c->seg_elems = rows * SPLIT_SEG_MAX;
            CTX_HIP(c, hipMalloc(&c->d_seg, sizeof(double) * c->seg_elems));
        }
    }
    if (c->nSI() > 0 && c->nNSI() + c->nNI() > 0) {
        CTX_HIP(c, hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
        CTX_HIP(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        CTX_HIP(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    }
    // upload the state (w of r = mass)
    {
        std::vector<d4> r(n), v(n);
        for (int i = 0; i < n; i++) {
            r[i] = d4{b->r[3 * i], b->r[3 * i + 1], b->r[3 * i + 2], b->mass[i]};
            v[i] = d4{b->v[3 * i], b->v[3 * i + 1], b->v[3 * i + 2], 0.0};
        }
        if (n > 0) {
            CTX_HIP(c, stream_copy(c->stream, c->d_r[0], r.data(), sizeof(d4) * n, hipMemcpyHostToDevice));
            CTX_HIP(c, stream_copy(c->stream, c->d_v[0], v.data(), sizeof(d4) * n, hipMemcpyHostToDevice));
        }
    }
    int rc = upload_all(c);
    if (rc) return rc;
    CTX_HIP(c, hipDeviceSynchronize());
    return RED_OK;
}

int red_get_counts(red_ctx* c, int counts[7])
{
    if (!c || !counts) return RED_ERR_ARG;
    for (int k = 0; k < 7; k++) counts[k] = c->cnt[k];
    return RED_OK;
}

int red_get_time(red_ctx* c, double* t) { if (!c || !t) return RED_ERR_ARG; *t = c->t; return RED_OK; }
int red_set_time(red_ctx* c, double t) { if (!c) return RED_ERR_ARG; c->t = t; return RED_OK; }
int red_get_dt(red_ctx* c, double* dt) { if (!c || !dt) return RED_ERR_ARG; *dt = c->dt_try; return RED_OK; }
int red_set_dt(red_ctx* c, double dt) { if (!c) return RED_ERR_ARG; c->dt_try = dt; return RED_OK; }

int red_get_state(red_ctx* c, double* r, double* v, double* mass, double* radius, double* density, double* cd,
                  int* id, int* type)
{
    if (!c) return RED_ERR_ARG;
    std::vector<d4> R, V;
    if (r || v) {
        int rc = download_state(c, R, V);
        if (rc) return rc;
    }
    for (int i = 0; i < c->n; i++) {
        if (r) { r[3 * i] = R[i].x; r[3 * i + 1] = R[i].y; r[3 * i + 2] = R[i].z; }
        if (v) { v[3 * i] = V[i].x; v[3 * i + 1] = V[i].y; v[3 * i + 2] = V[i].z; }
        if (mass) mass[i] = c->h_mass[i];
        if (radius) radius[i] = c->h_radius[i];
        if (density) density[i] = c->h_density[i];
        if (cd) cd[i] = c->h_cd[i];
        if (id) id[i] = c->h_id[i];
        if (type) type[i] = c->h_type[i];
    }
    return RED_OK;
}

int red_set_state(red_ctx* c, const double* r, const double* v)
{
    if (!c || !r || !v) return RED_ERR_ARG;
    std::vector<d4> R(c->n), V(c->n);
    for (int i = 0; i < c->n; i++) {
        R[i] = d4{r[3 * i], r[3 * i + 1], r[3 * i + 2], c->h_mass[i]};
        V[i] = d4{v[3 * i], v[3 * i + 1], v[3 * i + 2], 0.0};
    }
    CTX_HIP(c, hipStreamSynchronize(c->stream));
    if (c->n > 0) {
        CTX_HIP(c, stream_copy(c->stream, c->d_r[c->cur], R.data(), sizeof(d4) * c->n, hipMemcpyHostToDevice));
        CTX_HIP(c, stream_copy(c->stream, c->d_v[c->cur], V.data(), sizeof(d4) * c->n, hipMemcpyHostToDevice));
    }
    c->n_eta = 0;                 // a new state starts a new step history
    c->have_prev_feat = false;
    return RED_OK;
}

int red_state_device(red_ctx* c, double** r, double** v)
{
    if (!c) return RED_ERR_ARG;
    if (r) *r = (double*)c->d_r[c->cur];
    if (v) *v = (double*)c->d_v[c->cur];
    return RED_OK;
}

int red_sync(red_ctx* c)
{
    if (!c) return RED_ERR_ARG;
    CTX_HIP(c, hipStreamSynchronize(c->stream));
    return RED_OK;
}

int red_accel(red_ctx* c, double t, int stage, const double* r_dev, const double* v_dev, double* a_dev)
{
    if (!c || !r_dev || !v_dev || !a_dev) return RED_ERR_ARG;
    hipLaunchKernelGGL(k_set_ctl, dim3(1), dim3(1), 0, c->stream, c->d_ctl, t, 0.0, 0, 1, 0, 0);
    CTX_HIP(c, hipGetLastError());
    c->raw_valid = false;
    CTX_HIP(c, eval_forces(c, stage, 0.0, (const d4*)r_dev, (const d4*)v_dev, (d4*)a_dev, false));
    return RED_OK;
}

namespace {

float clampf(double x, double lo, double hi) { return (float)std::min(hi, std::max(lo, x)); }

int learn_buffers(red_ctx* c)
{
    if (c->d_learn_part && c->d_learn_counter && c->h_prop_out && c->learn_side) return RED_OK;
    SetupLock lk;
    if (!c->learn_side) {
        CTX_HIP(c, hipStreamCreateWithFlags(&c->learn_side, hipStreamNonBlocking));
        CTX_HIP(c, hipEventCreateWithFlags(&c->ev_learn_fork, hipEventDisableTiming));
        CTX_HIP(c, hipEventCreateWithFlags(&c->ev_learn_join, hipEventDisableTiming));
    }
    if (!c->d_learn_part) CTX_HIP(c, hipMalloc(&c->d_learn_part, sizeof(float) * 3 * LEARN_MAX_BLOCKS));
    if (!c->d_learn_counter) {
        CTX_HIP(c, hipMalloc(&c->d_learn_counter, sizeof(unsigned)));
        CTX_HIP(c, hipMemsetAsync(c->d_learn_counter, 0, sizeof(unsigned), c->stream));
        CTX_HIP(c, hipStreamSynchronize(c->stream));
    }
    if (!c->h_prop_out) {
        CTX_HIP(c, hipHostMalloc((void**)&c->h_prop_out, sizeof(float) * RED_HIST_FIELDS, hipHostMallocDefault));
        CTX_HIP(c, hipHostGetDevicePointer((void**)&c->h_prop_out_dev, c->h_prop_out, 0));
    }
    return RED_OK;
}

// Launches the step-start kernel: the state features of (y_n, f0), the network (if
// proposing) and the first trial step, written into the control block after the attempt's
// own fields.  The feature vector and the network output reach the pinned buffer with the
// attempt's read-back; nothing waits for them on the host.
int learn_launch(red_ctx* c, bool propose, double eta_n)
{
    int rc = learn_buffers(c);
    if (rc) return rc;
    ProposeArgs A;
    memset(&A, 0, sizeof(A));
    if (c->n_eta >= 4) {
        const double* e = c->eta_hist;
        A.x[0] = clampf(std::log10(c->last_dt), -30.0, 30.0);
        A.x[1] = clampf(std::log10(c->last_err_ratio), -30.0, 30.0);
        A.x[2] = clampf(std::log(e[0] / e[1]), -30.0, 30.0);
        A.x[3] = clampf(std::log(e[1] / e[2]), -30.0, 30.0);
        A.x[4] = (float)(c->last_attempts - 1);
        A.x[5] = (float)std::log1p((double)c->last_events);
        A.x[6] = clampf(std::log(e[2] / e[3]), -30.0, 30.0);
    }
    A.prev_log_min_sep = c->prev_log_min_sep;
    A.prev_log_max_pert = c->prev_log_max_pert;
    A.have_prev = c->have_prev_feat ? 1 : 0;
    LearnBuffers B{c->d_learn_part, c->d_learn_counter, c->h_prop_out_dev};
    if (c->learn_on_side) CTX_HIP(c, hipStreamWaitEvent(c->stream, c->ev_learn_join, 0));   // learn_pairs' launch
    CTX_HIP(c, timed(c, RED_PROF_LEARN, [&] {
        return launch_learn_tail(propose ? c->prop_net : nullptr, A, c->learn_plan_cur, c->d_r[c->cur], c->d_a[0],
                                 c->d_id, c->nSI(), c->nMassive(), c->n, B, c->d_ctl, propose ? 1 : 0, eta_n,
                                 c->prop_safety, c->prop_max_log, c->stream);
    }));
    return RED_OK;
}

// At the start of a step, before f0: the pair features need only y_n, so their launch runs
// on the proposer's stream beside the force evaluation (learn_launch joins it) when that
// evaluation is long enough to hide it; a small disk's runs in line, ahead of f0 (there the
// fork and join between the streams cost more than the launch itself).
constexpr long LEARN_SIDE_MIN_PAIRS = 1L << 22;

int learn_pairs(red_ctx* c)
{
    int rc = learn_buffers(c);
    if (rc) return rc;
    c->learn_plan_cur = learn_plan(c->nSI(), c->nMassive(), c->n);
    LearnBuffers B{c->d_learn_part, c->d_learn_counter, c->h_prop_out_dev};
    c->learn_on_side = (long)c->nMassive() * (long)c->n >= LEARN_SIDE_MIN_PAIRS;
    hipStream_t s = c->learn_on_side ? c->learn_side : c->stream;
    if (c->learn_on_side) {
        CTX_HIP(c, hipEventRecord(c->ev_learn_fork, c->stream));
        CTX_HIP(c, hipStreamWaitEvent(c->learn_side, c->ev_learn_fork, 0));
    }
    CTX_HIP(c, timed(c, RED_PROF_LEARN, [&] {
        return launch_learn_pairs(c->learn_plan_cur, c->d_r[c->cur], c->d_id, c->nSI(), c->nMassive(), c->n, B, s);
    }, s));
    if (c->learn_on_side) CTX_HIP(c, hipEventRecord(c->ev_learn_join, c->learn_side));
    return RED_OK;
}

// After a synchronisation following learn_launch: the features and the network output.
void learn_collect(red_ctx* c, float* rec)
{
    for (int k = 0; k <= MLP_NF; k++) rec[k < MLP_NF ? k : 13] = c->h_prop_out[k];
    c->prev_log_min_sep = c->h_prop_out[FEAT_LOG_MIN_SEP];
    c->prev_log_max_pert = c->h_prop_out[FEAT_LOG_MAX_PERT];
    c->have_prev_feat = true;
}

// Accepted step: shift the history (an unusable error estimate restarts it).
void learn_accept(red_ctx* c, double dt, double err, int attempts, int events)
{
    const double eta = tools::tolerance_step(dt, err, c->prm.tolerance, c->S.exponent);
    if (!(eta > 0.0) || !std::isfinite(eta) || !(dt > 0.0)) {
        c->n_eta = 0;
        return;
    }
    for (int k = 3; k > 0; k--) c->eta_hist[k] = c->eta_hist[k - 1];
    c->eta_hist[0] = eta;
    c->n_eta = std::min(4, c->n_eta + 1);
    c->last_dt = dt;
    c->last_err_ratio = err / c->prm.tolerance;
    c->last_attempts = attempts;
    c->last_events = events;
}

}  // namespace

int red_history_enable(red_ctx* c, int capacity)
{
    if (!c || capacity < 0) return RED_ERR_ARG;
    c->hist_cap = capacity;
    c->hist.clear();
    c->hist.reserve((size_t)std::min(capacity, 1 << 20) * RED_HIST_FIELDS);
    return RED_OK;
}

int red_history_get(red_ctx* c, float* records, int cap, int* n)
{
    if (!c || cap < 0 || (cap > 0 && !records)) return RED_ERR_ARG;
    const int have = (int)(c->hist.size() / RED_HIST_FIELDS);
    const int k = std::min(have, cap);
    if (k > 0) memcpy(records, c->hist.data(), sizeof(float) * (size_t)k * RED_HIST_FIELDS);
    c->hist.erase(c->hist.begin(), c->hist.begin() + (size_t)k * RED_HIST_FIELDS);
    if (n) *n = k;
    return RED_OK;
}

int red_set_proposer(red_ctx* c, red_mlp* net, double safety, double max_log_change)
{
    if (!c) return RED_ERR_ARG;
    if (net) {
        if (!(safety > 0.0 && safety <= 2.0) || !(max_log_change > 0.0 && max_log_change <= 20.0))
            return c->fail(RED_ERR_ARG, "proposer safety must lie in (0, 2] and max_log_change in (0, 20]");
        if (mlp_device(net) != c->prm.device)
            return c->fail(RED_ERR_ARG, "the network lives on another device");
        if (!(c->prm.adaptive && c->S.err_form >= 0))
            return c->fail(RED_ERR_UNSUPPORTED, "a step-size proposer needs an adaptive method");
    }
    c->prop_net = net;
    c->prop_safety = safety;
    c->prop_max_log = max_log_change;
    return RED_OK;
}

int red_proposer_stats(red_ctx* c, int out[2])
{
    if (!c || !out) return RED_ERR_ARG;
    out[0] = c->prop_count;
    out[1] = c->prop_first_ok;
    return RED_OK;
}

int red_step(red_ctx* c, red_step_info* info)
{
    if (!c) return RED_ERR_ARG;
    const Scheme& S = c->S;
    const bool adaptive = c->prm.adaptive && S.err_form >= 0;
    const bool learning = adaptive && (c->prop_net || c->hist_cap > 0);
    // f0 = f(t, y_n) (int_rungekutta8.cu:715-724)
    hipLaunchKernelGGL(k_set_ctl, dim3(1), dim3(1), 0, c->stream, c->d_ctl, c->t, c->dt_try, 0, 1, 0, 0);
    CTX_HIP(c, hipGetLastError());
    if (learning) {
        int rc = learn_pairs(c);
        if (rc) return rc;
    }
    CTX_HIP(c, run_maybe_graph(c, &c->g_f0[c->cur], enqueue_f0));
    c->raw_valid = false;
    // learned proposal of the first trial step (the reference controller's step otherwise):
    // the step-start kernel runs on the stream ahead of the first attempt, which takes its
    // step from the device; nothing waits for it on the host
    float rec[RED_HIST_FIELDS] = {0.0f};
    const bool rec_open = learning && c->n_eta >= 4;
    const bool proposed = rec_open && c->prop_net != nullptr;
    const double eta_n = rec_open ? c->eta_hist[0] : 0.0;
    double max_err = 0.0;
    int iter = 0, attempts = 0;
    bool done = false;
    while (!done) {
        if (iter >= c->attempts_cap)
            return c->fail(RED_ERR_STEP_LIMIT, "step-size control did not converge within " + std::to_string(iter) + " attempts");
        c->dt_did = c->dt_try;
        hipLaunchKernelGGL(k_set_ctl, dim3(1), dim3(1), 0, c->stream, c->d_ctl, c->t, c->dt_try, iter, 1,
                           S.clear_events_each_attempt ? 1 : 0, 0);
        CTX_HIP(c, hipGetLastError());
        if (iter == 0 && learning) {
            int rc = learn_launch(c, proposed, eta_n);
            if (rc) return rc;
        }
        CTX_HIP(c, run_maybe_graph(c, &c->g_attempt[c->cur], enqueue_attempt));
        if (!adaptive) {
            iter++;
            break;
        }
        int rc = read_ctl(c);
        if (rc) return rc;
        if (iter == 0 && learning) learn_collect(c, rec);
        if (iter == 0 && proposed) {
            c->dt_did = c->dt_try = c->h_ctl->dt;   // the step the device proposed and tried
            c->prop_count++;
        }
        if (c->h_ctl->nan_flag) return c->fail(RED_ERR_NAN, "non-finite error estimate");
        if (c->h_ctl->n_events > c->ev_cap)
            return c->fail(RED_ERR_EVENT_OVERFLOW, "pair events exceed the event capacity");
        // integrator.cu:255 and the controllers (int_rungekutta8.cu:776-777 etc.)
        max_err = tools::controller_update(&c->dt_try, S.lambda, bits_to_double(c->h_ctl->err_bits),
                                           c->prm.tolerance, S.exponent);
        c->last_err = max_err;
        if (attempts++ == 0 && rec_open) {
            const double eta1 = tools::tolerance_step(c->dt_did, max_err, c->prm.tolerance, S.exponent);
            rec[12] = clampf(std::log(eta1 / eta_n), -30.0, 30.0);
            rec[14] = clampf(std::log(c->dt_did / eta_n), -30.0, 30.0);
            rec[15] = max_err > c->prm.tolerance ? 0.0f : 1.0f;
            if (c->hist_cap > 0 && c->hist.size() < (size_t)c->hist_cap * RED_HIST_FIELDS)
                c->hist.insert(c->hist.end(), rec, rec + RED_HIST_FIELDS);
            if (proposed && !(max_err > c->prm.tolerance)) c->prop_first_ok++;
        }
        if (S.accept_on_event && c->h_ctl->n_events > 0) {
            if (c->dt_try < c->dt_did) c->dt_try = c->dt_did;
            break;   // leaves before iter++ as int_rungekutta4.cu:255-264 does
        }
        iter++;
        done = !(max_err > c->prm.tolerance);
    }
    if (adaptive && (c->prop_net || c->hist_cap > 0))
        learn_accept(c, c->dt_did, max_err, attempts, c->h_ctl->n_events);
    // integrator.cu:258-263
    c->n_tried += iter;
    c->n_failed += (iter - 1);
    c->n_passed++;
    c->t += c->dt_did;
    c->cur = 1 - c->cur;   // pp_disk::swap (pp_disk.cu:1043-1057)
    if (info) {
        info->dt_did = c->dt_did;
        info->dt_next = c->dt_try;
        info->err = max_err;
        info->tries = iter;
        info->n_events = adaptive ? c->h_ctl->n_events : -1;
    }
    return RED_OK;
}

int red_run_fixed(red_ctx* c, int max_steps, double t_stop, double output_interval, double* ps, int* steps_done,
                  int* stop_reason)
{
    if (!c || !ps || !steps_done || !stop_reason) return RED_ERR_ARG;
    *steps_done = 0;
    *stop_reason = 0;
    const Scheme& S = c->S;
    if (c->prm.adaptive && S.err_form >= 0)
        return c->fail(RED_ERR_UNSUPPORTED, "red_run_fixed: the integrator controls its step size");
    if (c->gcfg.precision != RED_PREC_FP64 || c->gcfg.fp64_split || c->gcfg.astro)
        return c->fail(RED_ERR_UNSUPPORTED,
                       "red_run_fixed: needs FP64 sums in the reference order and the barycentric frame");
    if (c->n > FUSED_MAX_BODIES || S.n_main > FUSED_MAX_ROWS)
        return c->fail(RED_ERR_UNSUPPORTED, "red_run_fixed: more than 4096 bodies");
    {
        Sweep sw[3];
        sweeps_of(c, sw);
        long pairs = 0;
        for (const Sweep& s : sw)
            if (s.sink1 > s.sink0) pairs += (long)(s.sink1 - s.sink0) * (s.src1 - s.src0);
        if (pairs > FUSED_MAX_PAIRS)
            return c->fail(RED_ERR_UNSUPPORTED, "red_run_fixed: more than " + std::to_string(FUSED_MAX_PAIRS) +
                                                    " interacting pairs per evaluation (stepwise is faster)");
    }
    if (!c->d_fused_rows || !c->d_fused_out) {
        SetupLock lk;
        if (!c->d_fused_rows) CTX_HIP(c, dmalloc(&c->d_fused_rows, 2 * FUSED_MAX_ROWS));
        if (!c->d_fused_out) CTX_HIP(c, dmalloc(&c->d_fused_out, 1));
    }
    // stage arguments of both state-buffer parities (stage 0 reads y_n of that parity):
    // rows 1 .. n_main-1 into S_k, then y_(n+1) into the other buffer, as enqueue_attempt
    FusedArgs A;
    memset(&A, 0, sizeof(A));
    std::vector<StageArgs> rows(2 * S.n_main);
    const int cur0 = c->cur;
    for (int p = 0; p < 2; p++) {
        c->cur = p;
        for (int r = 1; r < S.n_main; r++)
            rows[p * S.n_main + r - 1] = make_args(c, S.rows[r], c->d_r[p], c->d_v[p], c->d_sr[r], c->d_sv[r], c->n);
        rows[p * S.n_main + S.n_main - 1] = make_args(c, S.np1, c->d_r[p], c->d_v[p], c->d_r[1 - p], c->d_v[1 - p], c->n);
    }
    c->cur = cur0;
    for (int r = 1; r < S.n_main; r++) A.form[r - 1] = S.rows[r].form;
    A.form[S.n_main - 1] = S.np1.form;
    CTX_HIP(c, hipMemcpyAsync(c->d_fused_rows, rows.data(), sizeof(StageArgs) * rows.size(), hipMemcpyHostToDevice,
                              c->stream));
    for (int k = 0; k < 2; k++) { A.r[k] = c->d_r[k]; A.v[k] = c->d_v[k]; }
    for (int k = 0; k < S.n_main; k++) { A.a[k] = c->d_a[k]; A.c_stage[k] = S.c[k]; }
    A.rows = c->d_fused_rows;
    A.n_main = S.n_main;
    A.clear_events_after_f0 = S.clear_events_each_attempt ? 1 : 0;
    sweeps_of(c, A.sw);
    A.n = c->n;
    A.mass = c->d_mass; A.radius = c->d_radius; A.id = c->d_id; A.type = c->d_type;
    A.ej = c->prm.ejection; A.hc = c->prm.hit_centrum;
    A.ev_buf = c->d_events; A.ev_cap = c->ev_cap;
    A.ctl = c->d_ctl;
    A.dt = c->dt_try;
    A.t_stop = t_stop;
    A.interval = output_interval;
    A.out = c->d_fused_out;
    c->raw_valid = false;
    // bounded launches: each runs at most FUSED_STEPS_PER_LAUNCH steps
    while (true) {
        A.t = c->t;
        A.ps = *ps;
        A.cur = c->cur;
        A.max_steps = std::min(FUSED_STEPS_PER_LAUNCH, max_steps - *steps_done);
        CTX_HIP(c, launch_fused_fixed(A, c->ptest, c->stream));
        FusedOut o;
        CTX_HIP(c, stream_copy(c->stream, &o, c->d_fused_out, sizeof(o), hipMemcpyDeviceToHost));
        // integrator.cu:258-263 and pp_disk::swap for every step done
        c->t = o.t;
        c->cur = o.cur;
        c->dt_did = c->dt_try;
        c->n_tried += o.steps;
        c->n_passed += o.steps;
        *ps = o.ps;
        *steps_done += o.steps;
        *stop_reason = o.reason;
        if (o.reason != RED_FUSED_STOP_STEPS || *steps_done >= max_steps) break;
    }
    return RED_OK;
}

int red_clear_events(red_ctx* c)
{
    if (!c) return RED_ERR_ARG;
    hipLaunchKernelGGL(k_set_ctl, dim3(1), dim3(1), 0, c->stream, c->d_ctl, c->t, c->dt_try, 0, 0, 1, 1);
    CTX_HIP(c, hipGetLastError());
    c->raw.clear();
    c->raw_valid = true;
    return RED_OK;
}

int red_raw_events(red_ctx* c, red_event* buf, int cap, int* n, int* overflow)
{
    if (!c) return RED_ERR_ARG;
    int rc = read_ctl(c);
    if (rc) return rc;
    const int nd = c->h_ctl->n_events;
    if (overflow) *overflow = nd > c->ev_cap;
    if (nd > c->ev_cap) {
        if (n) *n = nd;
        return RED_OK;
    }
    rc = fetch_raw_events(c);
    if (rc) return rc;
    if (n) *n = (int)c->raw.size();
    for (int k = 0; k < (int)c->raw.size() && k < cap && buf; k++) buf[k] = c->raw[k];
    return RED_OK;
}

int red_sp_events(red_ctx* c, red_event* buf, int cap, int* n)
{
    if (!c) return RED_ERR_ARG;
    if (n) *n = (int)c->sp.size();
    for (int k = 0; k < (int)c->sp.size() && k < cap && buf; k++) buf[k] = c->sp[k];
    return RED_OK;
}

// pp_disk.cu:683-715, 934-996, 1091-1109
int red_check_ejection_hit_centrum(red_ctx* c, int* handled)
{
    if (!c) return RED_ERR_ARG;
    if (handled) *handled = 0;
    const double ej = c->prm.ejection, hc = c->prm.hit_centrum;
    if (!(ej > 0.0) && !(hc > 0.0)) return RED_OK;
    hipLaunchKernelGGL(k_set_ctl, dim3(1), dim3(1), 0, c->stream, c->d_ctl, c->t, c->dt_try, 0, 0, 0, 1);
    CTX_HIP(c, hipGetLastError());
    CTX_HIP(c, timed(c, RED_PROF_EJECT, [&] {
        return launch_ejection_scan(c->d_r[c->cur], c->d_v[c->cur], c->d_id, c->d_type, c->n, ej, hc, c->d_ctl,
                                    c->d_events_hc, std::max(1, c->n), c->stream);
    }));
    int rc = read_ctl(c);
    if (rc) return rc;
    const int ne = c->h_ctl->n_events_hc;
    if (ne <= 0) return RED_OK;
    std::vector<red_event> ev(ne);
    CTX_HIP(c, stream_copy(c->stream, ev.data(), c->d_events_hc, sizeof(red_event) * ne, hipMemcpyDeviceToHost));
    std::vector<BodyUpdate> ups;
    events::BodyView B = c->view();
    std::string msg;
    rc = events::ejection_pass(B, ev, c->evc, ups, msg);
    c->sp = ev;
    if (rc) return c->fail(rc, msg);
    // ids negated by the kernel are already on the device; updates carry the merge results
    rc = apply_updates(c, ups);
    if (rc) return rc;
    if (handled) *handled = 1;
    return RED_OK;
}

// pp_disk.cu:717-738, 1078-1149 (close-encounter records of the Hill test are reported, not merged)
int red_check_collision(red_ctx* c, int* handled)
{
    if (!c) return RED_ERR_ARG;
    if (handled) *handled = 0;
    int rc = fetch_raw_events(c);
    if (rc) return rc;
    if (c->raw.empty()) return RED_OK;
    std::vector<BodyUpdate> ups;
    events::BodyView B = c->view();
    std::string msg;
    rc = events::collision_pass(B, c->raw, c->sp, c->evc, ups, msg);
    if (rc) return c->fail(rc, msg);
    rc = apply_updates(c, ups);
    if (rc) return rc;
    if (handled) *handled = 1;
    return RED_OK;
}

// pp_disk.cu:740-770 (inner_steps == true)
int red_check_close_encounter(red_ctx* c, int* handled)
{
    if (!c) return RED_ERR_ARG;
    if (handled) *handled = 0;
    int rc = fetch_raw_events(c);
    if (rc) return rc;
    if (c->raw.empty()) return RED_OK;
    events::close_encounter_pass(c->view(), c->raw, c->sp, c->evc);
    if (handled) *handled = 1;
    return RED_OK;
}

// pp_disk.cu:893-907, 1324-1409; number_of_bodies.cpp:27-104
int red_check_rebuild(red_ctx* c, int threshold, int* rebuilt)
{
    if (!c) return RED_ERR_ARG;
    if (rebuilt) *rebuilt = 0;
    if (!(c->evc.event[events::CNT_LAST_CLEAR] >= threshold)) return RED_OK;
    std::vector<d4> R, V;
    int rc = download_state(c, R, V);
    if (rc) return rc;
    int removed[7] = {0};
    c->rebuild_map.clear();
    for (int i = 0; i < c->n; i++) {
        if (0 < c->h_id[i] && BT_PAD > c->h_type[i]) { c->rebuild_map.push_back(i); continue; }
        const int t = c->h_type[i];
        if (t < 0 || t > BT_PAD) return c->fail(RED_ERR_ARG, "Undefined body type!");
        if (t == BT_PAD) continue;
        removed[t]++;
        c->n_inactive[t]++;
    }
    for (int k = 0; k < 7; k++) c->cnt[k] -= removed[k];
    const int m = (int)c->rebuild_map.size();
    int nn = 0;
    for (int k = 0; k < 7; k++) nn += c->cnt[k];
    if (nn != m) return c->fail(RED_ERR_ARG, "Error: number of copied bodies does not equal to the number of active bodies.");
    std::vector<d4> R2(m), V2(m);
    for (int k = 0; k < m; k++) {
        const int i = c->rebuild_map[k];
        R2[k] = R[i]; V2[k] = V[i];
        c->h_mass[k] = c->h_mass[i]; c->h_radius[k] = c->h_radius[i]; c->h_density[k] = c->h_density[i];
        c->h_cd[k] = c->h_cd[i]; c->h_id[k] = c->h_id[i]; c->h_type[k] = c->h_type[i];
        c->h_epoch[k] = c->h_epoch[i]; c->h_mig_type[k] = c->h_mig_type[i]; c->h_mig_stop[k] = c->h_mig_stop[i];
    }
    c->n = m;
    for (auto* vec : {&c->h_mass, &c->h_radius, &c->h_density, &c->h_cd, &c->h_epoch, &c->h_mig_stop}) vec->resize(m);
    for (auto* vec : {&c->h_id, &c->h_type, &c->h_mig_type}) vec->resize(m);
    if (m > 0) {
        CTX_HIP(c, stream_copy(c->stream, c->d_r[c->cur], R2.data(), sizeof(d4) * m, hipMemcpyHostToDevice));
        CTX_HIP(c, stream_copy(c->stream, c->d_v[c->cur], V2.data(), sizeof(d4) * m, hipMemcpyHostToDevice));
    }
    rc = upload_all(c);
    if (rc) return rc;
    destroy_graphs(c);
    c->gcfg.n_fp64_src = c->cnt[0] + c->cnt[1];
    refresh_pair_test(c);
    c->evc.set(events::CNT_LAST_CLEAR, 0);
    if (rebuilt) *rebuilt = 1;
    return RED_OK;
}

int red_last_rebuild_map(red_ctx* c, int* old_index, int cap, int* n)
{
    if (!c) return RED_ERR_ARG;
    if (n) *n = (int)c->rebuild_map.size();
    for (int k = 0; k < (int)c->rebuild_map.size() && k < cap && old_index; k++) old_index[k] = c->rebuild_map[k];
    return RED_OK;
}

int red_counters(red_ctx* c, int out[8])
{
    if (!c || !out) return RED_ERR_ARG;
    out[0] = c->n_tried; out[1] = c->n_failed; out[2] = c->n_passed;
    const events::Counters& E = c->evc;
    out[3] = E.collision[events::CNT_TOTAL]; out[4] = E.ejection[events::CNT_TOTAL];
    out[5] = E.hit_centrum[events::CNT_TOTAL]; out[6] = E.close_encounter[events::CNT_TOTAL];
    out[7] = E.event[events::CNT_LAST_CLEAR];
    return RED_OK;
}

int red_energy(red_ctx* c, double* E, double L[3])
{
    if (!c) return RED_ERR_ARG;
    const int nm = c->nMassive(), nsi = c->nSI();
    const int nch = energy_chunks(nm);
    const size_t np = (size_t)std::max(nsi, 0) * nch;
    if (np > c->energy_part_cap) {
        SetupLock lk;
        (void)hipFree(c->d_energy_part);
        c->d_energy_part = nullptr;
        c->energy_part_cap = 0;
        CTX_HIP(c, dmalloc(&c->d_energy_part, np));
        c->energy_part_cap = np;
    }
    std::vector<double> o(6 * (size_t)std::max(nm, 1)), u(std::max<size_t>(np, 1));
    CTX_HIP(c, launch_energy(c->d_r[c->cur], c->d_v[c->cur], c->d_mass, c->d_id, nm, nsi, c->d_energy,
                             c->d_energy_part, c->stream));
    if (nm > 0) CTX_HIP(c, stream_copy(c->stream, o.data(), c->d_energy, sizeof(double) * 6 * nm, hipMemcpyDeviceToHost));
    if (np > 0) CTX_HIP(c, stream_copy(c->stream, u.data(), c->d_energy_part, sizeof(double) * np, hipMemcpyDeviceToHost));
    double T = 0.0, U = 0.0, lx = 0.0, ly = 0.0, lz = 0.0;
    for (int i = 0; i < nm; i++) {
        T += o[6 * i];
        lx += o[6 * i + 2]; ly += o[6 * i + 3]; lz += o[6 * i + 4];
    }
    for (size_t k = 0; k < np; k++) U += u[k];
    if (E) *E = T - U;
    if (L) { L[0] = lx; L[1] = ly; L[2] = lz; }
    return RED_OK;
}

int red_energy_noalloc(red_ctx* c, double* out_dev)
{
    if (!c || !out_dev) return RED_ERR_ARG;
    const int nm = c->nMassive(), nsi = c->nSI();
    const size_t need = (size_t)std::max(nsi, 0) * energy_chunks(nm);
    if (need > ENERGY_SCRATCH_DOUBLES)
        return c->fail(RED_ERR_ARG, "red_energy_noalloc: the disk needs " + std::to_string(need) +
                                        " doubles of scratch, more than the static buffer's " +
                                        std::to_string(ENERGY_SCRATCH_DOUBLES) + " (use red_energy)");
    double* scr = nullptr;
    CTX_HIP(c, energy_scratch_ptr(c->prm.device, &scr));
    CTX_HIP(c, launch_energy_noalloc(c->d_r[c->cur], c->d_v[c->cur], c->d_mass, c->d_id, nm, nsi, scr, out_dev,
                                     c->stream));
    return RED_OK;
}

int red_set_profiling(red_ctx* c, int on)
{
    if (!c) return RED_ERR_ARG;
    if (on && c->prm.use_graph) return c->fail(RED_ERR_ARG, "profiling needs use_graph = 0");
    c->profiling = on != 0;
    return RED_OK;
}

int red_get_profile(red_ctx* c, double ms[RED_PROF_CLASSES], int launches[RED_PROF_CLASSES], int reset)
{
    if (!c) return RED_ERR_ARG;
    CTX_HIP(c, hipStreamSynchronize(c->stream));
    for (auto& p : c->prof) {
        float t = 0.0f;
        CTX_HIP(c, hipEventElapsedTime(&t, p.a, p.b));
        c->prof_ms[p.cls] += t;
        c->prof_n[p.cls] += 1;
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    c->prof.clear();
    for (int k = 0; k < RED_PROF_CLASSES; k++) {
        if (ms) ms[k] = c->prof_ms[k];
        if (launches) launches[k] = c->prof_n[k];
        if (reset) { c->prof_ms[k] = 0.0; c->prof_n[k] = 0; }
    }
    return RED_OK;
}

int red_device_alloc(red_ctx* c, unsigned long long bytes, void** dev)
{
    if (!c || !dev) return RED_ERR_ARG;
    *dev = nullptr;
    SetupLock lk;
    CTX_HIP(c, hipMalloc(dev, bytes ? bytes : 1));
    c->user_buffers.push_back(*dev);
    return RED_OK;
}

int red_device_free(red_ctx* c, void* dev)
{
    if (!c) return RED_ERR_ARG;
    auto it = std::find(c->user_buffers.begin(), c->user_buffers.end(), dev);
    if (it == c->user_buffers.end()) return c->fail(RED_ERR_ARG, "pointer was not allocated by red_device_alloc");
    c->user_buffers.erase(it);
    CTX_HIP(c, hipStreamSynchronize(c->stream));
    SetupLock lk;
    CTX_HIP(c, hipFree(dev));
    return RED_OK;
}

int red_copy_to_device(red_ctx* c, void* dev, const void* host, unsigned long long bytes)
{
    if (!c || (bytes && (!dev || !host))) return RED_ERR_ARG;
    CTX_HIP(c, hipStreamSynchronize(c->stream));
    CTX_HIP(c, stream_copy(c->stream, dev, host, bytes, hipMemcpyHostToDevice));
    return RED_OK;
}

int red_copy_to_host(red_ctx* c, void* host, const void* dev, unsigned long long bytes)
{
    if (!c || (bytes && (!dev || !host))) return RED_ERR_ARG;
    CTX_HIP(c, hipStreamSynchronize(c->stream));
    CTX_HIP(c, stream_copy(c->stream, host, dev, bytes, hipMemcpyDeviceToHost));
    return RED_OK;
}

int red_host_transform_to_ac(int n, const int* type, double* r, double* v)
{
    if (n < 0 || (n > 0 && (!type || !r || !v))) return RED_ERR_ARG;
    if (n == 0) return RED_OK;
    if (type[0] != 0) return RED_ERR_NO_STAR;   // the star leads the body list
    const double R0[3] = {r[0], r[1], r[2]}, V0[3] = {v[0], v[1], v[2]};
    for (int j = n - 1; j >= 0; j--)
        for (int k = 0; k < 3; k++) { r[3 * j + k] -= R0[k]; v[3 * j + k] -= V0[k]; }
    return RED_OK;
}

int red_host_transform_to_bc(int n, const double* mass, double* r, double* v)
{
    if (n < 0 || (n > 0 && (!mass || !r || !v))) return RED_ERR_ARG;
    double R0[3] = {0, 0, 0}, V0[3] = {0, 0, 0};
    for (int j = n - 1; j >= 0; j--) {
        const double m = mass[j];
        for (int k = 0; k < 3; k++) { R0[k] += m * r[3 * j + k]; V0[k] += m * v[3 * j + k]; }
    }
    double M0 = 0.0;
    for (int j = n - 1; j >= 0; j--) M0 += mass[j];
    for (int k = 0; k < 3; k++) { R0[k] /= M0; V0[k] /= M0; }
    for (int j = n - 1; j >= 0; j--)
        for (int k = 0; k < 3; k++) { r[3 * j + k] -= R0[k]; v[3 * j + k] -= V0[k]; }
    return RED_OK;
}

int red_host_transform_velocity(int n, double* v)
{
    if (n < 0 || (n > 0 && !v)) return RED_ERR_ARG;
    for (int j = 0; j < 3 * n; j++) v[j] /= GAUSS;
    return RED_OK;
}

}  // extern "C"
