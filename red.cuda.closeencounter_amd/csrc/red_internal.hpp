// red_internal.hpp -- shared device/host definitions of the MI355X integration path.
//
// Data layout in HBM (one context):
//   state  Y[2]   : r[cap] then v[cap], each a 32-byte record {x, y, z, w}; the w lane of
//                   r carries the body mass (so one 32-byte scalar load hands a source body
//                   to a whole wavefront), the w lane of v is 0.  This is the reference's
//                   vec_t AoS (red_type.h:142-148) with the unused lane put to work.
//   stages S_k    : the same pair of records for every intermediate stage state.
//   accels A_k    : one record {ax, ay, az, 0} per body per stage.
//   mass, radius  : FP64 per body (canonical copies; r.w mirrors mass), id, type: int32.
//   Ctl           : per-context control block (step-start time, trial step, attempt,
//                   error maximum, NaN flag, event counter), read by every kernel so that a
//                   captured HIP graph replays with the current step values.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "../../include/rednative.h"
#include "host_tools.hpp"

namespace red {

struct alignas(32) d4 {
    double x, y, z, w;
};

// Device control block.  The host writes the first three fields before an attempt;
// kernels write the rest.
struct Ctl {
    double t;                      // step-start time (internal units)
    double dt;                     // trial step of the current attempt
    int attempt;                   // attempt index within the step
    int pad0;
    unsigned long long err_bits;   // max(0, max error component), IEEE bits (monotone for >= 0)
    int nan_flag;                  // a non-finite error component was seen
    int n_events;                  // pair events appended since the last clear
    int n_events_hc;               // ejection / hit-centrum events of the last scan
    int pad1;
};

// One force sweep: sinks [sink0, sink1) against sources [src0, src1).
struct Sweep {
    int sink0, sink1, src0, src1;
    // activity of sinks (> 0) and sources (>= 0): the body ids, or under padded storage an
    // array of ones (pp_disk.cu:144-215 tests no id: inactive bodies keep pulling and being
    // pulled until the next compaction); event records always carry the ids
    const int* gate;
};

// Pair-test parameters shared by every force kernel.
struct PairTest {
    double radii_f;                // collision when d < radii_f * (R_i + R_j)
    double hill_f;                 // close encounter when d < hill_f * R_Hill
    double three_mstar;            // 3 M_star, for R_Hill
    double rmax;                   // largest body radius (mixed-precision screen)
    double mmax;                   // largest source mass (mixed-precision Hill screen)
    double rmax_all;               // largest radius of any body (FP64 screen)
    double mmax_ns;                // largest mass of any non-star body (FP64 screen)
    int n_star;                    // bodies [0, n_star) are stars: no Hill test with them
    int mass_core;                 // every active massive body's mass (the w lane of the
                                   // swept positions) lies in [2^-300, 2^300]
    int close_encounter;           // 1: report CLOSE_ENCOUNTER, keep (i, j) order
    int enabled;                   // any pair test active
};

struct EventSink {
    red_event* buf;
    int cap;
    int stage;
    double c_stage;                // event time = ctl->t + c_stage * ctl->dt
};

// Packed source record for the mixed-precision sweep: FP64 position, FP32 mass.
struct alignas(32) SrcMixed {
    double x, y, z;
    float m;
    int active;
};

}  // namespace red

// ---- launchers implemented in the .hip files (host side) ----------------------------
namespace red {

// Per source chunk of the one-work-item-per-sink FP64 sweep (TPS_CHUNK packed sources):
// the largest |coordinate| of the chunk's packed sources, the largest radius of its active
// sources (the pair screen's collision reach) and whether every packed mass is 0 or inside
// the square-root / quotient cores' range [2^-300, 2^300].
constexpr int TPS_CHUNK = 256;
struct TpsChunk {
    double bound;
    double rmax;
    int mass_ok;
    int pad;
};

struct GravBuffers {
    const d4* r;                   // stage positions (w = mass)
    const d4* v;                   // stage velocities (only read for event records)
    d4* a;                         // output accelerations
    const double* mass;
    const double* radius;
    const int* id;
    Ctl* ctl;
    void* scratch;                 // split-sweep scratch
    size_t scratch_bytes;
    SrcMixed* src_mixed;           // packed sources for the mixed sweep (cap entries)
    d4* src_fp64 = nullptr;        // packed sources for the FP64 one-per-sink sweep (cap
    TpsChunk* tps_chunk = nullptr; // entries, inactive ones parked) and their chunks
    d4* partial;                   // per-part sums of source-partitioned sweeps
    size_t partial_elems;
    double* seg;                   // segment sums of split sweeps (own buffer: a split
    size_t seg_elems;              // sweep may run beside a partitioned one)
};

struct GravConfig {
    int precision;                 // RED_PREC_*
    int n_fp64_src;                // mixed: sources [0, n_fp64_src) stay FP64
    int block;                     // threads per block
    int spt;                       // sinks per thread
    int strategy;                  // 1 scalar broadcast, 2 LDS tiles
    int kahan_pairs;
    int split_max_sinks;           // sweeps with fewer sinks use the split path
    int fp64_split;                // FP64 sums may be partitioned over sources
    int source_parts;              // forced part count, 0 = automatic
    int block_seq;                 // producer/consumer block size (256, 512, 1024), 0 = automatic
    int seq_sinks;                 // producer/consumer sinks per block (4, 8, 16), 0 = automatic
    int astro;                     // astro-centric frame (FP64, self-interacting bodies only)
    double astro_mstar;            // the star's mass for the astro-centric start term
};

// Number of source parts a sweep is cut into.
int sweep_parts(const Sweep& sw, const GravConfig& C, size_t partial_elems);

// Launches one force sweep (all three classes are separate sweeps).
hipError_t launch_sweep(const Sweep& sw, const GravBuffers& B, const GravConfig& C, const PairTest& P,
                        const EventSink& E, hipStream_t s);
// Scratch bytes the split path needs for a sweep.
size_t split_scratch_bytes(const Sweep& sw);
// whether launch_sweep forms the sweep's terms in the split scratch (given room for them)
bool sweep_wants_split(const Sweep& sw, const GravConfig& C);
constexpr int SPLIT_SEG_MAX = 64;   // segments per row of a free-order split sum
// Packs SrcMixed records for bodies [0, n).
hipError_t launch_pack_mixed(const d4* r, const int* id, SrcMixed* out, int n, hipStream_t s);
// whether launch_sweep sweeps it with one work-item per sink (then the FP64 sources [0, n)
// must have been packed by launch_pack_fp64 for this evaluation)
bool sweep_takes_tps(const Sweep& sw, const GravConfig& C, const GravBuffers& B);
hipError_t launch_pack_fp64(const d4* r, const int* gate, const double* radius, d4* out, TpsChunk* chunks, int n,
                            hipStream_t s);

// Stage-sum forms (see kern_stages.hip).
enum StageForm { SF_AXPY = 0, SF_DTSUM = 1, SF_ACCUM = 2, SF_RK4NP1 = 3, SF_RKF78NP1 = 4 };
enum ErrForm { EF_RK4 = 0, EF_RKF78 = 1, EF_CK = 2 };

struct StageTerm {
    double coef;
    const d4* fv;                  // position derivative: velocity record of stage state k
    const d4* fa;                  // velocity derivative: acceleration of stage k
};

constexpr int MAX_TERMS = 13;

struct StageArgs {
    const d4* yr;
    const d4* yv;
    d4* outr;
    d4* outv;
    int n;
    int nterms;
    double c1, c2;
    const Ctl* ctl;
    // ensembles of independent disks (null for one disk): per-body disk index, per-disk
    // trial step and flags (bit 0: the disk takes part in this attempt), per-body error-norm
    // membership, per-disk error maximum and NaN flag
    const int* disk_of;
    const double* disk_dt;
    const int* disk_flags;
    const unsigned char* err_on;
    unsigned long long* disk_err;
    int* disk_nan;
    StageTerm t[MAX_TERMS];
};

hipError_t launch_stage(int form, const StageArgs& A, hipStream_t s);
hipError_t launch_err(int form, const StageArgs& A, hipStream_t s);
hipError_t launch_reset_ctl(Ctl* ctl, int clear_events, hipStream_t s);

__device__ __forceinline__ void put3(double* d, const d4& q) { d[0] = q.x; d[1] = q.y; d[2] = q.z; }

// Mutual Hill radius helper: cube root by bit-seeded Newton iterations using only
// +, -, *, / so that host and device produce identical results.
__device__ __forceinline__ double cbrt_ieee(double x)
{
    if (!(x > 0.0)) return 0.0;
    unsigned long long b = __double_as_longlong(x);
    unsigned long long e = (b >> 52) & 0x7ff;
    long long eb = (long long)e - 1023;
    long long q = eb >= 0 ? eb / 3 : -((-eb + 2) / 3);
    double y = __longlong_as_double((unsigned long long)(q + 1023) << 52);
    for (int k = 0; k < 60; k++) {
        double yn = y - (y * y * y - x) / (3.0 * y * y);
        if (yn == y) break;
        y = yn;
    }
    return y;
}

// Ejection / hit-centrum predicate of one active non-star body (pp_disk.cu:72-142):
// |r|^2 > R_ej^2 (R_ej > 0) is an ejection, otherwise |r|^2 < R_hc^2 (R_hc > 0) a hit.
__device__ __forceinline__ int ejection_name(const d4& q, double ej, double hc)
{
    const double r2 = (q.x * q.x) + (q.y * q.y) + (q.z * q.z);
    if (0.0 < ej && ej * ej < r2) return RED_EV_EJECTION;
    if (0.0 < hc && hc * hc > r2) return RED_EV_HIT_CENTRUM;
    return RED_EV_NONE;
}

// Ejection / hit-centrum scan (pp_disk.cu:72-142): appends events and negates ids.
hipError_t launch_ejection_scan(const d4* r, const d4* v, int* id, const int* type, int n, double ej,
                                double hc, Ctl* ctl, red_event* buf, int cap, hipStream_t s);

// Scatter of merged-body updates (BodyUpdate: host_tools.hpp).
hipError_t launch_apply_updates(const BodyUpdate* u, int nu, d4* r, d4* v, double* mass, double* radius,
                                int* id, hipStream_t s);

// Energy / angular momentum partials: per massive body (1/2 m v^2, 0, m r x v, 0) into out6,
// and the potential of the interacting pairs per (SI body, source chunk) into part
// (n_si x energy_chunks(n_massive) values, to be added in that order).
int energy_chunks(int n_massive);
hipError_t launch_energy(const d4* r, const d4* v, const double* mass, const int* id, int n_massive, int n_si,
                         double* out6, double* part, hipStream_t s);
// Allocation-free variant: the potential partials in a static device array of
// ENERGY_SCRATCH_DOUBLES (1 MiB per GPU), the sums of red_energy done on the device in the
// same order into res[4] = {E, Lx, Ly, Lz}.  energy_scratch_ptr looks the array's address
// up once per device.
constexpr size_t ENERGY_SCRATCH_DOUBLES = (size_t)1 << 17;
hipError_t energy_scratch_ptr(int device, double** out);
hipError_t launch_energy_noalloc(const d4* r, const d4* v, const double* mass, const int* id, int n_massive,
                                 int n_si, double* scratch_dev, double* res, hipStream_t s);

// ---- fused fixed-step loop: main-loop iterations of one small disk in one work-group ----
constexpr int FUSED_MAX_ROWS = MAX_TERMS;     // stage rows after f0, plus y_(n+1)
constexpr int FUSED_MAX_BODIES = 4096;
// One work-group runs the whole loop, so it pays only while the evaluations are small.
// Measured on MI355X (RK4 steps; stepwise = one graph replay per step, to which the
// driver loop adds about 50 us of per-step ejection and collision checks): 256-body disk,
// 1527 interacting pairs per evaluation: 25 us fused, 55 us stepwise; 128 self-interacting
// bodies, 16384 pairs: 95 against 40 (+50); 256 of them, 65536 pairs: 181 against 47 (+50);
// 1024 of them: 2.8 ms against 64 us.  Above 16384 pairs (sum over the sweeps of sinks x
// sources) the driver loop is faster launch by launch.
constexpr long FUSED_MAX_PAIRS = 16384;
constexpr int FUSED_STEPS_PER_LAUNCH = 4096;  // the host relaunches for longer runs
struct FusedOut {
    double t, ps;                  // time and accumulated output step after the last step
    int steps, cur, reason, pad;   // steps done, state buffer parity, RED_FUSED_STOP_*
};
struct FusedArgs {
    d4* r[2];                      // positions by buffer parity (y_n in r[cur])
    d4* v[2];
    d4* a[FUSED_MAX_ROWS];         // accelerations per stage
    const StageArgs* rows;         // [2][n_main]: rows 1 .. n_main-1 into S_k, then y_(n+1)
    int form[FUSED_MAX_ROWS];      // stage form of each entry of a parity's rows
    int n_main;
    int clear_events_after_f0;     // RKF7(8) clears the event list at the attempt start
    double c_stage[FUSED_MAX_ROWS];
    Sweep sw[3];
    int n;
    const double* mass;
    const double* radius;
    const int* id;
    const int* type;
    double ej, hc;
    red_event* ev_buf;
    int ev_cap;
    Ctl* ctl;
    double t, dt, t_stop, ps, interval;
    int cur, max_steps;
    FusedOut* out;
};
hipError_t launch_fused_fixed(const FusedArgs& A, const PairTest& P, hipStream_t s);

// ---- learned step-size proposer ----
constexpr int MLP_NF = RED_MLP_FEATURES, MLP_NH = 64;
constexpr int MLP_NPARAM = MLP_NH * MLP_NF + MLP_NH + MLP_NH * MLP_NH + MLP_NH + MLP_NH + 1;
// feature slots filled on the device (the rest come from the step history on the host)
enum {
    FEAT_LOG_MIN_SEP = 7,
    FEAT_D_LOG_MIN_SEP = 8,
    FEAT_LOG_MAX_ACC = 9,
    FEAT_LOG_MAX_PERT = 10,
    FEAT_D_LOG_MAX_PERT = 11
};
struct ProposeArgs {
    float x[MLP_NF];            // history features; device slots are overwritten
    float prev_log_min_sep;
    float prev_log_max_pert;
    int have_prev;
};
// Device buffers of the fused step-start kernel.
constexpr int LEARN_PAIR_BLOCKS = 2048;   // blocks of the pair launch (slots 0 ..)
constexpr int LEARN_BODY_BLOCKS = 512;    // blocks of the body launch (slots after those)
constexpr int LEARN_MAX_BLOCKS = LEARN_PAIR_BLOCKS + LEARN_BODY_BLOCKS;
struct LearnBuffers {
    float* part;                // 3 floats per block: min (d/R_H)^3, max |a|, max |a - a_kep|/|a_kep|
    unsigned* counter;          // blocks finished; the last block resets it to 0
    float* out;                 // host-mapped: features [0..11], network output [12]
};
// Work of the step-start kernel: pair items SI x massive and NSI x SI without the star
// (index 0), body items 1..n-1 with id > 0, and the block counts of its two launches.
struct LearnPlan {
    long n0, n1, nb;
    int ch0, ch1;
    int b0, b1;
};
LearnPlan learn_plan(int n_si, int n_massive, int n);
// First launch (state only, may run beside the step's first force evaluation): the pair
// extrema of r into slots [0, b0).
hipError_t launch_learn_pairs(const LearnPlan& L, const d4* r, const int* id, int n_si, int n_massive, int n,
                              const LearnBuffers& B, hipStream_t s);
// Second launch, after the first and after the evaluation of a: the body extrema, the
// reduction of all slots, the feature vector from A, the network (if net) and, with apply,
// ctl->dt = safety * eta_n * exp(clamp(y, +-max_log)).
hipError_t launch_learn_tail(const red_mlp* net, const ProposeArgs& A, const LearnPlan& L, const d4* r, const d4* a,
                             const int* id, int n_si, int n_massive, int n, const LearnBuffers& B, Ctl* ctl,
                             int apply, double eta_n, double safety, double max_log, hipStream_t s);
int mlp_device(const red_mlp* M);


}  // namespace red

namespace red {

// Serialises, across host threads, what must not overlap a HIP graph capture running on
// another thread's stream: device and pinned allocation and release, stream set-up and the
// capture itself.  Every other call stays on its context's own stream and takes no lock.
std::recursive_mutex& setup_mutex();
struct SetupLock {
    std::lock_guard<std::recursive_mutex> g{setup_mutex()};
};

}  // namespace red

#define RED_HIP_TRY(expr)                                   \
    do {                                                    \
        hipError_t _e = (expr);                             \
        if (_e != hipSuccess) return _e;                    \
    } while (0)

namespace red {
// Blocking copy on a context's own stream (no legacy-stream synchronisation).
inline hipError_t stream_copy(hipStream_t s, void* dst, const void* src, size_t bytes, hipMemcpyKind kind)
{
    RED_HIP_TRY(hipMemcpyAsync(dst, src, bytes, kind, s));
    return hipStreamSynchronize(s);
}
}  // namespace red
