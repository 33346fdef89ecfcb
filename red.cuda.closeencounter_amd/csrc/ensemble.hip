// ensemble.hip -- many independent disks of uneven size integrated together (gfx950).
//
// All disks live in one set of buffers (bodies of disk d occupy [b0_d, b1_d)).  Every
// stage of an attempt is ONE launch over all disks: the force launch walks a list of
// tiles (sink range x source range of one disk) and forms the stage sums in its epilogue,
// the error launch walks all bodies with a per-body disk index.  Each disk keeps its own
// time, trial step and attempt count in device arrays, and a work-item per disk runs the
// reference's controller after every attempt with the single-disk host controller's
// arithmetic (step_pow.hpp), so an ensemble member follows exactly the steps it would take
// alone (int_rungekutta8.cu:704-797 and siblings).  A batch of supersteps (one attempt of
// every disk still stepping, each) is captured once in a HIP graph and replayed; the host
// sees the disks once per batch.
//
// Events: the force launches run the pair tests of every disk (collision at
// f (R_i + R_j), close encounter at f mutual Hill radii; pp_disk.cu:265-291) into one
// record list per attempt, tagged with the disk and taken apart by the host; every attempt
// also evaluates the ejection / hit-centrum predicate (pp_disk.cu:72-142) on its result.
// Each member then runs the main loop's checks in the single-disk order (main.cu:201-255)
// with the single-disk host code (host_tools.cpp, namespace events): the collision or
// close-encounter pass after an accepted step, the compaction after 8 events (within the
// disk's slot of the buffers), and at the start of the next step the ejection scan, done
// exactly on the host whenever the predicate found a body or the state was changed since.
// The device holds a disk for the host where one of these changes its state (a merge, a
// scan, a compaction that moves bodies); the other steps' passes run after the batch.
//
// Force tiles (FP64, the reference's operation order and ascending-j summation):
//   kind 0: up to `block` sinks, one work-item each; sources staged through LDS in
//           chunks of `jtile` bodies;
//   kind 1: one sink (the star or a giant against its disk): the block computes `jtile`
//           pair terms in parallel into LDS and three lanes add them in source order.
// The tile shape (block, jtile) is picked by red_ensemble_autotune from measured launches.
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "host_tools.hpp"
#include "red_internal.hpp"
#include "scheme.hpp"
#include "stage_forms.hpp"
#include "step_pow.hpp"

using namespace red;

namespace {

struct EnsTile {
    int sink0, sink1, src0, src1;
    int disk, kind;
};

// Pair-test constants of one disk (the single-disk PairTest's per-disk part).
struct EnsDiskPair {
    double three_mstar;            // 3 M_star, for the mutual Hill radius
    double rmax;                   // largest body radius (screen)
    double mmax;                   // largest non-star mass (screen)
    int n_star;                    // bodies [b0, b0 + n_star) are stars: no Hill test with them
    int b0, b1;                    // the disk's bodies
};

// ---- the attempt's pair records, deduplicated on the device ---------------------------------
// The host keeps, of each step's records, one per (survivor, merger) pair: the one of
// smallest t, the first of equal times in canonical order (stage, sink, source), placed where
// the pair first occurs in that order (create_sp_events, pp_disk.cu:1111-1149).  Within one
// attempt a pair's records differ only in their stage, so three passes over the attempt's
// raw records through a pair table (keyed by superstep, survivor and merger) find each pair's
// smallest time, the smallest stage at that time (the kept record) and its smallest stage
// (where it first occurs); the kept records go to the batch's list with that first stage in
// the attempt field (superstep | first << 16), and the host only orders them by (first stage,
// sink, source).  In close-encounter mode every record is kept (no dedupe there).
struct EnsDedup {
    unsigned long long* key;    // pair table, 2^bits slots; ~0: empty
    unsigned long long* tmin;   // smallest time of the pair's records (order-preserving bits)
    unsigned int* first;        // smallest stage of the pair's records
    unsigned int* kept;         // smallest stage among its records at tmin
    int bits;
    const red_event* raw;       // the attempt's raw records
    int raw_cap;
    red_event* out;             // the batch's kept records
    int out_cap;
    const EnsDiskPair* dp;      // each disk's first body
};

constexpr unsigned long long EMPTY_KEY = ~0ull;

__device__ __forceinline__ unsigned long long time_key(double t)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(t);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

// superstep (6 bits) | survivor (29 bits) | merger (29 bits), global body indices; never
// ~0, as survivor and merger differ
__device__ __forceinline__ unsigned long long pair_key(const EnsDedup& D, const red_event& e, int s)
{
    const unsigned long long b0 = (unsigned long long)D.dp[e.pad].b0;
    return ((unsigned long long)s << 58) | ((b0 + (unsigned long long)e.idx1) << 29) | (b0 + (unsigned long long)e.idx2);
}

__device__ __forceinline__ unsigned int key_hash(unsigned long long k, int bits)
{
    return (unsigned int)((k * 0x9E3779B97F4A7C15ull) >> (64 - bits));
}

// The pair's slot, claimed if new; -1 when the table is full (every probe sequence ends).
__device__ int pair_slot(const EnsDedup& D, unsigned long long k)
{
    const unsigned int mask = (1u << D.bits) - 1u;
    const unsigned int h = key_hash(k, D.bits);
    for (unsigned int p = 0; p <= mask; p++) {
        const unsigned int i = (h + p) & mask;
        const unsigned long long old = atomicCAS(&D.key[i], EMPTY_KEY, k);
        if (old == EMPTY_KEY || old == k) return (int)i;
    }
    return -1;
}

// The pair's slot after k_dedup_insert (a read-only probe); -1 if it is not there.
__device__ int pair_find(const EnsDedup& D, unsigned long long k)
{
    const unsigned int mask = (1u << D.bits) - 1u;
    const unsigned int h = key_hash(k, D.bits);
    for (unsigned int p = 0; p <= mask; p++) {
        const unsigned int i = (h + p) & mask;
        const unsigned long long q = D.key[i];
        if (q == k) return (int)i;
        if (q == EMPTY_KEY) return -1;
    }
    return -1;
}

// What a force launch needs for the pair tests of every disk.
struct EnsPair {
    double radii_f, hill_f;
    int close_encounter, enabled;
    const EnsDiskPair* dp;
    const double* disk_t;          // step-start time per disk
    const double* disk_dt;         // trial step per disk
    double c_stage;
    int stage;
    int superstep;                 // the record's attempt field until the host stamps it
    red_event* buf;                // the attempt's raw records, all disks
    int cap;
    int* count;
    int* rec;                      // records of this attempt per disk, and of them collisions
    int* rec_col;
    const int* gate;               // activity: the ids, or ones under padded storage
};

constexpr int MAX_JTILE = 256;

__device__ __forceinline__ double fp64_term(const d4& q, double xi, double yi, double zi, bool use, double& tx,
                                            double& ty, double& tz)
{
    const double dx = q.x - xi;
    const double dy = q.y - yi;
    const double dz = q.z - zi;
    const double r2 = (dx * dx + dy * dy) + dz * dz;
    const double d = sqrt(r2);
    const double sc = use ? q.w / (d * r2) : 0.0;
    tx = sc * dx;
    ty = sc * dy;
    tz = sc * dz;
    return d;
}

// Distance below which ens_pair_test can report an event for sink i (as the single-disk
// pair_screen): the collision reach with the disk's largest radius and the Hill reach with
// its largest non-star mass, using |r_j| <= |r_i| + d; 0.1 % margin for rounding.
__device__ __forceinline__ double ens_screen(const EnsPair& P, const EnsDiskPair& D, double Ri, double mi,
                                             const d4& qi)
{
    double cut = P.radii_f * (Ri + D.rmax);
    if (P.hill_f > 0.0) {
        const double fc = P.hill_f * cbrt((mi + D.mmax) / D.three_mstar);
        const double hr = sqrt((qi.x * qi.x) + (qi.y * qi.y) + (qi.z * qi.z));
        cut = fc < 1.0 ? fmax(cut, fc * hr / (1.0 - 0.5 * fc)) : 1e300;
    }
    return cut * 1.001;
}

// The single-disk pair test (kern_gravity.hip pair_test / append_pair_event) for sink i and
// source j of one disk, with the disk's own indices in the record (the host sets the
// attempt; pad carries the disk).
__device__ void ens_pair_test(const EnsPair& P, int disk, int i, int j, double d, double Ri, const d4* r,
                              const d4* v, const double* mass, const double* radius, const int* id)
{
    const EnsDiskPair D = P.dp[disk];
    const int il = i - D.b0, jl = j - D.b0;
    if (!(0 < il && il > jl)) return;
    const bool collide = d < P.radii_f * (Ri + radius[j]);
    bool hill = false;
    if (!collide && P.hill_f > 0.0 && jl >= D.n_star) {
        const d4 qi = r[i], qj = r[j];
        double ri = sqrt((qi.x * qi.x) + (qi.y * qi.y) + (qi.z * qi.z));
        double rj = sqrt((qj.x * qj.x) + (qj.y * qj.y) + (qj.z * qj.z));
        double rh = cbrt_ieee((mass[i] + mass[j]) / D.three_mstar) * ((ri + rj) * 0.5);
        hill = d < P.hill_f * rh;
    }
    if (!(collide || hill)) return;
    int surv = i, merg = j;
    int name;
    if (collide && !P.close_encounter) {
        if (mass[merg] > mass[surv]) { int t = surv; surv = merg; merg = t; }
        name = RED_EV_COLLISION;
    } else {
        name = RED_EV_CLOSE_ENCOUNTER;
    }
    // one set of atomics per wavefront for the lanes that record here together (the lanes of a
    // wavefront share the disk: a tile holds one disk's sinks)
    const unsigned long long act = __ballot(1);
    const unsigned long long colm = __ballot(name == RED_EV_COLLISION);
    const int lane = __lane_id();
    const int leader = __ffsll((long long)act) - 1;
    const int rank = __popcll(act & ((1ull << lane) - 1ull));
    int base = 0;
    if (lane == leader) {
        const int cnt = __popcll(act);
        atomicAdd(&P.rec[disk], cnt);
        const int nc = __popcll(colm);
        if (nc) atomicAdd(&P.rec_col[disk], nc);
        base = atomicAdd(P.count, cnt);
    }
    const int k = __shfl(base, leader) + rank;
    if (k >= P.cap) return;   // counted but dropped: the host reports an overflow
    red_event* e = P.buf + k;
    e->name = name;
    e->id1 = id[surv]; e->idx1 = surv - D.b0;
    e->id2 = id[merg]; e->idx2 = merg - D.b0;
    e->attempt = P.superstep; e->stage = P.stage;
    e->sink = il; e->source = jl; e->pad = disk;
    e->t = P.disk_t[disk] + P.c_stage * P.disk_dt[disk];
    e->d = d;
    for (int c = 0; c < 4; c++) { e->p1[c] = 0.0; e->p2[c] = 0.0; e->ps[c] = 0.0; }
    put3(e->r1, r[surv]); put3(e->v1, v[surv]);
    put3(e->r2, r[merg]); put3(e->v2, v[merg]);
    if (name == RED_EV_CLOSE_ENCOUNTER) {
        put3(e->rs, r[D.b0]); put3(e->vs, v[D.b0]);
    } else {
        for (int c = 0; c < 3; c++) { e->rs[c] = 0.0; e->vs[c] = 0.0; }
    }
    for (int c = 0; c < 6; c++) { e->oe1[c] = 0.0; e->oe2[c] = 0.0; e->oes[c] = 0.0; }
}

// The stage combinations that follow a force evaluation, formed per sink in its epilogue:
// each needs only its own body's records, so the launch that finishes a sink's acceleration
// forms its next stage state (and y_(n+1) after the last main stage), with k_stage's
// arithmetic (combine_get), instead of a stage launch over all bodies.
struct EnsEpi {
    const StageArgs* rows;   // device: [0] y_(n+1), [r] stage state r
    int next, next_form;     // stage state formed after this evaluation (0: none)
    int np1, np1_form;       // y_(n+1) formed after it (np1 != 0)
    const d4* cur_a;         // this evaluation's accelerations, whose values are in registers
    // the ejection / hit-centrum predicate on y_(n+1) (pp_disk.cu:72-142), with np1: one flag
    // per disk when an active non-star body meets it
    int probe_on;
    double ej, hc;
    const int *id, *type;
    int* probe;
};

// One row for body i; the term that is this evaluation's acceleration comes from `cur`.
template <int FORM>
__device__ __forceinline__ d4 epi_row_f(const StageArgs& S, int i, double dt, const d4* cur_a, d4 cur)
{
    const d4 yr = S.yr[i];
    const d4 yv = S.yv[i];
    d4 orr = combine_get<FORM>(S, yr, dt, [&](int k) { return S.t[k].fv[i]; });
    d4 ovv = combine_get<FORM>(S, yv, dt, [&](int k) { return S.t[k].fa == cur_a ? cur : S.t[k].fa[i]; });
    orr.w = yr.w;   // mass rides along unchanged
    ovv.w = 0.0;
    S.outr[i] = orr;
    S.outv[i] = ovv;
    return orr;
}

__device__ __forceinline__ d4 epi_row(int form, const StageArgs& S, int i, double dt, const d4* cur_a, d4 cur)
{
    switch (form) {
    case SF_AXPY: return epi_row_f<SF_AXPY>(S, i, dt, cur_a, cur);
    case SF_DTSUM: return epi_row_f<SF_DTSUM>(S, i, dt, cur_a, cur);
    case SF_ACCUM: return epi_row_f<SF_ACCUM>(S, i, dt, cur_a, cur);
    case SF_RK4NP1: return epi_row_f<SF_RK4NP1>(S, i, dt, cur_a, cur);
    default: return epi_row_f<SF_RKF78NP1>(S, i, dt, cur_a, cur);
    }
}

__device__ __forceinline__ void ens_epilogue(const EnsEpi& X, int i, int disk, double dt, d4 cur)
{
    if (X.next > 0) (void)epi_row(X.next_form, X.rows[X.next], i, dt, X.cur_a, cur);
    if (X.np1) {
        const d4 o = epi_row(X.np1_form, X.rows[0], i, dt, X.cur_a, cur);
        if (X.probe_on && X.id[i] > 0 && X.type[i] != 0 && ejection_name(o, X.ej, X.hc) != RED_EV_NONE)
            atomicOr(&X.probe[disk], 1);
    }
}

// Disks in the attempt (bit 0); stage 0's forces only for disks starting a new step (bit 1):
// a retried disk's tiles form only the epilogue from the stored f0.  Inactive sources add +0,
// inactive sinks get a = 0, as in the single-disk sweeps.
template <bool EV>
__global__ __launch_bounds__(256) void k_ens_force(const EnsTile* __restrict__ tiles, const d4* __restrict__ r,
                                                   const d4* __restrict__ v, d4* __restrict__ a,
                                                   const double* __restrict__ mass, const double* __restrict__ radius,
                                                   const int* __restrict__ id, const int* __restrict__ flags,
                                                   int need_bit, int jtile, EnsPair P, EnsEpi X)
{
    __shared__ d4 src[MAX_JTILE];
    __shared__ int sact[MAX_JTILE];
    __shared__ double term[2][3][MAX_JTILE];
    const EnsTile T = tiles[blockIdx.x];
    const int fl = flags[T.disk];
    if (!(fl & 1)) return;
    const bool epi = X.next > 0 || X.np1 != 0;
    // each work-item carries at most one sink to the epilogue at the end (a tile holds at
    // most blockDim.x sinks)
    int ei = -1;
    d4 ecur{0.0, 0.0, 0.0, 0.0};
    if (!(fl & need_bit)) {
        // forces not due (a retry's f0): the epilogue from the stored accelerations
        const int i = T.sink0 + threadIdx.x;
        if (i < T.sink1) { ei = i; ecur = a[i]; }
    } else if (T.kind == 0) {
        const int i = T.sink0 + threadIdx.x;
        const bool in = i < T.sink1;
        const int ic = in ? i : T.sink0;
        const d4 p = r[ic];
        const bool live = in && P.gate[ic] > 0;
        const double Ri = EV ? radius[ic] : 0.0;
        const double scr = EV && live ? ens_screen(P, P.dp[T.disk], Ri, mass[ic], p) : 0.0;
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int j0 = T.src0; j0 < T.src1; j0 += jtile) {
            const int nj = min(jtile, T.src1 - j0);
            __syncthreads();
            for (int k = threadIdx.x; k < nj; k += blockDim.x) {
                src[k] = r[j0 + k];
                sact[k] = P.gate[j0 + k] >= 0;
            }
            __syncthreads();
            for (int k = 0; k < nj; ++k) {
                double tx, ty, tz;
                const bool use = sact[k] && (j0 + k) != i;
                const double d = fp64_term(src[k], p.x, p.y, p.z, use, tx, ty, tz);
                ax += tx;
                ay += ty;
                az += tz;
                if (EV && use && live && d < scr) ens_pair_test(P, T.disk, i, j0 + k, d, Ri, r, v, mass, radius, id);
            }
        }
        const d4 res = d4{live ? ax : 0.0, live ? ay : 0.0, live ? az : 0.0, 0.0};
        if (in) {
            a[i] = res;
            ei = i;
            ecur = res;
        }
    } else {
        // One sink against many sources (a disk's star): its three sums are serial chains in
        // ascending j, the kernel's critical path.  With more than one wave, waves 1.. form
        // the terms of chunk c + 1 into one half of a double-buffered LDS tile while lanes
        // 0-2 of wave 0 add chunk c, the next 16 terms read while the current 16 are added;
        // with one wave it forms, then adds, each chunk.
        const int i = T.sink0;
        const d4 p = r[i];
        const bool live = P.gate[i] > 0;
        const double Ri = EV ? radius[i] : 0.0;
        const int c = threadIdx.x;
        const int nsrc = T.src1 - T.src0;
        const int nch = (nsrc + jtile - 1) / jtile;
        const bool split = blockDim.x > 64;
        const int p0 = split ? 64 : 0;                       // first producer work-item
        const bool producer = (int)threadIdx.x >= p0;
        auto produce = [&](int ch) {
            const int j0 = T.src0 + ch * jtile;
            const int nj = min(jtile, T.src1 - j0);
            double* t0 = term[ch & 1][0];
            double* t1 = term[ch & 1][1];
            double* t2 = term[ch & 1][2];
            for (int k = (int)threadIdx.x - p0; k < nj; k += (int)blockDim.x - p0) {
                double tx, ty, tz;
                const bool use = P.gate[j0 + k] >= 0 && (j0 + k) != i;
                (void)fp64_term(r[j0 + k], p.x, p.y, p.z, use, tx, ty, tz);
                t0[k] = tx;
                t1[k] = ty;
                t2[k] = tz;
            }
        };
        double acc = 0.0;
        auto consume = [&](int ch) {
            const int nj = min(jtile, T.src1 - (T.src0 + ch * jtile));
            const double* t = term[ch & 1][c];
            constexpr int GQ = 16;
            const int ng = nj / GQ;
            double x[GQ], y[GQ];
            if (ng > 0) {
#pragma unroll
                for (int k = 0; k < GQ; ++k) x[k] = t[k];
            }
            for (int g = 0; g < ng; g += 2) {
                if (g + 1 < ng) {
#pragma unroll
                    for (int k = 0; k < GQ; ++k) y[k] = t[(g + 1) * GQ + k];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int k = 0; k < GQ; ++k) acc += x[k];
                __builtin_amdgcn_sched_barrier(0);
                if (g + 1 < ng) {
                    if (g + 2 < ng) {
#pragma unroll
                        for (int k = 0; k < GQ; ++k) x[k] = t[(g + 2) * GQ + k];
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int k = 0; k < GQ; ++k) acc += y[k];
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            for (int k = ng * GQ; k < nj; ++k) acc += t[k];
        };
        if (split) {
            // the adding wave issues ahead of the producers on its SIMD whenever its next
            // addition is ready; the producers hold the sources of the chunk they form next
            // in registers, requested one chunk period ahead
            if (!producer) __builtin_amdgcn_s_setprio(3);
            constexpr int MS = MAX_JTILE / 64;   // source slots per producer (at least 64 of them)
            const int tp = (int)threadIdx.x - p0;
            const int np = (int)blockDim.x - p0;
            d4 q[MS];
            int qa[MS];
            auto fetch = [&](int ch) {
                const int j0 = T.src0 + ch * jtile;
                const int nj = min(jtile, T.src1 - j0);
#pragma unroll
                for (int m = 0; m < MS; ++m) {
                    const int k = tp + m * np;
                    if (k < nj) {
                        q[m] = r[j0 + k];
                        qa[m] = P.gate[j0 + k];
                    }
                }
            };
            auto form = [&](int ch) {
                const int j0 = T.src0 + ch * jtile;
                const int nj = min(jtile, T.src1 - j0);
#pragma unroll
                for (int m = 0; m < MS; ++m) {
                    const int k = tp + m * np;
                    if (k < nj) {
                        double tx, ty, tz;
                        const bool use = qa[m] >= 0 && (j0 + k) != i;
                        (void)fp64_term(q[m], p.x, p.y, p.z, use, tx, ty, tz);
                        term[ch & 1][0][k] = tx;
                        term[ch & 1][1][k] = ty;
                        term[ch & 1][2][k] = tz;
                    }
                }
            };
            if (producer && nch > 0) {
                fetch(0);
                form(0);
                if (nch > 1) fetch(1);
            }
            __syncthreads();
            for (int ch = 0; ch < nch; ++ch) {
                if (producer) {
                    if (ch + 1 < nch) {
                        form(ch + 1);
                        if (ch + 2 < nch) fetch(ch + 2);
                    }
                } else if (c < 3) {
                    consume(ch);
                }
                __syncthreads();
            }
        } else {
            for (int ch = 0; ch < nch; ++ch) {
                produce(ch);
                __syncthreads();
                if (c < 3) consume(ch);
                __syncthreads();
            }
        }
        if (EV && live) {
            // the sink's pair tests (sources j < i only: i > j in the disk's order) after its
            // sums, out of the producers' loops so that their registers stay those of the
            // sums; d is the term's distance, formed by the same operations
            for (int j = T.src0 + (int)threadIdx.x; j < min(T.src1, i); j += blockDim.x) {
                if (P.gate[j] < 0) continue;
                double tx, ty, tz;
                const double d = fp64_term(r[j], p.x, p.y, p.z, true, tx, ty, tz);
                ens_pair_test(P, T.disk, i, j, d, Ri, r, v, mass, radius, id);
            }
        }
        if (c < 3) reinterpret_cast<double*>(a + i)[c] = live ? acc : 0.0;
        if (c == 3) reinterpret_cast<double*>(a + i)[3] = 0.0;
        if (epi && threadIdx.x < 64) {
            // the three sums of lanes 0-2 gathered in lane 0 for the sink's epilogue
            const double sx = __shfl(acc, 0), sy = __shfl(acc, 1), sz = __shfl(acc, 2);
            if (threadIdx.x == 0) {
                ei = i;
                ecur = d4{live ? sx : 0.0, live ? sy : 0.0, live ? sz : 0.0, 0.0};
            }
        }
    }
    if (epi && ei >= 0) ens_epilogue(X, ei, T.disk, P.disk_dt[T.disk], ecur);
}

// y <- y_(n+1) for the bodies of disks whose last attempt was accepted (bit 2)
__global__ __launch_bounds__(256) void k_ens_accept(d4* __restrict__ yr, d4* __restrict__ yv,
                                                    const d4* __restrict__ outr, const d4* __restrict__ outv,
                                                    const int* __restrict__ disk_of, const int* __restrict__ flags,
                                                    int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int d = disk_of[i];
    if (d < 0 || !(flags[d] & 4)) return;
    yr[i] = outr[i];
    yv[i] = outv[i];
}

// Slot compaction of the disks being rebuilt: map[i] is -2 for bodies of other disks (left
// alone), -1 for the slot tail (zeroed) and otherwise the index whose accepted state moves to
// i.  The state is gathered from y_(n+1) into y, then copied back (k_ens_compact_back), so
// no launch reads what it writes.
struct EnsBodies {
    d4 *yr, *yv, *outr, *outv;
    double *mass, *radius;
    int *id, *type, *disk_of;
    unsigned char* err_on;
    double *s_mass, *s_radius;   // scratch of the gather
    int *s_id, *s_type;
};

// map[i] >= 0 is (source << 1) | err_on of the moved body.
__global__ __launch_bounds__(256) void k_ens_compact(const int* __restrict__ map, EnsBodies B, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int m = map[i];
    if (m == -2) return;
    if (m == -1) {
        const d4 z{0.0, 0.0, 0.0, 0.0};
        B.yr[i] = z; B.yv[i] = z;
        B.s_mass[i] = 0.0; B.s_radius[i] = 0.0; B.s_id[i] = 0; B.s_type[i] = 7;
        B.disk_of[i] = -1; B.err_on[i] = 0;
        return;
    }
    const int j = m >> 1;
    B.yr[i] = B.outr[j]; B.yv[i] = B.outv[j];
    B.s_mass[i] = B.mass[j]; B.s_radius[i] = B.radius[j]; B.s_id[i] = B.id[j]; B.s_type[i] = B.type[j];
    B.err_on[i] = (unsigned char)(m & 1);
}

__global__ __launch_bounds__(256) void k_ens_compact_back(const int* __restrict__ map, EnsBodies B, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || map[i] == -2) return;
    B.outr[i] = B.yr[i]; B.outv[i] = B.yv[i];
    B.mass[i] = B.s_mass[i]; B.radius[i] = B.s_radius[i]; B.id[i] = B.s_id[i]; B.type[i] = B.s_type[i];
}

// ---- per-disk step control on the device -------------------------------------------------
// Each disk's controller state lives in device arrays (SoA), so a batch of supersteps runs
// without the host: every superstep starts with k_ens_begin (which disks attempt, from their
// phase) and ends with k_ens_control (the reference's controller on the attempt's error
// maximum), one work-item per disk.  The host mirrors the arrays between batches.
struct EnsWords {
    double t_stop;
    int k_limit;       // supersteps of this batch that may run
    int stop;          // set once the kept records fill half the record capacity: the batch's
                       // later supersteps idle
    int nan_disk;      // lowest disk with a non-finite error estimate (n_disks: none)
    int limit_disk;    // lowest disk that reached 1000 attempts in one step (n_disks: none)
    int ev_count;      // raw pair records of the current attempt
    int raw_max;       // the most raw records one attempt of the batch appended
    int kept_count;    // records kept in the batch's list (one per pair and step)
    int lost;          // records that found no room in the pair table or the kept list
};

struct EnsCtl {
    double *t, *dt_try, *dt_did, *last_err;
    int *iter, *tried, *failed, *passed, *phase, *accepted_prev, *need_scan, *hold_mode;
    EnsWords* w;
    int* steplog;      // [superstep][disk]: 0 no attempt, 1 rejected, 2 + attempt index: accepted
};

// Per-attempt values of every disk (device only).
struct EnsScratch {
    int* flags;                // bit 0 attempts, bit 1 new step (forces of f0 due), bit 2 y <- y_(n+1)
    unsigned long long* err;   // error maximum (bits of a non-negative double)
    int *nan, *probe, *rec, *rec_col;
};

enum { PH_NEW = 0, PH_RETRY = 1, PH_DONE = 2, PH_HELD = 3 };
constexpr int MAX_BATCH = 64;

// Which disks attempt in superstep s: the loop condition t <= t_stop at a new step
// (main.cu:176), a step that starts with an ejection scan waits for the host (held), a
// rejected attempt is retried.  Clears the attempt's per-disk values.
__global__ __launch_bounds__(256) void k_ens_begin(EnsCtl C, EnsScratch X, int s, int nd, int scan_on)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= nd) return;
    int f = 0;
    if (!C.w->stop && s < C.w->k_limit) {
        const int ph = C.phase[d];
        if (ph == PH_NEW) {
            if (!(C.t[d] <= C.w->t_stop)) C.phase[d] = PH_DONE;
            else if (scan_on && C.need_scan[d]) C.phase[d] = PH_HELD;
            else { f = 1 | 2; C.iter[d] = 0; }
        } else if (ph == PH_RETRY) {
            f = 1;
        }
        if (f & 1) {
            C.dt_did[d] = C.dt_try[d];
            if (C.accepted_prev[d]) { f |= 4; C.accepted_prev[d] = 0; }
        }
    }
    X.flags[d] = f;
    X.err[d] = 0ull; X.nan[d] = 0; X.probe[d] = 0; X.rec[d] = 0; X.rec_col[d] = 0;
    C.steplog[s * nd + d] = 0;
}

struct EnsCtlArgs {
    double lambda, tol, exponent;
    int adaptive, accept_on_event, ce_mode, half_cap;
};

// The controller of disk d after superstep s's attempt (integrator.cu:255,
// int_rungekutta8.cu:776-797, int_rungekutta4.cu:246-273 and the fixed-step loops), with the
// host controller's arithmetic (tools::controller_update: the same operations and step_pow).
// An accepted step is held for the host when its records need a merge (collisions), or may
// lead to a compaction that moves bodies (hold_mode, set by the host: bit 0 the slot holds
// inactive bodies, so any record may complete the 8 events; bit 1 the 8 are already reached).
__device__ void control_disk(const EnsCtl& C, const EnsScratch& X, int s, int nd, int d, const EnsCtlArgs& A)
{
    if (!(X.flags[d] & 1)) return;
    bool accept = true, counted = true;
    if (A.adaptive) {
        if (X.nan[d]) {
            atomicMin(&C.w->nan_disk, d);
            C.phase[d] = PH_HELD;
            return;
        }
        const double m = __longlong_as_double((long long)X.err[d]);
        const double max_err = fabs(C.dt_try[d] * A.lambda * m);
        C.dt_try[d] = C.dt_try[d] * (0.9 * step_pow(A.tol / max_err, A.exponent));
        C.last_err[d] = max_err;
        accept = !(max_err > A.tol);
        // RK4 and Cash-Karp accept an attempt that recorded pair events and leave the loop
        // before counting it (int_rungekutta4.cu:255-264)
        if (A.accept_on_event && X.rec[d] > 0) {
            if (C.dt_try[d] < C.dt_did[d]) C.dt_try[d] = C.dt_did[d];
            accept = true;
            counted = false;
        }
    }
    const int stamp = C.iter[d];
    if (counted) C.iter[d] = stamp + 1;
    if (accept) {
        const int it = C.iter[d];
        C.tried[d] += it;
        C.failed[d] += it - 1;
        C.passed[d] += 1;
        C.t[d] = C.t[d] + C.dt_did[d];
        C.accepted_prev[d] = 1;
        C.need_scan[d] = X.probe[d] ? 1 : 0;
        C.steplog[s * nd + d] = 2 + stamp;
        const int hm = C.hold_mode[d];
        const bool hold = (!A.ce_mode && X.rec_col[d] > 0) || (hm & 2) || ((hm & 1) && X.rec[d] > 0);
        C.phase[d] = hold ? PH_HELD : PH_NEW;
    } else {
        C.phase[d] = PH_RETRY;
        C.steplog[s * nd + d] = 1;
        if (C.iter[d] >= 1000) {
            atomicMin(&C.w->limit_disk, d);
            C.phase[d] = PH_HELD;
        }
    }
}

// The attempt's raw records are in the kept list once the attempt ends: the raw list starts
// empty for the next one; a batch whose kept records passed half the capacity idles on.
__device__ void control_words(const EnsCtl& C, int kept, const EnsCtlArgs& A)
{
    C.w->raw_max = max(C.w->raw_max, C.w->ev_count);
    C.w->ev_count = 0;
    if (kept > A.half_cap) C.w->stop = 1;
}

__global__ __launch_bounds__(256) void k_ens_control(EnsCtl C, EnsScratch X, int s, int nd, EnsCtlArgs A)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d == 0) control_words(C, C.w->kept_count, A);
    if (d < nd) control_disk(C, X, s, nd, d, A);
}

__global__ __launch_bounds__(256) void k_dedup_insert(EnsDedup D, EnsWords* w, int s)
{
    const int n = min(w->ev_count, D.raw_cap);   // records past the capacity were not written
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const red_event& e = D.raw[q];
        const int i = pair_slot(D, pair_key(D, e, s));
        if (i < 0) { atomicAdd(&w->lost, 1); continue; }
        atomicMin(&D.tmin[i], time_key(e.t));
        atomicMin(&D.first[i], (unsigned int)e.stage);
    }
}

__global__ __launch_bounds__(256) void k_dedup_mark(EnsDedup D, EnsWords* w, int s)
{
    const int n = min(w->ev_count, D.raw_cap);
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const red_event& e = D.raw[q];
        const int i = pair_find(D, pair_key(D, e, s));
        if (i >= 0 && time_key(e.t) == D.tmin[i]) atomicMin(&D.kept[i], (unsigned int)e.stage);
    }
}

// The kept records into the batch's list.
template <bool ALL>
__global__ __launch_bounds__(256) void k_dedup_emit(EnsDedup D, EnsWords* w, int s)
{
    const int n = min(w->ev_count, D.raw_cap);
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const red_event& e = D.raw[q];
        unsigned int first = (unsigned int)e.stage;
        if (!ALL) {
            const int i = pair_find(D, pair_key(D, e, s));
            if (i < 0 || D.kept[i] != (unsigned int)e.stage) continue;
            first = D.first[i];
        }
        const int k = atomicAdd(&w->kept_count, 1);
        if (k >= D.out_cap) { atomicAdd(&w->lost, 1); continue; }
        red_event* o = D.out + k;
        *o = e;
        o->attempt = s | (int)(first << 16);
    }
}

// step_pow over arrays, as the device controller evaluates it
__global__ __launch_bounds__(256) void k_step_pow(const double* __restrict__ x, const double* __restrict__ y,
                                                  double* __restrict__ out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = step_pow(x[i], y[i]);
}

// The controller state's device layout (the host twin has the same): per-disk arrays, the
// batch words, then the step log.  Returns the bytes of the part the host uploads.
size_t ctl_layout(int nd, size_t off[4])
{
    const size_t n = (size_t)std::max(nd, 1);
    off[0] = 0;                                              // 4 double arrays
    off[1] = off[0] + 4 * n * sizeof(double);                // 8 int arrays
    off[2] = (off[1] + 8 * n * sizeof(int) + 15) / 16 * 16;  // batch words
    off[3] = off[2] + sizeof(EnsWords);                      // step log
    return off[3];
}

size_t carve_ctl(char* base, int nd, EnsCtl* C)
{
    const size_t n = (size_t)std::max(nd, 1);
    size_t off[4];
    const size_t up = ctl_layout(nd, off);
    double* dp = (double*)(base + off[0]);
    C->t = dp; C->dt_try = dp + n; C->dt_did = dp + 2 * n; C->last_err = dp + 3 * n;
    int* ip = (int*)(base + off[1]);
    C->iter = ip; C->tried = ip + n; C->failed = ip + 2 * n; C->passed = ip + 3 * n;
    C->phase = ip + 4 * n; C->accepted_prev = ip + 5 * n; C->need_scan = ip + 6 * n; C->hold_mode = ip + 7 * n;
    C->w = (EnsWords*)(base + off[2]);
    C->steplog = (int*)(base + off[3]);
    return up;
}

size_t ctl_bytes(int nd)
{
    size_t off[4];
    return ctl_layout(nd, off) + sizeof(int) * (size_t)MAX_BATCH * (size_t)std::max(nd, 1);
}

template <typename T>
hipError_t dmalloc(T** p, size_t count)
{
    *p = nullptr;
    return hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T));
}

}  // namespace

struct red_ensemble {
    red_params prm{};
    Scheme S;
    hipStream_t stream = nullptr;
    std::string err;
    int nd = 0, n = 0;
    std::vector<int> b0, b1;                 // disk d's bodies are [b0, b1); its slot ends at the next b0
    std::vector<std::array<int, 7>> cnt;     // per disk: ns ngp nrp npp nspl npl ntp after removals
    std::vector<int> disk_of_h;              // -1 for the slot tail left by a compaction
    // host mirror of the per-body parameters (pp_disk::h_p, h_body_md), all disks
    std::vector<double> h_mass, h_radius, h_density, h_cd;
    std::vector<int> h_id, h_type;
    // device
    d4 *yr = nullptr, *yv = nullptr, *outr = nullptr, *outv = nullptr;
    std::vector<d4*> sr, sv, acc;
    int* disk_of = nullptr;
    unsigned char* err_on = nullptr;
    double *d_mass = nullptr, *d_radius = nullptr;
    int *d_id = nullptr, *d_type = nullptr;
    // per-disk controller state on the device and its pinned host twin (same layout): one
    // block each way per batch of supersteps
    char* d_ctl = nullptr;
    char* h_ctl = nullptr;
    size_t ctl_up = 0, ctl_all = 0;          // bytes uploaded / downloaded
    EnsCtl dc{}, hc{};
    // per-attempt values (device only)
    char* d_scr = nullptr;
    EnsScratch scr{};
    double *d_dt = nullptr, *d_t = nullptr;  // the attempt's step and step-start time (in d_ctl)
    int* d_flags = nullptr;
    unsigned long long* d_err = nullptr;
    int* d_nan = nullptr;
    int* d_map = nullptr;                    // slot compaction map (k_ens_compact)
    int batch = 8;                           // supersteps per graph launch
    // the attempts' records deduplicated on the device into the batch's kept list
    EnsDedup dedup{};
    char* d_table = nullptr;                 // the pair table, cleared before every batch
    size_t table_bytes = 0;
    red_event* d_kept = nullptr;
    int kept_cap = 0;
    int dedup_blocks = 1;
    // a batch's event passes left for the time the next batch runs: the disks, their records
    // (pointers into h_events) and the batch's step log
    std::vector<int> def_list;
    std::vector<std::vector<const red_event*>> def_rec;
    std::vector<int> def_log;
    int def_k = 0;
    std::vector<int> inactive;               // inactive bodies in each disk's slot (until compaction)
    StageArgs* d_rows = nullptr;             // the stage combinations' arguments ([0] y_(n+1))
    int* d_ones = nullptr;                   // padded storage: every body gates as active
    bool padded = false;
    double* d_cscratch = nullptr;            // its gather scratch: mass, radius, id, type
    EnsDiskPair* d_pair = nullptr;
    red_event* d_events = nullptr;
    int ev_cap = 0;
    int* d_ev_count = nullptr;               // in d_ctl's batch words
    BodyUpdate* d_updates = nullptr;
    int updates_cap = 0;
    EnsTile* d_tiles = nullptr;
    int n_tiles = 0;
    int block = 256, jtile = 256;
    bool pair_tests = false, scan_on = false;
    // host mirror of the per-disk controller state (the device's between batches)
    std::vector<double> t, dt_try, dt_did, last_err;
    std::vector<int> iter, tried, failed, passed, phase;   // PH_NEW, PH_RETRY, PH_DONE, PH_HELD
    std::vector<int> accepted_prev;
    std::vector<int> need_scan;              // the next step starts with an exact ejection scan
    std::vector<int> hold_mode;
    std::vector<events::Counters> evc;
    // pair records of one step of a disk, in the pinned read-back h_events: an attempt that
    // recorded pair events is accepted (int_rungekutta4.cu:255-264, and every attempt of a
    // fixed step) or its records are cleared with the next attempt (RKF7(8),
    // int_rungekutta8.cu:730-733), so a step's records are those of its accepted attempt
    std::vector<std::vector<const red_event*>> pending;
    std::vector<int> pending_attempt;        // the attempt counter when they were recorded
    std::vector<std::vector<red_event>> log;       // handled records (sp lists) not yet read
    red_event* h_events = nullptr;         // pinned staging of a batch's pair records
    int h_events_cap = 0;
    int* h_flags = nullptr;                // pinned: flags for the closing y <- y_(n+1) copy
    hipGraphExec_t graph = nullptr;
    int graph_batch = 0;                     // supersteps in the captured graph
    bool use_graph = true;
    std::vector<char> compact_due;           // the disk's slot is compacted after this attempt
    double event_s = 0.0, rebuild_s = 0.0;   // host time of the current advance call
    std::vector<unsigned char> h_err_on;     // host mirror of err_on
    int* h_map = nullptr;                    // pinned; -2 outside the slots being compacted
    std::vector<EnsDiskPair> h_pair;         // host copy of d_pair
    int tiles_cap = 0;                       // d_tiles holds tiles_cap entries, empty ones at the end
    int tiles_block = 0;                     // the launch shape the tiles and the graph were made for
    int host_threads = 1;
    d4* h_state = nullptr;                   // pinned copies of yr, yv, outr, outv (n each)

    int nSI(int d) const { return cnt[d][0] + cnt[d][1] + cnt[d][2] + cnt[d][3]; }
    int nMassive(int d) const { return nSI(d) + cnt[d][4] + cnt[d][5]; }
    events::BodyView view(int d)
    {
        const int o = b0[d];
        return events::BodyView{h_mass.data() + o, h_radius.data() + o, h_density.data() + o, h_cd.data() + o,
                                h_id.data() + o,   h_type.data() + o,   nMassive(d)};
    }
    int fail(int code, const std::string& m) { err = m; return code; }
    int hipfail(hipError_t e, const char* where) { err = std::string(where) + ": " + hipGetErrorString(e); return RED_ERR_HIP; }
};

#define ENS_HIP(E, expr)                                        \
    do {                                                        \
        hipError_t _e = (expr);                                 \
        if (_e != hipSuccess) return (E)->hipfail(_e, #expr);   \
    } while (0)

namespace {

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Tiles of every disk's three sweeps (number_of_bodies.cpp:196-258 per disk), largest first.
std::vector<EnsTile> build_tiles(const red_ensemble* E, int block)
{
    std::vector<EnsTile> tiles;
    for (int d = 0; d < E->nd; d++) {
        const int b0 = E->b0[d], b1 = E->b1[d], si = E->nSI(d), m = E->nMassive(d);
        auto add = [&](int s0, int s1, int j0, int j1) {
            if (s1 <= s0) return;
            if (s1 - s0 <= 4 && j1 - j0 >= 64) {
                for (int i = s0; i < s1; i++) tiles.push_back(EnsTile{i, i + 1, j0, j1, d, 1});
                return;
            }
            for (int i = s0; i < s1; i += block) tiles.push_back(EnsTile{i, std::min(s1, i + block), j0, j1, d, 0});
        };
        add(b0, b0 + si, b0, b0 + m);              // SI sinks x massive sources
        add(b0 + si, b0 + m, b0, b0 + si);         // NSI sinks x SI sources
        add(b0 + m, b1, b0, b0 + m);               // NI sinks x massive sources
    }
    std::stable_sort(tiles.begin(), tiles.end(), [](const EnsTile& x, const EnsTile& y) {
        const long cx = (long)(x.kind ? 64 : (x.sink1 - x.sink0)) * (x.src1 - x.src0);
        const long cy = (long)(y.kind ? 64 : (y.sink1 - y.sink0)) * (y.src1 - y.src0);
        return cx > cy;
    });
    return tiles;
}

// The pair-test constants of disk d from the host mirror (as refresh_pair_test of a single
// disk): 3 M_star, the largest radius, the largest non-star mass, the number of stars.
EnsDiskPair pair_consts(red_ensemble* E, int d)
{
    events::BodyView B = E->view(d);
    const double ms = events::mass_of_star(B);
    EnsDiskPair P;
    P.three_mstar = 3.0 * (ms > 0 ? ms : 1.0);
    double rall = 0.0, mns = 0.0;
    const int nb = E->b1[d] - E->b0[d];
    for (int i = 0; i < nb; i++) {
        rall = std::max(rall, B.radius[i]);
        if (i >= E->cnt[d][0]) mns = std::max(mns, B.mass[i]);
    }
    P.rmax = rall;
    P.mmax = mns;
    P.n_star = E->cnt[d][0];
    P.b0 = E->b0[d];
    P.b1 = E->b1[d];
    return P;
}

int upload_pair(red_ensemble* E, int d)
{
    E->h_pair[d] = pair_consts(E, d);
    ENS_HIP(E, stream_copy(E->stream, E->d_pair + d, &E->h_pair[d], sizeof(EnsDiskPair), hipMemcpyHostToDevice));
    return RED_OK;
}

// The pair-test constants of the disks in `list` recomputed, the whole table uploaded once.
int upload_pairs(red_ensemble* E, const std::vector<int>& list)
{
    if (E->nd == 0) return RED_OK;
    for (int d : list) E->h_pair[d] = pair_consts(E, d);
    ENS_HIP(E, stream_copy(E->stream, E->d_pair, E->h_pair.data(), sizeof(EnsDiskPair) * E->nd, hipMemcpyHostToDevice));
    return RED_OK;
}

const d4* ens_stage_v(red_ensemble* E, int k)
{
    if (k == 0) return E->yv;
    if (E->S.extra_on_yout && k == 4) return E->outv;
    return E->sv[k];
}

StageArgs ens_args(red_ensemble* E, const Row& row, d4* outr, d4* outv)
{
    StageArgs A;
    memset(&A, 0, sizeof(A));
    A.yr = E->yr; A.yv = E->yv; A.outr = outr; A.outv = outv; A.n = E->n;
    A.nterms = row.nterms; A.c1 = row.c1; A.c2 = row.c2; A.ctl = nullptr;
    A.disk_of = E->disk_of; A.disk_dt = E->d_dt; A.disk_flags = E->d_flags; A.err_on = E->err_on;
    A.disk_err = E->d_err; A.disk_nan = E->d_nan;
    for (int k = 0; k < row.nterms; k++) {
        A.t[k].coef = row.coef[k];
        A.t[k].fv = ens_stage_v(E, row.stage[k]);
        A.t[k].fa = E->acc[row.stage[k]];
    }
    return A;
}

hipError_t ens_force(red_ensemble* E, int stage, const d4* r, const d4* v, int superstep, bool epilogue = true)
{
    if (E->n_tiles == 0) return hipSuccess;
    EnsPair P;
    P.radii_f = E->prm.radii_enhance_factor;
    P.hill_f = E->prm.encounter_hill_factor;
    P.close_encounter = E->prm.close_encounter;
    P.enabled = E->pair_tests ? 1 : 0;
    P.dp = E->d_pair;
    P.disk_t = E->d_t;
    P.disk_dt = E->d_dt;
    P.c_stage = E->S.c[stage];
    P.stage = stage;
    P.buf = E->d_events;
    P.cap = E->ev_cap;
    P.count = E->d_ev_count;
    P.superstep = superstep;
    P.rec = E->scr.rec;
    P.rec_col = E->scr.rec_col;
    P.gate = E->padded ? E->d_ones : E->d_id;
    const int need = stage == 0 ? 2 : 1;
    // the stage combinations formed in the launch's epilogue: stage state stage + 1 (main
    // stages, and the error stages of schemes that form them from stage sums), y_(n+1) after
    // the last main stage
    EnsEpi X;
    X.rows = E->d_rows;
    X.cur_a = E->acc[stage];
    X.next = 0; X.next_form = 0; X.np1 = 0; X.np1_form = 0;
    X.probe_on = 0;
    X.ej = E->prm.ejection; X.hc = E->prm.hit_centrum;
    X.id = E->d_id; X.type = E->d_type;
    X.probe = E->scr.probe;
    if (epilogue) {
        const Scheme& S = E->S;
        const bool adaptive = E->prm.adaptive && S.err_form >= 0;
        const int nx = stage + 1;
        if (nx < (int)S.rows.size() && (nx < S.n_main || (adaptive && !S.extra_on_yout && nx < S.n_stage))) {
            X.next = nx;
            X.next_form = S.rows[nx].form;
        }
        if (stage == S.n_main - 1) {
            X.np1 = 1;
            X.np1_form = S.np1.form;
            X.probe_on = E->scan_on ? 1 : 0;
        }
    }
    // RKF7(8) clears the records at every attempt start (int_rungekutta8.cu:730-733), so f0's
    // would be dropped unread: its launch runs without pair tests
    const bool ev = E->pair_tests && !(stage == 0 && E->S.clear_events_each_attempt);
    if (ev)
        hipLaunchKernelGGL(k_ens_force<true>, dim3(E->n_tiles), dim3(E->block), 0, E->stream, E->d_tiles, r, v,
                           E->acc[stage], E->d_mass, E->d_radius, E->d_id, E->d_flags, need, E->jtile, P, X);
    else
        hipLaunchKernelGGL(k_ens_force<false>, dim3(E->n_tiles), dim3(E->block), 0, E->stream, E->d_tiles, r, v,
                           E->acc[stage], E->d_mass, E->d_radius, E->d_id, E->d_flags, need, E->jtile, P, X);
    return hipGetLastError();
}

// Superstep s of a batch: which disks attempt, their attempt, their controllers.
hipError_t ens_superstep(red_ensemble* E, int s)
{
    const Scheme& S = E->S;
    const dim3 bb(256), gd((E->nd + 255) / 256);
    hipLaunchKernelGGL(k_ens_begin, gd, bb, 0, E->stream, E->dc, E->scr, s, E->nd, E->scan_on ? 1 : 0);
    RED_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_ens_accept, dim3((E->n + 255) / 256), bb, 0, E->stream, E->yr, E->yv, E->outr, E->outv,
                       E->disk_of, E->d_flags, E->n);
    RED_HIP_TRY(hipGetLastError());
    // each force launch forms the next stage state (and y_(n+1), with the ejection predicate)
    // in its epilogue
    RED_HIP_TRY(ens_force(E, 0, E->yr, E->yv, s));
    for (int r = 1; r < S.n_main; r++) RED_HIP_TRY(ens_force(E, r, E->sr[r], E->sv[r], s));
    const bool adaptive = E->prm.adaptive && S.err_form >= 0;
    if (adaptive) {
        for (int r = S.n_main; r < S.n_stage; r++) {
            if (S.extra_on_yout) RED_HIP_TRY(ens_force(E, r, E->outr, E->outv, s));
            else RED_HIP_TRY(ens_force(E, r, E->sr[r], E->sv[r], s));
        }
        RED_HIP_TRY(launch_err(S.err_form, ens_args(E, S.err, nullptr, nullptr), E->stream));
    }
    EnsCtlArgs A;
    A.lambda = S.lambda; A.tol = E->prm.tolerance; A.exponent = S.exponent;
    A.adaptive = adaptive ? 1 : 0;
    A.accept_on_event = S.accept_on_event ? 1 : 0;
    A.ce_mode = E->prm.close_encounter ? 1 : 0;
    A.half_cap = E->ev_cap / 2;
    if (E->pair_tests) {
        // the attempt's records, one per pair (all of them in close-encounter mode), into the
        // batch's list
        const dim3 gq(E->dedup_blocks);
        if (!E->prm.close_encounter) {
            hipLaunchKernelGGL(k_dedup_insert, gq, bb, 0, E->stream, E->dedup, E->dc.w, s);
            RED_HIP_TRY(hipGetLastError());
            hipLaunchKernelGGL(k_dedup_mark, gq, bb, 0, E->stream, E->dedup, E->dc.w, s);
            RED_HIP_TRY(hipGetLastError());
            hipLaunchKernelGGL(k_dedup_emit<false>, gq, bb, 0, E->stream, E->dedup, E->dc.w, s);
        } else {
            hipLaunchKernelGGL(k_dedup_emit<true>, gq, bb, 0, E->stream, E->dedup, E->dc.w, s);
        }
        RED_HIP_TRY(hipGetLastError());
    }
    // one work-item per disk (64 per work-group: the controllers spread over more CUs)
    hipLaunchKernelGGL(k_ens_control, dim3((E->nd + 63) / 64), dim3(64), 0, E->stream, E->dc, E->scr, s, E->nd, A);
    return hipGetLastError();
}

// A batch: `batch` supersteps (the body of the captured graph).
hipError_t ens_batch(red_ensemble* E)
{
    for (int s = 0; s < E->batch; s++) RED_HIP_TRY(ens_superstep(E, s));
    return hipSuccess;
}

void destroy_graph(red_ensemble* E)
{
    SetupLock lk;
    if (E->graph) (void)hipGraphExecDestroy(E->graph);
    E->graph = nullptr;
}

hipError_t run_batch(red_ensemble* E)
{
    if (!E->use_graph) return ens_batch(E);
    if (E->graph && E->graph_batch != E->batch) destroy_graph(E);
    if (!E->graph) {
        SetupLock lk;
        hipGraph_t g = nullptr;
        RED_HIP_TRY(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = ens_batch(E);
        hipError_t e2 = hipStreamEndCapture(E->stream, &g);
        if (e != hipSuccess) { if (g) (void)hipGraphDestroy(g); return e; }
        RED_HIP_TRY(e2);
        e = hipGraphInstantiate(&E->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        RED_HIP_TRY(e);
        E->graph_batch = E->batch;
    }
    return hipGraphLaunch(E->graph, E->stream);
}

// The force launches cover tiles_cap tiles, the unused ones empty (no sinks, no sources), so
// a rebuild that does not need more tiles keeps the launch shape and the captured graph.
int upload_tiles(red_ensemble* E)
{
    std::vector<EnsTile> tiles = build_tiles(E, E->block);
    if (E->tiles_block != E->block) {
        // tiles hold at most `block` sinks, one per work-item of a launch of that shape
        destroy_graph(E);
        E->tiles_block = E->block;
    }
    // room for a quarter more tiles than needed; a table more than twice that (left by a
    // smaller launch shape, whose tiles hold fewer sinks) shrinks back, since every empty tile
    // is still a work-group of every force launch
    const int want = (int)tiles.size() + (int)tiles.size() / 4 + 64;
    if ((int)tiles.size() > E->tiles_cap || E->tiles_cap > 2 * want) {
        SetupLock lk;
        (void)hipStreamSynchronize(E->stream);
        if (E->d_tiles) (void)hipFree(E->d_tiles);
        E->d_tiles = nullptr;
        E->tiles_cap = want;
        ENS_HIP(E, dmalloc(&E->d_tiles, E->tiles_cap));
        destroy_graph(E);
    }
    E->n_tiles = (int)tiles.size() > 0 ? E->tiles_cap : 0;
    tiles.resize(E->tiles_cap, EnsTile{0, 0, 0, 0, 0, 0});
    ENS_HIP(E, stream_copy(E->stream, E->d_tiles, tiles.data(), sizeof(EnsTile) * tiles.size(), hipMemcpyHostToDevice));
    return RED_OK;
}

// The buffers holding disk d's current state: y_(n+1) of an accepted attempt until the next
// attempt's first launch moves it into y.
d4* cur_r(red_ensemble* E, int d) { return E->accepted_prev[d] ? E->outr : E->yr; }
d4* cur_v(red_ensemble* E, int d) { return E->accepted_prev[d] ? E->outv : E->yv; }

// Merge results (disk-local indices) into the current state and the device parameters.
int apply_updates(red_ensemble* E, int d, std::vector<BodyUpdate> ups)
{
    if (ups.empty()) return RED_OK;
    for (auto& u : ups) u.idx += E->b0[d];
    if ((int)ups.size() > E->updates_cap) {
        SetupLock lk;
        (void)hipStreamSynchronize(E->stream);
        if (E->d_updates) (void)hipFree(E->d_updates);
        E->d_updates = nullptr;
        E->updates_cap = (int)ups.size() * 2;
        ENS_HIP(E, dmalloc(&E->d_updates, E->updates_cap));
    }
    ENS_HIP(E, hipMemcpyAsync(E->d_updates, ups.data(), sizeof(BodyUpdate) * ups.size(), hipMemcpyHostToDevice,
                              E->stream));
    ENS_HIP(E, launch_apply_updates(E->d_updates, (int)ups.size(), cur_r(E, d), cur_v(E, d), E->d_mass, E->d_radius,
                                    E->d_id, E->stream));
    ENS_HIP(E, hipStreamSynchronize(E->stream));
    return upload_pair(E, d);
}

// Host copies of all four state buffers, for the scans of many disks at once (one transfer
// each instead of two per disk).
int fetch_states(red_ensemble* E)
{
    const size_t n = (size_t)std::max(E->n, 1);
    if (!E->h_state) {
        SetupLock lk;
        ENS_HIP(E, hipHostMalloc((void**)&E->h_state, sizeof(d4) * 4 * n, hipHostMallocDefault));
    }
    const d4* src[4] = {E->yr, E->yv, E->outr, E->outv};
    for (int k = 0; k < 4; k++)
        ENS_HIP(E, hipMemcpyAsync(E->h_state + k * n, src[k], sizeof(d4) * n, hipMemcpyDeviceToHost, E->stream));
    ENS_HIP(E, hipStreamSynchronize(E->stream));
    return RED_OK;
}

// pp_disk.cu:683-715 for disk d: the scan on its current state, its records and ids, the
// hit-centrum merges (main.cu:201-208); the pair records since the last clear are dropped.
// from_states: the state is read from the copies fetch_states made.
int ejection_check(red_ensemble* E, int d, bool from_states)
{
    const int o = E->b0[d], nb = E->b1[d] - o;
    if (nb <= 0) return RED_OK;
    std::vector<d4> Rl, Vl;
    const d4 *R, *V;
    if (from_states) {
        const size_t n = (size_t)std::max(E->n, 1);
        const size_t k = E->accepted_prev[d] ? 2 : 0;
        R = E->h_state + k * n + o;
        V = E->h_state + (k + 1) * n + o;
    } else {
        Rl.resize(nb); Vl.resize(nb);
        ENS_HIP(E, stream_copy(E->stream, Rl.data(), cur_r(E, d) + o, sizeof(d4) * nb, hipMemcpyDeviceToHost));
        ENS_HIP(E, stream_copy(E->stream, Vl.data(), cur_v(E, d) + o, sizeof(d4) * nb, hipMemcpyDeviceToHost));
        R = Rl.data(); V = Vl.data();
    }
    std::vector<int> ids(E->h_id.begin() + o, E->h_id.begin() + o + nb);   // the device's ids
    std::vector<red_event> ev;
    events::ejection_scan(nb, &R[0].x, &V[0].x, ids.data(), E->h_type.data() + o, E->t[d], E->prm.ejection,
                          E->prm.hit_centrum, ev);
    if (ev.empty()) return RED_OK;
    for (auto& e : ev) e.pad = d;
    events::BodyView B = E->view(d);
    std::vector<BodyUpdate> ups;
    std::string msg;
    const int rc = events::ejection_pass(B, ev, E->evc[d], ups, msg);
    if (rc) return E->fail(rc, "disk " + std::to_string(d) + ": " + msg);
    // the negated ids (the host mirror now), then the merge results
    ENS_HIP(E, stream_copy(E->stream, E->d_id + o, E->h_id.data() + o, sizeof(int) * nb, hipMemcpyHostToDevice));
    const int rc2 = apply_updates(E, d, ups);
    if (rc2) return rc2;
    E->log[d].insert(E->log[d].end(), ev.begin(), ev.end());
    E->pending[d].clear();   // clear_event_counter (main.cu:205)
    return upload_pair(E, d);
}

// pp_disk.cu:893-907, 1324-1409 and number_of_bodies.cpp:27-104 for every disk marked in
// compact_due (each accepted its last attempt, so its state is y_(n+1)): its active bodies
// move to the front of its slot in order, the rest of the slot is left out of every launch.
// The host mirror is compacted here; the device arrays by two gather launches over one map
// for all the disks rebuilt after an attempt.
int compact_all(red_ensemble* E)
{
    std::vector<int> list;
    for (int d = 0; d < E->nd; d++)
        if (E->compact_due[d]) list.push_back(d);
    if (list.empty()) return RED_OK;
    int* map = E->h_map;
    std::vector<int> moved;   // disks whose slot changes
    for (int d : list) {
        if (!E->accepted_prev[d]) return E->fail(RED_ERR_ARG, "compaction of a disk whose state is not y_(n+1)");
        const int o = E->b0[d], nb = E->b1[d] - o;
        int removed[7] = {0};
        std::vector<int> keep;
        for (int i = 0; i < nb; i++) {
            const int t = E->h_type[o + i];
            if (0 < E->h_id[o + i] && 7 > t) { keep.push_back(i); continue; }
            if (t < 0 || t > 7) return E->fail(RED_ERR_ARG, "Undefined body type!");
            if (t == 7) continue;
            removed[t]++;
        }
        for (int k = 0; k < 7; k++) E->cnt[d][k] -= removed[k];
        const int m = (int)keep.size();
        int nn = 0;
        for (int k = 0; k < 7; k++) nn += E->cnt[d][k];
        if (nn != m) return E->fail(RED_ERR_ARG, "Error: number of copied bodies does not equal to the number of active bodies.");
        E->evc[d].set(events::CNT_LAST_CLEAR, 0);
        if (m == nb) {   // every body kept in place: the rebuild leaves the disk as it is
            E->compact_due[d] = 0;
            continue;
        }
        moved.push_back(d);
        E->b1[d] = o + m;
        const int nm = E->nMassive(d);
        for (int k = 0; k < m; k++) {
            const int i = keep[k];
            const int eon = E->prm.error_check_for_tp || k < nm ? 1 : 0;
            map[o + k] = ((o + i) << 1) | eon;
            E->h_err_on[o + k] = (unsigned char)eon;
            E->h_mass[o + k] = E->h_mass[o + i]; E->h_radius[o + k] = E->h_radius[o + i];
            E->h_density[o + k] = E->h_density[o + i]; E->h_cd[o + k] = E->h_cd[o + i];
            E->h_id[o + k] = E->h_id[o + i]; E->h_type[o + k] = E->h_type[o + i];
        }
        // the slot's tail: massless, inactive padding that no launch visits
        for (int k = m; k < nb; k++) {
            map[o + k] = -1;
            E->h_err_on[o + k] = 0;
            E->h_mass[o + k] = 0.0; E->h_radius[o + k] = 0.0; E->h_density[o + k] = 0.0; E->h_cd[o + k] = 0.0;
            E->h_id[o + k] = 0; E->h_type[o + k] = 7;
            E->disk_of_h[o + k] = -1;
        }
    }
    if (moved.empty()) return RED_OK;
    const int n = E->n;
    if (n > 0) {
        EnsBodies B;
        B.yr = E->yr; B.yv = E->yv; B.outr = E->outr; B.outv = E->outv;
        B.mass = E->d_mass; B.radius = E->d_radius; B.id = E->d_id; B.type = E->d_type;
        B.disk_of = E->disk_of; B.err_on = E->err_on;
        B.s_mass = E->d_cscratch; B.s_radius = E->d_cscratch + n;
        B.s_id = (int*)(E->d_cscratch + 2 * (size_t)n); B.s_type = B.s_id + n;
        const dim3 gb((n + 255) / 256), bb(256);
        ENS_HIP(E, stream_copy(E->stream, E->d_map, map, sizeof(int) * n, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_ens_compact, gb, bb, 0, E->stream, E->d_map, B, n);
        ENS_HIP(E, hipGetLastError());
        hipLaunchKernelGGL(k_ens_compact_back, gb, bb, 0, E->stream, E->d_map, B, n);
        ENS_HIP(E, hipGetLastError());
        ENS_HIP(E, hipStreamSynchronize(E->stream));
    }
    // the map back to "left alone" for the next rebuild
    for (int d : moved) {
        const int end = d + 1 < E->nd ? E->b0[d + 1] : E->n;
        for (int i = E->b0[d]; i < end; i++) map[i] = -2;
        E->compact_due[d] = 0;
    }
    int rc = upload_pairs(E, moved);
    if (rc) return rc;
    return upload_tiles(E);
}

// The checks after an accepted step of disk d (main.cu:225-255), on its new state: the
// event passes on the host mirror of this disk alone (disks run on separate threads).
// ups: the merges' device updates (disk-local indices); *compact: the slot is rebuilt.
int after_step_host(red_ensemble* E, int d, std::vector<BodyUpdate>& ups, bool* compact, std::string& msg)
{
    *compact = false;
    const red_params& P = E->prm;
    std::vector<const red_event*>& raw = E->pending[d];
    std::vector<red_event>& log = E->log[d];
    // the pass's sp list is made in place at the end of the disk's log, the records stamped
    // with the step's attempt
    auto stamp = [&](size_t base) {
        for (size_t k = base; k < log.size(); k++) { log[k].attempt = E->pending_attempt[d]; log[k].pad = 0; }
    };
    if (P.close_encounter) {
        if (E->pair_tests && !raw.empty()) {
            events::canonical_sort(raw);
            const size_t base = log.size();
            log.resize(base + raw.size());
            for (size_t k = 0; k < raw.size(); k++) log[base + k] = *raw[k];
            stamp(base);
            events::close_encounter_sp_pass(E->view(d), log.data() + base, raw.size(), E->evc[d]);
            raw.clear();
        }
        return RED_OK;
    }
    if (E->pair_tests && !raw.empty()) {
        // one record per pair already (k_dedup_emit): put them where each pair first occurs
        events::kept_order(raw);
        const size_t base = log.size();
        log.resize(base + raw.size());
        for (size_t k = 0; k < raw.size(); k++) log[base + k] = *raw[k];
        stamp(base);
        events::BodyView B = E->view(d);
        const int rc = events::collision_sp_pass(B, log.data() + base, log.size() - base, E->evc[d], ups, msg);
        if (rc) return rc;
        raw.clear();
    }
    *compact = E->evc[d].event[events::CNT_LAST_CLEAR] >= 8;
    return RED_OK;
}

// f(d) for every disk of the list, on up to `threads` threads.
template <class F>
void for_disks(int threads, const std::vector<int>& list, F f)
{
    const int nt = std::min<int>(threads, (int)list.size());
    if (nt <= 1) {
        for (int d : list) f(d);
        return;
    }
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (size_t k; (k = next.fetch_add(1)) < list.size();) f(list[k]);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; t++) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
}

// Inactive bodies in disk d's slot (merged, ejected or hit the centre, until a compaction).
int count_inactive(const red_ensemble* E, int d)
{
    int k = 0;
    for (int i = E->b0[d]; i < E->b1[d]; i++) k += !(0 < E->h_id[i] && 7 > E->h_type[i]);
    return k;
}

// Whether the host must see disk d after its next accepted step (k_ens_control): without
// inactive bodies a compaction keeps the slot as it is, so only merges need the host; with
// them, any record may complete the 8 events of a compaction, or the 8 are already there.
int hold_mode_of(const red_ensemble* E, int d)
{
    if (E->prm.close_encounter || E->inactive[d] == 0) return 0;
    return E->evc[d].event[events::CNT_LAST_CLEAR] >= 8 ? 3 : 1;
}

// The checks after disk d's accepted steps of a batch (main.cu:225-255), in step order: each
// step's records are those its accepted attempt recorded (the superstep is in the record's
// attempt field, R holds the disk's records of the batch).  A step the device did not hold
// changes nothing on the device (no merge; a compaction, if due, keeps every body in place);
// a held disk's last step may merge (ups) and then compact (*comp).
int disk_steps_host(red_ensemble* E, int d, std::vector<const red_event*>& R, const int* log, int K,
                    std::vector<BodyUpdate>& ups, char* comp, std::string& msg)
{
    const int nd = E->nd;
    std::stable_sort(R.begin(), R.end(), [](const red_event* a, const red_event* b) {
        return (a->attempt & 0xffff) < (b->attempt & 0xffff);
    });
    int last = -1;
    for (int s = 0; s < K; s++)
        if (log[s * nd + d] >= 2) last = s;
    size_t q = 0;
    for (int s = 0; s <= last; s++) {
        const size_t q0 = q;
        while (q < R.size() && (R[q]->attempt & 0xffff) == s) q++;
        if (log[s * nd + d] < 2) continue;   // a rejected attempt's records are dropped
        E->pending[d].assign(R.begin() + q0, R.begin() + q);
        E->pending_attempt[d] = log[s * nd + d] - 2;
        std::vector<BodyUpdate> u;
        bool c = false;
        const int rc = after_step_host(E, d, u, &c, msg);
        if (rc) return rc;
        const bool held = s == last && E->phase[d] == PH_HELD;
        if (!u.empty() && !held) { msg = "a merge in a step the device did not hold"; return RED_ERR_ARG; }
        if (!u.empty()) ups = std::move(u);
        if (c) {
            if (count_inactive(E, d) == 0) {
                E->evc[d].set(events::CNT_LAST_CLEAR, 0);   // the rebuild keeps every body in place
            } else if (held) {
                *comp = 1;
            } else {
                msg = "a compaction in a step the device did not hold";
                return RED_ERR_ARG;
            }
        }
    }
    return RED_OK;
}

// After a batch: the disks whose state the host changes before they step again -- held by the
// device, or starting their next step with an ejection scan -- have their event passes now,
// then every merge's device update in one launch and the compactions.  The other disks'
// passes only add records and counts on the host: they are kept for after_batch_deferred,
// which runs while the next batch is on the device.
int after_batch(red_ensemble* E, int ne)
{
    const int nd = E->nd, K = E->batch;
    const auto t_ev = std::chrono::steady_clock::now();
    E->def_log.assign(E->hc.steplog, E->hc.steplog + (size_t)K * nd);
    E->def_k = K;
    const int* log = E->def_log.data();
    std::vector<int> now;
    E->def_list.clear();
    for (int d = 0; d < nd; d++) {
        bool stepped = false;
        for (int s = 0; s < K && !stepped; s++) stepped = log[s * nd + d] >= 2;
        if (!stepped) continue;
        if (E->phase[d] == PH_HELD || (E->scan_on && E->need_scan[d])) now.push_back(d);
        else E->def_list.push_back(d);
    }
    // records to their disks
    for (int d = 0; d < nd; d++) E->def_rec[d].clear();
    for (int q = 0; q < ne; q++) {
        const red_event& e = E->h_events[q];
        if (E->S.clear_events_each_attempt && e.stage == 0) continue;   // f0's records, cleared with the attempt
        E->def_rec[e.pad].push_back(&e);
    }
    if (now.empty()) {
        E->event_s += seconds_since(t_ev);
        return RED_OK;
    }
    std::vector<std::vector<BodyUpdate>> ups(nd);
    std::vector<int> rcs(nd, RED_OK);
    std::vector<std::string> msgs(nd);
    std::vector<char> comp(nd, 0);
    size_t records = 0;
    for (int d : now) records += E->def_rec[d].size();
    for_disks(records >= 1024 ? E->host_threads : 1, now, [&](int d) {
        rcs[d] = disk_steps_host(E, d, E->def_rec[d], log, K, ups[d], &comp[d], msgs[d]);
    });
    std::vector<BodyUpdate> all;
    std::vector<int> merged;
    for (int d : now) {
        if (rcs[d]) return E->fail(rcs[d], "disk " + std::to_string(d) + ": " + msgs[d]);
        if (!ups[d].empty()) {
            merged.push_back(d);
            for (BodyUpdate u : ups[d]) { u.idx += E->b0[d]; all.push_back(u); }
            // the step's ejection predicate holds for the state unless a merge moved a body; a
            // rebuild only drops inactive bodies, which no scan visits
            if (E->scan_on) E->need_scan[d] = 1;
        }
        if (comp[d]) E->compact_due[d] = 1;
    }
    if (!all.empty()) {
        if ((int)all.size() > E->updates_cap) {
            SetupLock lk;
            (void)hipStreamSynchronize(E->stream);
            if (E->d_updates) (void)hipFree(E->d_updates);
            E->d_updates = nullptr;
            E->updates_cap = (int)all.size() * 2;
            ENS_HIP(E, dmalloc(&E->d_updates, E->updates_cap));
        }
        ENS_HIP(E, hipMemcpyAsync(E->d_updates, all.data(), sizeof(BodyUpdate) * all.size(), hipMemcpyHostToDevice,
                                  E->stream));
        ENS_HIP(E, launch_apply_updates(E->d_updates, (int)all.size(), E->outr, E->outv, E->d_mass, E->d_radius,
                                        E->d_id, E->stream));
        ENS_HIP(E, hipStreamSynchronize(E->stream));
        const int rc = upload_pairs(E, merged);
        if (rc) return rc;
        for (int d : merged) E->inactive[d] = count_inactive(E, d);
    }
    E->event_s += seconds_since(t_ev);
    const auto t_rb = std::chrono::steady_clock::now();
    std::vector<int> compacted;
    for (int d = 0; d < nd; d++)
        if (E->compact_due[d]) compacted.push_back(d);
    const int rc = compact_all(E);
    if (rc) return rc;
    for (int d : compacted) E->inactive[d] = count_inactive(E, d);
    E->rebuild_s += seconds_since(t_rb);
    return RED_OK;
}

// The event passes after_batch left: host records and counts only, run while the next batch
// is on the device (they read the batch's records and step log from host copies).
int after_batch_deferred(red_ensemble* E)
{
    if (E->def_list.empty()) return RED_OK;
    const auto t_ev = std::chrono::steady_clock::now();
    const int nd = E->nd;
    std::vector<std::vector<BodyUpdate>> ups(nd);
    std::vector<int> rcs(nd, RED_OK);
    std::vector<std::string> msgs(nd);
    std::vector<char> comp(nd, 0);
    size_t records = 0;
    for (int d : E->def_list) records += E->def_rec[d].size();
    for_disks(records >= 1024 ? E->host_threads : 1, E->def_list, [&](int d) {
        rcs[d] = disk_steps_host(E, d, E->def_rec[d], E->def_log.data(), E->def_k, ups[d], &comp[d], msgs[d]);
    });
    for (int d : E->def_list) {
        if (!rcs[d] && (!ups[d].empty() || comp[d])) { rcs[d] = RED_ERR_ARG; msgs[d] = "a merge or compaction left for later"; }
        if (rcs[d]) return E->fail(rcs[d], "disk " + std::to_string(d) + ": " + msgs[d]);
    }
    E->def_list.clear();
    E->event_s += seconds_since(t_ev);
    return RED_OK;
}

// The host mirror of the controller state into the pinned twin of the device block, and back.
void pack_ctl(red_ensemble* E)
{
    const int nd = E->nd;
    const EnsCtl& H = E->hc;
    for (int d = 0; d < nd; d++) {
        H.t[d] = E->t[d]; H.dt_try[d] = E->dt_try[d]; H.dt_did[d] = E->dt_did[d]; H.last_err[d] = E->last_err[d];
        H.iter[d] = E->iter[d]; H.tried[d] = E->tried[d]; H.failed[d] = E->failed[d]; H.passed[d] = E->passed[d];
        H.phase[d] = E->phase[d]; H.accepted_prev[d] = E->accepted_prev[d]; H.need_scan[d] = E->need_scan[d];
        H.hold_mode[d] = E->hold_mode[d];
    }
}

void unpack_ctl(red_ensemble* E)
{
    const int nd = E->nd;
    const EnsCtl& H = E->hc;
    for (int d = 0; d < nd; d++) {
        E->t[d] = H.t[d]; E->dt_try[d] = H.dt_try[d]; E->dt_did[d] = H.dt_did[d]; E->last_err[d] = H.last_err[d];
        E->iter[d] = H.iter[d]; E->tried[d] = H.tried[d]; E->failed[d] = H.failed[d]; E->passed[d] = H.passed[d];
        E->phase[d] = H.phase[d]; E->accepted_prev[d] = H.accepted_prev[d]; E->need_scan[d] = H.need_scan[d];
    }
}

}  // namespace

extern "C" {

int red_ensemble_destroy(red_ensemble* E)
{
    if (!E) return RED_OK;
    if (E->stream) (void)hipStreamSynchronize(E->stream);
    destroy_graph(E);
    SetupLock lk;
    for (void* p : {(void*)E->yr, (void*)E->yv, (void*)E->outr, (void*)E->outv, (void*)E->disk_of, (void*)E->err_on,
                    (void*)E->d_ctl, (void*)E->d_scr, (void*)E->d_map, (void*)E->d_ones, (void*)E->d_cscratch, (void*)E->d_rows, (void*)E->d_tiles,
                    (void*)E->d_mass,
                    (void*)E->d_radius, (void*)E->d_id, (void*)E->d_type, (void*)E->d_pair, (void*)E->d_events,
                    (void*)E->d_kept, (void*)E->d_table,
                    (void*)E->d_updates})
        (void)hipFree(p);
    for (auto p : E->sr) (void)hipFree(p);
    for (auto p : E->sv) (void)hipFree(p);
    for (auto p : E->acc) (void)hipFree(p);
    for (void* p : {(void*)E->h_ctl, (void*)E->h_flags, (void*)E->h_events, (void*)E->h_state, (void*)E->h_map})
        if (p) (void)hipHostFree(p);
    if (E->stream) (void)hipStreamDestroy(E->stream);
    delete E;
    return RED_OK;
}

const char* red_ensemble_last_error(const red_ensemble* E) { return E ? E->err.c_str() : "no ensemble"; }

int red_ensemble_create(int n_disks, const red_bodies* disks, const red_params* p, red_ensemble** out)
{
    if (!out || !disks || !p || n_disks <= 0) return RED_ERR_ARG;
    *out = nullptr;
    SetupLock lk;
    if (p->integrator < RED_EULER || p->integrator > RED_RKF78 || p->precision != RED_PREC_FP64 ||
        p->frame_center != RED_FRAME_BARY)
        return RED_ERR_UNSUPPORTED;
    if (!p->inner_steps && (p->radii_enhance_factor > 0.0 || p->encounter_hill_factor > 0.0)) return RED_ERR_UNSUPPORTED;
    red_ensemble* E = new red_ensemble();
    E->prm = *p;
    E->S = make_scheme(p->integrator, p->adaptive != 0);
    E->nd = n_disks;
    E->use_graph = p->use_graph != 0;
    E->pair_tests = p->inner_steps && (p->radii_enhance_factor > 0.0 || p->encounter_hill_factor > 0.0);
    E->scan_on = p->ejection > 0.0 || p->hit_centrum > 0.0;
    std::vector<d4> r, v;
    std::vector<unsigned char> eon;
    for (int d = 0; d < n_disks; d++) {
        const red_bodies& b = disks[d];
        int nb = 0;
        std::array<int, 7> c{};
        for (int k = 0; k < 7; k++) {
            if (b.counts[k] < 0) { delete E; return RED_ERR_ARG; }
            c[k] = b.counts[k];
            nb += b.counts[k];
        }
        if (nb > 0 && (!b.id || !b.type || !b.mass || !b.radius || !b.density || !b.cd || !b.r || !b.v)) {
            delete E;
            return RED_ERR_ARG;
        }
        const int si = c[0] + c[1] + c[2] + c[3];
        const int m = si + c[4] + c[5];
        E->b0.push_back(E->n);
        E->b1.push_back(E->n + nb);
        E->cnt.push_back(c);
        for (int i = 0; i < nb; i++) {
            r.push_back(d4{b.r[3 * i], b.r[3 * i + 1], b.r[3 * i + 2], b.mass[i]});
            v.push_back(d4{b.v[3 * i], b.v[3 * i + 1], b.v[3 * i + 2], 0.0});
            E->disk_of_h.push_back(d);
            eon.push_back((unsigned char)(p->error_check_for_tp ? 1 : (i < m ? 1 : 0)));
            E->h_mass.push_back(b.mass[i]); E->h_radius.push_back(b.radius[i]);
            E->h_density.push_back(b.density[i]); E->h_cd.push_back(b.cd[i]);
            E->h_id.push_back(b.id[i]); E->h_type.push_back(b.type[i]);
        }
        E->n += nb;
        if (nb > 0 && (p->encounter_hill_factor > 0.0 || p->close_encounter) && events::mass_of_star(E->view(d)) < 0.0) {
            delete E;
            return RED_ERR_NO_STAR;
        }
    }
    *out = E;
    if (E->n >= (1 << 29)) return E->fail(RED_ERR_ARG, "more than 2^29 - 1 bodies in an ensemble");
    ENS_HIP(E, hipSetDevice(p->device));
    ENS_HIP(E, hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
    const int n = std::max(E->n, 1);
    for (d4** q : {&E->yr, &E->yv, &E->outr, &E->outv}) ENS_HIP(E, dmalloc(q, n));
    E->sr.assign(E->S.n_stage, nullptr); E->sv.assign(E->S.n_stage, nullptr); E->acc.assign(E->S.n_stage, nullptr);
    for (int k = 0; k < E->S.n_stage; k++) {
        ENS_HIP(E, dmalloc(&E->acc[k], n));
        ENS_HIP(E, hipMemsetAsync(E->acc[k], 0, sizeof(d4) * n, E->stream));
        if (k >= 1) { ENS_HIP(E, dmalloc(&E->sr[k], n)); ENS_HIP(E, dmalloc(&E->sv[k], n)); }
    }
    ENS_HIP(E, dmalloc(&E->disk_of, n));
    ENS_HIP(E, dmalloc(&E->err_on, n));
    ENS_HIP(E, dmalloc(&E->d_mass, n));
    ENS_HIP(E, dmalloc(&E->d_radius, n));
    ENS_HIP(E, dmalloc(&E->d_id, n));
    ENS_HIP(E, dmalloc(&E->d_type, n));
    ENS_HIP(E, dmalloc(&E->d_map, n));
    E->padded = p->padded_storage != 0;
    if (E->padded) {
        std::vector<int> ones(n, 1);
        ENS_HIP(E, dmalloc(&E->d_ones, n));
        ENS_HIP(E, stream_copy(E->stream, E->d_ones, ones.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    }
    ENS_HIP(E, dmalloc(&E->d_cscratch, 3 * (size_t)n));
    ENS_HIP(E, hipHostMalloc((void**)&E->h_map, sizeof(int) * n, hipHostMallocDefault));
    for (int i = 0; i < n; i++) E->h_map[i] = -2;
    const size_t nd = (size_t)std::max(E->nd, 1);
    // the controller state (device block and its pinned twin) and the per-attempt values
    E->ctl_all = ctl_bytes(E->nd);
    ENS_HIP(E, dmalloc(&E->d_ctl, E->ctl_all));
    ENS_HIP(E, hipHostMalloc((void**)&E->h_ctl, E->ctl_all, hipHostMallocDefault));
    memset(E->h_ctl, 0, E->ctl_all);
    E->ctl_up = carve_ctl(E->d_ctl, E->nd, &E->dc);
    (void)carve_ctl(E->h_ctl, E->nd, &E->hc);
    const size_t scr_bytes = nd * (sizeof(unsigned long long) + 5 * sizeof(int));
    ENS_HIP(E, dmalloc(&E->d_scr, scr_bytes));
    E->scr.err = (unsigned long long*)E->d_scr;
    E->scr.flags = (int*)(E->scr.err + nd);
    E->scr.nan = E->scr.flags + nd; E->scr.probe = E->scr.nan + nd;
    E->scr.rec = E->scr.probe + nd; E->scr.rec_col = E->scr.rec + nd;
    E->d_t = E->dc.t; E->d_dt = E->dc.dt_did;
    E->d_flags = E->scr.flags; E->d_err = E->scr.err; E->d_nan = E->scr.nan;
    E->d_ev_count = &E->dc.w->ev_count;
    ENS_HIP(E, hipHostMalloc((void**)&E->h_flags, sizeof(int) * nd, hipHostMallocDefault));
    ENS_HIP(E, hipMemsetAsync(E->d_scr, 0, scr_bytes, E->stream));
    ENS_HIP(E, stream_copy(E->stream, E->d_ctl, E->h_ctl, E->ctl_all, hipMemcpyHostToDevice));
    ENS_HIP(E, dmalloc(&E->d_pair, E->nd));
    E->ev_cap = p->event_capacity > 0 ? p->event_capacity : std::max(4096, E->n);   // as a single disk's
    ENS_HIP(E, dmalloc(&E->d_events, E->ev_cap));
    if (E->pair_tests) {
        // a batch keeps at most half the capacity before its last attempt, which adds at most
        // the capacity; the pair table holds twice that many keys
        E->kept_cap = E->ev_cap + E->ev_cap / 2 + 1;
        ENS_HIP(E, dmalloc(&E->d_kept, E->kept_cap));
        int bits = 10;
        while ((1ull << bits) < 2ull * (unsigned long long)E->kept_cap) bits++;
        if (bits > 31) return E->fail(RED_ERR_ARG, "event capacity too large for the pair table");
        const size_t slots = (size_t)1 << bits;
        E->table_bytes = slots * (2 * sizeof(unsigned long long) + 2 * sizeof(unsigned int));
        ENS_HIP(E, dmalloc(&E->d_table, E->table_bytes));
        EnsDedup& D = E->dedup;
        D.key = (unsigned long long*)E->d_table;
        D.tmin = D.key + slots;
        D.first = (unsigned int*)(D.tmin + slots);
        D.kept = D.first + slots;
        D.bits = bits;
        D.raw = E->d_events; D.raw_cap = E->ev_cap;
        D.out = E->d_kept; D.out_cap = E->kept_cap;
        D.dp = E->d_pair;
        E->dedup_blocks = std::max(1, std::min(2048, (E->ev_cap + 255) / 256));
    }
    if (E->n > 0) {
        ENS_HIP(E, stream_copy(E->stream, E->yr, r.data(), sizeof(d4) * E->n, hipMemcpyHostToDevice));
        ENS_HIP(E, stream_copy(E->stream, E->yv, v.data(), sizeof(d4) * E->n, hipMemcpyHostToDevice));
        ENS_HIP(E, stream_copy(E->stream, E->outr, r.data(), sizeof(d4) * E->n, hipMemcpyHostToDevice));
        ENS_HIP(E, stream_copy(E->stream, E->outv, v.data(), sizeof(d4) * E->n, hipMemcpyHostToDevice));
        ENS_HIP(E, stream_copy(E->stream, E->disk_of, E->disk_of_h.data(), sizeof(int) * E->n, hipMemcpyHostToDevice));
        ENS_HIP(E, stream_copy(E->stream, E->err_on, eon.data(), E->n, hipMemcpyHostToDevice));
        E->h_err_on = eon;
        ENS_HIP(E, stream_copy(E->stream, E->d_mass, E->h_mass.data(), sizeof(double) * E->n, hipMemcpyHostToDevice));
        ENS_HIP(E, stream_copy(E->stream, E->d_radius, E->h_radius.data(), sizeof(double) * E->n, hipMemcpyHostToDevice));
        ENS_HIP(E, stream_copy(E->stream, E->d_id, E->h_id.data(), sizeof(int) * E->n, hipMemcpyHostToDevice));
        ENS_HIP(E, stream_copy(E->stream, E->d_type, E->h_type.data(), sizeof(int) * E->n, hipMemcpyHostToDevice));
    }
    const double dt0 = p->dt_day * tools::GAUSS;   // integrator.cu:29
    E->t.assign(E->nd, 0.0); E->dt_try.assign(E->nd, dt0); E->dt_did.assign(E->nd, 0.0);
    E->last_err.assign(E->nd, 0.0);
    E->iter.assign(E->nd, 0); E->tried.assign(E->nd, 0); E->failed.assign(E->nd, 0); E->passed.assign(E->nd, 0);
    E->phase.assign(E->nd, PH_NEW); E->accepted_prev.assign(E->nd, 0);
    E->need_scan.assign(E->nd, 1);
    E->hold_mode.assign(E->nd, 0);
    E->inactive.assign(E->nd, 0);
    for (int d = 0; d < E->nd; d++) E->inactive[d] = count_inactive(E, d);
    E->evc.assign(E->nd, events::Counters());
    E->pending.assign(E->nd, {});
    E->def_rec.assign(E->nd, {});
    E->pending_attempt.assign(E->nd, 0);
    E->log.assign(E->nd, {});
    E->compact_due.assign(E->nd, 0);
    if (E->h_err_on.empty()) E->h_err_on.assign(n, 0);
    {
        // the stage combinations' arguments, read by the force launches' epilogues
        std::vector<StageArgs> rows(std::max(E->S.n_stage, 1));
        memset(rows.data(), 0, sizeof(StageArgs) * rows.size());
        rows[0] = ens_args(E, E->S.np1, E->outr, E->outv);
        for (int r = 1; r < (int)E->S.rows.size() && r < E->S.n_stage; r++)
            rows[r] = ens_args(E, E->S.rows[r], E->sr[r], E->sv[r]);
        ENS_HIP(E, dmalloc(&E->d_rows, rows.size()));
        ENS_HIP(E, stream_copy(E->stream, E->d_rows, rows.data(), sizeof(StageArgs) * rows.size(),
                               hipMemcpyHostToDevice));
    }
    E->h_pair.assign(E->nd, EnsDiskPair{});
    std::vector<int> all_disks(E->nd);
    for (int d = 0; d < E->nd; d++) all_disks[d] = d;
    int rc0 = upload_pairs(E, all_disks);
    if (rc0) return rc0;
    E->host_threads = p->host_threads > 0 ? p->host_threads
                                          : std::max(1, std::min(8, (int)std::thread::hardware_concurrency()));
    if (p->block >= 64) E->block = std::min(256, p->block / 64 * 64);
    int rc = upload_tiles(E);
    if (rc) return rc;
    ENS_HIP(E, hipDeviceSynchronize());
    return RED_OK;
}

int red_ensemble_set_time(red_ensemble* E, double t)
{
    if (!E) return RED_ERR_ARG;
    for (int d = 0; d < E->nd; d++) E->t[d] = t;
    return RED_OK;
}

// Integrates every disk with the reference loop `while (t <= t_stop) { ejection check; step;
// collision check; rebuild check }` (main.cu:176-255) until all disks are done or
// max_attempts supersteps have run.  The device runs batches of supersteps with each disk's
// controller (k_ens_begin / k_ens_control); between batches the host runs what the device
// held a disk for (ejection scans, merges, compactions) and every step's event passes.
int red_ensemble_advance(red_ensemble* E, double t_stop, int max_attempts, red_ensemble_info* info)
{
    if (!E) return RED_ERR_ARG;
    int attempts = 0, batches = 0, full_batches = 0;
    const auto t_begin = std::chrono::steady_clock::now();
    E->event_s = 0.0;
    E->rebuild_s = 0.0;
    const int nd = E->nd;
    {
        // passes a failed call left
        const int rc = after_batch_deferred(E);
        if (rc) return rc;
    }
    // a disk done against an earlier t_stop continues (a run to t1, then to t2, steps as one
    // run to t2); a disk left in a retry by max_attempts resumes it
    for (int d = 0; d < nd; d++)
        if (E->phase[d] == PH_DONE) E->phase[d] = PH_NEW;
    for (;;) {
        // the loop iteration's ejection / hit-centrum check, before the step, for the disks at
        // a step boundary that take a step; many disks' scans read one host copy of the state
        int scans = 0;
        for (int d = 0; d < nd; d++) {
            const bool boundary = E->phase[d] == PH_NEW || E->phase[d] == PH_HELD;
            scans += boundary && E->t[d] <= t_stop && E->scan_on && E->need_scan[d];
        }
        if (scans > 0) {
            const auto t_scan = std::chrono::steady_clock::now();
            const bool bulk = scans >= 8;
            if (bulk) {
                const int rc = fetch_states(E);
                if (rc) return rc;
            }
            for (int d = 0; d < nd; d++) {
                const bool boundary = E->phase[d] == PH_NEW || E->phase[d] == PH_HELD;
                if (!(boundary && E->t[d] <= t_stop && E->scan_on && E->need_scan[d])) continue;
                const int rc = ejection_check(E, d, bulk);
                if (rc) return rc;
                E->need_scan[d] = 0;
                E->inactive[d] = count_inactive(E, d);
            }
            E->event_s += seconds_since(t_scan);
        }
        int active = 0;
        for (int d = 0; d < nd; d++) {
            if (E->phase[d] == PH_HELD) E->phase[d] = PH_NEW;
            active += (E->phase[d] == PH_NEW && E->t[d] <= t_stop) || E->phase[d] == PH_RETRY;
            E->hold_mode[d] = hold_mode_of(E, d);
        }
        if (active == 0 || attempts >= max_attempts) break;
        pack_ctl(E);
        EnsWords& w = *E->hc.w;
        w.t_stop = t_stop;
        w.k_limit = std::min(E->batch, max_attempts - attempts);
        w.stop = 0;
        w.nan_disk = nd;
        w.limit_disk = nd;
        w.ev_count = 0;
        w.raw_max = 0;
        w.kept_count = 0;
        w.lost = 0;
        ENS_HIP(E, hipMemcpyAsync(E->d_ctl, E->h_ctl, E->ctl_up, hipMemcpyHostToDevice, E->stream));
        if (E->pair_tests) ENS_HIP(E, hipMemsetAsync(E->d_table, 0xff, E->table_bytes, E->stream));
        ENS_HIP(E, run_batch(E));
        ENS_HIP(E, hipMemcpyAsync(E->h_ctl, E->d_ctl, E->ctl_all, hipMemcpyDeviceToHost, E->stream));
        {
            // the previous batch's remaining event passes, while this batch runs
            const int rc = after_batch_deferred(E);
            if (rc) return rc;
        }
        ENS_HIP(E, hipStreamSynchronize(E->stream));
        unpack_ctl(E);
        batches++;
        full_batches += w.stop != 0;
        for (int s = 0; s < E->batch; s++) {
            bool any = false;
            for (int d = 0; d < nd && !any; d++) any = E->hc.steplog[s * nd + d] != 0;
            attempts += any;
        }
        if (w.nan_disk < nd) return E->fail(RED_ERR_NAN, "non-finite error estimate in disk " + std::to_string(w.nan_disk));
        if (w.limit_disk < nd) return E->fail(RED_ERR_STEP_LIMIT, "disk " + std::to_string(w.limit_disk) + " did not converge");
        // the batch's kept pair records, to their disks
        if (E->pair_tests && w.raw_max > E->ev_cap)
            return E->fail(RED_ERR_EVENT_OVERFLOW, "pair events exceed the event capacity (" +
                                                       std::to_string(w.raw_max) + " > " + std::to_string(E->ev_cap) + ")");
        if (E->pair_tests && w.lost > 0)
            return E->fail(RED_ERR_EVENT_OVERFLOW, std::to_string(w.lost) + " pair records found no room in the batch's list");
        const int ne = E->pair_tests ? w.kept_count : 0;
        if (ne > E->h_events_cap) {
            SetupLock lk;
            if (E->h_events) (void)hipHostFree(E->h_events);
            E->h_events = nullptr;
            E->h_events_cap = std::max(ne, 2 * E->h_events_cap);
            ENS_HIP(E, hipHostMalloc((void**)&E->h_events, sizeof(red_event) * E->h_events_cap, hipHostMallocDefault));
        }
        if (ne > 0)
            ENS_HIP(E, stream_copy(E->stream, E->h_events, E->d_kept, sizeof(red_event) * ne, hipMemcpyDeviceToHost));
        const int rc = after_batch(E, ne);
        if (rc) return rc;
    }
    {
        const int rc = after_batch_deferred(E);
        if (rc) return rc;
    }
    // apply the last accepted results so that the state buffers hold y
    bool pending = false;
    for (int d = 0; d < nd; d++) { E->h_flags[d] = E->accepted_prev[d] ? 4 : 0; pending |= E->accepted_prev[d] != 0; }
    if (pending) {
        ENS_HIP(E, hipMemcpyAsync(E->d_flags, E->h_flags, sizeof(int) * nd, hipMemcpyHostToDevice, E->stream));
        hipLaunchKernelGGL(k_ens_accept, dim3((E->n + 255) / 256), dim3(256), 0, E->stream, E->yr, E->yv, E->outr,
                           E->outv, E->disk_of, E->d_flags, E->n);
        ENS_HIP(E, hipGetLastError());
        ENS_HIP(E, hipStreamSynchronize(E->stream));
        for (int d = 0; d < nd; d++) E->accepted_prev[d] = 0;
    }
    if (info) {
        info->supersteps = attempts;
        int still = 0;
        for (int d = 0; d < nd; d++) still += E->phase[d] == PH_RETRY || (E->phase[d] == PH_NEW && E->t[d] <= t_stop);
        info->active_disks = still;
        info->seconds = seconds_since(t_begin);
        info->event_seconds = E->event_s;
        info->rebuild_seconds = E->rebuild_s;
        info->batches = batches;
        info->full_batches = full_batches;
    }
    return RED_OK;
}

int red_step_pow(const double* x, const double* y, double* out, int n, int on_device, int device)
{
    if (n < 0 || (n > 0 && (!x || !y || !out))) return RED_ERR_ARG;
    if (!on_device) {
        for (int i = 0; i < n; i++) out[i] = step_pow(x[i], y[i]);
        return RED_OK;
    }
    if (n == 0) return RED_OK;
    if (hipSetDevice(device) != hipSuccess) return RED_ERR_HIP;
    double* buf = nullptr;
    const size_t b = sizeof(double) * (size_t)n;
    if (hipMalloc((void**)&buf, 3 * b) != hipSuccess) return RED_ERR_HIP;
    hipError_t e = hipMemcpy(buf, x, b, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(buf + n, y, b, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_step_pow, dim3((n + 255) / 256), dim3(256), 0, 0, buf, buf + n, buf + 2 * (size_t)n, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, buf + 2 * (size_t)n, b, hipMemcpyDeviceToHost);
    (void)hipFree(buf);
    return e == hipSuccess ? RED_OK : RED_ERR_HIP;
}

// Supersteps per graph launch (1..64): each batch is one captured graph of that many
// supersteps and one control transfer each way.
int red_ensemble_set_batch(red_ensemble* E, int supersteps)
{
    if (!E || supersteps < 1 || supersteps > MAX_BATCH) return RED_ERR_ARG;
    if (supersteps != E->batch) destroy_graph(E);
    E->batch = supersteps;
    return RED_OK;
}

int red_ensemble_get_state(red_ensemble* E, int disk, double* r, double* v)
{
    if (!E || disk < 0 || disk >= E->nd) return RED_ERR_ARG;
    const int b0 = E->b0[disk], nb = E->b1[disk] - b0;
    std::vector<d4> R(std::max(nb, 1)), V(std::max(nb, 1));
    ENS_HIP(E, hipStreamSynchronize(E->stream));
    if (nb > 0) {
        ENS_HIP(E, stream_copy(E->stream, R.data(), cur_r(E, disk) + b0, sizeof(d4) * nb, hipMemcpyDeviceToHost));
        ENS_HIP(E, stream_copy(E->stream, V.data(), cur_v(E, disk) + b0, sizeof(d4) * nb, hipMemcpyDeviceToHost));
    }
    for (int i = 0; i < nb; i++) {
        if (r) { r[3 * i] = R[i].x; r[3 * i + 1] = R[i].y; r[3 * i + 2] = R[i].z; }
        if (v) { v[3 * i] = V[i].x; v[3 * i + 1] = V[i].y; v[3 * i + 2] = V[i].z; }
    }
    return RED_OK;
}

int red_ensemble_get_bodies(red_ensemble* E, int disk, int counts[7], double* mass, double* radius, double* density,
                            double* cd, int* id, int* type)
{
    if (!E || disk < 0 || disk >= E->nd) return RED_ERR_ARG;
    if (counts) for (int k = 0; k < 7; k++) counts[k] = E->cnt[disk][k];
    const int o = E->b0[disk], nb = E->b1[disk] - o;
    for (int i = 0; i < nb; i++) {
        if (mass) mass[i] = E->h_mass[o + i];
        if (radius) radius[i] = E->h_radius[o + i];
        if (density) density[i] = E->h_density[o + i];
        if (cd) cd[i] = E->h_cd[o + i];
        if (id) id[i] = E->h_id[o + i];
        if (type) type[i] = E->h_type[o + i];
    }
    return RED_OK;
}

// out[3] = tried failed passed; t, dt_try of the disk
int red_ensemble_disk_info(red_ensemble* E, int disk, int out[3], double* t, double* dt_try)
{
    if (!E || disk < 0 || disk >= E->nd) return RED_ERR_ARG;
    if (out) { out[0] = E->tried[disk]; out[1] = E->failed[disk]; out[2] = E->passed[disk]; }
    if (t) *t = E->t[disk];
    if (dt_try) *dt_try = E->dt_try[disk];
    return RED_OK;
}

int red_ensemble_disk_counters(red_ensemble* E, int disk, int out[8])
{
    if (!E || !out || disk < 0 || disk >= E->nd) return RED_ERR_ARG;
    const events::Counters& C = E->evc[disk];
    out[0] = E->tried[disk]; out[1] = E->failed[disk]; out[2] = E->passed[disk];
    out[3] = C.collision[events::CNT_TOTAL]; out[4] = C.ejection[events::CNT_TOTAL];
    out[5] = C.hit_centrum[events::CNT_TOTAL]; out[6] = C.close_encounter[events::CNT_TOTAL];
    out[7] = C.event[events::CNT_LAST_CLEAR];
    return RED_OK;
}

int red_ensemble_disk_events(red_ensemble* E, int disk, red_event* buf, int cap, int* n)
{
    if (!E || disk < 0 || disk >= E->nd || cap < 0 || (cap > 0 && !buf)) return RED_ERR_ARG;
    std::vector<red_event>& L = E->log[disk];
    const int k = std::min((int)L.size(), cap);
    for (int i = 0; i < k; i++) buf[i] = L[i];
    if (n) *n = (int)L.size();
    L.erase(L.begin(), L.begin() + k);
    return RED_OK;
}

// Times `reps` attempts' force launches of all disks for every (block, jtile) in
// {64,128,256}^2 on the current state, keeps the fastest and returns the 9 timings (ms per
// attempt's launches, row-major block x jtile).
int red_ensemble_autotune(red_ensemble* E, int reps, int* best_block, int* best_jtile, double ms[9])
{
    if (!E || reps <= 0) return RED_ERR_ARG;
    const int choices[3] = {64, 128, 256};
    {
        // a pending y_(n+1) (left by a failed advance) moves into y first: the timed
        // launches' epilogues write the y_(n+1) buffers
        bool pending = false;
        for (int d = 0; d < E->nd; d++) { E->h_flags[d] = E->accepted_prev[d] ? 4 : 0; pending |= E->accepted_prev[d] != 0; }
        if (pending) {
            ENS_HIP(E, hipMemcpyAsync(E->d_flags, E->h_flags, sizeof(int) * E->nd, hipMemcpyHostToDevice, E->stream));
            hipLaunchKernelGGL(k_ens_accept, dim3((E->n + 255) / 256), dim3(256), 0, E->stream, E->yr, E->yv, E->outr,
                               E->outv, E->disk_of, E->d_flags, E->n);
            ENS_HIP(E, hipGetLastError());
            ENS_HIP(E, hipStreamSynchronize(E->stream));
            for (int d = 0; d < E->nd; d++) E->accepted_prev[d] = 0;
        }
    }
    // the timed attempts cover the quarter of the disks with the most bodies (at least one):
    // in a run only some disks step in a superstep, and its time is set by the largest of
    // them, whose star and giant sums are the launches' serial chains (with every disk
    // stepping at once, as the first supersteps do, the choice favours throughput instead:
    // on the 512-disk workload that picked 128 threads per block, 0.69 ms per superstep
    // against 0.62 ms with 256)
    std::vector<int> order(E->nd);
    for (int d = 0; d < E->nd; d++) order[d] = d;
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return E->b1[x] - E->b0[x] > E->b1[y] - E->b0[y]; });
    std::vector<int> flags(E->nd, 0);
    for (int k = 0; k < std::max(1, E->nd / 4); k++) flags[order[k]] = 3;
    ENS_HIP(E, stream_copy(E->stream, E->d_flags, flags.data(), sizeof(int) * E->nd, hipMemcpyHostToDevice));
    hipEvent_t a, b;
    ENS_HIP(E, hipEventCreate(&a));
    ENS_HIP(E, hipEventCreate(&b));
    double best = 1e300;
    int bb = E->block, bj = E->jtile;
    for (int x = 0; x < 3; x++) {
        E->block = choices[x];
        int rc = upload_tiles(E);
        if (rc) return rc;
        for (int y = 0; y < 3; y++) {
            E->jtile = choices[y];
            // an attempt's launch sequence with its epilogues, every stage with its own
            // pair-test rule, on the current state: stage 0 recomputes f0 of that state
            // (identical values, so a retry that reuses f0 is unaffected); the other stages'
            // accelerations and the stage states, y_(n+1) and ejection flags the epilogues
            // form are recomputed by the next attempt before it reads them (no y_(n+1) is
            // pending here); the pair records are dropped when the next batch starts
            auto attempt = [&]() -> hipError_t {
                for (int r = 0; r < E->S.n_stage; r++) RED_HIP_TRY(ens_force(E, r, E->yr, E->yv, 0, true));
                return hipSuccess;
            };
            ENS_HIP(E, attempt());   // warm-up
            ENS_HIP(E, hipEventRecord(a, E->stream));
            for (int k = 0; k < reps; k++) ENS_HIP(E, attempt());
            ENS_HIP(E, hipEventRecord(b, E->stream));
            ENS_HIP(E, hipEventSynchronize(b));
            float t = 0.0f;
            ENS_HIP(E, hipEventElapsedTime(&t, a, b));
            const double per = t / reps;
            if (ms) ms[3 * x + y] = per;
            if (per < best) { best = per; bb = choices[x]; bj = choices[y]; }
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    E->block = bb;
    E->jtile = bj;
    int rc = upload_tiles(E);
    if (rc) return rc;
    destroy_graph(E);   // captured launches hold the launch shape and the source tile
    if (best_block) *best_block = bb;
    if (best_jtile) *best_jtile = bj;
    return RED_OK;
}

}  // extern "C"
