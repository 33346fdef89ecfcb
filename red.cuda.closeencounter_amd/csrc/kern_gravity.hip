// kern_gravity.hip -- all-pairs gravitational acceleration with fused pair tests (gfx950).
//
// Replaces kernel_pp_disk::calc_grav_accel and its padded twin
// (red.cuda/src/red.cuda/pp_disk.cu:144-296) and the launcher
// pp_disk::call_kernel_calc_grav_accel (pp_disk.cu:609-681).
//
// Per sweep (SI x massive, NSI x SI, NI x massive) one of three paths runs.  Sweeps whose
// sums may be reordered (mixed mode, or FP64 with fp64_split) are also cut into source
// ranges ("parts", blockIdx.y) so that even 65536 sinks fill 256 CUs with about eight
// wavefronts per SIMD; the parts are added in order by k_reduce_parts.
//
//  * broadcast sweep: a work-item owns SPT sinks, the j loop is wave-uniform so every
//    source record {x, y, z, m} arrives through one 32-byte scalar load into SGPRs and
//    feeds all 64 lanes as a scalar operand -- no LDS traffic and no VGPRs for sources.
//    The mixed sweep's four wavefronts of a work-group share their sinks and split the
//    sources of the group's part, adding their sums in LDS (one partial per work-group).
//  * LDS-tile sweep: the classic i-block x j-tile scheme, sources staged in LDS per tile
//    (kept as an autotuner alternative).
//  * split sweep for sweeps with few sinks (the star and giants against the whole disk):
//    pair terms for all (sink, source) pairs in parallel, then per-sink sums.  In FP64
//    mode the sum runs sequentially in ascending j, so the result is bit-identical to
//    the reference loop; in mixed mode it is chunked (deterministic, not sequential).
//
// FP64 mode keeps the reference's operation order with no fused multiply-adds
// (build with -ffp-contract=off): dx = xj - xi; r2 = (dx*dx + dy*dy) + dz*dz;
// d = sqrt(r2); s = m / (d*r2); a += s*dx (pp_disk.cu:247-260).  sqrt and '/' are
// correctly rounded on gfx950, so the accelerations equal the CPU oracle's bit for bit.
//
// Mixed mode: FP64 position differences, FP32 r^-3 via v_rsq_f32, FP32 partial sums per
// source chunk folded into FP64 accumulators with Kahan compensation (optionally a
// Kahan-compensated FP32 sum per pair).  Sources [0, n_fp64_src) (the star and giant
// planets) stay in the FP64 path.  Pair tests stay exact: a conservative FP32 screen on
// r^2 selects candidates and the reference predicate d < f (R_i + R_j) is re-evaluated
// in FP64 for them, so event lists equal the FP64 path's for a given state.
#include <algorithm>

#include "fp64_core.hpp"
#include "red_internal.hpp"
#include "stage_forms.hpp"

namespace red {
namespace {

constexpr int CHUNK = 64;   // mixed sweep: sources per FP32 partial sum



// Appends one pair event (pp_disk.cu:32-67, 265-291).
__device__ void append_pair_event(const PairTest& P, const EventSink& E, Ctl* ctl, int i, int j, double d,
                                  bool collide, const d4* r, const d4* v, const double* mass, const int* id)
{
    int surv = i, merg = j;
    int name;
    if (collide && !P.close_encounter) {
        if (mass[merg] > mass[surv]) { int t = surv; surv = merg; merg = t; }
        name = RED_EV_COLLISION;
    } else {
        name = RED_EV_CLOSE_ENCOUNTER;
    }
    int k = atomicAdd(&ctl->n_events, 1);
    if (k >= E.cap) return;   // counted but dropped: the host reports an overflow
    red_event* e = E.buf + k;
    e->name = name;
    e->id1 = id[surv]; e->idx1 = surv;
    e->id2 = id[merg]; e->idx2 = merg;
    e->attempt = ctl->attempt; e->stage = E.stage;
    e->sink = i; e->source = j; e->pad = 0;
    e->t = ctl->t + E.c_stage * ctl->dt;
    e->d = d;
    for (int c = 0; c < 4; c++) { e->p1[c] = 0.0; e->p2[c] = 0.0; e->ps[c] = 0.0; }
    put3(e->r1, r[surv]); put3(e->v1, v[surv]);
    put3(e->r2, r[merg]); put3(e->v2, v[merg]);
    if (name == RED_EV_CLOSE_ENCOUNTER) {
        // the host derives astro-centric elements from the star's stage state
        put3(e->rs, r[0]); put3(e->vs, v[0]);
    } else {
        for (int c = 0; c < 3; c++) { e->rs[c] = 0.0; e->vs[c] = 0.0; }
    }
    for (int c = 0; c < 6; c++) { e->oe1[c] = 0.0; e->oe2[c] = 0.0; e->oes[c] = 0.0; }
}

// The reference pair predicate (pp_disk.cu:265) plus the optional Hill criterion.
__device__ __forceinline__ void pair_test(const PairTest& P, const EventSink& E, Ctl* ctl, int i, int j, double d,
                                          double Ri, double Rj, const d4* r, const d4* v, const double* mass,
                                          const int* id)
{
    if (!(0 < i && i > j)) return;
    bool collide = d < P.radii_f * (Ri + Rj);
    bool hill = false;
    if (!collide && P.hill_f > 0.0 && j >= P.n_star) {   // both bodies orbit the star
        const d4 qi = r[i], qj = r[j];
        double ri = sqrt((qi.x * qi.x) + (qi.y * qi.y) + (qi.z * qi.z));
        double rj = sqrt((qj.x * qj.x) + (qj.y * qj.y) + (qj.z * qj.z));
        double rh = cbrt_ieee((mass[i] + mass[j]) / P.three_mstar) * ((ri + rj) * 0.5);
        hill = d < P.hill_f * rh;
    }
    if (collide || hill) append_pair_event(P, E, ctl, i, j, d, collide, r, v, mass, id);
}


// Distance below which pair_test can report an event for sink i, so that farther pairs skip
// it without changing any result: the collision reach with the largest body radius, and the
// Hill reach with the largest non-star mass, using |r_j| <= |r_i| + d (which turns
// d < f c (|r_i| + |r_j|) / 2 into d < f c |r_i| / (1 - f c / 2)); 0.1 % margin for rounding.
__device__ __forceinline__ double pair_screen_hill(const PairTest& P, double mi, const d4& qi)
{
    if (!(P.hill_f > 0.0)) return 0.0;
    const double fc = P.hill_f * cbrt((mi + P.mmax_ns) / P.three_mstar);
    const double hr = sqrt((qi.x * qi.x) + (qi.y * qi.y) + (qi.z * qi.z));
    return fc < 1.0 ? fc * hr / (1.0 - 0.5 * fc) : 1e300;
}

__device__ __forceinline__ double pair_screen(const PairTest& P, double Ri, double mi, const d4& qi)
{
    return fmax(P.radii_f * (Ri + P.rmax_all), pair_screen_hill(P, mi, qi)) * 1.001;
}

// ---------------------------------------------------------------------------------------
// Source-range partition: part p of `parts` covers [lo, hi) of [j0, j1), slices rounded
// to `align` sources.
__device__ __forceinline__ void part_range(int j0, int j1, int parts, int p, int align, int& lo, int& hi)
{
    const int n = j1 - j0;
    int len = (n + parts - 1) / parts;
    len = (len + align - 1) / align * align;
    lo = min(j1, j0 + p * len);
    hi = min(j1, lo + len);
}

// Writes a sink's sum: straight to a (masked by activity) or to its part slot.  Part sums
// are three planes (x, y, z) of FP64 per part, 24 bytes per sink and part.
__device__ __forceinline__ void store_acc(d4* a, d4* partial, int parts, int p, const Sweep& sw, int i, bool live,
                                          double x, double y, double z)
{
    if (parts == 1) {
        a[i] = d4{live ? x : 0.0, live ? y : 0.0, live ? z : 0.0, 0.0};
    } else {
        const size_t nsink = (size_t)(sw.sink1 - sw.sink0);
        double* plane = reinterpret_cast<double*>(partial) + 3 * (size_t)p * nsink + (i - sw.sink0);
        plane[0] = x;
        plane[nsink] = y;
        plane[2 * nsink] = z;
    }
}

// One FP64 pair in the reference's operation order.  Inactive sources (id < 0) and
// i == j are skipped in the reference (pp_disk.cu:242); here their term is replaced by
// +0, which leaves the sum bit-identical: the accumulator starts at +0 and
// x + (+-0) == x for every other value it can hold.
template <bool OVERLAP, bool EV>
__device__ __forceinline__ void fp64_pair(const d4& q, int j, bool act, int i, double xi, double yi, double zi,
                                          double Ri, bool live, double& ax, double& ay, double& az,
                                          const PairTest& P, const EventSink& E, Ctl* ctl, const d4* r, const d4* v,
                                          const double* mass, const double* radius, const int* id)
{
    const double dx = q.x - xi;
    const double dy = q.y - yi;
    const double dz = q.z - zi;
    const double r2 = (dx * dx + dy * dy) + dz * dz;
    const double d = sqrt(r2);
    const bool use = OVERLAP ? (act && j != i) : act;
    const double sc = use ? q.w / (d * r2) : 0.0;
    ax += sc * dx;
    ay += sc * dy;
    az += sc * dz;
    if (EV && use && live) pair_test(P, E, ctl, i, j, d, Ri, radius[j], r, v, mass, id);
}

// FP64 sweep.  LDS = false: every source record reaches the wavefront through a scalar
// load (wave-uniform j); LDS = true: sources are staged in LDS tiles of TILE bodies.
template <int SPT, bool OVERLAP, bool EV, bool LDS>
__global__ __launch_bounds__(256) void k_grav_fp64(const d4* __restrict__ r, const d4* __restrict__ v,
                                                   d4* __restrict__ a, d4* __restrict__ partial,
                                                   const double* __restrict__ mass,
                                                   const double* __restrict__ radius, const int* __restrict__ id,
                                                   Sweep sw, int parts, PairTest P, EventSink E,
                                                   Ctl* __restrict__ ctl)
{
    constexpr int TILE = 256;
    __shared__ d4 tile[LDS ? TILE : 1];
    __shared__ int tact[LDS ? TILE : 1];
    const int nt = blockDim.x;
    const int p = blockIdx.y;
    const int base = sw.sink0 + blockIdx.x * nt * SPT + threadIdx.x;
    int jlo, jhi;
    part_range(sw.src0, sw.src1, parts, p, 1, jlo, jhi);
    int ii[SPT];
    bool live[SPT];
    double xi[SPT], yi[SPT], zi[SPT], Ri[SPT], ax[SPT], ay[SPT], az[SPT];
#pragma unroll
    for (int s = 0; s < SPT; s++) {
        ii[s] = base + s * nt;
        const bool in = ii[s] < sw.sink1;
        const int k = in ? ii[s] : sw.sink0;
        const d4 q = r[k];
        xi[s] = q.x; yi[s] = q.y; zi[s] = q.z;
        live[s] = in && sw.gate[k] > 0;
        Ri[s] = EV ? radius[k] : 0.0;
        ax[s] = ay[s] = az[s] = 0.0;
    }
    if (!LDS) {
        // Sources in groups of G: the next group's records are requested before the current
        // group's G independent pair terms are computed, so load latency and the long FP64
        // sqrt/divide sequences overlap; the sums still take j in ascending order.
        constexpr int G = 4;
        int j = jlo;
        d4 qn[G];
        bool an[G];
        if (j + G <= jhi) {
#pragma unroll
            for (int k = 0; k < G; ++k) { qn[k] = r[j + k]; an[k] = sw.gate[j + k] >= 0; }
        }
        for (; j + G <= jhi; j += G) {
            d4 q[G];
            bool act[G];
#pragma unroll
            for (int k = 0; k < G; ++k) { q[k] = qn[k]; act[k] = an[k]; }
            if (j + 2 * G <= jhi) {
#pragma unroll
                for (int k = 0; k < G; ++k) { qn[k] = r[j + G + k]; an[k] = sw.gate[j + G + k] >= 0; }
            }
#pragma unroll
            for (int k = 0; k < G; ++k)
#pragma unroll
                for (int s = 0; s < SPT; s++)
                    fp64_pair<OVERLAP, EV>(q[k], j + k, act[k], ii[s], xi[s], yi[s], zi[s], Ri[s], live[s], ax[s],
                                           ay[s], az[s], P, E, ctl, r, v, mass, radius, id);
        }
        for (; j < jhi; ++j) {
            const bool act = sw.gate[j] >= 0;
            const d4 q = r[j];
#pragma unroll
            for (int s = 0; s < SPT; s++)
                fp64_pair<OVERLAP, EV>(q, j, act, ii[s], xi[s], yi[s], zi[s], Ri[s], live[s], ax[s], ay[s], az[s], P,
                                       E, ctl, r, v, mass, radius, id);
        }
    } else {
        for (int j0 = jlo; j0 < jhi; j0 += TILE) {
            const int nj = min(TILE, jhi - j0);
            __syncthreads();
            for (int k = threadIdx.x; k < nj; k += nt) {
                tile[k] = r[j0 + k];
                tact[k] = sw.gate[j0 + k] >= 0;
            }
            __syncthreads();
#pragma unroll 2
            for (int k = 0; k < nj; ++k) {
                const d4 q = tile[k];
                const bool act = tact[k] != 0;
#pragma unroll
                for (int s = 0; s < SPT; s++)
                    fp64_pair<OVERLAP, EV>(q, j0 + k, act, ii[s], xi[s], yi[s], zi[s], Ri[s], live[s], ax[s], ay[s],
                                           az[s], P, E, ctl, r, v, mass, radius, id);
            }
        }
    }
#pragma unroll
    for (int s = 0; s < SPT; s++)
        if (ii[s] < sw.sink1) store_acc(a, partial, parts, p, sw, ii[s], live[s], ax[s], ay[s], az[s]);
}

// ---------------------------------------------------------------------------------------
// FP64 sweep with sequential sums, producer/consumer through LDS.  The reference adds the
// source terms of a sink one after another in ascending j (pp_disk.cu:144-215); only those
// additions are a serial chain -- the pair terms themselves (differences, IEEE sqrt and
// divide) are independent.  A block of 1 + NP waves takes SG sinks: NP producer waves
// compute the terms of the next J = TILE / SG sources into one half of a double-buffered
// LDS tile while the consumer wave adds the previous chunk in ascending j, one lane per
// (sink, component) row, starting from 0.  Every term is formed and added exactly as in the
// fused loop, so the sums are bit-identical to the reference order, while the long FP64
// sequences run on all SIMDs.  A producer lane keeps one sink in registers and takes
// M sources of every chunk (independent work for the scheduler to interleave); the chunk
// holds exactly M x SLOTS sources, so no producer lane idles on a full chunk, and the
// source records of the chunk after next are requested before a chunk is computed.  Longer
// chunks (fewer sinks per block) spread the per-chunk barrier and producer latency over
// more sources; the tile stays within 2 x TILE x 3 doubles, TILE = 1024, or 2048 for sweeps
// of at most 256 blocks of 4 sinks (one block per CU, so LDS does not limit occupancy).
constexpr int SQ_TILE = 1024;         // sinks x sources per chunk, at most
constexpr int SQ_TILE_FEW = 2048;     // the same for sweeps of at most one block per CU
constexpr int BUF_RSRC_DW3 = 0x00020000;   // raw buffer descriptor word 3 on CDNA (32-bit data format)
constexpr long SEQ_MAX_SOURCES = (1L << 31) / 32;   // buffer byte offsets stay below 2^31

template <bool OVERLAP, bool EV, int NP, int SQ_SG, int TILE_SQ = SQ_TILE>
__global__ __launch_bounds__(64 * (NP + 1)) void k_grav_fp64_seq(const d4* __restrict__ r, const d4* __restrict__ v,
                                                                 d4* __restrict__ a, const double* __restrict__ mass,
                                                                 const double* __restrict__ radius,
                                                                 const int* __restrict__ id, Sweep sw, PairTest P,
                                                                 EventSink E, Ctl* __restrict__ ctl)
{
    constexpr int SQ_ROWS = 3 * SQ_SG;                // consumer lanes in use
    constexpr int SLOTS = NP * 64 / SQ_SG;            // source slots per pass over the chunk
    constexpr int M = TILE_SQ / SQ_SG / SLOTS > 0 ? TILE_SQ / SQ_SG / SLOTS : 1;   // sources per lane and chunk
    constexpr int SQ_J = M * SLOTS;                   // sources per chunk
    static_assert(SQ_J * SQ_SG <= TILE_SQ, "tile");
    static_assert(SQ_ROWS <= 64, "one consumer wave");
    __shared__ double buf[2][SQ_J][SQ_ROWS];
    const int wv = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int i0 = sw.sink0 + blockIdx.x * SQ_SG;
    const int ns = min(SQ_SG, sw.sink1 - i0);
    const int nsrc = sw.src1 - sw.src0;
    const int nch = (nsrc + SQ_J - 1) / SQ_J;
    double sum = 0.0;

    if (wv > 0) {
        // producer lane: sink sk, source slots slot + m * SLOTS of every chunk
        const int t = (wv - 1) * 64 + lane;
        const int sk = t % SQ_SG;
        const int slot = t / SQ_SG;
        const bool sink_in = sk < ns;
        const int i = i0 + (sink_in ? sk : 0);
        const d4 p = r[i];
        const bool live = sink_in && sw.gate[i] > 0;
        const double Ri = EV ? radius[i] : 0.0;
        const double scr = EV ? pair_screen(P, Ri, mass[i], p) : 0.0;
        // Sources reach the lanes through buffer loads: the lane's byte offset is fixed for the
        // sweep and each chunk gets its own descriptor, based at the chunk's first source and
        // ending at the sweep's last one (wave-uniform, so built with scalar instructions).  The
        // hardware range check compares the lane offset with the descriptor's size, so slots
        // past the last source in the last chunk read zeros (they are never used) and no load
        // leaves the sweep's sources; the host keeps nsrc * 32 bytes below 2^31.
        unsigned off_r[M], off_i[M];
#pragma unroll
        for (int m = 0; m < M; ++m) {
            off_r[m] = (unsigned)(slot + m * SLOTS) * (unsigned)sizeof(d4);
            off_i[m] = (unsigned)(slot + m * SLOTS) * (unsigned)sizeof(int);
        }
        // raw gate values are kept: comparing them here would wait for the loads
        auto fetch = [&](int c, d4* qd, int* idd) {
            const int sb = c * SQ_J;
            const __amdgpu_buffer_rsrc_t rs_r = __builtin_amdgcn_make_buffer_rsrc(
                (void*)(r + sw.src0 + sb), 0, (nsrc - sb) * (int)sizeof(d4), BUF_RSRC_DW3);
            const __amdgpu_buffer_rsrc_t rs_i = __builtin_amdgcn_make_buffer_rsrc(
                (void*)(sw.gate + sw.src0 + sb), 0, (nsrc - sb) * (int)sizeof(int), BUF_RSRC_DW3);
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const double2 lo =
                    __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(rs_r, off_r[m], 0, 0));
                const double2 hi =
                    __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(rs_r, off_r[m] + 16u, 0, 0));
                qd[m] = d4{lo.x, lo.y, hi.x, hi.y};
                idd[m] = (int)__builtin_amdgcn_raw_buffer_load_b32(rs_i, off_i[m], 0, 0);
            }
        };
        // one term: formed for every slot (the quotient is masked afterwards, so the M
        // dependency chains of a lane interleave), written when the slot is in the chunk
        auto term = [&](int bsel, int jj, int j, const d4& qm, int idm, bool in, double& dd, bool& hit) {
            const double dx = qm.x - p.x;
            const double dy = qm.y - p.y;
            const double dz = qm.z - p.z;
            const double r2 = (dx * dx + dy * dy) + dz * dz;
            const double d = sqrt(r2);
            const bool act = idm >= 0;
            const bool use = sink_in && (OVERLAP ? (act && j != i) : act);
            const double qt = qm.w / (d * r2);
            const double sc = use ? qt : 0.0;
            if (in) {
                buf[bsel][jj][3 * sk + 0] = sc * dx;
                buf[bsel][jj][3 * sk + 1] = sc * dy;
                buf[bsel][jj][3 * sk + 2] = sc * dz;
            }
            dd = d;
            hit = EV && in && use && live && d < scr;
        };
        // full chunk with every operand of the wave in the core range: the same terms through
        // sqrt_core / div_core
        auto compute_core = [&](int j0, int bsel, const d4* q, const int* qid, double* dd, bool* hit) {
            double dx[M], dy[M], dz[M], r2[M];
            bool use[M];
            bool ok = true;
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const int j = j0 + slot + m * SLOTS;
                dx[m] = q[m].x - p.x;
                dy[m] = q[m].y - p.y;
                dz[m] = q[m].z - p.z;
                r2[m] = (dx[m] * dx[m] + dy[m] * dy[m]) + dz[m] * dz[m];
                const bool act = qid[m] >= 0;
                use[m] = sink_in && (OVERLAP ? (act && j != i) : act);
            }
            // source masses need no test per pair when the host has found them all in range
            if (P.mass_core) {
#pragma unroll
                for (int m = 0; m < M; ++m) ok = ok && (!use[m] || in_core_range(r2[m]));
            } else {
#pragma unroll
                for (int m = 0; m < M; ++m) ok = ok && (!use[m] || (in_core_range(r2[m]) && in_core_range(q[m].w)));
            }
            if (!__all(ok)) return false;
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const int jj = slot + m * SLOTS;
                const double d = sqrt_core(r2[m]);
                const double qt = div_core(q[m].w, d * r2[m]);
                const double sc = use[m] ? qt : 0.0;
                buf[bsel][jj][3 * sk + 0] = sc * dx[m];
                buf[bsel][jj][3 * sk + 1] = sc * dy[m];
                buf[bsel][jj][3 * sk + 2] = sc * dz[m];
                dd[m] = d;
                hit[m] = EV && use[m] && live && d < scr;
            }
            return true;
        };
        auto compute = [&](int c, int bsel, const d4* q, const int* qid) {
            const int j0 = sw.src0 + c * SQ_J;
            const int nj = min(SQ_J, sw.src1 - j0);
            double dd[M];
            bool hit[M];
            if (nj == SQ_J && compute_core(j0, bsel, q, qid, dd, hit)) {
                // done on the core path
            } else if (nj == SQ_J) {   // full chunk (uniform): straight-line code
#pragma unroll
                for (int m = 0; m < M; ++m)
                    term(bsel, slot + m * SLOTS, j0 + slot + m * SLOTS, q[m], qid[m], true, dd[m], hit[m]);
            } else {
#pragma unroll
                for (int m = 0; m < M; ++m)
                    term(bsel, slot + m * SLOTS, j0 + slot + m * SLOTS, q[m], qid[m], slot + m * SLOTS < nj, dd[m],
                         hit[m]);
            }
            if (EV) {
#pragma unroll
                for (int m = 0; m < M; ++m) {
                    if (hit[m]) {
                        const int j = j0 + slot + m * SLOTS;
                        pair_test(P, E, ctl, i, j, dd[m], Ri, radius[j], r, v, mass, id);
                    }
                }
            }
        };
        // two register sets: even chunks in A, odd chunks in B, each refilled with the chunk
        // two ahead as soon as it has been computed (no copies between the sets); the
        // producers pass the same 1 + nch barriers as the consumer
        d4 qA[M], qB[M];
        int iA[M], iB[M];
        if (nch > 0) {
            fetch(0, qA, iA);
            if (nch > 1) fetch(1, qB, iB);
            compute(0, 0, qA, iA);
            if (nch > 2) fetch(2, qA, iA);
        }
        __syncthreads();
        for (int c = 0; c < nch; c += 2) {
            if (c + 1 < nch) {
                compute(c + 1, 1, qB, iB);
                if (c + 3 < nch) fetch(c + 3, qB, iB);
            }
            __syncthreads();
            if (c + 1 < nch) {
                if (c + 2 < nch) {
                    compute(c + 2, 0, qA, iA);
                    if (c + 4 < nch) fetch(c + 4, qA, iA);
                }
                __syncthreads();
            }
        }
        return;
    }
    // consumer wave: its additions are the block's serial chain, each waiting on the last,
    // so it issues ahead of the producer waves on its SIMD whenever its next addition is
    // ready (the producers fill the cycles in between)
    __builtin_amdgcn_s_setprio(3);
    __syncthreads();
    for (int c = 0; c < nch; ++c) {
        const int b = c & 1;
        if (lane < 3 * ns) {
            const int nj = min(SQ_J, nsrc - c * SQ_J);
            if (nj == SQ_J) {
                // full chunk: the reads of the next GQ terms are in flight while the current
                // GQ are added (scheduling barriers keep the compiler from sinking the reads
                // next to their uses, which would put the LDS latency on the serial chain)
                constexpr int GQ = SQ_J % 16 == 0 ? 16 : (SQ_J % 8 == 0 ? 8 : 4);
                static_assert(SQ_J % GQ == 0, "consumer groups");
                double x[GQ], y[GQ];
#pragma unroll
                for (int k = 0; k < GQ; ++k) x[k] = buf[b][k][lane];
#pragma unroll
                for (int g = 0; g < SQ_J / GQ; ++g) {
                    if (g + 1 < SQ_J / GQ) {
#pragma unroll
                        for (int k = 0; k < GQ; ++k) y[k] = buf[b][(g + 1) * GQ + k][lane];
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int k = 0; k < GQ; ++k) sum += x[k];
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int k = 0; k < GQ; ++k) x[k] = y[k];
                }
            } else {
                for (int jj = 0; jj < nj; ++jj) sum += buf[b][jj][lane];
            }
        }
        __syncthreads();
    }
    if (lane < 3 * ns) {
        const int i = i0 + lane / 3;
        const int comp = lane % 3;
        double* o = reinterpret_cast<double*>(a + i);
        o[comp] = sw.gate[i] > 0 ? sum : 0.0;
        if (comp == 0) o[3] = 0.0;
    }
}

template <bool OV, bool EV, int SG>
hipError_t go_fp64_seq_sg(int np, const Sweep& sw, const GravBuffers& B, const PairTest& P, const EventSink& E,
                          hipStream_t s)
{
    dim3 g((sw.sink1 - sw.sink0 + SG - 1) / SG);
    // at most one block per CU (the SI sweep of a few hundred sinks): LDS is no constraint,
    // and longer chunks mean fewer barriers on the serial chain
    if constexpr (SG == 4) {
        if (g.x <= 256 && np > 7) {
            hipLaunchKernelGGL((k_grav_fp64_seq<OV, EV, 15, SG, SQ_TILE_FEW>), g, dim3(1024), 0, s, B.r, B.v, B.a,
                               B.mass, B.radius, B.id, sw, P, E, B.ctl);
            return hipGetLastError();
        }
        if (g.x <= 256 && np > 3) {
            hipLaunchKernelGGL((k_grav_fp64_seq<OV, EV, 7, SG, SQ_TILE_FEW>), g, dim3(512), 0, s, B.r, B.v, B.a,
                               B.mass, B.radius, B.id, sw, P, E, B.ctl);
            return hipGetLastError();
        }
    }
    if (np <= 3)
        hipLaunchKernelGGL((k_grav_fp64_seq<OV, EV, 3, SG>), g, dim3(256), 0, s, B.r, B.v, B.a, B.mass, B.radius, B.id,
                           sw, P, E, B.ctl);
    else if (np <= 7)
        hipLaunchKernelGGL((k_grav_fp64_seq<OV, EV, 7, SG>), g, dim3(512), 0, s, B.r, B.v, B.a, B.mass, B.radius, B.id,
                           sw, P, E, B.ctl);
    else
        hipLaunchKernelGGL((k_grav_fp64_seq<OV, EV, 15, SG>), g, dim3(1024), 0, s, B.r, B.v, B.a, B.mass, B.radius,
                           B.id, sw, P, E, B.ctl);
    return hipGetLastError();
}

template <bool OV, bool EV>
hipError_t go_fp64_seq(int np, int sg, const Sweep& sw, const GravBuffers& B, const PairTest& P, const EventSink& E,
                       hipStream_t s)
{
    if (sg >= 16) return go_fp64_seq_sg<OV, EV, 16>(np, sw, B, P, E, s);
    if (sg >= 8) return go_fp64_seq_sg<OV, EV, 8>(np, sw, B, P, E, s);
    return go_fp64_seq_sg<OV, EV, 4>(np, sw, B, P, E, s);
}

// ---------------------------------------------------------------------------------------
// FP64 sweep with one work-item per sink: the reference's own loop (pp_disk.cu:217-296),
// each sink's sums formed in ascending j by its work-item, for sweeps whose sink and source
// ranges are disjoint (NSI x SI, NI x massive) and whose sinks fill the chip (65536 sinks are
// one wavefront per SIMD).  Sources reach the work-group through LDS, chunk by chunk, from a
// packed copy (k_pack_fp64) in which an inactive source (id < 0) sits at (1e30, 0, 0)
// with mass 0: its term is +0 (a zero quotient times the difference), which leaves every sum
// as the reference's skip of that source does (the sums start at +0 and x + (+-0) == x), so
// the loop has no per-source test.  Groups of four sources give independent terms, so one
// wavefront keeps the FP64 pipeline busy while its three sums wait on their last addition.
// The terms use the IEEE square-root and quotient cores: a chunk whose packed masses are not
// all 0 or inside [2^-300, 2^300], whose coordinates with the sink's could push r^2 to 2^300,
// or in which some r^2 falls below 2^-300 is formed again from its starting sums with sqrt()
// and the division operator.  Pair tests: a chunk whose smallest r^2 is below the sink's
// screen for that chunk (collision reach with the chunk's largest radius, Hill reach with the
// largest non-star mass; squared, with margin) is swept again over the real positions and
// ids for the reference predicate with the exact distance.
constexpr int TPS_MIN_SINKS = 49152;      // automatic choice from this many sinks
constexpr double TPS_PARK = 1e30;         // where the packed copy parks inactive sources

__global__ __launch_bounds__(TPS_CHUNK) void k_pack_fp64(const d4* __restrict__ r, const int* __restrict__ gate,
                                                         const double* __restrict__ radius, d4* __restrict__ out,
                                                         TpsChunk* __restrict__ chunks, int n)
{
    __shared__ double wmax[2][TPS_CHUNK / 64];
    const int j = blockIdx.x * TPS_CHUNK + threadIdx.x;
    double b = 0.0, rm = 0.0;
    int ok = 1;
    if (j < n) {
        const d4 q = r[j];
        const bool act = gate[j] >= 0;
        const d4 o = act ? q : d4{TPS_PARK, 0.0, 0.0, 0.0};
        out[j] = o;
        b = fmax(fabs(o.x), fmax(fabs(o.y), fabs(o.z)));
        rm = act ? radius[j] : 0.0;
        ok = o.w == 0.0 || in_core_range(o.w);
    }
    ok = __syncthreads_and(ok);
    for (int off = 32; off > 0; off >>= 1) {
        b = fmax(b, __shfl_xor(b, off));
        rm = fmax(rm, __shfl_xor(rm, off));
    }
    if ((threadIdx.x & 63) == 0) {
        wmax[0][threadIdx.x >> 6] = b;
        wmax[1][threadIdx.x >> 6] = rm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < TPS_CHUNK / 64; w++) {
            b = fmax(b, wmax[0][w]);
            rm = fmax(rm, wmax[1][w]);
        }
        chunks[blockIdx.x] = TpsChunk{b, rm, ok, 0};
    }
}

__device__ __forceinline__ void tps_term(const d4& q, const d4& p, double& ax, double& ay, double& az, double& lo)
{
    const double dx = q.x - p.x;
    const double dy = q.y - p.y;
    const double dz = q.z - p.z;
    const double r2 = (dx * dx + dy * dy) + dz * dz;
    const double d = sqrt_core(r2);
    const double sc = div_core(q.w, d * r2);
    ax += sc * dx;
    ay += sc * dy;
    az += sc * dz;
    lo = fmin(lo, r2);
}

// One wavefront per SIMD: the occupancy hint lets the scheduler interleave a group's four
// terms (with the default target it may form them one after another: 8.2 against 6.6 ms on a
// 65536 x 65536 sweep).  A work-group is TPS_CHUNK work-items.
template <bool EV>
__global__ __launch_bounds__(TPS_CHUNK) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_grav_fp64_tps(
    const d4* __restrict__ r, const d4* __restrict__ v,
                                                       d4* __restrict__ a, const double* __restrict__ mass,
                                                       const double* __restrict__ radius,
                                                       const int* __restrict__ id, const d4* __restrict__ src,
                                                       const TpsChunk* __restrict__ chunk, Sweep sw, PairTest P,
                                                       EventSink E, Ctl* __restrict__ ctl)
{
    constexpr double LO = 0x1p-300, COORD = 0x1p149;   // |dx|, |dy|, |dz| < 2^149: r^2 < 2^300
    const int i = sw.sink0 + blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = i < sw.sink1;
    const int ic = in ? i : sw.sink0;
    const d4 p = r[ic];
    const double pb = fmax(fabs(p.x), fmax(fabs(p.y), fabs(p.z)));
    const bool live = in && sw.gate[ic] > 0;
    const double Ri = EV ? radius[ic] : 0.0;
    // the screen's Hill reach for this sink; its collision reach is taken per chunk, with the
    // chunk's largest radius (a disk's star is in one chunk, not in every sink's screen)
    const double hill = EV && live ? pair_screen_hill(P, mass[ic], p) : 0.0;
    double ax = 0.0, ay = 0.0, az = 0.0;
    // Sources pass through LDS, a chunk per work-item each: the next chunk's loads are issued
    // when a chunk starts and stored when it ends (freshly written positions come from memory,
    // about a microsecond, which scalar loads in the loop would wait for group by group)
    __shared__ d4 stage[2][TPS_CHUNK];
    const int t = threadIdx.x;
    const int nch = (sw.src1 - sw.src0 + TPS_CHUNK - 1) / TPS_CHUNK;
    if (nch > 0 && sw.src0 + t < sw.src1) stage[0][t] = src[sw.src0 + t];
    __syncthreads();
    for (int c = 0; c < nch; ++c) {
        const int c0 = sw.src0 + c * TPS_CHUNK;
        const int c1 = min(sw.src1, c0 + TPS_CHUNK);
        const int nc = c1 - c0;
        const d4* S = stage[c & 1];
        const int jn = c0 + TPS_CHUNK + t;
        const bool has_next = c + 1 < nch && jn < sw.src1;
        d4 next{0.0, 0.0, 0.0, 0.0};
        if (has_next) next = src[jn];
        const TpsChunk ch = chunk[c0 / TPS_CHUNK];
        const double scr = EV && live ? fmax(P.radii_f * (Ri + ch.rmax), hill) * 1.001 : -1.0;
        // d < scr implies r^2 < scr^2 (1 + 2^-51): the squared screen with a margin
        const double scr2 = scr > 0.0 ? scr * scr * 1.0001 : -1.0;
        const double bx = ax, by = ay, bz = az;
        double lo = 0x1p300;        // smallest r^2 of the chunk
        // two groups of four per pass: the second group's reads are in flight while the first
        // group's terms are formed, and the next pass's first group while the second's are
        int k = 0;
        if (nc >= 8) {
            d4 a0 = S[0], a1 = S[1], a2 = S[2], a3 = S[3];
            for (; k + 8 <= nc; k += 8) {
                const d4 b0 = S[k + 4], b1 = S[k + 5], b2 = S[k + 6], b3 = S[k + 7];
                tps_term(a0, p, ax, ay, az, lo);
                tps_term(a1, p, ax, ay, az, lo);
                tps_term(a2, p, ax, ay, az, lo);
                tps_term(a3, p, ax, ay, az, lo);
                if (k + 12 <= nc) {
                    a0 = S[k + 8];
                    a1 = S[k + 9];
                    a2 = S[k + 10];
                    a3 = S[k + 11];
                }
                tps_term(b0, p, ax, ay, az, lo);
                tps_term(b1, p, ax, ay, az, lo);
                tps_term(b2, p, ax, ay, az, lo);
                tps_term(b3, p, ax, ay, az, lo);
            }
        }
        for (; k < nc; ++k) tps_term(S[k], p, ax, ay, az, lo);
        if (!ch.mass_ok || !__all(lo >= LO && pb + ch.bound < COORD)) {
            // outside the cores' range somewhere in the wavefront: the chunk again, general
            ax = bx;
            ay = by;
            az = bz;
            for (k = 0; k < nc; ++k) {
                const d4 q = S[k];
                const double dx = q.x - p.x;
                const double dy = q.y - p.y;
                const double dz = q.z - p.z;
                const double r2 = (dx * dx + dy * dy) + dz * dz;
                const double d = sqrt(r2);
                const double sc = q.w / (d * r2);
                ax += sc * dx;
                ay += sc * dy;
                az += sc * dz;
            }
        }
        if (EV && __any(lo < scr2)) {
            for (int j = c0; j < c1; ++j) {
                if (sw.gate[j] < 0) continue;
                const d4 q = r[j];
                const double dx = q.x - p.x;
                const double dy = q.y - p.y;
                const double dz = q.z - p.z;
                const double r2 = (dx * dx + dy * dy) + dz * dz;
                const double d = sqrt(r2);
                if (live && d < scr) pair_test(P, E, ctl, i, j, d, Ri, radius[j], r, v, mass, id);
            }
        }
        // the buffer being filled was last read in the previous chunk, before its barrier
        if (has_next) stage[(c + 1) & 1][t] = next;
        __syncthreads();
    }
    if (in) a[i] = d4{live ? ax : 0.0, live ? ay : 0.0, live ? az : 0.0, 0.0};
}

// ---------------------------------------------------------------------------------------
// Astro-centric frame (the reference's CPU path, pp_disk.cu:409-454; its GPU kernel has no
// such term).  The star sits at the origin and is neither sink nor source of the SI sweep
// (number_of_bodies.cpp:205-209, 218-221); every sink starts from the star's pull
// -(m_star + m_i) r_i / |r_i|^3 and each active source j != i adds its direct term, then
// subtracts its indirect term m_j r_j / |r_j|^3, in the reference's operation order.  The
// star's own row (below sink0) is written as zero.
template <bool EV>
__global__ __launch_bounds__(256) void k_grav_fp64_astro(const d4* __restrict__ r, const d4* __restrict__ v,
                                                         d4* __restrict__ a, const double* __restrict__ mass,
                                                         const double* __restrict__ radius,
                                                         const int* __restrict__ id, Sweep sw, double mstar,
                                                         PairTest P, EventSink E, Ctl* __restrict__ ctl)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sw.sink1) return;
    if (i < sw.sink0) {   // the star rests at the origin of this frame
        a[i] = d4{0.0, 0.0, 0.0, 0.0};
        return;
    }
    const d4 qi = r[i];
    const double mu = mstar + mass[i];
    const double r2 = (qi.x * qi.x) + (qi.y * qi.y) + (qi.z * qi.z);
    double ax = -mu / (r2 * sqrt(r2)) * qi.x;
    double ay = -mu / (r2 * sqrt(r2)) * qi.y;
    double az = -mu / (r2 * sqrt(r2)) * qi.z;
    if (sw.gate[i] > 0) {
        const double Ri = EV ? radius[i] : 0.0;
        for (int j = sw.src0; j < sw.src1; ++j) {
            if (i == j || sw.gate[j] < 0) continue;
            const d4 q = r[j];
            const double dx = q.x - qi.x;
            const double dy = q.y - qi.y;
            const double dz = q.z - qi.z;
            const double dr2 = (dx * dx) + (dy * dy) + (dz * dz);
            const double d = sqrt(dr2);
            const double sc = mass[j] / (d * dr2);
            ax += sc * dx;
            ay += sc * dy;
            az += sc * dz;
            const double rj2 = (q.x * q.x) + (q.y * q.y) + (q.z * q.z);
            ax -= mass[j] / (rj2 * sqrt(rj2)) * q.x;
            ay -= mass[j] / (rj2 * sqrt(rj2)) * q.y;
            az -= mass[j] / (rj2 * sqrt(rj2)) * q.z;
            if (EV) pair_test(P, E, ctl, i, j, d, Ri, radius[j], r, v, mass, id);
        }
    }
    a[i] = d4{ax, ay, az, 0.0};
}

// ---------------------------------------------------------------------------------------
// Mixed-precision broadcast sweep.
__device__ __forceinline__ void kahan_add(double& sum, double& comp, double x)
{
    const double y = x - comp;
    const double t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

// A work-group's four wavefronts share 64 x SPT sinks (SPT per work-item) and split every
// 64-source chunk of the group's part into four 16-source slices, wavefront w taking slice
// w: they read the same source lines through the scalar cache at about the same time, and
// the work-group writes one partial sum per sink (a quarter of the part sums that one
// wavefront per part would write: 25 instead of 75 MB per launch on the 65533 x 65539
// test-particle sweep, for 2 % more time).  Each wavefront keeps FP32 sums over its slices
// of four consecutive chunks (64 pair terms per sink), Kahan-adds them into FP64, and the
// four FP64 sums meet in LDS in wavefront order.
template <int SPT, bool OVERLAP, bool EV, bool KAHAN>
__global__ __launch_bounds__(256) void k_grav_mixed_bcast(const d4* __restrict__ r, const d4* __restrict__ v,
                                                          d4* __restrict__ a, d4* __restrict__ partial,
                                                          const double* __restrict__ mass,
                                                          const double* __restrict__ radius,
                                                          const int* __restrict__ id,
                                                          const SrcMixed* __restrict__ src, Sweep sw, int parts,
                                                          int n_fp64_src, PairTest P, EventSink E,
                                                          Ctl* __restrict__ ctl)
{
    constexpr int NS = 64 * SPT;               // sinks per work-group
    __shared__ double wsum[3][3][NS];          // wavefronts 1-3: x, y, z sums of every sink
    const int lane = threadIdx.x & 63;
    // wave-uniform, so that the source index stays scalar and sources arrive by scalar loads
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int p = blockIdx.y;
    const int base = sw.sink0 + blockIdx.x * NS + lane;
    int ii[SPT];
    bool live[SPT];
    double xi[SPT], yi[SPT], zi[SPT], Ri[SPT], ax[SPT], ay[SPT], az[SPT], cx[SPT], cy[SPT], cz[SPT];
    float cut2[SPT];
#pragma unroll
    for (int s = 0; s < SPT; s++) {
        ii[s] = base + s * 64;
        const bool in = ii[s] < sw.sink1;
        const int k = in ? ii[s] : sw.sink0;
        const d4 q = r[k];
        xi[s] = q.x; yi[s] = q.y; zi[s] = q.z;
        live[s] = in && sw.gate[k] > 0;
        Ri[s] = EV ? radius[k] : 0.0;
        ax[s] = ay[s] = az[s] = 0.0;
        cx[s] = cy[s] = cz[s] = 0.0;
        if (EV) {
            // conservative screen: FP32 r^2 is within ~5e-7 relative of the FP64 value
            double cut = P.radii_f * (Ri[s] + P.rmax);
            if (P.hill_f > 0.0) {
                // d < f c (|r_i| + |r_j|)/2 with c = ((m_i + m_j)/(3 M_star))^(1/3) and
                // |r_j| <= |r_i| + d implies d < f c |r_i| / (1 - f c / 2), c taken at m_max
                const double fc = P.hill_f * cbrt((q.w + P.mmax) / P.three_mstar);
                const double hr = sqrt((q.x * q.x) + (q.y * q.y) + (q.z * q.z));
                cut = fc < 1.0 ? fmax(cut, fc * hr / (1.0 - 0.5 * fc)) : 1e18;
            }
            cut *= 1.001;
            cut2[s] = (float)fmin(cut * cut, 3.0e38);
        }
    }
    // FP64 head (the star and giant planets), processed by wavefront 0 of part 0
    const int jsplit = min(sw.src1, max(sw.src0, n_fp64_src));
    if (p == 0 && w == 0) {
        for (int j = sw.src0; j < jsplit; ++j) {
            const bool act = sw.gate[j] >= 0;
            const d4 q = r[j];
#pragma unroll
            for (int s = 0; s < SPT; s++)
                fp64_pair<OVERLAP, EV>(q, j, act, ii[s], xi[s], yi[s], zi[s], Ri[s], live[s], ax[s], ay[s], az[s], P,
                                       E, ctl, r, v, mass, radius, id);
        }
    }
    int jlo, jhi;
    part_range(jsplit, sw.src1, parts, p, CHUNK, jlo, jhi);
    for (int g0 = jlo; g0 < jhi; g0 += 4 * CHUNK) {
        float sx[SPT], sy[SPT], sz[SPT], kx[SPT], ky[SPT], kz[SPT], mn[SPT];
#pragma unroll
        for (int s = 0; s < SPT; s++) {
            sx[s] = sy[s] = sz[s] = 0.0f;
            kx[s] = ky[s] = kz[s] = 0.0f;
            mn[s] = 3.0e38f;
        }
        for (int c = 0; c < 4; c++) {
            const int j0 = g0 + c * CHUNK + w * (CHUNK / 4);
            const int j1 = min(jhi, j0 + CHUNK / 4);
#pragma unroll 4
            for (int j = j0; j < j1; ++j) {
                // one 32-byte scalar load; the low word of w is the FP32 mass (0 if inactive)
                const d4 q = reinterpret_cast<const d4*>(src)[j];
                const float qm =
                    __uint_as_float((unsigned)((unsigned long long)__double_as_longlong(q.w) & 0xffffffffull));
#pragma unroll
                for (int s = 0; s < SPT; s++) {
                    const float dx = (float)(q.x - xi[s]);
                    const float dy = (float)(q.y - yi[s]);
                    const float dz = (float)(q.z - zi[s]);
                    const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                    // one v_rsq_f32 (half rate: two issue slots) and three multiplies; an
                    // added v_rcp_f32 for r^-2 in place of inv * inv measured 11 % slower
                    const float inv = __builtin_amdgcn_rsqf(r2);
                    float sc = (qm * inv) * (inv * inv);
                    if (OVERLAP) sc = (j == ii[s]) ? 0.0f : sc;
                    if (KAHAN) {
                        float y, t;
                        y = __builtin_fmaf(sc, dx, -kx[s]); t = sx[s] + y; kx[s] = (t - sx[s]) - y; sx[s] = t;
                        y = __builtin_fmaf(sc, dy, -ky[s]); t = sy[s] + y; ky[s] = (t - sy[s]) - y; sy[s] = t;
                        y = __builtin_fmaf(sc, dz, -kz[s]); t = sz[s] + y; kz[s] = (t - sz[s]) - y; sz[s] = t;
                    } else {
                        sx[s] = __builtin_fmaf(sc, dx, sx[s]);
                        sy[s] = __builtin_fmaf(sc, dy, sy[s]);
                        sz[s] = __builtin_fmaf(sc, dz, sz[s]);
                    }
                    if (EV) mn[s] = fminf(mn[s], (OVERLAP && j == ii[s]) ? 3.0e38f : r2);
                }
            }
        }
#pragma unroll
        for (int s = 0; s < SPT; s++) {
            kahan_add(ax[s], cx[s], (double)sx[s]);
            kahan_add(ay[s], cy[s], (double)sy[s]);
            kahan_add(az[s], cz[s], (double)sz[s]);
            if (EV && live[s] && mn[s] < cut2[s]) {
                // exact FP64 re-evaluation of the reference predicate for this wavefront's chunk
                for (int c = 0; c < 4; c++) {
                    const int j0 = g0 + c * CHUNK + w * (CHUNK / 4);
                    const int j1 = min(jhi, j0 + CHUNK / 4);
                    for (int j = j0; j < j1; ++j) {
                        if (sw.gate[j] < 0 || j == ii[s]) continue;
                        const d4 q = r[j];
                        const double dx = q.x - xi[s];
                        const double dy = q.y - yi[s];
                        const double dz = q.z - zi[s];
                        const double r2 = (dx * dx + dy * dy) + dz * dz;
                        pair_test(P, E, ctl, ii[s], j, sqrt(r2), Ri[s], radius[j], r, v, mass, id);
                    }
                }
            }
        }
    }
    // the four wavefronts' sums (each with its compensation) added in wavefront order
    if (w > 0) {
#pragma unroll
        for (int s = 0; s < SPT; s++) {
            wsum[w - 1][0][s * 64 + lane] = ax[s] - cx[s];
            wsum[w - 1][1][s * 64 + lane] = ay[s] - cy[s];
            wsum[w - 1][2][s * 64 + lane] = az[s] - cz[s];
        }
    }
    __syncthreads();
    if (w == 0) {
#pragma unroll
        for (int s = 0; s < SPT; s++) {
            double x = ax[s] - cx[s], y = ay[s] - cy[s], z = az[s] - cz[s];
            for (int k = 0; k < 3; k++) {
                x += wsum[k][0][s * 64 + lane];
                y += wsum[k][1][s * 64 + lane];
                z += wsum[k][2][s * 64 + lane];
            }
            if (ii[s] < sw.sink1) store_acc(a, partial, parts, p, sw, ii[s], live[s], x, y, z);
        }
    }
}

// Sums the parts of a split sweep in part order (deterministic).
__global__ __launch_bounds__(256) void k_reduce_parts(const d4* __restrict__ partial, d4* __restrict__ a,
                                                      const int* __restrict__ id, Sweep sw, int parts)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int nsink = sw.sink1 - sw.sink0;
    if (k >= nsink) return;
    double x = 0.0, y = 0.0, z = 0.0;
    const double* planes = reinterpret_cast<const double*>(partial);
    for (int p = 0; p < parts; p++) {
        const double* q = planes + 3 * (size_t)p * nsink + k;
        x += q[0]; y += q[nsink]; z += q[2 * (size_t)nsink];
    }
    const int i = sw.sink0 + k;
    const bool live = sw.gate[i] > 0;
    a[i] = d4{live ? x : 0.0, live ? y : 0.0, live ? z : 0.0, 0.0};
}

__global__ void k_zero_acc(d4* __restrict__ a, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = d4{0.0, 0.0, 0.0, 0.0};
}

__global__ void k_pack_mixed(const d4* __restrict__ r, const int* __restrict__ id, SrcMixed* __restrict__ out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const d4 q = r[i];
    SrcMixed s;
    s.x = q.x; s.y = q.y; s.z = q.z;
    s.active = id[i] >= 0;
    s.m = s.active ? (float)q.w : 0.0f;
    out[i] = s;
}

// ---------------------------------------------------------------------------------------
// Split sweep, phase 1: one work-item per source j, SPLIT_SG sinks per block row.
constexpr int SPLIT_SG = 16;

template <bool EV>
__global__ __launch_bounds__(256) void k_split_terms(const d4* __restrict__ r, const d4* __restrict__ v,
                                                     const double* __restrict__ mass,
                                                     const double* __restrict__ radius,
                                                     const int* __restrict__ id, Sweep sw, PairTest P,
                                                     EventSink E, Ctl* __restrict__ ctl, double* __restrict__ T)
{
    const int nsrc = sw.src1 - sw.src0;
    const int jj = blockIdx.x * blockDim.x + threadIdx.x;
    const int i0 = sw.sink0 + blockIdx.y * SPLIT_SG;
    const int i1 = min(sw.sink1, i0 + SPLIT_SG);
    if (jj >= nsrc) return;
    const int j = sw.src0 + jj;
    const bool act = sw.gate[j] >= 0;
    const d4 q = r[j];
    const double Rj = EV ? radius[j] : 0.0;
    for (int i = i0; i < i1; ++i) {
        const d4 p = r[i];
        const double dx = q.x - p.x;
        const double dy = q.y - p.y;
        const double dz = q.z - p.z;
        const double r2 = (dx * dx + dy * dy) + dz * dz;
        const double d = sqrt(r2);
        const bool use = act && j != i;
        const double sc = use ? q.w / (d * r2) : 0.0;
        double* row = T + (size_t)(i - sw.sink0) * 3 * nsrc;
        row[jj] = sc * dx;
        row[nsrc + jj] = sc * dy;
        row[2 * nsrc + jj] = sc * dz;
        if (EV && use && sw.gate[i] > 0) pair_test(P, E, ctl, i, j, d, radius[i], Rj, r, v, mass, id);
    }
}

// Split sweep, phase 2 where the order is free and no segment buffer is at hand: one block
// per (sink, component); contiguous runs are summed in parallel and the run sums added in
// order.
__global__ __launch_bounds__(256) void k_split_sum(const double* __restrict__ T, d4* __restrict__ a,
                                                   const int* __restrict__ id, Sweep sw)
{
    __shared__ double part[256];
    const int nsrc = sw.src1 - sw.src0;
    const int row = blockIdx.x;            // (sink - sink0) * 3 + comp
    const int i = sw.sink0 + row / 3;
    const int c = row % 3;
    const double* t = T + (size_t)row * nsrc;
    const int per = (nsrc + blockDim.x - 1) / blockDim.x;
    const int b = threadIdx.x * per;
    const int e = min(nsrc, b + per);
    double acc = 0.0;
    for (int j = b; j < e; ++j) acc += t[j];
    part[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0;
        for (int k = 0; k < (int)blockDim.x; ++k) sum += part[k];
        double* o = reinterpret_cast<double*>(a + i);
        o[c] = sw.gate[i] > 0 ? sum : 0.0;
        if (c == 0) o[3] = 0.0;
    }
}

__device__ __forceinline__ double readlane_f64(double x, int l)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Split sweep, phase 2 in the reference order, one wave per (sink, component) row: lane k
// loads term 64 b + k of block b (one coalesced load per block, issued SPLIT_PF blocks
// ahead) and the wave adds the block's terms lane by lane in ascending j, starting from 0,
// so the serial chain of additions does not wait on memory.  Loads past the row end are
// clamped to its last term and never added.
constexpr int SPLIT_PF = 4;

__global__ __launch_bounds__(64) void k_split_chain(const double* __restrict__ T, d4* __restrict__ a,
                                                    const int* __restrict__ id, Sweep sw)
{
    const int nsrc = sw.src1 - sw.src0;
    const int row = blockIdx.x;
    const int lane = threadIdx.x;
    const double* t = T + (size_t)row * nsrc;
    const int nb = (nsrc + 63) / 64;
    double buf[SPLIT_PF];
#pragma unroll
    for (int k = 0; k < SPLIT_PF; ++k) buf[k] = t[min(k * 64 + lane, nsrc - 1)];
    double sum = 0.0;
    for (int b0 = 0; b0 < nb; b0 += SPLIT_PF) {
#pragma unroll
        for (int k = 0; k < SPLIT_PF; ++k) {
            const double cur = buf[k];
            buf[k] = t[min((b0 + k + SPLIT_PF) * 64 + lane, nsrc - 1)];
            const int cnt = nsrc - (b0 + k) * 64;
            if (cnt >= 64) {
#pragma unroll
                for (int l = 0; l < 64; ++l) sum += readlane_f64(cur, l);
            } else {
                for (int l = 0; l < cnt; ++l) sum += readlane_f64(cur, l);
            }
        }
    }
    if (lane == 0) {
        const int i = sw.sink0 + row / 3, c = row % 3;
        double* o = reinterpret_cast<double*>(a + i);
        o[c] = sw.gate[i] > 0 ? sum : 0.0;
        if (c == 0) o[3] = 0.0;
    }
}

// Sums of the split-sweep terms where the summation order is free (mixed precision,
// source-partitioned FP64): every row (sink, component) is cut into nseg contiguous
// segments, one block per segment with coalesced loads and a tree in LDS; k_split_sum_fin
// then adds each row's segments in segment order, so the result is run-to-run identical.
__global__ __launch_bounds__(256) void k_split_sum_seg(const double* __restrict__ T, double* __restrict__ seg,
                                                       Sweep sw, int nseg)
{
    __shared__ double red_lds[256];
    const int nsrc = sw.src1 - sw.src0;
    const int row = blockIdx.x, g = blockIdx.y;
    const int len = (nsrc + nseg - 1) / nseg;
    const int b = g * len, e = min(nsrc, b + len);
    const double* t = T + (size_t)row * nsrc;
    double acc = 0.0;
    for (int j = b + (int)threadIdx.x; j < e; j += 256) acc += t[j];
    red_lds[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red_lds[threadIdx.x] += red_lds[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) seg[(size_t)row * nseg + g] = red_lds[0];
}

__global__ __launch_bounds__(256) void k_split_sum_fin(const double* __restrict__ seg, d4* __restrict__ a,
                                                       const int* __restrict__ id, Sweep sw, int nseg)
{
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= (sw.sink1 - sw.sink0) * 3) return;
    double acc = 0.0;
    for (int k = 0; k < nseg; ++k) acc += seg[(size_t)row * nseg + k];
    const int i = sw.sink0 + row / 3, c = row % 3;
    double* o = reinterpret_cast<double*>(a + i);
    o[c] = sw.gate[i] > 0 ? acc : 0.0;
    if (c == 0) o[3] = 0.0;
}

template <int SPT, bool OV, bool EV>
hipError_t go_fp64(int strategy, dim3 g, dim3 b, const Sweep& sw, const GravBuffers& B, int parts,
                   const PairTest& P, const EventSink& E, hipStream_t s)
{
    if (strategy == 2)
        hipLaunchKernelGGL((k_grav_fp64<SPT, OV, EV, true>), g, b, 0, s, B.r, B.v, B.a, B.partial, B.mass, B.radius,
                           B.id, sw, parts, P, E, B.ctl);
    else
        hipLaunchKernelGGL((k_grav_fp64<SPT, OV, EV, false>), g, b, 0, s, B.r, B.v, B.a, B.partial, B.mass, B.radius,
                           B.id, sw, parts, P, E, B.ctl);
    return hipGetLastError();
}

template <int SPT, bool OV, bool EV, bool KH>
hipError_t go_mixed(dim3 g, dim3 b, const Sweep& sw, const GravBuffers& B, const GravConfig& C, int parts,
                    const PairTest& P, const EventSink& E, hipStream_t s)
{
    hipLaunchKernelGGL((k_grav_mixed_bcast<SPT, OV, EV, KH>), g, b, 0, s, B.r, B.v, B.a, B.partial, B.mass,
                       B.radius, B.id, B.src_mixed, sw, parts, C.n_fp64_src, P, E, B.ctl);
    return hipGetLastError();
}

template <int SPT>
hipError_t dispatch_tile(const Sweep& sw, const GravBuffers& B, const GravConfig& C, int parts, const PairTest& P,
                         const EventSink& E, hipStream_t s)
{
    const int nsink = sw.sink1 - sw.sink0;
    const int bs = C.block > 0 ? C.block : 256;
    dim3 b(bs), g((nsink + bs * SPT - 1) / (bs * SPT), parts);
    const bool ov = sw.src0 < sw.sink1 && sw.sink0 < sw.src1;
    const bool ev = P.enabled != 0;
    if (C.precision == RED_PREC_MIXED) {
        // four wavefronts per work-group on 64 x SPT sinks
        b = dim3(256);
        g = dim3((nsink + 64 * SPT - 1) / (64 * SPT), parts);
        if (C.kahan_pairs) {
            if (ov) return ev ? go_mixed<SPT, true, true, true>(g, b, sw, B, C, parts, P, E, s)
                              : go_mixed<SPT, true, false, true>(g, b, sw, B, C, parts, P, E, s);
            return ev ? go_mixed<SPT, false, true, true>(g, b, sw, B, C, parts, P, E, s)
                      : go_mixed<SPT, false, false, true>(g, b, sw, B, C, parts, P, E, s);
        }
        if (ov) return ev ? go_mixed<SPT, true, true, false>(g, b, sw, B, C, parts, P, E, s)
                          : go_mixed<SPT, true, false, false>(g, b, sw, B, C, parts, P, E, s);
        return ev ? go_mixed<SPT, false, true, false>(g, b, sw, B, C, parts, P, E, s)
                  : go_mixed<SPT, false, false, false>(g, b, sw, B, C, parts, P, E, s);
    }
    if (ov) return ev ? go_fp64<SPT, true, true>(C.strategy, g, b, sw, B, parts, P, E, s)
                      : go_fp64<SPT, true, false>(C.strategy, g, b, sw, B, parts, P, E, s);
    return ev ? go_fp64<SPT, false, true>(C.strategy, g, b, sw, B, parts, P, E, s)
              : go_fp64<SPT, false, false>(C.strategy, g, b, sw, B, parts, P, E, s);
}

}  // namespace

size_t split_scratch_bytes(const Sweep& sw)
{
    return (size_t)(sw.sink1 - sw.sink0) * 3 * (size_t)(sw.src1 - sw.src0) * sizeof(double);
}

hipError_t launch_pack_mixed(const d4* r, const int* id, SrcMixed* out, int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_mixed, dim3((n + 255) / 256), dim3(256), 0, s, r, id, out, n);
    return hipGetLastError();
}

int sweep_parts(const Sweep& sw, const GravConfig& C, size_t partial_elems)
{
    const int nsink = sw.sink1 - sw.sink0;
    const int nsrc = sw.src1 - sw.src0;
    if (nsink <= 0 || nsrc <= 0) return 1;
    const bool sequential = C.precision == RED_PREC_FP64 && !C.fp64_split;
    if (sequential) return 1;
    int parts = C.source_parts;
    if (parts <= 0) {
        // aim for about 24 wavefronts per SIMD (256 CUs x 4 SIMDs) for one wavefront per
        // part, 32 for the mixed sweep's four per part: measured on the 65533 x 65539
        // test-particle sweep (mixed, 2 sinks per work-item), one wavefront per part: 16
        // parts 1.70 ms, 32 parts 1.60, 48 parts 1.57, 64 1.57; four: 12 parts 1.605 ms,
        // 16 parts 1.597, 24 parts 1.63
        const int spt = C.spt > 0 ? C.spt : 1;
        const int groups = (nsink + 64 * spt - 1) / (64 * spt);
        const int per_part = C.precision == RED_PREC_MIXED ? 4 : 1;
        const int target = C.precision == RED_PREC_MIXED ? 32 * 1024 : 24 * 1024;
        parts = (target + groups * per_part - 1) / (groups * per_part);
        parts = std::min(parts, std::max(1, nsrc / (512 * per_part)));
    }
    parts = std::max(1, std::min(parts, 64));
    while (parts > 1 && (size_t)parts * nsink > partial_elems) parts--;
    return parts;
}

bool sweep_wants_split(const Sweep& sw, const GravConfig& C)
{
    const int nsink = sw.sink1 - sw.sink0;
    const int nsrc = sw.src1 - sw.src0;
    if (nsink <= 0 || nsrc <= 0 || C.astro) return false;
    // FP64 sequential sums over many sources for a handful of sinks: the producer/consumer
    // sweep would run a few blocks, so the terms are formed by the whole chip and each
    // (sink, component) row is added by a wave of its own
    if (C.precision == RED_PREC_FP64 && !C.fp64_split && (C.strategy == 3 || C.strategy == 4))
        return nsink <= 64 && nsrc >= 4096;
    return nsink <= C.split_max_sinks && nsrc >= 1024;
}

// One work-item per sink where the sink and source ranges are disjoint and the sources start
// at body 0 (the packed copy's chunks): asked for (strategy 4), or chosen when the sinks fill
// the chip and no producer/consumer shape was asked for.
bool sweep_takes_tps(const Sweep& sw, const GravConfig& C, const GravBuffers& B)
{
    const int nsink = sw.sink1 - sw.sink0;
    const int nsrc = sw.src1 - sw.src0;
    if (!B.src_fp64 || !B.tps_chunk || nsink <= 0 || nsrc <= 0 || nsrc >= SEQ_MAX_SOURCES || C.astro) return false;
    if (!(C.precision == RED_PREC_FP64 && !C.fp64_split && (C.strategy == 3 || C.strategy == 4))) return false;
    if (sweep_wants_split(sw, C)) return false;
    const bool ov = sw.src0 < sw.sink1 && sw.sink0 < sw.src1;
    if (ov || sw.src0 != 0) return false;
    return C.strategy == 4 || (C.seq_sinks <= 0 && C.block_seq <= 0 && nsink >= TPS_MIN_SINKS);
}

hipError_t launch_pack_fp64(const d4* r, const int* gate, const double* radius, d4* out, TpsChunk* chunks, int n,
                            hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_fp64, dim3((n + TPS_CHUNK - 1) / TPS_CHUNK), dim3(TPS_CHUNK), 0, s, r, gate, radius, out,
                       chunks, n);
    return hipGetLastError();
}

hipError_t launch_sweep(const Sweep& sw, const GravBuffers& B, const GravConfig& C, const PairTest& P,
                        const EventSink& E, hipStream_t s)
{
    const int nsink = sw.sink1 - sw.sink0;
    const int nsrc = sw.src1 - sw.src0;
    if (nsink <= 0) return hipSuccess;
    if (C.astro) {
        dim3 g((sw.sink1 + 255) / 256), b(256);   // also clears the star's row below sink0
        if (P.enabled)
            hipLaunchKernelGGL(k_grav_fp64_astro<true>, g, b, 0, s, B.r, B.v, B.a, B.mass, B.radius, B.id, sw,
                               C.astro_mstar, P, E, B.ctl);
        else
            hipLaunchKernelGGL(k_grav_fp64_astro<false>, g, b, 0, s, B.r, B.v, B.a, B.mass, B.radius, B.id, sw,
                               C.astro_mstar, P, E, B.ctl);
        return hipGetLastError();
    }
    const bool use_split = sweep_wants_split(sw, C) && split_scratch_bytes(sw) <= B.scratch_bytes;
    if (C.precision == RED_PREC_FP64 && !C.fp64_split && (C.strategy == 3 || C.strategy == 4) && nsrc > 0 &&
        nsrc < SEQ_MAX_SOURCES && !use_split) {
        const bool ov = sw.src0 < sw.sink1 && sw.sink0 < sw.src1;
        const bool ev = P.enabled != 0;
        if (sweep_takes_tps(sw, C, B)) {
            const dim3 g((nsink + TPS_CHUNK - 1) / TPS_CHUNK);
            if (ev)
                hipLaunchKernelGGL(k_grav_fp64_tps<true>, g, dim3(TPS_CHUNK), 0, s, B.r, B.v, B.a, B.mass, B.radius, B.id,
                                   B.src_fp64, B.tps_chunk, sw, P, E, B.ctl);
            else
                hipLaunchKernelGGL(k_grav_fp64_tps<false>, g, dim3(TPS_CHUNK), 0, s, B.r, B.v, B.a, B.mass, B.radius,
                                   B.id, B.src_fp64, B.tps_chunk, sw, P, E, B.ctl);
            return hipGetLastError();
        }
        // sinks per block (chunk length 1024 / sg) and producer waves per block: as asked for,
        // else long chunks for few sinks (the serial chain per sink is the critical path), and
        // 16 sinks (fewer consumer additions per term) once that still gives 3 blocks per CU;
        // many producers when there are few blocks to spread over the chip
        const int sg = C.seq_sinks > 0 ? C.seq_sinks : (nsink <= 4096 ? 4 : (nsink >= 12288 ? 16 : 8));
        const int nblk = (nsink + sg - 1) / sg;
        const int np = C.block_seq > 0 ? C.block_seq / 64 - 1 : (nblk >= 128 ? 7 : 15);
        if (ov) return ev ? go_fp64_seq<true, true>(np, sg, sw, B, P, E, s) : go_fp64_seq<true, false>(np, sg, sw, B, P, E, s);
        return ev ? go_fp64_seq<false, true>(np, sg, sw, B, P, E, s) : go_fp64_seq<false, false>(np, sg, sw, B, P, E, s);
    }
    if (use_split) {
        double* T = (double*)B.scratch;
        dim3 g1((nsrc + 255) / 256, (nsink + SPLIT_SG - 1) / SPLIT_SG);
        if (P.enabled)
            hipLaunchKernelGGL(k_split_terms<true>, g1, dim3(256), 0, s, B.r, B.v, B.mass, B.radius, B.id, sw, P, E,
                               B.ctl, T);
        else
            hipLaunchKernelGGL(k_split_terms<false>, g1, dim3(256), 0, s, B.r, B.v, B.mass, B.radius, B.id, sw, P,
                               E, B.ctl, T);
        RED_HIP_TRY(hipGetLastError());
        dim3 g2(nsink * 3);
        const int nrow = nsink * 3;
        const int nseg = std::min(SPLIT_SEG_MAX, std::max(1, nsrc / 2048));
        if (C.precision == RED_PREC_FP64 && !C.fp64_split) {
            hipLaunchKernelGGL(k_split_chain, g2, dim3(64), 0, s, T, B.a, B.id, sw);
        } else if (B.seg && (size_t)nrow * nseg <= B.seg_elems) {
            double* seg = B.seg;
            hipLaunchKernelGGL(k_split_sum_seg, dim3(nrow, nseg), dim3(256), 0, s, T, seg, sw, nseg);
            RED_HIP_TRY(hipGetLastError());
            hipLaunchKernelGGL(k_split_sum_fin, dim3((nrow + 255) / 256), dim3(256), 0, s, seg, B.a, B.id, sw, nseg);
        } else {
            hipLaunchKernelGGL(k_split_sum, g2, dim3(256), 0, s, T, B.a, B.id, sw);
        }
        return hipGetLastError();
    }
    if (nsrc <= 0) {
        // no sources: accelerations are zero (the reference zeroes a before the j loop)
        hipLaunchKernelGGL(k_zero_acc, dim3((nsink + 255) / 256), dim3(256), 0, s, B.a + sw.sink0, nsink);
        return hipGetLastError();
    }
    const int parts = sweep_parts(sw, C, B.partial_elems);
    hipError_t e;
    switch (C.spt) {
    case 2: e = dispatch_tile<2>(sw, B, C, parts, P, E, s); break;
    case 4: e = dispatch_tile<4>(sw, B, C, parts, P, E, s); break;
    default: e = dispatch_tile<1>(sw, B, C, parts, P, E, s); break;
    }
    RED_HIP_TRY(e);
    if (parts > 1) {
        hipLaunchKernelGGL(k_reduce_parts, dim3((nsink + 255) / 256), dim3(256), 0, s, B.partial, B.a, B.id, sw, parts);
        return hipGetLastError();
    }
    return hipSuccess;
}

}  // namespace red

// ---------------------------------------------------------------------------------------
// Fused fixed-step loop.  One work-group runs main-loop iterations of one small disk:
// the ejection / hit-centrum predicate (detection only), f0, every stage state and its
// forces, y_(n+1), then t += dt and ps += |dt|, with barriers where the stream of separate
// launches has its kernel boundaries.  The stage states come from combine<FORM> and the
// forces from the same pair terms as the sweep kernels, with every sink's sum in ascending
// j, so the state equals that of red_step bit for bit.
namespace red {
namespace {

constexpr int FUSED_BLOCK = 256;
constexpr int FUSED_TILE = 1536;   // pair terms per LDS tile (3 doubles each: 36 KiB)
constexpr int FUSED_TILE_SINKS = 16;

// A source as a pair needs it, staged in LDS: position, mass, radius, activity.
struct FusedSrc {
    double x, y, z, m, R;
    int act, pad;
};
// A sink of the few-sinks mode, staged in LDS.
struct FusedSink {
    double x, y, z, R;
    int live, pad;
};

struct FusedLds {
    double terms[3 * FUSED_TILE];
    FusedSrc src[FUSED_TILE];
    FusedSink sink[FUSED_TILE_SINKS];
    int flag;
    int eject[2];   // small disks: the ejection predicate of the next step, by step parity
};

// U pair terms of one work-item in the reference's operation order (as fp64_pair), from
// staged records, their dependency chains interleaved: the square roots and quotients go
// through the core sequences when every active lane's operands are in range (the wave
// votes), else through sqrt() and '/', with identical results.  Returns d and whether the
// pair counts; the caller runs the pair tests afterwards (few values live across them).
template <int U, bool OV>
__device__ __forceinline__ void fused_terms(const FusedSrc* q, const int* j, const int* i, const double* px,
                                            const double* py, const double* pz, const bool* val, double* tx,
                                            double* ty, double* tz, double* d, bool* use)
{
    double dx[U], dy[U], dz[U], r2[U], qt[U];
    bool ok = true;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        dx[u] = q[u].x - px[u];
        dy[u] = q[u].y - py[u];
        dz[u] = q[u].z - pz[u];
        r2[u] = (dx[u] * dx[u] + dy[u] * dy[u]) + dz[u] * dz[u];
        use[u] = val[u] && (OV ? (q[u].act && j[u] != i[u]) : q[u].act != 0);
        ok = ok && (!use[u] || (in_core_range(r2[u]) && in_core_range(q[u].m)));
    }
    if (__all(ok)) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            d[u] = sqrt_core(r2[u]);
            qt[u] = div_core(q[u].m, d[u] * r2[u]);
        }
    } else {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            d[u] = sqrt(r2[u]);
            qt[u] = q[u].m / (d[u] * r2[u]);
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const double sc = use[u] ? qt[u] : 0.0;
        tx[u] = sc * dx[u];
        ty[u] = sc * dy[u];
        tz[u] = sc * dz[u];
    }
}

// Stages sources [j0, j0 + nj) of a sweep in LDS (one global round trip for the tile).
__device__ __forceinline__ void fused_stage_sources(FusedLds& L, int j0, int nj, const d4* r, const double* radius,
                                                    const int* gate)
{
    for (int k = threadIdx.x; k < nj; k += FUSED_BLOCK) {
        const d4 q = r[j0 + k];
        L.src[k] = FusedSrc{q.x, q.y, q.z, q.w, radius[j0 + k], gate[j0 + k] >= 0 ? 1 : 0, 0};
    }
}

// One sweep by the whole work-group, every pair's operands staged in LDS.  Few sinks
// against many sources: all work-items form the pair terms of a tile (up to 16 sinks x
// FUSED_TILE / 16 sources) in LDS, then one lane per (sink, component) row adds them in
// ascending j; otherwise one sink per work-item adds its terms as they are formed, over
// source tiles.
template <bool OV, bool EV>
__device__ void fused_sweep(const Sweep& sw, const d4* r, const d4* v, d4* a, const double* mass,
                            const double* radius, const int* id, const PairTest& P, const EventSink& E, Ctl* ctl,
                            FusedLds& L)
{
    const int tid = threadIdx.x;
    const int nsink = sw.sink1 - sw.sink0;
    const int nsrc = sw.src1 - sw.src0;
    if (nsrc > 64 && nsink * 4 <= FUSED_BLOCK) {
        for (int g0 = 0; g0 < nsink; g0 += FUSED_TILE_SINKS) {
            const int S = min(FUSED_TILE_SINKS, nsink - g0);
            const int J = FUSED_TILE / S;
            if (tid < S) {
                const int i = sw.sink0 + g0 + tid;
                const d4 p = r[i];
                L.sink[tid] = FusedSink{p.x, p.y, p.z, EV ? radius[i] : 0.0, sw.gate[i] > 0 ? 1 : 0, 0};
            }
            double sum = 0.0;   // lane tid < 3 S: row (sink g0 + tid / 3, component tid % 3)
            for (int j0 = 0; j0 < nsrc; j0 += J) {
                const int nj = min(J, nsrc - j0);
                fused_stage_sources(L, sw.src0 + j0, nj, r, radius, sw.gate);
                __syncthreads();
                constexpr int TU = 2;   // terms per work-item and pass
                for (int k0 = 0; k0 < S * nj; k0 += TU * FUSED_BLOCK) {
                    FusedSrc q[TU];
                    int jg[TU], ig[TU], kk[TU], sk[TU];
                    double px[TU], py[TU], pz[TU], tx[TU], ty[TU], tz[TU], d[TU];
                    bool val[TU], use[TU];
#pragma unroll
                    for (int u = 0; u < TU; ++u) {
                        kk[u] = k0 + u * FUSED_BLOCK + tid;
                        val[u] = kk[u] < S * nj;
                        const int kc = val[u] ? kk[u] : 0;
                        sk[u] = kc / nj;
                        const int jj = kc - sk[u] * nj;
                        q[u] = L.src[jj];
                        jg[u] = sw.src0 + j0 + jj;
                        ig[u] = sw.sink0 + g0 + sk[u];
                        px[u] = L.sink[sk[u]].x; py[u] = L.sink[sk[u]].y; pz[u] = L.sink[sk[u]].z;
                    }
                    fused_terms<TU, OV>(q, jg, ig, px, py, pz, val, tx, ty, tz, d, use);
#pragma unroll
                    for (int u = 0; u < TU; ++u) {
                        if (val[u]) {
                            double* t3 = L.terms + 3 * kk[u];
                            t3[0] = tx[u]; t3[1] = ty[u]; t3[2] = tz[u];
                        }
                    }
                    if (EV) {
#pragma unroll
                        for (int u = 0; u < TU; ++u)
                            if (use[u] && L.sink[sk[u]].live)
                                pair_test(P, E, ctl, ig[u], jg[u], d[u], L.sink[sk[u]].R, q[u].R, r, v, mass, id);
                    }
                }
                __syncthreads();
                if (tid < 3 * S) {
                    // the row in ascending j; the next 16 terms are read while the current
                    // 16 are added (two register sets), so the LDS latency stays off the
                    // serial chain
                    const int s = tid / 3;
                    const double* t3 = L.terms + 3 * (s * nj) + (tid - 3 * s);
                    constexpr int GQ = 16;
                    const int ng = nj / GQ;
                    double x[GQ], y[GQ];
                    if (ng > 0) {
#pragma unroll
                        for (int k = 0; k < GQ; ++k) x[k] = t3[3 * k];
                    }
                    for (int g = 0; g < ng; g += 2) {
                        if (g + 1 < ng) {
#pragma unroll
                            for (int k = 0; k < GQ; ++k) y[k] = t3[3 * ((g + 1) * GQ + k)];
                        }
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int k = 0; k < GQ; ++k) sum += x[k];
                        __builtin_amdgcn_sched_barrier(0);
                        if (g + 1 < ng) {
                            if (g + 2 < ng) {
#pragma unroll
                                for (int k = 0; k < GQ; ++k) x[k] = t3[3 * ((g + 2) * GQ + k)];
                            }
                            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                            for (int k = 0; k < GQ; ++k) sum += y[k];
                            __builtin_amdgcn_sched_barrier(0);
                        }
                    }
                    for (int jj = ng * GQ; jj < nj; ++jj) sum += t3[3 * jj];
                }
                __syncthreads();
            }
            if (tid < 3 * S) {
                const int i = sw.sink0 + g0 + tid / 3;
                const int comp = tid % 3;
                double* o = reinterpret_cast<double*>(a + i);
                o[comp] = L.sink[tid / 3].live ? sum : 0.0;
                if (comp == 0) o[3] = 0.0;
            }
            __syncthreads();   // the sink records are rewritten by the next group
        }
        return;
    }
    for (int i0 = sw.sink0; i0 < sw.sink1; i0 += FUSED_BLOCK) {
        const int i = i0 + tid;
        const bool in = i < sw.sink1;
        const int ic = in ? i : sw.sink0;
        const d4 p = r[ic];
        const bool live = in && sw.gate[ic] > 0;
        const double Ri = EV ? radius[ic] : 0.0;
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int j0 = sw.src0; j0 < sw.src1; j0 += FUSED_TILE) {
            const int nj = min(FUSED_TILE, sw.src1 - j0);
            fused_stage_sources(L, j0, nj, r, radius, sw.gate);
            __syncthreads();
            if (in) {
                // four sources at a time, their terms added in ascending j
                constexpr int TU = 4;
                for (int k0 = 0; k0 < nj; k0 += TU) {
                    FusedSrc q[TU];
                    int jg[TU], ig[TU];
                    double px[TU], py[TU], pz[TU], tx[TU], ty[TU], tz[TU], d[TU];
                    bool val[TU], use[TU];
#pragma unroll
                    for (int u = 0; u < TU; ++u) {
                        val[u] = k0 + u < nj;
                        q[u] = L.src[val[u] ? k0 + u : k0];
                        jg[u] = j0 + k0 + u;
                        ig[u] = i;
                        px[u] = p.x; py[u] = p.y; pz[u] = p.z;
                    }
                    fused_terms<TU, OV>(q, jg, ig, px, py, pz, val, tx, ty, tz, d, use);
#pragma unroll
                    for (int u = 0; u < TU; ++u) {
                        if (val[u]) {
                            ax += tx[u];
                            ay += ty[u];
                            az += tz[u];
                        }
                    }
                    if (EV && live) {
#pragma unroll
                        for (int u = 0; u < TU; ++u)
                            if (use[u]) pair_test(P, E, ctl, i, jg[u], d[u], Ri, q[u].R, r, v, mass, id);
                    }
                }
            }
            __syncthreads();
        }
        if (in) a[i] = d4{live ? ax : 0.0, live ? ay : 0.0, live ? az : 0.0, 0.0};
    }
}

// The three sweeps of one force evaluation (as eval_forces launches them).
// (inlined: a call would copy the kernel's argument block to scratch to pass it by reference)
__device__ __forceinline__ void fused_forces(const FusedArgs& A, const PairTest& P, int stage, const d4* r,
                                             const d4* v, d4* a, FusedLds& L)
{
    EventSink E;
    E.buf = A.ev_buf;
    E.cap = A.ev_cap;
    E.stage = stage;
    E.c_stage = A.c_stage[stage];
    for (int k = 0; k < 3; ++k) {
        const Sweep sw = A.sw[k];
        if (sw.sink1 <= sw.sink0) continue;
        const bool ov = sw.src0 < sw.sink1 && sw.sink0 < sw.src1;
        if (P.enabled) {
            if (ov) fused_sweep<true, true>(sw, r, v, a, A.mass, A.radius, A.id, P, E, A.ctl, L);
            else fused_sweep<false, true>(sw, r, v, a, A.mass, A.radius, A.id, P, E, A.ctl, L);
        } else {
            if (ov) fused_sweep<true, false>(sw, r, v, a, A.mass, A.radius, A.id, P, E, A.ctl, L);
            else fused_sweep<false, false>(sw, r, v, a, A.mass, A.radius, A.id, P, E, A.ctl, L);
        }
    }
}

// ---- small disks: every massive source in one LDS tile ---------------------------------
// The pass that forms a stage state also stages it (position, mass, radius, activity) as the
// next force evaluation's sources, so an evaluation needs no staging pass and no barrier of
// its own before the pair terms; the SI sweep's serial sums (wave 0) run beside the other
// sinks' sums (waves 1-3).  Terms and sums are the same operations in the same order as the
// tiled sweeps above, so the results are identical.

struct SmallPlan {
    int nsi, nm, n;
    bool few;       // few SI sinks: their pair terms are formed by the whole work-group
};

// FusedSrc.pad carries the sink activity (gate > 0) of a staged massive body.
__device__ __forceinline__ void small_stage_src(FusedLds& L, int i, const d4& q, const double* radius,
                                                const int* gate)
{
    const int g = gate[i];
    L.src[i] = FusedSrc{q.x, q.y, q.z, q.w, radius[i], g >= 0 ? 1 : 0, g > 0 ? 1 : 0};
}

// One sink against staged sources [0, nsrc): four at a time, added in ascending j.
template <bool OV, bool EV>
__device__ __forceinline__ void small_sink(const FusedArgs& A, const PairTest& P, const EventSink& E, int i,
                                           int nsrc, const d4* r, const d4* v, d4* a, const FusedLds& L, int nm)
{
    // a massive sink's record is staged among the sources
    const bool staged = i < nm;
    const d4 p = staged ? d4{L.src[i].x, L.src[i].y, L.src[i].z, L.src[i].m} : r[i];
    const bool live = staged ? L.src[i].pad != 0 : A.sw[0].gate[i] > 0;
    const double Ri = EV ? (staged ? L.src[i].R : A.radius[i]) : 0.0;
    double ax = 0.0, ay = 0.0, az = 0.0;
    constexpr int TU = 4;
    for (int k0 = 0; k0 < nsrc; k0 += TU) {
        FusedSrc q[TU];
        int jg[TU], ig[TU];
        double px[TU], py[TU], pz[TU], tx[TU], ty[TU], tz[TU], d[TU];
        bool val[TU], use[TU];
#pragma unroll
        for (int u = 0; u < TU; ++u) {
            val[u] = k0 + u < nsrc;
            q[u] = L.src[val[u] ? k0 + u : k0];
            jg[u] = k0 + u;
            ig[u] = i;
            px[u] = p.x; py[u] = p.y; pz[u] = p.z;
        }
        fused_terms<TU, OV>(q, jg, ig, px, py, pz, val, tx, ty, tz, d, use);
#pragma unroll
        for (int u = 0; u < TU; ++u) {
            if (val[u]) {
                ax += tx[u];
                ay += ty[u];
                az += tz[u];
            }
        }
        if (EV && live) {
#pragma unroll
            for (int u = 0; u < TU; ++u)
                if (use[u]) pair_test(P, E, A.ctl, i, jg[u], d[u], Ri, q[u].R, r, v, A.mass, A.id);
        }
    }
    a[i] = d4{live ? ax : 0.0, live ? ay : 0.0, live ? az : 0.0, 0.0};
}

// One force evaluation of a small disk whose massive sources [0, nm) are staged in L.src;
// ends with a barrier.
template <bool EV>
__device__ __forceinline__ void small_forces(const FusedArgs& A, const PairTest& P, const SmallPlan& SP, int stage,
                                             const d4* r, const d4* v, d4* a, FusedLds& L)
{
    EventSink E;
    E.buf = A.ev_buf;
    E.cap = A.ev_cap;
    E.stage = stage;
    E.c_stage = A.c_stage[stage];
    const int tid = threadIdx.x;
    int first = 0, lane0 = 0, stride = FUSED_BLOCK;   // sinks [first, n) one per work-item
    if (SP.few) {
        // the SI sweep's pair terms, S sinks x nm sources, by every work-item
        const int S = SP.nsi, nj = SP.nm;
        constexpr int TU = 2;
        for (int k0 = 0; k0 < S * nj; k0 += TU * FUSED_BLOCK) {
            FusedSrc q[TU];
            int jg[TU], ig[TU], kk[TU];
            double px[TU], py[TU], pz[TU], tx[TU], ty[TU], tz[TU], d[TU];
            bool val[TU], use[TU];
#pragma unroll
            for (int u = 0; u < TU; ++u) {
                kk[u] = k0 + u * FUSED_BLOCK + tid;
                val[u] = kk[u] < S * nj;
                const int kc = val[u] ? kk[u] : 0;
                ig[u] = kc / nj;
                jg[u] = kc - ig[u] * nj;
                q[u] = L.src[jg[u]];
                px[u] = L.src[ig[u]].x; py[u] = L.src[ig[u]].y; pz[u] = L.src[ig[u]].z;
            }
            fused_terms<TU, true>(q, jg, ig, px, py, pz, val, tx, ty, tz, d, use);
#pragma unroll
            for (int u = 0; u < TU; ++u) {
                if (val[u]) {
                    double* t3 = L.terms + 3 * kk[u];
                    t3[0] = tx[u]; t3[1] = ty[u]; t3[2] = tz[u];
                }
            }
            if (EV) {
#pragma unroll
                for (int u = 0; u < TU; ++u)
                    if (use[u] && L.src[ig[u]].pad)
                        pair_test(P, E, A.ctl, ig[u], jg[u], d[u], L.src[ig[u]].R, q[u].R, r, v, A.mass, A.id);
            }
        }
        __syncthreads();
        first = S; lane0 = 64; stride = FUSED_BLOCK - 64;
        if (tid < 3 * S) {
            // row (sink tid / 3, component tid % 3) in ascending j, the next 16 terms read
            // while the current 16 are added
            const int sk = tid / 3;
            const double* t3 = L.terms + 3 * (sk * nj) + (tid - 3 * sk);
            constexpr int GQ = 16;
            const int ng = nj / GQ;
            double sum = 0.0, x[GQ], y[GQ];
            if (ng > 0) {
#pragma unroll
                for (int k = 0; k < GQ; ++k) x[k] = t3[3 * k];
            }
            for (int g = 0; g < ng; g += 2) {
                if (g + 1 < ng) {
#pragma unroll
                    for (int k = 0; k < GQ; ++k) y[k] = t3[3 * ((g + 1) * GQ + k)];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int k = 0; k < GQ; ++k) sum += x[k];
                __builtin_amdgcn_sched_barrier(0);
                if (g + 1 < ng) {
                    if (g + 2 < ng) {
#pragma unroll
                        for (int k = 0; k < GQ; ++k) x[k] = t3[3 * ((g + 2) * GQ + k)];
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int k = 0; k < GQ; ++k) sum += y[k];
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            for (int jj = ng * GQ; jj < nj; ++jj) sum += t3[3 * jj];
            double* o = reinterpret_cast<double*>(a + sk);
            const int comp = tid - 3 * sk;
            o[comp] = L.src[sk].pad ? sum : 0.0;
            if (comp == 0) o[3] = 0.0;
        }
    }
    if (tid >= lane0) {
        for (int i = first + tid - lane0; i < SP.n; i += stride) {
            if (i < SP.nsi) small_sink<true, EV>(A, P, E, i, SP.nm, r, v, a, L, SP.nm);        // SI x massive
            else if (i < SP.nm) small_sink<false, EV>(A, P, E, i, SP.nsi, r, v, a, L, SP.nm);  // NSI x SI
            else small_sink<false, EV>(A, P, E, i, SP.nm, r, v, a, L, SP.nm);                  // NI x massive
        }
    }
    __syncthreads();
}

// One body of a stage combination (k_stage's work for body i); returns its position record.
template <int FORM>
__device__ __forceinline__ d4 fused_stage_body(const StageArgs& S, int i, double dt)
{
    const d4 yr = S.yr[i];
    const d4 yv = S.yv[i];
    d4 orr = combine<FORM>(S, yr, i, false, dt);
    d4 ovv = combine<FORM>(S, yv, i, true, dt);
    orr.w = yr.w;
    ovv.w = 0.0;
    S.outr[i] = orr;
    S.outv[i] = ovv;
    return orr;
}

// A stage combination of a small disk that also stages the massive bodies' new states as
// the next evaluation's sources; with eject, the ejection predicate of the new states (the
// next step's check at its start) is raised in *eject.
__device__ __forceinline__ void small_stage(int form, const StageArgs& S, int n, int nm, double dt, FusedLds& L,
                                            const FusedArgs& A, int* eject)
{
    for (int i = threadIdx.x; i < n; i += FUSED_BLOCK) {
        d4 q;
        switch (form) {
        case SF_AXPY: q = fused_stage_body<SF_AXPY>(S, i, dt); break;
        case SF_DTSUM: q = fused_stage_body<SF_DTSUM>(S, i, dt); break;
        case SF_ACCUM: q = fused_stage_body<SF_ACCUM>(S, i, dt); break;
        case SF_RK4NP1: q = fused_stage_body<SF_RK4NP1>(S, i, dt); break;
        default: q = fused_stage_body<SF_RKF78NP1>(S, i, dt); break;
        }
        if (i < nm) small_stage_src(L, i, q, A.radius, A.sw[0].gate);
        if (eject && A.id[i] > 0 && A.type[i] != 0 && ejection_name(q, A.ej, A.hc) != RED_EV_NONE) *eject = 1;
    }
}

__device__ void fused_stage(int form, const StageArgs& S, int n, double dt)
{
    for (int i = threadIdx.x; i < n; i += FUSED_BLOCK) {
        switch (form) {
        case SF_AXPY: fused_stage_body<SF_AXPY>(S, i, dt); break;
        case SF_DTSUM: fused_stage_body<SF_DTSUM>(S, i, dt); break;
        case SF_ACCUM: fused_stage_body<SF_ACCUM>(S, i, dt); break;
        case SF_RK4NP1: fused_stage_body<SF_RK4NP1>(S, i, dt); break;
        default: fused_stage_body<SF_RKF78NP1>(S, i, dt); break;
        }
    }
}

// Every loop decision is taken by work-item 0 and read by all through LDS between barriers,
// so the whole work-group leaves the loop together; the loop ends after at most
// A.max_steps steps.
__global__ __launch_bounds__(FUSED_BLOCK) void k_fused_fixed(FusedArgs A, PairTest P)
{
    __shared__ FusedLds L;
    __shared__ StageArgs rows_l[2 * FUSED_MAX_ROWS];
    int& flag = L.flag;
    {
        // the stage arguments of both parities, read once from global memory
        const int words = 2 * A.n_main * (int)(sizeof(StageArgs) / sizeof(int));
        const int* src = reinterpret_cast<const int*>(A.rows);
        int* dst = reinterpret_cast<int*>(rows_l);
        for (int w = threadIdx.x; w < words; w += FUSED_BLOCK) dst[w] = src[w];
        __syncthreads();
    }
    const int tid = threadIdx.x;
    double t = A.t, ps = A.ps;
    int cur = A.cur, steps = 0, reason = 0;
    const bool scan = 0.0 < A.ej || 0.0 < A.hc;
    SmallPlan SP;
    SP.nsi = A.sw[0].sink1;
    SP.nm = A.sw[0].src1;
    SP.n = A.n;
    SP.few = SP.nsi > 0 && SP.nsi <= FUSED_TILE_SINKS && SP.nm > 64 && SP.nsi * SP.nm <= FUSED_TILE;
    // the sweeps as the small path assumes them (barycentric bounds of number_of_bodies.cpp)
    const bool small = SP.nm <= FUSED_TILE && A.sw[0].sink0 == 0 && A.sw[0].src0 == 0 &&
                       A.sw[1].sink0 == SP.nsi && A.sw[1].sink1 == SP.nm && A.sw[1].src0 == 0 &&
                       A.sw[1].src1 == SP.nsi && A.sw[2].sink0 == SP.nm && A.sw[2].sink1 == SP.n &&
                       A.sw[2].src0 == 0 && A.sw[2].src1 == SP.nm;
    if (small) {
        if (tid == 0) {
            L.eject[0] = 0;
            L.eject[1] = 0;
            A.ctl->t = t;
            A.ctl->dt = A.dt;
            A.ctl->attempt = 0;
            A.ctl->err_bits = 0ull;
            A.ctl->nan_flag = 0;
        }
        __syncthreads();
        // y_n staged as the first evaluation's sources, and the first step's ejection check
        for (int i = tid; i < A.n; i += FUSED_BLOCK) {
            const d4 q = A.r[cur][i];
            if (i < SP.nm) small_stage_src(L, i, q, A.radius, A.sw[0].gate);
            if (scan && A.id[i] > 0 && A.type[i] != 0 && ejection_name(q, A.ej, A.hc) != RED_EV_NONE) L.eject[0] = 1;
        }
        __syncthreads();
        for (;;) {
            if (!(t <= A.t_stop)) { reason = RED_FUSED_STOP_TIME; break; }
            if (steps >= A.max_steps) { reason = RED_FUSED_STOP_STEPS; break; }
            if (A.interval <= fabs(ps)) { reason = RED_FUSED_STOP_OUTPUT; break; }
            if (L.eject[steps & 1] != 0) { reason = RED_FUSED_STOP_EJECTION; break; }
            const StageArgs* rows = rows_l + cur * A.n_main;
            int nev = 0;
            // evaluation k - 1 (f0 from y_n, then each stage state), then stage state k
            for (int k = 1; k <= A.n_main; ++k) {
                const d4* fr = k == 1 ? A.r[cur] : rows[k - 2].outr;
                const d4* fv = k == 1 ? A.v[cur] : rows[k - 2].outv;
                if (P.enabled) small_forces<true>(A, P, SP, k - 1, fr, fv, A.a[k - 1], L);
                else small_forces<false>(A, P, SP, k - 1, fr, fv, A.a[k - 1], L);
                if (k == 1 && A.clear_events_after_f0 && tid == 0)
                    __hip_atomic_store(&A.ctl->n_events, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const StageArgs& S = rows[k - 1];
                const bool last = k == A.n_main;
                if (last && tid == 0) {
                    // every pair test of the step is behind the last barrier (with one force
                    // evaluation, f0's records were cleared); the count reaches every
                    // work-item through LDS behind this pass's barrier
                    L.flag = (A.clear_events_after_f0 && A.n_main == 1)
                                 ? 0
                                 : __hip_atomic_load(&A.ctl->n_events, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    L.eject[steps & 1] = 0;   // read at this step's start, behind barriers since
                    A.ctl->t = t + A.dt;      // the next step's event times
                }
                small_stage(A.form[k - 1], S, A.n, SP.nm, A.dt, L, A, last && scan ? &L.eject[(steps + 1) & 1] : nullptr);
                __syncthreads();
                if (last) nev = L.flag;
            }
            t = t + A.dt;
            ps = ps + fabs(A.dt);
            ++steps;
            cur = 1 - cur;
            if (nev > 0) { reason = RED_FUSED_STOP_EVENTS; break; }
        }
        if (tid == 0) *A.out = FusedOut{t, ps, steps, cur, reason, 0};
        return;
    }
    for (;;) {
        if (!(t <= A.t_stop)) { reason = RED_FUSED_STOP_TIME; break; }
        if (steps >= A.max_steps) { reason = RED_FUSED_STOP_STEPS; break; }
        if (A.interval <= fabs(ps)) { reason = RED_FUSED_STOP_OUTPUT; break; }
        if (tid == 0) flag = 0;
        __syncthreads();
        if (scan) {
            for (int i = tid; i < A.n; i += FUSED_BLOCK)
                if (A.id[i] > 0 && A.type[i] != 0 && ejection_name(A.r[cur][i], A.ej, A.hc) != RED_EV_NONE) flag = 1;
        }
        if (tid == 0) {
            A.ctl->t = t;
            A.ctl->dt = A.dt;
            A.ctl->attempt = 0;
            A.ctl->err_bits = 0ull;
            A.ctl->nan_flag = 0;
        }
        __syncthreads();
        if (flag != 0) { reason = RED_FUSED_STOP_EJECTION; break; }
        fused_forces(A, P, 0, A.r[cur], A.v[cur], A.a[0], L);
        __syncthreads();
        if (A.clear_events_after_f0 && tid == 0)
            __hip_atomic_store(&A.ctl->n_events, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const StageArgs* rows = rows_l + cur * A.n_main;
        for (int k = 1; k <= A.n_main; ++k) {
            const StageArgs& S = rows[k - 1];
            fused_stage(A.form[k - 1], S, A.n, A.dt);
            __syncthreads();
            if (k < A.n_main) {
                fused_forces(A, P, k, S.outr, S.outv, A.a[k], L);
                __syncthreads();
            }
        }
        t = t + A.dt;
        ps = ps + fabs(A.dt);
        ++steps;
        cur = 1 - cur;
        if (tid == 0) flag = __hip_atomic_load(&A.ctl->n_events, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const bool events = flag > 0;
        __syncthreads();
        if (events) { reason = RED_FUSED_STOP_EVENTS; break; }
    }
    if (tid == 0) *A.out = FusedOut{t, ps, steps, cur, reason, 0};
}

}  // namespace

hipError_t launch_fused_fixed(const FusedArgs& A, const PairTest& P, hipStream_t s)
{
    hipLaunchKernelGGL(k_fused_fixed, dim3(1), dim3(FUSED_BLOCK), 0, s, A, P);
    return hipGetLastError();
}

}  // namespace red
