// ensemble.hip -- many independent disks of uneven size integrated together (gfx950).
//
// All disks live in one set of buffers (bodies of disk d occupy [b0_d, b1_d)).  Every
// stage of an attempt is ONE launch over all disks: the force launch walks a list of
// tiles (sink range x source range of one disk), the stage-sum and error launches walk
// all bodies with a per-body disk index, and the whole attempt is captured once in a HIP
// graph and replayed.  Each disk keeps its own time, trial step and attempt count; the
// host reads one error maximum per disk per attempt and runs the reference's controller
// for every disk with the same arithmetic as a single disk, so an ensemble member
// follows exactly the steps it would take alone (int_rungekutta8.cu:704-797 and
// siblings).  Pair tests (collisions, ejections) are not applied inside an ensemble.
//
// Force tiles (FP64, the reference's operation order and ascending-j summation):
//   kind 0: up to `block` sinks, one work-item each; sources staged through LDS in
//           chunks of `jtile` bodies;
//   kind 1: one sink (the star against its disk): the block computes `jtile` pair terms in
//           parallel into LDS and three lanes add them in source order.
// The tile shape (block, jtile) is picked by red_ensemble_autotune from measured launches.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "host_tools.hpp"
#include "red_internal.hpp"
#include "scheme.hpp"

using namespace red;

namespace {

struct EnsTile {
    int sink0, sink1, src0, src1;
    int disk, kind;
};

constexpr int MAX_JTILE = 256;

__device__ __forceinline__ void fp64_term(const d4& q, double xi, double yi, double zi, bool use, double& tx,
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
}

// stage 0 runs only for disks starting a new step (bit 1), the other stages for every
// disk in the attempt (bit 0)
__global__ __launch_bounds__(256) void k_ens_force(const EnsTile* __restrict__ tiles, const d4* __restrict__ r,
                                                   d4* __restrict__ a, const int* __restrict__ flags, int need_bit,
                                                   int jtile)
{
    __shared__ d4 src[MAX_JTILE];
    __shared__ double term[2][3][MAX_JTILE];
    const EnsTile T = tiles[blockIdx.x];
    if (!(flags[T.disk] & need_bit)) return;
    if (T.kind == 0) {
        const int i = T.sink0 + threadIdx.x;
        const bool in = i < T.sink1;
        const d4 p = r[in ? i : T.sink0];
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int j0 = T.src0; j0 < T.src1; j0 += jtile) {
            const int nj = min(jtile, T.src1 - j0);
            __syncthreads();
            for (int k = threadIdx.x; k < nj; k += blockDim.x) src[k] = r[j0 + k];
            __syncthreads();
            for (int k = 0; k < nj; ++k) {
                double tx, ty, tz;
                fp64_term(src[k], p.x, p.y, p.z, (j0 + k) != i, tx, ty, tz);
                ax += tx;
                ay += ty;
                az += tz;
            }
        }
        if (in) a[i] = d4{ax, ay, az, 0.0};
    } else {
        // One sink against many sources (a disk's star): its three sums are serial chains in
        // ascending j, the kernel's critical path.  With more than one wave, waves 1.. form
        // the terms of chunk c + 1 into one half of a double-buffered LDS tile while lanes
        // 0-2 of wave 0 add chunk c, the next 16 terms read while the current 16 are added;
        // with one wave it forms, then adds, each chunk.
        const int i = T.sink0;
        const d4 p = r[i];
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
                fp64_term(r[j0 + k], p.x, p.y, p.z, (j0 + k) != i, tx, ty, tz);
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
            auto fetch = [&](int ch) {
                const int j0 = T.src0 + ch * jtile;
                const int nj = min(jtile, T.src1 - j0);
#pragma unroll
                for (int m = 0; m < MS; ++m) {
                    const int k = tp + m * np;
                    if (k < nj) q[m] = r[j0 + k];
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
                        fp64_term(q[m], p.x, p.y, p.z, (j0 + k) != i, tx, ty, tz);
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
        if (c < 3) reinterpret_cast<double*>(a + i)[c] = acc;
        if (c == 3) reinterpret_cast<double*>(a + i)[3] = 0.0;
    }
}

// y <- y_(n+1) for the bodies of disks whose last attempt was accepted (bit 2)
__global__ __launch_bounds__(256) void k_ens_accept(d4* __restrict__ yr, d4* __restrict__ yv,
                                                    const d4* __restrict__ outr, const d4* __restrict__ outv,
                                                    const int* __restrict__ disk_of, const int* __restrict__ flags,
                                                    int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !(flags[disk_of[i]] & 4)) return;
    yr[i] = outr[i];
    yv[i] = outv[i];
}

__global__ void k_ens_reset(unsigned long long* err, int* nan, int n)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d < n) { err[d] = 0ull; nan[d] = 0; }
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
    std::vector<int> b0, b1, nsi, nm;
    std::vector<int> disk_of_h;
    // device
    d4 *yr = nullptr, *yv = nullptr, *outr = nullptr, *outv = nullptr;
    std::vector<d4*> sr, sv, acc;
    int* disk_of = nullptr;
    unsigned char* err_on = nullptr;
    double* d_dt = nullptr;
    int* d_flags = nullptr;
    unsigned long long* d_err = nullptr;
    int* d_nan = nullptr;
    EnsTile* d_tiles = nullptr;
    int n_tiles = 0;
    int block = 256, jtile = 256;
    // host per-disk controller state
    std::vector<double> t, dt_try, dt_did, last_err;
    std::vector<int> iter, tried, failed, passed, phase;   // phase 0: new step, 1: retry, 2: done
    std::vector<int> accepted_prev;
    double t_stop = 0.0;
    double* h_dt = nullptr;                // pinned staging
    int* h_flags = nullptr;
    unsigned long long* h_err = nullptr;
    int* h_nan = nullptr;
    hipGraphExec_t graph = nullptr;
    bool use_graph = true;

    int fail(int code, const std::string& m) { err = m; return code; }
    int hipfail(hipError_t e, const char* where) { err = std::string(where) + ": " + hipGetErrorString(e); return RED_ERR_HIP; }
};

#define ENS_HIP(E, expr)                                        \
    do {                                                        \
        hipError_t _e = (expr);                                 \
        if (_e != hipSuccess) return (E)->hipfail(_e, #expr);   \
    } while (0)

namespace {

// Tiles of every disk's three sweeps (number_of_bodies.cpp:196-258 per disk), largest first.
std::vector<EnsTile> build_tiles(const red_ensemble* E, int block)
{
    std::vector<EnsTile> tiles;
    for (int d = 0; d < E->nd; d++) {
        const int b0 = E->b0[d], b1 = E->b1[d], si = E->nsi[d], m = E->nm[d];
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

hipError_t ens_force(red_ensemble* E, int stage, const d4* r)
{
    if (E->n_tiles == 0) return hipSuccess;
    hipLaunchKernelGGL(k_ens_force, dim3(E->n_tiles), dim3(E->block), 0, E->stream, E->d_tiles, r, E->acc[stage],
                       E->d_flags, stage == 0 ? 2 : 1, E->jtile);
    return hipGetLastError();
}

// One attempt for every disk (the body of the captured graph).
hipError_t ens_attempt(red_ensemble* E)
{
    const Scheme& S = E->S;
    const dim3 gb((E->n + 255) / 256), bb(256);
    hipLaunchKernelGGL(k_ens_accept, gb, bb, 0, E->stream, E->yr, E->yv, E->outr, E->outv, E->disk_of, E->d_flags, E->n);
    RED_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_ens_reset, dim3((E->nd + 255) / 256), dim3(256), 0, E->stream, E->d_err, E->d_nan, E->nd);
    RED_HIP_TRY(hipGetLastError());
    RED_HIP_TRY(ens_force(E, 0, E->yr));
    for (int r = 1; r < S.n_main; r++) {
        RED_HIP_TRY(launch_stage(S.rows[r].form, ens_args(E, S.rows[r], E->sr[r], E->sv[r]), E->stream));
        RED_HIP_TRY(ens_force(E, r, E->sr[r]));
    }
    RED_HIP_TRY(launch_stage(S.np1.form, ens_args(E, S.np1, E->outr, E->outv), E->stream));
    if (E->prm.adaptive && S.err_form >= 0) {
        for (int r = S.n_main; r < S.n_stage; r++) {
            if (S.extra_on_yout) {
                RED_HIP_TRY(ens_force(E, r, E->outr));
            } else {
                RED_HIP_TRY(launch_stage(S.rows[r].form, ens_args(E, S.rows[r], E->sr[r], E->sv[r]), E->stream));
                RED_HIP_TRY(ens_force(E, r, E->sr[r]));
            }
        }
        RED_HIP_TRY(launch_err(S.err_form, ens_args(E, S.err, nullptr, nullptr), E->stream));
    }
    return hipSuccess;
}

void destroy_graph(red_ensemble* E)
{
    SetupLock lk;
    if (E->graph) (void)hipGraphExecDestroy(E->graph);
    E->graph = nullptr;
}

hipError_t run_attempt(red_ensemble* E)
{
    if (!E->use_graph) return ens_attempt(E);
    if (!E->graph) {
        SetupLock lk;
        hipGraph_t g = nullptr;
        RED_HIP_TRY(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = ens_attempt(E);
        hipError_t e2 = hipStreamEndCapture(E->stream, &g);
        if (e != hipSuccess) { if (g) (void)hipGraphDestroy(g); return e; }
        RED_HIP_TRY(e2);
        e = hipGraphInstantiate(&E->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        RED_HIP_TRY(e);
    }
    return hipGraphLaunch(E->graph, E->stream);
}

int upload_tiles(red_ensemble* E)
{
    SetupLock lk;
    std::vector<EnsTile> tiles = build_tiles(E, E->block);
    if (E->d_tiles) (void)hipFree(E->d_tiles);
    E->d_tiles = nullptr;
    E->n_tiles = (int)tiles.size();
    ENS_HIP(E, dmalloc(&E->d_tiles, tiles.size()));
    if (!tiles.empty())
        ENS_HIP(E, stream_copy(E->stream, E->d_tiles, tiles.data(), sizeof(EnsTile) * tiles.size(), hipMemcpyHostToDevice));
    destroy_graph(E);
    return RED_OK;
}

}  // namespace

extern "C" {

int red_ensemble_destroy(red_ensemble* E)
{
    if (!E) return RED_OK;
    SetupLock lk;
    destroy_graph(E);
    if (E->stream) (void)hipStreamSynchronize(E->stream);
    for (void* p : {(void*)E->yr, (void*)E->yv, (void*)E->outr, (void*)E->outv, (void*)E->disk_of, (void*)E->err_on,
                    (void*)E->d_dt, (void*)E->d_flags, (void*)E->d_err, (void*)E->d_nan, (void*)E->d_tiles})
        (void)hipFree(p);
    for (auto p : E->sr) (void)hipFree(p);
    for (auto p : E->sv) (void)hipFree(p);
    for (auto p : E->acc) (void)hipFree(p);
    for (void* p : {(void*)E->h_dt, (void*)E->h_flags, (void*)E->h_err, (void*)E->h_nan}) if (p) (void)hipHostFree(p);
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
    // no pair tests, ejection or hit-centrum checks run inside an ensemble: asking for them
    // is refused rather than ignored
    if (p->radii_enhance_factor > 0.0 || p->encounter_hill_factor > 0.0 || p->close_encounter ||
        p->ejection > 0.0 || p->hit_centrum > 0.0)
        return RED_ERR_UNSUPPORTED;
    red_ensemble* E = new red_ensemble();
    E->prm = *p;
    E->S = make_scheme(p->integrator, p->adaptive != 0);
    E->nd = n_disks;
    E->use_graph = p->use_graph != 0;
    std::vector<d4> r, v;
    std::vector<unsigned char> eon;
    for (int d = 0; d < n_disks; d++) {
        const red_bodies& b = disks[d];
        int nb = 0;
        for (int k = 0; k < 7; k++) nb += b.counts[k];
        const int si = b.counts[0] + b.counts[1] + b.counts[2] + b.counts[3];
        const int m = si + b.counts[4] + b.counts[5];
        E->b0.push_back(E->n);
        E->b1.push_back(E->n + nb);
        E->nsi.push_back(si);
        E->nm.push_back(m);
        for (int i = 0; i < nb; i++) {
            r.push_back(d4{b.r[3 * i], b.r[3 * i + 1], b.r[3 * i + 2], b.mass[i]});
            v.push_back(d4{b.v[3 * i], b.v[3 * i + 1], b.v[3 * i + 2], 0.0});
            E->disk_of_h.push_back(d);
            eon.push_back((unsigned char)(p->error_check_for_tp ? 1 : (i < m ? 1 : 0)));
        }
        E->n += nb;
    }
    *out = E;
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
    ENS_HIP(E, dmalloc(&E->d_dt, E->nd));
    ENS_HIP(E, dmalloc(&E->d_flags, E->nd));
    ENS_HIP(E, dmalloc(&E->d_err, E->nd));
    ENS_HIP(E, dmalloc(&E->d_nan, E->nd));
    ENS_HIP(E, hipHostMalloc((void**)&E->h_dt, sizeof(double) * E->nd, hipHostMallocDefault));
    ENS_HIP(E, hipHostMalloc((void**)&E->h_flags, sizeof(int) * E->nd, hipHostMallocDefault));
    ENS_HIP(E, hipHostMalloc((void**)&E->h_err, sizeof(unsigned long long) * E->nd, hipHostMallocDefault));
    ENS_HIP(E, hipHostMalloc((void**)&E->h_nan, sizeof(int) * E->nd, hipHostMallocDefault));
    if (E->n > 0) {
        ENS_HIP(E, stream_copy(E->stream, E->yr, r.data(), sizeof(d4) * E->n, hipMemcpyHostToDevice));
        ENS_HIP(E, stream_copy(E->stream, E->yv, v.data(), sizeof(d4) * E->n, hipMemcpyHostToDevice));
        ENS_HIP(E, stream_copy(E->stream, E->outr, r.data(), sizeof(d4) * E->n, hipMemcpyHostToDevice));
        ENS_HIP(E, stream_copy(E->stream, E->outv, v.data(), sizeof(d4) * E->n, hipMemcpyHostToDevice));
        ENS_HIP(E, stream_copy(E->stream, E->disk_of, E->disk_of_h.data(), sizeof(int) * E->n, hipMemcpyHostToDevice));
        ENS_HIP(E, stream_copy(E->stream, E->err_on, eon.data(), E->n, hipMemcpyHostToDevice));
    }
    const double dt0 = p->dt_day * tools::GAUSS;   // integrator.cu:29
    E->t.assign(E->nd, 0.0); E->dt_try.assign(E->nd, dt0); E->dt_did.assign(E->nd, 0.0);
    E->last_err.assign(E->nd, 0.0);
    E->iter.assign(E->nd, 0); E->tried.assign(E->nd, 0); E->failed.assign(E->nd, 0); E->passed.assign(E->nd, 0);
    E->phase.assign(E->nd, 0); E->accepted_prev.assign(E->nd, 0);
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

// Integrates every disk with the reference loop condition `while (t <= t_stop) step()`
// (main.cu:176) until all disks are done or max_attempts attempts have been launched.
int red_ensemble_advance(red_ensemble* E, double t_stop, int max_attempts, red_ensemble_info* info)
{
    if (!E) return RED_ERR_ARG;
    const Scheme& S = E->S;
    const bool adaptive = E->prm.adaptive && S.err_form >= 0;
    int attempts = 0, active = 0;
    const auto t_begin = std::chrono::steady_clock::now();
    // a disk done against an earlier t_stop continues (a run to t1, then to t2, steps as one
    // run to t2); a disk left in a retry by max_attempts resumes it
    for (int d = 0; d < E->nd; d++)
        if (E->phase[d] == 2) E->phase[d] = 0;
    for (;;) {
        active = 0;
        for (int d = 0; d < E->nd; d++) {
            int f = 0;
            if (E->accepted_prev[d]) f |= 4;
            if (E->phase[d] == 0 && !(E->t[d] <= t_stop)) E->phase[d] = 2;
            if (E->phase[d] == 0) { f |= 1 | 2; E->iter[d] = 0; }
            else if (E->phase[d] == 1) f |= 1;
            if (f & 1) { active++; E->dt_did[d] = E->dt_try[d]; }
            E->h_flags[d] = f;
            E->h_dt[d] = E->dt_try[d];
        }
        if (active == 0 || attempts >= max_attempts) break;
        ENS_HIP(E, hipMemcpyAsync(E->d_flags, E->h_flags, sizeof(int) * E->nd, hipMemcpyHostToDevice, E->stream));
        ENS_HIP(E, hipMemcpyAsync(E->d_dt, E->h_dt, sizeof(double) * E->nd, hipMemcpyHostToDevice, E->stream));
        ENS_HIP(E, run_attempt(E));
        ENS_HIP(E, hipMemcpyAsync(E->h_err, E->d_err, sizeof(unsigned long long) * E->nd, hipMemcpyDeviceToHost, E->stream));
        ENS_HIP(E, hipMemcpyAsync(E->h_nan, E->d_nan, sizeof(int) * E->nd, hipMemcpyDeviceToHost, E->stream));
        ENS_HIP(E, hipStreamSynchronize(E->stream));
        attempts++;
        for (int d = 0; d < E->nd; d++) {
            E->accepted_prev[d] = 0;
            if (!(E->h_flags[d] & 1)) continue;
            bool accept = true;
            if (adaptive) {
                if (E->h_nan[d]) return E->fail(RED_ERR_NAN, "non-finite error estimate in disk " + std::to_string(d));
                double m;
                memcpy(&m, &E->h_err[d], 8);
                const double max_err = tools::controller_update(&E->dt_try[d], S.lambda, m, E->prm.tolerance, S.exponent);
                E->last_err[d] = max_err;
                accept = !(max_err > E->prm.tolerance);
            }
            E->iter[d]++;
            if (accept) {
                E->tried[d] += E->iter[d];
                E->failed[d] += E->iter[d] - 1;
                E->passed[d]++;
                E->t[d] += E->dt_did[d];
                E->phase[d] = 0;
                E->accepted_prev[d] = 1;
            } else {
                E->phase[d] = 1;
                if (E->iter[d] >= 1000) return E->fail(RED_ERR_STEP_LIMIT, "disk " + std::to_string(d) + " did not converge");
            }
        }
    }
    // apply the last accepted results so that the state buffers hold y
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
    if (info) {
        info->supersteps = attempts;
        int still = 0;
        for (int d = 0; d < E->nd; d++) still += E->phase[d] == 1 || (E->phase[d] == 0 && E->t[d] <= t_stop);
        info->active_disks = still;
        info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_begin).count();
    }
    return RED_OK;
}

int red_ensemble_get_state(red_ensemble* E, int disk, double* r, double* v)
{
    if (!E || disk < 0 || disk >= E->nd) return RED_ERR_ARG;
    const int b0 = E->b0[disk], nb = E->b1[disk] - b0;
    std::vector<d4> R(std::max(nb, 1)), V(std::max(nb, 1));
    ENS_HIP(E, hipStreamSynchronize(E->stream));
    if (nb > 0) {
        ENS_HIP(E, stream_copy(E->stream, R.data(), E->yr + b0, sizeof(d4) * nb, hipMemcpyDeviceToHost));
        ENS_HIP(E, stream_copy(E->stream, V.data(), E->yv + b0, sizeof(d4) * nb, hipMemcpyDeviceToHost));
    }
    for (int i = 0; i < nb; i++) {
        if (r) { r[3 * i] = R[i].x; r[3 * i + 1] = R[i].y; r[3 * i + 2] = R[i].z; }
        if (v) { v[3 * i] = V[i].x; v[3 * i + 1] = V[i].y; v[3 * i + 2] = V[i].z; }
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

// Times `reps` force evaluations of all disks for every (block, jtile) in {64,128,256}^2
// on the current state, keeps the fastest and returns the 9 timings (ms per evaluation,
// row-major block x jtile).
int red_ensemble_autotune(red_ensemble* E, int reps, int* best_block, int* best_jtile, double ms[9])
{
    if (!E || reps <= 0) return RED_ERR_ARG;
    const int choices[3] = {64, 128, 256};
    std::vector<int> flags(E->nd, 3);
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
            // stage 0 recomputes f0 of the current state: identical values, so a retry
            // that reuses f0 is unaffected
            ENS_HIP(E, ens_force(E, 0, E->yr));   // warm-up
            ENS_HIP(E, hipEventRecord(a, E->stream));
            for (int k = 0; k < reps; k++) ENS_HIP(E, ens_force(E, 0, E->yr));
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
    if (best_block) *best_block = bb;
    if (best_jtile) *best_jtile = bj;
    return RED_OK;
}

}  // extern "C"
