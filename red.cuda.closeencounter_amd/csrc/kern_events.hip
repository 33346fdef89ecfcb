// kern_events.hip -- ejection / hit-centrum scan, merge scatter, energy partials (gfx950).
//
// Ejection and hit-centrum replace kernel_pp_disk::check_for_ejection_hit_centrum
// (red.cuda/src/red.cuda/pp_disk.cu:72-142): for every active non-star body,
// |r|^2 > R_ej^2 (with R_ej > 0) is an EJECTION, otherwise |r|^2 < R_hc^2 (R_hc > 0) a
// HIT_CENTRUM; the body's id is negated.  Distances are taken from the frame origin as
// in the reference.  Events are appended atomically and put into body order by the host.
//
// The merge scatter applies the host-computed results of pp_disk::handle_collision_pair
// (pp_disk.cu:1151-1202) in one launch instead of seven small copies per pair.
//
// The energy kernels are not in the reference: per massive body (1/2 m v^2, m r x v), and
// the potential sum_{j>i, interacting} m_i m_j / r_ij in partials per (SI body, source
// chunk), all added on the host in a fixed order.
#include <algorithm>
#include <mutex>

#include "red_internal.hpp"

// Scratch of the allocation-free energy entry point (red_energy_noalloc): the potential
// partials per (SI body, source chunk) that red_energy keeps in a buffer the library
// allocates.  One copy per GPU, held until the process ends; every call on a GPU reuses it,
// so calls that overlap on one GPU race.  Host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device (energy_scratch_ptr below).  It carries
// an initialiser, so the array is stored in the code object's file image like the rest of
// its data: the code object then has no zero-filled segment larger than its file.
static __device__ __align__(256) double scratch[red::ENERGY_SCRATCH_DOUBLES] = {0.0, 0.0, 0.0, 1.0};

namespace red {
namespace {

__global__ __launch_bounds__(256) void k_ejection(const d4* __restrict__ r, const d4* __restrict__ v,
                                                  int* __restrict__ id, const int* __restrict__ type, int n,
                                                  double ej, double hc, Ctl* __restrict__ ctl,
                                                  red_event* __restrict__ buf, int cap)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!(id[i] > 0 && type[i] != 0)) return;
    const d4 q = r[i];
    const int name = ejection_name(q, ej, hc);
    if (name == RED_EV_NONE) return;
    const double r2 = (q.x * q.x) + (q.y * q.y) + (q.z * q.z);
    const int k = atomicAdd(&ctl->n_events_hc, 1);
    if (k < cap) {
        red_event* e = buf + k;
        e->name = name;
        e->id1 = id[0]; e->idx1 = 0;
        e->id2 = id[i]; e->idx2 = i;
        e->attempt = -1; e->stage = -1; e->sink = i; e->source = 0; e->pad = 0;
        e->t = ctl->t;
        e->d = sqrt(r2);
        const d4 r0 = r[0], v0 = v[0], vi = v[i];
        e->r1[0] = r0.x; e->r1[1] = r0.y; e->r1[2] = r0.z;
        e->v1[0] = v0.x; e->v1[1] = v0.y; e->v1[2] = v0.z;
        e->r2[0] = q.x; e->r2[1] = q.y; e->r2[2] = q.z;
        e->v2[0] = vi.x; e->v2[1] = vi.y; e->v2[2] = vi.z;
        for (int c = 0; c < 4; c++) { e->p1[c] = 0.0; e->p2[c] = 0.0; e->ps[c] = 0.0; }
        for (int c = 0; c < 3; c++) { e->rs[c] = 0.0; e->vs[c] = 0.0; }
        for (int c = 0; c < 6; c++) { e->oe1[c] = 0.0; e->oe2[c] = 0.0; e->oes[c] = 0.0; }
    }
    id[i] = -id[i];
}

__global__ void k_apply_updates(const BodyUpdate* __restrict__ u, int nu, d4* __restrict__ r, d4* __restrict__ v,
                                double* __restrict__ mass, double* __restrict__ radius, int* __restrict__ id)
{
    // Updates are applied in list order by a single work-item: a body may appear twice
    // (survivor of one pair, merger of the next) and the later write must win.
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int k = 0; k < nu; k++) {
        const BodyUpdate& b = u[k];
        r[b.idx] = d4{b.r[0], b.r[1], b.r[2], b.r[3]};
        v[b.idx] = d4{b.v[0], b.v[1], b.v[2], 0.0};
        mass[b.idx] = b.mass;
        radius[b.idx] = b.radius;
        id[b.idx] = b.id;
    }
}

// Per massive body: kinetic energy and angular momentum (o[1], the potential, is filled by
// the pair kernel below).
__global__ __launch_bounds__(256) void k_energy(const d4* __restrict__ r, const d4* __restrict__ v,
                                                const double* __restrict__ mass, const int* __restrict__ id,
                                                int n_massive, double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_massive) return;
    double* o = out + 6 * (size_t)i;
    for (int c = 0; c < 6; c++) o[c] = 0.0;
    if (id[i] <= 0) return;
    const d4 q = r[i], w = v[i];
    const double m = mass[i];
    o[0] = 0.5 * m * ((w.x * w.x) + (w.y * w.y) + (w.z * w.z));
    o[2] = m * (q.y * w.z - q.z * w.y);
    o[3] = m * (q.z * w.x - q.x * w.z);
    o[4] = m * (q.x * w.y - q.y * w.x);
}

// Potential energy of the interacting pairs: every pair has a self-interacting body i
// (SI x massive, NSI-NSI pairs do not interact), counted once as (i, j > i).  Block
// (chunk, i) sums m_i m_j / r_ij over sources j of one ECH-source chunk and reduces in LDS
// in a fixed order; the host adds the partials in (i, chunk) order.
constexpr int ECH = 4096;

__global__ __launch_bounds__(256) void k_energy_pairs(const d4* __restrict__ r, const double* __restrict__ mass,
                                                      const int* __restrict__ id, int n_massive,
                                                      double* __restrict__ part)
{
    __shared__ double red[256];
    const int i = blockIdx.y;
    const int nch = gridDim.x;
    const int j0 = blockIdx.x * ECH;
    const int j1 = min(n_massive, j0 + ECH);
    double U = 0.0;
    if (id[i] > 0) {
        const d4 q = r[i];
        const double m = mass[i];
        for (int j = max(j0, i + 1) + threadIdx.x; j < j1; j += blockDim.x) {
            if (id[j] <= 0) continue;
            const d4 p = r[j];
            const double dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
            U += m * mass[j] / sqrt((dx * dx) + (dy * dy) + (dz * dz));
        }
    }
    red[threadIdx.x] = U;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)i * nch + blockIdx.x] = red[0];
}

__device__ __forceinline__ double lane_value(double x, int l)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// red_energy's sums on the device, in the same order: the kinetic energy and angular
// momentum of every massive body in index order (formed here from the state with k_energy's
// expressions, so no per-body partials are stored), the potential partials in (SI body,
// chunk) order, E = T - U.  One wave: lane k forms the terms of body b + k of a block of 64
// (the next block is loaded before the current one is added), then every lane adds the
// block's terms lane by lane.
__global__ __launch_bounds__(64) void k_energy_sum(const d4* __restrict__ r, const d4* __restrict__ v,
                                                   const double* __restrict__ mass, const int* __restrict__ id,
                                                   int n_massive, const double* __restrict__ part, long n_part,
                                                   double* __restrict__ res)
{
    const int lane = threadIdx.x;
    double T = 0.0, lx = 0.0, ly = 0.0, lz = 0.0, U = 0.0;
    struct Body {
        d4 q, w;
        double m;
        bool on;
    };
    auto load = [&](int i) {
        Body b{};
        b.on = i < n_massive && id[i] > 0;
        if (b.on) {
            b.q = r[i];
            b.w = v[i];
            b.m = mass[i];
        }
        return b;
    };
    Body cur = load(lane);
    for (int b = 0; b < n_massive; b += 64) {
        const Body nxt = load(b + 64 + lane);
        double t = 0.0, x = 0.0, y = 0.0, z = 0.0;
        if (cur.on) {
            const d4 q = cur.q, w = cur.w;
            const double m = cur.m;
            t = 0.5 * m * ((w.x * w.x) + (w.y * w.y) + (w.z * w.z));
            x = m * (q.y * w.z - q.z * w.y);
            y = m * (q.z * w.x - q.x * w.z);
            z = m * (q.x * w.y - q.y * w.x);
        }
        const int cnt = min(64, n_massive - b);
        for (int l = 0; l < cnt; ++l) {
            T += lane_value(t, l);
            lx += lane_value(x, l);
            ly += lane_value(y, l);
            lz += lane_value(z, l);
        }
        cur = nxt;
    }
    double u = lane < n_part ? part[lane] : 0.0;
    for (long b = 0; b < n_part; b += 64) {
        const double un = b + 64 + lane < n_part ? part[b + 64 + lane] : 0.0;
        const int cnt = (int)min(64L, n_part - b);
        for (int l = 0; l < cnt; ++l) U += lane_value(u, l);
        u = un;
    }
    if (lane == 0) {
        res[0] = T - U;
        res[1] = lx;
        res[2] = ly;
        res[3] = lz;
    }
}

}  // namespace

hipError_t energy_scratch_ptr(int device, double** out)
{
    static std::mutex m;
    static double* ptrs[64] = {};
    *out = nullptr;
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(m);
    if (!ptrs[device]) {
        int prev = 0;
        RED_HIP_TRY(hipGetDevice(&prev));
        if (prev != device) RED_HIP_TRY(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev != device) (void)hipSetDevice(prev);
        RED_HIP_TRY(e);
        ptrs[device] = static_cast<double*>(p);
    }
    *out = ptrs[device];
    return hipSuccess;
}

hipError_t launch_energy_noalloc(const d4* r, const d4* v, const double* mass, const int* id, int n_massive,
                                 int n_si, double* scratch_dev, double* res, hipStream_t s)
{
    const int nm = std::max(n_massive, 0);
    const long n_part = nm > 0 && n_si > 0 ? (long)n_si * energy_chunks(nm) : 0;
    if (n_part > 0) {
        hipLaunchKernelGGL(k_energy_pairs, dim3(energy_chunks(nm), n_si), dim3(256), 0, s, r, mass, id, nm,
                           scratch_dev);
        RED_HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_energy_sum, dim3(1), dim3(64), 0, s, r, v, mass, id, nm, scratch_dev, n_part, res);
    return hipGetLastError();
}

hipError_t launch_ejection_scan(const d4* r, const d4* v, int* id, const int* type, int n, double ej, double hc,
                                Ctl* ctl, red_event* buf, int cap, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ejection, dim3((n + 255) / 256), dim3(256), 0, s, r, v, id, type, n, ej, hc, ctl, buf, cap);
    return hipGetLastError();
}

hipError_t launch_apply_updates(const BodyUpdate* u, int nu, d4* r, d4* v, double* mass, double* radius, int* id,
                                hipStream_t s)
{
    if (nu <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_apply_updates, dim3(1), dim3(64), 0, s, u, nu, r, v, mass, radius, id);
    return hipGetLastError();
}

int energy_chunks(int n_massive) { return (n_massive + ECH - 1) / ECH; }

hipError_t launch_energy(const d4* r, const d4* v, const double* mass, const int* id, int n_massive, int n_si,
                         double* out6, double* part, hipStream_t s)
{
    if (n_massive <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_energy, dim3((n_massive + 255) / 256), dim3(256), 0, s, r, v, mass, id, n_massive, out6);
    RED_HIP_TRY(hipGetLastError());
    if (n_si <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_energy_pairs, dim3(energy_chunks(n_massive), n_si), dim3(256), 0, s, r, mass, id, n_massive,
                       part);
    return hipGetLastError();
}

}  // namespace red
