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
// The energy kernel is not in the reference: per massive body it writes
// (1/2 m v^2, sum_{j>i, interacting} m_i m_j / r_ij, m r x v) for a fixed-order host sum.
#include "red_internal.hpp"

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
    const double r2 = (q.x * q.x) + (q.y * q.y) + (q.z * q.z);
    int name = RED_EV_NONE;
    if (0.0 < ej && ej * ej < r2) name = RED_EV_EJECTION;
    else if (0.0 < hc && hc * hc > r2) name = RED_EV_HIT_CENTRUM;
    if (name == RED_EV_NONE) return;
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

__global__ __launch_bounds__(256) void k_energy(const d4* __restrict__ r, const d4* __restrict__ v,
                                                const double* __restrict__ mass, const int* __restrict__ id,
                                                int n_massive, int n_si, double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_massive) return;
    double* o = out + 6 * (size_t)i;
    if (id[i] <= 0) {
        for (int c = 0; c < 6; c++) o[c] = 0.0;
        return;
    }
    const d4 q = r[i], w = v[i];
    const double m = mass[i];
    double U = 0.0;
    const int jend = i < n_si ? n_massive : n_si;   // NSI-NSI pairs do not interact
    for (int j = i + 1; j < jend; j++) {
        if (id[j] <= 0) continue;
        const d4 p = r[j];
        const double dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        U += m * mass[j] / sqrt((dx * dx) + (dy * dy) + (dz * dz));
    }
    o[0] = 0.5 * m * ((w.x * w.x) + (w.y * w.y) + (w.z * w.z));
    o[1] = U;
    o[2] = m * (q.y * w.z - q.z * w.y);
    o[3] = m * (q.z * w.x - q.x * w.z);
    o[4] = m * (q.x * w.y - q.y * w.x);
    o[5] = 0.0;
}

// ---- state features of a step start (learned step-size proposer; not in the reference) ----

constexpr int FEAT_SRC = 256;   // sources per wave in the pair pass (4 per lane)

__global__ void k_feat_reset(StepFeat* f)
{
    f->min_q3_bits = 0x7f800000u;   // +inf
    f->max_a_bits = 0u;
    f->max_pert_bits = 0u;
    f->pad = 0u;
}

__device__ __forceinline__ float wave_min(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_max(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// One wave per work item, four waves per block.  Items [0, n0): (SI sink 1..n_si-1,
// 256-source chunk of 1..n_massive-1); [n0, n0+n1): (NSI sink, chunk of SI sources
// 1..n_si-1); then one item per 64 bodies 1..n-1.  Lanes run over sources, so a sink with
// many sources is spread over waves; min and max are exact in any order.  The pair
// quantity is q^3 = d^3 3 m_star / ((m_i + m_j) ((h_i + h_j)/2)^3) with heliocentric h:
// differences in FP64, the ratio in FP32.
__global__ __launch_bounds__(256) void k_step_features(const d4* __restrict__ r, const d4* __restrict__ a,
                                                       const int* __restrict__ id, int n_si, int n_massive, int n,
                                                       long n0, int ch0, long n1, int ch1, long n_items,
                                                       StepFeat* __restrict__ f)
{
    const int lane = threadIdx.x & 63;
    long w = (long)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (w >= n_items) return;
    const d4 star = r[0];
    const double ms = star.w;
    if (w < n0 + n1) {
        int i, src0, src1;
        long c;
        if (w < n0) {
            i = 1 + (int)(w / ch0); c = w % ch0; src0 = 1; src1 = n_massive;
        } else {
            w -= n0;
            i = n_si + (int)(w / ch1); c = w % ch1; src0 = 1; src1 = n_si;
        }
        const int j0 = src0 + (int)c * FEAT_SRC;
        const int j1 = min(src1, j0 + FEAT_SRC);
        float best = __uint_as_float(0x7f800000u);
        if (id[i] > 0) {
            const d4 ri = r[i];
            const double hx = ri.x - star.x, hy = ri.y - star.y, hz = ri.z - star.z;
            const float hi = sqrtf((float)(hx * hx + hy * hy + hz * hz));
            for (int j = j0 + lane; j < j1; j += 64) {
                if (j == i || id[j] <= 0) continue;
                const d4 rj = r[j];
                const double mm = ri.w + rj.w;
                if (!(mm > 0.0)) continue;
                const double dx = rj.x - ri.x, dy = rj.y - ri.y, dz = rj.z - ri.z;
                const double gx = rj.x - star.x, gy = rj.y - star.y, gz = rj.z - star.z;
                const float d2 = (float)(dx * dx + dy * dy + dz * dz);
                const float hj = sqrtf((float)(gx * gx + gy * gy + gz * gz));
                const float hm = 0.5f * (hi + hj);
                const float q3 = d2 * sqrtf(d2) * (float)(3.0 * ms) / ((float)mm * hm * hm * hm);
                best = fminf(best, q3);
            }
        }
        best = wave_min(best);
        if (lane == 0) atomicMin(&f->min_q3_bits, __float_as_uint(best));
        return;
    }
    w -= n0 + n1;
    const int i = 1 + (int)w * 64 + lane;
    float amax = 0.0f, pmax = 0.0f;
    if (i < n && id[i] > 0) {
        const d4 ri = r[i];
        const d4 ai = a[i];
        const double hx = ri.x - star.x, hy = ri.y - star.y, hz = ri.z - star.z;
        const double h2 = hx * hx + hy * hy + hz * hz;
        const double s = -ms / (h2 * sqrt(h2));
        const double kx = s * hx, ky = s * hy, kz = s * hz;
        const double ex = ai.x - kx, ey = ai.y - ky, ez = ai.z - kz;
        const double an = sqrt(ai.x * ai.x + ai.y * ai.y + ai.z * ai.z);
        const double kn = sqrt(kx * kx + ky * ky + kz * kz);
        const double en = sqrt(ex * ex + ey * ey + ez * ez);
        amax = (float)an;
        pmax = kn > 0.0 ? (float)(en / kn) : 0.0f;
    }
    amax = wave_max(amax);
    pmax = wave_max(pmax);
    if (lane == 0) {
        atomicMax(&f->max_a_bits, __float_as_uint(amax));
        atomicMax(&f->max_pert_bits, __float_as_uint(pmax));
    }
}

}  // namespace

hipError_t launch_step_features(const d4* r, const d4* a, const int* id, int n_si, int n_massive, int n,
                                StepFeat* f, hipStream_t s)
{
    hipLaunchKernelGGL(k_feat_reset, dim3(1), dim3(1), 0, s, f);
    RED_HIP_TRY(hipGetLastError());
    const long c0 = n_si > 1 ? n_si - 1 : 0;                 // SI sinks without the star
    const long c1 = n_si > 1 ? n_massive - n_si : 0;         // NSI sinks (need SI sources)
    const int ch0 = n_massive > 1 ? (n_massive - 1 + FEAT_SRC - 1) / FEAT_SRC : 0;
    const int ch1 = n_si > 1 ? (n_si - 1 + FEAT_SRC - 1) / FEAT_SRC : 0;
    const long n0 = c0 * ch0, n1 = c1 * ch1;
    const long nb2 = n > 1 ? (n - 1 + 63) / 64 : 0;
    const long items = n0 + n1 + nb2;
    if (items <= 0) return hipSuccess;
    const long blocks = (items + 3) / 4;
    hipLaunchKernelGGL(k_step_features, dim3((unsigned)blocks), dim3(256), 0, s, r, a, id, n_si, n_massive, n, n0,
                       ch0, n1, ch1, items, f);
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

hipError_t launch_energy(const d4* r, const d4* v, const double* mass, const int* id, int n_massive, int n_si,
                         double* out6, hipStream_t s)
{
    if (n_massive <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_energy, dim3((n_massive + 255) / 256), dim3(256), 0, s, r, v, mass, id, n_massive, n_si,
                       out6);
    return hipGetLastError();
}

}  // namespace red
