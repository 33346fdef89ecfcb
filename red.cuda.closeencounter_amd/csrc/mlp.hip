// mlp.hip -- the learned step-size proposer's network: 12 -> 64 -> 64 -> 1 with GELU
// (erf form), BF16 weights and FP32 accumulation; training and inference kernels (gfx950).
//
// Not in the reference.  The network maps a 12-feature summary of the integrator's recent
// steps and of the current state to the log-change of the tolerance-matching step size
// (red_step proposes dt from it; the embedded error test still accepts or rejects every
// attempt and the reference controller computes every retry).
//
// Parameters: one FP32 master vector and its BF16 image (round to nearest even),
//   W1[64][12] b1[64] W2[64][64] b2[64] W3[64] b3          (5057 values)
// plus the input standardisation z = clip((x - mean) * inv_std, -8, 8) kept with the model,
// applied in-kernel so training and inference see identical inputs (the clip bounds what an
// input far outside the training records can do to the output).
// Forward (per sample, FP32, fused multiply-adds in index order):
//   p1[u] = b1[u] + sum_f W1[u][f] z[f];     h1 = gelu(p1)
//   p2[u] = b2[u] + sum_v W2[u][v] h1[v];    h2 = gelu(p2)
//   y     = b3    + sum_u W3[u] h2[u]
// Loss: mean squared error, or the pinball (quantile) loss of level tau,
//   L = mean(max(tau (t - y), (tau - 1) (t - y))),
// whose minimiser is the tau-quantile of the target given the input.  Gradients are taken
// at the BF16 weights
// (the FP32 master receives them), per 32-sample block, summed in block order: results
// are deterministic.  Adam updates the master copy and re-rounds the BF16 image.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "red_internal.hpp"

namespace {

using red::MLP_NF;
using red::MLP_NH;
using red::MLP_NPARAM;
constexpr int NF = MLP_NF, NH = MLP_NH, NPARAM = MLP_NPARAM;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NH * NF, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH;
static_assert(OFF_B3 + 1 == NPARAM, "parameter layout");
constexpr int SB = 32;         // samples per block
constexpr int W2S = NH + 2;    // padded W2 row (BF16 elements): rows start in different LDS banks
constexpr int HS = NH + 1;     // padded activation row (floats)
constexpr float Z_CLIP = 8.0f; // standardised inputs are clipped to [-8, 8]

__device__ __forceinline__ float bf(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }

__device__ __forceinline__ unsigned short to_bf16(float f)
{
    unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40);   // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);                                                  // nearest, ties to even
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x)
{
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

struct Lds {
    unsigned short w1[NH * NF];
    unsigned short w2[NH * W2S];
    float b1[NH], b2[NH], w3[NH], b3;
    float z[SB][NF];
    float p1[SB][HS], h1[SB][HS], p2[SB][HS], h2[SB][HS];
    float out[SB], dy[SB];
};

// Loads the BF16 weights and the standardised inputs of samples [s0, s0 + ns).
__device__ void load_block(Lds& L, const unsigned short* __restrict__ p, const float* __restrict__ norm,
                           const float* X, int s0, int ns)
{
    for (int k = threadIdx.x; k < NH * NF; k += blockDim.x) L.w1[k] = p[OFF_W1 + k];
    for (int k = threadIdx.x; k < NH * NH; k += blockDim.x) L.w2[(k / NH) * W2S + k % NH] = p[OFF_W2 + k];
    for (int k = threadIdx.x; k < NH; k += blockDim.x) {
        L.b1[k] = bf(p[OFF_B1 + k]);
        L.b2[k] = bf(p[OFF_B2 + k]);
        L.w3[k] = bf(p[OFF_W3 + k]);
    }
    if (threadIdx.x == 0) L.b3 = bf(p[OFF_B3]);
    for (int k = threadIdx.x; k < ns * NF; k += blockDim.x) {
        const int s = k / NF, f = k % NF;
        const float z = (X[(size_t)(s0 + s) * NF + f] - norm[f]) * norm[NF + f];
        L.z[s][f] = fminf(fmaxf(z, -Z_CLIP), Z_CLIP);
    }
    __syncthreads();
}

__device__ void forward(Lds& L, int ns)
{
    for (int k = threadIdx.x; k < ns * NH; k += blockDim.x) {
        const int s = k / NH, u = k % NH;
        float acc = L.b1[u];
        for (int f = 0; f < NF; f++) acc = fmaf(bf(L.w1[u * NF + f]), L.z[s][f], acc);
        L.p1[s][u] = acc;
        L.h1[s][u] = gelu(acc);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < ns * NH; k += blockDim.x) {
        const int s = k / NH, u = k % NH;
        float acc = L.b2[u];
        for (int v = 0; v < NH; v++) acc = fmaf(bf(L.w2[u * W2S + v]), L.h1[s][v], acc);
        L.p2[s][u] = acc;
        L.h2[s][u] = gelu(acc);
    }
    __syncthreads();
    for (int s = threadIdx.x; s < ns; s += blockDim.x) {
        float acc = L.b3;
        for (int u = 0; u < NH; u++) acc = fmaf(L.w3[u], L.h2[s][u], acc);
        L.out[s] = acc;
    }
    __syncthreads();
}

// The forward pass of one sample with the weights read straight from global memory into
// registers (one row of W1 and W2 per unit, issued before the inputs are known so that the
// loads overlap the feature reduction): the same multiply-adds in the same order as
// forward(), so the output is bit-identical, with a few hundred bytes of LDS instead of a
// 32-sample tile (the step-start kernel carries this LDS in every block of its grid).
struct Lds1 {
    float z[NF];
    float h1[NH], h2[NH], w3[NH];
    float out;
};

struct RowW {
    uint4 w2[NH / 8];          // W2[u][0..63], 8 BF16 per uint4
    uint2 w1[NF / 4];          // W1[u][0..11], 4 BF16 per uint2
    unsigned short b1, b2, w3;
};
static_assert(NH % 8 == 0 && NF % 4 == 0 && OFF_W2 % 8 == 0 && OFF_W1 % 4 == 0, "row loads");

__device__ __forceinline__ void load_row(RowW& R, const unsigned short* __restrict__ p, int u)
{
    const uint4* w2 = reinterpret_cast<const uint4*>(p + OFF_W2 + u * NH);
#pragma unroll
    for (int k = 0; k < NH / 8; k++) R.w2[k] = w2[k];
    const uint2* w1 = reinterpret_cast<const uint2*>(p + OFF_W1 + u * NF);
#pragma unroll
    for (int k = 0; k < NF / 4; k++) R.w1[k] = w1[k];
    R.b1 = p[OFF_B1 + u];
    R.b2 = p[OFF_B2 + u];
    R.w3 = p[OFF_W3 + u];
}

__device__ __forceinline__ float bf_lo(unsigned v) { return __uint_as_float(v << 16); }
__device__ __forceinline__ float bf_hi(unsigned v) { return __uint_as_float(v & 0xffff0000u); }

// All threads call it; threads u < NH hold row u in R (loaded by load_row).
__device__ void forward_one(Lds1& S, const RowW& R, const unsigned short* __restrict__ p,
                            const float* __restrict__ norm, const float* x)
{
    const int u = threadIdx.x;
    if (u < NF) {
        const float z = (x[u] - norm[u]) * norm[NF + u];
        S.z[u] = fminf(fmaxf(z, -Z_CLIP), Z_CLIP);
    }
    __syncthreads();
    if (u < NH) {
        float acc = bf(R.b1);
#pragma unroll
        for (int k = 0; k < NF / 4; k++) {
            acc = fmaf(bf_lo(R.w1[k].x), S.z[4 * k + 0], acc);
            acc = fmaf(bf_hi(R.w1[k].x), S.z[4 * k + 1], acc);
            acc = fmaf(bf_lo(R.w1[k].y), S.z[4 * k + 2], acc);
            acc = fmaf(bf_hi(R.w1[k].y), S.z[4 * k + 3], acc);
        }
        S.h1[u] = gelu(acc);
    }
    __syncthreads();
    if (u < NH) {
        float acc = bf(R.b2);
#pragma unroll
        for (int k = 0; k < NH / 8; k++) {
            const uint4 q = R.w2[k];
            acc = fmaf(bf_lo(q.x), S.h1[8 * k + 0], acc);
            acc = fmaf(bf_hi(q.x), S.h1[8 * k + 1], acc);
            acc = fmaf(bf_lo(q.y), S.h1[8 * k + 2], acc);
            acc = fmaf(bf_hi(q.y), S.h1[8 * k + 3], acc);
            acc = fmaf(bf_lo(q.z), S.h1[8 * k + 4], acc);
            acc = fmaf(bf_hi(q.z), S.h1[8 * k + 5], acc);
            acc = fmaf(bf_lo(q.w), S.h1[8 * k + 6], acc);
            acc = fmaf(bf_hi(q.w), S.h1[8 * k + 7], acc);
        }
        S.h2[u] = gelu(acc);
        S.w3[u] = bf(R.w3);
    }
    __syncthreads();
    if (u == 0) {
        float acc = bf(p[OFF_B3]);
        for (int k = 0; k < NH; k++) acc = fmaf(S.w3[k], S.h2[k], acc);
        S.out = acc;
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_mlp_predict(const unsigned short* __restrict__ p, const float* __restrict__ norm,
                                                     const float* __restrict__ X, float* __restrict__ Y, int n)
{
    __shared__ Lds L;
    const int s0 = blockIdx.x * SB;
    const int ns = min(SB, n - s0);
    load_block(L, p, norm, X, s0, ns);
    forward(L, ns);
    for (int s = threadIdx.x; s < ns; s += blockDim.x) Y[s0 + s] = L.out[s];
}

// Per-block gradient sums of the mean loss over n samples (the 1/n is applied here) into
// G[block][NPARAM] and the block's sum of per-sample losses into Lsum[block].
__global__ __launch_bounds__(256) void k_mlp_grad(const unsigned short* __restrict__ p, const float* __restrict__ norm,
                                                  const float* __restrict__ X, const float* __restrict__ T, int n,
                                                  int pinball, float tau, float* __restrict__ G,
                                                  float* __restrict__ Lsum)
{
    __shared__ Lds L;
    const int s0 = blockIdx.x * SB;
    const int ns = min(SB, n - s0);
    load_block(L, p, norm, X, s0, ns);
    forward(L, ns);
    const float inv_n = 1.0f / (float)n;
    for (int s = threadIdx.x; s < ns; s += blockDim.x) {
        const float e = T[s0 + s] - L.out[s];
        L.dy[s] = pinball ? (e > 0.0f ? -tau : 1.0f - tau) * inv_n : 2.0f * inv_n * (L.out[s] - T[s0 + s]);
    }
    __syncthreads();
    float* g = G + (size_t)blockIdx.x * NPARAM;
    if (threadIdx.x == 0) {
        float se = 0.0f, db3 = 0.0f;
        for (int s = 0; s < ns; s++) {
            const float e = T[s0 + s] - L.out[s];
            se += pinball ? fmaxf(tau * e, (tau - 1.0f) * e) : e * e;
            db3 += L.dy[s];
        }
        Lsum[blockIdx.x] = se;
        g[OFF_B3] = db3;
    }
    for (int u = threadIdx.x; u < NH; u += blockDim.x) {
        float acc = 0.0f;
        for (int s = 0; s < ns; s++) acc = fmaf(L.dy[s], L.h2[s][u], acc);
        g[OFF_W3 + u] = acc;
    }
    // layer-2 delta d2 = dy * W3 * gelu'(p2), kept in p2
    for (int k = threadIdx.x; k < ns * NH; k += blockDim.x) {
        const int s = k / NH, u = k % NH;
        L.p2[s][u] = L.dy[s] * L.w3[u] * gelu_grad(L.p2[s][u]);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < NH * NH; k += blockDim.x) {
        const int u = k / NH, v = k % NH;
        float acc = 0.0f;
        for (int s = 0; s < ns; s++) acc = fmaf(L.p2[s][u], L.h1[s][v], acc);
        g[OFF_W2 + k] = acc;
    }
    for (int u = threadIdx.x; u < NH; u += blockDim.x) {
        float acc = 0.0f;
        for (int s = 0; s < ns; s++) acc += L.p2[s][u];
        g[OFF_B2 + u] = acc;
    }
    // layer-1 delta d1 = (W2^T d2) * gelu'(p1), kept in p1
    for (int k = threadIdx.x; k < ns * NH; k += blockDim.x) {
        const int s = k / NH, v = k % NH;
        float acc = 0.0f;
        for (int u = 0; u < NH; u++) acc = fmaf(L.p2[s][u], bf(L.w2[u * W2S + v]), acc);
        L.p1[s][v] = acc * gelu_grad(L.p1[s][v]);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < NH * NF; k += blockDim.x) {
        const int v = k / NF, f = k % NF;
        float acc = 0.0f;
        for (int s = 0; s < ns; s++) acc = fmaf(L.p1[s][v], L.z[s][f], acc);
        g[OFF_W1 + k] = acc;
    }
    for (int v = threadIdx.x; v < NH; v += blockDim.x) {
        float acc = 0.0f;
        for (int s = 0; s < ns; s++) acc += L.p1[s][v];
        g[OFF_B1 + v] = acc;
    }
}

// g[k] = sum over blocks in block order; *loss = (sum of block SSE) / n.
__global__ void k_mlp_reduce(const float* __restrict__ G, int nb, float* __restrict__ g, const float* __restrict__ Lsum,
                             float* __restrict__ loss, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < NPARAM) {
        float acc = 0.0f;
        for (int b = 0; b < nb; b++) acc += G[(size_t)b * NPARAM + k];
        g[k] = acc;
    }
    if (k == 0) {
        float se = 0.0f;
        for (int b = 0; b < nb; b++) se += Lsum[b];
        *loss = se / (float)n;
    }
}

__global__ void k_mlp_adam(float* __restrict__ w, unsigned short* __restrict__ wb, const float* __restrict__ g,
                           float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2, float eps,
                           float c1, float c2)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= NPARAM) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.0f - b1) * gk;
    const float vk = b2 * v[k] + (1.0f - b2) * (gk * gk);
    m[k] = mk;
    v[k] = vk;
    const float wk = w[k] - lr * (mk / c1) / (sqrtf(vk / c2) + eps);
    w[k] = wk;
    wb[k] = to_bf16(wk);
}

__global__ void k_mlp_to_bf16(const float* __restrict__ w, unsigned short* __restrict__ wb)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < NPARAM) wb[k] = to_bf16(w[k]);
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

constexpr int FEAT_SRC = 64;    // sources per work item of the pair pass

// The step-start kernel of the proposer.
//
// State features.  Work items, one wave each: [0, n0): (64 SI sinks of 1..n_si-1, 64-source
// chunk of 1..n_massive-1); [n0, n0+n1): (64 NSI sinks, chunk of SI sources 1..n_si-1); then
// one item per 64 bodies 1..n-1.  Lanes run over sinks; each lane stages one source of the
// chunk and the wave broadcasts them with readlane, so the pair loop issues no memory
// operations; waves take items grid-stride and keep running extrema (exact in any order).
// The pair quantity is q^3 = d^3 3 m_star / ((m_i + m_j) ((h_i + h_j)/2)^3), heliocentric h,
// from FP32 star-relative positions (the star subtracted in FP64) and the hardware square
// root and reciprocal: a few 1e-4 relative in d at 1e-3 AU separations 5 AU out, ample for
// a logarithmic feature.  Per body: |a| and |a - a_kep| / |a_kep|, a_kep = -m_star (r - r_star)/|r - r_star|^3.
// Each block stores its extrema in its own slot.  The kernel runs in two launches: the pair
// items [it0, it1) need only the state, so they run on a stream of their own while the step's
// first force evaluation is computed (tail = 0: slots [0, blocks)); the body items need that
// evaluation and run after it (tail = 1: slots from slot0), and the last of their blocks to
// finish (counter) reduces all final_slots slots, completes the feature vector, runs the
// network and applies the proposal.
__global__ __launch_bounds__(256) void k_learn_step(const red::d4* __restrict__ r, const red::d4* __restrict__ a,
                                                    const int* __restrict__ id, int n_si, int n_massive, int n,
                                                    long n0, int ch0, long n1, int ch1, long it0, long it1,
                                                    int slot0, int final_slots, int tail,
                                                    float* __restrict__ part, unsigned* __restrict__ counter,
                                                    float* __restrict__ out, const unsigned short* __restrict__ wb,
                                                    const float* __restrict__ norm, red::ProposeArgs A,
                                                    red::Ctl* __restrict__ ctl, int apply, double eta_n,
                                                    double safety, double max_log)
{
    __shared__ Lds1 S;
    __shared__ float red3[3][4];
    __shared__ float x[NF];
    __shared__ int last;
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const long nwaves = (long)gridDim.x * (blockDim.x >> 6);
    const red::d4 star = r[0];
    const double ms = star.w;
    float best = __uint_as_float(0x7f800000u), amax = 0.0f, pmax = 0.0f;
    for (long w = it0 + (long)blockIdx.x * (blockDim.x >> 6) + wid; w < it1; w += nwaves) {
        if (w < n0 + n1) {
            int i, i1, src0, src1;
            long c;
            if (w < n0) {
                i = 1 + (int)(w / ch0) * 64 + lane; i1 = n_si; c = w % ch0; src0 = 1; src1 = n_massive;
            } else {
                const long u = w - n0;
                i = n_si + (int)(u / ch1) * 64 + lane; i1 = n_massive; c = u % ch1; src0 = 1; src1 = n_si;
            }
            const int j0 = src0 + (int)c * FEAT_SRC;
            const int nj = min(src1 - j0, FEAT_SRC);
            // lane k stages source j0 + k: star-relative FP32 position, heliocentric distance,
            // mass (-1 when inactive); the loop below broadcasts it with readlane
            float sx = 0.0f, sy = 0.0f, sz = 0.0f, sh = 0.0f, sm = -1.0f;
            if (lane < nj) {
                const red::d4 q = r[j0 + lane];
                sx = (float)(q.x - star.x); sy = (float)(q.y - star.y); sz = (float)(q.z - star.z);
                sh = sqrtf(sx * sx + sy * sy + sz * sz);
                sm = id[j0 + lane] > 0 ? (float)q.w : -1.0f;
            }
            const bool live = i < i1 && id[i] > 0;
            const red::d4 ri = r[live ? i : 0];
            const float xi = (float)(ri.x - star.x), yi = (float)(ri.y - star.y), zi = (float)(ri.z - star.z);
            const float hi = sqrtf(xi * xi + yi * yi + zi * zi);
            const float mi = (float)ri.w;
            const float three_ms = (float)(3.0 * ms);
            auto pair = [&](int k) {
                const float qx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sx), k));
                const float qy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sy), k));
                const float qz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sz), k));
                const float qh = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sh), k));
                const float qm = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sm), k));
                const float dx = qx - xi, dy = qy - yi, dz = qz - zi;
                const float d2 = dx * dx + dy * dy + dz * dz;
                const float mm = qm + mi;
                const float hm = 0.5f * (hi + qh);
                const float q3 = d2 * __builtin_amdgcn_sqrtf(d2) * three_ms * __builtin_amdgcn_rcpf(mm * hm * hm * hm);
                const bool ok = live && qm >= 0.0f && j0 + k != i && mm > 0.0f;
                best = ok ? fminf(best, q3) : best;
            };
            if (nj == FEAT_SRC) {
                // full chunk: unrolled, so that several pairs' sqrt / rcp chains overlap
#pragma unroll 8
                for (int k = 0; k < FEAT_SRC; ++k) pair(k);
            } else {
                for (int k = 0; k < nj; ++k) pair(k);
            }
        } else {
            const int i = 1 + (int)(w - n0 - n1) * 64 + lane;
            if (i < n && id[i] > 0) {
                const red::d4 ri = r[i];
                const red::d4 ai = a[i];
                const double hx = ri.x - star.x, hy = ri.y - star.y, hz = ri.z - star.z;
                const double h2 = hx * hx + hy * hy + hz * hz;
                const double s = -ms / (h2 * sqrt(h2));
                const double kx = s * hx, ky = s * hy, kz = s * hz;
                const double ex = ai.x - kx, ey = ai.y - ky, ez = ai.z - kz;
                const double an = sqrt(ai.x * ai.x + ai.y * ai.y + ai.z * ai.z);
                const double kn = sqrt(kx * kx + ky * ky + kz * kz);
                const double en = sqrt(ex * ex + ey * ey + ez * ez);
                amax = fmaxf(amax, (float)an);
                if (kn > 0.0) pmax = fmaxf(pmax, (float)(en / kn));
            }
        }
    }
    best = wave_min(best);
    amax = wave_max(amax);
    pmax = wave_max(pmax);
    if (lane == 0) { red3[0][wid] = best; red3[1][wid] = amax; red3[2][wid] = pmax; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
            best = fminf(best, red3[0][k]);
            amax = fmaxf(amax, red3[1][k]);
            pmax = fmaxf(pmax, red3[2][k]);
        }
        part[3 * (slot0 + blockIdx.x) + 0] = best;
        part[3 * (slot0 + blockIdx.x) + 1] = amax;
        part[3 * (slot0 + blockIdx.x) + 2] = pmax;
        if (tail) {
            __threadfence();
            const unsigned done = atomicAdd(counter, 1u);
            last = done == gridDim.x - 1;
        } else {
            last = 0;
        }
    }
    __syncthreads();
    if (!last) return;
    // last block: every other block's slot is visible (release fence before its increment);
    // all its threads read the slots and the block reduces them, while the network's weight
    // rows are on their way into registers
    __threadfence();
    RowW R;
    if (wb && threadIdx.x < NH) load_row(R, wb, threadIdx.x);
    {
        float q3 = __uint_as_float(0x7f800000u), ma = 0.0f, mp = 0.0f;
        for (int b = threadIdx.x; b < final_slots; b += blockDim.x) {
            q3 = fminf(q3, __hip_atomic_load(&part[3 * b + 0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            ma = fmaxf(ma, __hip_atomic_load(&part[3 * b + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            mp = fmaxf(mp, __hip_atomic_load(&part[3 * b + 2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        }
        q3 = wave_min(q3);
        ma = wave_max(ma);
        mp = wave_max(mp);
        if (lane == 0) { red3[0][wid] = q3; red3[1][wid] = ma; red3[2][wid] = mp; }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float q3 = red3[0][0], ma = red3[1][0], mp = red3[2][0];
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
            q3 = fminf(q3, red3[0][k]);
            ma = fmaxf(ma, red3[1][k]);
            mp = fmaxf(mp, red3[2][k]);
        }
        *counter = 0u;   // ready for the next launch on this stream
        for (int k = 0; k < NF; k++) x[k] = A.x[k];
        // log10(d / R_H) = log10(q^3) / 3; every value is clamped to keep the features finite
        const float lq = fminf(fmaxf(log10f(fmaxf(q3, 1e-30f)) / 3.0f, -10.0f), 10.0f);
        const float la = fminf(fmaxf(log10f(fmaxf(ma, 1e-30f)), -30.0f), 30.0f);
        const float lp = fminf(fmaxf(log10f(fmaxf(mp, 1e-30f)), -30.0f), 30.0f);
        x[red::FEAT_LOG_MIN_SEP] = lq;
        x[red::FEAT_D_LOG_MIN_SEP] = A.have_prev ? lq - A.prev_log_min_sep : 0.0f;
        x[red::FEAT_LOG_MAX_ACC] = la;
        x[red::FEAT_LOG_MAX_PERT] = lp;
        x[red::FEAT_D_LOG_MAX_PERT] = A.have_prev ? lp - A.prev_log_max_pert : 0.0f;
        for (int k = 0; k < NF; k++) out[k] = x[k];
        out[NF] = 0.0f;
    }
    __syncthreads();
    if (!wb) return;
    forward_one(S, R, wb, norm, x);
    if (threadIdx.x == 0) {
        out[NF] = S.out;
        if (apply) {
            const double y = fmin(max_log, fmax(-max_log, (double)S.out));
            ctl->dt = safety * eta_n * exp(y);
        }
    }
}

}  // namespace

struct red_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    float* w = nullptr;            // FP32 master
    unsigned short* wb = nullptr;  // BF16 image
    float* norm = nullptr;         // mean[12], inv_std[12]
    float *m = nullptr, *v = nullptr, *g = nullptr, *G = nullptr, *Ls = nullptr, *loss = nullptr;
    float *X = nullptr, *T = nullptr, *P = nullptr;
    int cap_n = 0, cap_blocks = 0, cap_losses = 0;
    long step = 0;
    int pinball = 0;               // RED_MLP_LOSS_*
    float tau = 0.5f;
    int hipfail(hipError_t e, const char* where)
    {
        err = std::string(where) + ": " + hipGetErrorString(e);
        return RED_ERR_HIP;
    }
};

#define MLP_HIP(M, expr)                                        \
    do {                                                        \
        hipError_t _e = (expr);                                 \
        if (_e != hipSuccess) return (M)->hipfail(_e, #expr);   \
    } while (0)

namespace red {

int mlp_device(const red_mlp* M) { return M->device; }

LearnPlan learn_plan(int n_si, int n_massive, int n)
{
    LearnPlan L;
    const long c0 = n_si > 1 ? (n_si - 1 + 63) / 64 : 0;                // SI sink groups (no star)
    const long c1 = n_si > 1 ? (n_massive - n_si + 63) / 64 : 0;        // NSI sink groups
    L.ch0 = n_massive > 1 ? (n_massive - 1 + FEAT_SRC - 1) / FEAT_SRC : 0;
    L.ch1 = n_si > 1 ? (n_si - 1 + FEAT_SRC - 1) / FEAT_SRC : 0;
    L.n0 = c0 * L.ch0;
    L.n1 = c1 * L.ch1;
    L.nb = n > 1 ? (n - 1 + 63) / 64 : 0;
    L.b0 = (int)std::min<long>((L.n0 + L.n1 + 3) / 4, LEARN_PAIR_BLOCKS);
    L.b1 = (int)std::max<long>(1, std::min<long>((L.nb + 3) / 4, LEARN_BODY_BLOCKS));
    return L;
}

hipError_t launch_learn_pairs(const LearnPlan& L, const d4* r, const int* id, int n_si, int n_massive, int n,
                              const LearnBuffers& B, hipStream_t s)
{
    if (L.b0 <= 0) return hipSuccess;
    ProposeArgs A;
    memset(&A, 0, sizeof(A));
    hipLaunchKernelGGL(k_learn_step, dim3((unsigned)L.b0), dim3(256), 0, s, r, nullptr, id, n_si, n_massive, n, L.n0,
                       L.ch0, L.n1, L.ch1, 0L, L.n0 + L.n1, 0, 0, 0, B.part, B.counter, B.out, nullptr, nullptr, A,
                       nullptr, 0, 0.0, 0.0, 0.0);
    return hipGetLastError();
}

hipError_t launch_learn_tail(const red_mlp* net, const ProposeArgs& A, const LearnPlan& L, const d4* r, const d4* a,
                             const int* id, int n_si, int n_massive, int n, const LearnBuffers& B, Ctl* ctl,
                             int apply, double eta_n, double safety, double max_log, hipStream_t s)
{
    const long it0 = L.n0 + L.n1;
    hipLaunchKernelGGL(k_learn_step, dim3((unsigned)L.b1), dim3(256), 0, s, r, a, id, n_si, n_massive, n, L.n0, L.ch0,
                       L.n1, L.ch1, it0, it0 + L.nb, L.b0, L.b0 + L.b1, 1, B.part, B.counter, B.out,
                       net ? net->wb : nullptr, net ? net->norm : nullptr, A, ctl, apply, eta_n, safety, max_log);
    return hipGetLastError();
}

}  // namespace red

namespace {

int ensure(red_mlp* M, int n, int nblocks, int nlosses)
{
    if (n <= M->cap_n && nblocks <= M->cap_blocks && nlosses <= M->cap_losses) return RED_OK;
    red::SetupLock lk;
    if (n > M->cap_n) {
        (void)hipFree(M->X); (void)hipFree(M->T); (void)hipFree(M->P);
        M->X = M->T = M->P = nullptr;
        M->cap_n = 0;
        MLP_HIP(M, hipMalloc(&M->X, sizeof(float) * (size_t)n * NF));
        MLP_HIP(M, hipMalloc(&M->T, sizeof(float) * (size_t)n));
        MLP_HIP(M, hipMalloc(&M->P, sizeof(float) * (size_t)n));
        M->cap_n = n;
    }
    if (nblocks > M->cap_blocks) {
        (void)hipFree(M->G); (void)hipFree(M->Ls);
        M->G = M->Ls = nullptr;
        M->cap_blocks = 0;
        MLP_HIP(M, hipMalloc(&M->G, sizeof(float) * (size_t)nblocks * NPARAM));
        MLP_HIP(M, hipMalloc(&M->Ls, sizeof(float) * (size_t)nblocks));
        M->cap_blocks = nblocks;
    }
    if (nlosses > M->cap_losses) {
        (void)hipFree(M->loss);
        M->loss = nullptr;
        M->cap_losses = 0;
        MLP_HIP(M, hipMalloc(&M->loss, sizeof(float) * (size_t)nlosses));
        M->cap_losses = nlosses;
    }
    return RED_OK;
}

constexpr int PB = 256;
inline dim3 pgrid() { return dim3((NPARAM + PB - 1) / PB); }

}  // namespace

extern "C" {

int red_mlp_param_count(void) { return NPARAM; }

int red_mlp_destroy(red_mlp* M)
{
    if (!M) return RED_OK;
    red::SetupLock lk;
    (void)hipSetDevice(M->device);
    if (M->stream) (void)hipStreamSynchronize(M->stream);
    for (void* p : {(void*)M->w, (void*)M->wb, (void*)M->norm, (void*)M->m, (void*)M->v, (void*)M->g, (void*)M->G,
                    (void*)M->Ls, (void*)M->loss, (void*)M->X, (void*)M->T, (void*)M->P})
        (void)hipFree(p);
    if (M->stream) (void)hipStreamDestroy(M->stream);
    delete M;
    return RED_OK;
}

const char* red_mlp_last_error(const red_mlp* M) { return M ? M->err.c_str() : "no network"; }

int red_mlp_create(int device, const float* params, red_mlp** out)
{
    if (!out) return RED_ERR_ARG;
    *out = nullptr;
    red::SetupLock lk;
    red_mlp* M = new red_mlp();
    M->device = device;
    *out = M;
    MLP_HIP(M, hipSetDevice(device));
    MLP_HIP(M, hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking));
    MLP_HIP(M, hipMalloc(&M->w, sizeof(float) * NPARAM));
    MLP_HIP(M, hipMalloc(&M->wb, sizeof(unsigned short) * NPARAM));
    MLP_HIP(M, hipMalloc(&M->norm, sizeof(float) * 2 * NF));
    MLP_HIP(M, hipMalloc(&M->m, sizeof(float) * NPARAM));
    MLP_HIP(M, hipMalloc(&M->v, sizeof(float) * NPARAM));
    MLP_HIP(M, hipMalloc(&M->g, sizeof(float) * NPARAM));
    MLP_HIP(M, hipMemsetAsync(M->m, 0, sizeof(float) * NPARAM, M->stream));
    MLP_HIP(M, hipMemsetAsync(M->v, 0, sizeof(float) * NPARAM, M->stream));
    std::vector<float> w(NPARAM, 0.0f), nrm(2 * NF, 0.0f);
    if (params) memcpy(w.data(), params, sizeof(float) * NPARAM);
    for (int f = 0; f < NF; f++) nrm[NF + f] = 1.0f;
    MLP_HIP(M, red::stream_copy(M->stream, M->w, w.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice));
    MLP_HIP(M, red::stream_copy(M->stream, M->norm, nrm.data(), sizeof(float) * 2 * NF, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mlp_to_bf16, pgrid(), dim3(PB), 0, M->stream, M->w, M->wb);
    MLP_HIP(M, hipGetLastError());
    MLP_HIP(M, hipStreamSynchronize(M->stream));
    return RED_OK;
}

int red_mlp_set_normalizer(red_mlp* M, const float* mean, const float* inv_std)
{
    if (!M || !mean || !inv_std) return RED_ERR_ARG;
    float nrm[2 * NF];
    for (int f = 0; f < NF; f++) {
        if (!std::isfinite(mean[f]) || !std::isfinite(inv_std[f])) {
            M->err = "normaliser values must be finite";
            return RED_ERR_ARG;
        }
        nrm[f] = mean[f];
        nrm[NF + f] = inv_std[f];
    }
    MLP_HIP(M, hipSetDevice(M->device));
    MLP_HIP(M, hipMemcpyAsync(M->norm, nrm, sizeof(nrm), hipMemcpyHostToDevice, M->stream));
    MLP_HIP(M, hipStreamSynchronize(M->stream));
    return RED_OK;
}

int red_mlp_set_loss(red_mlp* M, int kind, float tau)
{
    if (!M) return RED_ERR_ARG;
    if (kind != RED_MLP_LOSS_MSE && kind != RED_MLP_LOSS_PINBALL) {
        M->err = "unknown loss";
        return RED_ERR_ARG;
    }
    if (kind == RED_MLP_LOSS_PINBALL && !(tau > 0.0f && tau < 1.0f)) {
        M->err = "the quantile level must lie in (0, 1)";
        return RED_ERR_ARG;
    }
    M->pinball = kind == RED_MLP_LOSS_PINBALL;
    M->tau = tau;
    return RED_OK;
}

int red_mlp_get_params(red_mlp* M, float* params, unsigned short* bf16, float* norm)
{
    if (!M) return RED_ERR_ARG;
    MLP_HIP(M, hipSetDevice(M->device));
    MLP_HIP(M, hipStreamSynchronize(M->stream));
    if (params) MLP_HIP(M, red::stream_copy(M->stream, params, M->w, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
    if (bf16) MLP_HIP(M, red::stream_copy(M->stream, bf16, M->wb, sizeof(unsigned short) * NPARAM, hipMemcpyDeviceToHost));
    if (norm) MLP_HIP(M, red::stream_copy(M->stream, norm, M->norm, sizeof(float) * 2 * NF, hipMemcpyDeviceToHost));
    return RED_OK;
}

int red_mlp_predict(red_mlp* M, const float* X, int n, float* out)
{
    if (!M || n < 0 || (n && (!X || !out))) return RED_ERR_ARG;
    if (n == 0) return RED_OK;
    MLP_HIP(M, hipSetDevice(M->device));
    const int nb = (n + SB - 1) / SB;
    int rc = ensure(M, n, 1, 1);
    if (rc) return rc;
    MLP_HIP(M, hipMemcpyAsync(M->X, X, sizeof(float) * (size_t)n * NF, hipMemcpyHostToDevice, M->stream));
    hipLaunchKernelGGL(k_mlp_predict, dim3(nb), dim3(256), 0, M->stream, M->wb, M->norm, M->X, M->P, n);
    MLP_HIP(M, hipGetLastError());
    MLP_HIP(M, hipMemcpyAsync(out, M->P, sizeof(float) * n, hipMemcpyDeviceToHost, M->stream));
    MLP_HIP(M, hipStreamSynchronize(M->stream));
    return RED_OK;
}

int red_mlp_gradient(red_mlp* M, const float* X, const float* T, int n, float* grad, float* loss)
{
    if (!M || n <= 0 || !X || !T) return RED_ERR_ARG;
    MLP_HIP(M, hipSetDevice(M->device));
    const int nb = (n + SB - 1) / SB;
    int rc = ensure(M, n, nb, 1);
    if (rc) return rc;
    MLP_HIP(M, hipMemcpyAsync(M->X, X, sizeof(float) * (size_t)n * NF, hipMemcpyHostToDevice, M->stream));
    MLP_HIP(M, hipMemcpyAsync(M->T, T, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, M->stream));
    hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(256), 0, M->stream, M->wb, M->norm, M->X, M->T, n, M->pinball, M->tau,
                       M->G, M->Ls);
    MLP_HIP(M, hipGetLastError());
    hipLaunchKernelGGL(k_mlp_reduce, pgrid(), dim3(PB), 0, M->stream, M->G, nb, M->g, M->Ls, M->loss, n);
    MLP_HIP(M, hipGetLastError());
    if (grad) MLP_HIP(M, hipMemcpyAsync(grad, M->g, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, M->stream));
    if (loss) MLP_HIP(M, hipMemcpyAsync(loss, M->loss, sizeof(float), hipMemcpyDeviceToHost, M->stream));
    MLP_HIP(M, hipStreamSynchronize(M->stream));
    return RED_OK;
}

int red_mlp_train(red_mlp* M, const float* X, const float* T, int n, int epochs, int batch, float lr, float* losses)
{
    if (!M || n <= 0 || !X || !T || epochs <= 0 || batch <= 0 || !(lr > 0.0f)) return RED_ERR_ARG;
    MLP_HIP(M, hipSetDevice(M->device));
    const int nbatch = (n + batch - 1) / batch;
    const long nl = (long)nbatch * epochs;
    if (nl > (1L << 26)) {
        M->err = "too many minibatches";
        return RED_ERR_ARG;
    }
    int rc = ensure(M, n, (std::min(batch, n) + SB - 1) / SB, (int)nl);
    if (rc) return rc;
    MLP_HIP(M, hipMemcpyAsync(M->X, X, sizeof(float) * (size_t)n * NF, hipMemcpyHostToDevice, M->stream));
    MLP_HIP(M, hipMemcpyAsync(M->T, T, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, M->stream));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    long li = 0;
    for (int e = 0; e < epochs; e++) {
        for (int s = 0; s < n; s += batch, li++) {
            const int nbs = std::min(batch, n - s);
            const int nb = (nbs + SB - 1) / SB;
            hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(256), 0, M->stream, M->wb, M->norm, M->X + (size_t)s * NF,
                               M->T + s, nbs, M->pinball, M->tau, M->G, M->Ls);
            MLP_HIP(M, hipGetLastError());
            hipLaunchKernelGGL(k_mlp_reduce, pgrid(), dim3(PB), 0, M->stream, M->G, nb, M->g, M->Ls, M->loss + li, nbs);
            MLP_HIP(M, hipGetLastError());
            M->step++;
            const float c1 = 1.0f - (float)std::pow((double)b1, (double)M->step);
            const float c2 = 1.0f - (float)std::pow((double)b2, (double)M->step);
            hipLaunchKernelGGL(k_mlp_adam, pgrid(), dim3(PB), 0, M->stream, M->w, M->wb, M->g, M->m, M->v, lr, b1, b2,
                               eps, c1, c2);
            MLP_HIP(M, hipGetLastError());
        }
    }
    if (losses) {
        std::vector<float> all((size_t)nl);
        MLP_HIP(M, hipMemcpyAsync(all.data(), M->loss, sizeof(float) * (size_t)nl, hipMemcpyDeviceToHost, M->stream));
        MLP_HIP(M, hipStreamSynchronize(M->stream));
        for (int e = 0; e < epochs; e++) {
            double sum = 0.0;
            for (int b = 0; b < nbatch; b++) sum += all[(size_t)e * nbatch + b];
            losses[e] = (float)(sum / nbatch);
        }
    }
    MLP_HIP(M, hipStreamSynchronize(M->stream));
    return RED_OK;
}

int red_mlp_adam_state(red_mlp* M, float* m, float* v, long* step)
{
    if (!M) return RED_ERR_ARG;
    MLP_HIP(M, hipSetDevice(M->device));
    MLP_HIP(M, hipStreamSynchronize(M->stream));
    if (m) MLP_HIP(M, red::stream_copy(M->stream, m, M->m, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
    if (v) MLP_HIP(M, red::stream_copy(M->stream, v, M->v, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
    if (step) *step = M->step;
    return RED_OK;
}

}  // extern "C"
