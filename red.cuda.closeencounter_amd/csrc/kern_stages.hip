// kern_stages.hip -- Runge-Kutta stage combinations and the embedded-error norm (gfx950).
//
// One work-item per body handles both halves of the state: the position record (whose
// derivative at stage k is the velocity record of stage state k) and the velocity record
// (whose derivative is the acceleration of stage k).  Records are 32 bytes, so a wavefront
// streams 2 KiB per operand with fully coalesced 16-byte loads.  The trial step and stage
// time come from the control block, so a captured graph replays with current values.
//
// Expression shapes follow the reference kernels term for term (built with
// -ffp-contract=off, so each product and sum is rounded separately as there):
//   SF_AXPY     out = y + (a*dt) f                 Euler/RK2/RK4 sum_vector
//               (int_euler.cu:21-33, int_rungekutta2.cu:19-31, int_rungekutta4.cu:36-48)
//   SF_DTSUM    out = y + dt (a0 f0 + a1 f1 + ...)   RKF78 calc_ytemp_for_fN
//               (int_rungekutta8.cu:137-305)
//   SF_ACCUM    out = y; out += (dt*a_k) f_k         Cash-Karp calc_ytemp / calc_y_np1
//               (int_rungekutta5.cu:41-77)
//   SF_RK4NP1   out = y + (b0 dt)(f0+f3) + (b1 dt)(f1+f2)   (int_rungekutta4.cu:50-61)
//   SF_RKF78NP1 out = y + dt (b0 f0 + b5 f5 + b6 (f6+f7) + b8 (f8+f9) + b10 f10)
//               (int_rungekutta8.cu:321-333)
// Error forms (fused with the max-reduction that integrator::get_max_error performs
// with thrust::max_element, integrator.cu:211-256):
//   EF_RK4      e = f3 - f4                         (int_rungekutta4.cu:63-74)
//   EF_RKF78    e = f0 + f10 - f11 - f12            (int_rungekutta8.cu:307-319)
//   EF_CK       e = sum_k (bh_k - b_k) f_k          (int_rungekutta5.cu:79-92); the zero
//               coefficient of f1 is left out (0 f1 adds a zero, which cannot change a maximum
//               taken from 0)
// The maximum starts from 0 and ignores NaN like the reference CPU loop; a NaN sets a
// flag the host turns into an error.
#include "red_internal.hpp"
#include "stage_forms.hpp"

namespace red {
namespace {

// NT > 0: the form's term count, known at compile time (combine_nt); 0: any count.
template <int FORM, int NT>
__global__ __launch_bounds__(256) void k_stage(StageArgs A)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    double dt;
    if (A.disk_of) {
        const int d = A.disk_of[i];
        // a slot's tail left by a compaction, or a disk done or waiting: buffers stay as they are
        if (d < 0 || !(A.disk_flags[d] & 1)) return;
        dt = A.disk_dt[d];
    } else {
        dt = A.ctl->dt;
    }
    const d4 yr = A.yr[i];
    const d4 yv = A.yv[i];
    d4 orr, ovv;
    if constexpr (NT > 0) {
        orr = combine_nt<FORM, NT>(A, yr, i, false, dt);
        ovv = combine_nt<FORM, NT>(A, yv, i, true, dt);
    } else {
        orr = combine<FORM>(A, yr, i, false, dt);
        ovv = combine<FORM>(A, yv, i, true, dt);
    }
    orr.w = yr.w;   // mass rides along unchanged
    ovv.w = 0.0;
    A.outr[i] = orr;
    A.outv[i] = ovv;
}

template <int EF, int NT = 0>
__device__ __forceinline__ double3 err_rec(const StageArgs& A, int i, bool vel)
{
    double3 e;
    if constexpr (EF == EF_CK && NT > 0) {
        // the term count known at compile time: every record requested before the first is used
        d4 f[NT];
#pragma unroll
        for (int k = 0; k < NT; ++k) f[k] = vel ? A.t[k].fa[i] : A.t[k].fv[i];
        e.x = A.t[0].coef * f[0].x; e.y = A.t[0].coef * f[0].y; e.z = A.t[0].coef * f[0].z;
#pragma unroll
        for (int k = 1; k < NT; ++k) {
            e.x += A.t[k].coef * f[k].x;
            e.y += A.t[k].coef * f[k].y;
            e.z += A.t[k].coef * f[k].z;
        }
        return e;
    }
    if (EF == EF_RK4) {
        const d4 f3 = vel ? A.t[0].fa[i] : A.t[0].fv[i];
        const d4 f4 = vel ? A.t[1].fa[i] : A.t[1].fv[i];
        e.x = f3.x - f4.x; e.y = f3.y - f4.y; e.z = f3.z - f4.z;
    } else if (EF == EF_RKF78) {
        const d4 f0 = vel ? A.t[0].fa[i] : A.t[0].fv[i];
        const d4 f10 = vel ? A.t[1].fa[i] : A.t[1].fv[i];
        const d4 f11 = vel ? A.t[2].fa[i] : A.t[2].fv[i];
        const d4 f12 = vel ? A.t[3].fa[i] : A.t[3].fv[i];
        e.x = (f0.x + f10.x - f11.x - f12.x);
        e.y = (f0.y + f10.y - f11.y - f12.y);
        e.z = (f0.z + f10.z - f11.z - f12.z);
    } else {
        d4 f = vel ? A.t[0].fa[i] : A.t[0].fv[i];
        e.x = A.t[0].coef * f.x; e.y = A.t[0].coef * f.y; e.z = A.t[0].coef * f.z;
        for (int k = 1; k < A.nterms; k++) {
            f = vel ? A.t[k].fa[i] : A.t[k].fv[i];
            e.x += A.t[k].coef * f.x;
            e.y += A.t[k].coef * f.y;
            e.z += A.t[k].coef * f.z;
        }
    }
    return e;
}

__device__ __forceinline__ void upd_max(double& m, int& nan, double e)
{
    if (e != e) nan = 1;
    else if (m < e) m = e;
}

// Ensemble form: every body's maximum goes to its disk's slot.  A wavefront whose lanes
// all belong to one disk reduces first and issues one atomic.
template <int EF>
__global__ __launch_bounds__(256) void k_err_disks(StageArgs A)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    int nan = 0;
    int d = -1;
    if (i < A.n) {
        d = A.disk_of[i];
        if (d >= 0 && (A.disk_flags[d] & 1) && A.err_on[i]) {
            const double3 er = err_rec<EF>(A, i, false);
            const double3 ev = err_rec<EF>(A, i, true);
            upd_max(m, nan, er.x); upd_max(m, nan, er.y); upd_max(m, nan, er.z);
            upd_max(m, nan, ev.x); upd_max(m, nan, ev.y); upd_max(m, nan, ev.z);
        }
    }
    const int d0 = __shfl(d, 0, 64);
    const bool uniform = __all(d == d0 || d < 0);
    if (uniform) {
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(m, off, 64);
            m = m < o ? o : m;
            nan |= __shfl_xor(nan, off, 64);
        }
        if ((threadIdx.x & 63) == 0 && d0 >= 0) {
            if (m > 0.0) atomicMax(&A.disk_err[d0], (unsigned long long)__double_as_longlong(m));
            if (nan) atomicOr(&A.disk_nan[d0], 1);
        }
    } else if (d >= 0) {
        if (m > 0.0) atomicMax(&A.disk_err[d], (unsigned long long)__double_as_longlong(m));
        if (nan) atomicOr(&A.disk_nan[d], 1);
    }
}

template <int EF, int NT>
__global__ __launch_bounds__(256) void k_err(StageArgs A, Ctl* ctl)
{
    __shared__ double wmax[4];
    __shared__ int wnan[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    int nan = 0;
    if (i < A.n) {
        const double3 er = err_rec<EF, NT>(A, i, false);
        const double3 ev = err_rec<EF, NT>(A, i, true);
        upd_max(m, nan, er.x); upd_max(m, nan, er.y); upd_max(m, nan, er.z);
        upd_max(m, nan, ev.x); upd_max(m, nan, ev.y); upd_max(m, nan, ev.z);
    }
    // wavefront max via cross-lane shuffles, then across the block's waves
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off, 64);
        m = m < o ? o : m;
        nan |= __shfl_xor(nan, off, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) { wmax[wave] = m; wnan[wave] = nan; }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int w = 1; w < nw; w++) { m = m < wmax[w] ? wmax[w] : m; nan |= wnan[w]; }
        // m >= +0.0, whose IEEE bit patterns are ordered like the values
        if (m > 0.0) atomicMax(&ctl->err_bits, (unsigned long long)__double_as_longlong(m));
        if (nan) atomicOr(&ctl->nan_flag, 1);
    }
}

__global__ void k_reset_ctl(Ctl* ctl, int clear_events)
{
    ctl->err_bits = 0ull;
    ctl->nan_flag = 0;
    if (clear_events) ctl->n_events = 0;
}

}  // namespace

namespace {

template <int FORM, int NT>
hipError_t go_stage(const StageArgs& A, dim3 g, dim3 b, hipStream_t s)
{
    hipLaunchKernelGGL((k_stage<FORM, NT>), g, b, 0, s, A);
    return hipGetLastError();
}

// the forms with a term loop, specialised on the row's term count
template <int FORM>
hipError_t go_stage_n(const StageArgs& A, dim3 g, dim3 b, hipStream_t s)
{
    switch (A.nterms) {
    case 1: return go_stage<FORM, 1>(A, g, b, s);
    case 2: return go_stage<FORM, 2>(A, g, b, s);
    case 3: return go_stage<FORM, 3>(A, g, b, s);
    case 4: return go_stage<FORM, 4>(A, g, b, s);
    case 5: return go_stage<FORM, 5>(A, g, b, s);
    case 6: return go_stage<FORM, 6>(A, g, b, s);
    case 7: return go_stage<FORM, 7>(A, g, b, s);
    case 8: return go_stage<FORM, 8>(A, g, b, s);
    case 9: return go_stage<FORM, 9>(A, g, b, s);
    default: return go_stage<FORM, 0>(A, g, b, s);
    }
}

}  // namespace

hipError_t launch_stage(int form, const StageArgs& A, hipStream_t s)
{
    if (A.n <= 0) return hipSuccess;
    dim3 g((A.n + 255) / 256), b(256);
    switch (form) {
    case SF_AXPY: return go_stage<SF_AXPY, 0>(A, g, b, s);
    case SF_DTSUM: return go_stage_n<SF_DTSUM>(A, g, b, s);
    case SF_ACCUM: return go_stage_n<SF_ACCUM>(A, g, b, s);
    case SF_RK4NP1: return go_stage<SF_RK4NP1, 0>(A, g, b, s);
    case SF_RKF78NP1: return go_stage<SF_RKF78NP1, 0>(A, g, b, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_err(int form, const StageArgs& A, hipStream_t s)
{
    if (A.n <= 0) return hipSuccess;
    dim3 g((A.n + 255) / 256), b(256);
    if (A.disk_of) {
        switch (form) {
        case EF_RK4: hipLaunchKernelGGL(k_err_disks<EF_RK4>, g, b, 0, s, A); break;
        case EF_RKF78: hipLaunchKernelGGL(k_err_disks<EF_RKF78>, g, b, 0, s, A); break;
        case EF_CK: hipLaunchKernelGGL(k_err_disks<EF_CK>, g, b, 0, s, A); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    Ctl* ctl = const_cast<Ctl*>(A.ctl);
    switch (form) {
    case EF_RK4: hipLaunchKernelGGL((k_err<EF_RK4, 0>), g, b, 0, s, A, ctl); break;
    case EF_RKF78: hipLaunchKernelGGL((k_err<EF_RKF78, 0>), g, b, 0, s, A, ctl); break;
    case EF_CK:
        if (A.nterms == 5) hipLaunchKernelGGL((k_err<EF_CK, 5>), g, b, 0, s, A, ctl);
        else if (A.nterms == 6) hipLaunchKernelGGL((k_err<EF_CK, 6>), g, b, 0, s, A, ctl);
        else hipLaunchKernelGGL((k_err<EF_CK, 0>), g, b, 0, s, A, ctl);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_reset_ctl(Ctl* ctl, int clear_events, hipStream_t s)
{
    hipLaunchKernelGGL(k_reset_ctl, dim3(1), dim3(1), 0, s, ctl, clear_events);
    return hipGetLastError();
}

}  // namespace red
