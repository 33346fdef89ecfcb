// stage_forms.hpp -- the Runge-Kutta stage combinations of one body record, shared by the
// stage kernels and the fused fixed-step loop, so both compute every stage state with the
// same expressions (the forms and their reference sources are listed in kern_stages.hip).
#pragma once

#include "red_internal.hpp"

namespace red {

__device__ __forceinline__ void axpy_rec(d4& o, const d4& y, double c, const d4& f)
{
    o.x = y.x + c * f.x;
    o.y = y.y + c * f.y;
    o.z = y.z + c * f.z;
}

// The combination of one half of a body's record (positions or velocities) from its term
// records, which get(k) returns; combine reads them from the stage arguments.
template <int FORM, class Get>
__device__ __forceinline__ d4 combine_get(const StageArgs& A, const d4& y, double dt, Get get)
{
    d4 o;
    o.w = y.w;
    if (FORM == SF_AXPY) {
        const double c = A.t[0].coef * dt;
        const d4 f = get(0);
        axpy_rec(o, y, c, f);
    } else if (FORM == SF_DTSUM) {
        d4 f = get(0);
        double sx = A.t[0].coef * f.x, sy = A.t[0].coef * f.y, sz = A.t[0].coef * f.z;
        for (int k = 1; k < A.nterms; k++) {
            f = get(k);
            sx = sx + A.t[k].coef * f.x;
            sy = sy + A.t[k].coef * f.y;
            sz = sz + A.t[k].coef * f.z;
        }
        o.x = y.x + dt * sx;
        o.y = y.y + dt * sy;
        o.z = y.z + dt * sz;
    } else if (FORM == SF_ACCUM) {
        o.x = y.x; o.y = y.y; o.z = y.z;
        for (int k = 0; k < A.nterms; k++) {
            const double c = dt * A.t[k].coef;
            const d4 f = get(k);
            o.x += c * f.x;
            o.y += c * f.y;
            o.z += c * f.z;
        }
    } else if (FORM == SF_RK4NP1) {
        const double b0 = A.c1 * dt, b1 = A.c2 * dt;
        const d4 f0 = get(0);
        const d4 f1 = get(1);
        const d4 f2 = get(2);
        const d4 f3 = get(3);
        o.x = y.x + b0 * (f0.x + f3.x) + b1 * (f1.x + f2.x);
        o.y = y.y + b0 * (f0.y + f3.y) + b1 * (f1.y + f2.y);
        o.z = y.z + b0 * (f0.z + f3.z) + b1 * (f1.z + f2.z);
    } else {  // SF_RKF78NP1: terms f0 f5 f6 f7 f8 f9 f10, coefficients b0 b5 b6 - b8 - b10
        const d4 f0 = get(0);
        const d4 f5 = get(1);
        const d4 f6 = get(2);
        const d4 f7 = get(3);
        const d4 f8 = get(4);
        const d4 f9 = get(5);
        const d4 f10 = get(6);
        const double b0 = A.t[0].coef, b5 = A.t[1].coef, b6 = A.t[2].coef, b8 = A.t[4].coef, b10 = A.t[6].coef;
        o.x = y.x + dt * (b0 * f0.x + b5 * f5.x + b6 * (f6.x + f7.x) + b8 * (f8.x + f9.x) + b10 * f10.x);
        o.y = y.y + dt * (b0 * f0.y + b5 * f5.y + b6 * (f6.y + f7.y) + b8 * (f8.y + f9.y) + b10 * f10.y);
        o.z = y.z + dt * (b0 * f0.z + b5 * f5.z + b6 * (f6.z + f7.z) + b8 * (f8.z + f9.z) + b10 * f10.z);
    }
    return o;
}

template <int FORM>
__device__ __forceinline__ d4 combine(const StageArgs& A, const d4& y, int i, bool vel, double dt)
{
    return combine_get<FORM>(A, y, dt, [&](int k) { return vel ? A.t[k].fa[i] : A.t[k].fv[i]; });
}

// SF_ACCUM and SF_DTSUM with the term count known at compile time: every derivative record
// is requested before the first is used (the loop of combine waits for each in turn), then
// the terms are added with the same operations in the same order as combine.
template <int FORM, int NT>
__device__ __forceinline__ d4 combine_nt(const StageArgs& A, const d4& y, int i, bool vel, double dt)
{
    static_assert(FORM == SF_ACCUM || FORM == SF_DTSUM, "forms with a term loop");
    d4 f[NT];
#pragma unroll
    for (int k = 0; k < NT; ++k) f[k] = vel ? A.t[k].fa[i] : A.t[k].fv[i];
    d4 o;
    o.w = y.w;
    if (FORM == SF_ACCUM) {
        o.x = y.x; o.y = y.y; o.z = y.z;
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const double c = dt * A.t[k].coef;
            o.x += c * f[k].x;
            o.y += c * f[k].y;
            o.z += c * f[k].z;
        }
    } else {
        double sx = A.t[0].coef * f[0].x, sy = A.t[0].coef * f[0].y, sz = A.t[0].coef * f[0].z;
#pragma unroll
        for (int k = 1; k < NT; ++k) {
            sx = sx + A.t[k].coef * f[k].x;
            sy = sy + A.t[k].coef * f[k].y;
            sz = sz + A.t[k].coef * f[k].z;
        }
        o.x = y.x + dt * sx;
        o.y = y.y + dt * sy;
        o.z = y.z + dt * sz;
    }
    return o;
}

}  // namespace red
