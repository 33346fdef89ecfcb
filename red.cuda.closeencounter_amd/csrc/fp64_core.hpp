// IEEE FP64 square root and quotient cores for the force sweeps (device code).
#pragma once

#include <hip/hip_runtime.h>

namespace red {

// IEEE square root and quotient for operands well inside the normal range.  For such
// operands the compiler's correctly rounded sequences reduce to these cores: the square
// root's range scaling is by 2^0, its special-value select keeps the core result, the
// quotient's v_div_scale returns its operands unchanged with no scale flag (exponent gap
// < 768, no denormal numerator, reciprocal or quotient), v_div_fmas is then a plain fused
// multiply-add and v_div_fixup passes a normal quotient through.  So the results are
// bit-identical, with about ten fewer instructions per pair.  Callers take this path only
// when every lane's operands lie in [2^-300, 2^300] (in_core_range); otherwise they use
// sqrt() and the division operator.
__device__ __forceinline__ bool in_core_range(double x)
{
    return (unsigned)(__double2hiint(x) - 0x2D300000) < 0x25800000u;   // 2^-300 <= x < 2^300
}

__device__ __forceinline__ double sqrt_core(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double e = fma(-g, g, x);
    g = fma(e, h, g);
    e = fma(-g, g, x);
    return fma(e, h, g);
}

__device__ __forceinline__ double div_core(double num, double den)
{
    double rc = __builtin_amdgcn_rcp(den);
    double e = fma(-den, rc, 1.0);
    rc = fma(rc, e, rc);
    e = fma(-den, rc, 1.0);
    rc = fma(rc, e, rc);
    const double q = num * rc;
    const double rem = fma(-den, q, num);
    return fma(rem, rc, q);
}

}  // namespace red
