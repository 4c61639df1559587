// ---------------------------------------------------------------------------
// bh_mathprobe.hip -- bhMathProbe: the library's own approximations of
// elementary functions, evaluated on the GPU exactly as the fast trace
// kernels evaluate them (this object is compiled with the fast object's
// flags), so their error bounds can be tested against high-precision
// references on the host.
// ---------------------------------------------------------------------------
#include "bh_band.h"
#include "bh_internal.h"

namespace bhk {

__global__ __launch_bounds__(256) void mathProbeKernel(int fn, const double* __restrict__ in, int n,
                                                       double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = in[2 * i], b = in[2 * i + 1];
    double r0 = 0.0, r1 = 0.0;
    switch (fn)
    {
    case BH_MATH_ATAN2_POLY: r0 = atan2Poly((float)a, (float)b); break;
    case BH_MATH_FAST_SINCOS: r0 = MathFast::sin_((float)a); r1 = MathFast::cos_((float)a); break;
    case BH_MATH_FAST_RCP: r0 = MathFast::rcp((float)a); break;
    case BH_MATH_FAST_RSQ: r0 = MathFast::rsq((float)a); break;
    case BH_MATH_FAST_SQRT: r0 = MathFast::sqrt_((float)a); break;
    case BH_MATH_FAST_EXP: r0 = MathFast::exp_((float)a); break;
    case BH_MATH_FAST_LOG: r0 = MathFast::log_((float)a); break;
    case BH_MATH_FAST_POW: r0 = MathFast::pow_((float)a, (float)b); break;
    case BH_MATH_D64_SINCOS: D64Fast::sincos_(a, r0, r1); break;
    case BH_MATH_D64_SINCOS_NEAR:
    {
        double s0, c0;
        D64Fast::sincos_(a, s0, c0);
        D64Fast::sincosNear(a, s0, c0, b, r0, r1);
        break;
    }
    case BH_MATH_D64_RCP: r0 = D64Fast::rcp(a); break;
    case BH_MATH_BLACKBODY:
    {
        const v3 f = blackbody<MathFast>((float)a), e = blackbody<MathExact>((float)a);
        const int c = (int)b;
        r0 = c == 0 ? f.x : (c == 1 ? f.y : f.z);
        r1 = c == 0 ? e.x : (c == 1 ? e.y : e.z);
        break;
    }
    case BH_MATH_WRAP_PHI: r0 = wrapDeltaPhiFast((float)a); r1 = wrapDeltaPhi((float)a); break;
    default: r0 = r1 = __builtin_nan("");
    }
    out[2 * i] = r0;
    out[2 * i + 1] = r1;
}

}  // namespace bhk

extern "C" bh_error_t bhMathProbe(int fn, const double* devIn, int n, double* devOut, bh_stream_t stream)
{
    if (n < 0 || (n > 0 && (!devIn || !devOut)) || fn < 0 || fn > BH_MATH_WRAP_PHI) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(bhk::mathProbeKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       fn, devIn, n, devOut);
    return hipGetLastError();
}
