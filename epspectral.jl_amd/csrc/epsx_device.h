// Device helpers shared by the kernels of libepsx.
#pragma once
#include <hip/hip_runtime.h>

namespace epsx {

// e^y = M 2^k with M = e^r, |r| <= ln2/2 (Cody-Waite reduction, fdlibm's split of ln 2).
__device__ inline void split_exp(double y, double &M, int &k)
{
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    const double log2e = 1.44269504088896338700e+00;
    if (!(y == y)) { M = y; k = 0; return; }
    double yk = rint(y * log2e);
    if (yk > 1.0e6) yk = 1.0e6;
    if (yk < -1.0e6) yk = -1.0e6;
    const double r = (y - yk * ln2_hi) - yk * ln2_lo;
    M = exp(r);
    k = (int)yk;
}

// FP64 -> BF16 (as FP32 first, then round to nearest even on the upper 16 bits; NaN stays NaN).
__device__ inline unsigned short to_bf16(double x)
{
    const float f = (float)x;
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (unsigned short)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
    return (unsigned short)(u >> 16);
}

}  // namespace epsx
