// wr_math.h -- device elementary functions of the radiation kernels.
//
// FP64: exp/log evaluated with a fixed sequence of IEEE basic operations (Cody-Waite reduction
// with a 42-bit ln2 head, Taylor series of e^r through r^13; log via s = f/(2+f) and the atanh
// series through s^20).  Compiled without contraction, so every table index and McICA bit derived
// from them is reproducible bit for bit on any IEEE host that evaluates the same sequence.
// FP32 (mixed path): hardware v_exp_f32 / v_log_f32 through __expf/__logf.
// Philox4x32-10 (Salmon et al., SC'11) keys the McICA sub-column stream by (seed, column).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wr {

__device__ __forceinline__ double pow2i(int k)
{
    return __longlong_as_double((long long)(k + 1023) << 52);
}

__device__ __forceinline__ double exp_d(double x)
{
#pragma clang fp contract(off)
    if (x != x) return x;
    if (x > 709.782712893384) return __longlong_as_double(0x7ff0000000000000LL);
    if (x < -745.1332191019412) return 0.0;
    const double kf = __builtin_rint(x * 0x1.71547652b82fep+0);
    const double r = (x - kf * 0x1.62e42fefa3800p-1) - kf * 0x1.ef35793c76730p-45;
    double q = 0x1.6124613a86d09p-33;
    q = q * r + 0x1.1eed8eff8d898p-29;
    q = q * r + 0x1.ae64567f544e4p-26;
    q = q * r + 0x1.27e4fb7789f5cp-22;
    q = q * r + 0x1.71de3a556c734p-19;
    q = q * r + 0x1.a01a01a01a01ap-16;
    q = q * r + 0x1.a01a01a01a01ap-13;
    q = q * r + 0x1.6c16c16c16c17p-10;
    q = q * r + 0x1.1111111111111p-7;
    q = q * r + 0x1.5555555555555p-5;
    q = q * r + 0x1.5555555555555p-3;
    q = q * r + 0x1.0000000000000p-1;
    const double p = 1.0 + (r + (r * r) * q);
    const int k = (int)kf;
    if (k > 1023) return (p * pow2i(1023)) * 2.0;
    if (k < -1022) return (p * pow2i(k + 600)) * pow2i(-600);
    return p * pow2i(k);
}

__device__ __forceinline__ double log_d(double x)
{
#pragma clang fp contract(off)
    if (x != x || x < 0.0) return __longlong_as_double(0x7ff8000000000000LL);
    if (x == 0.0) return __longlong_as_double((long long)0xfff0000000000000ULL);
    if (x == __longlong_as_double(0x7ff0000000000000LL)) return x;
    unsigned long long b = (unsigned long long)__double_as_longlong(x);
    int eadj = 0;
    if ((b >> 52) == 0) {
        x = x * 0x1p54;
        b = (unsigned long long)__double_as_longlong(x);
        eadj = -54;
    }
    int e = (int)((b >> 52) & 0x7ff) - 1023 + eadj;
    double m = __longlong_as_double((long long)((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL));
    if (m > 0x1.6a09e667f3bcdp+0) {
        m = m * 0.5;
        e += 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double R = 0x1.8618618618618p-4;
    R = R * z + 0x1.af286bca1af28p-4;
    R = R * z + 0x1.e1e1e1e1e1e1ep-4;
    R = R * z + 0x1.1111111111111p-3;
    R = R * z + 0x1.3b13b13b13b14p-3;
    R = R * z + 0x1.745d1745d1746p-3;
    R = R * z + 0x1.c71c71c71c71cp-3;
    R = R * z + 0x1.2492492492492p-2;
    R = R * z + 0x1.999999999999ap-2;
    R = R * z + 0x1.5555555555555p-1;
    R = R * z;
    const double hfsq = (0.5 * f) * f;
    const double ef = (double)e;
    return ef * 0x1.62e42fefa3800p-1 + (f - (hfsq - (s * (hfsq + R) + ef * 0x1.ef35793c76730p-45)));
}

__device__ __forceinline__ double pow_d(double x, double y) { return exp_d(y * log_d(x)); }

// precision dispatch used by templated kernels
template <typename R> struct M;
template <> struct M<double> {
    __device__ static __forceinline__ double exp(double x) { return exp_d(x); }
    __device__ static __forceinline__ double log(double x) { return log_d(x); }
    __device__ static __forceinline__ double pow(double x, double y) { return pow_d(x, y); }
};
template <> struct M<float> {
    __device__ static __forceinline__ float exp(float x) { return __expf(x); }
    __device__ static __forceinline__ float log(float x) { return __logf(x); }
    __device__ static __forceinline__ float pow(float x, float y) { return __expf(y * __logf(x)); }
};

__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                              uint32_t k1, uint32_t out[4])
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

}  // namespace wr
