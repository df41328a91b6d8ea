// Workgroup-cooperative ("team") building blocks shared by the HIP kernels
// and by a host build used to test the same algorithms on the CPU.
//
// A team function is a sequence of phases.  Each phase is a parallel loop
// MSP_PFOR(i, count) whose iterations are independent, and phases are
// separated by MSP_SYNC().  On the device the loop is strided over the
// workgroup's threads and MSP_SYNC is __syncthreads(); in the host build
// (MSP_HOST_BUILD) the team is one thread, the loop runs sequentially and
// MSP_SYNC is a no-op, so a host run executes exactly the device schedule.
#pragma once

#include <math.h>
#include <stdint.h>

#ifdef MSP_HOST_BUILD
#define MSP_HD inline
#define MSP_DEV inline
#define MSP_TID 0
#define MSP_NT 1
#define MSP_SYNC() ((void)0)
#else
#include <hip/hip_runtime.h>
#define MSP_HD __host__ __device__ inline
#define MSP_DEV __device__ inline
#define MSP_TID ((int)threadIdx.x)
#define MSP_NT ((int)blockDim.x)
#define MSP_SYNC() __syncthreads()
#endif

#define MSP_PFOR(i, count) for (int i = MSP_TID; i < (count); i += MSP_NT)

// Complex double, interleaved (re, im), the layout of numpy complex128.
struct cdbl {
    double re, im;
};

MSP_HD cdbl cmk(double r, double i) { cdbl z; z.re = r; z.im = i; return z; }
MSP_HD cdbl cadd(cdbl a, cdbl b) { return cmk(a.re + b.re, a.im + b.im); }
MSP_HD cdbl csub(cdbl a, cdbl b) { return cmk(a.re - b.re, a.im - b.im); }
MSP_HD cdbl cmul(cdbl a, cdbl b) { return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
// conj(a) * b
MSP_HD cdbl cmulc(cdbl a, cdbl b) { return cmk(a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re); }
MSP_HD cdbl cscale(cdbl a, double s) { return cmk(a.re * s, a.im * s); }
MSP_HD cdbl cconj(cdbl a) { return cmk(a.re, -a.im); }
MSP_HD double cabs2(cdbl a) { return a.re * a.re + a.im * a.im; }
MSP_HD cdbl cdivc(cdbl a, cdbl b)
{
    // a / b with Smith's scaling
    if (fabs(b.re) >= fabs(b.im)) {
        double r = b.im / b.re, den = b.re + b.im * r;
        return cmk((a.re + a.im * r) / den, (a.im - a.re * r) / den);
    } else {
        double r = b.re / b.im, den = b.re * r + b.im;
        return cmk((a.re * r + a.im) / den, (a.im * r - a.re) / den);
    }
}
// acc += a * b
MSP_HD void cfma(cdbl &acc, cdbl a, cdbl b)
{
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(-a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(a.im, b.re, acc.im);
}
// acc += conj(a) * b
MSP_HD void cfmac(cdbl &acc, cdbl a, cdbl b)
{
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(-a.im, b.re, acc.im);
}

#ifndef MSP_HOST_BUILD
// One DPP lane move of a double (two 32-bit halves); lanes outside ROWMASK
// receive 0.  CTRL is a gfx9 DPP control: quad_perm, row mirrors or
// row_bcast15/31.
template <int CTRL, int ROWMASK>
__device__ inline double dpp_f64(double v)
{
    const long long x = __double_as_longlong(v);
    int lo = (int)(x & 0xffffffffll), hi = (int)(x >> 32);
    lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, ROWMASK, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, ROWMASK, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 64 lanes of a wave on the VALU (no LDS traffic); the total is
// exact in lane 63 only: pairs, quads, half rows, rows, then rows 0+1 and
// 2+3 through row_bcast15 and the four rows through row_bcast31.
__device__ inline double wave_sum_to63(double v)
{
    v += dpp_f64<0xB1, 0xF>(v);    // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E, 0xF>(v);    // quad_perm [2,3,0,1]
    v += dpp_f64<0x141, 0xF>(v);   // row_half_mirror
    v += dpp_f64<0x140, 0xF>(v);   // row_mirror
    v += dpp_f64<0x142, 0xA>(v);   // row_bcast15 -> rows 1, 3
    v += dpp_f64<0x143, 0xC>(v);   // row_bcast31 -> rows 2, 3
    return v;
}

// Wave sum broadcast to every lane.
__device__ inline double wave_sum_all(double v)
{
    const double t = wave_sum_to63(v);
    const long long x = __double_as_longlong(t);
    const int lo = __builtin_amdgcn_readlane((int)(x & 0xffffffffll), 63);
    const int hi = __builtin_amdgcn_readlane((int)(x >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
#endif

// Team-wide sum of one double per thread.  `red` is team scratch of at least
// MSP_NT doubles (LDS on the device).  Every thread returns the total.
MSP_HD double team_sum(double v, double *red)
{
#ifdef MSP_HOST_BUILD
    (void)red;
    return v;
#else
    const int tid = MSP_TID, nt = MSP_NT;
    // wave-level reduction first (64 lanes), then across waves through LDS
    v = wave_sum_to63(v);
    const int wave = tid >> 6, nw = (nt + 63) >> 6;
    __syncthreads();
    if ((tid & 63) == 63) red[wave] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += red[w];
    __syncthreads();
    return t;
#endif
}

MSP_HD double team_max(double v, double *red)
{
#ifdef MSP_HOST_BUILD
    (void)red;
    return v;
#else
    const int tid = MSP_TID, nt = MSP_NT;
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    const int wave = tid >> 6, nw = (nt + 63) >> 6;
    __syncthreads();
    if ((tid & 63) == 0) red[wave] = v;
    __syncthreads();
    double t = red[0];
    for (int w = 1; w < nw; ++w) t = fmax(t, red[w]);
    __syncthreads();
    return t;
#endif
}
