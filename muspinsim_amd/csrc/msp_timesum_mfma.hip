// Time sum of the real-coefficient routes on matrix cores (gfx950): FP32
// (mixed precision) and FP64 (plain) forms.
//
// P(t) = sum_ij C_ij cos(2 pi (E_i - E_j) t) for a real symmetric C (the
// fast_evolve / isotropic routes) is, with c_j(t) = cos(2 pi (E_j - Ec) t)
// and s_j(t) = sin(2 pi (E_j - Ec) t),
//   P(t) = sum_i c_i(t) (C c(t))_i + s_i(t) (C s(t))_i,
// a real GEMM Y = C [c | s] (n x n times n x 2T) followed by a contraction
// with the same phases.  Precision split:
//   - the phase arguments (E_j - Ec) t are formed and range-reduced to
//     [-1/2, 1/2] turns in FP64 (k_phase_table), so that only the reduced
//     angle's sine and cosine are FP32: phase errors stay ~1e-7 rad however
//     long the time axis;
//   - the GEMM runs on v_mfma_f32_16x16x4f32 (FP32 products and FP32
//     accumulation over j);
//   - the contraction over i of each 64-row block is summed in FP64, and so
//     are the blocks, the weight and the powder accumulation after it.
// MFMA lane layout (checked on the device with an in-bounds probe): A lane
// holds A[l % 16][l / 16], B holds B[l / 16][l % 16], D slot r holds
// D[4 (l / 16) + r][l % 16].
#include <hip/hip_runtime.h>

#include "msp_kernels.h"
#include "msp_mfma.h"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ inline f4 f4zero()
{
    f4 z = {0.f, 0.f, 0.f, 0.f};
    return z;
}

constexpr int kTT = 64;     // times per workgroup tile

inline unsigned cdivu(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// Phase table of one instance: Fc[j][t], Fs[j][t] for t < ldt (zero beyond
// ntimes), grid (ceil(n * ldt / 256), batch).
__global__ void __launch_bounds__(256) k_phase_table(int n, int ntimes, int ldt, const double *E,
                                                     const double *times, float *Fc, float *Fs)
{
    const int b = blockIdx.y;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)n * ldt) return;
    const int j = (int)(idx / ldt), t = (int)(idx - (size_t)j * ldt);
    const double *Eb = E + (size_t)b * n;
    const double Ec = 0.5 * (Eb[0] + Eb[n - 1]);
    float c = 0.f, s = 0.f;
    if (t < ntimes) {
        const double x = (Eb[j] - Ec) * times[t];         // turns
        const double r = x - rint(x);                       // [-1/2, 1/2]
        sincosf((float)(6.283185307179586 * r), &s, &c);
    }
    const size_t o = (size_t)b * n * ldt + idx;
    Fc[o] = c;
    Fs[o] = s;
}

// Y = C [c | s] on MFMA and P(t) = sum_i c_i Y^c_i + s_i Y^s_i for one
// 64-time tile of one instance, looping over 64-row blocks of C.  Wave w
// owns rows 16 w .. 16 w + 15 of each block and all 64 times (four 16-time
// column tiles for cos and four for sin: eight 16 x 16 accumulators).  The
// operands go through LDS in chunks of kKC columns of C (FP64 -> FP32 on
// the way) and kKC rows of the phase tables, shared by the four waves; the
// next chunk's loads are issued into registers before the current chunk's
// matrix-core work, so their latency overlaps it.
constexpr int kKC = 32;          // K (j) per staged chunk
constexpr int kLA = kKC + 1;     // padded row of the staged C block
constexpr int kLB = kTT + 4;     // padded row of a staged phase block

__global__ void __launch_bounds__(256) k_timesum_mfma(int n, int ntimes, int ldt, const double *C,
                                                      const float *Fc, const float *Fs,
                                                      const double *w, double *out)
{
    __shared__ float As[64 * kLA];
    __shared__ float Bc[kKC * kLB], Bs[kKC * kLB];
    __shared__ double red[4][kTT];
    int tt, b;
    msp_xcd_tile(gridDim.x, gridDim.y, tt, b);   // an instance's time tiles share one L2
    const int t0 = tt * kTT;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lq = lane >> 4;
    const double *Cb = C + (size_t)b * n * n;
    const float *Fcb = Fc + (size_t)b * n * ldt;
    const float *Fsb = Fs + (size_t)b * n * ldt;
    // staging assignment: C block 64 x kKC (8 per thread: row tid / 4,
    // columns 8 (tid % 4) ..), phase blocks kKC x 64 each (8 per thread:
    // row tid / 8, times 8 (tid % 8) ..)
    const int ar = tid >> 2, ac = (tid & 3) * 8;
    const int br = tid >> 3, bc = (tid & 7) * 8;
    double part[4] = {0.0, 0.0, 0.0, 0.0};        // per column tile, lane's time t0 + 16 ct + li
    const int nk = (n + kKC - 1) / kKC;
    for (int i0 = 0; i0 < n; i0 += 64) {
        f4 yc[4], ys[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            yc[ct] = f4zero();
            ys[ct] = f4zero();
        }
        const int ia = i0 + ar;
        const double *Crow = Cb + (size_t)((ia < n) ? ia : n - 1) * n;
        double ra[8];
        float rbc[8], rbs[8];
        auto load = [&](int j0) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int j = j0 + ac + u;
                ra[u] = (ia < n && j < n) ? Crow[j] : 0.0;
            }
            const int jb = j0 + br;
            const bool jok = jb < n;
            const float *pc = Fcb + (size_t)(jok ? jb : 0) * ldt + t0 + bc;
            const float *ps = Fsb + (size_t)(jok ? jb : 0) * ldt + t0 + bc;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                rbc[u] = jok ? pc[u] : 0.f;
                rbs[u] = jok ? ps[u] : 0.f;
            }
        };
        load(0);
        for (int kc = 0; kc < nk; ++kc) {
            __syncthreads();                      // the previous chunk's reads are done
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                As[ar * kLA + ac + u] = (float)ra[u];
                Bc[br * kLB + bc + u] = rbc[u];
                Bs[br * kLB + bc + u] = rbs[u];
            }
            __syncthreads();
            if (kc + 1 < nk) load((kc + 1) * kKC);   // in flight during the MFMAs
#pragma unroll
            for (int h = 0; h < kKC / 4; ++h) {
                const float a = As[(16 * wave + li) * kLA + 4 * h + lq];
#pragma unroll
                for (int ct = 0; ct < 4; ++ct) {
                    const float vc = Bc[(4 * h + lq) * kLB + 16 * ct + li];
                    const float vs = Bs[(4 * h + lq) * kLB + 16 * ct + li];
                    yc[ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, vc, yc[ct], 0, 0, 0);
                    ys[ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, vs, ys[ct], 0, 0, 0);
                }
            }
        }
        // contraction: slot r of tile ct is Y[i0 + 16 wave + 4 lq + r][t0 + 16 ct + li]
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            const int t = t0 + 16 * ct + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + 16 * wave + 4 * lq + r;
                if (i < n) {
                    const size_t o = (size_t)i * ldt + t;
                    part[ct] += (double)Fcb[o] * (double)yc[ct][r] + (double)Fsb[o] * (double)ys[ct][r];
                }
            }
        }
    }
    // sum the four row groups of the wave (lanes li, li + 16, li + 32, li + 48)
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        double v = part[ct];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        if (lq == 0) red[wave][16 * ct + li] = v;
    }
    __syncthreads();
    if (tid < kTT) {
        const int t = t0 + tid;
        if (t < ntimes) {
            const double P = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
            out[(size_t)b * ntimes + t] = (w ? w[b] : 1.0) * P;
        }
    }
}

// ---------------------------------------------------------------------------
// The same time sum in FP64 on v_mfma_f64_16x16x4f64, for the plain path: the
// phase table holds cos / sin of 2 pi (E_j - Ec) t in FP64, formed exactly as
// k_timesum forms them (same expression, same rounding), but once per (j, t)
// instead of once per 64-row block of C; products, sums and the contraction
// are FP64.  D slot r of the FP64 MFMA holds D[4 r + l / 16][l % 16] (A and
// B lanes as in the FP32 form).

__global__ void __launch_bounds__(256) k_phase_table64(int n, int ntimes, int ldt, const double *E,
                                                       const double *times, double *Fc, double *Fs)
{
    const int b = blockIdx.y;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)n * ldt) return;
    const int j = (int)(idx / ldt), t = (int)(idx - (size_t)j * ldt);
    const double *Eb = E + (size_t)b * n;
    const double Ec = 0.5 * (Eb[0] + Eb[n - 1]);
    double c = 0.0, s = 0.0;
    if (t < ntimes) {
        const double phi = 6.283185307179586 * (Eb[j] - Ec) * times[t];
        sincos(phi, &s, &c);
    }
    const size_t o = (size_t)b * n * ldt + idx;
    Fc[o] = c;
    Fs[o] = s;
}

constexpr int kKC64 = 16;          // K (j) per staged chunk
constexpr int kLA64 = kKC64 + 1;   // padded row of the staged C block
constexpr int kLB64 = kTT + 2;     // padded row of a staged phase block

// Y = C [c | s] on FP64 MFMA and P(t) = sum_i c_i Y^c_i + s_i Y^s_i for one
// 64-time tile of one instance (layout and staging as k_timesum_mfma, with
// 16-deep chunks: four elements per thread for the C block and for each
// phase block).
// (three waves per SIMD: 152 VGPRs with the accumulators among them, no
// spills; the compiler's default form held them in 64 AGPRs, two waves)
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_timesum_mfma64(int n, int ntimes, int ldt, const double *C,
                                                        const double *Fc, const double *Fs,
                                                        const double *w, double *out)
{
    __shared__ double As[64 * kLA64];
    __shared__ double Bc[kKC64 * kLB64], Bs[kKC64 * kLB64];
    __shared__ double red[4][kTT];
    int tt, b;
    msp_xcd_tile(gridDim.x, gridDim.y, tt, b);   // an instance's time tiles share one L2
    const int t0 = tt * kTT;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lq = lane >> 4;
    const double *Cb = C + (size_t)b * n * n;
    const double *Fcb = Fc + (size_t)b * n * ldt;
    const double *Fsb = Fs + (size_t)b * n * ldt;
    // staging: C block 64 x 16 (row tid / 4, columns 4 (tid % 4) ..), phase
    // blocks 16 x 64 (row tid / 16, times 4 (tid % 16) ..)
    const int ar = tid >> 2, ac = (tid & 3) * 4;
    const int br = tid >> 4, bc = (tid & 15) * 4;
    double part[4] = {0.0, 0.0, 0.0, 0.0};        // per column tile, lane's time t0 + 16 ct + li
    const int nk = (n + kKC64 - 1) / kKC64;
    for (int i0 = 0; i0 < n; i0 += 64) {
        msp_d4 yc[4], ys[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            yc[ct] = d4zero();
            ys[ct] = d4zero();
        }
        const int ia = i0 + ar;
        const double *Crow = Cb + (size_t)((ia < n) ? ia : n - 1) * n;
        double ra[4], rbc[4], rbs[4];
        auto load = [&](int j0) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = j0 + ac + u;
                ra[u] = (ia < n && j < n) ? Crow[j] : 0.0;
            }
            const int jb = j0 + br;
            const bool jok = jb < n;
            const double *pc = Fcb + (size_t)(jok ? jb : 0) * ldt + t0 + bc;
            const double *ps = Fsb + (size_t)(jok ? jb : 0) * ldt + t0 + bc;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                rbc[u] = jok ? pc[u] : 0.0;
                rbs[u] = jok ? ps[u] : 0.0;
            }
        };
        load(0);
        for (int kc = 0; kc < nk; ++kc) {
            __syncthreads();                      // the previous chunk's reads are done
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                As[ar * kLA64 + ac + u] = ra[u];
                Bc[br * kLB64 + bc + u] = rbc[u];
                Bs[br * kLB64 + bc + u] = rbs[u];
            }
            __syncthreads();
            if (kc + 1 < nk) load((kc + 1) * kKC64);   // in flight during the MFMAs
#pragma unroll
            for (int h = 0; h < kKC64 / 4; ++h) {
                const double a = As[(16 * wave + li) * kLA64 + 4 * h + lq];
#pragma unroll
                for (int ct = 0; ct < 4; ++ct) {
                    const double vc = Bc[(4 * h + lq) * kLB64 + 16 * ct + li];
                    const double vs = Bs[(4 * h + lq) * kLB64 + 16 * ct + li];
                    yc[ct] = mfma_f64(a, vc, yc[ct]);
                    ys[ct] = mfma_f64(a, vs, ys[ct]);
                }
            }
        }
        // contraction: slot r of tile ct is Y[i0 + 16 wave + 4 r + lq][t0 + 16 ct + li]
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            const int t = t0 + 16 * ct + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + 16 * wave + 4 * r + lq;
                if (i < n) {
                    const size_t o = (size_t)i * ldt + t;
                    part[ct] += Fcb[o] * yc[ct][r] + Fsb[o] * ys[ct][r];
                }
            }
        }
    }
    // sum the four row groups of the wave (lanes li, li + 16, li + 32, li + 48)
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        double v = part[ct];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        if (lq == 0) red[wave][16 * ct + li] = v;
    }
    __syncthreads();
    if (tid < kTT) {
        const int t = t0 + tid;
        if (t < ntimes) {
            const double P = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
            out[(size_t)b * ntimes + t] = (w ? w[b] : 1.0) * P;
        }
    }
}

}  // namespace

int timesum_mfma_ldt(int ntimes) { return (ntimes + kTT - 1) / kTT * kTT; }

size_t timesum_mfma_workspace(int n, int batch, int ntimes)
{
    return (size_t)2 * n * timesum_mfma_ldt(ntimes) * sizeof(float) * (size_t)batch;
}

hipError_t launch_timesum_mfma(int n, int batch, int ntimes, const double *E, const double *C,
                               const double *times, const double *w, double *out, void *work,
                               hipStream_t s)
{
    const int ldt = timesum_mfma_ldt(ntimes);
    float *Fc = (float *)work;
    float *Fs = Fc + (size_t)n * ldt * batch;
    {
        ProfScope ps_("k_phase_table", s);
        hipLaunchKernelGGL(k_phase_table, dim3(cdivu((size_t)n * ldt, 256), batch), dim3(256), 0, s,
                           n, ntimes, ldt, E, times, Fc, Fs);
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    {
        ProfScope ps_("k_timesum_mfma", s);
        hipLaunchKernelGGL(k_timesum_mfma, dim3(ldt / kTT, batch), dim3(256), 0, s, n, ntimes, ldt,
                           C, Fc, Fs, w, out);
    }
    return hipGetLastError();
}

size_t timesum_mfma64_workspace(int n, int batch, int ntimes)
{
    return (size_t)2 * n * timesum_mfma_ldt(ntimes) * sizeof(double) * (size_t)batch;
}

hipError_t launch_timesum_mfma64(int n, int batch, int ntimes, const double *E, const double *C,
                                 const double *times, const double *w, double *out, void *work,
                                 hipStream_t s)
{
    const int ldt = timesum_mfma_ldt(ntimes);
    double *Fc = (double *)work;
    double *Fs = Fc + (size_t)n * ldt * batch;
    {
        ProfScope ps_("k_phase_table", s);
        hipLaunchKernelGGL(k_phase_table64, dim3(cdivu((size_t)n * ldt, 256), batch), dim3(256), 0, s,
                           n, ntimes, ldt, E, times, Fc, Fs);
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    {
        ProfScope ps_("k_timesum_mfma64", s);
        hipLaunchKernelGGL(k_timesum_mfma64, dim3(ldt / kTT, batch), dim3(256), 0, s, n, ntimes, ldt,
                           C, Fc, Fs, w, out);
    }
    return hipGetLastError();
}
