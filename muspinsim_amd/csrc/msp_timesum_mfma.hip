// Time sum of the real-coefficient routes on FP32 matrix cores (gfx950).
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
// column tiles for cos and four for sin: eight 16 x 16 accumulators).
__global__ void __launch_bounds__(256) k_timesum_mfma(int n, int ntimes, int ldt, const double *C,
                                                      const float *Fc, const float *Fs,
                                                      const double *w, double *out)
{
    __shared__ double red[4][kTT];
    int tt, b;
    msp_xcd_tile(gridDim.x, gridDim.y, tt, b);   // an instance's time tiles share one L2
    const int t0 = tt * kTT;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lq = lane >> 4;
    const double *Cb = C + (size_t)b * n * n;
    const float *Fcb = Fc + (size_t)b * n * ldt;
    const float *Fsb = Fs + (size_t)b * n * ldt;
    double part[4] = {0.0, 0.0, 0.0, 0.0};        // per column tile, lane's time t0 + 16 ct + li
    for (int i0 = 0; i0 < n; i0 += 64) {
        const int ia = i0 + 16 * wave + li;       // A-operand row of this lane
        const bool iok = ia < n;
        const double *Crow = Cb + (size_t)(iok ? ia : 0) * n;
        f4 yc[4], ys[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            yc[ct] = f4zero();
            ys[ct] = f4zero();
        }
        int j0 = 0;
        for (; j0 + 16 <= n; j0 += 16) {
            float a[4], bc[4][4], bs[4][4];
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const int j = j0 + 4 * h + lq;
                a[h] = iok ? (float)Crow[j] : 0.f;
                const float *pc = Fcb + (size_t)j * ldt + t0 + li;
                const float *ps = Fsb + (size_t)j * ldt + t0 + li;
#pragma unroll
                for (int ct = 0; ct < 4; ++ct) {
                    bc[h][ct] = pc[16 * ct];
                    bs[h][ct] = ps[16 * ct];
                }
            }
#pragma unroll
            for (int h = 0; h < 4; ++h)
#pragma unroll
                for (int ct = 0; ct < 4; ++ct) {
                    yc[ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[h], bc[h][ct], yc[ct], 0, 0, 0);
                    ys[ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[h], bs[h][ct], ys[ct], 0, 0, 0);
                }
        }
        for (; j0 < n; j0 += 4) {
            const int j = j0 + lq;
            const bool jok = j < n;
            const float a = (iok && jok) ? (float)Crow[j] : 0.f;
            const int jc = jok ? j : n - 1;
            const float *pc = Fcb + (size_t)jc * ldt + t0 + li;
            const float *ps = Fsb + (size_t)jc * ldt + t0 + li;
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                const float vc = jok ? pc[16 * ct] : 0.f, vs = jok ? ps[16 * ct] : 0.f;
                yc[ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, vc, yc[ct], 0, 0, 0);
                ys[ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, vs, ys[ct], 0, 0, 0);
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
    if (threadIdx.x < kTT) {
        const int t = t0 + threadIdx.x;
        if (t < ntimes) {
            const double P = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] +
                             red[3][threadIdx.x];
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
