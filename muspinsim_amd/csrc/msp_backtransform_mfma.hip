// U = Q Z on FP64 matrix cores: the blocked back-transformation of
// msp_tridiag.h (U <- (I - V T V^H) U for reflector blocks of MSP_BT_NB = 16,
// last block first) with both block products on v_mfma_f64_16x16x4f64.
//
// One workgroup (4 waves) owns a 16-column chunk of U in LDS (split real and
// imaginary planes).  Per reflector block:
//   W  = V^H U      16 x 16, the k (row) dimension split over the 4 waves,
//                   partial tiles summed through LDS
//   W2 = T W        16 x 16, VALU
//   U -= V W2       16-row tiles distributed over the waves, k = 16
#include <hip/hip_runtime.h>

#include "msp_kernels.h"
#include "msp_mfma.h"
#include "msp_tridiag.h"

namespace {

constexpr int CW = 16;           // columns per workgroup
constexpr int LDU = CW + 1;      // padded LDS row stride (doubles)
static_assert(MSP_BT_NB == 16, "the MFMA back-transformation tiles reflector blocks of 16");

__global__ void __launch_bounds__(256) k_backtransform_mfma(int n, const double *H, double *work,
                                                            double *U)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int b = blockIdx.y;
    const int c0 = blockIdx.x * CW;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lq = lane >> 4;
    double *Ur = (double *)smem;                 // n x LDU
    double *Ui = Ur + (size_t)n * LDU;           // n x LDU
    double *Wp = Ui + (size_t)n * LDU;           // 4 waves x 2 x 256 partial W
    double *W2 = Wp + 4 * 2 * 256;               // 2 x 256 (re plane, im plane)

    const cdbl *A = (const cdbl *)H + (size_t)b * n * n;
    double *base = work + (size_t)b * dc_work_doubles(n);
    const double *Z = dc_result_ptr(n, base);
    const cdbl *Tall = dc_tblocks(n, base);

    for (int idx = tid; idx < n * CW; idx += 256) {
        const int r = idx / CW, c = idx % CW;
        Ur[r * LDU + c] = (c0 + c < n) ? Z[(size_t)r * n + c0 + c] : 0.0;
        Ui[r * LDU + c] = 0.0;
    }
    __syncthreads();

    const int nref = n - 1;
    const int nblk = (nref + MSP_BT_NB - 1) / MSP_BT_NB;
    for (int blk = nblk - 1; blk >= 0; --blk) {
        const int k0 = blk * MSP_BT_NB;
        const int nbk = (nref - k0 < MSP_BT_NB) ? nref - k0 : MSP_BT_NB;
        const int m = n - k0 - 1;                // rows k0+1 .. n-1
        const int nks = (m + 3) / 4;             // k-steps of 4 rows
        // ---- W = V^H U(k0+1.., :), partial per wave
        msp_d4 wr = d4zero(), wi = d4zero();
        for (int s = wave; s < nks; s += 4) {
            const int r = k0 + 1 + 4 * s + lq;   // row of this lane's operands
            double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
            if (r < n) {
                if (li < nbk) {
                    cdbl v = vblock(n, A, k0, r, li);
                    ar = v.re;
                    ai = -v.im;                  // conj(V)
                }
                br = Ur[r * LDU + li];
                bi = Ui[r * LDU + li];
            }
            mfma_cmac(ar, ai, br, bi, wr, wi);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = 4 * q + lq;          // reflector index j
            Wp[(wave * 2 + 0) * 256 + row * 16 + li] = wr[q];
            Wp[(wave * 2 + 1) * 256 + row * 16 + li] = wi[q];
        }
        __syncthreads();
        // ---- W2 = T W (VALU), one element per thread
        {
            const int j = tid / 16, c = tid % 16;
            const cdbl *T = Tall + (size_t)blk * MSP_BT_NB * MSP_BT_NB;
            double sr = 0.0, si = 0.0;
            for (int l = j; l < nbk; ++l) {
                double wre = Wp[0 * 256 + l * 16 + c] + Wp[2 * 256 + l * 16 + c] +
                             Wp[4 * 256 + l * 16 + c] + Wp[6 * 256 + l * 16 + c];
                double wim = Wp[1 * 256 + l * 16 + c] + Wp[3 * 256 + l * 16 + c] +
                             Wp[5 * 256 + l * 16 + c] + Wp[7 * 256 + l * 16 + c];
                const cdbl t = T[j * MSP_BT_NB + l];
                sr += t.re * wre - t.im * wim;
                si += t.re * wim + t.im * wre;
            }
            W2[j * 16 + c] = (j < nbk) ? sr : 0.0;
            W2[256 + j * 16 + c] = (j < nbk) ? si : 0.0;
        }
        __syncthreads();
        // ---- U(rows, :) -= V W2 on 16-row tiles
        const int ntile = (m + 15) / 16;
        for (int t = wave; t < ntile; t += 4) {
            const int rt = k0 + 1 + 16 * t;
            msp_d4 dr, di;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = rt + 4 * q + lq;
                dr[q] = (r < n) ? Ur[r * LDU + li] : 0.0;
                di[q] = (r < n) ? Ui[r * LDU + li] : 0.0;
            }
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int r = rt + li;           // A row (tile row li)
                const int j = 4 * s + lq;        // A column / B row
                double ar = 0.0, ai = 0.0;
                if (r < n && j < nbk) {
                    cdbl v = vblock(n, A, k0, r, j);
                    ar = v.re;
                    ai = v.im;
                }
                const double br = W2[j * 16 + li], bi = W2[256 + j * 16 + li];
                // D -= A B
                mfma_cmac(-ar, -ai, br, bi, dr, di);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = rt + 4 * q + lq;
                if (r < n) {
                    Ur[r * LDU + li] = dr[q];
                    Ui[r * LDU + li] = di[q];
                }
            }
        }
        __syncthreads();
    }
    cdbl *Ub = (cdbl *)U + (size_t)b * n * n;
    for (int idx = tid; idx < n * CW; idx += 256) {
        const int r = idx / CW, c = idx % CW;
        if (c0 + c < n) Ub[(size_t)r * n + c0 + c] = cmk(Ur[r * LDU + c], Ui[r * LDU + c]);
    }
}

}  // namespace

hipError_t launch_backtransform_mfma(int n, int batch, const double *H, double *work, double *U,
                                     hipStream_t s)
{
    size_t lds = sizeof(double) * (2 * (size_t)n * LDU + 4 * 2 * 256 + 2 * 256);
    hipError_t err = hipFuncSetAttribute((const void *)k_backtransform_mfma,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_backtransform_mfma, dim3((n + CW - 1) / CW, batch), dim3(256), lds, s, n,
                       H, work, U);
    return hipGetLastError();
}
