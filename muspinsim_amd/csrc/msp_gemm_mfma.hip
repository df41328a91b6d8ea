// Batched complex C = A^H B (n x n, row-major) on FP64 matrix cores.
// A workgroup (4 waves) owns a 64 x 64 output block of one instance, each
// wave a 32 x 32 quarter as 2 x 2 tiles of 16 x 16 accumulators (real and
// imaginary planes).  Operands are read in place: the A operand of a k-step,
// (A^H)[i][k] = conj(A[k][i]), and the B operand B[k][j] are both rows of
// 16 consecutive complex entries, so a k-step is 4 coalesced row segments
// per operand; complex products use four real MFMAs (msp_mfma.h).
#include <hip/hip_runtime.h>

#include "../../include/muspin_hip.h"
#include "msp_kernels.h"
#include "msp_mfma.h"
#include "msp_team.h"

namespace {

// UPPER: only the 64 x 64 blocks on or above the block diagonal (for a
// Hermitian product whose consumer reads those blocks only).
// MIX: B = (f v.S_mu (x) 1) A formed while loading, for a spin-1/2 muon at
// the stride its system gives: B[k] = M[a][a] A[k] + M[a][1-a] A[k'] with a
// the muon digit of row k and k' the row with the other digit
// (the product k_apply_muon writes, f = 2 for Pauli matrices).
// (three waves per SIMD: the accumulators then stay in VGPRs, 149 registers
// with MIX, no spills)
template <bool UPPER, bool MIX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
k_gemm_cn_mfma(int n, const double *A, const double *B, double *C, const MspSystem *systems,
               const int32_t *sys_id, const double *vec, int pauli)
{
    int tile, b;
    msp_xcd_tile(gridDim.x * gridDim.y, gridDim.z, tile, b);
    const int bx = tile % gridDim.x, by = tile / gridDim.x;
    if (UPPER && by > bx) return;
    int mstride = 1;
    bool mix_ok = true;
    cdbl m_same[2], m_other[2];
    if constexpr (MIX) {
        const MspSystem &S = systems[sys_id ? sys_id[b] : 0];
        mix_ok = S.dims[S.muon_index] == 2;   // otherwise the block writes NaN
        if (mix_ok) {
            for (int k = S.n_spins - 1; k > S.muon_index; --k) mstride *= S.dims[k];
        } else {
            mstride = n;                      // (kc / n) & 1 == 0; the partner row is kc itself
        }
        const double f = pauli ? 1.0 : 0.5;        // v.sigma or v.S
        const double vx = vec[3 * b], vy = vec[3 * b + 1], vz = vec[3 * b + 2];
        m_same[0] = cmk(f * vz, 0.0);              // M[0][0]
        m_same[1] = cmk(-f * vz, 0.0);             // M[1][1]
        m_other[0] = cmk(f * vx, -f * vy);         // M[0][1]
        m_other[1] = cmk(f * vx, f * vy);          // M[1][0]
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lq = lane >> 4;
    const int i0 = by * 64 + 32 * (wave >> 1);
    const int j0 = bx * 64 + 32 * (wave & 1);
    const cdbl *Ab = (const cdbl *)A + (size_t)b * n * n;
    const cdbl *Bb = (const cdbl *)B + (size_t)b * n * n;
    msp_d4 cr[2][2], ci[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            cr[x][y] = d4zero();
            ci[x][y] = d4zero();
        }
    const int ia = i0 + li, ib = i0 + 16 + li, ja = j0 + li, jb = j0 + 16 + li;
    const int iac = (ia < n) ? ia : n - 1, ibc = (ib < n) ? ib : n - 1;
    const int jac = (ja < n) ? ja : n - 1, jbc = (jb < n) ? jb : n - 1;
    for (int k0 = 0; k0 < n; k0 += 8) {
        cdbl av[2][2], bv[2][2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = k0 + 4 * h + lq;
            const int kc = (k < n) ? k : n - 1;
            av[h][0] = Ab[(size_t)kc * n + iac];
            av[h][1] = Ab[(size_t)kc * n + ibc];
            if constexpr (MIX) {
                const int a = (kc / mstride) & 1;
                const int k2 = mix_ok ? kc + (a ? -mstride : mstride) : kc;
                const cdbl x0 = Bb[(size_t)kc * n + jac], x1 = Bb[(size_t)kc * n + jbc];
                const cdbl y0 = Bb[(size_t)k2 * n + jac], y1 = Bb[(size_t)k2 * n + jbc];
                bv[h][0] = cadd(cmul(m_same[a], x0), cmul(m_other[a], y0));
                bv[h][1] = cadd(cmul(m_same[a], x1), cmul(m_other[a], y1));
            } else {
                bv[h][0] = Bb[(size_t)kc * n + jac];
                bv[h][1] = Bb[(size_t)kc * n + jbc];
            }
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const bool kok = k0 + 4 * h + lq < n;
            double ar[2], ai[2], br[2], bi[2];
            ar[0] = (kok && ia < n) ? av[h][0].re : 0.0;
            ai[0] = (kok && ia < n) ? -av[h][0].im : 0.0;
            ar[1] = (kok && ib < n) ? av[h][1].re : 0.0;
            ai[1] = (kok && ib < n) ? -av[h][1].im : 0.0;
            br[0] = (kok && ja < n) ? bv[h][0].re : 0.0;
            bi[0] = (kok && ja < n) ? bv[h][0].im : 0.0;
            br[1] = (kok && jb < n) ? bv[h][1].re : 0.0;
            bi[1] = (kok && jb < n) ? bv[h][1].im : 0.0;
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) mfma_cmac(ar[x], ai[x], br[y], bi[y], cr[x][y], ci[x][y]);
        }
    }
    cdbl *Cb = (cdbl *)C + (size_t)b * n * n;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = i0 + 16 * x + 4 * q + lq, j = j0 + 16 * y + li;
                if (i < n && j < n)
                    Cb[(size_t)i * n + j] = mix_ok ? cmk(cr[x][y][q], ci[x][y][q])
                                                   : cmk(__builtin_nan(""), __builtin_nan(""));
            }
}

}  // namespace

hipError_t launch_gemm_cn_mfma(int n, int batch, const double *A, const double *B, double *C,
                               hipStream_t s, bool upper)
{
    const unsigned t = (unsigned)((n + 63) / 64);
    if (upper)
        hipLaunchKernelGGL((k_gemm_cn_mfma<true, false>), dim3(t, t, batch), dim3(256), 0, s, n, A,
                           B, C, nullptr, nullptr, nullptr, 0);
    else
        hipLaunchKernelGGL((k_gemm_cn_mfma<false, false>), dim3(t, t, batch), dim3(256), 0, s, n,
                           A, B, C, nullptr, nullptr, nullptr, 0);
    return hipGetLastError();
}

hipError_t launch_gemm_muon_upper(int n, int batch, const MspSystem *systems,
                                  const int32_t *sys_id, const double *vec, int pauli,
                                  const double *U, double *O, hipStream_t s)
{
    ProfScope ps_("k_gemm_cn", s);
    const unsigned t = (unsigned)((n + 63) / 64);
    hipLaunchKernelGGL((k_gemm_cn_mfma<true, true>), dim3(t, t, batch), dim3(256), 0, s, n, U, U,
                       O, systems, sys_id, vec, pauli);
    return hipGetLastError();
}
