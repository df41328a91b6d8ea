// U = Q Z on FP64 matrix cores: the blocked back-transformation of
// msp_tridiag.h (U <- (I - V T V^H) U for reflector blocks of MSP_BT_NB = 16,
// last block first) with both block products on v_mfma_f64_16x16x4f64.
//
// One workgroup (NW waves: 4 up to n = 512, 8 up to n = 1024) owns a
// 16-column chunk of U for the whole back-transformation, held in registers
// as MFMA accumulators: wave w keeps the 16 x 16 row tiles w, w + NW, ...
// (real and imaginary planes).
// The accumulator of a tile is also the B operand of W = V^H U, because the
// accumulator slot r of lane l holds U[4 r + l/16][l%16] -- the rows of the
// k-step r of that tile.  Per reflector block:
//   W  = V^H U      16 x 16, partial per wave over its tiles, into LDS
//   U -= (V T) W    in the accumulators of every tile below the block; each
//                   wave sums the W partials itself, V T comes from the WY
//                   factor kernel
// One barrier per block (the W partials alternate between two LDS buffers;
// with 8 waves a second one, after the partials are summed once per element).
// U never leaves registers, so LDS holds only the W partials, and several
// workgroups share a CU.
#include <hip/hip_runtime.h>

#include "msp_kernels.h"
#include "msp_mfma.h"
#include "msp_tridiag.h"

namespace {

constexpr int CW = 16;           // columns per workgroup
static_assert(MSP_BT_NB == 16, "the MFMA back-transformation tiles reflector blocks of 16");

// Reflector j of block k0 at row r (unit at r = k0 + j + off, stored tail
// below, zero above), read from a clamped address so that no branch
// separates the loads of a k-step.
__device__ inline cdbl vload(int n, const cdbl *A, int k0, int nbk, int r, int j, int off)
{
    const int k = k0 + j;
    const bool tail = (r > k + off) && (r < n) && (j < nbk);
    const int rc = tail ? r : n - 1, kc = (k < n) ? k : n - 1;
    const cdbl a = A[(size_t)rc * n + kc];
    if (tail) return a;
    return cmk((r == k + off && j < nbk) ? 1.0 : 0.0, 0.0);
}

// WY factors of reflector block blk (reflectors k0 = 16 blk .. k0 + 15) on
// FP64 MFMA, operands read from global memory (8 KB of LDS, so that the
// kernel shares CUs with the divide and conquer it runs beside):
//   G = V^H V: each wave sums the tiles t = wave, wave + 4, ... of rows
//       (B operand V(16 t + 4 kk + q, l), A operand its conjugate); the wave
//       partials are added into LDS one wave after the other (fixed order);
//   T = (diag(1/tau) + striu(G))^{-1} scaled by tau row-wise, one column per
//       thread by back substitution (as the one-stage WY kernel);
//   V T per tile on MFMA, written column by column for the back-transformation.
template <int NW>
__global__ void __launch_bounds__(64 * NW) k_larft_mfma(int n, const double *H, double *work)
{
    __shared__ cdbl G[MSP_BT_NB * MSP_BT_NB];
    __shared__ cdbl T[MSP_BT_NB * MSP_BT_NB];
    int blk, b;
    msp_xcd_tile(gridDim.x, gridDim.y, blk, b);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lq = lane >> 4;
    const cdbl *A = (const cdbl *)H + (size_t)b * n * n;
    double *base = work + (size_t)b * dc_work_doubles(n);
    const cdbl *tau = dc_tau_ptr(n, base);
    const int k0 = blk * MSP_BT_NB;
    const int nref = n - 1;
    const int nbk = (nref - k0 < MSP_BT_NB) ? nref - k0 : MSP_BT_NB;
    const int ntiles = (n + 15) / 16;
    const int t0 = (k0 + 1) / 16;          // first tile with rows >= k0 + 1
    // ---- Gram partial of this wave
    msp_d4 gr = d4zero(), gi = d4zero();
    for (int t = t0 + wave; t < ntiles; t += NW) {
        cdbl v[4];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) v[kk] = vload(n, A, k0, nbk, 16 * t + 4 * kk + lq, li, 1);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) mfma_cmac(v[kk].re, -v[kk].im, v[kk].re, v[kk].im, gr, gi);
    }
    for (int w2 = 0; w2 < NW; ++w2) {
        if (wave == w2) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int o = (4 * r + lq) * 16 + li;
                cdbl g = cmk(gr[r], gi[r]);
                if (w2 > 0) g = cadd(g, G[o]);
                G[o] = g;
            }
        }
        __syncthreads();
    }
    // ---- T by back substitution (lanes 0..15 of wave 0)
    if (tid < MSP_BT_NB) {
        const int j = tid;
        cdbl tc[MSP_BT_NB];
#pragma unroll
        for (int i = MSP_BT_NB - 1; i >= 0; --i) {
            cdbl t = cmk(0.0, 0.0);
            if (j < nbk && i <= j) {
                if (i == j) {
                    t = tau[k0 + j];
                } else {
                    cdbl acc = cmk(0.0, 0.0);
#pragma unroll
                    for (int k = i + 1; k < MSP_BT_NB; ++k) cfma(acc, G[i * MSP_BT_NB + k], tc[k]);
                    const cdbl ti = tau[k0 + i];
                    t = cmul(cmk(-ti.re, -ti.im), acc);
                }
            }
            tc[i] = t;
        }
#pragma unroll
        for (int i = 0; i < MSP_BT_NB; ++i) T[i * MSP_BT_NB + j] = tc[i];
    }
    __syncthreads();
    cdbl *Tout = dc_tblocks(n, base) + (size_t)blk * MSP_BT_NB * MSP_BT_NB;
    if (tid < MSP_BT_NB * MSP_BT_NB) Tout[tid] = T[tid];
    // ---- V T per tile: D = V_t T (A operand V(16 t + l, 4 kk + q), B operand
    // T(4 kk + q, l)); rows above k0 + 1 are zero in V and are written as zeros
    cdbl *VT = dc_vt(n, base) + (size_t)blk * MSP_BT_NB * n;
    for (int t = wave; t < ntiles; t += NW) {
        msp_d4 dr = d4zero(), di = d4zero();
        if (t >= t0) {
            cdbl v[4], tt[4];
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                v[kk] = vload(n, A, k0, nbk, 16 * t + li, 4 * kk + lq, 1);
                tt[kk] = T[(4 * kk + lq) * MSP_BT_NB + li];
            }
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) mfma_cmac(v[kk].re, v[kk].im, tt[kk].re, tt[kk].im, dr, di);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * t + 4 * r + lq;
            if (row < n) VT[(size_t)li * n + row] = cmk(dr[r], di[r]);
        }
    }
}

// Workgroups per CU the register budget is set for: four up to three tiles per
// wave; two for the 4-wave, 8-tile form (n <= 512: 247 VGPRs, no spills,
// 11.8 -> 6.9 ms per 256 matrices against one); one otherwise (the 4-tile
// form measured 4 % slower at two)
template <int RT, int NW>
__global__ void __launch_bounds__(64 * NW, (NW == 4 && RT <= 3) ? 4 : (NW == 4 && RT > 4 ? 2 : 1)) k_backtransform_reg(int n, const double *H, double *work,
                                                               double *U, int off, int nref,
                                                               int zc)
{
    __shared__ double Wp[2][NW][2][256];      // per-wave partial W (re, im), by block parity
    __shared__ double Ws[(NW > 4) ? 512 : 1];  // NW > 4: W summed over the waves (re, im)
    int chunk, b;
    msp_xcd_tile(gridDim.x, gridDim.y, chunk, b);
    const int c0 = chunk * CW;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lq = lane >> 4;
    const int ntiles = (n + 15) / 16;

    const cdbl *A = (const cdbl *)H + (size_t)b * n * n;
    double *base = work + (size_t)b * dc_work_doubles(n);
    const double *Z = dc_result_ptr(n, base);
    const cdbl *VTall = dc_vt(n, base);

    const cdbl *Zcb = zc ? dc_zc_ptr(n, base) : nullptr;
    msp_d4 ur[RT], ui[RT];
#pragma unroll
    for (int i = 0; i < RT; ++i) {
        const int t = wave + NW * i;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * t + 4 * r + lq, col = c0 + li;
            const bool in = t < ntiles && row < n && col < n;
            if (Zcb) {
                const cdbl z = in ? Zcb[(size_t)col * n + row] : cmk(0.0, 0.0);
                ur[i][r] = z.re;
                ui[i][r] = z.im;
            } else {
                ur[i][r] = in ? Z[(size_t)col * n + row] : 0.0;
                ui[i][r] = 0.0;
            }
        }
    }

    const int nblk = (nref + MSP_BT_NB - 1) / MSP_BT_NB;
    for (int blk = nblk - 1; blk >= 0; --blk) {
        const int k0 = blk * MSP_BT_NB;
        const int nbk = (nref - k0 < MSP_BT_NB) ? nref - k0 : MSP_BT_NB;
        // an opaque zero added to the tile rows keeps the compiler from
        // hoisting every tile's row offsets out of the block loop (64-bit
        // addresses for all tiles, live across it: the 8-wave form spilled
        // 380 bytes per lane and ran n = 1024 at 41.0 ms per 128 matrices, 29.1
        // without; the 4-wave form fits 128 registers, four workgroups per CU)
        int zo = 0;
        asm volatile("" : "+s"(zo));
        // ---- W = V^H U over this wave's tiles (rows >= k0 + off only matter)
        msp_d4 wr = d4zero(), wi = d4zero();
#pragma unroll
        for (int i = 0; i < RT; ++i) {
            const int t = wave + NW * i;
            if (t < ntiles && 16 * t + 15 >= k0 + off) {
                cdbl v[4];
#pragma unroll
                for (int s = 0; s < 4; ++s) v[s] = vload(n, A, k0, nbk, 16 * t + 4 * s + lq + zo, li, off);
#pragma unroll
                for (int s = 0; s < 4; ++s)
                    mfma_cmac(v[s].re, -v[s].im, ur[i][s], ui[i][s], wr, wi);   // conj(V)
            }
            // with 8 waves (2 per SIMD, 256 registers) keep one tile's loads
            // in flight at a time
            if constexpr (NW > 4) __builtin_amdgcn_sched_barrier(0);
        }
        const int par = blk & 1;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            Wp[par][wave][0][(4 * q + lq) * 16 + li] = wr[q];
            Wp[par][wave][1][(4 * q + lq) * 16 + li] = wi[q];
        }
        __syncthreads();
        // ---- U -= (V T) W in the accumulators of the tiles below the block
        double br[4], bi[4];
        if constexpr (NW > 4) {
            // one thread per element of W sums the wave partials (in wave
            // order) and every lane then reads its four: summing in every
            // lane would hold 8 NW loads in flight, which the 8-wave form
            // has no registers for (n = 1024: 54.9 -> 45.1 ms per 140)
            if (tid < 512) {
                const int part = tid >> 8, o = tid & 255;
                double sum = 0.0;
                for (int w2 = 0; w2 < NW; ++w2) sum += Wp[par][w2][part][o];
                Ws[tid] = sum;
            }
            __syncthreads();
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int o = (4 * s + lq) * 16 + li;
                br[s] = Ws[o];
                bi[s] = Ws[256 + o];
            }
        } else {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int o = (4 * s + lq) * 16 + li;
                double sr = 0.0, si = 0.0;
#pragma unroll
                for (int w2 = 0; w2 < NW; ++w2) {
                    sr += Wp[par][w2][0][o];
                    si += Wp[par][w2][1][o];
                }
                br[s] = sr;
                bi[s] = si;
            }
        }
        const cdbl *VT = VTall + (size_t)blk * MSP_BT_NB * n;
#pragma unroll
        for (int i = 0; i < RT; ++i) {
            const int t = wave + NW * i;
            if (t < ntiles && 16 * t + 15 >= k0 + off) {
                const int r = 16 * t + li + zo;
                const int rc = (r < n) ? r : n - 1;
                cdbl v[4];
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const cdbl x = VT[(size_t)(4 * s + lq) * n + rc];
                    v[s] = (r < n) ? x : cmk(0.0, 0.0);
                }
#pragma unroll
                for (int s = 0; s < 4; ++s)
                    mfma_cmac(-v[s].re, -v[s].im, br[s], bi[s], ur[i], ui[i]);
            }
            if constexpr (NW > 4) __builtin_amdgcn_sched_barrier(0);
        }
    }
    cdbl *Ub = (cdbl *)U + (size_t)b * n * n;
#pragma unroll
    for (int i = 0; i < RT; ++i) {
        const int t = wave + NW * i;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * t + 4 * r + lq, col = c0 + li;
            if (t < ntiles && row < n && col < n)
                Ub[(size_t)row * n + col] = cmk(ur[i][r], ui[i][r]);
        }
    }
}

template <int RT, int NW>
hipError_t bt_launch(int n, int batch, const double *H, double *work, double *U, hipStream_t s,
                     int off, int nref, int zc)
{
    hipLaunchKernelGGL((k_backtransform_reg<RT, NW>), dim3((n + CW - 1) / CW, batch), dim3(64 * NW),
                       0, s, n, H, work, U, off, nref, zc);
    return hipGetLastError();
}

}  // namespace

// WY factors and V T of every reflector block of the one-stage reduction
hipError_t launch_larft_mfma(int n, int batch, const double *H, double *work, hipStream_t s)
{
    hipLaunchKernelGGL(k_larft_mfma<4>, dim3(bt_nblocks(n), batch), dim3(256), 0, s, n, H, work);
    return hipGetLastError();
}

// n <= 1024 (16-row tiles, at most 8 per wave; 8 waves above n = 512)
hipError_t launch_backtransform_mfma(int n, int batch, const double *H, double *work, double *U,
                                     hipStream_t s, int off, int nref, int zc)
{
    if (n > 1024) return hipErrorNotSupported;
    if (nref < 0) nref = n - off;
    if (n > 512) return bt_launch<8, 8>(n, batch, H, work, U, s, off, nref, zc);
    const int per_wave = ((n + 15) / 16 + 3) / 4;
    if (per_wave <= 1) return bt_launch<1, 4>(n, batch, H, work, U, s, off, nref, zc);
    if (per_wave <= 2) return bt_launch<2, 4>(n, batch, H, work, U, s, off, nref, zc);
    if (per_wave <= 3) return bt_launch<3, 4>(n, batch, H, work, U, s, off, nref, zc);
    if (per_wave <= 4) return bt_launch<4, 4>(n, batch, H, work, U, s, off, nref, zc);
    return bt_launch<8, 4>(n, batch, H, work, U, s, off, nref, zc);
}
