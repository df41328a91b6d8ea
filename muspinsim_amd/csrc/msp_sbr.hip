// Two-stage Hermitian tridiagonalisation (gfx950): dense -> band -> tridiagonal.
//
// The one-stage reduction reads the whole trailing triangle once per column
// (A22 v); at n = 192 that is 8 n^3 / 3 bytes per matrix and HBM-bound.  Here
// the first stage reduces to a lower band of width SB = 16 with panel QR and
// two-sided block updates on FP64 matrix cores, touching the trailing matrix
// three times per 16 columns; the second stage chases the band down to a real
// tridiagonal inside LDS.  Eigenvectors are U = Q1 Q2 Z: Q2 (the chase
// reflectors) is applied to the divide-and-conquer eigenvectors Z by
// k_chase_apply, Q1 (the panel reflectors, stored like the one-stage ones but
// with their unit element SB rows below the column instead of one) by the MFMA
// back-transformation, which takes the V T blocks this stage writes.
//
// Conventions are those of the one-stage reduction (msp_tridiag.h): reflector
// H = I - tau v v^H with v[0] = 1, beta = -sign(Re alpha) ||x||,
// tau = (beta - alpha) / beta, and A <- H^H A H.
#include <hip/hip_runtime.h>

#include "msp_kernels.h"
#include "msp_mfma.h"
#include "msp_tridiag.h"

namespace {

constexpr int SB = 16;          // band width of the first stage
constexpr int LDP = SB + 1;     // LDS row stride (complex) of the panel buffers
constexpr int BD = 2 * SB;      // stored band diagonals in the chase (offsets 0 .. 2 SB - 1)
constexpr int kChaseLag = 3;    // steps sweep j - 1 must be ahead of sweep j

// sum over the four lanes l, l ^ 16, l ^ 32, l ^ 48 (same column of a 16-lane row)
__device__ inline double xsum_rows(double v)
{
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}

// sum over the 16 lanes of each DPP row
__device__ inline double row16_sum(double v)
{
    v += dpp_f64<0xB1, 0xF>(v);    // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E, 0xF>(v);    // quad_perm [2,3,0,1]
    v += dpp_f64<0x141, 0xF>(v);   // row_half_mirror
    v += dpp_f64<0x140, 0xF>(v);   // row_mirror
    return v;
}

// LDS writes of this wave visible to its own later LDS reads (other lanes)
__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Householder scalars of x = (alpha, tail) with ||tail||^2 = xn2.
__device__ inline void reflector(cdbl alpha, double xn2, double &beta, cdbl &tau, cdbl &scale)
{
    if (xn2 == 0.0 && alpha.im == 0.0) {
        tau = cmk(0.0, 0.0);
        beta = alpha.re;
        scale = cmk(0.0, 0.0);
    } else {
        const double nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn2);
        beta = alpha.re >= 0.0 ? -nrm : nrm;
        tau = cmk((beta - alpha.re) / beta, -alpha.im / beta);
        scale = cdivc(cmk(1.0, 0.0), csub(alpha, cmk(beta, 0.0)));
    }
}

// V(t, s) of the current panel (unit lower trapezoidal, tails in P)
__device__ inline cdbl vpanel(const cdbl *P, int m, int t, int s)
{
    if (t < 0 || t >= m || t < s) return cmk(0.0, 0.0);
    if (t == s) return cmk(1.0, 0.0);
    return P[t * LDP + s];
}

// Sum of one double per thread over the workgroup; every thread gets it.
template <int NW>
__device__ inline double block_sum(double v, double *red)
{
    v = wave_sum_all(v);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[wave] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) t += red[w];
    __syncthreads();
    return t;
}

// ---------------------------------------------------------------------------
// Stage 1: dense -> lower band of width SB, one workgroup of NW waves per
// matrix.  Per panel of SB columns c0 .. c0 + 15 (rows r0 = c0 + 16 .. n - 1):
//   QR of the panel in LDS (reflector jj: unit at row r0 + jj, tail below,
//   stored in column c0 + jj as the one-stage reflectors are);
//   T (WY factor) and W = V T, written to the workspace for the
//   back-transformation;
//   X = A22 W on MFMA (A22 Hermitian, lower triangle read, upper blocks as
//   conjugate transposes), M = T^H V^H X, W' = X - V M / 2;
//   A22 -= W' V^H + V W'^H on the lower 16 x 16 blocks, on MFMA.
// LDS: the panel (V), the W / W' buffer (m x 17 complex each), the wave
// partials of V^H X, T and M.
template <int NW>
__global__ void __launch_bounds__(64 * NW) k_sbr_band(int n, double *H, double *work)
{
    constexpr int NT = 64 * NW;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ cdbl Ts[SB * SB];
    __shared__ cdbl Ms[SB * SB];
    __shared__ cdbl tauS[SB];
    __shared__ double red[2 * NW * SB + 8];
    const int b = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lq = lane >> 4;
    cdbl *A = (cdbl *)H + (size_t)b * n * n;
    double *base = work + (size_t)b * dc_work_doubles(n);
    cdbl *tauG = dc_tau_ptr(n, base);
    cdbl *VTg = dc_vt(n, base);
    const int mmax = n - SB;
    cdbl *P = (cdbl *)smem;                       // mmax x LDP
    cdbl *Wb = P + (size_t)mmax * LDP;            // mmax x LDP
    cdbl *G2 = Wb + (size_t)mmax * LDP;           // NW x 256 partials of V^H X

    for (int c0 = 0; c0 + SB + 1 < n; c0 += SB) {
        const int r0 = c0 + SB, m = n - r0, p = c0 / SB;
        const int TR = (m + 15) / 16;
        // ---- panel into LDS
        for (int idx = tid; idx < m * SB; idx += NT) {
            const int t = idx >> 4, s = idx & 15;
            P[t * LDP + s] = A[(size_t)(r0 + t) * n + c0 + s];
        }
        __syncthreads();
        // ---- panel QR: column jj reflects rows jj .. m - 1 of the panel
        const int cg = tid & 15, rp = tid >> 4;
        for (int jj = 0; jj < SB; ++jj) {
            if (m - jj < 2) {
                if (tid == 0) tauS[jj] = cmk(0.0, 0.0);
                continue;
            }
            double part = 0.0;
            for (int t = jj + 1 + tid; t < m; t += NT) part += cabs2(P[t * LDP + jj]);
            const double xn2 = block_sum<NW>(part, red);
            const cdbl alpha = P[jj * LDP + jj];
            double beta;
            cdbl tk, scale;
            reflector(alpha, xn2, beta, tk, scale);
            // w_s = sum_t conj(v_t) P[t][s] over the columns s = cg > jj (v from the
            // unscaled column, scaled on the fly)
            cdbl w = cmk(0.0, 0.0);
            if (cg > jj) {
                for (int t = jj + rp; t < m; t += NT / 16) {
                    const cdbl vt = (t == jj) ? cmk(1.0, 0.0) : cmul(P[t * LDP + jj], scale);
                    cfmac(w, vt, P[t * LDP + cg]);
                }
            }
            w.re = xsum_rows(w.re);
            w.im = xsum_rows(w.im);
            if (lq == 0) {
                red[wave * SB + cg] = w.re;
                red[NW * SB + wave * SB + cg] = w.im;
            }
            __syncthreads();
            cdbl ws = cmk(0.0, 0.0);
#pragma unroll
            for (int w2 = 0; w2 < NW; ++w2) {
                ws.re += red[w2 * SB + cg];
                ws.im += red[NW * SB + w2 * SB + cg];
            }
            // P[t][s] -= conj(tau) v_t w_s; column jj becomes (beta, v tail)
            const cdbl f = cmul(cconj(tk), ws);
            if (cg > jj) {
                for (int t = jj + rp; t < m; t += NT / 16) {
                    const cdbl vt = (t == jj) ? cmk(1.0, 0.0) : cmul(P[t * LDP + jj], scale);
                    const cdbl d = cmul(vt, f);
                    P[t * LDP + cg] = csub(P[t * LDP + cg], d);
                }
            }
            __syncthreads();
            for (int t = jj + tid; t < m; t += NT)
                P[t * LDP + jj] = (t == jj) ? cmk(beta, 0.0) : cmul(P[t * LDP + jj], scale);
            if (tid == 0) tauS[jj] = tk;
            __syncthreads();
        }
        // ---- panel back to A (band block R and reflector tails), tau out
        for (int idx = tid; idx < m * SB; idx += NT) {
            const int t = idx >> 4, s = idx & 15;
            A[(size_t)(r0 + t) * n + c0 + s] = P[t * LDP + s];
        }
        if (tid < SB) tauG[c0 + tid] = tauS[tid];
        // ---- T: Gram G = V^H V on MFMA (wave partials over its tile rows; the
        // A and B operands of a k-step are the same element of V), then the
        // upper triangular T = (diag(1/tau) + striu(G))^{-1} scaled by tau
        // row-wise, one column per thread by back substitution
        constexpr int RT = (256 / 16 + NW - 1) / NW;   // tile rows per wave, n <= 272
        {
            msp_d4 gr = d4zero(), gi = d4zero();
#pragma unroll
            for (int u = 0; u < RT; ++u) {
                const int i = wave + NW * u;
                if (i >= TR) continue;
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const cdbl v = vpanel(P, m, 16 * i + 4 * kk + lq, li);
                    mfma_cmac(v.re, -v.im, v.re, v.im, gr, gi);
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) G2[wave * 256 + (4 * r + lq) * 16 + li] = cmk(gr[r], gi[r]);
        }
        __syncthreads();
        if (tid < SB * SB) {
            cdbl g = cmk(0.0, 0.0);
#pragma unroll
            for (int w2 = 0; w2 < NW; ++w2) g = cadd(g, G2[w2 * 256 + tid]);
            Ms[tid] = g;
        }
        __syncthreads();
        if (tid < SB) {
            const int j = tid;
            cdbl tc[SB];
#pragma unroll
            for (int i = SB - 1; i >= 0; --i) {
                cdbl t = cmk(0.0, 0.0);
                if (i <= j) {
                    if (i == j) {
                        t = tauS[j];
                    } else {
                        cdbl acc = cmk(0.0, 0.0);
#pragma unroll
                        for (int k = i + 1; k < SB; ++k) cfma(acc, Ms[i * SB + k], tc[k]);
                        const cdbl ti = tauS[i];
                        t = cmul(cmk(-ti.re, -ti.im), acc);
                    }
                }
                tc[i] = t;
            }
#pragma unroll
            for (int i = 0; i < SB; ++i) Ts[i * SB + j] = tc[i];
        }
        __syncthreads();
        // ---- W = V T on MFMA, into LDS and the workspace (block p, rows r0 ..)
        {
            cdbl *VT = VTg + (size_t)p * SB * n;
#pragma unroll
            for (int u = 0; u < RT; ++u) {
                const int i = wave + NW * u;
                if (i >= TR) continue;
                msp_d4 wr = d4zero(), wi = d4zero();
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const cdbl v = vpanel(P, m, 16 * i + li, 4 * kk + lq);
                    const cdbl t = Ts[(4 * kk + lq) * SB + li];
                    mfma_cmac(v.re, v.im, t.re, t.im, wr, wi);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int t = 16 * i + 4 * r + lq;
                    if (t < m) {
                        const cdbl x = cmk(wr[r], wi[r]);
                        Wb[t * LDP + li] = x;
                        VT[(size_t)li * n + r0 + t] = x;
                    }
                }
            }
        }
        __syncthreads();
        // ---- X = A22 W: wave `wave` owns tile rows i = wave, wave + NW, ...
        msp_d4 xr[RT], xi[RT];
#pragma unroll
        for (int u = 0; u < RT; ++u) {
            xr[u] = d4zero();
            xi[u] = d4zero();
            const int i = wave + NW * u;
            if (i >= TR) continue;
            const int ra = r0 + 16 * i + li;      // row of the A operand
            for (int j = 0; j < TR; ++j) {
                cdbl av[4];
                const int cb = r0 + 16 * j + 4 * lq;   // k index c = 4 lq + kk
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const int c = cb + kk;
                    const bool in = (ra < n) && (c < n);
                    const int rr = in ? ra : r0, cc = in ? c : r0;
                    // lower element (ra, c) or the conjugate of (c, ra)
                    const bool low = (j < i) || (j == i && li >= 4 * lq + kk);
                    const cdbl a = low ? A[(size_t)rr * n + cc] : cconj(A[(size_t)cc * n + rr]);
                    av[kk] = in ? a : cmk(0.0, 0.0);
                }
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const int t = 16 * j + 4 * lq + kk;
                    const cdbl bw = (t < m) ? Wb[t * LDP + li] : cmk(0.0, 0.0);
                    mfma_cmac(av[kk].re, av[kk].im, bw.re, bw.im, xr[u], xi[u]);
                }
            }
        }
        // ---- V^H X partial of this wave (k = 4 kk + lq: slot kk of X is the B operand)
        {
            msp_d4 gr = d4zero(), gi = d4zero();
#pragma unroll
            for (int u = 0; u < RT; ++u) {
                const int i = wave + NW * u;
                if (i >= TR) continue;
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const cdbl v = vpanel(P, m, 16 * i + 4 * kk + lq, li);
                    mfma_cmac(v.re, -v.im, xr[u][kk], xi[u][kk], gr, gi);
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) G2[wave * 256 + (4 * r + lq) * 16 + li] = cmk(gr[r], gi[r]);
        }
        __syncthreads();
        // M = T^H (V^H X)
        if (tid < SB * SB) {
            const int a = tid >> 4, c = tid & 15;
            cdbl acc = cmk(0.0, 0.0);
            for (int k = 0; k <= a; ++k) {
                cdbl g = cmk(0.0, 0.0);
#pragma unroll
                for (int w2 = 0; w2 < NW; ++w2) g = cadd(g, G2[w2 * 256 + k * 16 + c]);
                cfmac(acc, Ts[k * SB + a], g);
            }
            Ms[tid] = acc;
        }
        __syncthreads();
        // W' = X - V M / 2 into the W buffer (W is no longer read)
#pragma unroll
        for (int u = 0; u < RT; ++u) {
            const int i = wave + NW * u;
            if (i >= TR) continue;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const cdbl v = vpanel(P, m, 16 * i + li, 4 * kk + lq);
                const cdbl mm = Ms[(4 * kk + lq) * SB + li];
                mfma_cmac(-0.5 * v.re, -0.5 * v.im, mm.re, mm.im, xr[u], xi[u]);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int t = 16 * i + 4 * r + lq;
                if (t < m) Wb[t * LDP + li] = cmk(xr[u][r], xi[u][r]);
            }
        }
        __syncthreads();
        // ---- A22 -= W' V^H + V W'^H on the lower blocks (i >= j)
        const int nblk = TR * (TR + 1) / 2;
        for (int q = wave; q < nblk; q += NW) {
            int i = 0, rem = q;
            while (rem > i) {
                rem -= i + 1;
                ++i;
            }
            const int j = rem;
            msp_d4 dr, di;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = r0 + 16 * i + 4 * r + lq, col = r0 + 16 * j + li;
                const bool in = row < n && col < n;
                const cdbl a = A[(size_t)(in ? row : r0) * n + (in ? col : r0)];
                dr[r] = in ? a.re : 0.0;
                di[r] = in ? a.im : 0.0;
            }
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const int k = 4 * kk + lq;
                const int ti = 16 * i + li, tj = 16 * j + li;
                const cdbl wi_ = (ti < m) ? Wb[ti * LDP + k] : cmk(0.0, 0.0);
                const cdbl vj = vpanel(P, m, tj, k);
                const cdbl vi = vpanel(P, m, ti, k);
                const cdbl wj = (tj < m) ? Wb[tj * LDP + k] : cmk(0.0, 0.0);
                // D -= W'_i conj(V_j)^T and D -= V_i conj(W'_j)^T
                mfma_cmac(-wi_.re, -wi_.im, vj.re, -vj.im, dr, di);
                mfma_cmac(-vi.re, -vi.im, wj.re, -wj.im, dr, di);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = r0 + 16 * i + 4 * r + lq, col = r0 + 16 * j + li;
                if (row < n && col < n && (i > j || row >= col))
                    A[(size_t)row * n + col] = cmk(dr[r], di[r]);
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Stage 2: band (width SB, lower, complex) -> real symmetric tridiagonal by
// bulge chasing, the band in LDS (AB[i][o] = A(i, i - o), o < 2 SB).  Sweep j
// reflects column j's entries j + 2 .. j + SB onto row j + 1, then chases the
// bulge down: step s >= 1 reflects rows r = j + 1 + SB s .. r + SB - 1 to
// clear column r - SB below row r.  Each step applies its reflector to the
// left block (rows of the step, columns col + 1 .. r - 1), the diagonal
// block (two-sided) and the block below (rows r + L .. r + L + SB - 1); the
// three are disjoint.  One wave runs one sweep at a time (wave w takes sweeps
// w, w + NW, ...); step s of sweep j waits until sweep j - 1 has finished
// step s + kChaseLag - 1 (the regions of the two are then disjoint; checked
// against the serial order on random interleavings).  The first reflector of
// every sweep makes e_j real (length-1 reflectors included).
// Chase reflectors go to the workspace in (sweep, step) order: SB complex
// entries of v (zero past the length) and tau per step.
template <int NW>
__global__ void __launch_bounds__(64 * NW) k_band_chase(int n, const double *H, double *E,
                                                        double *work, double *chase)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int prog[1024];
    __shared__ cdbl vq[NW][2][SB];
    const int b = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lq = lane >> 4;
    const cdbl *A = (const cdbl *)H + (size_t)b * n * n;
    double *base = work + (size_t)b * dc_work_doubles(n);
    const int S2 = chase_steps(n);
    cdbl *CV = (cdbl *)(chase + (size_t)b * chase_doubles(n));
    cdbl *AB = (cdbl *)smem;
    for (int idx = tid; idx < n * BD; idx += blockDim.x) {
        const int i = idx / BD, o = idx - (idx / BD) * BD;
        AB[idx] = (o <= SB && i - o >= 0) ? A[(size_t)i * n + i - o] : cmk(0.0, 0.0);
    }
    for (int j = tid; j < n; j += blockDim.x) prog[j] = 0;
    __syncthreads();

    for (int j = wave; j < n - 1; j += NW) {
        int r = j + 1, col = j;
        for (int s = 0;; ++s) {
            if (s > 0) {
                col = r;
                r += SB;
                if (r + 1 >= n) break;
            }
            const int L = (n - r < SB) ? n - r : SB;
            if (j > 0) {
                // wait for sweep j - 1 (a wave that never waits on this one)
                while (true) {
                    const int pr = __hip_atomic_load(&prog[j - 1], __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (pr >= s + kChaseLag) break;
                    __builtin_amdgcn_s_sleep(1);
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
            }
            // ---- reflector from x_t = A(r + t, col)
            const int ox = r - col;
            const cdbl x = (li < L) ? AB[(r + li) * BD + ox + li] : cmk(0.0, 0.0);
            const double xn2 = row16_sum(li >= 1 ? cabs2(x) : 0.0);
            const cdbl alpha = cmk(__shfl(x.re, lane & 48), __shfl(x.im, lane & 48));
            double beta;
            cdbl tau, scale;
            reflector(alpha, xn2, beta, tau, scale);
            const cdbl v = (li == 0) ? cmk(1.0, 0.0) : (li < L ? cmul(x, scale) : cmk(0.0, 0.0));
            if (lq == 0) {
                if (li < L) AB[(r + li) * BD + ox + li] = (li == 0) ? cmk(beta, 0.0) : cmk(0.0, 0.0);
                vq[wave][0][li] = v;
                cdbl *dst = CV + ((size_t)j * S2 + s) * (SB + 1);
                dst[li] = v;
                if (li == 0) dst[SB] = tau;
            }
            wave_lds_sync();
            cdbl vs[4];
#pragma unroll
            for (int sg = 0; sg < 4; ++sg) vs[sg] = vq[wave][0][lq + 4 * sg];
            // ---- left block (chase steps): columns k = col + 1 + li, li < SB - 1
            if (s > 0) {
                const bool lact = li < SB - 1;
                const int k = col + 1 + li;
                cdbl w = cmk(0.0, 0.0);
                cdbl av[4];
#pragma unroll
                for (int sg = 0; sg < 4; ++sg) {
                    const int t = lq + 4 * sg;
                    av[sg] = (lact && t < L) ? AB[(r + t) * BD + r + t - k] : cmk(0.0, 0.0);
                    cfmac(w, vs[sg], av[sg]);
                }
                w.re = xsum_rows(w.re);
                w.im = xsum_rows(w.im);
                const cdbl f = cmul(cconj(tau), w);
#pragma unroll
                for (int sg = 0; sg < 4; ++sg) {
                    const int t = lq + 4 * sg;
                    if (lact && t < L) AB[(r + t) * BD + r + t - k] = csub(av[sg], cmul(vs[sg], f));
                }
            }
            // ---- diagonal block: M(a, c) with c = li, a = lq + 4 sg
            {
                cdbl mv[4];
                cdbl pa[4];
                const cdbl vc = v;   // v_c, c = li
#pragma unroll
                for (int sg = 0; sg < 4; ++sg) {
                    const int a = lq + 4 * sg;
                    const bool in = a < L && li < L;
                    cdbl m_ = cmk(0.0, 0.0);
                    if (in) m_ = (a >= li) ? AB[(r + a) * BD + a - li] : cconj(AB[(r + li) * BD + li - a]);
                    mv[sg] = m_;
                    cdbl pr = cmul(m_, vc);
                    pa[sg] = cmk(row16_sum(pr.re), row16_sum(pr.im));
                }
                // mu = v^H p (one row holds p_a for a = lq + 4 sg)
                cdbl mu = cmk(0.0, 0.0);
#pragma unroll
                for (int sg = 0; sg < 4; ++sg) cfmac(mu, vs[sg], pa[sg]);
                const double mur = xsum_rows(mu.re);
                // q_a = tau p_a - |tau|^2 mu v_a / 2
                const double h = 0.5 * cabs2(tau) * mur;
                cdbl qa[4];
#pragma unroll
                for (int sg = 0; sg < 4; ++sg)
                    qa[sg] = csub(cmul(tau, pa[sg]), cscale(vs[sg], h));
                if (li < 4) {
                    // lane li = sg of row lq publishes q_{lq + 4 sg}
                    cdbl qs = qa[0];
#pragma unroll
                    for (int sg = 1; sg < 4; ++sg)
                        if (li == sg) qs = qa[sg];
                    vq[wave][1][lq + 4 * li] = qs;
                }
                wave_lds_sync();
                const cdbl qc = vq[wave][1][li];
#pragma unroll
                for (int sg = 0; sg < 4; ++sg) {
                    const int a = lq + 4 * sg;
                    if (a < L && li <= a) {
                        // M(a, c) - q_a conj(v_c) - v_a conj(q_c)
                        cdbl d = mv[sg];
                        d = csub(d, cmulc(vc, qa[sg]));
                        d = csub(d, cmulc(qc, vs[sg]));
                        AB[(r + a) * BD + a - li] = d;
                    }
                }
            }
            // ---- block below: rows i = r + L + u (u = lq + 4 sg), columns r + t (t = li)
            {
                cdbl bv[4];
#pragma unroll
                for (int sg = 0; sg < 4; ++sg) {
                    const int i = r + L + lq + 4 * sg;
                    const bool in = i < n && li < L;
                    bv[sg] = in ? AB[i * BD + i - r - li] : cmk(0.0, 0.0);
                    cdbl pr = cmul(bv[sg], v);
                    const cdbl y = cmk(row16_sum(pr.re), row16_sum(pr.im));
                    const cdbl f = cmul(tau, y);
                    if (in) AB[i * BD + i - r - li] = csub(bv[sg], cmulc(v, f));
                }
            }
            // publish the step: LDS writes before the counter (LDS-only order)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            if (lane == 0)
                __hip_atomic_store(&prog[j], s + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __builtin_amdgcn_wave_barrier();
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        if (lane == 0)
            __hip_atomic_store(&prog[j], 1 << 20, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
    // d and e (real after the sweeps)
    double *e = base;
    for (int i = tid; i < n; i += blockDim.x) {
        E[(size_t)b * n + i] = AB[i * BD].re;
        if (i + 1 < n) e[i] = AB[(i + 1) * BD + 1].re;
    }
}

// ---------------------------------------------------------------------------
// Z' = Q2 Z for a 16-column chunk of Z: the chase reflectors in reverse
// order, a sweep at a time (the steps of one sweep touch disjoint rows and
// commute; step s goes to wave s % NW).  Z (real, column-major as the divide
// and conquer leaves it) in, Z' complex column-major out over the S and QK
// arrays of the workspace (dc_zc_ptr), which the divide and conquer no longer
// needs.
template <int NW>
__global__ void __launch_bounds__(64 * NW) k_chase_apply(int n, double *work, const double *chase)
{
    extern __shared__ __align__(16) unsigned char smem[];
    int chunk, b;
    msp_xcd_tile(gridDim.x, gridDim.y, chunk, b);
    const int c0 = chunk * 16;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lq = lane >> 4;
    double *base = work + (size_t)b * dc_work_doubles(n);
    const double *Z = dc_result_ptr(n, base);
    const int S2 = chase_steps(n);
    const cdbl *CV = (const cdbl *)(chase + (size_t)b * chase_doubles(n));
    cdbl *Zc = (cdbl *)smem;   // n x 16
    for (int idx = tid; idx < n * 16; idx += blockDim.x) {
        const int cc = idx / n, row = idx - (idx / n) * n;
        const int col = c0 + cc;
        Zc[row * 16 + cc] = cmk(col < n ? Z[(size_t)col * n + row] : 0.0, 0.0);
    }
    __syncthreads();
    for (int j = n - 2; j >= 0; --j) {
        for (int s = wave;; s += NW) {
            const int r = j + 1 + SB * s;
            if (s > 0 && r + 1 >= n) break;
            if (r >= n) break;
            const int L = (n - r < SB) ? n - r : SB;
            const cdbl *src = CV + ((size_t)j * S2 + s) * (SB + 1);
            const cdbl tau = src[SB];
            cdbl vs[4], zv[4];
            cdbl w = cmk(0.0, 0.0);
#pragma unroll
            for (int sg = 0; sg < 4; ++sg) {
                const int t = lq + 4 * sg;
                vs[sg] = (t < L) ? src[t] : cmk(0.0, 0.0);
                zv[sg] = (t < L) ? Zc[(r + t) * 16 + li] : cmk(0.0, 0.0);
                cfmac(w, vs[sg], zv[sg]);
            }
            w.re = xsum_rows(w.re);
            w.im = xsum_rows(w.im);
            const cdbl f = cmul(tau, w);
#pragma unroll
            for (int sg = 0; sg < 4; ++sg) {
                const int t = lq + 4 * sg;
                if (t < L) Zc[(r + t) * 16 + li] = csub(zv[sg], cmul(vs[sg], f));
            }
        }
        __syncthreads();
    }
    cdbl *Zo = dc_zc_ptr(n, base);
    for (int idx = tid; idx < n * 16; idx += blockDim.x) {
        const int cc = idx / n, row = idx - (idx / n) * n;
        const int col = c0 + cc;
        if (col < n) Zo[(size_t)col * n + row] = Zc[row * 16 + cc];
    }
}

}  // namespace

size_t sbr_lds_bytes(int n) { return sizeof(cdbl) * ((size_t)2 * (n - SB) * LDP + 4 * 256); }
size_t chase_lds_bytes(int n) { return sizeof(cdbl) * (size_t)n * BD; }

// Two-stage reduction for 2 SB + 2 <= n <= kSbrMax: band, chase (d, e and the
// chase reflectors out); `chase` holds chase_doubles(n) per instance.
hipError_t launch_sbr(int n, int batch, double *H, double *E, double *work, double *chase,
                      hipStream_t s)
{
    if (n < kSbrMin || n > kSbrMax) return hipErrorNotSupported;
    const size_t l1 = sbr_lds_bytes(n), l2 = chase_lds_bytes(n);
    hipError_t err = msp_set_lds((const void *)k_sbr_band<4>, l1);
    if (err != hipSuccess) return err;
    {
        ProfScope ps_("k_sbr_band", s);
        hipLaunchKernelGGL(k_sbr_band<4>, dim3(batch), dim3(256), l1, s, n, H, work);
    }
    if ((err = hipGetLastError()) != hipSuccess) return err;
    if ((err = msp_set_lds((const void *)k_band_chase<4>, l2)) != hipSuccess) return err;
    {
        ProfScope ps_("k_band_chase", s);
        hipLaunchKernelGGL(k_band_chase<4>, dim3(batch), dim3(256), l2, s, n, (const double *)H, E,
                           work, chase);
    }
    return hipGetLastError();
}

// Z' = Q2 Z into each instance's dc_zc_ptr
hipError_t launch_chase_apply(int n, int batch, double *work, const double *chase, hipStream_t s)
{
    const size_t lds = sizeof(cdbl) * (size_t)n * 16;
    hipError_t err = msp_set_lds((const void *)k_chase_apply<4>, lds);
    if (err != hipSuccess) return err;
    ProfScope ps_("k_chase_apply", s);
    hipLaunchKernelGGL(k_chase_apply<4>, dim3((n + 15) / 16, batch), dim3(256), lds, s, n, work,
                       chase);
    return hipGetLastError();
}
