// Panel-blocked Householder tridiagonalisation that reads and writes only the
// lower triangle (gfx950, one workgroup of W = 4 or 8 waves per matrix).
//
// Same reduction as tridiag_team in msp_tridiag.h (A = Q T Q^H, reflector
// tails stored below the subdiagonal, d/e/tau out), organised for HBM:
//   - the trailing update is deferred to the end of each panel of NB columns
//     (zlatrd form): inside a panel, column k and the product A22 v use the
//     panel-start matrix corrected by the panel's earlier vectors, kept in LDS;
//   - A22 v is one pass over the lower triangle: each wave walks rows, lanes
//     own columns; a row's left part is reduced across the wave (row sum)
//     while conj(a_ic) v_i accumulates per lane (column sum), so every
//     element is read once per step;
//   - the end-of-panel update A -= Vp Wp^H + Wp Vp^H touches c <= i only,
//     with each lane's column of Vp and Wp held in registers.
#include <hip/hip_runtime.h>

#include "msp_kernels.h"
#include "msp_tridiag.h"

namespace {

constexpr int kCpRows = 4;      // LDS rows of per-wave column partials
constexpr int kPanelRows = 8;   // rows per batched pass of the end-of-panel update
#ifndef MSP_TD_RB
#define MSP_TD_RB 4     // rows per batched pass of the A22 v product (64-lane rows)
#endif
#ifndef MSP_TD_ROWS16
#define MSP_TD_ROWS16 1 // n <= 192: A22 v with one row per 16-lane group
#endif
#ifndef MSP_TD_RP
#define MSP_TD_RP 2     // rows per 16-lane group in one pass of that product
#endif
#ifndef MSP_TD_WL
#define MSP_TD_WL 8     // waves per workgroup above n = 192
#endif
#ifndef MSP_TD_RBMAX
#define MSP_TD_RBMAX 4  // most rows in one pass of the 64-lane-row A22 v product
#endif
#ifndef MSP_TD_SPL
#define MSP_TD_SPL 2    // staged panel entries per lane in the workspace-panel update (1: 2-3 % slower)
#endif
#ifndef MSP_TD_NBL
#define MSP_TD_NBL 8    // panel width above n = 512 (panels in the workspace; 4 was 5-13 % slower)
#endif
#ifndef MSP_TD_RB1024
#define MSP_TD_RB1024 1 // rows per pass of the A22 v product for n <= 1024 (2 spills at 8 waves)
#endif

// reductions across a wave run on DPP lane moves (msp_team.h)
__device__ inline double wave_sum(double v) { return wave_sum_all(v); }

// Sum over the 16 lanes of each DPP row; every lane of the row gets it.
__device__ inline double row16_sum(double v)
{
    v += dpp_f64<0xB1, 0xF>(v);    // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E, 0xF>(v);    // quad_perm [2,3,0,1]
    v += dpp_f64<0x141, 0xF>(v);   // row_half_mirror
    v += dpp_f64<0x140, 0xF>(v);   // row_mirror
    return v;
}

// One pass of the 16-lane-row A22 v product over rows ib + g + 4 W r
// (r < RP), with the first QA column slots (QA is a compile-time constant so
// that every load of the pass is issued before the first use).  Slots past a
// row's diagonal, and rows past n, load A[0][n-1] instead, an element of the
// strict upper triangle that the kernel zeroes at its start and never writes
// again, so no loaded value needs a mask.  Every column sum then also takes
// its diagonal term conj(a_ii) v_i, which the row sum gives back.
template <int W, int Q16, int QA, int RP>
__device__ inline void rows16_pass(int n, int k, int ib, int g, int li, const cdbl *A,
                                   const cdbl *vb, cdbl *pb, cdbl *nxt, cdbl (&colacc)[Q16])
{
    cdbl av[RP][QA];
    cdbl dg[RP];
#pragma unroll
    for (int r = 0; r < RP; ++r) {
        const int i = ib + g + 4 * W * r;
        const bool rowok = i < n;
        const int ro = i * n;
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const int c = k + 1 + li + 16 * q;
            av[r][q] = A[(rowok && c <= i) ? ro + c : n - 1];
        }
        dg[r] = A[rowok ? ro + i : n - 1];
    }
#pragma unroll
    for (int r = 0; r < RP; ++r) {
        const int i = ib + g + 4 * W * r;
        const bool rowok = i < n;
        const cdbl vi = vb[rowok ? i : n - 1];
        double rr = 0.0, ri = 0.0;
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const int c = k + 1 + li + 16 * q;
            const cdbl a = av[r][q];
            const cdbl vc = vb[(c < n) ? c : n - 1];
            rr += a.re * vc.re - a.im * vc.im;
            ri += a.re * vc.im + a.im * vc.re;
            cfmac(colacc[q], a, vi);
            if (q == 0 && li == 0 && rowok) nxt[i] = a;   // column k+1
        }
        rr = row16_sum(rr);
        ri = row16_sum(ri);
        if (li == 0 && rowok) {
            const cdbl d = dg[r];
            pb[i] = cmk(rr - (d.re * vi.re + d.im * vi.im), ri - (d.re * vi.im - d.im * vi.re));
        }
    }
}

// Column partials of one step, every thread of the workgroup calling: wave w
// stores its per-lane sums acc[q] (column c0 + CS q, lanes with `writer`) in
// LDS row w % 4 of cp; with W = 8 the waves w and w + 4 share a row, the
// second adding to the first after a barrier, so the order of the sum is
// fixed.  (The caller's barrier ends the last phase.)
template <int W, int Q, int CS>
__device__ inline void store_colpart(int n, int c0, int wave, bool writer, const cdbl (&acc)[Q],
                                     cdbl *cp)
{
    static_assert(W % kCpRows == 0, "W is a multiple of the partial rows");
    for (int ph = 0; ph < W / kCpRows; ++ph) {
        if (ph > 0) __syncthreads();
        if (wave / kCpRows == ph && writer) {
            cdbl *row = cp + (size_t)(wave % kCpRows) * n;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int c = c0 + CS * q;
                if (c < n) row[c] = (ph > 0) ? cadd(row[c], acc[q]) : acc[q];
            }
        }
    }
}

#ifndef MSP_TD_QSTEP
#define MSP_TD_QSTEP 2  // slot-count granularity of the 16-lane-row passes (1 and 4 measured slower)
#endif

// rows16_pass with the smallest QA = Q16 - MSP_TD_QSTEP m >= qact
// (qact wave-uniform), so that a pass loads no slot wholly past its rows.
template <int W, int Q16, int QA, int RP>
__device__ __forceinline__ void rows16_dispatch(int qact, int n, int k, int ib, int g, int li,
                                                const cdbl *A, const cdbl *vb, cdbl *pb,
                                                cdbl *nxt, cdbl (&colacc)[Q16])
{
    constexpr int S = MSP_TD_QSTEP;
    if constexpr (QA > S) {
        if (qact <= QA - S) {
            rows16_dispatch<W, Q16, QA - S, RP>(qact, n, k, ib, g, li, A, vb, pb, nxt, colacc);
            return;
        }
    }
    rows16_pass<W, Q16, QA, RP>(n, k, ib, g, li, A, vb, pb, nxt, colacc);
}

// A22 v for step k with one matrix row per 16-lane group: wave w takes rows
// k+1 + 4 W t + 4 w + g (g = lane / 16), a lane the columns k+1 + (lane % 16)
// + 16 q.  A row sum is a 16-lane DPP reduction shared by the wave's four
// rows; the column sums conj(a_ic) v_i stay in registers and are combined
// across the four groups once per step.  pb[i] receives the row sums, nxt[i]
// the column k+1 (next step's column), cp the waves' column sums.
template <int W, int Q16>
__device__ inline void matvec_rows16(int n, int k, int wave, int lane, const cdbl *A, const cdbl *vb,
                                     cdbl *pb, cdbl *nxt, cdbl *cp)
{
    static_assert(Q16 == 4 || Q16 == 8 || Q16 == 12, "16-lane rows cover n <= 192");
    const int g = lane >> 4, li = lane & 15;
    cdbl colacc[Q16];
#pragma unroll
    for (int q = 0; q < Q16; ++q) colacc[q] = cmk(0.0, 0.0);
    constexpr int RP = MSP_TD_RP;
    for (int ib = k + 1 + 4 * wave; ib < n; ib += 4 * W * RP) {
        const int ilast0 = ib + 4 * W * (RP - 1) + 3;
        const int ilast = (ilast0 < n) ? ilast0 : n - 1;
        const int qact = ((ilast - (k + 1)) >> 4) + 1;          // wave-uniform, <= Q16
        rows16_dispatch<W, Q16, Q16, RP>(qact, n, k, ib, g, li, A, vb, pb, nxt, colacc);
    }
    const int qstep = ((n - 1 - (k + 1)) >> 4) + 1;
#pragma unroll
    for (int q = 0; q < Q16; ++q) {
        if (q < qstep) {
            cdbl t = colacc[q];
            t.re += __shfl_xor(t.re, 16, 64);
            t.im += __shfl_xor(t.im, 16, 64);
            t.re += __shfl_xor(t.re, 32, 64);
            t.im += __shfl_xor(t.im, 32, 64);
            colacc[q] = t;
        }
    }
    store_colpart<W, Q16, 16>(n, k + 1 + li, wave, g == 0, colacc, cp);
}

// End-of-panel rank-2k update with lanes owning columns: for each 64-column
// block the lane's column of Vp and Wp stays in registers while the wave walks
// the rows at or below the block (rows wave, wave + W, ..., RB per pass, all
// loads of a pass issued first); a row's Vp and Wp entries are wave-uniform
// LDS reads.  A(i, c) -= sum_l Vp_il conj(Wp_cl) + Wp_il conj(Vp_cl), c <= i.
// (always_inline: as a call, the ABI's saved registers spill to scratch)
template <int W, int NB, int RB, bool GP>
__device__ __attribute__((always_inline)) inline void panel_update_cols(
    int n, int t0, int wave, int lane, cdbl *A, const cdbl *Vp, const cdbl *Wp, cdbl *col,
    cdbl *stage)
{
    constexpr int LDP = NB + 1;
    // With the panel pair in global memory (GP), a pass's 2 NB panel entries
    // of each of its RB rows are fetched by one load per lane, issued with
    // the pass's matrix loads, and read back from the wave's LDS block, so
    // no row waits on a global round trip of its own.
    static_assert(!GP || 2 * NB * RB == 64 * MSP_TD_SPL, "MSP_TD_SPL staged panel entries per lane");
    cdbl *stw = stage + 64 * MSP_TD_SPL * wave;
    for (int cb = t0; cb < n; cb += 64) {
        const int c = cb + lane;
        const bool cok = c < n;
        const int cc = cok ? c : n - 1;
        cdbl vc[NB], wc[NB];
#pragma unroll
        for (int l = 0; l < NB; ++l) {
            vc[l] = cconj(Vp[cc * LDP + l]);
            wc[l] = cconj(Wp[cc * LDP + l]);
        }
        for (int i0 = cb + wave; i0 < n; i0 += RB * W) {
            cdbl av[RB];
#pragma unroll
            for (int u = 0; u < RB; ++u) {
                const int i = i0 + u * W;
                const int ic = (i < n) ? i : n - 1;
                av[u] = A[(size_t)ic * n + ((cc <= ic) ? cc : ic)];
            }
            if constexpr (GP) {
                // entry e = 2 NB u + m: row i0 + u W, Vp column m (m < NB)
                // or Wp column m - NB
                cdbl pv[MSP_TD_SPL];
#pragma unroll
                for (int h = 0; h < MSP_TD_SPL; ++h) {
                    const int e = lane + 64 * h;
                    const int u = e / (2 * NB), m = e % (2 * NB);
                    const int i = i0 + u * W;
                    const int ic = (i < n) ? i : n - 1;
                    pv[h] = (m < NB) ? Vp[ic * LDP + m] : Wp[ic * LDP + m - NB];
                }
#pragma unroll
                for (int h = 0; h < MSP_TD_SPL; ++h) stw[lane + 64 * h] = pv[h];
                __builtin_amdgcn_wave_barrier();
            }
#pragma unroll
            for (int u = 0; u < RB; ++u) {
                const int i = i0 + u * W;
                if (i >= n) break;                       // wave-uniform
                cdbl a = av[u];
#pragma unroll
                for (int l = 0; l < NB; ++l) {
                    const cdbl vp = GP ? stw[2 * NB * u + l] : Vp[i * LDP + l];
                    const cdbl wp = GP ? stw[2 * NB * u + NB + l] : Wp[i * LDP + l];
                    a = csub(a, cmul(vp, wc[l]));
                    a = csub(a, cmul(wp, vc[l]));
                }
                if (cok && c <= i) {
                    A[(size_t)i * n + c] = a;
                    if (c == t0) col[i] = a;             // next step's column
                }
            }
            if constexpr (GP) __builtin_amdgcn_wave_barrier();
        }
    }
}

// One pass of the 64-lane-row A22 v product over the rows i0 + u W (u < RB)
// of step k, with the first QN column slots (columns k+1 + lane + 64 q,
// q < QN).  QN is a compile-time constant, so every load of the pass is
// issued, from clamped addresses with the value masked, before the first
// use; the caller picks the smallest QN that reaches the pass's longest row.
template <int W, int QMAX, int RB, int QN>
__device__ __forceinline__ void a22v_pass(int n, int k, int i0, int lane, const cdbl *A,
                                          const cdbl *vb, cdbl *pb, cdbl *nxt,
                                          cdbl (&colacc)[QMAX])
{
    cdbl av[RB][QN];
#pragma unroll
    for (int u = 0; u < RB; ++u) {
        const int i = i0 + u * W;
        const int ic = (i < n) ? i : n - 1;
        const cdbl *row = A + (size_t)ic * n;
#pragma unroll
        for (int q = 0; q < QN; ++q) {
            const int c = k + 1 + lane + 64 * q;
            const int cc = (c <= ic) ? c : ic;
            av[u][q] = row[cc];
        }
    }
    double rr[RB], ri[RB];
#pragma unroll
    for (int u = 0; u < RB; ++u) {
        const int i = i0 + u * W;
        const bool rowok = i < n;
        const cdbl vi = vb[rowok ? i : n - 1];
        rr[u] = 0.0;
        ri[u] = 0.0;
#pragma unroll
        for (int q = 0; q < QN; ++q) {
            const int c = k + 1 + lane + 64 * q;
            if (rowok && c <= i) {
                const cdbl a = av[u][q];
                if (q == 0 && lane == 0) nxt[i] = a;   // column k+1
                const cdbl vc = vb[c];
                rr[u] += a.re * vc.re - a.im * vc.im;
                ri[u] += a.re * vc.im + a.im * vc.re;
                if (c < i) cfmac(colacc[q], a, vi);
            }
        }
    }
#pragma unroll
    for (int u = 0; u < RB; ++u) {
        rr[u] = wave_sum_to63(rr[u]);
        ri[u] = wave_sum_to63(ri[u]);
    }
    if (lane == 63) {
#pragma unroll
        for (int u = 0; u < RB; ++u) {
            const int i = i0 + u * W;
            if (i < n) pb[i] = cmk(rr[u], ri[u]);
        }
    }
}

// a22v_pass with QN = qn, 1 <= qn <= Q (qn wave-uniform).
template <int W, int QMAX, int RB, int Q>
__device__ __forceinline__ void a22v_pass_q(int qn, int n, int k, int i0, int lane, const cdbl *A,
                                            const cdbl *vb, cdbl *pb, cdbl *nxt,
                                            cdbl (&colacc)[QMAX])
{
    if constexpr (Q > 1) {
        if (qn < Q) {
            a22v_pass_q<W, QMAX, RB, Q - 1>(qn, n, k, i0, lane, A, vb, pb, nxt, colacc);
            return;
        }
    }
    a22v_pass<W, QMAX, RB, Q>(n, k, i0, lane, A, vb, pb, nxt, colacc);
}

// One pass of the A22 v product at row i0, returning the rows it covered.
// Passes whose rows are short (few column slots) take 2 RB or 4 RB rows
// within the same slot budget SB = RB QMAX, so a wave keeps as many loads in
// flight near the diagonal as in the full-width passes.
template <int W, int QMAX, int RB>
__device__ __forceinline__ int a22v_rows(int n, int k, int i0, int lane, const cdbl *A,
                                         const cdbl *vb, cdbl *pb, cdbl *nxt,
                                         cdbl (&colacc)[QMAX])
{
    constexpr int SB = RB * QMAX;
    auto slots = [&](int rows) {
        const int il = (i0 + (rows - 1) * W < n) ? i0 + (rows - 1) * W : n - 1;
        return __builtin_amdgcn_readfirstlane(((il - k - 1) >> 6) + 1);
    };
    if constexpr (SB / (4 * RB) >= 1 && 4 * RB <= MSP_TD_RBMAX) {
        const int qn = slots(4 * RB);
        if (qn <= SB / (4 * RB)) {
            a22v_pass_q<W, QMAX, 4 * RB, SB / (4 * RB)>(qn, n, k, i0, lane, A, vb, pb, nxt, colacc);
            return 4 * RB * W;
        }
    }
    if constexpr (SB / (2 * RB) >= 1 && 2 * RB <= MSP_TD_RBMAX) {
        const int qn = slots(2 * RB);
        if (qn <= SB / (2 * RB)) {
            a22v_pass_q<W, QMAX, 2 * RB, SB / (2 * RB)>(qn, n, k, i0, lane, A, vb, pb, nxt, colacc);
            return 2 * RB * W;
        }
    }
    a22v_pass_q<W, QMAX, RB, QMAX>(slots(RB), n, k, i0, lane, A, vb, pb, nxt, colacc);
    return RB * W;
}

template <int W, int NB, int QMAX, int RB, int Q16, bool GP>
__global__ void __launch_bounds__(64 * W) k_tridiag_lp(int n, double *H, double *E, double *work)
{
    static_assert(W == 4 || W == 8, "4 or 8 waves per matrix");
    constexpr int kThreads = 64 * W;
    extern __shared__ __align__(16) unsigned char smem[];
    const int b = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int LDP = NB + 1;                     // padded panel row stride
    static_assert(!GP || LDP <= MSP_TD_NB, "global panels hold MSP_TD_NB columns");
    cdbl *A = (cdbl *)H + (size_t)b * n * n;
    double *base = work + (size_t)b * dc_work_doubles(n);
    double *e;
    cdbl *tau;
    dc_carve(n, base, &e, &tau);
    double *d = E + (size_t)b * n;
    // the panel pair lives in LDS, or (GP, large n) in the workspace
    cdbl *Vp = GP ? dc_panels(n, base) : (cdbl *)smem;                 // n x LDP
    cdbl *Wp = Vp + (size_t)n * LDP;                                     // n x LDP
    cdbl *vb = GP ? (cdbl *)smem : Wp + (size_t)n * LDP;                // n
    cdbl *pb = vb + n;                               // n (row sums, then p)
    cdbl *cp = pb + n;                               // kCpRows x n column partials
    cdbl *y = cp + (size_t)kCpRows * n;              // 2 NB
    cdbl *colA = y + 2 * NB;                         // n: current column k
    cdbl *colB = colA + n;                           // n: column k+1, stashed by the matvec
    double *red = (double *)(colB + n);              // 2 W
    cdbl *stage = (cdbl *)(red + 2 * W);             // GP: 64 MSP_TD_SPL per wave

    cdbl *col = colA, *nxt = colB;
    // A[0][n-1] (strict upper triangle, never written below) is the zero that
    // the 16-lane-row A22 v passes load in place of elements past a row's
    // diagonal; the first barrier below publishes it
    if (Q16 > 0 && tid == 0) A[n - 1] = cmk(0.0, 0.0);
    for (int i = tid; i < n; i += kThreads) col[i] = A[(size_t)i * n];
    for (int k0 = 0; k0 < n - 1; k0 += NB) {
        const int nbk = (n - 1 - k0 < NB) ? n - 1 - k0 : NB;
        // (rows above the panel are never read)
        for (int idx = k0 * LDP + tid; idx < n * LDP; idx += kThreads) {
            Vp[idx] = cmk(0.0, 0.0);
            Wp[idx] = cmk(0.0, 0.0);
        }
        __syncthreads();
        for (int j = 0; j < nbk; ++j) {
            const int k = k0 + j;
            const int m = n - k - 1;
            // (a) column k, rows k..n-1, brought up to date within the panel,
            // with the squared norm of rows k+2.. summed on the way
            double part = 0.0;
            if (j > 0) {
                for (int i = k + tid; i < n; i += kThreads) {
                    cdbl a = col[i];
                    for (int l = 0; l < j; ++l) {
                        a = csub(a, cmul(Vp[i * LDP + l], cconj(Wp[k * LDP + l])));
                        a = csub(a, cmul(Wp[i * LDP + l], cconj(Vp[k * LDP + l])));
                    }
                    col[i] = a;
                    if (i >= k + 2) part += cabs2(a);
                }
            } else {
                for (int i = k + 2 + tid; i < n; i += kThreads) part += cabs2(col[i]);
            }
            // (b) reflector of A[k+1.., k]
            part = wave_sum(part);
            if (lane == 0) red[wave] = part;
            __syncthreads();
            double xn2 = red[0];
            for (int w = 1; w < W; ++w) xn2 += red[w];
            const cdbl alpha = col[k + 1];
            double beta;
            cdbl tk, scale;
            if (xn2 == 0.0 && alpha.im == 0.0) {
                tk = cmk(0.0, 0.0);
                beta = alpha.re;
                scale = cmk(0.0, 0.0);
            } else {
                double nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn2);
                beta = alpha.re >= 0.0 ? -nrm : nrm;
                // two Newton-refined reciprocals instead of IEEE divisions
                // (this chain is serial in every column step):
                // tau = (beta - alpha)/beta, 1/(alpha - beta) = (1/beta)/d
                // with d = alpha/beta - 1, 1 <= |d| <= 2
                const double ib = msp_recip(beta);
                tk = cmk((beta - alpha.re) * ib, -alpha.im * ib);
                const cdbl dd = cmk(alpha.re * ib - 1.0, alpha.im * ib);
                const double id = ib * msp_recip(dd.re * dd.re + dd.im * dd.im);
                scale = cmk(dd.re * id, -dd.im * id);
            }
            const bool act = (tk.re != 0.0 || tk.im != 0.0);
            for (int i = k + 1 + tid; i < n; i += kThreads) {
                cdbl vi;
                if (i == k + 1) vi = cmk(1.0, 0.0);
                else vi = cmul(col[i], scale);
                vb[i] = vi;
                Vp[i * LDP + j] = act ? vi : cmk(0.0, 0.0);
            }
            __syncthreads();
            if (act) {
                // (c) y[l] = Wp(:, l)^H v, y[j + l] = Vp(:, l)^H v over rows k+1..
                for (int l = wave; l < 2 * j; l += W) {
                    const cdbl *M = (l < j) ? Wp : Vp;
                    const int c = (l < j) ? l : l - j;
                    double sr = 0.0, si = 0.0;
                    for (int i = k + 1 + lane; i < n; i += 64) {
                        cdbl t = cmulc(M[i * LDP + c], vb[i]);
                        sr += t.re;
                        si += t.im;
                    }
                    sr = wave_sum(sr);
                    si = wave_sum(si);
                    if (lane == 0) y[l] = cmk(sr, si);
                }
                // (d) A22 v over the lower triangle of rows/cols k+1..n-1
                if constexpr (Q16 > 0) {
                    matvec_rows16<W, Q16>(n, k, wave, lane, A, vb, pb, nxt, cp);
                } else {
                    cdbl colacc[QMAX];
    #pragma unroll
                    for (int q = 0; q < QMAX; ++q) colacc[q] = cmk(0.0, 0.0);
                    // RB rows per pass (i, i+W, ... of this wave): every load of
                    // the pass is issued first, from clamped addresses with the
                    // value masked, so no branch separates the loads
                    // (slots past a pass's longest row hold c > i for every
                    // lane and row: neither loaded nor used)
                    for (int i0 = k + 1 + wave; i0 < n;)
                        i0 += a22v_rows<W, QMAX, RB>(n, k, i0, lane, A, vb, pb, nxt, colacc);
                    store_colpart<W, QMAX, 64>(n, k + 1 + lane, wave, true, colacc, cp);
                }
                __syncthreads();
                // (e) p = tau (A22 v - Vp y1 - Wp y2), and p^H v
                double pre = 0.0, pim = 0.0;
                for (int i = k + 1 + tid; i < n; i += kThreads) {
                    cdbl acc = pb[i];
                    for (int w = 0; w < kCpRows; ++w) acc = cadd(acc, cp[w * n + i]);
                    for (int l = 0; l < j; ++l) {
                        acc = csub(acc, cmul(Vp[i * LDP + l], y[l]));
                        acc = csub(acc, cmul(Wp[i * LDP + l], y[j + l]));
                    }
                    acc = cmul(tk, acc);
                    pb[i] = acc;
                    cdbl t = cmulc(acc, vb[i]);
                    pre += t.re;
                    pim += t.im;
                }
                pre = wave_sum(pre);
                pim = wave_sum(pim);
                // (red was last read before the barrier that follows the
                // v write, so it can be overwritten here)
                if (lane == 0) {
                    red[wave] = pre;
                    red[W + wave] = pim;
                }
                __syncthreads();
                double sre = red[0], sim = red[W];
                for (int w = 1; w < W; ++w) {
                    sre += red[w];
                    sim += red[W + w];
                }
                const cdbl a2 = cmul(cmk(-0.5 * tk.re, -0.5 * tk.im), cmk(sre, sim));
                // (f) w = p + a2 v
                for (int i = k + 1 + tid; i < n; i += kThreads) Wp[i * LDP + j] = cadd(pb[i], cmul(a2, vb[i]));
            }
            if (!act)
                for (int i = k + 1 + tid; i < n; i += kThreads) nxt[i] = A[(size_t)i * n + k + 1];
            if (tid == 0) {
                d[k] = col[k].re;
                e[k] = beta;
                tau[k] = tk;
            }
            __syncthreads();
            {
                cdbl *t = col;
                col = nxt;
                nxt = t;
            }
            (void)m;
        }
        // end of panel: reflector tails of the panel's columns go to memory
        // (rows k+2.., below the subdiagonal), and the trailing lower triangle
        // takes A(i, c) -= sum_l Vp_il conj(Wp_cl) + Wp_il conj(Vp_cl)
        for (int idx = tid; idx < n * nbk; idx += kThreads) {
            const int i = idx / nbk, l = idx % nbk;
            if (i >= k0 + l + 2) A[(size_t)i * n + k0 + l] = Vp[i * LDP + l];
        }
        panel_update_cols<W, NB, GP ? 64 * MSP_TD_SPL / (2 * NB) : kPanelRows, GP>(n, k0 + nbk, wave, lane, A, Vp, Wp,
                                                                     col, stage);
        __syncthreads();
    }
    if (tid == 0) d[n - 1] = col[n - 1].re;
}

template <int W, int NB, bool GP>
size_t lp_lds(int n)
{
    return sizeof(cdbl) * ((GP ? 0 : (size_t)2 * n * (NB + 1)) + 4 * (size_t)n +
                           (size_t)kCpRows * n + 2 * NB + (GP ? 64 * MSP_TD_SPL * W : 0)) +
           16 * sizeof(double);
}

template <int W, int NB, int QMAX, int RB, int Q16, bool GP = false>
hipError_t lp_launch(int n, int batch, double *H, double *E, double *work, hipStream_t s)
{
    size_t lds = lp_lds<W, NB, GP>(n);
    hipError_t err = msp_set_lds((const void *)k_tridiag_lp<W, NB, QMAX, RB, Q16, GP>, lds);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL((k_tridiag_lp<W, NB, QMAX, RB, Q16, GP>), dim3(batch), dim3(64 * W), lds, s, n,
                       H, E, work);
    return hipGetLastError();
}

}  // namespace

// Lower-triangle panel kernel for n <= 1024 (panels in the workspace above
// n = 512); returns hipErrorNotSupported above.
hipError_t launch_tridiag_lp(int n, int batch, double *H, double *E, double *work, hipStream_t s)
{
    // n <= 64: the register-resident kernel (msp_tridiag_rs.hip) is faster
    if (n >= 2 && n <= 64) return launch_tridiag_rs(n, batch, H, E, work, s);
    if (n <= 64) return lp_launch<4, 8, 1, MSP_TD_RB, MSP_TD_ROWS16 ? 4 : 0>(n, batch, H, E, work, s);
    if (n <= 128) return lp_launch<4, 8, 2, MSP_TD_RB, MSP_TD_ROWS16 ? 8 : 0>(n, batch, H, E, work, s);
    if (n <= 192) return lp_launch<4, 8, 3, MSP_TD_RB, MSP_TD_ROWS16 ? 12 : 0>(n, batch, H, E, work, s);
    // above n = 192 one workgroup fills a CU's LDS: 8 waves put two on each SIMD
    if (n <= 256) return lp_launch<MSP_TD_WL, 8, 4, 4, 0>(n, batch, H, E, work, s);
    if (n <= 512) return lp_launch<MSP_TD_WL, 4, 8, 4, 0>(n, batch, H, E, work, s);
    if (n <= 1024) return lp_launch<MSP_TD_WL, MSP_TD_NBL, 16, MSP_TD_RB1024, 0, true>(n, batch, H, E, work, s);
    return hipErrorNotSupported;
}
