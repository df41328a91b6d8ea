// Householder tridiagonalisation with the matrix resident on the CU (gfx950,
// one workgroup of 8 waves per matrix).  The template covers n <= 192; the
// library instantiates it for n <= 64, where the whole lower triangle sits in
// VGPRs and it beats the panel kernel (msp_tridiag_lp.hip) by 7-25 %.  Above
// n = 64 the 64-column lane slots leave about half of the lanes idle (columns
// <= k and the upper part of the diagonal block) and the panel kernel is
// faster.
//
// Same reduction and outputs as k_tridiag_lp (A = Q T Q^H, reflector
// H_k = I - tau_k v v^H with v_{k+1} = 1, tails below the subdiagonal of the
// row-major matrix, d to E, e and tau to the eigensolver workspace), organised
// so that the trailing matrix never leaves the CU:
//   - columns 0..63 of the lower triangle live in VGPRs: lane l holds column
//     l, slot a of wave w holds row w + 8a (24 slots at n = 192);
//   - columns 64.. live in LDS as a packed lower triangle (132 KB at n = 192);
//   - vectors of length n (v, the pending update pair, the current column,
//     p) are lane-distributed, entry i in lane i % 64, slot i / 64, and every
//     wave computes them redundantly from the same inputs in the same order,
//     so reflector, p^H v and w need no exchange between waves.
// Step k makes ONE pass over the trailing triangle that applies the previous
// step's rank-2 update (A22 -= v' w'^H + w' v'^H) and accumulates A22 v_k:
// row sums are reduced across the wave four rows at a time by a butterfly,
// column sums stay in the lane and go to LDS once per step, one row per wave;
// column k+1, brought up to date on the way, is stashed for the next step.
// Three barriers per column.
#include <hip/hip_runtime.h>

#include "msp_kernels.h"
#include "msp_tridiag.h"

namespace {

constexpr int kW = 8;            // waves per matrix
constexpr int kT = 64 * kW;

__device__ inline double readlane_f64(double v, int l)
{
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(x & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(x >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Entry i (wave-uniform) of a lane-distributed vector, in every lane.  Every
// slot is read and the choice is made on the scalar side (selecting the slot
// first would index the array dynamically and move it to scratch).
template <int NS>
__device__ inline cdbl bcast(const cdbl (&x)[NS], int i)
{
    const int s = i >> 6, l = i & 63;
    cdbl v = cmk(readlane_f64(x[0].re, l), readlane_f64(x[0].im, l));
#pragma unroll
    for (int t = 1; t < NS; ++t) {
        const cdbl q = cmk(readlane_f64(x[t].re, l), readlane_f64(x[t].im, l));
        if (s == t) v = q;
    }
    return v;
}

// Lane l's entry of one slot, in every lane (l wave-uniform).
__device__ inline cdbl bcast1(cdbl x, int l) { return cmk(readlane_f64(x.re, l), readlane_f64(x.im, l)); }

// a - vp_i conj(wp_c) - wp_i conj(vp_c), as FMA chains
__device__ inline cdbl rank2(cdbl a, cdbl vpi, cdbl wpi, cdbl vpc, cdbl wpc)
{
    a.re = fma(-vpi.re, wpc.re, a.re);
    a.re = fma(-vpi.im, wpc.im, a.re);
    a.re = fma(-wpi.re, vpc.re, a.re);
    a.re = fma(-wpi.im, vpc.im, a.re);
    a.im = fma(-vpi.im, wpc.re, a.im);
    a.im = fma(vpi.re, wpc.im, a.im);
    a.im = fma(-wpi.im, vpc.re, a.im);
    a.im = fma(wpi.re, vpc.im, a.im);
    return a;
}

// Cross-lane swaps of gfx950 on the two halves of a double:
//   swap32: x <- [x lanes 0-31, y lanes 0-31], y <- [x lanes 32-63, y lanes 32-63]
//   swap16: x <- [x row 0, y row 0, x row 2, y row 2], y <- [x row 1, y row 1, x row 3, y row 3]
__device__ inline void swap32(double &x, double &y)
{
    const unsigned long long a = (unsigned long long)__double_as_longlong(x);
    const unsigned long long b = (unsigned long long)__double_as_longlong(y);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)a, (unsigned)b, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(a >> 32), (unsigned)(b >> 32), false, false);
    x = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
    y = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}
__device__ inline void swap16(double &x, double &y)
{
    const unsigned long long a = (unsigned long long)__double_as_longlong(x);
    const unsigned long long b = (unsigned long long)__double_as_longlong(y);
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)a, (unsigned)b, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(a >> 32), (unsigned)(b >> 32), false, false);
    x = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
    y = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}

// Sums over the 64 lanes of eight doubles per lane (four complex row sums):
// a butterfly that halves the values held per lane at strides 32 and 16
// (lane swaps, then one add: lanes of the lower half keep the lower values),
// then at the row mirror (partner 15 - l), then sums the 8 lanes of each half
// row.  On return every lane holds the total of value (lane >> 3) & 7.
__device__ inline double butterfly8(double (&x)[8], int lane)
{
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        swap32(x[j], x[j + 4]);
        x[j] += x[j + 4];
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        swap16(x[j], x[j + 2]);
        x[j] += x[j + 2];
    }
    const bool hi = lane & 8;
    const double send = hi ? x[0] : x[1];
    double t = (hi ? x[1] : x[0]) + dpp_f64<0x140, 0xF>(send);   // row_mirror
    t += dpp_f64<0xB1, 0xF>(t);    // quad_perm [1,0,3,2]
    t += dpp_f64<0x4E, 0xF>(t);    // quad_perm [2,3,0,1]
    t += dpp_f64<0x141, 0xF>(t);   // row_half_mirror
    return t;
}

// Workgroup barrier ordering LDS only: the reflector tails this kernel
// stores to global memory are read by later kernels, not by this one, so the
// barrier need not wait for them.
__device__ inline void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ inline int tri_off(int i) { return ((i - 64) * (i - 63)) >> 1; }   // row i >= 64 of the LDS part

// NA: register slots per lane (rows per wave), NS: 64-column slots (n <= 64 NS).
template <int NA, int NS>
__global__ void __launch_bounds__(kT) k_tridiag_rs(int n, double *H, double *E, double *work)
{
    static_assert(NA == 8 * NS, "one register slot per row of the wave");
    extern __shared__ __align__(16) unsigned char smem[];
    const int b = blockIdx.x;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    cdbl *A = (cdbl *)H + (size_t)b * n * n;
    double *base = work + (size_t)b * dc_work_doubles(n);
    double *e;
    cdbl *tau;
    dc_carve(n, base, &e, &tau);
    double *d = E + (size_t)b * n;
    const int nl = (n > 64) ? n - 64 : 0;
    cdbl *L = (cdbl *)smem;                              // packed rows 64.., columns 64..
    cdbl *cp = L + ((nl * (nl + 1)) >> 1);               // kW x n column partials
    double *rs = (double *)(cp + (size_t)kW * n);        // 2n: row sums (re, im)
    cdbl *cst = (cdbl *)(rs + 2 * n);                    // n: column k+1 for the next step

    cdbl R[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const int i = wave + kW * a;
        R[a] = (i < n && lane <= i) ? A[(size_t)i * n + lane] : cmk(0.0, 0.0);
    }
    for (int i = 64 + wave; i < n; i += kW)
        for (int c = 64 + lane; c <= i; c += 64) L[tri_off(i) + c - 64] = A[(size_t)i * n + c];
    cdbl vp[NS], wp[NS], vc[NS], colv[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int i = 64 * s + lane;
        vp[s] = cmk(0.0, 0.0);
        wp[s] = cmk(0.0, 0.0);
        vc[s] = cmk(0.0, 0.0);
        colv[s] = (i < n) ? A[(size_t)i * n] : cmk(0.0, 0.0);   // column 0
    }
    __syncthreads();

    for (int k = 0; k < n - 1; ++k) {
        // z is 0, opaque to the compiler: row indices wave + 8a + z are then
        // recomputed each step instead of hoisted out of the loop as 24 x 3
        // row predicates and LDS addresses, which would not fit in registers
        int z = 0;
        asm volatile("" : "+s"(z));
        cdbl tk;
        // (1) column k with the pending update (vp, wp), its reflector and
        // v_k, in every wave
        {
            const cdbl vpk = bcast(vp, k), wpk = bcast(wp, k);
            double part = 0.0;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const int i = 64 * s + lane;
                const cdbl x = rank2(colv[s], vp[s], wp[s], vpk, wpk);
                colv[s] = (i >= k && i < n) ? x : cmk(0.0, 0.0);
                if (i >= k + 2 && i < n) part += cabs2(x);
            }
            const double xn2 = wave_sum_all(part);
            const cdbl alpha = bcast(colv, k + 1);
            const double dk = bcast(colv, k).re;
            double beta;
            cdbl scale;
            if (xn2 == 0.0 && alpha.im == 0.0) {
                tk = cmk(0.0, 0.0);
                beta = alpha.re;
                scale = cmk(0.0, 0.0);
            } else {
                const double nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn2);
                beta = alpha.re >= 0.0 ? -nrm : nrm;
                // Newton-refined reciprocals (serial in every column step):
                // tau = (beta - alpha)/beta, 1/(alpha - beta) = (1/beta)/d with
                // d = alpha/beta - 1, 1 <= |d| <= 2 (beta has the sign opposite
                // to Re alpha)
                const double rb = msp_recip(beta);
                tk = cmk((beta - alpha.re) * rb, -alpha.im * rb);
                const double dr = alpha.re * rb - 1.0, di = alpha.im * rb;
                const double rd = rb * msp_recip(dr * dr + di * di);
                scale = cmk(dr * rd, -di * rd);
            }
            const bool act = (tk.re != 0.0 || tk.im != 0.0);
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const int i = 64 * s + lane;
                cdbl v = cmk(0.0, 0.0);
                if (act && i > k && i < n) v = (i == k + 1) ? cmk(1.0, 0.0) : cmul(colv[s], scale);
                vc[s] = v;
                if (wave == 0 && i >= k + 2 && i < n) A[(size_t)i * n + k] = v;
            }
            if (wave == 0 && lane == 0) {
                d[k] = dk;
                e[k] = beta;
                tau[k] = tk;
            }
        }

        // (2) one pass over rows k+1.. of this wave: pending update applied,
        // A22 v_k accumulated, column k+1 stashed.  Rows go four at a time
        // (rows i0 + 8j of group g, j < 4; their entries of the lane-
        // distributed vectors lie in slot g / 2, their elements in column
        // slots 0 .. g / 2) as straight-line code, so that the four rows'
        // loads and FMA chains overlap.  Only the last column slot (the
        // diagonal block) holds columns c > i, which are masked.  No other
        // mask is needed: v_k, vp and wp vanish at indices < k and at
        // indices >= n (vp_k = 1 only touches the finished row and column
        // k), so finished rows and columns and the padding rows add nothing
        // to live sums; the stores they make only touch finished storage.
        {
            const int sk = (k + 1) >> 6, lk = (k + 1) & 63;   // slot and lane of column k+1
            cdbl cacc[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s) cacc[s] = cmk(0.0, 0.0);
#pragma unroll
            for (int g = 0; g < NA / 4; ++g) {
                const int sd = g >> 1;                       // (wave < 8)
                const int i0 = wave + kW * 4 * g + z;
                if (i0 + 3 * kW <= k || i0 >= n) continue;   // no live row in the group
                cdbl vpi[4], wpi[4], vci[4], rr[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int li = (i0 + kW * j) & 63;
                    vpi[j] = bcast1(vp[sd], li);
                    wpi[j] = bcast1(wp[sd], li);
                    vci[j] = bcast1(vc[sd], li);
                    rr[j] = cmk(0.0, 0.0);
                }
#pragma unroll
                for (int s = 0; s <= sd; ++s) {
                    if (64 * s + 63 <= k) continue;          // no live column in the slot
                    const int c = 64 * s + lane;
                    const bool diag = (s == sd);
                    cdbl x[4];
                    int idx[4];
                    bool ok[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int i = i0 + kW * j;
                        ok[j] = !diag || c <= i;
                        if (s == 0) {
                            x[j] = R[4 * g + j];
                        } else {
                            idx[j] = (i < n) ? tri_off(i) + (ok[j] ? c : 64 * s) - 64 : 0;
                            x[j] = L[idx[j]];
                        }
                    }
#pragma unroll
                    for (int j = 0; j < 4; ++j) x[j] = rank2(x[j], vpi[j], wpi[j], vp[s], wp[s]);
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int i = i0 + kW * j;
                        if (s == 0) {
                            if (ok[j]) R[4 * g + j] = x[j];
                        } else {
                            if (ok[j] && i < n) L[idx[j]] = x[j];
                        }
                        if (s == sk && lane == lk && i > k && i < n) cst[i] = x[j];
                        // in the diagonal block: elements c > i drop out and
                        // a_ii (real) goes half to the row sum, half to the
                        // column sum of column i, which is the same entry p_i
                        if (diag) {
                            const double m = (c < i) ? 1.0 : ((c == i) ? 0.5 : 0.0);
                            x[j] = cscale(x[j], m);
                        }
                        cfma(rr[j], x[j], vc[s]);
                        cfmac(cacc[s], x[j], vci[j]);
                    }
                }
                double rb[8];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    rb[2 * j] = rr[j].re;
                    rb[2 * j + 1] = rr[j].im;
                }
                const double t = butterfly8(rb, lane);
                const int v = (lane >> 3) & 7;
                const int row = i0 + kW * (v >> 1);
                if ((lane & 7) == 0 && row > k && row < n) rs[2 * row + (v & 1)] = t;
            }
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const int c = 64 * s + lane;
                if (c > k && c < n) cp[wave * n + c] = cacc[s];
            }
        }
        lds_barrier();

        // (3) p = tau (A22 v): each entry summed once (row sum and the
        // waves' column partials), in place of its row sum
        {
            const int i = k + 1 + (int)threadIdx.x;
            if (i < n) {
                cdbl q = cmk(rs[2 * i], rs[2 * i + 1]);
#pragma unroll
                for (int w = 0; w < kW; ++w) q = cadd(q, cp[w * n + i]);
                q = cmul(tk, q);
                rs[2 * i] = q.re;
                rs[2 * i + 1] = q.im;
            }
        }
        lds_barrier();
        // p^H v and w = p - (tau/2)(p^H v) v in every wave; column k+1 comes
        // back from the stash
        {
            const cdbl *pv = (const cdbl *)rs;
            cdbl p[NS];
            double sre = 0.0, sim = 0.0;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const int i = 64 * s + lane;
                const bool ok = (i > k) && (i < n);
                const int ic = ok ? i : n - 1;
                p[s] = ok ? pv[ic] : cmk(0.0, 0.0);
                colv[s] = ok ? cst[ic] : cmk(0.0, 0.0);
                const cdbl t = cmulc(p[s], vc[s]);
                sre += t.re;
                sim += t.im;
            }
            sre = wave_sum_all(sre);
            sim = wave_sum_all(sim);
            const cdbl a2 = cmul(cmk(-0.5 * tk.re, -0.5 * tk.im), cmk(sre, sim));
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                wp[s] = cadd(p[s], cmul(a2, vc[s]));
                vp[s] = vc[s];
            }
        }
        lds_barrier();
    }
    // d[n-1]: the last diagonal entry with the last pending update
    {
        const cdbl x = rank2(bcast(colv, n - 1), bcast(vp, n - 1), bcast(wp, n - 1), bcast(vp, n - 1),
                             bcast(wp, n - 1));
        if (wave == 0 && lane == 0) d[n - 1] = x.re;
    }
}

template <int NA, int NS>
hipError_t rs_launch(int n, int batch, double *H, double *E, double *work, hipStream_t s)
{
    const size_t lds = tridiag_rs_lds_bytes(n);
    hipError_t err = msp_set_lds((const void *)k_tridiag_rs<NA, NS>, lds);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL((k_tridiag_rs<NA, NS>), dim3(batch), dim3(kT), lds, s, n, H, E, work);
    return hipGetLastError();
}

}  // namespace

size_t tridiag_rs_lds_bytes(int n)
{
    const size_t nl = (n > 64) ? (size_t)n - 64 : 0;
    return sizeof(cdbl) * (nl * (nl + 1) / 2 + (size_t)kW * n + 2 * (size_t)n);
}

// Register-resident tridiagonalisation for 2 <= n <= 64 (hipErrorNotSupported
// otherwise).
hipError_t launch_tridiag_rs(int n, int batch, double *H, double *E, double *work, hipStream_t s)
{
    if (n < 2 || n > 64) return hipErrorNotSupported;
    return rs_launch<8, 1>(n, batch, H, E, work, s);
}
