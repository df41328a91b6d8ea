// Mixed-precision eigensolver pieces (gfx950).
//
// k_tridiag32: Householder tridiagonalisation of a Hermitian matrix in FP32
// with the whole lower triangle resident in LDS (n <= 192: 148 KB packed),
// one workgroup of 256 threads per matrix.  The reduction is the one
// k_tridiag_lp performs in FP64 (A = Q T Q^H, reflector H_k = I - tau v v^H
// with v_{k+1} = 1, p = tau A22 v, w = p - (tau/2)(p^H v) v,
// A22 -= v w^H + w v^H), and it writes the same outputs, promoted to FP64:
// the reflector tails below the subdiagonal of a row-major n x n matrix,
// d (n), e (n-1) and tau (n-1) in the eigensolver workspace, so that the
// divide and conquer and the back-transformation run on them unchanged.
//
// Organisation for LDS instead of HBM: step k first brings column k up to
// date with the previous step's rank-2 update (only that column), forms the
// reflector, then makes ONE pass over the trailing triangle that both applies
// the previous step's update (A22 -= v' w'^H + w' v'^H) and accumulates
// A22 v for the new reflector.  Every element is read and written once per
// step, from LDS.  The A22 v product has lanes owning columns and 16-lane
// groups owning rows: a row sum is a 16-lane DPP reduction, the column sums
// conj(a_ic) v_i stay in registers and are combined across the four row
// groups of a wave by lane shuffles and across the four waves through LDS.
#include <hip/hip_runtime.h>

#include "msp_kernels.h"
#include "msp_tridiag.h"

namespace {

struct cf {
    float re, im;
};

__device__ inline cf cfk(float r, float i) { cf z; z.re = r; z.im = i; return z; }

template <int CTRL, int ROWMASK>
__device__ inline float dpp_f32(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWMASK, 0xF, false));
}

// Sum over the 16 lanes of each DPP row, every lane of the row gets it.
__device__ inline float row16_sumf(float v)
{
    v += dpp_f32<0xB1, 0xF>(v);    // quad_perm [1,0,3,2]
    v += dpp_f32<0x4E, 0xF>(v);    // quad_perm [2,3,0,1]
    v += dpp_f32<0x141, 0xF>(v);   // row_half_mirror
    v += dpp_f32<0x140, 0xF>(v);   // row_mirror
    return v;
}

// Wave sum broadcast to every lane.
__device__ inline float wave_sumf(float v)
{
    v = row16_sumf(v);
    v += dpp_f32<0x142, 0xA>(v);   // row_bcast15 -> rows 1, 3
    v += dpp_f32<0x143, 0xC>(v);   // row_bcast31 -> rows 2, 3
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ inline int tri(int i) { return (i * (i + 1)) >> 1; }

// One pass over rows ib + g + 16 r (r < RP) of the trailing triangle for
// step k: a(i, c) -= vp_i conj(wp_c) + wp_i conj(vp_c), written back; row
// sums of a(i, c) vc_c into pb[i]; column sums conj(a(i, c)) vc_i (c < i)
// into colacc.  QA: column slots live in this pass (compile-time, so that
// every load is issued before the first use).
template <int Q16, int QA, int RP>
__device__ __attribute__((always_inline)) inline void fused_rows(
    int n, int k, int ib, int g, int li, cf *T, const cf *vp, const cf *wp, const cf *vc,
    const cf (&vpc)[Q16], const cf (&wpc)[Q16], const cf (&vcc)[Q16], cf *pb, cf (&colacc)[Q16])
{
    cf av[RP][QA];
#pragma unroll
    for (int r = 0; r < RP; ++r) {
        const int i = ib + g + 16 * r;
        const int ic = (i < n) ? i : n - 1;
        const cf *row = T + tri(ic);
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const int c = k + 1 + li + 16 * q;
            av[r][q] = row[(c <= ic) ? c : ic];
        }
    }
#pragma unroll
    for (int r = 0; r < RP; ++r) {
        const int i = ib + g + 16 * r;
        const bool rowok = i < n;
        const int ic = rowok ? i : n - 1;
        cf *row = T + tri(ic);
        const cf vpi = vp[ic], wpi = wp[ic], vci = vc[ic];
        float rr = 0.f, ri = 0.f;
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const int c = k + 1 + li + 16 * q;
            const bool ok = rowok && c <= i;
            cf a = av[r][q];
            // a -= vp_i conj(wp_c) + wp_i conj(vp_c)
            a.re -= vpi.re * wpc[q].re + vpi.im * wpc[q].im + wpi.re * vpc[q].re + wpi.im * vpc[q].im;
            a.im -= vpi.im * wpc[q].re - vpi.re * wpc[q].im + wpi.im * vpc[q].re - wpi.re * vpc[q].im;
            if (ok) row[c] = a;
            if (!ok) a = cfk(0.f, 0.f);
            rr += a.re * vcc[q].re - a.im * vcc[q].im;
            ri += a.re * vcc[q].im + a.im * vcc[q].re;
            if (c < i) {   // conj(a) vc_i
                colacc[q].re += a.re * vci.re + a.im * vci.im;
                colacc[q].im += a.re * vci.im - a.im * vci.re;
            }
        }
        rr = row16_sumf(rr);
        ri = row16_sumf(ri);
        if (li == 0 && rowok) pb[i] = cfk(rr, ri);
    }
}

template <int Q16>
__global__ void __launch_bounds__(256) k_tridiag32(int n, double *H, double *R, double *E,
                                                   double *work)
{
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int RP = 2;
    const int b = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, li = lane & 15;
    const cdbl *Ag = (const cdbl *)H + (size_t)b * n * n;
    cdbl *Rg = (cdbl *)R + (size_t)b * n * n;
    double *base = work + (size_t)b * dc_work_doubles(n);
    double *e;
    cdbl *tau;
    dc_carve(n, base, &e, &tau);
    double *d = E + (size_t)b * n;
    cf *T = (cf *)smem;                      // packed lower triangle, row i at i(i+1)/2
    cf *vp = T + tri(n);                     // previous step's v (pending update)
    cf *wp = vp + n;                         // previous step's w
    cf *vc = wp + n;                         // this step's v
    cf *col = vc + n;                        // column k
    cf *pb = col + n;                        // row sums, then p
    cf *cp = pb + n;                         // 4 x n column partials
    float *red = (float *)(cp + 4 * n);      // 16

    for (int i = wave; i < n; i += 4) {
        const cdbl *row = Ag + (size_t)i * n;
        cf *trow = T + tri(i);
        for (int c = lane; c <= i; c += 64) {
            const cdbl a = row[c];
            trow[c] = cfk((float)a.re, (float)a.im);
        }
    }
    for (int i = tid; i < n; i += 256) {
        vp[i] = cfk(0.f, 0.f);
        wp[i] = cfk(0.f, 0.f);
    }
    __syncthreads();
    for (int k = 0; k < n - 1; ++k) {
        // (a) column k (rows k..) with the pending update, |a|^2 of rows k+2..
        const cf vpk = vp[k], wpk = wp[k];
        float part = 0.f;
        for (int i = k + tid; i < n; i += 256) {
            cf a = T[tri(i) + k];
            const cf vpi = vp[i], wpi = wp[i];
            a.re -= vpi.re * wpk.re + vpi.im * wpk.im + wpi.re * vpk.re + wpi.im * vpk.im;
            a.im -= vpi.im * wpk.re - vpi.re * wpk.im + wpi.im * vpk.re - wpi.re * vpk.im;
            col[i] = a;
            if (i >= k + 2) part += a.re * a.re + a.im * a.im;
        }
        part = wave_sumf(part);
        if (lane == 0) red[wave] = part;
        __syncthreads();
        // (b) reflector of column k below the diagonal
        const float xn2 = red[0] + red[1] + red[2] + red[3];
        const cf alpha = col[k + 1];
        float beta;
        cf tk, scale;
        if (xn2 == 0.f && alpha.im == 0.f) {
            tk = cfk(0.f, 0.f);
            beta = alpha.re;
            scale = cfk(0.f, 0.f);
        } else {
            const float nrm = sqrtf(alpha.re * alpha.re + alpha.im * alpha.im + xn2);
            beta = alpha.re >= 0.f ? -nrm : nrm;
            tk = cfk((beta - alpha.re) / beta, -alpha.im / beta);
            // 1 / (alpha - beta)
            const float dr = alpha.re - beta, di = alpha.im;
            const float den = dr * dr + di * di;
            scale = cfk(dr / den, -di / den);
        }
        const bool act = (tk.re != 0.f || tk.im != 0.f);
        for (int i = k + 1 + tid; i < n; i += 256) {
            cf vi;
            if (!act) vi = cfk(0.f, 0.f);
            else if (i == k + 1) vi = cfk(1.f, 0.f);
            else {
                const cf x = col[i];
                vi = cfk(x.re * scale.re - x.im * scale.im, x.re * scale.im + x.im * scale.re);
            }
            vc[i] = vi;
            if (i >= k + 2) Rg[(size_t)i * n + k] = cmk((double)vi.re, (double)vi.im);
        }
        if (tid == 0) {
            d[k] = (double)col[k].re;
            e[k] = (double)beta;
            tau[k] = cmk((double)tk.re, (double)tk.im);
        }
        __syncthreads();
        // (c) one pass: pending update applied, A22 v accumulated
        {
            cf vpc[Q16], wpc[Q16], vcc[Q16], colacc[Q16];
#pragma unroll
            for (int q = 0; q < Q16; ++q) {
                const int c = k + 1 + li + 16 * q;
                const int cc = (c < n) ? c : n - 1;
                vpc[q] = vp[cc];
                wpc[q] = wp[cc];
                vcc[q] = vc[cc];
                colacc[q] = cfk(0.f, 0.f);
            }
            for (int ib = k + 1 + 4 * wave; ib < n; ib += 16 * RP) {
                const int ilast0 = ib + 16 * (RP - 1) + 3;
                const int ilast = (ilast0 < n) ? ilast0 : n - 1;
                const int qact = ((ilast - (k + 1)) >> 4) + 1;   // wave-uniform, <= Q16
                if (qact <= 4)
                    fused_rows<Q16, 4, RP>(n, k, ib, g, li, T, vp, wp, vc, vpc, wpc, vcc, pb, colacc);
                else if (Q16 >= 8 && qact <= 8)
                    fused_rows<Q16, (Q16 >= 8 ? 8 : 4), RP>(n, k, ib, g, li, T, vp, wp, vc, vpc, wpc,
                                                            vcc, pb, colacc);
                else
                    fused_rows<Q16, Q16, RP>(n, k, ib, g, li, T, vp, wp, vc, vpc, wpc, vcc, pb, colacc);
            }
            const int qstep = ((n - 1 - (k + 1)) >> 4) + 1;
#pragma unroll
            for (int q = 0; q < Q16; ++q) {
                if (q < qstep) {
                    cf t = colacc[q];
                    t.re += __shfl_xor(t.re, 16, 64);
                    t.im += __shfl_xor(t.im, 16, 64);
                    t.re += __shfl_xor(t.re, 32, 64);
                    t.im += __shfl_xor(t.im, 32, 64);
                    const int c = k + 1 + li + 16 * q;
                    if (g == 0 && c < n) cp[wave * n + c] = t;
                }
            }
        }
        __syncthreads();
        // (d) p = tau (A22 v), p^H v
        float sre = 0.f, sim = 0.f;
        for (int i = k + 1 + tid; i < n; i += 256) {
            cf acc = pb[i];
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                acc.re += cp[w * n + i].re;
                acc.im += cp[w * n + i].im;
            }
            const cf p = cfk(tk.re * acc.re - tk.im * acc.im, tk.re * acc.im + tk.im * acc.re);
            pb[i] = p;
            const cf v = vc[i];
            sre += p.re * v.re + p.im * v.im;     // conj(p) v
            sim += p.re * v.im - p.im * v.re;
        }
        sre = wave_sumf(sre);
        sim = wave_sumf(sim);
        if (lane == 0) {
            red[4 + wave] = sre;
            red[8 + wave] = sim;
        }
        __syncthreads();
        // (e) w = p - (tau/2)(p^H v) v becomes the pending update with v
        {
            const float s_re = red[4] + red[5] + red[6] + red[7];
            const float s_im = red[8] + red[9] + red[10] + red[11];
            const cf a2 = cfk(-0.5f * (tk.re * s_re - tk.im * s_im), -0.5f * (tk.re * s_im + tk.im * s_re));
            for (int i = k + 1 + tid; i < n; i += 256) {
                const cf v = vc[i], p = pb[i];
                wp[i] = cfk(p.re + a2.re * v.re - a2.im * v.im, p.im + a2.re * v.im + a2.im * v.re);
                vp[i] = v;
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        const int i = n - 1;
        cf a = T[tri(i) + i];
        const cf vpi = vp[i], wpi = wp[i];
        a.re -= 2.f * (vpi.re * wpi.re + vpi.im * wpi.im);
        d[n - 1] = (double)a.re;
    }
}

}  // namespace

size_t tridiag32_lds_bytes(int n)
{
    return sizeof(float) * 2 * ((size_t)n * (n + 1) / 2 + 9 * (size_t)n) + 16 * sizeof(float);
}

// FP32 LDS-resident tridiagonalisation for n <= 192; H (lower triangle) is
// read, the reflector tails go to R (may equal H), d to E, e and tau to the
// eigensolver workspace.  hipErrorNotSupported above n = 192.
hipError_t launch_tridiag32(int n, int batch, const double *H, double *R, double *E, double *work,
                            hipStream_t s)
{
    if (n > 192) return hipErrorNotSupported;
    const size_t lds = tridiag32_lds_bytes(n);
    hipError_t err;
#define MSP_T32(Q)                                                                              \
    err = msp_set_lds((const void *)k_tridiag32<Q>, lds);                                       \
    if (err != hipSuccess) return err;                                                          \
    hipLaunchKernelGGL(k_tridiag32<Q>, dim3(batch), dim3(256), lds, s, n, (double *)H, R, E, work);
    if (n <= 65) { MSP_T32(4) }
    else if (n <= 129) { MSP_T32(8) }
    else { MSP_T32(12) }
#undef MSP_T32
    return hipGetLastError();
}
