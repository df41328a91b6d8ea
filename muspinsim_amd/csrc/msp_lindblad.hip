// Lindbladian path on gfx950 (muspinsim/lindbladian.py:19-173):
//
//   k_liouvillian     L = -i (H (x) 1 - 1 (x) H^T)
//                         + sum_k g_k (A_k (x) A_k* - 1/2 (A_k^H A_k (x) 1 + 1 (x) (A_k^H A_k)^T))
//                     on row-major vec(rho), one workgroup per instance
//   k_lindblad_modes  eigendecomposition L = V diag(lambda) V^{-1} (msp_nonsym.h) and
//                     the mode amplitudes amp_km = (u_k^T V)_m (V^{-1} r)_m; the
//                     matrix lives in LDS when N <= 64, else in global memory
//   k_mode_sum        P_k(t) = sum_m amp_km exp(2 pi lambda_m t)
//   k_mode_integral   I_k = sum_m amp_km / (1/tau - 2 pi lambda_m)
//   k_mode_states     rho(t) = V (c exp(2 pi lambda t)) (states without operators)
#include <hip/hip_runtime.h>

#include "msp_kernels.h"
#include "msp_nonsym.h"

namespace {

constexpr int kNT = 256;
constexpr int kLdsMaxN = 64;    // N at or below this keeps L and Z in LDS

__global__ void __launch_bounds__(kNT) k_liouvillian(int n, const double *H, int n_jump,
                                                     const double *Aj, const double *rate,
                                                     double *Lout)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int b = blockIdx.x;
    const int N = n * n;
    cdbl *Hs = (cdbl *)smem;                 // n x n
    cdbl *M = Hs + N;                        // n x n: sum_k g_k A_k^H A_k
    const cdbl *Hb = (const cdbl *)H + (size_t)b * N;
    const cdbl *Ab = (const cdbl *)Aj + (size_t)b * n_jump * N;
    const double *gb = rate + (size_t)b * n_jump;
    MSP_PFOR(idx, N)
    {
        Hs[idx] = Hb[idx];
        const int a = idx / n, c = idx - (idx / n) * n;
        cdbl s = cmk(0.0, 0.0);
        for (int k = 0; k < n_jump; ++k) {
            const cdbl *A = Ab + (size_t)k * N;
            cdbl t = cmk(0.0, 0.0);
            for (int m = 0; m < n; ++m) cfmac(t, A[m * n + a], A[m * n + c]);
            s = cadd(s, cscale(t, gb[k]));
        }
        M[idx] = s;
    }
    __syncthreads();
    cdbl *L = (cdbl *)Lout + (size_t)b * N * N;
    // element ((a,b2),(c,d)) at row a*n+b2, column c*n+d
    for (size_t e = threadIdx.x; e < (size_t)N * N; e += kNT) {
        const int row = (int)(e / N), col = (int)(e - (e / N) * N);
        const int a = row / n, b2 = row - (row / n) * n;
        const int c = col / n, d = col - (col / n) * n;
        cdbl v = cmk(0.0, 0.0);
        // -i (H_ac delta_bd - delta_ac H_db)
        cdbl h = cmk(0.0, 0.0);
        if (b2 == d) h = Hs[a * n + c];
        if (a == c) h = csub(h, Hs[d * n + b2]);
        v = cmk(h.im, -h.re);
        // dissipators
        for (int k = 0; k < n_jump; ++k) {
            const cdbl *A = Ab + (size_t)k * N;
            v = cadd(v, cscale(cmul(A[a * n + c], cconj(A[b2 * n + d])), gb[k]));
        }
        if (b2 == d) v = csub(v, cscale(M[a * n + c], 0.5));
        if (a == c) v = csub(v, cscale(M[d * n + b2], 0.5));
        L[e] = v;
    }
}

// Per-instance scratch of the modes kernel: vectors v, w, y, t, r, u (N each),
// plus Z (N x N) when the matrix is not in LDS.
__host__ __device__ inline size_t modes_vec_elems(int N) { return 6 * (size_t)N; }

// NT threads per instance: one wave when the matrix is in LDS (N <= 64: a
// sweep step's rows and columns are one element per lane, and the phase
// boundaries are those of a single wave), four above.  LDS and VMODE are
// template parameters so that the LDS form's pointers derive from the LDS
// block alone and compile to LDS instructions (a pointer that may be LDS or
// global is accessed with slower flat instructions).  VMODE (LDS, no
// eigenvectors to return): Z^H r and u_q^T Z are carried through the
// reduction instead of Z, which halves the LDS.
template <int NT, bool LDS, bool VMODE>
__global__ void __launch_bounds__(NT) k_lindblad_modes(int N, double *Lg, const double *r,
                                                       int n_ops, const double *u, double *lam,
                                                       double *amp, double *cout, double *Vout,
                                                       int32_t *info, double *work, int max_its)
{
    static_assert(LDS || !VMODE, "the vector form is the LDS form");
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double red[NT];
    const int b = blockIdx.x;
    cdbl *A, *Z, *vec;
    int ld;
    constexpr bool vmode = VMODE;
    if constexpr (LDS) {
        ld = N + 1;
        A = (cdbl *)smem;
        Z = vmode ? nullptr : A + (size_t)N * ld;
        vec = vmode ? A + (size_t)N * ld : Z + (size_t)N * ld;
        const cdbl *Lb = (const cdbl *)Lg + (size_t)b * N * N;
        MSP_PFOR(idx, N * N)
        {
            const int i = idx / N, j = idx - (idx / N) * N;
            A[i * ld + j] = Lb[idx];
        }
    } else {
        ld = N;
        A = (cdbl *)Lg + (size_t)b * N * N;
        cdbl *wb = (cdbl *)work + (size_t)b * ((size_t)N * N + modes_vec_elems(N));
        Z = wb;
        vec = wb + (size_t)N * N;
    }
    cdbl *v = vec, *w = vec + N, *y = vec + 2 * N, *t = vec + 3 * N, *rv = vec + 4 * N,
         *lv = vec + 5 * N;
    ModeVecs mv{nullptr, nullptr, 0};
    if (vmode) {
        mv.rho = rv;
        mv.omega = vec + 6 * N;
        mv.n_omega = n_ops;
        MSP_PFOR(i, N) rv[i] = ((const cdbl *)r)[(size_t)b * N + i];
        MSP_PFOR(i, n_ops * N) mv.omega[i] = ((const cdbl *)u)[(size_t)b * n_ops * N + i];
    }
    MSP_SYNC();
    hessenberg_team(N, A, ld, Z, ld, v, w, red, mv);
    const int st = schur_team(N, A, ld, Z, ld, lv, max_its, red, mv);
    if (threadIdx.x == 0) info[b] = st;
    if (st != 0) return;                      // uniform across the workgroup
    schur_vectors_team(N, A, ld, lv);
    if (!vmode) {
        MSP_PFOR(i, N) rv[i] = ((const cdbl *)r)[(size_t)b * N + i];
        MSP_SYNC();
    }
    modes_coeff_team(N, A, ld, Z, ld, rv, y);          // y = c
    for (int k = 0; k < n_ops; ++k) {
        const cdbl *uk = vmode ? mv.omega + (size_t)k * N : rv;
        if (!vmode) {
            MSP_PFOR(i, N) rv[i] = ((const cdbl *)u)[((size_t)b * n_ops + k) * N + i];
            MSP_SYNC();
        }
        modes_proj_team(N, A, ld, Z, ld, uk, t, w);    // w = o
        MSP_PFOR(m, N) ((cdbl *)amp)[((size_t)b * n_ops + k) * N + m] = cmul(w[m], y[m]);
        MSP_SYNC();
    }
    MSP_PFOR(m, N)
    {
        ((cdbl *)lam)[(size_t)b * N + m] = lv[m];
        if (cout) ((cdbl *)cout)[(size_t)b * N + m] = y[m];
    }
    if (Vout) {
        modes_vectors_team(N, A, ld, Z, ld);
        cdbl *Vb = (cdbl *)Vout + (size_t)b * N * N;
        MSP_PFOR(idx, N * N)
        {
            const int i = idx / N, j = idx - (idx / N) * N;
            Vb[idx] = Z[i * ld + j];
        }
    }
}

// exp(2 pi lambda t)
__device__ inline cdbl mode_phase(cdbl l, double t)
{
    const double twopi = 6.283185307179586;
    const double mag = exp(twopi * l.re * t);
    double sn, cs;
    sincos(twopi * l.im * t, &sn, &cs);
    return cmk(mag * cs, mag * sn);
}

// one thread per (instance, operator, time)
__global__ void k_mode_sum(int N, int batch, int n_ops, int ntimes, const double *lam,
                           const double *amp, const double *times, const double *w,
                           int out_complex, double *out)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)batch * n_ops * ntimes;
    if (e >= total) return;
    const int ti = (int)(e % ntimes);
    const size_t bk = e / ntimes;
    const int b = (int)(bk / n_ops);
    const cdbl *l = (const cdbl *)lam + (size_t)b * N;
    const cdbl *a = (const cdbl *)amp + bk * N;
    const double tt = times[ti];
    cdbl s = cmk(0.0, 0.0);
    for (int m = 0; m < N; ++m) s = cadd(s, cmul(a[m], mode_phase(l[m], tt)));
    if (out_complex) ((cdbl *)out)[e] = s;
    else out[e] = (w ? w[b] : 1.0) * s.re;
}

__global__ void k_mode_integral(int N, int batch, int n_ops, const double *lam, const double *amp,
                                double tau, double out_scale, const double *w, int out_complex,
                                double *out)
{
    const int bk = blockIdx.x * blockDim.x + threadIdx.x;
    if (bk >= batch * n_ops) return;
    const int b = bk / n_ops;
    const cdbl *l = (const cdbl *)lam + (size_t)b * N;
    const cdbl *a = (const cdbl *)amp + (size_t)bk * N;
    const double twopi = 6.283185307179586;
    cdbl s = cmk(0.0, 0.0);
    for (int m = 0; m < N; ++m)
        s = cadd(s, cdivc(a[m], cmk(1.0 / tau - twopi * l[m].re, -twopi * l[m].im)));
    s = cscale(s, out_scale);
    if (out_complex) ((cdbl *)out)[bk] = s;
    else out[bk] = (w ? w[b] : 1.0) * s.re;
}

// out[t][i] = sum_m V_im c_m exp(2 pi lambda_m t), one instance
__global__ void k_mode_states(int N, int ntimes, const double *lam, const double *c,
                              const double *V, const double *times, double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int ti = blockIdx.y;
    if (i >= N || ti >= ntimes) return;
    const cdbl *l = (const cdbl *)lam, *cc = (const cdbl *)c, *Vm = (const cdbl *)V;
    cdbl s = cmk(0.0, 0.0);
    for (int m = 0; m < N; ++m)
        s = cadd(s, cmul(Vm[(size_t)i * N + m], cmul(cc[m], mode_phase(l[m], times[ti]))));
    ((cdbl *)out)[(size_t)ti * N + i] = s;
}

inline unsigned cdivu(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

size_t lindblad_modes_workspace(int N, int batch)
{
    if (N <= kLdsMaxN) return 0;
    return sizeof(cdbl) * ((size_t)N * N + modes_vec_elems(N)) * (size_t)batch;
}

hipError_t launch_liouvillian(int n, int batch, const double *H, int n_jump, const double *A,
                              const double *rate, double *L, hipStream_t s)
{
    const size_t lds = sizeof(cdbl) * 2 * (size_t)n * n;
    ProfScope ps_("k_liouvillian", s);
    hipLaunchKernelGGL(k_liouvillian, dim3(batch), dim3(kNT), lds, s, n, H, n_jump, A, rate, L);
    return hipGetLastError();
}

hipError_t launch_lindblad_modes(int N, int batch, double *L, const double *r, int n_ops,
                                 const double *u, double *lam, double *amp, double *c, double *V,
                                 int32_t *info, double *work, int max_its, hipStream_t s)
{
    const int in_lds = N <= kLdsMaxN;
    // (without V the LDS form holds the matrix, the vectors and the n_ops rows
    // u_q^T Z, not Z)
    const size_t lds = !in_lds ? 0
                       : V ? sizeof(cdbl) * (2 * (size_t)N * (N + 1) + modes_vec_elems(N))
                           : sizeof(cdbl) * ((size_t)N * (N + 1) + modes_vec_elems(N) + (size_t)n_ops * N);
    const void *fn = !in_lds ? (const void *)k_lindblad_modes<kNT, false, false>
                     : V     ? (const void *)k_lindblad_modes<64, true, false>
                             : (const void *)k_lindblad_modes<64, true, true>;
    hipError_t err = msp_set_lds(fn, lds);
    if (err != hipSuccess) return err;
    ProfScope ps_("k_lindblad_modes", s);
    if (!in_lds)
        hipLaunchKernelGGL((k_lindblad_modes<kNT, false, false>), dim3(batch), dim3(kNT), lds, s, N, L,
                           r, n_ops, u, lam, amp, c, V, info, work, max_its);
    else if (V)
        hipLaunchKernelGGL((k_lindblad_modes<64, true, false>), dim3(batch), dim3(64), lds, s, N, L, r,
                           n_ops, u, lam, amp, c, V, info, work, max_its);
    else
        hipLaunchKernelGGL((k_lindblad_modes<64, true, true>), dim3(batch), dim3(64), lds, s, N, L, r,
                           n_ops, u, lam, amp, c, V, info, work, max_its);
    return hipGetLastError();
}

hipError_t launch_mode_sum(int N, int batch, int n_ops, int ntimes, const double *lam,
                           const double *amp, const double *times, const double *w,
                           int out_complex, double *out, hipStream_t s)
{
    const size_t total = (size_t)batch * n_ops * ntimes;
    ProfScope ps_("k_mode_sum", s);
    hipLaunchKernelGGL(k_mode_sum, dim3(cdivu(total, 256)), dim3(256), 0, s, N, batch, n_ops,
                       ntimes, lam, amp, times, w, out_complex, out);
    return hipGetLastError();
}

hipError_t launch_mode_integral(int N, int batch, int n_ops, const double *lam, const double *amp,
                                double tau, double out_scale, const double *w, int out_complex,
                                double *out, hipStream_t s)
{
    ProfScope ps_("k_mode_integral", s);
    hipLaunchKernelGGL(k_mode_integral, dim3(cdivu((size_t)batch * n_ops, 256)), dim3(256), 0, s,
                       N, batch, n_ops, lam, amp, tau, out_scale, w, out_complex, out);
    return hipGetLastError();
}

hipError_t launch_mode_states(int N, int ntimes, const double *lam, const double *c,
                              const double *V, const double *times, double *out, hipStream_t s)
{
    ProfScope ps_("k_mode_states", s);
    hipLaunchKernelGGL(k_mode_states, dim3(cdivu(N, 256), ntimes), dim3(256), 0, s, N, ntimes,
                       lam, c, V, times, out);
    return hipGetLastError();
}
