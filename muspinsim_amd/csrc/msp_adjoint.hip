// Adjoint (vector-Jacobian product) of the fast-path muon polarisation
//   P_b(t) = (1/4d) sum_jk |O_jk|^2 cos(2 pi (E_j - E_k) t),  O = U^H sigma U,
// with respect to the Hamiltonian, for fitting (the reference fits by finite
// differences, muspinsim/fitting.py:99-105).  For upstream weights r_b(t) =
// dL/dP_b(t), L = sum_bt r_b(t) P_b(t) has the eigenbasis gradient G'
// (dL = Re tr(G' U^H dH U)):
//   G'_jj = -(pi/d) sum_k |O_jk|^2 S_jk                      (eigenvalue part)
//   G'_jm = (1/2d) Y_jm / (E_j - E_m),  Y = [R o O, O]       (eigenvector part)
//   G'_jm = (1/2d) sum_k O_jk O_km (-2 pi S_jk)  for E_j ~ E_m (the limit of
//           the divided difference, so degenerate levels are well defined)
// with R_jk = sum_t r(t) cos(w_jk t), S_jk = sum_t r(t) t sin(w_jk t),
// w_jk = 2 pi (E_j - E_k).  G = U G' U^H, summed over instances, gives
// dL/dtheta = Re tr(G dH/dtheta) for any parameter of the field-independent H.
#include <hip/hip_runtime.h>

#include "msp_kernels.h"
#include "msp_team.h"

namespace {

constexpr int TJ = 16;           // (j, k) tile edge
constexpr int TT = 64;           // times per LDS chunk

// R and S of one (j, k) tile of one instance; phases of the tile's rows and
// columns for a chunk of times are formed once in LDS.
__global__ void __launch_bounds__(256) k_adjoint_freq(int n, int ntimes, const double *E,
                                                      const double *times, const double *r,
                                                      const int32_t *slot, const double *w,
                                                      double scale, double *R, double *S)
{
    __shared__ cdbl Pj[TJ][TT + 1];
    __shared__ cdbl Pk[TJ][TT + 1];
    __shared__ double rt[TT], tt[TT];
    const int b = blockIdx.z;
    const int j0 = blockIdx.y * TJ, k0 = blockIdx.x * TJ;
    const int tid = threadIdx.x, jj = tid / TJ, kk = tid % TJ;
    const double *Eb = E + (size_t)b * n;
    const double Ec = 0.5 * (Eb[0] + Eb[n - 1]);
    const double *rb = r + (size_t)(slot ? slot[b] : 0) * ntimes;
    const double twopi = 6.283185307179586;
    double sr = 0.0, ss = 0.0;
    for (int t0 = 0; t0 < ntimes; t0 += TT) {
        __syncthreads();
        for (int l = tid; l < TT; l += 256) {
            const int t = t0 + l;
            rt[l] = (t < ntimes) ? rb[t] : 0.0;
            tt[l] = (t < ntimes) ? times[t] : 0.0;
        }
        for (int l = tid; l < 2 * TJ * TT; l += 256) {
            const int which = l / (TJ * TT), rem = l % (TJ * TT);
            const int a = rem / TT, c = rem % TT;
            const int t = t0 + c;
            const int idx = (which ? k0 : j0) + a;
            double sn = 0.0, cs = 1.0;
            if (idx < n && t < ntimes) sincos(twopi * (Eb[idx] - Ec) * times[t], &sn, &cs);
            if (which) Pk[a][c] = cmk(cs, sn);
            else Pj[a][c] = cmk(cs, sn);
        }
        __syncthreads();
        const int lim = (ntimes - t0 < TT) ? ntimes - t0 : TT;
        for (int c = 0; c < lim; ++c) {
            // exp(i w_jk t) = Pj conj(Pk)
            const cdbl pj = Pj[jj][c], pk = Pk[kk][c];
            const double re = pj.re * pk.re + pj.im * pk.im;
            const double im = pj.im * pk.re - pj.re * pk.im;
            sr = fma(rt[c], re, sr);
            ss = fma(rt[c] * tt[c], im, ss);
        }
    }
    const int j = j0 + jj, k = k0 + kk;
    if (j < n && k < n) {
        const double f = scale * (w ? w[b] : 1.0);
        R[(size_t)b * n * n + (size_t)j * n + k] = f * sr;
        S[(size_t)b * n * n + (size_t)j * n + k] = f * ss;
    }
}

// M = R o O (elementwise, R real, O complex)
__global__ void k_hadamard_real(int n, int batch, const double *R, const double *O, double *M)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)batch * n * n) return;
    ((cdbl *)M)[e] = cscale(((const cdbl *)O)[e], R[e]);
}

// G'_jm from Y = M O - O M, the eigenvalues, S and O (see the file header);
// the result is made Hermitian, G' <- (G' + G'^H)/2, which leaves Re tr(G' D)
// unchanged for Hermitian D.
__global__ void k_adjoint_gprime(int n, int batch, const double *E, const double *O,
                                 const double *MO, const double *OM, const double *S, double d,
                                 double degtol, double *G)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)batch * n * n) return;
    const int b = (int)(e / ((size_t)n * n));
    const int jm = (int)(e % ((size_t)n * n));
    const int j = jm / n, m = jm % n;
    const double *Eb = E + (size_t)b * n;
    const cdbl *Ob = (const cdbl *)O + (size_t)b * n * n;
    const double *Sb = S + (size_t)b * n * n;
    const double twopi = 6.283185307179586;
    const double scaleE = fmax(fabs(Eb[0]), fabs(Eb[n - 1]));
    cdbl g[2];
    // entry (j, m) and its mirror (m, j), averaged conjugated
    for (int side = 0; side < 2; ++side) {
        const int p = side ? m : j, q = side ? j : m;
        cdbl v;
        if (p == q) {
            double a = 0.0;
            for (int k = 0; k < n; ++k) a += cabs2(Ob[(size_t)p * n + k]) * Sb[(size_t)p * n + k];
            v = cmk(-0.5 * twopi / d * a, 0.0);
        } else {
            const double dE = Eb[p] - Eb[q];
            if (fabs(dE) > degtol * fmax(scaleE, 1.0)) {
                const size_t o = (size_t)b * n * n + (size_t)p * n + q;
                const cdbl y = csub(((const cdbl *)MO)[o], ((const cdbl *)OM)[o]);
                v = cscale(y, 0.5 / (d * dE));
            } else {
                cdbl acc = cmk(0.0, 0.0);
                for (int k = 0; k < n; ++k)
                    acc = cadd(acc, cscale(cmul(Ob[(size_t)p * n + k], Ob[(size_t)k * n + q]),
                                           -twopi * Sb[(size_t)p * n + k]));
                v = cscale(acc, 0.5 / d);
            }
        }
        g[side] = v;
    }
    // G'_jm <- (G'_jm + conj(G'_mj)) / 2
    ((cdbl *)G)[e] = cscale(cadd(g[0], cconj(g[1])), 0.5);
}

// out[i] = sum_b X[b][i] (deterministic order), complex
__global__ void k_batch_sum(int len, int batch, const double *X, double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    cdbl acc = cmk(0.0, 0.0);
    for (int b = 0; b < batch; ++b) acc = cadd(acc, ((const cdbl *)X)[(size_t)b * len + i]);
    ((cdbl *)out)[i] = cadd(((cdbl *)out)[i], acc);
}

// out[p] = Re sum_ab G_ab dH_p[b][a], one workgroup per parameter
__global__ void __launch_bounds__(256) k_trace_products(int n, const double *G, const double *dH,
                                                        double *out)
{
    __shared__ double red[8];
    const int p = blockIdx.x;
    const cdbl *Gm = (const cdbl *)G;
    const cdbl *D = (const cdbl *)dH + (size_t)p * n * n;
    double s = 0.0;
    for (int idx = threadIdx.x; idx < n * n; idx += 256) {
        const int a = idx / n, c = idx % n;
        const cdbl g = Gm[idx], h = D[(size_t)c * n + a];
        s += g.re * h.re - g.im * h.im;
    }
    s = team_sum(s, red);
    if (threadIdx.x == 0) out[p] = s;
}

inline unsigned cdivu(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

hipError_t launch_adjoint_freq(int n, int batch, int ntimes, const double *E, const double *times,
                               const double *r, const int32_t *slot, const double *w, double scale,
                               double *R, double *S, hipStream_t s)
{
    ProfScope ps_("k_adjoint_freq", s);
    dim3 grid(cdivu(n, TJ), cdivu(n, TJ), batch);
    hipLaunchKernelGGL(k_adjoint_freq, grid, dim3(256), 0, s, n, ntimes, E, times, r, slot, w,
                       scale, R, S);
    return hipGetLastError();
}

hipError_t launch_hadamard_real(int n, int batch, const double *R, const double *O, double *M,
                                hipStream_t s)
{
    ProfScope ps_("k_hadamard_real", s);
    hipLaunchKernelGGL(k_hadamard_real, dim3(cdivu((size_t)batch * n * n, 256)), dim3(256), 0, s,
                       n, batch, R, O, M);
    return hipGetLastError();
}

hipError_t launch_adjoint_gprime(int n, int batch, const double *E, const double *O,
                                 const double *MO, const double *OM, const double *S, double d,
                                 double degtol, double *G, hipStream_t s)
{
    ProfScope ps_("k_adjoint_gprime", s);
    hipLaunchKernelGGL(k_adjoint_gprime, dim3(cdivu((size_t)batch * n * n, 256)), dim3(256), 0, s,
                       n, batch, E, O, MO, OM, S, d, degtol, G);
    return hipGetLastError();
}

hipError_t launch_batch_sum(int len, int batch, const double *X, double *out, hipStream_t s)
{
    ProfScope ps_("k_batch_sum", s);
    hipLaunchKernelGGL(k_batch_sum, dim3(cdivu(len, 256)), dim3(256), 0, s, len, batch, X, out);
    return hipGetLastError();
}

hipError_t launch_trace_products(int n, int nparam, const double *G, const double *dH, double *out,
                                 hipStream_t s)
{
    ProfScope ps_("k_trace_products", s);
    hipLaunchKernelGGL(k_trace_products, dim3(nparam), dim3(256), 0, s, n, G, dH, out);
    return hipGetLastError();
}
