// Complex non-Hermitian eigendecomposition as team functions (device: one
// workgroup per matrix; host: one thread), for the Lindbladian path
// (muspinsim/lindbladian.py:43-173, which calls numpy.linalg.eig and solve).
//
//   L = Z T Z^H                Householder Hessenberg reduction, then complex
//                              single-shift QR sweeps (Givens bulge chase) to
//                              upper-triangular Schur form T, Z accumulated
//   T X = X diag(lambda)       X unit upper triangular, column k by back
//                              substitution (T - lambda_k) x = 0, x_k = 1
//   V = Z X                    right eigenvectors (columns)
//
// The Lindblad expectation values only need the mode amplitudes
//   c = V^{-1} r = X^{-1} (Z^H r)          (r = vec(rho0), row-major)
//   o = u^T V   = (u^T Z) X                (u = vec(O^T))
// so V^{-1} is never formed: Z is unitary and X triangular.
//
// Storage: A (n x n, row stride lda) holds L on entry and T on exit with the
// strictly lower triangle reused for X: X[i][k] (i < k) is kept at A[k][i].
// Z (n x n, stride ldz) receives the Schur vectors.
#pragma once

#include "msp_team.h"

// Phase boundary of the team functions below.  A team of one wave (the
// LDS-resident form: every shared array is in LDS, whose operations a wave
// executes in issue order) needs only the compiler not to move memory
// operations across it; larger teams take a workgroup barrier.
#ifdef MSP_HOST_BUILD
#define NS_SYNC() ((void)0)
#else
#define NS_SYNC() ns_sync_dev()
__device__ inline void ns_sync_dev()
{
    if (blockDim.x <= 64) __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    else __syncthreads();
}
#endif

// Team sum; a one-wave team reduces with DPP lane moves (no LDS).
MSP_HD double ns_team_sum(double v, double *red)
{
#ifdef MSP_HOST_BUILD
    return team_sum(v, red);
#else
    if (blockDim.x > 64) return team_sum(v, red);
    return wave_sum_all(v);
#endif
}

// Team maximum; a one-wave team reduces with DPP lane moves (no LDS).
MSP_HD double ns_team_max(double v, double *red)
{
#ifdef MSP_HOST_BUILD
    return team_max(v, red);
#else
    if (blockDim.x > 64) return team_max(v, red);
    v = fmax(v, dpp_f64<0xB1, 0xF>(v));    // quad_perm [1,0,3,2]
    v = fmax(v, dpp_f64<0x4E, 0xF>(v));    // quad_perm [2,3,0,1]
    v = fmax(v, dpp_f64<0x141, 0xF>(v));   // row_half_mirror
    v = fmax(v, dpp_f64<0x140, 0xF>(v));   // row_mirror
    v = fmax(v, dpp_f64<0x142, 0xA>(v));   // row_bcast15 -> rows 1, 3
    v = fmax(v, dpp_f64<0x143, 0xC>(v));   // row_bcast31 -> rows 2, 3
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(x & 0xffffffffll), 63);
    const int hi = __builtin_amdgcn_readlane((int)(x >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
#endif
}

// |re| + |im|, the LAPACK cabs1 magnitude used in deflation tests
MSP_HD double cabs1(cdbl a) { return fabs(a.re) + fabs(a.im); }
MSP_HD double cabs(cdbl a) { return hypot(a.re, a.im); }

// ---------------------------------------------------------------------------
// Vectors carried through the reduction instead of Z (when the eigenvectors
// themselves are not wanted): rho = Z^H r and the rows omega_q = u_q^T Z,
// updated with every reflector and rotation that would update Z.
struct ModeVecs {
    cdbl *rho;      // n (starts as r)
    cdbl *omega;    // n_omega x n, row q at q n (start as u_q)
    int n_omega;
};

// rho <- H^H rho and omega_q <- omega_q H for H = I - tau v v^H (v on rows
// k+1 .. n-1): one thread per vector forms its dot, then every thread updates.
MSP_HD void modevecs_reflect(int n, int k, const cdbl *v, cdbl tau, ModeVecs mv, cdbl *dots)
{
    const int nv = 1 + mv.n_omega;
    MSP_PFOR(q, nv)
    {
        cdbl s = cmk(0.0, 0.0);
        if (q == 0) {
#pragma unroll 8
            for (int r = k + 1; r < n; ++r) cfmac(s, v[r], mv.rho[r]);      // v^H rho
            dots[0] = cmul(cconj(tau), s);
        } else {
            const cdbl *om = mv.omega + (size_t)(q - 1) * n;
#pragma unroll 8
            for (int r = k + 1; r < n; ++r) cfma(s, om[r], v[r]);           // omega v
            dots[q] = cmul(s, tau);
        }
    }
    NS_SYNC();
    MSP_PFOR(idx, nv * (n - k - 1))
    {
        const int q = idx / (n - k - 1), r = k + 1 + idx % (n - k - 1);
        if (q == 0) mv.rho[r] = csub(mv.rho[r], cmul(v[r], dots[0]));
        else {
            cdbl *om = mv.omega + (size_t)(q - 1) * n;
            om[r] = csub(om[r], cmul(dots[q], cconj(v[r])));
        }
    }
    NS_SYNC();
}

// ---------------------------------------------------------------------------
// Hessenberg reduction: A <- Q^H A Q (upper Hessenberg), and Z <- Q, or
// (Z null) the mode vectors mv transformed with Q.
// v, w: team scratch of n complex each; red: team scratch of MSP_NT doubles.
// Every update loop gives a thread whole columns or rows (no index division),
// with the inner dots unrolled so several LDS loads are in flight.
MSP_HD void hessenberg_team(int n, cdbl *A, int lda, cdbl *Z, int ldz, cdbl *v, cdbl *w,
                            double *red, ModeVecs mv = ModeVecs{nullptr, nullptr, 0})
{
    if (Z) {
        MSP_PFOR(i, n)
        {
            for (int j = 0; j < n; ++j) Z[i * ldz + j] = cmk(i == j ? 1.0 : 0.0, 0.0);
        }
    }
    NS_SYNC();
    for (int k = 0; k + 2 < n; ++k) {
        // reflector H = I - tau v v^H with H^H [alpha; x] = [beta; 0]
        double part = 0.0;
        MSP_PFOR(i, n - k - 2) part += cabs2(A[(k + 2 + i) * lda + k]);
        const double xn2 = ns_team_sum(part, red);
        const cdbl alpha = A[(k + 1) * lda + k];
        if (xn2 == 0.0 && alpha.im == 0.0) {
            NS_SYNC();
            continue;                                  // already reduced: tau = 0
        }
        const double anorm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn2);
        const double beta = (alpha.re >= 0.0) ? -anorm : anorm;
        const cdbl tau = cmk((beta - alpha.re) / beta, -alpha.im / beta);
        const cdbl scal = cdivc(cmk(1.0, 0.0), cmk(alpha.re - beta, alpha.im));
        MSP_PFOR(i, n - k - 1)
        {
            const int r = k + 1 + i;
            v[r] = (i == 0) ? cmk(1.0, 0.0) : cmul(A[r * lda + k], scal);
        }
        NS_SYNC();
        // left: A[k+1.., j] -= conj(tau) v (v^H A[k+1.., j]), columns j > k
        MSP_PFOR(jj, n - k - 1)
        {
            const int j = k + 1 + jj;
            cdbl s = cmk(0.0, 0.0);
#pragma unroll 8
            for (int r = k + 1; r < n; ++r) cfmac(s, v[r], A[r * lda + j]);
            w[j] = cmul(cconj(tau), s);
        }
        NS_SYNC();
        MSP_PFOR(jj, n - k)
        {
            const int j = k + jj;
            if (j == k) {
                A[(k + 1) * lda + k] = cmk(beta, 0.0);
                for (int r = k + 2; r < n; ++r) A[r * lda + k] = cmk(0.0, 0.0);
            } else {
                const cdbl wj = w[j];
#pragma unroll 8
                for (int r = k + 1; r < n; ++r) A[r * lda + j] = csub(A[r * lda + j], cmul(v[r], wj));
            }
        }
        NS_SYNC();
        // right: A[i, k+1..] -= (A[i, k+1..] v) tau v^H for every row
        MSP_PFOR(i, n)
        {
            cdbl s = cmk(0.0, 0.0);
#pragma unroll 8
            for (int r = k + 1; r < n; ++r) cfma(s, A[i * lda + r], v[r]);
            const cdbl wi = cmul(s, tau);
#pragma unroll 8
            for (int r = k + 1; r < n; ++r) A[i * lda + r] = csub(A[i * lda + r], cmul(wi, cconj(v[r])));
        }
        NS_SYNC();
        if (Z) {
            MSP_PFOR(i, n)
            {
                cdbl s = cmk(0.0, 0.0);
#pragma unroll 8
                for (int r = k + 1; r < n; ++r) cfma(s, Z[i * ldz + r], v[r]);
                const cdbl wi = cmul(s, tau);
#pragma unroll 8
                for (int r = k + 1; r < n; ++r) Z[i * ldz + r] = csub(Z[i * ldz + r], cmul(wi, cconj(v[r])));
            }
            NS_SYNC();
        } else {
            modevecs_reflect(n, k, v, tau, mv, w);
        }
    }
}

// 1/x and 1/sqrt(x) from the hardware estimates refined by two Newton steps
// (within an ulp or two of the IEEE result for normal x, a fraction of the
// instructions of an IEEE division or square root); exact on the host.
MSP_HD double ns_recip(double x)
{
#ifdef MSP_HOST_BUILD
    return 1.0 / x;
#else
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
#endif
}
MSP_HD double ns_rsqrt(double x)
{
#ifdef MSP_HOST_BUILD
    return 1.0 / sqrt(x);
#else
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    y = y * fma(-h * y, y, 1.5);
    y = y * fma(-h * y, y, 1.5);
    return y;
#endif
}

// Givens rotation G = [c s; -conj(s) c] with G [f; g] = [r; 0], c real:
// c = |f| / nrm, s = f conj(g) / (|f| nrm), r = f nrm / |f| with
// nrm = sqrt(|f|^2 + |g|^2), all from f and g scaled by their largest
// component and one reciprocal square root, 1 / sqrt(|f|^2 (|f|^2 + |g|^2))
// (the rotation's step is in the dependent chain of every QR sweep step).
MSP_HD void givens(cdbl f, cdbl g, double &c, cdbl &s, cdbl &r)
{
    const double sc = fmax(fmax(fabs(f.re), fabs(f.im)), fmax(fabs(g.re), fabs(g.im)));
    if (g.re == 0.0 && g.im == 0.0) {
        c = 1.0;
        s = cmk(0.0, 0.0);
        r = f;
        return;
    }
    const double isc = ns_recip(sc);
    const cdbl fs = cscale(f, isc), gs = cscale(g, isc);
    const double f2 = fs.re * fs.re + fs.im * fs.im, g2 = gs.re * gs.re + gs.im * gs.im;
    if (f2 < 1e-300) {                                 // f = 0, or |f| < 1e-150 |g|
        const double ig = ns_rsqrt(g2);
        c = 0.0;
        s = cmk(gs.re * ig, -gs.im * ig);
        r = cmk(g2 * ig * sc, 0.0);
        return;
    }
    const double t = f2 + g2;
    const double inv = ns_rsqrt(f2 * t);               // 1 / (|f| nrm), scaled
    c = f2 * inv;                                      // |f| / nrm
    s = cscale(cmul(fs, cconj(gs)), inv);
    r = cscale(f, t * inv);                            // f nrm / |f|
}

// ---------------------------------------------------------------------------
// Upper Hessenberg A -> upper triangular T = Z^H A Z by single-shift QR with
// Wilkinson shifts (exceptional shifts after 10 and 20 stalled sweeps).
// lam receives the diagonal.  Returns 0, or the 1-based index of an
// unconverged eigenvalue after max_its sweeps per eigenvalue.  With Z null
// the rotations go to the mode vectors mv instead.  red: team scratch of
// MSP_NT doubles (the split search is a team maximum over the subdiagonal).
MSP_HD int schur_team(int n, cdbl *A, int lda, cdbl *Z, int ldz, cdbl *lam, int max_its,
                      double *red, ModeVecs mv = ModeVecs{nullptr, nullptr, 0})
{
    const double ulp = 1.1102230246251565e-16;
    const double smlnum = 2.2250738585072014e-308 * ((double)n / ulp);
    int ihi = n - 1, its = 0, total = 0;
    while (ihi >= 0) {
        // split point: largest l <= ihi with A[l][l-1] negligible (0 if none),
        // every l tested at once
        double best = 0.0;
        MSP_PFOR(t, ihi)
        {
            const int l = t + 1;
            const double sub = cabs1(A[l * lda + l - 1]);
            double tst = cabs1(A[(l - 1) * lda + l - 1]) + cabs1(A[l * lda + l]);
            if (tst == 0.0) {
                if (l >= 2) tst += fabs(A[(l - 1) * lda + l - 2].re);
                if (l + 1 <= ihi) tst += fabs(A[(l + 1) * lda + l].re);
            }
            if (sub <= smlnum || sub <= ulp * tst) best = fmax(best, (double)l);
        }
        const int l = (int)ns_team_max(best, red);
        if (l > 0) {
            if (MSP_TID == 0) A[l * lda + l - 1] = cmk(0.0, 0.0);
        }
        if (l == ihi) {
            if (MSP_TID == 0) lam[ihi] = A[ihi * lda + ihi];
            NS_SYNC();
            --ihi;
            its = 0;
            continue;
        }
        if (++total > max_its * n) return ihi + 1;
        // shift
        cdbl sft;
        if (its == 10) {
            const cdbl a = A[l * lda + l];
            sft = cmk(a.re + 0.75 * fabs(A[(l + 1) * lda + l].re), a.im);
        } else if (its == 20) {
            const cdbl a = A[ihi * lda + ihi];
            sft = cmk(a.re + 0.75 * fabs(A[ihi * lda + ihi - 1].re), a.im);
        } else {
            // eigenvalue of the trailing 2 x 2 block closest to its last diagonal
            const cdbl a = A[(ihi - 1) * lda + ihi - 1], b = A[(ihi - 1) * lda + ihi];
            const cdbl c = A[ihi * lda + ihi - 1], d = A[ihi * lda + ihi];
            const cdbl hm = cscale(csub(a, d), 0.5);
            const cdbl disc = cadd(cmul(hm, hm), cmul(b, c));
            // principal square root
            const double m = cabs(disc);
            double sr = sqrt(0.5 * (m + fabs(disc.re)));
            double si = (sr > 0.0) ? 0.5 * disc.im / sr : 0.0;
            if (disc.re < 0.0) {
                const double t = sr;
                sr = fabs(si);
                si = (disc.im >= 0.0) ? t : -t;
            }
            cdbl root = cmk(sr, si);
            // lambda = d + hm -/+ root; pick the one nearer d: d + (hm - root) or d + (hm + root)
            const cdbl e1 = csub(hm, root), e2 = cadd(hm, root);
            sft = cadd(d, (cabs1(e1) <= cabs1(e2)) ? e1 : e2);
        }
        ++its;
        // single-shift sweep over rows/cols l .. ihi, applied to the full
        // matrix (columns up to n-1, rows from 0) and to Z or the mode vectors
        for (int k = l; k < ihi; ++k) {
            cdbl f, g;
            if (k == l) {
                f = csub(A[l * lda + l], sft);
                g = A[(l + 1) * lda + l];
            } else {
                f = A[k * lda + k - 1];
                g = A[(k + 1) * lda + k - 1];
            }
            double c;
            cdbl s, r;
            givens(f, g, c, s, r);
            NS_SYNC();
            // rows k, k+1: [a; b] <- [c a + s b; -conj(s) a + c b]
            const int j0 = (k == l) ? k : k - 1;
            MSP_PFOR(jj, n - j0)
            {
                const int j = j0 + jj;
                const cdbl a = A[k * lda + j], b = A[(k + 1) * lda + j];
                if (k > l && j == k - 1) {
                    A[k * lda + j] = r;
                    A[(k + 1) * lda + j] = cmk(0.0, 0.0);
                } else {
                    A[k * lda + j] = cadd(cscale(a, c), cmul(s, b));
                    A[(k + 1) * lda + j] = csub(cscale(b, c), cmul(cconj(s), a));
                }
            }
            NS_SYNC();
            // columns k, k+1 of A (rows 0 .. min(k+2, ihi)) and of Z (all rows):
            // [p q] <- [c p + conj(s) q, -s p + c q]; without Z, rho takes the
            // row rotation and each omega_q the column rotation
            const int iend = (k + 2 < ihi) ? k + 2 : ihi;
            MSP_PFOR(i, iend + 1)
            {
                const cdbl p = A[i * lda + k], q = A[i * lda + k + 1];
                A[i * lda + k] = cadd(cscale(p, c), cmul(cconj(s), q));
                A[i * lda + k + 1] = csub(cscale(q, c), cmul(s, p));
            }
            if (Z) {
                MSP_PFOR(i, n)
                {
                    const cdbl p = Z[i * ldz + k], q = Z[i * ldz + k + 1];
                    Z[i * ldz + k] = cadd(cscale(p, c), cmul(cconj(s), q));
                    Z[i * ldz + k + 1] = csub(cscale(q, c), cmul(s, p));
                }
            } else {
                MSP_PFOR(q2, 1 + mv.n_omega)
                {
                    if (q2 == 0) {
                        const cdbl a = mv.rho[k], b = mv.rho[k + 1];
                        mv.rho[k] = cadd(cscale(a, c), cmul(s, b));
                        mv.rho[k + 1] = csub(cscale(b, c), cmul(cconj(s), a));
                    } else {
                        cdbl *om = mv.omega + (size_t)(q2 - 1) * n;
                        const cdbl p = om[k], q = om[k + 1];
                        om[k] = cadd(cscale(p, c), cmul(cconj(s), q));
                        om[k + 1] = csub(cscale(q, c), cmul(s, p));
                    }
                }
            }
            NS_SYNC();
        }
    }
    NS_SYNC();
    return 0;
}

// ---------------------------------------------------------------------------
// Eigenvectors of the Schur form: for each k, X[i][k] (i < k) with
// (T - lambda_k) x = 0, x_k = 1, stored at A[k][i].  Near-zero pivots are
// replaced by smin = max(ulp |lambda_k|, smlnum) as LAPACK's ztrevc does.
MSP_HD void schur_vectors_team(int n, cdbl *A, int lda, const cdbl *lam)
{
    const double ulp = 1.1102230246251565e-16;
    const double smlnum = 2.2250738585072014e-308 * ((double)n / ulp);
    MSP_PFOR(k, n)
    {
        const cdbl lk = lam[k];
        const double smin = fmax(ulp * cabs1(lk), smlnum);
        // x_i for i = k-1 .. 0; x stored in the lower triangle row k
        for (int i = k - 1; i >= 0; --i) {
            cdbl s = A[i * lda + k];                  // T_ik * x_k (x_k = 1)
#pragma unroll 8
            for (int j = i + 1; j < k; ++j) cfma(s, A[i * lda + j], A[k * lda + j]);
            cdbl den = csub(lam[i], lk);
            if (cabs1(den) < smin) den = cmk(smin, 0.0);
            A[k * lda + i] = cscale(cdivc(s, den), -1.0);
        }
    }
    NS_SYNC();
}

// X[i][k] accessor (unit diagonal, zero below)
MSP_HD cdbl xel(const cdbl *A, int lda, int i, int k)
{
    if (i == k) return cmk(1.0, 0.0);
    if (i > k) return cmk(0.0, 0.0);
    return A[k * lda + i];
}

// c = X^{-1} (Z^H r).  y: team scratch (n); out: c (n).
MSP_HD void modes_coeff_team(int n, const cdbl *A, int lda, const cdbl *Z, int ldz, const cdbl *r,
                             cdbl *y)
{
    MSP_PFOR(m, n)
    {
        cdbl s = cmk(0.0, 0.0);
        if (Z) {
#pragma unroll 8
            for (int i = 0; i < n; ++i) cfmac(s, Z[i * ldz + m], r[i]);
        } else {
            s = r[m];                       // r is already Z^H r
        }
        y[m] = s;
    }
    NS_SYNC();
    // back substitution with the unit upper triangular X, column oriented
    for (int j = n - 1; j > 0; --j) {
        const cdbl cj = y[j];
        MSP_PFOR(i, j) y[i] = csub(y[i], cmul(A[j * lda + i], cj));
        NS_SYNC();
    }
}

// o = (u^T Z) X.  t: team scratch (n); out o (n).
MSP_HD void modes_proj_team(int n, const cdbl *A, int lda, const cdbl *Z, int ldz, const cdbl *u,
                            cdbl *t, cdbl *o)
{
    MSP_PFOR(m, n)
    {
        cdbl s = cmk(0.0, 0.0);
        if (Z) {
#pragma unroll 8
            for (int i = 0; i < n; ++i) cfma(s, u[i], Z[i * ldz + m]);
        } else {
            s = u[m];                       // u is already u^T Z
        }
        t[m] = s;
    }
    NS_SYNC();
    MSP_PFOR(k, n)
    {
        cdbl s = t[k];
#pragma unroll 8
        for (int j = 0; j < k; ++j) cfma(s, t[j], A[k * lda + j]);
        o[k] = s;
    }
    NS_SYNC();
}

// V = Z X in place in Z (row by row, columns in descending order).
MSP_HD void modes_vectors_team(int n, const cdbl *A, int lda, cdbl *Z, int ldz)
{
    MSP_PFOR(i, n)
    {
        for (int k = n - 1; k >= 0; --k) {
            cdbl s = Z[i * ldz + k];
            for (int j = 0; j < k; ++j) cfma(s, Z[i * ldz + j], A[k * lda + j]);
            Z[i * ldz + k] = s;
        }
    }
    NS_SYNC();
}
