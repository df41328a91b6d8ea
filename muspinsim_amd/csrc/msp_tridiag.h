// Hermitian eigensolver for dimensions above the LDS-resident Jacobi:
//   (1) Householder reduction A = Q T Q^H to a real symmetric tridiagonal T,
//   (2) divide and conquer on T (Cuppen's split, Gu-Eisenstat eigenvectors),
//   (3) back-transformation U = Q Z.
// Replaces numpy.linalg.eigh (LAPACK zheevd) in Hermitian.diag
// (muspinsim/spinop.py:69) with the same output contract: eigenvalues
// ascending, orthonormal eigenvector columns.
//
// All three stages are team functions (see msp_team.h): one team owns one
// matrix; the matrix and the divide-and-conquer workspaces live in global
// memory, the back-transformation works on LDS-resident column chunks.
#pragma once
#include "msp_team.h"
#ifndef MSP_HOST_BUILD
#include "msp_mfma.h"
#endif

// ---------------------------------------------------------------------------
// (1) Householder tridiagonalisation, full Hermitian storage (row-major,
// row stride n).  Step k reflects column k below the diagonal:
//   v = [1, x(1:)/(alpha - beta)], beta = -sign(Re alpha) ||x||, tau = (beta - alpha)/beta
//   A22 <- H^H A22 H with H = I - tau v v^H, applied as the rank-2 update
//   A22 -= v w^H + w v^H, p = tau A22 v, w = p - (tau/2)(p^H v) v.
// Both triangles of A22 are kept, so p_i = sum_j conj(A22[j][i]) v_j reads
// columns: consecutive threads touch consecutive addresses.
// On exit d (n), e (n-1) hold T, tau (n-1) the reflector scalars and the
// strictly-lower part of column k (rows k+2..n-1) the tail of v_k.
// vbuf/wbuf: team-shared complex scratch of n entries each.
MSP_HD void tridiag_team(int n, cdbl *A, double *d, double *e, cdbl *tau,
                         cdbl *vbuf, cdbl *wbuf, double *red)
{
    for (int k = 0; k < n - 1; ++k) {
        const int m = n - k - 1;          // size of the trailing block
        cdbl *col = A + (size_t)(k + 1) * n + k;  // A[k+1][k], stride n
        // ||x(1:)||^2
        double part = 0.0;
        MSP_PFOR(i, m - 1) part += cabs2(col[(size_t)(i + 1) * n]);
        const double xn2 = team_sum(part, red);
        const cdbl alpha = col[0];
        double beta;
        cdbl tk;
        cdbl scale;
        if (xn2 == 0.0 && alpha.im == 0.0) {
            tk = cmk(0.0, 0.0);
            beta = alpha.re;
            scale = cmk(0.0, 0.0);
        } else {
            double nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn2);
            beta = alpha.re >= 0.0 ? -nrm : nrm;
            tk = cmk((beta - alpha.re) / beta, -alpha.im / beta);
            scale = cdivc(cmk(1.0, 0.0), csub(alpha, cmk(beta, 0.0)));
        }
        // v into vbuf and into the column (tail only)
        MSP_PFOR(i, m)
        {
            cdbl vi;
            if (i == 0) vi = cmk(1.0, 0.0);
            else {
                vi = cmul(col[(size_t)i * n], scale);
                col[(size_t)i * n] = vi;
            }
            vbuf[i] = vi;
        }
        MSP_SYNC();
        cdbl *A22 = A + (size_t)(k + 1) * n + (k + 1);
        if (tk.re != 0.0 || tk.im != 0.0) {
            // p = tau A22 v
            MSP_PFOR(i, m)
            {
                cdbl acc = cmk(0.0, 0.0);
                for (int j = 0; j < m; ++j) cfmac(acc, A22[(size_t)j * n + i], vbuf[j]);
                wbuf[i] = cmul(tk, acc);
            }
            MSP_SYNC();
            // alpha2 = -1/2 tau (p^H v)
            double pre = 0.0, pim = 0.0;
            MSP_PFOR(i, m)
            {
                cdbl t = cmulc(wbuf[i], vbuf[i]);
                pre += t.re;
                pim += t.im;
            }
            double sre = team_sum(pre, red);
            double sim = team_sum(pim, red);
            cdbl a2 = cmul(cmk(-0.5 * tk.re, -0.5 * tk.im), cmk(sre, sim));
            MSP_PFOR(i, m) wbuf[i] = cadd(wbuf[i], cmul(a2, vbuf[i]));
            MSP_SYNC();
            // A22 -= v w^H + w v^H
            MSP_PFOR(idx, m * m)
            {
                int i = idx / m, j = idx % m;
                cdbl a = A22[(size_t)i * n + j];
                cdbl vi = vbuf[i], wi = wbuf[i], vj = vbuf[j], wj = wbuf[j];
                a = csub(a, cmul(vi, cconj(wj)));
                a = csub(a, cmul(wi, cconj(vj)));
                A22[(size_t)i * n + j] = a;
            }
            MSP_SYNC();
        }
#ifdef MSP_HOST_BUILD
        d[k] = A[(size_t)k * n + k].re;
        e[k] = beta;
        tau[k] = tk;
#else
        if (MSP_TID == 0) {
            d[k] = A[(size_t)k * n + k].re;
            e[k] = beta;
            tau[k] = tk;
        }
#endif
        MSP_SYNC();
    }
#ifdef MSP_HOST_BUILD
    d[n - 1] = A[(size_t)(n - 1) * n + (n - 1)].re;
#else
    if (MSP_TID == 0) d[n - 1] = A[(size_t)(n - 1) * n + (n - 1)].re;
#endif
    MSP_SYNC();
}

// ---------------------------------------------------------------------------
// (2) Divide and conquer for the symmetric tridiagonal (d, e).
//
// The split tree halves [lo, hi) at mid = lo + (hi - lo)/2 down to single
// entries, so every off-diagonal e_i is a split coupling: first
// d_i -= |e_i| and d_{i+1} -= |e_i| for all i (Cuppen), then blocks are
// merged bottom-up, all merges of one tree level together.  A merge of
// T1 = Q1 D1 Q1^T and T2 = Q2 D2 Q2^T solves D + rho z z^T with
// z = [last row of Q1, sign(e) first row of Q2]/sqrt(2), rho = 2|e|:
//   - deflation of tiny rho|z_i| and of near-equal d pairs (Givens),
//   - secular equation 1 + rho sum z_i^2/(d_i - lambda) = 0 per root,
//     solved with the origin at the nearer pole by a two-pole rational
//     model (Li's "middle way") safeguarded by bisection,
//   - Loewner-recomputed z (Gu & Eisenstat) for orthogonal eigenvectors,
//   - block multiply Q_new = Q_old S.
// Q is n x n, row-major, column j = eigenvector j.

struct DcWork {
    double *Q0, *Q1, *S, *QK;  // n*n each (QK: gathered kept columns)
    double *dv, *zv;           // current eigenvalues / merge z (by column)
    double *ds, *zs;           // sorted d / z
    int *perm;                 // sorted slot -> column
    double *dl, *zl, *zh;      // kept (non-deflated) values, Loewner z
    int *kcol;                 // kept slot -> column
    int *korg;                 // origin pole per root
    double *ktau;              // root offset from its origin
    int *defl;                 // 1 = deflated (by sorted slot)
    int *nkept;                // per block (indexed by lo)
    int *rp, *rn;              // Givens pairs (columns), per block at lo..
    double *rc, *rs;
    int *nrot;
    int *dfl;                  // deflated sorted slots, per block at lo..
    double *item_v;            // merge items: kept roots then deflated slots
    int *item_src;             // root index j >= 0, or ~slot for a deflated one
    double *outv;              // merged eigenvalue per sorted output slot
    int *outsrc;               // item source of each output column
    int *rootpos;              // kept root j -> output column
    int *blo, *bhi;            // block [lo, hi) of every position at the current level
    int *plo, *phi;            // parent block (level L - 1; whole range at L = 0)
    int *sup;                  // rows a column may be nonzero in: 1 first child, 2 second, 3 both
    int *kperm;                // per block: product index k -> kept index, kept columns
                               // supported on the first child only, on both, on the second only
    int *nka, *kb0;            // per block at lo: rows of the first child sum k < nka,
                               // rows of the second k >= kb0
};

MSP_HD int dc_depth(int n)
{
    int dep = 0, s = n;
    while (s > 1) { s = s - s / 2; ++dep; }  // ceil halving
    return dep;
}

// Block [lo, hi) containing position i at tree level L (root = level 0).
MSP_HD void dc_block(int n, int L, int i, int &lo, int &hi)
{
    lo = 0;
    hi = n;
    for (int l = 0; l < L && hi - lo > 1; ++l) {
        int mid = lo + (hi - lo) / 2;
        if (i < mid) hi = mid; else lo = mid;
    }
}

// Block number bi (left to right) of tree level L, for blocks whose
// ancestors all have two or more positions.
MSP_HD void dc_block_index(int n, int L, int bi, int &lo, int &hi)
{
    lo = 0;
    hi = n;
    for (int l = 0; l < L && hi - lo > 1; ++l) {
        const int mid = lo + (hi - lo) / 2;
        if ((bi >> (L - 1 - l)) & 1) lo = mid; else hi = mid;
    }
}

// Largest block at tree level L (the rightmost path takes the ceiling half).
MSP_HD int dc_maxblock(int n, int L)
{
    int lo, hi;
    dc_block(n, L, n - 1, lo, hi);
    return hi - lo;
}

// 1/x: on the device the hardware reciprocal refined by two Newton steps
// r <- r + r (1 - x r) (within an ulp of the IEEE quotient for normal x,
// about a fifth of the instructions of a division); exact division on the host.
MSP_HD double msp_recip(double x)
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

MSP_HD double secular_root(int K, const double *dl, const double *zl, const double *zz,
                           double rho, int j, int &org_out)
{
    // Root j of 1 + rho sum_i zl_i^2/(dl_i - x) = 0, dl ascending, rho > 0;
    // zz_i = rho zl_i zl_i (same rounding as the product written out).
    const double eps = 2.220446049250313e-16;
    int org;
    double lo, hi;
    // the secular function at the midpoint of an interior gap, split into the
    // poles left (psi) and right (phi) of the root, picks the origin pole and
    // is the first iterate of the model steps below
    double m_psi = 0.0, m_dpsi = 0.0, m_phi = 0.0, m_dphi = 0.0, m_tau = 0.0;
    const bool interior = j < K - 1;
    if (interior) {
        double gap = dl[j + 1] - dl[j];
        double mid = 0.5 * gap;
        // (unrolled: the reciprocal chains of four poles overlap; the sums
        // keep their order)
#pragma unroll 4
        for (int i = 0; i < K; ++i) {
            const double r = msp_recip((dl[i] - dl[j]) - mid);
            const double t = zz[i] * r;
            if (i <= j) { m_psi += t; m_dpsi += t * r; }
            else { m_phi += t; m_dphi += t * r; }
        }
        double f = 1.0 + m_psi + m_phi;
        if (f > 0.0) { org = j; lo = 0.0; hi = mid; m_tau = mid; }
        else { org = j + 1; lo = -mid; hi = 0.0; m_tau = -mid; }
    } else {
        double z2 = 0.0;
        for (int i = 0; i < K; ++i) z2 += zl[i] * zl[i];
        org = K - 1;
        lo = 0.0;
        hi = rho * z2;
    }
    org_out = org;
    if (K == 1) return zz[0];
    const double dorg = dl[org];
    // the model keeps two poles exact: j and j+1 around an interior root, the
    // last two poles for the last root (right of both), the other poles of
    // each side lumped into one rational term
    const int pl = interior ? j : K - 2;
    const int pr = interior ? j + 1 : K - 1;
    double tau = interior ? m_tau : 0.5 * (lo + hi);
    for (int it = 0; it < 80; ++it) {
        // psi / phi: poles left / right of the root; with r = 1/del,
        // term = rho z^2 r and its derivative rho z^2 r^2 (the first interior
        // iterate reuses the midpoint evaluation)
        double psi = 0.0, dpsi = 0.0, phi = 0.0, dphi = 0.0;
        if (it == 0 && interior) {
            psi = m_psi;
            dpsi = m_dpsi;
            phi = m_phi;
            dphi = m_dphi;
        } else {
#pragma unroll 4
            for (int i = 0; i < K; ++i) {
                const double r = msp_recip((dl[i] - dorg) - tau);
                const double t = zz[i] * r;
                if (i <= pl) { psi += t; dpsi += t * r; }
                else { phi += t; dphi += t * r; }
            }
        }
        double w = 1.0 + psi + phi;
        // |w| within its rounding error: tau is the root to working precision
        // (a further model step would round onto the bracket and fall back to
        // bisection)
        if (fabs(w) <= 8.0 * eps * K * (1.0 + fabs(psi) + fabs(phi))) break;
        if (w < 0.0) lo = tau; else hi = tau;
        // model c + sl/(Dl - eta) + sr/(Dr - eta) = 0, i.e. c eta^2 - a eta
        // + b = 0 with a = c (Dl + Dr) + sl + sr, b = c Dl Dr + sl Dr + sr Dl
        // and c = w - sl/Dl - sr/Dr.  The first step keeps the two poles'
        // own weights (sl, sr = rho z^2; every other pole goes into the
        // constant), which lands near a root that sits close to a pole of
        // small weight; later steps match each side's slope at tau (sl =
        // Dl^2 psi', sr = Dr^2 phi').
        double eta;
        const double Dl = (dl[pl] - dorg) - tau;
        const double Dr = (dl[pr] - dorg) - tau;
        double sl = Dl * Dl * dpsi, sr = Dr * Dr * dphi;
        double c = w - Dl * dpsi - Dr * dphi;
        if (it == 0) {
            const double cf = w - zz[pl] / Dl - zz[pr] / Dr;
            if (interior || cf > 0.0) {
                sl = zz[pl];
                sr = zz[pr];
                c = cf;
            }
        }
        const double a = c * (Dl + Dr) + sl + sr;
        const double b = c * Dl * Dr + sl * Dr + sr * Dl;
        if (interior) {
            // the root between the two poles
            const double disc = sqrt(fabs(a * a - 4.0 * b * c));
            if (c == 0.0) eta = b / a;
            else if (a <= 0.0) eta = (a - disc) / (2.0 * c);
            else eta = 2.0 * b / (a + disc);
        } else {
            // the root right of both poles: f -> 1 > 0 there, so the lumped
            // constant is taken positive and the larger root is used; a step
            // against the sign of w falls back to Newton's
            c = fabs(c);
            const double disc = sqrt(fabs(a * a - 4.0 * b * c));
            if (c == 0.0) eta = b / a;
            else if (a >= 0.0) eta = (a + disc) / (2.0 * c);
            else eta = 2.0 * b / (a - disc);
            if (w * eta > 0.0) eta = -w / (dpsi + dphi);
        }
        double tn = tau + eta;
        bool inside = interior ? (tn > lo && tn < hi) : (tn > lo && tn <= hi);
        if (!inside) tn = 0.5 * (lo + hi);
        double step = fabs(tn - tau);
        tau = tn;
        double tol = 4.0 * eps * fmax(fabs(tau), 1e-300);
        if (step <= tol || hi - lo <= tol) break;
    }
    return tau;
}

// Qn[lo+r][rootpos[lo+j]] = sum_i Q[lo+r][kcol[lo+i]] S[lo+i][lo+j] for every
// merge block of level L (the kept columns are read through kcol in place;
// Q and Qn hold eigenvector columns contiguously, element (row, col) at
// col * n + row).  A kept column is nonzero only on its support, so a row
// of the first child sums the product indices k < nka and a row of the
// second k >= kb0 (kperm orders them); without rotations that halves the
// work.  Large blocks (>= kMergeMfmaMin) run on FP64 matrix cores as Qn^T = S^T Q_K^T,
// each wave computing 32 x 32 output tiles from operands read in place,
// every operand and result access along a contiguous 16-run; small ones are
// direct dots.
#ifndef MSP_MERGE_MFMA_MIN
#define MSP_MERGE_MFMA_MIN 12
#endif
constexpr int kMergeMfmaMin = MSP_MERGE_MFMA_MIN;   // smallest merge block on MFMA

MSP_HD void dc_merge_gemm(int n, int L, const DcWork &W, const double *Q, double *Qn, double *red)
{
    // small blocks: one output per iteration, (column, row in its block) items
    const int smax = dc_maxblock(n, L);
    int sh = 0;
    while ((1 << sh) < smax) ++sh;
    MSP_PFOR(idx, n << sh)
    {
        const int col = idx >> sh;
        const int lo = W.blo[col], hi = W.bhi[col];
        const int row = lo + (idx & ((1 << sh) - 1));
        if (hi - lo >= 2 && hi - lo < kMergeMfmaMin && row < hi) {
            int K = W.nkept[lo];
            int j = col - lo;
            if (j < K) {
                const bool first = row < lo + (hi - lo) / 2;
                const int kb = first ? 0 : W.kb0[lo], ke = first ? W.nka[lo] : K;
                double acc = 0.0;
                for (int k = kb; k < ke; ++k) {
                    const int i = W.kperm[lo + k];
                    acc += Q[(size_t)W.kcol[lo + i] * n + row] * W.S[(size_t)(lo + i) * n + lo + j];
                }
                Qn[(size_t)W.rootpos[lo + j] * n + row] = acc;
            }
        }
    }
    // large blocks
#ifdef MSP_HOST_BUILD
    (void)red;
    for (int lo = 0; lo < n;) {
        const int blo = W.blo[lo], bhi = W.bhi[lo];
        const int sz = bhi - blo;
        if (sz >= kMergeMfmaMin) {
            const int K = W.nkept[blo];
            for (int r = 0; r < sz; ++r) {
                const bool first = r < sz / 2;
                const int kb = first ? 0 : W.kb0[blo], ke = first ? W.nka[blo] : K;
                for (int j = 0; j < K; ++j) {
                    double acc = 0.0;
                    for (int k = kb; k < ke; ++k) {
                        const int i = W.kperm[blo + k];
                        acc += Q[(size_t)W.kcol[blo + i] * n + blo + r] * W.S[(size_t)(blo + i) * n + blo + j];
                    }
                    Qn[(size_t)W.rootpos[blo + j] * n + blo + r] = acc;
                }
            }
        }
        lo = (bhi > lo) ? bhi : lo + 1;
    }
#else
    {
        // 32 x 32 tiles of Qn^T = S^T Q_K^T (2 x 2 MFMA tiles of 16 x 16)
        // per wave task: rows j = kept roots, columns r = block rows.
        // A operand S^T[j][k] = S[k][j] (16 consecutive j of a row of S),
        // B operand Q_K^T[k][r] = Q(blo + r, kcol[k]) (16 consecutive r of
        // an eigenvector column); no LDS staging and no barriers
        const int lane = MSP_TID & 63, wave = MSP_TID >> 6, nwaves = MSP_NT >> 6;
        const int li = lane & 15, lq = lane >> 4;
        // the tasks of every large block of the level share the waves
        const int tside = (smax + 31) / 32, tmax = tside * tside;
        const int ntask = (smax >= kMergeMfmaMin) ? (1 << L) * tmax : 0;
        for (int task = wave; task < ntask; task += nwaves) {
            const int bi = task / tmax, tt = task - bi * tmax;
            int blo, bhi;
            dc_block_index(n, L, bi, blo, bhi);
            const int sz = bhi - blo;
            if (sz < kMergeMfmaMin) continue;
            const int K = W.nkept[blo];
            const int tj = (K + 31) / 32, tr = (sz + 31) / 32;
            if (tt >= tj * tr) continue;
            const int midl = sz / 2, nka = W.nka[blo], kb0 = W.kb0[blo];
            const int jt = (tt / tr) * 32, rt = (tt % tr) * 32;
            // product range of this row tile: one child's rows sum their
            // supported columns; a tile across the middle sums all with a
            // support test per element
            const int rend = (rt + 32 < sz) ? rt + 32 : sz;
            const bool allA = rend <= midl, allB = rt >= midl;
            const int kb = allB ? kb0 : 0, ke = allA ? nka : K;
            msp_d4 acc[2][2];
            for (int x = 0; x < 2; ++x)
                for (int y2 = 0; y2 < 2; ++y2) acc[x][y2] = d4zero();
            for (int k0 = kb; k0 < ke; k0 += 4) {
                const int kk = k0 + lq;
                const bool kok = kk < ke;
                const int ki = W.kperm[blo + (kok ? kk : ke - 1)];
                const int qc = W.kcol[blo + ki];
                const int sp = (allA || allB) ? 3 : W.sup[qc];
                const double *Qk = Q + (size_t)qc * n + blo;
                const double *Sk = W.S + (size_t)(blo + ki) * n + blo;
                double a[2], bv[2];
#pragma unroll
                for (int x = 0; x < 2; ++x) {
                    const int j = jt + 16 * x + li;
                    const int jc = (j < K) ? j : K - 1;
                    const double va = Sk[jc];
                    a[x] = (j < K && kok) ? va : 0.0;
                    const int r = rt + 16 * x + li;
                    const int rc = (r < sz) ? r : sz - 1;
                    const double vb = Qk[rc];
                    const bool on = sp & ((r < midl) ? 1 : 2);
                    bv[x] = (r < sz && kok && on) ? vb : 0.0;
                }
#pragma unroll
                for (int x = 0; x < 2; ++x)
#pragma unroll
                    for (int y2 = 0; y2 < 2; ++y2) acc[x][y2] = mfma_f64(a[x], bv[y2], acc[x][y2]);
            }
            for (int x = 0; x < 2; ++x)
                for (int y2 = 0; y2 < 2; ++y2)
                    for (int q = 0; q < 4; ++q) {
                        const int j = jt + 16 * x + 4 * q + lq;
                        const int r = rt + 16 * y2 + li;
                        if (j < K && r < sz) Qn[(size_t)W.rootpos[blo + j] * n + blo + r] = acc[x][y2][q];
                    }
        }
        (void)red;
    }
#endif
    MSP_SYNC();
}

// d (n): diagonal, overwritten by ascending eigenvalues; e (n-1): off-diagonal.
// Q: on exit eigenvectors, one contiguous column each: component r of
// eigenvector c at Q[c * n + r] (points to W.Q0 or W.Q1, returned).
MSP_HD double *stedc_team(int n, double *d, const double *e, DcWork W, double *red)
{
    const double eps = 2.220446049250313e-16;
    // scale to unit max norm
    double part = 0.0;
    MSP_PFOR(i, n) part = fmax(part, fabs(d[i]));
    MSP_PFOR(i, n - 1) part = fmax(part, fabs(e[i]));
    double scl = team_max(part, red);
    if (scl == 0.0) scl = 1.0;
    const double inv = 1.0 / scl;
    MSP_PFOR(i, n)
    {
        double di = d[i] * inv;
        if (i > 0) di -= fabs(e[i - 1] * inv);
        if (i < n - 1) di -= fabs(e[i] * inv);
        W.dv[i] = di;
    }
    double *Q = W.Q0, *Qn = W.Q1;
    MSP_PFOR(idx, n * n) Q[idx] = (idx / n == idx % n) ? 1.0 : 0.0;
    MSP_SYNC();

    const int depth = dc_depth(n);
    for (int L = depth - 1; L >= 0; --L) {
        // block table of this level and its parent level
        MSP_PFOR(i, n)
        {
            int lo, hi;
            dc_block(n, L, i, lo, hi);
            W.blo[i] = lo;
            W.bhi[i] = hi;
            if (L > 0) dc_block(n, L - 1, i, lo, hi);
            W.plo[i] = lo;
            W.phi[i] = hi;
        }
        MSP_SYNC();
        // --- z and sorted order per block (by position) ---
        MSP_PFOR(i, n)
        {
            const int lo = W.blo[i], hi = W.bhi[i];
            if (hi - lo >= 2) {
                int mid = lo + (hi - lo) / 2;
                double ec = e[mid - 1] * inv;
                double zi;
                if (i < mid) zi = Q[(size_t)i * n + mid - 1];
                else zi = (ec < 0.0 ? -1.0 : 1.0) * Q[(size_t)i * n + mid];
                W.zv[i] = zi * 0.7071067811865476;
                W.sup[i] = (i < mid) ? 1 : 2;      // support: the column's child
                // rank within the merged (two sorted halves) order
                double di = W.dv[i];
                int r;
                if (i < mid) {
                    int c = 0; // count in right half strictly less
                    for (int k = mid; k < hi; ++k) c += (W.dv[k] < di);
                    r = (i - lo) + c;
                } else {
                    int c = 0; // count in left half less or equal
                    for (int k = lo; k < mid; ++k) c += (W.dv[k] <= di);
                    r = (i - mid) + c;
                }
                W.perm[lo + r] = i;
            }
        }
        MSP_SYNC();
        MSP_PFOR(s, n)
        {
            const int lo = W.blo[s], hi = W.bhi[s];
            if (hi - lo >= 2) {
                W.ds[s] = W.dv[W.perm[s]];
                W.zs[s] = W.zv[W.perm[s]];
            }
        }
        MSP_SYNC();
        // --- deflation tolerance, one thread per block (kept in item_v[lo],
        // which is not read before the merge items below) ---
        MSP_PFOR(lo0, n)
        {
            const int lo = W.blo[lo0], hi = W.bhi[lo0];
            if (lo == lo0 && hi - lo >= 2) {
                double dmax = 0.0, zmax = 0.0;
                for (int s = lo; s < hi; ++s) {
                    dmax = fmax(dmax, fabs(W.ds[s]));
                    zmax = fmax(zmax, fabs(W.zs[s]));
                }
                W.item_v[lo] = 8.0 * eps * fmax(dmax, zmax);
                W.item_src[lo] = hi;      // first rotation slot: none yet
                W.nkept[lo] = 0;
                W.nrot[lo] = 0;
            }
        }
        MSP_SYNC();
        // --- per slot: z deflation (defl) and whether the slot would rotate
        // with the previous undeflated slot (korg, rewritten by the secular
        // step), both as the scan below tests them while no rotation has
        // happened yet in the block ---
        MSP_PFOR(s, n)
        {
            const int lo = W.blo[s], hi = W.bhi[s];
            if (hi - lo >= 2) {
                const int mid = lo + (hi - lo) / 2;
                const double rho = 2.0 * fabs(e[mid - 1] * inv);
                const double tol = W.item_v[lo];
                const int zd = (rho * fabs(W.zs[s]) <= tol) ? 1 : 0;
                int rot = 0;
                if (!zd) {
                    int pj = s - 1;
                    while (pj >= lo && rho * fabs(W.zs[pj]) <= tol) --pj;
                    if (pj >= lo) {
                        double zs_ = W.zs[pj], zc = W.zs[s];
                        double tu = sqrt(zs_ * zs_ + zc * zc);
                        double c = zc / tu, sn = -zs_ / tu;
                        double t = W.ds[s] - W.ds[pj];
                        rot = (fabs(t * c * sn) <= tol) ? 1 : 0;
                    }
                }
                W.defl[s] = zd;
                W.korg[s] = rot;
            }
        }
        MSP_SYNC();
        // --- per slot, up to the block's first rotation: z-deflated slots
        // to dfl, kept ones to dl/zl/kcol, at the positions the scan gives
        // them (counts of the deflated / undeflated slots before); the slot
        // that pairs with the first rotation leaves the scan state in
        // item_src[lo] (start), outsrc[lo] (K), rootpos[lo] (nd) and
        // kperm[lo] (pj), all rewritten later in the level ---
        MSP_PFOR(s, n)
        {
            const int lo = W.blo[s], hi = W.bhi[s];
            if (hi - lo >= 2) {
                int nz = 0, nk = 0;
                bool clean = true;
                for (int t = lo; t < s; ++t) {
                    const int dt = W.defl[t];
                    clean = clean && (dt || !W.korg[t]);
                    nz += dt;
                    nk += 1 - dt;
                }
                if (clean && (W.defl[s] || !W.korg[s])) {
                    if (W.defl[s]) {
                        W.dfl[lo + nz] = s;
                    } else {
                        int t = s + 1;
                        while (t < hi && W.defl[t]) ++t;
                        if (t == hi || !W.korg[t]) {
                            W.dl[lo + nk] = W.ds[s];
                            W.zl[lo + nk] = W.zs[s];
                            W.kcol[lo + nk] = W.perm[s];
                            if (t == hi) W.nkept[lo] = nk + 1;
                        } else {
                            W.item_src[lo] = t;
                            W.outsrc[lo] = nk;
                            W.rootpos[lo] = nz + (t - s - 1);
                            W.kperm[lo] = s;
                        }
                    }
                }
            }
        }
        MSP_SYNC();
        // --- deflation scan from the first rotation on, one thread per
        // block (blocks without a rotation are complete) ---
        MSP_PFOR(lo0, n)
        {
            const int lo = W.blo[lo0], hi = W.bhi[lo0];
            if (lo == lo0 && hi - lo >= 2 && W.item_src[lo] < hi) {
                int mid = lo + (hi - lo) / 2;
                double rho = 2.0 * fabs(e[mid - 1] * inv);
                const double tol = W.item_v[lo];
                const int s0 = W.item_src[lo];
                int K = W.outsrc[lo], nr = 0, nd = W.rootpos[lo], pj = W.kperm[lo];
                for (int s = s0; s < hi; ++s) {
                    W.defl[s] = 0;
                    if (rho * fabs(W.zs[s]) <= tol) {
                        W.defl[s] = 1;
                        W.dfl[lo + nd++] = s;
                        continue;
                    }
                    if (pj < 0) { pj = s; continue; }
                    double zs_ = W.zs[pj], zc = W.zs[s];
                    double tu = sqrt(zs_ * zs_ + zc * zc);
                    double c = zc / tu, sn = -zs_ / tu;
                    double t = W.ds[s] - W.ds[pj];
                    if (fabs(t * c * sn) <= tol) {
                        // rotate columns perm[pj], perm[s]; pj deflates
                        W.zs[s] = tu;
                        W.zs[pj] = 0.0;
                        double dp = W.ds[pj], dn = W.ds[s];
                        W.ds[pj] = dp * c * c + dn * sn * sn;
                        W.ds[s] = dp * sn * sn + dn * c * c;
                        W.defl[pj] = 1;
                        W.dfl[lo + nd++] = pj;
                        W.rp[lo + nr] = W.perm[pj];
                        W.rn[lo + nr] = W.perm[s];
                        W.rc[lo + nr] = c;
                        W.rs[lo + nr] = sn;
                        ++nr;
                        pj = s;
                    } else {
                        W.dl[lo + K] = W.ds[pj];
                        W.zl[lo + K] = W.zs[pj];
                        W.kcol[lo + K] = W.perm[pj];
                        ++K;
                        pj = s;
                    }
                }
                if (pj >= 0) {
                    W.dl[lo + K] = W.ds[pj];
                    W.zl[lo + K] = W.zs[pj];
                    W.kcol[lo + K] = W.perm[pj];
                    ++K;
                }
                W.nkept[lo] = K;
                W.nrot[lo] = nr;
                // a rotated column is supported on both children (its other
                // half is zeroed before the rotations)
                for (int r = 0; r < nr; ++r) {
                    W.sup[W.rp[lo + r]] = 3;
                    W.sup[W.rn[lo + r]] = 3;
                }
            }
        }
        MSP_SYNC();
        // --- apply the deflation rotations to Q rows (in scan order) ---
        // A column holds valid entries on its child's rows only; a rotated
        // one first gets zeros on the other child's rows.
        MSP_PFOR(row, n)
        {
            const int lo = W.blo[row], hi = W.bhi[row];
            if (hi - lo >= 2) {
                int nr = W.nrot[lo];
                const int mid = lo + (hi - lo) / 2;
                const bool first = row < mid;
                for (int r = 0; r < nr; ++r) {
                    const int a = W.rp[lo + r], b = W.rn[lo + r];
                    if ((a < mid) != first) Q[(size_t)a * n + row] = 0.0;
                    if ((b < mid) != first) Q[(size_t)b * n + row] = 0.0;
                }
                for (int r = 0; r < nr; ++r) {
                    int a = W.rp[lo + r], b = W.rn[lo + r];
                    double c = W.rc[lo + r], s = W.rs[lo + r];
                    double qa = Q[(size_t)a * n + row], qb = Q[(size_t)b * n + row];
                    Q[(size_t)a * n + row] = c * qa + s * qb;
                    Q[(size_t)b * n + row] = c * qb - s * qa;
                }
                // rho z^2 of the kept poles for the secular solves (zh is
                // written with the Loewner z only after them)
                if (row - lo < W.nkept[lo]) {
                    const double rho = 2.0 * fabs(e[mid - 1] * inv);
                    W.zh[row] = rho * W.zl[row] * W.zl[row];
                }
            }
        }
        MSP_SYNC();
        // --- secular roots ---
        MSP_PFOR(s, n)
        {
            const int lo = W.blo[s], hi = W.bhi[s];
            if (hi - lo >= 2) {
                int mid = lo + (hi - lo) / 2;
                double rho = 2.0 * fabs(e[mid - 1] * inv);
                int K = W.nkept[lo];
                int j = s - lo;
                if (j < K) {
                    int org;
                    double t = secular_root(K, W.dl + lo, W.zl + lo, W.zh + lo, rho, j, org);
                    W.korg[s] = org;
                    W.ktau[s] = t;
                    // product order of the kept columns: first child only,
                    // both, second child only, so that each child's rows sum
                    // a range of it
                    const int cls = (W.sup[W.kcol[s]] == 1) ? 0 : (W.sup[W.kcol[s]] == 3) ? 1 : 2;
                    int pos = 0, c1 = 0, c3 = 0;
                    for (int i = 0; i < K; ++i) {
                        const int sp = W.sup[W.kcol[lo + i]];
                        const int ci = (sp == 1) ? 0 : (sp == 3) ? 1 : 2;
                        pos += (ci < cls) || (ci == cls && i < j);
                        c1 += (ci == 0);
                        c3 += (ci == 1);
                    }
                    W.kperm[lo + pos] = j;
                    if (j == 0) {
                        W.nka[lo] = c1 + c3;
                        W.kb0[lo] = c1;
                    }
                }
            }
        }
        MSP_SYNC();
        // --- Loewner z ---
        MSP_PFOR(s, n)
        {
            const int lo = W.blo[s], hi = W.bhi[s];
            if (hi - lo >= 2) {
                int K = W.nkept[lo];
                int i = s - lo;
                if (i < K) {
                    const double *dl = W.dl + lo;
                    double di = dl[i];
                    double wv = 1.0;
#pragma unroll 4
                    for (int j = 0; j < K; ++j) {
                        int org = W.korg[lo + j];
                        double del = (di - dl[org]) - W.ktau[lo + j];  // d_i - lambda_j
                        if (j == i) wv *= del;
                        else wv *= del * msp_recip(di - dl[j]);
                    }
                    double zz = sqrt(fabs(wv));
                    W.zh[s] = (W.zl[s] < 0.0) ? -zz : zz;
                }
            }
        }
        MSP_SYNC();
        // --- eigenvectors of the rank-one problem: S[i][j] (block-local square) ---
        MSP_PFOR(s, n)
        {
            const int lo = W.blo[s], hi = W.bhi[s];
            if (hi - lo >= 2) {
                int K = W.nkept[lo];
                int j = s - lo;
                if (j < K) {
                    const double *dl = W.dl + lo;
                    int org = W.korg[s];
                    double tj = W.ktau[s];
                    // the norm first, then each entry once (recomputed, not
                    // read back)
                    double nrm = 0.0;
#pragma unroll 4
                    for (int i = 0; i < K; ++i) {
                        const double u = W.zh[lo + i] * msp_recip((dl[i] - dl[org]) - tj);
                        nrm += u * u;
                    }
                    nrm = 1.0 / sqrt(nrm);
#pragma unroll 4
                    for (int i = 0; i < K; ++i)
                        W.S[(size_t)(lo + i) * n + (lo + j)] =
                            W.zh[lo + i] * nrm * msp_recip((dl[i] - dl[org]) - tj);
                }
            }
        }
        MSP_SYNC();
        // --- merge items (kept roots, then deflated slots) ---
        MSP_PFOR(s, n)
        {
            const int lo = W.blo[s], hi = W.bhi[s];
            if (hi - lo >= 2) {
                int K = W.nkept[lo];
                int t = s - lo;
                if (t < K) {
                    W.item_v[s] = W.dl[lo + W.korg[s]] + W.ktau[s];
                    W.item_src[s] = t;
                } else {
                    int sl = W.dfl[lo + (t - K)];
                    W.item_v[s] = W.ds[sl];
                    W.item_src[s] = ~sl;
                }
            }
        }
        MSP_SYNC();
        // --- ascending order of the merged eigenvalues, ties by item index ---
        MSP_PFOR(s, n)
        {
            const int lo = W.blo[s], hi = W.bhi[s];
            if (hi - lo >= 2) {
                double v = W.item_v[s];
                int r = 0;
                for (int t = lo; t < hi; ++t) {
                    double vt = W.item_v[t];
                    r += (vt < v || (vt == v && t < s));
                }
                W.outv[lo + r] = v;
                W.outsrc[lo + r] = W.item_src[s];
                if (W.item_src[s] >= 0) W.rootpos[lo + W.item_src[s]] = lo + r;
            }
        }
        MSP_SYNC();
        // --- new Q columns ---
        // Q is block diagonal, and only the rows of a column's block are
        // written: deflated columns (on their support, zero on the other
        // child's rows) and untouched blocks are copied here; kept roots are
        // the products Q[:, kcol] S below.  Rows of the parent block outside
        // the column's block keep stale values, which the next level never
        // reads (its supports).
        // (items are (column, row lane) pairs with at most 64 lanes per
        // column, each lane striding over its block's rows, so a kept
        // column costs one item per lane at every level)
        {
            const int smax = dc_maxblock(n, L);
            int sh = 0;
            while ((1 << sh) < smax) ++sh;
            const int cs = (sh > 6) ? 6 : sh;
            MSP_PFOR(idx, n << cs)
            {
                const int col = idx >> cs;
                const int lo = W.blo[col], hi = W.bhi[col];
                const int r0 = lo + (idx & ((1 << cs) - 1));
                if (hi - lo >= 2) {
                    const int src = W.outsrc[col];
                    if (src < 0) {
                        // four rows per lane in flight: loads first, then stores
                        const int c = W.perm[~src];
                        const int mid = lo + (hi - lo) / 2, sp = W.sup[c];
                        for (int rb = r0; rb < hi; rb += 4 << cs) {
                            double v[4];
#pragma unroll
                            for (int u = 0; u < 4; ++u) {
                                const int row = rb + (u << cs);
                                const int rc = (row < hi) ? row : hi - 1;
                                const int sd = (rc < mid) ? 1 : 2;
                                v[u] = (sp & sd) ? Q[(size_t)c * n + rc] : 0.0;
                            }
#pragma unroll
                            for (int u = 0; u < 4; ++u) {
                                const int row = rb + (u << cs);
                                if (row < hi) Qn[(size_t)col * n + row] = v[u];
                            }
                        }
                    }
                } else {
                    for (int row = r0; row < hi; row += 1 << cs)
                        Qn[(size_t)col * n + row] = Q[(size_t)col * n + row];
                }
            }
        }
        // no barrier: the products read Q and write only the kept-root
        // columns of Qn, which the copy above leaves alone
        dc_merge_gemm(n, L, W, Q, Qn, red);
        MSP_PFOR(s, n)
        {
            const int lo = W.blo[s], hi = W.bhi[s];
            if (hi - lo >= 2) W.dv[s] = W.outv[s];
        }
        MSP_SYNC();
        double *t = Q;
        Q = Qn;
        Qn = t;
    }
    MSP_PFOR(i, n) d[i] = W.dv[i] * scl;
    MSP_SYNC();
    return Q;
}

// ---------------------------------------------------------------------------
// (3) U = Q Z, Q = H(0) H(1) ... H(n-2), one column chunk [c0, c0+cw) held
// in LDS (Uc: n x cw complex, row stride cw).  Z is real row-major; U complex
// row-major (row stride n).  A holds the reflector tails (tridiag_team).
// Chunks are independent, so the device runs one workgroup per chunk.
MSP_HD void backtransform_chunk_team(int n, const cdbl *A, const cdbl *tau, const double *Z,
                                     cdbl *U, cdbl *Uc, int cw, int c0, cdbl *wsum)
{
    const int cwk = (n - c0 < cw) ? n - c0 : cw;
    MSP_PFOR(idx, n * cwk)
    {
        int i = idx / cwk, c = idx % cwk;
        Uc[i * cw + c] = cmk(Z[(size_t)(c0 + c) * n + i], 0.0);
    }
    MSP_SYNC();
    for (int k = n - 2; k >= 0; --k) {
        const cdbl tk = tau[k];
        if (tk.re == 0.0 && tk.im == 0.0) continue;
        const int m = n - k - 1;
        const cdbl *col = A + (size_t)(k + 1) * n + k;
        // w_c = tau sum_i conj(v_i) U[k+1+i][c]
        MSP_PFOR(c, cwk)
        {
            cdbl acc = Uc[(k + 1) * cw + c];
            for (int i = 1; i < m; ++i) cfmac(acc, col[(size_t)i * n], Uc[(k + 1 + i) * cw + c]);
            wsum[c] = cmul(tk, acc);
        }
        MSP_SYNC();
        MSP_PFOR(idx, m * cwk)
        {
            int i = idx / cwk, c = idx % cwk;
            cdbl vi = (i == 0) ? cmk(1.0, 0.0) : col[(size_t)i * n];
            cdbl u = Uc[(k + 1 + i) * cw + c];
            Uc[(k + 1 + i) * cw + c] = csub(u, cmul(vi, wsum[c]));
        }
        MSP_SYNC();
    }
    MSP_PFOR(idx, n * cwk)
    {
        int i = idx / cwk, c = idx % cwk;
        U[(size_t)i * n + c0 + c] = Uc[i * cw + c];
    }
    MSP_SYNC();
}

MSP_HD void backtransform_team(int n, const cdbl *A, const cdbl *tau, const double *Z,
                               cdbl *U, cdbl *Uc, int cw, cdbl *wsum)
{
    for (int c0 = 0; c0 < n; c0 += cw) backtransform_chunk_team(n, A, tau, Z, U, Uc, cw, c0, wsum);
}

// Reflector block size of the blocked back-transformation.
#define MSP_BT_NB 16
// Panel width of the blocked tridiagonalisation.
#define MSP_TD_NB 16

MSP_HD int bt_nblocks(int n) { return (n - 1 + MSP_BT_NB - 1) / MSP_BT_NB; }

// Doubles of per-matrix global workspace used by the divide-and-conquer path:
// e (n), tau (2n), DcWork doubles (4n^2 + 12n) and ints (20n, padded).
MSP_HD size_t dc_work_doubles(int n)
{
    return (size_t)4 * n * n + 15 * (size_t)n + 11 * (size_t)n + 8 +
           2 * (size_t)(bt_nblocks(n) + 1) * MSP_BT_NB * MSP_BT_NB + 4 * (size_t)n * MSP_TD_NB +
           2 * (size_t)bt_nblocks(n) * MSP_BT_NB * n;
}

// Chase reflectors of the two-stage reduction: sweep j (0 <= j <= n - 2) has
// steps s = 0 and s >= 1 while j + 1 + 16 s <= n - 2; each stores 16 complex
// entries of v and tau at index j * chase_steps(n) + s.
MSP_HD int chase_steps(int n) { return 1 + (n >= 3 ? (n - 3) / 16 : 0); }
MSP_HD size_t chase_doubles(int n)
{
    return (size_t)2 * (n > 1 ? n - 1 : 1) * chase_steps(n) * 17;
}

// Carve a DcWork (and e, tau) out of one contiguous per-matrix block.
MSP_HD DcWork dc_carve(int n, double *base, double **e, cdbl **tau)
{
    DcWork W;
    double *p = base;
    *e = p; p += n;
    *tau = (cdbl *)p; p += 2 * n;
    W.Q0 = p; p += (size_t)n * n;
    W.Q1 = p; p += (size_t)n * n;
    W.S = p; p += (size_t)n * n;
    W.QK = p; p += (size_t)n * n;
    W.dv = p; p += n; W.zv = p; p += n; W.ds = p; p += n; W.zs = p; p += n;
    W.dl = p; p += n; W.zl = p; p += n; W.zh = p; p += n; W.ktau = p; p += n;
    W.rc = p; p += n; W.rs = p; p += n; W.item_v = p; p += n; W.outv = p; p += n;
    int *q = (int *)p;
    W.perm = q; q += n; W.kcol = q; q += n; W.korg = q; q += n; W.defl = q; q += n;
    W.nkept = q; q += n; W.rp = q; q += n; W.rn = q; q += n; W.nrot = q; q += n;
    W.dfl = q; q += n; W.item_src = q; q += n; W.outsrc = q; q += n; W.rootpos = q; q += n;
    W.blo = q; q += n; W.bhi = q; q += n; W.plo = q; q += n; W.phi = q; q += n;
    W.sup = q; q += n; W.kperm = q; q += n; W.nka = q; q += n; W.kb0 = q; q += n;
    return W;
}

// T factors of the blocked back-transformation (after the DcWork arrays).
MSP_HD cdbl *dc_tblocks(int n, double *base)
{
    return (cdbl *)(base + (size_t)4 * n * n + 15 * (size_t)n + 11 * (size_t)n + 8);
}

// Panel buffers Vp, Wp (n x MSP_TD_NB complex each) of the blocked reduction.
MSP_HD cdbl *dc_panels(int n, double *base)
{
    return dc_tblocks(n, base) + (size_t)(bt_nblocks(n) + 1) * MSP_BT_NB * MSP_BT_NB;
}

// (V T) of every reflector block, column by column: element (r, c) of block
// blk at [(blk * MSP_BT_NB + c) * n + r].
MSP_HD cdbl *dc_vt(int n, double *base)
{
    return dc_panels(n, base) + (size_t)2 * n * MSP_TD_NB;
}

// Direct offsets into the per-matrix workspace (same layout as dc_carve), for
// kernels that need only a few of the arrays.
MSP_HD cdbl *dc_tau_ptr(int n, double *base) { return (cdbl *)(base + n); }
// Complex n x n (column-major) over the S and QK arrays, free once the divide
// and conquer has finished: Z' = Q2 Z of the two-stage reduction.
MSP_HD cdbl *dc_zc_ptr(int n, double *base) { return (cdbl *)(base + 3 * (size_t)n + 2 * (size_t)n * n); }
MSP_HD double *dc_result_ptr(int n, double *base)
{
    return base + 3 * (size_t)n + ((dc_depth(n) & 1) ? (size_t)n * n : 0);
}

// The eigenvector matrix stedc_team leaves after dc_depth(n) buffer swaps.
MSP_HD double *dc_result(int n, const DcWork &W)
{
    return (dc_depth(n) & 1) ? W.Q1 : W.Q0;
}

// ---------------------------------------------------------------------------
// Blocked back-transformation.  Reflectors k0 .. k0+nb-1 combine into
// H(k0) ... H(k0+nb-1) = I - V T V^H with V the (rows k0+1..n-1) x nb
// unit-lower-trapezoidal matrix of reflector vectors and T upper triangular:
//   T_jj = tau_j,  T(0:j, j) = -tau_j T(0:j, 0:j) (V(:, 0:j)^H v_j).
// Tb (nb x nb complex, row-major) for block b is built by larft_block_team;
// backtransform_blocked_team then applies U <- (I - V T V^H) U blockwise,
// last block first, on an LDS column chunk of U.

// V(r, j) of block k0 (absolute row r), from the reflector tails in A.
MSP_HD cdbl vblock(int n, const cdbl *A, int k0, int r, int j)
{
    const int k = k0 + j;
    if (r < k + 1) return cmk(0.0, 0.0);
    if (r == k + 1) return cmk(1.0, 0.0);
    return A[(size_t)r * n + k];
}

// T for reflectors [k0, k0 + nbk); G is team scratch of nb*nb complex.
MSP_HD void larft_block_team(int n, const cdbl *A, const cdbl *tau, int k0, int nbk, int nb,
                             cdbl *T, cdbl *G)
{
    // Gram G(i, j) = V(:, i)^H V(:, j), i < j
    MSP_PFOR(idx, nbk * nbk)
    {
        int i = idx / nbk, j = idx % nbk;
        cdbl acc = cmk(0.0, 0.0);
        if (i < j) {
            for (int r = k0 + j + 1; r < n; ++r)
                cfmac(acc, vblock(n, A, k0, r, i), vblock(n, A, k0, r, j));
        }
        G[i * nb + j] = acc;
    }
    MSP_PFOR(idx, nb * nb) T[idx] = cmk(0.0, 0.0);
    MSP_SYNC();
    for (int j = 0; j < nbk; ++j) {
        const cdbl tj = tau[k0 + j];
        // T(0:j, j) = -tau_j T(0:j, 0:j) G(0:j, j)
        MSP_PFOR(i, j + 1)
        {
            if (i == j) {
                T[j * nb + j] = tj;
            } else {
                cdbl acc = cmk(0.0, 0.0);
                for (int l = i; l < j; ++l) cfma(acc, T[i * nb + l], G[l * nb + j]);
                T[i * nb + j] = cmul(cmk(-tj.re, -tj.im), acc);
            }
        }
        MSP_SYNC();
    }
}

// Apply all blocks to the column chunk [c0, c0 + cw) of U = Q Z.
// Uc: n x ldu complex (LDS), Tl: nb x nb, W, W2: nb x cw.
MSP_HD void backtransform_blocked_team(int n, const cdbl *A, const cdbl *Tall, const double *Z,
                                       cdbl *U, cdbl *Uc, int ldu, int cw, int c0, int nb,
                                       cdbl *W, cdbl *W2)
{
    const int cwk = (n - c0 < cw) ? n - c0 : cw;
    MSP_PFOR(idx, n * cwk)
    {
        int i = idx / cwk, c = idx % cwk;
        Uc[i * ldu + c] = cmk(Z[(size_t)(c0 + c) * n + i], 0.0);
    }
    MSP_SYNC();
    const int nref = n - 1;
    const int nblk = (nref + nb - 1) / nb;
    for (int b = nblk - 1; b >= 0; --b) {
        const int k0 = b * nb;
        const int nbk = (nref - k0 < nb) ? nref - k0 : nb;
        const cdbl *T = Tall + (size_t)b * nb * nb;
        // W = V^H Uc(k0+1 .., :)
        MSP_PFOR(idx, nbk * cwk)
        {
            int j = idx / cwk, c = idx % cwk;
            cdbl acc = cmk(0.0, 0.0);
            for (int r = k0 + j + 1; r < n; ++r)
                cfmac(acc, vblock(n, A, k0, r, j), Uc[r * ldu + c]);
            W[j * cw + c] = acc;
        }
        MSP_SYNC();
        // W2 = T W (T upper triangular)
        MSP_PFOR(idx, nbk * cwk)
        {
            int j = idx / cwk, c = idx % cwk;
            cdbl acc = cmk(0.0, 0.0);
            for (int l = j; l < nbk; ++l) cfma(acc, T[j * nb + l], W[l * cw + c]);
            W2[j * cw + c] = acc;
        }
        MSP_SYNC();
        // Uc -= V W2
        MSP_PFOR(idx, (n - k0 - 1) * cwk)
        {
            int r = k0 + 1 + idx / cwk, c = idx % cwk;
            int jmax = r - k0 - 1;  // V(r, j) = 0 for j > r - k0 - 1
            if (jmax > nbk - 1) jmax = nbk - 1;
            cdbl acc = Uc[r * ldu + c];
            for (int j = 0; j <= jmax; ++j) {
                cdbl v = vblock(n, A, k0, r, j);
                cdbl t = cmul(v, W2[j * cw + c]);
                acc = csub(acc, t);
            }
            Uc[r * ldu + c] = acc;
        }
        MSP_SYNC();
    }
    MSP_PFOR(idx, n * cwk)
    {
        int i = idx / cwk, c = idx % cwk;
        U[(size_t)i * n + c0 + c] = Uc[i * ldu + c];
    }
    MSP_SYNC();
}

// ---------------------------------------------------------------------------
// Panel-blocked tridiagonalisation (same reduction as tridiag_team, with the
// trailing-matrix update deferred to the end of each panel of nb columns).
// Inside a panel, column k and the product A22 v use the panel-start matrix
// corrected by the panel's earlier vectors:
//   a(:, k) -= Vp W_p(k, :)^H + Wp Vp(k, :)^H,
//   p = tau (A22 v - Vp (Wp^H v) - Wp (Vp^H v)),  w = p - (tau/2)(p^H v) v,
// and after the panel the trailing block takes A -= Vp Wp^H + Wp Vp^H.
// Vp, Wp: n x nb complex (row stride nb) per-matrix scratch (global memory);
// y: 2*nb complex team scratch; vbuf, wbuf: n complex team scratch.
MSP_HD void tridiag_blocked_team(int n, cdbl *A, double *d, double *e, cdbl *tau, int nb,
                                 cdbl *Vp, cdbl *Wp, cdbl *vbuf, cdbl *wbuf, cdbl *y, double *red)
{
    for (int k0 = 0; k0 < n - 1; k0 += nb) {
        const int nbk = (n - 1 - k0 < nb) ? n - 1 - k0 : nb;
        MSP_PFOR(idx, n * nb)
        {
            Vp[idx] = cmk(0.0, 0.0);
            Wp[idx] = cmk(0.0, 0.0);
        }
        MSP_SYNC();
        for (int j = 0; j < nbk; ++j) {
            const int k = k0 + j;
            const int m = n - k - 1;
            // 1. bring column k (rows k..n-1) up to date with the panel
            if (j > 0) {
                MSP_PFOR(ii, n - k)
                {
                    const int i = k + ii;
                    cdbl a = A[(size_t)i * n + k];
                    for (int l = 0; l < j; ++l) {
                        a = csub(a, cmul(Vp[i * nb + l], cconj(Wp[k * nb + l])));
                        a = csub(a, cmul(Wp[i * nb + l], cconj(Vp[k * nb + l])));
                    }
                    A[(size_t)i * n + k] = a;
                }
                MSP_SYNC();
            }
            cdbl *col = A + (size_t)(k + 1) * n + k;
            // 2. reflector
            double part = 0.0;
            MSP_PFOR(i, m - 1) part += cabs2(col[(size_t)(i + 1) * n]);
            const double xn2 = team_sum(part, red);
            const cdbl alpha = col[0];
            double beta;
            cdbl tk, scale;
            if (xn2 == 0.0 && alpha.im == 0.0) {
                tk = cmk(0.0, 0.0);
                beta = alpha.re;
                scale = cmk(0.0, 0.0);
            } else {
                double nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn2);
                beta = alpha.re >= 0.0 ? -nrm : nrm;
                tk = cmk((beta - alpha.re) / beta, -alpha.im / beta);
                scale = cdivc(cmk(1.0, 0.0), csub(alpha, cmk(beta, 0.0)));
            }
            MSP_PFOR(i, m)
            {
                cdbl vi;
                if (i == 0) vi = cmk(1.0, 0.0);
                else {
                    vi = cmul(col[(size_t)i * n], scale);
                    col[(size_t)i * n] = vi;
                }
                vbuf[i] = vi;
                Vp[(k + 1 + i) * nb + j] = vi;
            }
            MSP_SYNC();
            const bool act = (tk.re != 0.0 || tk.im != 0.0);
            if (act) {
                // y1 = Wp^H v, y2 = Vp^H v over rows k+1.. (earlier panel vectors)
                MSP_PFOR(l, 2 * j)
                {
                    const cdbl *M = (l < j) ? Wp : Vp;
                    const int c = (l < j) ? l : l - j;
                    cdbl acc = cmk(0.0, 0.0);
                    for (int i = 0; i < m; ++i) cfmac(acc, M[(k + 1 + i) * nb + c], vbuf[i]);
                    y[l] = acc;
                }
                MSP_SYNC();
                // p = tau (A22 v - Vp y1 - Wp y2)
                const cdbl *A22 = A + (size_t)(k + 1) * n + (k + 1);
                MSP_PFOR(i, m)
                {
                    cdbl acc = cmk(0.0, 0.0);
                    for (int c = 0; c < m; ++c) cfmac(acc, A22[(size_t)c * n + i], vbuf[c]);
                    for (int l = 0; l < j; ++l) {
                        acc = csub(acc, cmul(Vp[(k + 1 + i) * nb + l], y[l]));
                        acc = csub(acc, cmul(Wp[(k + 1 + i) * nb + l], y[j + l]));
                    }
                    wbuf[i] = cmul(tk, acc);
                }
                MSP_SYNC();
                double pre = 0.0, pim = 0.0;
                MSP_PFOR(i, m)
                {
                    cdbl t = cmulc(wbuf[i], vbuf[i]);
                    pre += t.re;
                    pim += t.im;
                }
                double sre = team_sum(pre, red);
                double sim = team_sum(pim, red);
                cdbl a2 = cmul(cmk(-0.5 * tk.re, -0.5 * tk.im), cmk(sre, sim));
                MSP_PFOR(i, m) Wp[(k + 1 + i) * nb + j] = cadd(wbuf[i], cmul(a2, vbuf[i]));
                MSP_SYNC();
            } else {
                MSP_PFOR(i, m) Vp[(k + 1 + i) * nb + j] = cmk(0.0, 0.0);
                MSP_SYNC();
            }
#ifdef MSP_HOST_BUILD
            d[k] = A[(size_t)k * n + k].re;
            e[k] = beta;
            tau[k] = tk;
#else
            if (MSP_TID == 0) {
                d[k] = A[(size_t)k * n + k].re;
                e[k] = beta;
                tau[k] = tk;
            }
#endif
            MSP_SYNC();
        }
        // trailing update A(t:, t:) -= Vp Wp^H + Wp Vp^H, t = k0 + nbk
        const int t0 = k0 + nbk;
        const int mt = n - t0;
        MSP_PFOR(idx, mt * mt)
        {
            const int i = t0 + idx / mt, c = t0 + idx % mt;
            cdbl a = A[(size_t)i * n + c];
            for (int l = 0; l < nbk; ++l) {
                a = csub(a, cmul(Vp[i * nb + l], cconj(Wp[c * nb + l])));
                a = csub(a, cmul(Wp[i * nb + l], cconj(Vp[c * nb + l])));
            }
            A[(size_t)i * n + c] = a;
        }
        MSP_SYNC();
    }
#ifdef MSP_HOST_BUILD
    d[n - 1] = A[(size_t)(n - 1) * n + (n - 1)].re;
#else
    if (MSP_TID == 0) d[n - 1] = A[(size_t)(n - 1) * n + (n - 1)].re;
#endif
    MSP_SYNC();
}
