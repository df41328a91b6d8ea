// Cyclic two-sided Jacobi for a complex Hermitian matrix held by one team.
//
// Replaces, for dimensions that fit LDS, the LAPACK zheevd call made by
// Hermitian.diag (muspinsim/spinop.py:69).  Output matches numpy.linalg.eigh's
// contract: eigenvalues ascending, eigenvectors as orthonormal columns.
//
// Schedule: round-robin ("circle") ordering, m/2 disjoint (p, q) pairs per
// step, m = n rounded up to even.  Each step is two phases: (1) one thread
// per pair builds the 2x2 unitary J = [[c, s e], [-s conj(e), c]] with
// e = b/|b| that diagonalises [[a, b], [conj(b), d]]; (2) every 2x2 block
// (pair P rows, pair Q columns) of A is replaced by J_P^H X J_Q and every
// (row, pair) of V by [v_p v_q] J_Q.  All blocks are disjoint, so phase (2)
// needs no ordering.  Sweeps stop when a whole sweep applies no rotation.
#pragma once
#include "msp_team.h"

struct JacRot {
    int p, q;     // p < q, q >= n marks an idle pair
    double c;
    cdbl s;       // s * e
    double dp, dq; // new diagonal entries
};

MSP_HD void jac_pair(int k, int r, int m, int &p, int &q)
{
    int a = (k == 0) ? 0 : 1 + (k - 1 + r) % (m - 1);
    int b = 1 + (m - 2 - k + r) % (m - 1);
    p = a < b ? a : b;
    q = a < b ? b : a;
}

// A: n x n (row stride lda), V: n x n (row stride ldv), both team-shared.
// rot: m/2 JacRot, red: team scratch (>= MSP_NT doubles), flag: one int.
// On exit A's diagonal holds the (unsorted) eigenvalues, V the eigenvectors.
// Returns the number of sweeps used (> max_sweeps means not converged).
MSP_HD int jacobi_team(int n, cdbl *A, int lda, cdbl *V, int ldv, JacRot *rot,
                       double *red, int *flag, int max_sweeps)
{
    const int m = n + (n & 1);
    const int half = m / 2;

    // V = I
    MSP_PFOR(idx, n * n)
    {
        int i = idx / n, j = idx % n;
        V[i * ldv + j] = cmk(i == j ? 1.0 : 0.0, 0.0);
    }
    // Frobenius norm sets the rotation threshold.
    double part = 0.0;
    MSP_PFOR(idx, n * n)
    {
        int i = idx / n, j = idx % n;
        part += cabs2(A[i * lda + j]);
    }
    const double fro = sqrt(team_sum(part, red));
    const double thresh = 1.1102230246251565e-16 * fro; // 2^-53 ||A||_F
    MSP_SYNC();

    int sweep = 0;
    for (; sweep < max_sweeps; ++sweep) {
#ifdef MSP_HOST_BUILD
        *flag = 0;
#else
        if (MSP_TID == 0) *flag = 0;
        MSP_SYNC();
#endif
        for (int r = 0; r < m - 1; ++r) {
            MSP_PFOR(k, half)
            {
                int p, q;
                jac_pair(k, r, m, p, q);
                JacRot R;
                R.p = p;
                R.q = q;
                R.c = 1.0;
                R.s = cmk(0.0, 0.0);
                R.dp = 0.0;
                R.dq = 0.0;
                if (q < n) {
                    double a = A[p * lda + p].re, d = A[q * lda + q].re;
                    cdbl b = A[p * lda + q];
                    double beta = sqrt(cabs2(b));
                    R.dp = a;
                    R.dq = d;
                    if (beta > thresh) {
                        double zeta = (d - a) / (2.0 * beta);
                        double t = 1.0 / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                        if (zeta < 0.0) t = -t;
                        double c = 1.0 / sqrt(1.0 + t * t);
                        double s = t * c;
                        R.c = c;
                        R.s = cmk(s * b.re / beta, s * b.im / beta);
                        R.dp = a - t * beta;
                        R.dq = d + t * beta;
#ifdef MSP_HOST_BUILD
                        *flag = 1;
#else
                        *flag = 1; // benign race: every writer stores 1
#endif
                    }
                }
                rot[k] = R;
            }
            MSP_SYNC();
            // A <- J^H A J on disjoint 2x2 blocks
            MSP_PFOR(bi, half * half)
            {
                const JacRot P = rot[bi / half];
                const JacRot Q = rot[bi % half];
                if (P.q < n && Q.q < n) {
                    if (P.p == Q.p) {
                        A[P.p * lda + P.p] = cmk(P.dp, 0.0);
                        A[P.q * lda + P.q] = cmk(P.dq, 0.0);
                        A[P.p * lda + P.q] = cmk(0.0, 0.0);
                        A[P.q * lda + P.p] = cmk(0.0, 0.0);
                    } else {
                        cdbl x00 = A[P.p * lda + Q.p], x01 = A[P.p * lda + Q.q];
                        cdbl x10 = A[P.q * lda + Q.p], x11 = A[P.q * lda + Q.q];
                        // right: [y0 y1] = [x0 x1] J_Q, J = [[c, s], [-conj(s), c]]
                        cdbl sq = Q.s, sqc = cconj(Q.s);
                        cdbl y00 = csub(cscale(x00, Q.c), cmul(x01, sqc));
                        cdbl y01 = cadd(cmul(x00, sq), cscale(x01, Q.c));
                        cdbl y10 = csub(cscale(x10, Q.c), cmul(x11, sqc));
                        cdbl y11 = cadd(cmul(x10, sq), cscale(x11, Q.c));
                        // left: J_P^H = [[c, -s], [conj(s), c]]
                        cdbl sp = P.s, spc = cconj(P.s);
                        A[P.p * lda + Q.p] = csub(cscale(y00, P.c), cmul(sp, y10));
                        A[P.p * lda + Q.q] = csub(cscale(y01, P.c), cmul(sp, y11));
                        A[P.q * lda + Q.p] = cadd(cmul(spc, y00), cscale(y10, P.c));
                        A[P.q * lda + Q.q] = cadd(cmul(spc, y01), cscale(y11, P.c));
                    }
                } else if (P.q < n && Q.q >= n && Q.p < n) {
                    // column Q.p alone (its partner is the idle pad index)
                    cdbl x0 = A[P.p * lda + Q.p], x1 = A[P.q * lda + Q.p];
                    A[P.p * lda + Q.p] = csub(cscale(x0, P.c), cmul(P.s, x1));
                    A[P.q * lda + Q.p] = cadd(cmul(cconj(P.s), x0), cscale(x1, P.c));
                } else if (Q.q < n && P.q >= n && P.p < n) {
                    // row P.p alone
                    cdbl x0 = A[P.p * lda + Q.p], x1 = A[P.p * lda + Q.q];
                    A[P.p * lda + Q.p] = csub(cscale(x0, Q.c), cmul(x1, cconj(Q.s)));
                    A[P.p * lda + Q.q] = cadd(cmul(x0, Q.s), cscale(x1, Q.c));
                }
            }
            // V <- V J on (row, pair)
            MSP_PFOR(vi, n * half)
            {
                const JacRot Q = rot[vi % half];
                const int i = vi / half;
                if (Q.q < n && Q.s.re * Q.s.re + Q.s.im * Q.s.im > 0.0) {
                    cdbl v0 = V[i * ldv + Q.p], v1 = V[i * ldv + Q.q];
                    V[i * ldv + Q.p] = csub(cscale(v0, Q.c), cmul(v1, cconj(Q.s)));
                    V[i * ldv + Q.q] = cadd(cmul(v0, Q.s), cscale(v1, Q.c));
                }
            }
            MSP_SYNC();
        }
        int any = *flag;
        MSP_SYNC();
        if (!any) break;
    }
    return sweep;
}

// Ascending order with ties broken by index: rank of entry i among diag(A).
MSP_HD int jac_rank(int n, const cdbl *A, int lda, int i)
{
    double di = A[i * lda + i].re;
    int r = 0;
    for (int j = 0; j < n; ++j) {
        double dj = A[j * lda + j].re;
        if (dj < di || (dj == di && j < i)) ++r;
    }
    return r;
}
