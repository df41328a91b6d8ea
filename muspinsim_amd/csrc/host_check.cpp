// Host build of the device eigensolver team functions (one thread, same
// phase schedule), exported for CPU tests of the algorithms the HIP kernels
// run.  Not part of the product library.
#define MSP_HOST_BUILD 1
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "msp_jacobi.h"
#include "msp_nonsym.h"
#include "msp_tridiag.h"

extern "C" {

// H: n x n complex row-major (copied); E ascending; U columns = eigenvectors.
int hc_jacobi(int n, const double *H, double *E, double *U, int max_sweeps)
{
    std::vector<cdbl> A((size_t)n * n), V((size_t)n * n);
    memcpy(A.data(), H, sizeof(cdbl) * n * n);
    std::vector<JacRot> rot(n / 2 + 1);
    double red[1];
    int flag = 0;
    int sweeps = jacobi_team(n, A.data(), n, V.data(), n, rot.data(), red, &flag, max_sweeps);
    for (int i = 0; i < n; ++i) {
        int r = jac_rank(n, A.data(), n, i);
        E[r] = A[(size_t)i * n + i].re;
        for (int row = 0; row < n; ++row) {
            U[2 * ((size_t)row * n + r)] = V[(size_t)row * n + i].re;
            U[2 * ((size_t)row * n + r) + 1] = V[(size_t)row * n + i].im;
        }
    }
    return sweeps;
}

int hc_tridiag(int n, double *A, double *d, double *e, double *tau)
{
    std::vector<cdbl> vb(n), wb(n);
    double red[1];
    tridiag_team(n, (cdbl *)A, d, e, (cdbl *)tau, vb.data(), wb.data(), red);
    return 0;
}

struct DcAlloc {
    std::vector<double> dbl;
    std::vector<int> in;
    DcWork W;
    // Workspace starts as NaN (doubles) / -1 (ints), as reused device memory
    // would hold stale values: every read must follow a write.
    DcAlloc(int n) : dbl((size_t)4 * n * n + 12 * (size_t)n, NAN), in(24 * (size_t)n, -1)
    {
        double *p = dbl.data();
        W.Q0 = p; p += (size_t)n * n;
        W.Q1 = p; p += (size_t)n * n;
        W.S = p; p += (size_t)n * n;
        W.QK = p; p += (size_t)n * n;
        W.dv = p; p += n; W.zv = p; p += n; W.ds = p; p += n; W.zs = p; p += n;
        W.dl = p; p += n; W.zl = p; p += n; W.zh = p; p += n; W.ktau = p; p += n;
        W.rc = p; p += n; W.rs = p; p += n; W.item_v = p; p += n; W.outv = p; p += n;
        int *q = in.data();
        W.perm = q; q += n; W.kcol = q; q += n; W.korg = q; q += n; W.defl = q; q += n;
        W.nkept = q; q += n; W.rp = q; q += n; W.rn = q; q += n; W.nrot = q; q += n;
        W.dfl = q; q += n; W.item_src = q; q += n; W.outsrc = q; q += n; W.rootpos = q; q += n;
        W.blo = q; q += n; W.bhi = q; q += n; W.plo = q; q += n; W.phi = q; q += n;
        W.sup = q; q += n; W.kperm = q; q += n; W.nka = q; q += n; W.kb0 = q; q += n;
    }
};

// d: in diagonal, out ascending eigenvalues; Z: n x n real eigenvectors (columns).
int hc_stedc(int n, double *d, const double *e, double *Z)
{
    DcAlloc al(n);
    double red[1];
    double *Q = stedc_team(n, d, e, al.W, red);
    for (int i = 0; i < n; ++i)              // Q holds eigenvector columns contiguously
        for (int j = 0; j < n; ++j) Z[(size_t)i * n + j] = Q[(size_t)j * n + i];
    return 0;
}

int hc_tridiag_blocked(int n, double *A, double *d, double *e, double *tau, int nb)
{
    std::vector<cdbl> vb(n), wb(n), Vp((size_t)n * nb), Wp((size_t)n * nb), y(2 * nb);
    double red[1];
    tridiag_blocked_team(n, (cdbl *)A, d, e, (cdbl *)tau, nb, Vp.data(), Wp.data(), vb.data(),
                         wb.data(), y.data(), red);
    return 0;
}

// Full path: tridiagonalise, divide and conquer, back-transform.
int hc_eigh_dc_blocked(int n, const double *H, double *E, double *U, int cw, int nb)
{
    std::vector<cdbl> A((size_t)n * n);
    memcpy(A.data(), H, sizeof(cdbl) * n * n);
    std::vector<double> e(n);
    std::vector<cdbl> tau(n), vb(n), wb(n);
    double red[1];
    tridiag_team(n, A.data(), E, e.data(), tau.data(), vb.data(), wb.data(), red);
    DcAlloc al(n);
    double *Q = stedc_team(n, E, e.data(), al.W, red);
    int nblk = (n - 1 + nb - 1) / nb;
    std::vector<cdbl> T((size_t)(nblk > 0 ? nblk : 1) * nb * nb), G((size_t)nb * nb);
    for (int b = 0; b < nblk; ++b) {
        int k0 = b * nb, nbk = (n - 1 - k0 < nb) ? n - 1 - k0 : nb;
        larft_block_team(n, A.data(), tau.data(), k0, nbk, nb, T.data() + (size_t)b * nb * nb, G.data());
    }
    std::vector<cdbl> Uc((size_t)n * (cw + 1)), W((size_t)nb * cw), W2((size_t)nb * cw);
    for (int c0 = 0; c0 < n; c0 += cw)
        backtransform_blocked_team(n, A.data(), T.data(), Q, (cdbl *)U, Uc.data(), cw + 1, cw, c0, nb,
                                   W.data(), W2.data());
    return 0;
}

int hc_eigh_dc(int n, const double *H, double *E, double *U, int cw)
{
    std::vector<cdbl> A((size_t)n * n);
    memcpy(A.data(), H, sizeof(cdbl) * n * n);
    std::vector<double> e(n);
    std::vector<cdbl> tau(n), vb(n), wb(n), Uc((size_t)n * cw), ws(cw);
    double red[1];
    tridiag_team(n, A.data(), E, e.data(), tau.data(), vb.data(), wb.data(), red);
    DcAlloc al(n);
    double *Q = stedc_team(n, E, e.data(), al.W, red);
    backtransform_team(n, A.data(), tau.data(), Q, (cdbl *)U, Uc.data(), cw, ws.data());
    return 0;
}

// Non-Hermitian eigendecomposition: L (n x n complex, copied) -> lam (n
// complex), V (n x n complex, columns = right eigenvectors), and the mode
// amplitudes c = V^{-1} r, o = u^T V.  Returns the Schur status (0 = ok).
int hc_eig(int n, const double *L, double *lam, double *V, const double *r, const double *u,
           double *c, double *o, int max_its)
{
    std::vector<cdbl> A((size_t)n * n), Z((size_t)n * n), v(n), w(n), y(n), t(n);
    for (auto &x : A) x = cmk(NAN, NAN);
    for (auto &x : Z) x = cmk(NAN, NAN);
    memcpy(A.data(), L, sizeof(cdbl) * n * n);
    double red[1];
    hessenberg_team(n, A.data(), n, Z.data(), n, v.data(), w.data(), red);
    int info = schur_team(n, A.data(), n, Z.data(), n, (cdbl *)lam, max_its, red);
    if (info) return info;
    schur_vectors_team(n, A.data(), n, (const cdbl *)lam);
    modes_coeff_team(n, A.data(), n, Z.data(), n, (const cdbl *)r, y.data());
    memcpy(c, y.data(), sizeof(cdbl) * n);
    modes_proj_team(n, A.data(), n, Z.data(), n, (const cdbl *)u, t.data(), (cdbl *)o);
    modes_vectors_team(n, A.data(), n, Z.data(), n);
    memcpy(V, Z.data(), sizeof(cdbl) * n * n);
    return 0;
}

// The mode amplitudes without Z: r and u carried through the reduction as
// Z^H r and u^T Z (the device's LDS form when no eigenvectors are wanted).
int hc_eig_vec(int n, const double *L, double *lam, const double *r, const double *u, double *c,
               double *o, int max_its)
{
    std::vector<cdbl> A((size_t)n * n), v(n), w(n), y(n), t(n), rho(n), om(n);
    memcpy(A.data(), L, sizeof(cdbl) * n * n);
    memcpy(rho.data(), r, sizeof(cdbl) * n);
    memcpy(om.data(), u, sizeof(cdbl) * n);
    double red[1];
    ModeVecs mv{rho.data(), om.data(), 1};
    hessenberg_team(n, A.data(), n, nullptr, 0, v.data(), w.data(), red, mv);
    int info = schur_team(n, A.data(), n, nullptr, 0, (cdbl *)lam, max_its, red, mv);
    if (info) return info;
    schur_vectors_team(n, A.data(), n, (const cdbl *)lam);
    modes_coeff_team(n, A.data(), n, nullptr, 0, rho.data(), y.data());
    memcpy(c, y.data(), sizeof(cdbl) * n);
    modes_proj_team(n, A.data(), n, nullptr, 0, om.data(), t.data(), (cdbl *)o);
    return 0;
}

// Hessenberg + Schur only: T (upper triangular, lower part zeroed) and Z.
int hc_schur(int n, const double *L, double *T, double *Zout, double *lam, int max_its)
{
    std::vector<cdbl> A((size_t)n * n), Z((size_t)n * n), v(n), w(n);
    memcpy(A.data(), L, sizeof(cdbl) * n * n);
    double red[1];
    hessenberg_team(n, A.data(), n, Z.data(), n, v.data(), w.data(), red);
    int info = schur_team(n, A.data(), n, Z.data(), n, (cdbl *)lam, max_its, red);
    memcpy(T, A.data(), sizeof(cdbl) * n * n);
    memcpy(Zout, Z.data(), sizeof(cdbl) * n * n);
    return info;
}
}
