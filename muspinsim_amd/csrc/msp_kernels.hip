// HIP kernels of the spin-dynamics path, gfx950.  Launch wrappers and
// argument checks live in msp_capi.hip; the eigensolver algorithms are the
// team functions of msp_jacobi.h / msp_tridiag.h.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/muspin_hip.h"
#include "msp_jacobi.h"
#include "msp_kernels.h"
#include "msp_tridiag.h"

// <r|S^ax|c> for one spin of multiplicity d (m_k = I - k), ax = 0,1,2 = x,y,z.
// Raising elements <k|S+|k+1> = sqrt((k+1)(2I - k)) (muspinsim/spinop.py:18-41).
__device__ inline cdbl spin_el(int d, int ax, int r, int c)
{
    if (ax == 2) return (r == c) ? cmk(0.5 * (d - 1) - r, 0.0) : cmk(0.0, 0.0);
    if (r + 1 == c) {
        double sp = sqrt((double)(r + 1) * (double)(d - 1 - r));
        return ax == 0 ? cmk(0.5 * sp, 0.0) : cmk(0.0, -0.5 * sp);
    }
    if (c + 1 == r) {
        double sp = sqrt((double)(c + 1) * (double)(d - 1 - c));
        return ax == 0 ? cmk(0.5 * sp, 0.0) : cmk(0.0, 0.5 * sp);
    }
    return cmk(0.0, 0.0);
}

// <r|S^a S^b|c> (same-spin product, quadrupolar terms)
__device__ inline cdbl spin_el2(int d, int a, int b, int r, int c)
{
    cdbl acc = cmk(0.0, 0.0);
    for (int m = r - 1; m <= r + 1; ++m) {
        if (m < 0 || m >= d) continue;
        acc = cadd(acc, cmul(spin_el(d, a, r, m), spin_el(d, b, m, c)));
    }
    return acc;
}

// ---------------------------------------------------------------------------
__global__ void k_rotate(int batch, const double *q, const double *B, const double *Bint,
                         const double *p, double *Br, double *pr)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    const double w = q[4 * b], x = q[4 * b + 1], y = q[4 * b + 2], z = q[4 * b + 3];
    // v' = q v q*: v + 2w (u x v) + 2 u x (u x v), u = (x, y, z)
    for (int which = 0; which < 2; ++which) {
        const double *v = which ? p + 3 * b : B + 3 * b;
        double cx = y * v[2] - z * v[1], cy = z * v[0] - x * v[2], cz = x * v[1] - y * v[0];
        double ddx = y * cz - z * cy, ddy = z * cx - x * cz, ddz = x * cy - y * cx;
        double o0 = v[0] + 2.0 * (w * cx + ddx);
        double o1 = v[1] + 2.0 * (w * cy + ddy);
        double o2 = v[2] + 2.0 * (w * cz + ddz);
        if (which) {
            pr[3 * b] = o0; pr[3 * b + 1] = o1; pr[3 * b + 2] = o2;
        } else {
            Br[3 * b] = o0 + Bint[3 * b];
            Br[3 * b + 1] = o1 + Bint[3 * b + 1];
            Br[3 * b + 2] = o2 + Bint[3 * b + 2];
        }
    }
}

// ---------------------------------------------------------------------------
// H[b][a][c]: terms (or a precomputed base) + Zeeman, one thread per element.
__device__ cdbl h_element(int n, const MspSystem &S, const MspTerm *terms, const cdbl *bm,
                          const double *B, int zeeman, int idx)
{
    const int ar = idx / n, ac = idx % n;
    int dr[MSP_MAX_SPINS], dc[MSP_MAX_SPINS];
    int ndiff = 0, p0 = -1, p1 = -1;
    {
        int x = ar, y = ac;
        for (int k = S.n_spins - 1; k >= 0; --k) {
            int d = S.dims[k];
            dr[k] = x % d;
            dc[k] = y % d;
            x /= d;
            y /= d;
            if (dr[k] != dc[k]) {
                if (ndiff == 0) p0 = k; else if (ndiff == 1) p1 = k;
                ++ndiff;
            }
        }
    }
    cdbl acc = cmk(0.0, 0.0);
    if (bm) acc = bm[idx];
    if (ndiff <= 2) {
        if (!bm) {
            for (int t = 0; t < S.n_terms; ++t) {
                const MspTerm &T = terms[S.term_offset + t];
                if (T.kind == 0) {
                    if (ndiff == 0) acc.re += T.t[0];
                } else if (T.kind == 1) {
                    int i = T.i;
                    if (ndiff == 0 || (ndiff == 1 && p0 == i)) {
                        for (int a = 0; a < 3; ++a)
                            acc = cadd(acc, cscale(spin_el(S.dims[i], a, dr[i], dc[i]), T.t[a]));
                    }
                } else {
                    int i = T.i, j = T.j;
                    if (i != j) {
                        bool ok = (ndiff == 0) || (ndiff == 1 && (p0 == i || p0 == j)) ||
                                  (ndiff == 2 && ((p0 == i && p1 == j) || (p0 == j && p1 == i)));
                        if (ok) {
                            cdbl ei[3], ej[3];
                            for (int a = 0; a < 3; ++a) {
                                ei[a] = spin_el(S.dims[i], a, dr[i], dc[i]);
                                ej[a] = spin_el(S.dims[j], a, dr[j], dc[j]);
                            }
                            for (int a = 0; a < 3; ++a)
                                for (int c = 0; c < 3; ++c) {
                                    double m = T.t[3 * a + c];
                                    if (m != 0.0) acc = cadd(acc, cscale(cmul(ei[a], ej[c]), m));
                                }
                        }
                    } else if (ndiff == 0 || (ndiff == 1 && p0 == i)) {
                        for (int a = 0; a < 3; ++a)
                            for (int c = 0; c < 3; ++c) {
                                double m = T.t[3 * a + c];
                                if (m != 0.0)
                                    acc = cadd(acc, cscale(spin_el2(S.dims[i], a, c, dr[i], dc[i]), m));
                            }
                    }
                }
            }
        }
        if (zeeman && ndiff <= 1) {
            for (int i = 0; i < S.n_spins; ++i) {
                if (ndiff == 1 && p0 != i) continue;
                double g = S.gammas[i];
                for (int a = 0; a < 3; ++a)
                    acc = cadd(acc, cscale(spin_el(S.dims[i], a, dr[i], dc[i]), g * B[a]));
            }
        }
    }
    return acc;
}

// H[b][a][c]: the system's terms (or a precomputed base) + Zeeman, one thread
// per element, index digits by division (any multiplicities).
__global__ void k_build_h(int n, const MspSystem *systems, const MspTerm *terms,
                          const int32_t *sys_id, const double *Br, const double *base,
                          int zeeman, int lower, double *H)
{
    const int b = blockIdx.y;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * n) return;
    if (lower && idx % n > idx / n) return;
    const int s = sys_id ? sys_id[b] : 0;
    const cdbl *bm = base ? (const cdbl *)base + (size_t)s * n * n : nullptr;
    ((cdbl *)H)[(size_t)b * n * n + idx] = h_element(n, systems[s], terms, bm, Br + 3 * b, zeeman, idx);
}

// H[b] = base[sys[b]] + sum_i gamma_i Br[b].S_i.  The spin digits of every
// basis index are packed 4 bits per spin into one LDS word per index, so an
// element needs one XOR: zero means diagonal (only S_z contributes), a single
// non-zero nibble whose digits differ by one means a one-spin flip (S_x, S_y),
// anything else is the base element alone.  Systems with a multiplicity above
// 16 take the per-element path.
constexpr int kBuildEPT = 16;
__global__ void __launch_bounds__(256) k_build_h_base(int n, const MspSystem *systems,
                                                      const MspTerm *terms, const int32_t *sys_id,
                                                      const double *Br, const double *base,
                                                      int lower, double *H)
{
    extern __shared__ unsigned long long digs[];
    const int b = blockIdx.y;
    const int s = sys_id ? sys_id[b] : 0;
    const MspSystem &S = systems[s];
    const int ns = S.n_spins;
    bool packable = true;
    for (int k = 0; k < ns; ++k) packable = packable && (S.dims[k] <= 16);
    const cdbl *bm = (const cdbl *)base + (size_t)s * n * n;
    cdbl *Hb = (cdbl *)H + (size_t)b * n * n;
    const double *B = Br + 3 * b;
    // element t of the launch: row-major over the matrix, or (lower) over
    // its packed lower triangle, row r holding t = r (r + 1) / 2 + c
    const int cnt = lower ? n * (n + 1) / 2 : n * n;
    const int t0 = blockIdx.x * 256 * kBuildEPT + threadIdx.x;
    auto element = [&](int t) {
        if (!lower) return t;
        int r = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
        while ((r + 1) * (r + 2) / 2 <= t) ++r;
        while (r * (r + 1) / 2 > t) --r;
        return r * n + (t - r * (r + 1) / 2);
    };
    if (!packable) {
        for (int e = 0; e < kBuildEPT; ++e) {
            const int t = t0 + e * 256;
            if (t < cnt) {
                const int idx = element(t);
                Hb[idx] = h_element(n, S, terms, bm, B, 1, idx);
            }
        }
        return;
    }
    for (int r = threadIdx.x; r < n; r += 256) {
        unsigned long long w = 0;
        int x = r;
        for (int k = ns - 1; k >= 0; --k) {
            const int d = S.dims[k];
            w |= (unsigned long long)(x % d) << (4 * k);
            x /= d;
        }
        digs[r] = w;
    }
    __syncthreads();
    for (int e = 0; e < kBuildEPT; ++e) {
        const int t = t0 + e * 256;
        if (t >= cnt) break;
        const int idx = element(t);
        const int r = idx / n, c = idx - r * n;
        cdbl acc = bm[idx];
        const unsigned long long dr = digs[r], x = dr ^ digs[c];
        if (x == 0) {
            double z = 0.0;
            for (int k = 0; k < ns; ++k)
                z += S.gammas[k] * (0.5 * (S.dims[k] - 1) - (double)((dr >> (4 * k)) & 15));
            acc.re += B[2] * z;
        } else {
            const int k = __builtin_ctzll(x) >> 2;
            if ((x >> (4 * k)) <= 15) {
                const int a = (int)((dr >> (4 * k)) & 15), c2 = (int)((digs[c] >> (4 * k)) & 15);
                const int d = S.dims[k];
                const double g = S.gammas[k];
                acc = cadd(acc, cscale(spin_el(d, 0, a, c2), g * B[0]));
                acc = cadd(acc, cscale(spin_el(d, 1, a, c2), g * B[1]));
            }
        }
        Hb[idx] = acc;
    }
}

// ---------------------------------------------------------------------------
// Jacobi eigensolver: one workgroup per matrix, A and V in LDS.
__global__ void k_eigh_jacobi(int n, const double *H, double *E, double *U, int32_t *info,
                              int max_sweeps)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int b = blockIdx.x;
    const int ld = n + 1;
    const int m = n + (n & 1);
    cdbl *A = (cdbl *)smem;
    cdbl *V = A + (size_t)n * ld;
    JacRot *rot = (JacRot *)(V + (size_t)n * ld);
    double *red = (double *)(rot + m / 2);
    int *flag = (int *)(red + 16);
    int *rank = flag + 1;
    const cdbl *Hb = (const cdbl *)H + (size_t)b * n * n;
    MSP_PFOR(idx, n * n) A[(idx / n) * ld + idx % n] = Hb[idx];
    __syncthreads();
    int sweeps = jacobi_team(n, A, ld, V, ld, rot, red, flag, max_sweeps);
    __syncthreads();
    MSP_PFOR(i, n)
    {
        int r = jac_rank(n, A, ld, i);
        rank[i] = r;
        E[(size_t)b * n + r] = A[i * ld + i].re;
    }
    __syncthreads();
    cdbl *Ub = (cdbl *)U + (size_t)b * n * n;
    MSP_PFOR(idx, n * n)
    {
        int row = idx / n, j = idx % n;
        Ub[(size_t)row * n + rank[j]] = V[row * ld + j];
    }
    if (threadIdx.x == 0) info[b] = (sweeps >= max_sweeps) ? 1 : 0;
}

size_t jacobi_lds_bytes(int n)
{
    const int ld = n + 1, m = n + (n & 1);
    return 2 * sizeof(cdbl) * (size_t)n * ld + sizeof(JacRot) * (m / 2) + 16 * sizeof(double) +
           sizeof(int) * (1 + n) + 64;
}

// ---------------------------------------------------------------------------
// Divide-and-conquer path, three kernels per batch.
__global__ void __launch_bounds__(256) k_tridiag2(int n, double *H, double *E, double *work)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int b = blockIdx.x;
    cdbl *vb = (cdbl *)smem;
    cdbl *wb = vb + n;
    cdbl *y = wb + n;
    double *red = (double *)(y + 2 * MSP_TD_NB);
    double *base = work + (size_t)b * dc_work_doubles(n);
    double *e;
    cdbl *tau;
    dc_carve(n, base, &e, &tau);
    cdbl *Vp = dc_panels(n, base);
    cdbl *Wp = Vp + (size_t)n * MSP_TD_NB;
    tridiag_blocked_team(n, (cdbl *)H + (size_t)b * n * n, E + (size_t)b * n, e, tau, MSP_TD_NB,
                         Vp, Wp, vb, wb, y, red);
}

__global__ void k_tridiag(int n, double *H, double *E, double *work)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int b = blockIdx.x;
    cdbl *vb = (cdbl *)smem;
    cdbl *wb = vb + n;
    double *red = (double *)(wb + n);
    double *e;
    cdbl *tau;
    dc_carve(n, work + (size_t)b * dc_work_doubles(n), &e, &tau);
    tridiag_team(n, (cdbl *)H + (size_t)b * n * n, E + (size_t)b * n, e, tau, vb, wb, red);
}

// Divide and conquer on one tridiagonal per workgroup.  The n-sized arrays of
// the merges live in LDS when `small_in_lds` (dynamic LDS holds them), the
// n x n matrices (Q0, Q1, S, QK) in the global workspace.
// (four waves per SIMD: the compiler fits the kernel in 128 registers
// without spilling, so four 256-thread instances share a CU instead of three.
// NT = 1024 puts one instance on a whole CU, for the large dimensions whose
// instances would otherwise leave one wave per SIMD.)
// The placement of the n-arrays is a template parameter, not a run-time flag,
// so that each LDS pointer derives from the LDS block alone and the compiler
// emits LDS instructions for it (a pointer that may be either LDS or global
// memory is accessed with flat instructions, slower for LDS).  MODE 0: all in
// the global workspace; 1: all in LDS; 2: all in LDS but the four arrays of
// the deflation rotations (rc, rs, rp, rn), which stay in the workspace, so
// that n up to 1024 fits one CU's LDS (152 n bytes).
template <int NT, int MODE>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4)))
k_stedc(int n, double *E, double *work, int32_t *info)
{
    __shared__ double red[256];
    extern __shared__ __align__(16) unsigned char smem[];
    const int b = blockIdx.x;
    double *e;
    cdbl *tau;
    DcWork W = dc_carve(n, work + (size_t)b * dc_work_doubles(n), &e, &tau);
    if constexpr (MODE == 1) {
        double *p = (double *)smem;
        W.dv = p; p += n; W.zv = p; p += n; W.ds = p; p += n; W.zs = p; p += n;
        W.dl = p; p += n; W.zl = p; p += n; W.zh = p; p += n; W.ktau = p; p += n;
        W.rc = p; p += n; W.rs = p; p += n; W.item_v = p; p += n; W.outv = p; p += n;
        int *q = (int *)p;
        W.perm = q; q += n; W.kcol = q; q += n; W.korg = q; q += n; W.defl = q; q += n;
        W.nkept = q; q += n; W.rp = q; q += n; W.rn = q; q += n; W.nrot = q; q += n;
        W.dfl = q; q += n; W.item_src = q; q += n; W.outsrc = q; q += n; W.rootpos = q; q += n;
        W.blo = q; q += n; W.bhi = q; q += n; W.plo = q; q += n; W.phi = q; q += n;
        W.sup = q; q += n; W.kperm = q; q += n; W.nka = q; q += n; W.kb0 = q; q += n;
    } else if constexpr (MODE == 2) {
        double *p = (double *)smem;
        W.dv = p; p += n; W.zv = p; p += n; W.ds = p; p += n; W.zs = p; p += n;
        W.dl = p; p += n; W.zl = p; p += n; W.zh = p; p += n; W.ktau = p; p += n;
        W.item_v = p; p += n; W.outv = p; p += n;
        int *q = (int *)p;
        W.perm = q; q += n; W.kcol = q; q += n; W.korg = q; q += n; W.defl = q; q += n;
        W.nkept = q; q += n; W.nrot = q; q += n;
        W.dfl = q; q += n; W.item_src = q; q += n; W.outsrc = q; q += n; W.rootpos = q; q += n;
        W.blo = q; q += n; W.bhi = q; q += n; W.plo = q; q += n; W.phi = q; q += n;
        W.sup = q; q += n; W.kperm = q; q += n; W.nka = q; q += n; W.kb0 = q; q += n;
    }
    stedc_team(n, E + (size_t)b * n, e, W, red);
    if (threadIdx.x == 0) info[b] = 0;
}

size_t stedc_lds_bytes(int n) { return (size_t)n * (12 * sizeof(double) + 20 * sizeof(int)) + 16; }
// LDS of MODE 2 (without the rotation arrays)
static size_t stedc_lds2_bytes(int n) { return (size_t)n * (10 * sizeof(double) + 18 * sizeof(int)) + 16; }

constexpr int kLarftStageMaxDev = 256;   // n up to which k_larft stages the block in LDS

// T factors of every reflector block, and V T for the MFMA back-transformation:
// grid (blocks, instances).  With STAGE the block's reflectors are first
// copied to LDS (n x 16, rows padded to 17), so that the Gram matrix and V T
// read them from there instead of from strided global rows.
template <bool STAGE>
__global__ void __launch_bounds__(256) k_larft(int n, const double *H, double *work)
{
    __shared__ cdbl G[MSP_BT_NB * MSP_BT_NB];
    __shared__ cdbl T[MSP_BT_NB * MSP_BT_NB];
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int LDV = MSP_BT_NB + 1;
    const int b = blockIdx.y, blk = blockIdx.x;
    double *base = work + (size_t)b * dc_work_doubles(n);
    const cdbl *tau = dc_tau_ptr(n, base);
    const int k0 = blk * MSP_BT_NB;
    const int nbk = (n - 1 - k0 < MSP_BT_NB) ? n - 1 - k0 : MSP_BT_NB;
    const cdbl *A = (const cdbl *)H + (size_t)b * n * n;
    cdbl *Vs = (cdbl *)smem;
    if constexpr (STAGE) {
        // all of a thread's loads are issued before any LDS store (n <= 256,
        // so at most 16 elements per thread)
        constexpr int PER = (kLarftStageMaxDev * MSP_BT_NB + 255) / 256;
        cdbl tmp[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int idx = threadIdx.x + 256 * q;
            const int r = idx / MSP_BT_NB, l = idx % MSP_BT_NB;
            tmp[q] = (idx < n * MSP_BT_NB && l < nbk) ? vblock(n, A, k0, r, l) : cmk(0.0, 0.0);
        }
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int idx = threadIdx.x + 256 * q;
            if (idx < n * MSP_BT_NB) Vs[(idx / MSP_BT_NB) * LDV + idx % MSP_BT_NB] = tmp[q];
        }
        __syncthreads();
        // Gram G(i, j) = V(:, i)^H V(:, j), i < j: each of the 120 pairs is
        // split over two threads (rows halved, four accumulators each); the
        // second half's partial goes through T, which is rebuilt below
        {
            const int t = threadIdx.x;
            const int q = t & 127, h = t >> 7;
            int i = 0, rem = q;
            while (i < MSP_BT_NB - 1 && rem >= MSP_BT_NB - 1 - i) {
                rem -= MSP_BT_NB - 1 - i;
                ++i;
            }
            const int j = i + 1 + rem;
            cdbl part = cmk(0.0, 0.0);
            if (q < MSP_BT_NB * (MSP_BT_NB - 1) / 2 && j < nbk) {
                const int r0 = k0 + j + 1;
                const int mid = r0 + (n - r0) / 2;
                const int ra = h ? mid : r0, rb = h ? n : mid;
                cdbl a0 = cmk(0.0, 0.0), a1 = a0, a2 = a0, a3 = a0;
                int r = ra;
                for (; r + 3 < rb; r += 4) {
                    cfmac(a0, Vs[r * LDV + i], Vs[r * LDV + j]);
                    cfmac(a1, Vs[(r + 1) * LDV + i], Vs[(r + 1) * LDV + j]);
                    cfmac(a2, Vs[(r + 2) * LDV + i], Vs[(r + 2) * LDV + j]);
                    cfmac(a3, Vs[(r + 3) * LDV + i], Vs[(r + 3) * LDV + j]);
                }
                for (; r < rb; ++r) cfmac(a0, Vs[r * LDV + i], Vs[r * LDV + j]);
                part = cmk(a0.re + a1.re + a2.re + a3.re, a0.im + a1.im + a2.im + a3.im);
            }
            if (t < 2 * 128 && q < MSP_BT_NB * (MSP_BT_NB - 1) / 2)
                (h ? T : G)[i * MSP_BT_NB + j] = part;
        }
        __syncthreads();
        {
            const int t = threadIdx.x;
            const int i = t / MSP_BT_NB, j = t % MSP_BT_NB;
            if (t < MSP_BT_NB * MSP_BT_NB && i < j) {
                const cdbl a = G[t], b = T[t];
                G[t] = cmk(a.re + b.re, a.im + b.im);
            }
        }
        __syncthreads();
        // T = (diag(1/tau) + strictly upper part of V^H V)^{-1}, by back
        // substitution one column per thread, multiplied through by tau_i so
        // that tau_i = 0 (no reflection) gives a zero row:
        //   T(j, j) = tau_j,  T(i, j) = -tau_i sum_{k = i+1..j} G(i, k) T(k, j)
        // (each thread's column is kept in registers: static indices after
        // unrolling, terms with k > j are zero since tc[k] = 0 there)
        if (threadIdx.x < MSP_BT_NB) {
            const int j = threadIdx.x;
            cdbl tc[MSP_BT_NB];
#pragma unroll
            for (int i = MSP_BT_NB - 1; i >= 0; --i) {
                cdbl t = cmk(0.0, 0.0);
                if (j < nbk && i <= j) {
                    if (i == j) {
                        t = tau[k0 + j];
                    } else {
                        cdbl acc = cmk(0.0, 0.0);
#pragma unroll
                        for (int k = i + 1; k < MSP_BT_NB; ++k) cfma(acc, G[i * MSP_BT_NB + k], tc[k]);
                        const cdbl ti = tau[k0 + i];
                        t = cmul(cmk(-ti.re, -ti.im), acc);
                    }
                }
                tc[i] = t;
            }
#pragma unroll
            for (int i = 0; i < MSP_BT_NB; ++i) T[i * MSP_BT_NB + j] = tc[i];
        }
        __syncthreads();
    } else {
        larft_block_team(n, A, tau, k0, nbk, MSP_BT_NB, T, G);
    }
    cdbl *Tout = dc_tblocks(n, base) + (size_t)blk * MSP_BT_NB * MSP_BT_NB;
    MSP_PFOR(idx, MSP_BT_NB * MSP_BT_NB) Tout[idx] = T[idx];
    // V T (T upper triangular), so that the MFMA back-transformation applies
    // U -= (V T)(V^H U) without a T W step
    cdbl *VT = dc_vt(n, base) + (size_t)blk * MSP_BT_NB * n;
    if constexpr (STAGE) {
        // one row of V per thread, held in registers; T entries are LDS
        // broadcasts; the VT columns are written along r (coalesced)
        MSP_PFOR(r, n)
        {
            cdbl v[MSP_BT_NB];
#pragma unroll
            for (int l = 0; l < MSP_BT_NB; ++l) v[l] = Vs[r * LDV + l];
            // (T is upper triangular: the terms l > c add zeros, and the
            // full l loop keeps v in registers with static indices)
#pragma unroll 1
            for (int c = 0; c < MSP_BT_NB; ++c) {
                cdbl acc = cmk(0.0, 0.0);
#pragma unroll
                for (int l = 0; l < MSP_BT_NB; ++l) cfma(acc, v[l], T[l * MSP_BT_NB + c]);
                VT[(size_t)c * n + r] = acc;
            }
        }
    } else {
        MSP_PFOR(idx, MSP_BT_NB * n)
        {
            const int c = idx / n, r = idx % n;
            cdbl acc = cmk(0.0, 0.0);
            const int lmax = (c < nbk) ? c : nbk - 1;
            for (int l = 0; l <= lmax; ++l) cfma(acc, vblock(n, A, k0, r, l), T[l * MSP_BT_NB + c]);
            VT[idx] = acc;
        }
    }
}

// LDS-staged WY factors up to this order (n x 17 complex, 52 KB at n = 192)
constexpr int kLarftStageMax = kLarftStageMaxDev;

hipError_t launch_larft(int n, int batch, const double *H, double *work, hipStream_t s)
{
    if (n <= 1024) return launch_larft_mfma(n, batch, H, work, s);
    if (n <= kLarftStageMax) {
        const size_t lds = sizeof(cdbl) * (size_t)n * (MSP_BT_NB + 1);
        hipError_t err = msp_set_lds((const void *)k_larft<true>, lds);
        if (err != hipSuccess) return err;
        hipLaunchKernelGGL(k_larft<true>, dim3(bt_nblocks(n), batch), dim3(256), lds, s, n, H, work);
    } else {
        hipLaunchKernelGGL(k_larft<false>, dim3(bt_nblocks(n), batch), dim3(256), 0, s, n, H, work);
    }
    return hipGetLastError();
}

// U = Q Z on column chunks of width cw: grid (chunks, instances).
__global__ void __launch_bounds__(256) k_backtransform2(int n, const double *H, double *work,
                                                        double *U, int cw)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int b = blockIdx.y;
    const int c0 = blockIdx.x * cw;
    const int ldu = cw + 1;
    cdbl *Uc = (cdbl *)smem;
    cdbl *W = Uc + (size_t)n * ldu;
    cdbl *W2 = W + MSP_BT_NB * cw;
    double *base = work + (size_t)b * dc_work_doubles(n);
    double *e;
    cdbl *tau;
    DcWork Wk = dc_carve(n, base, &e, &tau);
    backtransform_blocked_team(n, (const cdbl *)H + (size_t)b * n * n, dc_tblocks(n, base),
                               dc_result(n, Wk), (cdbl *)U + (size_t)b * n * n, Uc, ldu, cw, c0,
                               MSP_BT_NB, W, W2);
}

__global__ void k_backtransform(int n, const double *H, double *work, double *U, int cw)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int b = blockIdx.y;
    const int c0 = blockIdx.x * cw;
    cdbl *Uc = (cdbl *)smem;
    cdbl *ws = Uc + (size_t)n * cw;
    double *e;
    cdbl *tau;
    DcWork W = dc_carve(n, work + (size_t)b * dc_work_doubles(n), &e, &tau);
    backtransform_chunk_team(n, (const cdbl *)H + (size_t)b * n * n, tau, dc_result(n, W),
                             (cdbl *)U + (size_t)b * n * n, Uc, cw, c0, ws);
}

// ---------------------------------------------------------------------------
// Y = (v.S_mu (x) 1) U (pauli: sigma = 2S).
__global__ void k_apply_muon(int n, const MspSystem *systems, const int32_t *sys_id,
                             const double *vec, int pauli, const double *U, double *Y)
{
    const int b = blockIdx.y;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * n) return;
    const MspSystem &S = systems[sys_id ? sys_id[b] : 0];
    const int mu = S.muon_index;
    int stride = 1;
    for (int k = S.n_spins - 1; k > mu; --k) stride *= S.dims[k];
    const int dm = S.dims[mu];
    const int row = idx / n, col = idx % n;
    const int am = (row / stride) % dm;
    const double *v = vec + 3 * b;
    const double f = pauli ? 2.0 : 1.0;
    const cdbl *Ub = (const cdbl *)U + (size_t)b * n * n;
    cdbl acc = cmk(0.0, 0.0);
    for (int mp = 0; mp < dm; ++mp) {
        cdbl mel = cmk(0.0, 0.0);
        for (int a = 0; a < 3; ++a) mel = cadd(mel, cscale(spin_el(dm, a, am, mp), f * v[a]));
        if (mel.re == 0.0 && mel.im == 0.0) continue;
        int r2 = row + (mp - am) * stride;
        acc = cadd(acc, cmul(mel, Ub[(size_t)r2 * n + col]));
    }
    ((cdbl *)Y)[(size_t)b * n * n + idx] = acc;
}

// One mode product: Y[a][c] = sum_m F[a_k][m] X[a | digit k = m][c].
__global__ void k_mode_product(int n, int dk, int stride, int foff, int fstride, const double *F,
                               const double *X, double *Y)
{
    const int b = blockIdx.y;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * n) return;
    const int row = idx / n, col = idx % n;
    const int ak = (row / stride) % dk;
    const cdbl *Fb = (const cdbl *)F + (size_t)b * fstride + foff;
    const cdbl *Xb = (const cdbl *)X + (size_t)b * n * n;
    cdbl acc = cmk(0.0, 0.0);
    for (int mp = 0; mp < dk; ++mp) {
        int r2 = row + (mp - ak) * stride;
        cfma(acc, Fb[ak * dk + mp], Xb[(size_t)r2 * n + col]);
    }
    ((cdbl *)Y)[(size_t)b * n * n + idx] = acc;
}


// ---------------------------------------------------------------------------
__global__ void k_form_c(int n, const double *rho, const double *obs, int real_mode,
                         double alpha, double beta, double *C)
{
    const int b = blockIdx.y;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * n) return;
    const int i = idx / n, j = idx % n;
    const cdbl *O = (const cdbl *)obs + (size_t)b * n * n;
    if (real_mode) {
        double v = alpha * cabs2(O[idx]);
        if (i == j) v += beta * O[idx].re;
        C[(size_t)b * n * n + idx] = v;
    } else {
        const cdbl *R = (const cdbl *)rho + (size_t)b * n * n;
        ((cdbl *)C)[(size_t)b * n * n + idx] = cmul(R[idx], O[(size_t)j * n + i]);
    }
}

// ---------------------------------------------------------------------------
// Time sum P(t) = sum_i conj(F_i(t)) sum_j C_ij F_j(t), F_j(t) = exp(i phi_j(t)),
// phi_j = 2 pi (E_j - E_c) t with E_c the spectrum midpoint (phases formed in
// FP64 from centred energies).  Workgroup = (instance, 64 times); 256 threads
// = 16 row groups x 16 time groups, 4 x 4 complex accumulators each.
template <bool REAL>
__global__ void __launch_bounds__(256) k_timesum(int n, int ntimes, const double *E,
                                                 const double *C, const double *times,
                                                 const double *w, int out_complex,
                                                 double *out)
{
    constexpr int TI = 64, TT = 64, TJ = 16;
    __shared__ double Cs_re[TI][TJ + 1];
    __shared__ double Cs_im[REAL ? 1 : TI][TJ + 1];
    __shared__ cdbl Ph[TJ][TT + 1];
    __shared__ cdbl red[16][TT + 1];
    int tt, b;
    msp_xcd_tile(gridDim.x, gridDim.y, tt, b);   // an instance's time tiles share one L2
    const int t0 = tt * TT;
    const int tg = threadIdx.x % 16, ig = threadIdx.x / 16;
    const double *Eb = E + (size_t)b * n;
    const double Ec = 0.5 * (Eb[0] + Eb[n - 1]);
    const double twopi = 6.283185307179586;
    double tloc[4];
    for (int s = 0; s < 4; ++s) {
        int t = t0 + tg * 4 + s;
        tloc[s] = (t < ntimes) ? times[t] : 0.0;
    }
    cdbl Pacc[4];
    for (int s = 0; s < 4; ++s) Pacc[s] = cmk(0.0, 0.0);
    for (int i0 = 0; i0 < n; i0 += TI) {
        cdbl acc[4][4];
        for (int r = 0; r < 4; ++r)
            for (int s = 0; s < 4; ++s) acc[r][s] = cmk(0.0, 0.0);
        for (int j0 = 0; j0 < n; j0 += TJ) {
            for (int l = threadIdx.x; l < TI * TJ; l += 256) {
                int r = l / TJ, c = l % TJ;
                int i = i0 + r, j = j0 + c;
                bool ok = (i < n && j < n);
                if (REAL) {
                    Cs_re[r][c] = ok ? C[(size_t)b * n * n + (size_t)i * n + j] : 0.0;
                } else {
                    cdbl v = ok ? ((const cdbl *)C)[(size_t)b * n * n + (size_t)i * n + j] : cmk(0.0, 0.0);
                    Cs_re[r][c] = v.re;
                    Cs_im[r][c] = v.im;
                }
            }
            for (int l = threadIdx.x; l < TJ * TT; l += 256) {
                int jj = l / TT, tt = l % TT;
                int j = j0 + jj, t = t0 + tt;
                cdbl ph = cmk(0.0, 0.0);
                if (j < n && t < ntimes) {
                    double phi = twopi * (Eb[j] - Ec) * times[t];
                    double sn, cs;
                    sincos(phi, &sn, &cs);
                    ph = cmk(cs, sn);
                }
                Ph[jj][tt] = ph;
            }
            __syncthreads();
#pragma unroll 4
            for (int jj = 0; jj < TJ; ++jj) {
                cdbl p[4];
                for (int s = 0; s < 4; ++s) p[s] = Ph[jj][tg * 4 + s];
                for (int r = 0; r < 4; ++r) {
                    double cr = Cs_re[ig * 4 + r][jj];
                    if (REAL) {
                        for (int s = 0; s < 4; ++s) {
                            acc[r][s].re = fma(cr, p[s].re, acc[r][s].re);
                            acc[r][s].im = fma(cr, p[s].im, acc[r][s].im);
                        }
                    } else {
                        cdbl c = cmk(cr, Cs_im[ig * 4 + r][jj]);
                        for (int s = 0; s < 4; ++s) cfma(acc[r][s], c, p[s]);
                    }
                }
            }
            __syncthreads();
        }
        // contract the rows with conj(F_i(t))
        for (int s = 0; s < 4; ++s) {
            cdbl part = cmk(0.0, 0.0);
            for (int r = 0; r < 4; ++r) {
                int i = i0 + ig * 4 + r;
                if (i >= n) continue;
                double phi = twopi * (Eb[i] - Ec) * tloc[s];
                double sn, cs;
                sincos(phi, &sn, &cs);
                cfmac(part, cmk(cs, sn), acc[r][s]);
            }
            Pacc[s] = cadd(Pacc[s], part);
        }
    }
    for (int s = 0; s < 4; ++s) red[ig][tg * 4 + s] = Pacc[s];
    __syncthreads();
    if (threadIdx.x < TT) {
        int tt = threadIdx.x, t = t0 + tt;
        cdbl P = cmk(0.0, 0.0);
        for (int g = 0; g < 16; ++g) P = cadd(P, red[g][tt]);
        if (t < ntimes) {
            if (out_complex) {
                out[2 * ((size_t)b * ntimes + t)] = P.re;
                out[2 * ((size_t)b * ntimes + t) + 1] = P.im;
            } else {
                double wb = w ? w[b] : 1.0;
                out[(size_t)b * ntimes + t] = wb * P.re;
            }
        }
    }
}

template __global__ void k_timesum<true>(int, int, const double *, const double *, const double *,
                                         const double *, int, double *);
template __global__ void k_timesum<false>(int, int, const double *, const double *, const double *,
                                          const double *, int, double *);

// ---------------------------------------------------------------------------
// out[t]_ij = exp(-2 pi i (E_i - E_j) t) rho_ij: the eigenbasis density matrix
// of one instance at every time.
__global__ void k_phase_evolve(int n, int ntimes, const double *E, const double *rho,
                               const double *times, double *out)
{
    const int t = blockIdx.y;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * n || t >= ntimes) return;
    const int i = idx / n, j = idx - (idx / n) * n;
    double sn, cs;
    sincos(-6.283185307179586 * (E[i] - E[j]) * times[t], &sn, &cs);
    const cdbl r = ((const cdbl *)rho)[idx];
    ((cdbl *)out)[(size_t)t * n * n + idx] = cmul(cmk(cs, sn), r);
}

// ---------------------------------------------------------------------------
// I = sum_ij C_ij / (1/tau + 2 pi i (E_i - E_j)), one workgroup per instance.
__global__ void __launch_bounds__(256) k_integral(int n, const double *E, const double *C,
                                                  int real_mode, double tau, double out_scale,
                                                  const double *w, int out_complex, double *out)
{
    __shared__ double red[16];
    const int b = blockIdx.x;
    const double *Eb = E + (size_t)b * n;
    const double a = 1.0 / tau;
    const double twopi = 6.283185307179586;
    double sre = 0.0, sim = 0.0;
    MSP_PFOR(idx, n * n)
    {
        int i = idx / n, j = idx % n;
        double bb = twopi * (Eb[i] - Eb[j]);
        double den = a * a + bb * bb;
        if (real_mode) {
            double c = C[(size_t)b * n * n + idx];
            sre += c * a / den;
            sim -= c * bb / den;
        } else {
            cdbl c = ((const cdbl *)C)[(size_t)b * n * n + idx];
            // c / (a + i bb) = c (a - i bb) / den
            sre += (c.re * a + c.im * bb) / den;
            sim += (c.im * a - c.re * bb) / den;
        }
    }
    sre = team_sum(sre, red);
    sim = team_sum(sim, red);
    if (threadIdx.x == 0) {
        if (out_complex) {
            out[2 * b] = sre * out_scale;
            out[2 * b + 1] = sim * out_scale;
        } else {
            out[b] = (w ? w[b] : 1.0) * sre * out_scale;
        }
    }
}

// Integral of C_ij = alpha |O_ij|^2 + beta delta_ij Re O_ii against
// 1 / (1/tau + 2 pi i (E_i - E_j)), real part, read straight from O's 64 x 64
// blocks on and above the block diagonal: O is Hermitian, so |O_ij| = |O_ji|,
// and the real part of the kernel is even in E_i - E_j, so a block above the
// diagonal counts twice and the blocks below are never read.
__global__ void __launch_bounds__(256) k_integral_obs(int n, const double *E, const double *O,
                                                      double alpha, double beta, double tau,
                                                      double out_scale, const double *w,
                                                      double *out)
{
    __shared__ double red[16];
    const int b = blockIdx.x;
    const double *Eb = E + (size_t)b * n;
    const cdbl *Ob = (const cdbl *)O + (size_t)b * n * n;
    const double a = 1.0 / tau;
    const double twopi = 6.283185307179586;
    double sre = 0.0;
    MSP_PFOR(idx, n * n)
    {
        const int i = idx / n, j = idx % n;
        const int bi = i >> 6, bj = j >> 6;
        if (bi <= bj) {
            const cdbl o = Ob[idx];
            double c = alpha * (o.re * o.re + o.im * o.im);
            if (i == j) c += beta * o.re;
            const double bb = twopi * (Eb[i] - Eb[j]);
            sre += ((bi < bj) ? 2.0 : 1.0) * c * a / (a * a + bb * bb);
        }
    }
    sre = team_sum(sre, red);
    if (threadIdx.x == 0) out[b] = (w ? w[b] : 1.0) * sre * out_scale;
}

// ---------------------------------------------------------------------------
// results[s][t] += scale * sum of rows[idx[k]][t] over the instances k of
// slot s.  A workgroup owns 64 output elements; its kAccParts waves sum
// contiguous parts of each slot's instance range (four loads in flight per
// lane) and wave 0 adds the parts in order, so the sum order is fixed.
constexpr int kAccParts = 4;

__global__ void __launch_bounds__(64 * kAccParts) k_accumulate(int nslots, int width, const int32_t *ptr,
                                                               const int32_t *idx, const double *rows,
                                                               double scale, double *results)
{
    __shared__ double part[kAccParts][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long long g = (long long)blockIdx.x * 64 + lane;
    const bool ok = g < (long long)nslots * width;
    double acc = 0.0;
    if (ok) {
        const int s = (int)(g / width), t = (int)(g % width);
        const int k0 = ptr[s], len = ptr[s + 1] - k0;
        const int a = k0 + (int)((long long)len * wave / kAccParts);
        const int b = k0 + (int)((long long)len * (wave + 1) / kAccParts);
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        int k = a;
        for (; k + 3 < b; k += 4) {
            a0 += rows[(size_t)idx[k] * width + t];
            a1 += rows[(size_t)idx[k + 1] * width + t];
            a2 += rows[(size_t)idx[k + 2] * width + t];
            a3 += rows[(size_t)idx[k + 3] * width + t];
        }
        for (; k < b; ++k) a0 += rows[(size_t)idx[k] * width + t];
        acc = (a0 + a1) + (a2 + a3);
    }
    part[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && ok) {
        double sum = part[0][lane];
#pragma unroll
        for (int w = 1; w < kAccParts; ++w) sum += part[w][lane];
        results[g] += scale * sum;
    }
}

// ---------------------------------------------------------------------------
// Optional per-kernel timing: when enabled, every launch below is bracketed by
// two hipEvents on its stream; msp_profile_report sums them per kernel name.
#include <string>
#include <vector>

namespace {
struct ProfRec {
    const char *name;
    hipEvent_t a, b;
};
bool g_prof = false;
std::vector<ProfRec> g_recs;
std::vector<hipEvent_t> g_pool;

hipEvent_t prof_event()
{
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

}  // namespace

ProfScope::ProfScope(const char *n, hipStream_t st) : name(n), s(st)
{
    if (g_prof && (a = prof_event())) hipEventRecord(a, s);
}

ProfScope::~ProfScope()
{
    if (!a) return;
    hipEvent_t b = prof_event();
    if (!b) return;
    hipEventRecord(b, s);
    g_recs.push_back({name, a, b});
}


void prof_enable(int on) { g_prof = on != 0; }

// "name calls total_ms\n" lines; synchronises on the recorded events and resets.
std::string prof_report()
{
    std::vector<std::string> names;
    std::vector<double> tot;
    std::vector<long> cnt;
    for (auto &r : g_recs) {
        hipEventSynchronize(r.b);
        float ms = 0.f;
        hipEventElapsedTime(&ms, r.a, r.b);
        size_t k = 0;
        for (; k < names.size(); ++k)
            if (names[k] == r.name) break;
        if (k == names.size()) {
            names.push_back(r.name);
            tot.push_back(0.0);
            cnt.push_back(0);
        }
        tot[k] += ms;
        cnt[k] += 1;
        g_pool.push_back(r.a);
        g_pool.push_back(r.b);
    }
    g_recs.clear();
    std::string out;
    char line[160];
    for (size_t k = 0; k < names.size(); ++k) {
        snprintf(line, sizeof(line), "%s %ld %.6f\n", names[k].c_str(), cnt[k], tot[k]);
        out += line;
    }
    return out;
}

// ---------------------------------------------------------------------------
// Launchers
static inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

size_t dc_workspace_doubles(int n) { return dc_work_doubles(n); }

// Blocked back-transformation chunk width: U chunk (n x (cw+1)) plus the two
// nb x cw block products within ~76 KB of LDS (two workgroups per CU).
int bt2_chunk_width(int n)
{
    int cw = 64;
    while (cw > 2 && ((size_t)n * (cw + 1) + 2 * MSP_BT_NB * cw) * sizeof(cdbl) > 76 * 1024) cw /= 2;
    return cw;
}

// Back-transformation chunk width: as many columns as fit ~128 KB of LDS.
int dc_chunk_width(int n)
{
    int cw = 64;
    while (cw > 4 && (size_t)n * cw * sizeof(cdbl) > 128 * 1024) cw /= 2;
    return cw;
}

hipError_t launch_rotate(int batch, const double *q, const double *B, const double *Bint,
                         const double *p, double *Br, double *pr, hipStream_t s)
{
    { ProfScope ps_("k_rotate", s);
    hipLaunchKernelGGL(k_rotate, dim3(cdiv(batch, 256)), dim3(256), 0, s, batch, q, B, Bint, p, Br, pr);
    }
    return hipGetLastError();
}

hipError_t launch_build_h(int n, int batch, const MspSystem *systems, const MspTerm *terms,
                          const int32_t *sys_id, const double *Br, const double *base, int zeeman,
                          int lower, double *H, hipStream_t s)
{
    ProfScope ps_("k_build_h", s);
    if (base && zeeman) {
        const size_t cnt = lower ? (size_t)n * (n + 1) / 2 : (size_t)n * n;
        hipLaunchKernelGGL(k_build_h_base, dim3(cdiv(cnt, 256 * kBuildEPT), batch),
                           dim3(256), sizeof(unsigned long long) * n, s, n, systems, terms, sys_id,
                           Br, base, lower, H);
    } else {
        hipLaunchKernelGGL(k_build_h, dim3(cdiv((size_t)n * n, 256), batch), dim3(256), 0, s, n,
                           systems, terms, sys_id, Br, base, zeeman, lower, H);
    }
    return hipGetLastError();
}

hipError_t launch_eigh_jacobi(int n, int batch, const double *H, double *E, double *U,
                              int32_t *info, int max_sweeps, hipStream_t s)
{
    const int m = n + (n & 1);
    int work = (m / 2) * (m / 2);
    int nt = ((work + 63) / 64) * 64;
    if (nt < 64) nt = 64;
    if (nt > 1024) nt = 1024;
    size_t lds = jacobi_lds_bytes(n);
    hipError_t err = msp_set_lds((const void *)k_eigh_jacobi, lds);
    if (err != hipSuccess) return err;
    { ProfScope ps_("k_eigh_jacobi", s);
    hipLaunchKernelGGL(k_eigh_jacobi, dim3(batch), dim3(nt), lds, s, n, H, E, U, info, max_sweeps);
    }
    return hipGetLastError();
}

hipError_t msp_set_lds(const void *func, size_t bytes)
{
    struct Rec {
        const void *f;
        int dev;
        size_t bytes;
    };
    static std::vector<Rec> done;
    static std::mutex mu;
    int dev = 0;
    hipError_t err = hipGetDevice(&dev);
    if (err != hipSuccess) return err;
    std::lock_guard<std::mutex> lock(mu);
    for (Rec &r : done)
        if (r.f == func && r.dev == dev) {
            if (bytes <= r.bytes) return hipSuccess;
            err = hipFuncSetAttribute(func, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
            if (err == hipSuccess) r.bytes = bytes;
            return err;
        }
    err = hipFuncSetAttribute(func, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (err == hipSuccess) done.push_back({func, dev, bytes});
    return err;
}

// Second stream of each device and calling thread, for work that overlaps
// the divide and conquer (created on first use; default priority; nullptr if
// it cannot be made).  Per thread, so that one thread's msp_eigh never
// enqueues onto a stream that another thread's graph capture has joined.
static hipStream_t side_stream()
{
    thread_local hipStream_t streams[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    if (!streams[dev] && hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking) != hipSuccess)
        streams[dev] = nullptr;
    return streams[dev];
}

// The divide and conquer on d (E) and e (workspace), eigenvectors into the
// workspace (dc_result_ptr).
static hipError_t launch_stedc(int n, int batch, double *E, double *work, int32_t *info,
                               hipStream_t s)
{
    const int in_lds = (n <= 512) ? 1 : 0;
    // 512 < n <= 1024: LDS but for the rotation arrays (152 n bytes)
    const bool part = !in_lds && n <= 1024;
    const size_t lds2 = in_lds ? stedc_lds_bytes(n) : part ? stedc_lds2_bytes(n) : 0;
    // one 1024-thread workgroup per instance where the instance's LDS
    // arrays leave room for one workgroup per CU (n >= 466) or live in
    // global memory (n > 512): 256 threads would leave one wave per SIMD
    // (n = 512: 9.8 -> 5.2 ms, n = 1024: 64 -> 28 ms per 140 instances)
    const bool big = !in_lds || 2 * lds2 > 160 * 1024;
    const void *fn = part    ? (const void *)k_stedc<1024, 2>
                     : !in_lds ? (const void *)k_stedc<1024, 0>
                     : big   ? (const void *)k_stedc<1024, 1>
                             : (const void *)k_stedc<256, 1>;
    hipError_t err = msp_set_lds(fn, lds2);
    if (err != hipSuccess) return err;
    ProfScope ps_("k_stedc", s);
    if (part)
        hipLaunchKernelGGL((k_stedc<1024, 2>), dim3(batch), dim3(1024), lds2, s, n, E, work, info);
    else if (!in_lds)
        hipLaunchKernelGGL((k_stedc<1024, 0>), dim3(batch), dim3(1024), 0, s, n, E, work, info);
    else if (big)
        hipLaunchKernelGGL((k_stedc<1024, 1>), dim3(batch), dim3(1024), lds2, s, n, E, work, info);
    else
        hipLaunchKernelGGL((k_stedc<256, 1>), dim3(batch), dim3(256), lds2, s, n, E, work, info);
    return hipGetLastError();
}

// The tridiagonal eigensolver on s and, when s2 is given, the WY factors on
// s2 (forked from s after the tridiagonalisation, joined back before the
// back-transformation); every exit after the fork joins s2 into s.
static hipError_t eigh_dc_middle(int n, int batch, double *H, double *E, int32_t *info,
                                 double *work, hipStream_t s, hipStream_t s2,
                                 hipEvent_t ev_td, hipEvent_t ev_ft)
{
    hipError_t err = hipSuccess;
    bool forked = false;
    if (s2) {
        if ((err = hipEventRecord(ev_td, s)) != hipSuccess) return err;
        if ((err = hipStreamWaitEvent(s2, ev_td, 0)) != hipSuccess) return err;
        forked = true;
        ProfScope ps_("k_larft", s2);
        err = launch_larft(n, batch, (const double *)H, work, s2);
    }
    if (err == hipSuccess) err = launch_stedc(n, batch, E, work, info, s);
    if (forked) {
        // join even after a failure, so that no capture ends with s2 unjoined
        hipError_t e1 = hipEventRecord(ev_ft, s2);
        hipError_t e2 = (e1 == hipSuccess) ? hipStreamWaitEvent(s, ev_ft, 0) : e1;
        if (err == hipSuccess) err = e2;
    }
    if (err == hipSuccess && !s2 && bt_nblocks(n) > 0) {
        ProfScope ps_("k_larft", s);
        err = launch_larft(n, batch, (const double *)H, work, s);
    }
    return err;
}

hipError_t launch_eigh_dc(int n, int batch, double *H, double *E, double *U, int32_t *info,
                          double *work, hipStream_t s, bool overlap)
{
    hipError_t err = hipSuccess;
    if (n <= 1024) {
        ProfScope ps_("k_tridiag", s);
        err = launch_tridiag_lp(n, batch, H, E, work, s);
    } else {
        size_t lds1 = sizeof(cdbl) * 2 * (size_t)n + 16 * sizeof(double);
        err = msp_set_lds((const void *)k_tridiag, lds1);
        if (err != hipSuccess) return err;
        ProfScope ps_("k_tridiag", s);
        hipLaunchKernelGGL(k_tridiag, dim3(batch), dim3(256), lds1, s, n, H, E, work);
        err = hipGetLastError();
    }
    if (err != hipSuccess) return err;
    if ((err = hipGetLastError()) != hipSuccess) return err;
    return launch_eigh_from_tridiag(n, batch, H, E, U, info, work, s, overlap);
}

size_t dc2_workspace_doubles(int n) { return dc_work_doubles(n) + chase_doubles(n); }

// Two-stage path: band and chase (d, e, the chase reflectors after the
// per-instance arrays), divide and conquer, Z' = Q2 Z, then U = Q1 Z' with the
// panel reflectors (unit element 16 rows below their column; the band stage
// has already written their V T blocks, so no WY-factor kernel runs).
hipError_t launch_eigh_dc2(int n, int batch, double *H, double *E, double *U, int32_t *info,
                           double *work, hipStream_t s)
{
    if (n < kSbrMin || n > kSbrMax) return hipErrorNotSupported;
    double *chase = work + (size_t)batch * dc_work_doubles(n);
    hipError_t err = launch_sbr(n, batch, H, E, work, chase, s);
    if (err != hipSuccess) return err;
    if ((err = launch_stedc(n, batch, E, work, info, s)) != hipSuccess) return err;
    if ((err = launch_chase_apply(n, batch, work, chase, s)) != hipSuccess) return err;
    ProfScope ps_("k_backtransform", s);
    return launch_backtransform_mfma(n, batch, (const double *)H, work, U, s, 16, n - 17, 1);
}

hipError_t launch_eigh_from_tridiag(int n, int batch, const double *R, double *E, double *U,
                                    int32_t *info, double *work, hipStream_t s, bool overlap)
{
    double *H = (double *)R;     // read only from here on
    hipError_t err = hipSuccess;
    // The WY factors need only the reflectors, so they are formed on the side
    // stream while the divide and conquer runs (it depends on d and e only);
    // the back-transformation joins both.  Events are per call, so
    // concurrent callers on other streams keep their own dependencies.
    hipStream_t s2 = (overlap && bt_nblocks(n) > 0) ? side_stream() : nullptr;
    hipEvent_t ev_td = nullptr, ev_ft = nullptr;
    if (s2 && (hipEventCreateWithFlags(&ev_td, hipEventDisableTiming) != hipSuccess ||
               hipEventCreateWithFlags(&ev_ft, hipEventDisableTiming) != hipSuccess)) {
        if (ev_td) (void)hipEventDestroy(ev_td);
        ev_td = nullptr;
        s2 = nullptr;
    }
    err = eigh_dc_middle(n, batch, H, E, info, work, s, s2, ev_td, ev_ft);
    if (ev_td) (void)hipEventDestroy(ev_td);
    if (ev_ft) (void)hipEventDestroy(ev_ft);
    if (err != hipSuccess) return err;
    if ((err = hipGetLastError()) != hipSuccess) return err;
    if (n <= 1024) {
        ProfScope ps_("k_backtransform", s);
        return launch_backtransform_mfma(n, batch, (const double *)H, work, U, s);
    }
    const int cw = bt2_chunk_width(n);
    size_t lds3 = sizeof(cdbl) * ((size_t)n * (cw + 1) + 2 * MSP_BT_NB * cw);
    err = msp_set_lds((const void *)k_backtransform2, lds3);
    if (err != hipSuccess) return err;
    { ProfScope ps_("k_backtransform", s);
    hipLaunchKernelGGL(k_backtransform2, dim3(cdiv(n, cw), batch), dim3(256), lds3, s, n,
                       (const double *)H, work, U, cw);
    }
    return hipGetLastError();
}

hipError_t launch_apply_muon(int n, int batch, const MspSystem *systems, const int32_t *sys_id,
                             const double *vec, int pauli, const double *U, double *Y,
                             hipStream_t s)
{
    { ProfScope ps_("k_apply_muon", s);
    hipLaunchKernelGGL(k_apply_muon, dim3(cdiv((size_t)n * n, 256), batch), dim3(256), 0, s, n,
                       systems, sys_id, vec, pauli, U, Y);
    }
    return hipGetLastError();
}

hipError_t launch_mode_product(int n, int batch, int dk, int stride, int foff, int fstride,
                               const double *F, const double *X, double *Y, hipStream_t s)
{
    { ProfScope ps_("k_mode_product", s);
    hipLaunchKernelGGL(k_mode_product, dim3(cdiv((size_t)n * n, 256), batch), dim3(256), 0, s, n,
                       dk, stride, foff, fstride, F, X, Y);
    }
    return hipGetLastError();
}

hipError_t launch_gemm_cn(int n, int batch, const double *A, const double *B, double *C,
                          hipStream_t s)
{
    ProfScope ps_("k_gemm_cn", s);
    return launch_gemm_cn_mfma(n, batch, A, B, C, s);
}

hipError_t launch_gemm_cn_upper(int n, int batch, const double *A, const double *B, double *C,
                                hipStream_t s)
{
    ProfScope ps_("k_gemm_cn", s);
    return launch_gemm_cn_mfma(n, batch, A, B, C, s, true);
}

hipError_t launch_integral_obs(int n, int batch, const double *E, const double *O, double alpha,
                               double beta, double tau, double out_scale, const double *w,
                               double *out, hipStream_t s)
{
    ProfScope ps_("k_integral", s);
    hipLaunchKernelGGL(k_integral_obs, dim3(batch), dim3(256), 0, s, n, E, O, alpha, beta, tau,
                       out_scale, w, out);
    return hipGetLastError();
}

hipError_t launch_form_c(int n, int batch, const double *rho, const double *obs, int real_mode,
                         double alpha, double beta, double *C, hipStream_t s)
{
    { ProfScope ps_("k_form_c", s);
    hipLaunchKernelGGL(k_form_c, dim3(cdiv((size_t)n * n, 256), batch), dim3(256), 0, s, n, rho,
                       obs, real_mode, alpha, beta, C);
    }
    return hipGetLastError();
}

hipError_t launch_timesum(int n, int batch, int ntimes, const double *E, const double *C,
                          int real_mode, const double *times, const double *w, int out_complex,
                          double *out, hipStream_t s)
{
    dim3 grid(cdiv(ntimes, 64), batch);
    if (real_mode)
        { ProfScope ps_("k_timesum", s);
        hipLaunchKernelGGL(k_timesum<true>, grid, dim3(256), 0, s, n, ntimes, E, C, times, w,
                           out_complex, out);
        }
    else
        { ProfScope ps_("k_timesum", s);
        hipLaunchKernelGGL(k_timesum<false>, grid, dim3(256), 0, s, n, ntimes, E, C, times, w,
                           out_complex, out);
        }
    return hipGetLastError();
}

hipError_t launch_phase_evolve(int n, int ntimes, const double *E, const double *rho,
                              const double *times, double *out, hipStream_t s)
{
    ProfScope ps_("k_phase_evolve", s);
    hipLaunchKernelGGL(k_phase_evolve, dim3(cdiv((size_t)n * n, 256), ntimes), dim3(256), 0, s, n,
                       ntimes, E, rho, times, out);
    return hipGetLastError();
}

hipError_t launch_integral(int n, int batch, const double *E, const double *C, int real_mode,
                           double tau, double out_scale, const double *w, int out_complex,
                           double *out, hipStream_t s)
{
    { ProfScope ps_("k_integral", s);
    hipLaunchKernelGGL(k_integral, dim3(batch), dim3(256), 0, s, n, E, C, real_mode, tau,
                       out_scale, w, out_complex, out);
    }
    return hipGetLastError();
}

hipError_t launch_accumulate(int nslots, int width, const int32_t *ptr, const int32_t *idx,
                             const double *rows, double scale, double *results, hipStream_t s)
{
    { ProfScope ps_("k_accumulate", s);
    hipLaunchKernelGGL(k_accumulate, dim3(cdiv((size_t)nslots * width, 64)), dim3(64 * kAccParts), 0,
                       s, nslots, width, ptr, idx, rows, scale, results);
    }
    return hipGetLastError();
}
