// Celio's Trotter evolution of state vectors (muspinsim/celio.py:210-476,
// cpp/celio.cpp:23-68, cpp/parallel.cpp:130-266).
//
// One workgroup owns one state vector for the whole time loop.  For
// n <= 8192 the state lives in LDS (128 KB at n = 8192); above that it lives
// in a per-state slice of the workspace.  Each time step measures
// <psi|O|psi> for the muon operator O and then applies k times the
// sequence of Trotter factors c_1..c_m, each a small m x m matrix acting on
// one or two spins: (c (x) 1) psi over the "fibers" of the other spins.
// The fiber base offsets and the element offsets inside a fiber come from
// host-built tables, so a factor on any spins (any order) is one gather /
// small matrix product / scatter.
//
// Factors with m in {2..10, 12, 14, 16} use one thread per fiber with the
// fiber in registers (it reads and writes only its own elements, so no
// barrier is needed inside a factor).  Other m <= 64 use one lane per output row,
// the m lanes of a fiber inside one wavefront: every lane issues all its
// reads before its single write, and a wavefront executes in lockstep, so
// no lane overwrites an element another lane of its fiber has yet to read.
// A block barrier separates consecutive factors.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "msp_kernels.h"

namespace {

constexpr int kLdsStateMax = 8192;
constexpr int kRedSlots = 8;    // reduction slots (one per wavefront, 512 threads)

struct CelioDims {
    int n_spins;
    int d[16];
};

__device__ inline double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// <x|O|x> for an operator on one spin (M = its dimension), this thread's part
template <int M>
__device__ __attribute__((always_inline)) double measure_part(const double2 *psi, int n, int4 md, const double2 *__restrict__ mats,
                               const int *__restrict__ tables)
{
    const int nf = n / M;
    const double2 *O = mats + md.y;
    const int *fib = tables + md.z;
    const int *el = tables + md.w;
    double acc = 0.0;
    for (int f = threadIdx.x; f < nf; f += blockDim.x) {
        const int base = fib[f];
        double2 x[M];
#pragma unroll
        for (int j = 0; j < M; ++j) x[j] = psi[base + el[j]];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            double yr = 0.0, yi = 0.0;
#pragma unroll
            for (int j = 0; j < M; ++j) {
                const double2 a = O[i * M + j];
                yr = fma(a.x, x[j].x, fma(-a.y, x[j].y, yr));
                yi = fma(a.x, x[j].y, fma(a.y, x[j].x, yi));
            }
            // Re(conj(x_i) y_i)
            acc = fma(x[i].x, yr, fma(x[i].y, yi, acc));
        }
    }
    return acc;
}

// one thread per fiber, the fiber in registers (compile-time M)
template <int M>
__device__ __attribute__((always_inline)) void apply_reg(double2 *psi, int n, int4 d,
                                                         const double2 *Mx,
                                                         const int *__restrict__ tables)
{
    const int nf = n / M;
    const int *fib = tables + d.z;
    const int *el = tables + d.w;
    int off[M];
#pragma unroll
    for (int j = 0; j < M; ++j) off[j] = el[j];
    for (int f = threadIdx.x; f < nf; f += blockDim.x) {
        // keeps the matrix reads inside the loop: hoisted, its M^2 entries
        // would be held in registers across fibers
        __asm__ volatile("" ::: "memory");
        const int base = fib[f];
        double2 x[M];
#pragma unroll
        for (int j = 0; j < M; ++j) x[j] = psi[base + off[j]];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            double yr = 0.0, yi = 0.0;
#pragma unroll
            for (int j = 0; j < M; ++j) {
                const double2 a = Mx[i * M + j];
                yr = fma(a.x, x[j].x, fma(-a.y, x[j].y, yr));
                yi = fma(a.x, x[j].y, fma(a.y, x[j].x, yi));
            }
            psi[base + off[i]] = make_double2(yr, yi);
        }
    }
}

__device__ __attribute__((always_inline)) void apply_rows(double2 *psi, int n, int4 d, const double2 *__restrict__ mats,
                           const int *__restrict__ tables)
{
    const int m = d.x;
    const int nf = n / m;
    const double2 *M = mats + d.y;
    const int *fib = tables + d.z;
    const int *el = tables + d.w;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    const int fpw = 64 / m;                      // whole fibers per wavefront
    if (fpw == 0) return;                        // m > 64 is rejected by the host
    const int fl = lane / m;
    const int row = lane - fl * m;
    for (int g = wave; g * fpw < nf; g += nwaves) {
        const int f = g * fpw + fl;
        const bool ok = fl < fpw && f < nf;
        double yr = 0.0, yi = 0.0;
        int base = 0;
        if (ok) {
            base = fib[f];
            for (int j = 0; j < m; ++j) {
                const double2 a = M[row * m + j];
                const double2 x = psi[base + el[j]];
                yr = fma(a.x, x.x, fma(-a.y, x.y, yr));
                yi = fma(a.x, x.y, fma(a.y, x.x, yi));
            }
        }
        if (ok) psi[base + el[row]] = make_double2(yr, yi);
    }
}

__device__ __attribute__((always_inline)) double block_sum(double v, double *red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    __syncthreads();
    return s;
}

template <bool GLOBAL>
__global__ __launch_bounds__(512) void k_celio_evolve(
    int n, int ntimes, int k, CelioDims dims, const int *__restrict__ state_cfg,
    const int *__restrict__ cfg_steps, const int4 *__restrict__ desc,
    const int4 *__restrict__ meas, const double2 *__restrict__ mats,
    const int *__restrict__ tables, int init_mode, const double2 *__restrict__ init_vec,
    const int *__restrict__ init_idx, const double *__restrict__ phases,
    double *__restrict__ out, double2 *__restrict__ gstate)
{
    extern __shared__ double2 lds[];
    const int s = blockIdx.x;
    double2 *psi = GLOBAL ? gstate + (size_t)s * n : lds;
    double *red = GLOBAL ? (double *)lds : (double *)(lds + n);
    const int cfg = state_cfg[s];

    if (init_mode == 0) {
        // muon first (dimension 2): psi = mu (x) exp(2 pi i phi) / sqrt(n/2)
        const int half = n / 2;
        const double2 *mu = init_vec + init_idx[s];
        const double *ph = phases + (size_t)s * half;
        const double sc = 1.0 / sqrt((double)half);
        for (int I = threadIdx.x; I < n; I += blockDim.x) {
            const int j = I >= half ? 1 : 0;
            double sn, cs;
            sincospi(2.0 * ph[I - j * half], &sn, &cs);
            const double2 a = mu[j];
            psi[I] = make_double2((a.x * cs - a.y * sn) * sc, (a.x * sn + a.y * cs) * sc);
        }
    } else if (init_mode == 2) {
        // explicit state vector
        const double2 *v = init_vec + init_idx[s];
        for (int I = threadIdx.x; I < n; I += blockDim.x) psi[I] = v[I];
    } else {
        // product state: psi[I] = prod_q v_q[digit_q(I)], spin 0 slowest
        const int *idx = init_idx + (size_t)s * dims.n_spins;
        for (int I = threadIdx.x; I < n; I += blockDim.x) {
            int rem = I;
            double2 acc = make_double2(1.0, 0.0);
            for (int q = dims.n_spins - 1; q >= 0; --q) {
                const int dq = dims.d[q];
                const int dig = rem % dq;
                rem /= dq;
                acc = cmul(acc, init_vec[idx[q] + dig]);
            }
            psi[I] = acc;
        }
    }
    __syncthreads();

    const int first = cfg_steps[2 * cfg], count = cfg_steps[2 * cfg + 1];
    const int4 md = meas[cfg];
    for (int t = 0; t < ntimes; ++t) {
        double part;
        switch (md.x) {
        case 3: part = measure_part<3>(psi, n, md, mats, tables); break;
        case 4: part = measure_part<4>(psi, n, md, mats, tables); break;
        case 5: part = measure_part<5>(psi, n, md, mats, tables); break;
        case 6: part = measure_part<6>(psi, n, md, mats, tables); break;
        case 7: part = measure_part<7>(psi, n, md, mats, tables); break;
        case 8: part = measure_part<8>(psi, n, md, mats, tables); break;
        case 9: part = measure_part<9>(psi, n, md, mats, tables); break;
        case 10: part = measure_part<10>(psi, n, md, mats, tables); break;
        default: part = measure_part<2>(psi, n, md, mats, tables); break;
        }
        const double v = block_sum(part, red);
        if (threadIdx.x == 0) out[(size_t)s * ntimes + t] = v;
        if (t == ntimes - 1) break;
        for (int r = 0; r < k; ++r)
            for (int c = 0; c < count; ++c) {
                // the descriptor and the matrix address are uniform, so the
                // matrix entries are read through the scalar cache
                const int4 d = desc[first + c];
                const double2 *cur = mats + d.y;
                switch (d.x) {
                case 2: apply_reg<2>(psi, n, d, cur, tables); break;
                case 3: apply_reg<3>(psi, n, d, cur, tables); break;
                case 4: apply_reg<4>(psi, n, d, cur, tables); break;
                case 5: apply_reg<5>(psi, n, d, cur, tables); break;
                case 6: apply_reg<6>(psi, n, d, cur, tables); break;
                case 7: apply_reg<7>(psi, n, d, cur, tables); break;
                case 8: apply_reg<8>(psi, n, d, cur, tables); break;
                case 9: apply_reg<9>(psi, n, d, cur, tables); break;
                case 10: apply_reg<10>(psi, n, d, cur, tables); break;
                case 12: apply_reg<12>(psi, n, d, cur, tables); break;
                case 14: apply_reg<14>(psi, n, d, cur, tables); break;
                case 16: apply_reg<16>(psi, n, d, cur, tables); break;
                default: apply_rows(psi, n, d, mats, tables); break;
                }
                __syncthreads();
            }
    }
}

}  // namespace

int celio_block_threads(int n)
{
    if (n >= 1024) return 512;
    const int t = ((n / 2 + 63) / 64) * 64;
    return t < 64 ? 64 : t;
}

size_t celio_workspace_bytes(int n, int nstates)
{
    return n > kLdsStateMax ? (size_t)nstates * n * 16 : 0;
}

hipError_t launch_celio_evolve(int n, int n_spins, const int32_t *dims, int nstates, int ntimes,
                               int k, const int32_t *state_cfg, const int32_t *cfg_steps,
                               const int32_t *desc, const int32_t *meas, const double *mats,
                               const int32_t *tables, int init_mode, const double *init_vec,
                               const int32_t *init_idx, const double *phases, double *out,
                               double *work, hipStream_t s)
{
    ProfScope prof("k_celio_evolve", s);
    CelioDims cd{};
    cd.n_spins = n_spins;
    for (int q = 0; q < n_spins; ++q) cd.d[q] = dims[q];
    const int bs = celio_block_threads(n);
    const size_t extra = kRedSlots * sizeof(double);     // reduction slots
    const bool global = n > kLdsStateMax;
    const size_t lds = global ? extra : (size_t)n * 16 + extra;
    const void *f = global ? (const void *)k_celio_evolve<true> : (const void *)k_celio_evolve<false>;
    hipError_t e = msp_set_lds(f, lds);
    if (e != hipSuccess) return e;
    const int4 *dsc = reinterpret_cast<const int4 *>(desc);
    const int4 *msr = reinterpret_cast<const int4 *>(meas);
    const double2 *mt = reinterpret_cast<const double2 *>(mats);
    const double2 *iv = reinterpret_cast<const double2 *>(init_vec);
    double2 *gs = reinterpret_cast<double2 *>(work);
    if (global)
        hipLaunchKernelGGL(k_celio_evolve<true>, dim3(nstates), dim3(bs), lds, s, n, ntimes, k, cd,
                           state_cfg, cfg_steps, dsc, msr, mt, tables, init_mode, iv, init_idx,
                           phases, out, gs);
    else
        hipLaunchKernelGGL(k_celio_evolve<false>, dim3(nstates), dim3(bs), lds, s, n, ntimes, k,
                           cd, state_cfg, cfg_steps, dsc, msr, mt, tables, init_mode, iv, init_idx,
                           phases, out, gs);
    return hipGetLastError();
}
