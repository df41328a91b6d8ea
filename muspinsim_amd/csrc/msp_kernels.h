// Host-side launchers of the kernels in msp_kernels.hip (no argument checks;
// msp_capi.hip validates before calling).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

struct MspSystem;
struct MspTerm;

// While profiling is enabled (msp_profile_enable), brackets the launches in
// its scope with two hipEvents on stream s, reported under `name`.
struct ProfScope {
    const char *name;
    hipStream_t s;
    hipEvent_t a = nullptr;
    ProfScope(const char *n, hipStream_t st);
    ~ProfScope();
};

// Raise a kernel's dynamic LDS limit to `bytes` unless an earlier call on
// this device already did (the attribute is not a stream operation; after a
// first eager run, launches under stream capture make no such call).
hipError_t msp_set_lds(const void *func, size_t bytes);

size_t jacobi_lds_bytes(int n);
size_t dc_workspace_doubles(int n);
int dc_chunk_width(int n);

hipError_t launch_rotate(int batch, const double *q, const double *B, const double *Bint,
                         const double *p, double *Br, double *pr, hipStream_t s);
hipError_t launch_build_h(int n, int batch, const MspSystem *systems, const MspTerm *terms,
                          const int32_t *sys_id, const double *Br, const double *base, int zeeman,
                          int lower, double *H, hipStream_t s);
hipError_t launch_eigh_jacobi(int n, int batch, const double *H, double *E, double *U,
                              int32_t *info, int max_sweeps, hipStream_t s);
// overlap: form the WY factors on a per-thread side stream while the divide
// and conquer runs (forked from and joined back into s); false keeps every
// launch on s and creates no events.
hipError_t launch_eigh_dc(int n, int batch, double *H, double *E, double *U, int32_t *info,
                          double *work, hipStream_t s, bool overlap = true);
// The divide and conquer and the back-transformation after a tridiagonal
// reduction: d in E, e and tau in the workspace, reflector tails below the
// subdiagonal of R (row-major n x n per instance, read only).
hipError_t launch_eigh_from_tridiag(int n, int batch, const double *R, double *E, double *U,
                                    int32_t *info, double *work, hipStream_t s, bool overlap = true);
// FP32 LDS-resident tridiagonalisation (n <= 192) with the same outputs as
// the FP64 one: reflector tails to R, d to E, e and tau to the workspace.
// Time sum of a real symmetric C on FP32 matrix cores: phase table (FP64
// range reduction, FP32 sin/cos) in `work`, then Y = C [cos | sin] and the
// contraction with the phases per 64-time tile.
int timesum_mfma_ldt(int ntimes);
size_t timesum_mfma_workspace(int n, int batch, int ntimes);
size_t timesum_mfma64_workspace(int n, int batch, int ntimes);
hipError_t launch_timesum_mfma64(int n, int batch, int ntimes, const double *E, const double *C,
                                 const double *times, const double *w, double *out, void *work,
                                 hipStream_t s);
hipError_t launch_timesum_mfma(int n, int batch, int ntimes, const double *E, const double *C,
                               const double *times, const double *w, double *out, void *work,
                               hipStream_t s);
size_t tridiag32_lds_bytes(int n);
hipError_t launch_tridiag32(int n, int batch, const double *H, double *R, double *E, double *work,
                            hipStream_t s);
hipError_t launch_apply_muon(int n, int batch, const MspSystem *systems, const int32_t *sys_id,
                             const double *vec, int pauli, const double *U, double *Y,
                             hipStream_t s);
hipError_t launch_mode_product(int n, int batch, int dk, int stride, int foff, int fstride,
                               const double *F, const double *X, double *Y, hipStream_t s);
hipError_t launch_gemm_cn(int n, int batch, const double *A, const double *B, double *C,
                          hipStream_t s);
hipError_t launch_form_c(int n, int batch, const double *rho, const double *obs, int real_mode,
                         double alpha, double beta, double *C, hipStream_t s);
hipError_t launch_timesum(int n, int batch, int ntimes, const double *E, const double *C,
                          int real_mode, const double *times, const double *w, int out_complex,
                          double *out, hipStream_t s);
hipError_t launch_phase_evolve(int n, int ntimes, const double *E, const double *rho,
                              const double *times, double *out, hipStream_t s);
hipError_t launch_integral(int n, int batch, const double *E, const double *C, int real_mode,
                           double tau, double out_scale, const double *w, int out_complex,
                           double *out, hipStream_t s);
hipError_t launch_accumulate(int nslots, int width, const int32_t *ptr, const int32_t *idx,
                             const double *rows, double scale, double *results, hipStream_t s);

void prof_enable(int on);
std::string prof_report();
hipError_t launch_tridiag_lp(int n, int batch, double *H, double *E, double *work, hipStream_t s);
// Register-resident reduction (matrix in VGPRs, one workgroup of 8 waves per
// matrix) for 2 <= n <= 64, same outputs as launch_tridiag_lp.
size_t tridiag_rs_lds_bytes(int n);
hipError_t launch_tridiag_rs(int n, int batch, double *H, double *E, double *work, hipStream_t s);
// U = Q Z on MFMA.  off: rows between a reflector's column and its unit
// element (1 for the one-stage reduction, 16 for the band stage's panels);
// nref reflectors; zc: start from the complex Z' = Q2 Z of the two-stage
// reduction (dc_zc_ptr) instead of the divide and conquer's real Z.
hipError_t launch_backtransform_mfma(int n, int batch, const double *H, double *work, double *U,
                                     hipStream_t s, int off = 1, int nref = -1, int zc = 0);

// WY factors (T blocks) and V T of the one-stage reflectors on MFMA, reading
// the reflectors from H directly (no LDS staging).
hipError_t launch_larft_mfma(int n, int batch, const double *H, double *work, hipStream_t s);

// Two-stage reduction (msp_sbr.hip) for kSbrMin <= n <= kSbrMax: dense ->
// band of 16 (panel reflectors, tau and the V T blocks into the workspace as
// the one-stage reduction leaves them, unit element 16 rows down), then band ->
// real tridiagonal (d to E, e to the workspace, the chase reflectors to
// `chase`: chase_doubles(n) per instance).
constexpr int kSbrMin = 34;
constexpr int kSbrMax = 256;
hipError_t launch_sbr(int n, int batch, double *H, double *E, double *work, double *chase,
                      hipStream_t s);
// Z' = Q2 Z (the chase reflectors in reverse order) into each instance's
// dc_zc_ptr (complex, column-major).
hipError_t launch_chase_apply(int n, int batch, double *work, const double *chase, hipStream_t s);
// msp_eigh with the two-stage reduction: workspace doubles per instance and
// the whole pipeline (band, chase, divide and conquer, Q2, Q1).
size_t dc2_workspace_doubles(int n);
hipError_t launch_eigh_dc2(int n, int batch, double *H, double *E, double *U, int32_t *info,
                           double *work, hipStream_t s);

// Lindbladian path (msp_lindblad.hip)
size_t lindblad_modes_workspace(int N, int batch);
hipError_t launch_liouvillian(int n, int batch, const double *H, int n_jump, const double *A,
                              const double *rate, double *L, hipStream_t s);
hipError_t launch_lindblad_modes(int N, int batch, double *L, const double *r, int n_ops,
                                 const double *u, double *lam, double *amp, double *c, double *V,
                                 int32_t *info, double *work, int max_its, hipStream_t s);
hipError_t launch_mode_sum(int N, int batch, int n_ops, int ntimes, const double *lam,
                           const double *amp, const double *times, const double *w,
                           int out_complex, double *out, hipStream_t s);
hipError_t launch_mode_integral(int N, int batch, int n_ops, const double *lam, const double *amp,
                                double tau, double out_scale, const double *w, int out_complex,
                                double *out, hipStream_t s);
hipError_t launch_mode_states(int N, int ntimes, const double *lam, const double *c,
                              const double *V, const double *times, double *out, hipStream_t s);

// Adjoint of the fast-path time sum (msp_adjoint.hip)
hipError_t launch_adjoint_freq(int n, int batch, int ntimes, const double *E, const double *times,
                               const double *r, const int32_t *slot, const double *w, double scale,
                               double *R, double *S, hipStream_t s);
hipError_t launch_hadamard_real(int n, int batch, const double *R, const double *O, double *M,
                                hipStream_t s);
hipError_t launch_adjoint_gprime(int n, int batch, const double *E, const double *O,
                                 const double *MO, const double *OM, const double *S, double d,
                                 double degtol, double *G, hipStream_t s);
hipError_t launch_batch_sum(int len, int batch, const double *X, double *out, hipStream_t s);
hipError_t launch_trace_products(int n, int nparam, const double *G, const double *dH, double *out,
                                 hipStream_t s);

// Batched complex C = A^H B on FP64 matrix cores (msp_gemm_mfma.hip)
hipError_t launch_gemm_cn_mfma(int n, int batch, const double *A, const double *B, double *C,
                               hipStream_t s, bool upper = false);

// Celio/Trotter evolution (msp_celio.hip)
int celio_block_threads(int n);
size_t celio_workspace_bytes(int n, int nstates);
hipError_t launch_celio_evolve(int n, int n_spins, const int32_t *dims, int nstates, int ntimes,
                               int k, const int32_t *state_cfg, const int32_t *cfg_steps,
                               const int32_t *desc, const int32_t *meas, const double *mats,
                               const int32_t *tables, int init_mode, const double *init_vec,
                               const int32_t *init_idx, const double *phases, double *out,
                               double *work, hipStream_t s);

// Hermitian products on the upper block triangle and the decaying integral
// read from them (msp_kernels.hip)
hipError_t launch_gemm_cn_upper(int n, int batch, const double *A, const double *B, double *C,
                                hipStream_t s);
hipError_t launch_integral_obs(int n, int batch, const double *E, const double *O, double alpha,
                               double beta, double tau, double out_scale, const double *w,
                               double *out, hipStream_t s);
hipError_t launch_gemm_muon_upper(int n, int batch, const MspSystem *systems,
                                  const int32_t *sys_id, const double *vec, int pauli,
                                  const double *U, double *O, hipStream_t s);
