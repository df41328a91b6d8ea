/* muspin_hip.h -- C ABI of the MI355X (gfx950) spin-dynamics engine.
 *
 * One shared library, libmuspin_hip.so, exporting plain-pointer entry points
 * for the muspinsim hot path: per (orientation x field) instance, build H,
 * diagonalise it, change the density matrix and observable into the
 * eigenbasis, and evaluate the time sum / decaying integral with the
 * powder-weighted accumulation fused into the last stage.
 *
 * Conventions
 *   - Every pointer argument named d_* is a DEVICE pointer (HIP global
 *     memory, e.g. torch.Tensor.data_ptr()); h_* pointers are host memory.
 *   - Complex numbers are interleaved (re, im) doubles: numpy complex128.
 *   - Matrices are dense, row-major, n x n, batched with stride n*n.
 *   - Calls are asynchronous on `stream` (a hipStream_t; 0 = default
 *     stream) and return 0 on success or an MSP_ERR_* code; the message of
 *     the last failure on the calling thread is read with msp_last_error.
 *   - The library allocates no device memory: scratch space comes from the
 *     caller through d_work, sized by the matching *_workspace function.
 *
 * Reference interfaces replaced (paths in the muspinsim tree, v2.3.1):
 *   msp_rotate_fields   ExperimentRunner.load_config rotations, experiment.py:384-411
 *   msp_build_h         SpinSystem.hamiltonian + ExperimentRunner.Hz/Htot,
 *   msp_build_h_lower   spinsys.py:34-69,613-626; experiment.py:238-269,331-341
 *   msp_eigh            Hermitian.diag -> numpy.linalg.eigh, spinop.py:51-82
 *   msp_apply_muon_op   MuonSpinSystem.muon_operator / sigma_mu (x) 1,
 *                       spinsys.py:707-760; hamiltonian.py:204
 *   msp_apply_kron      DensityOperator kron of per-spin states, experiment.py:169-236
 *   msp_gemm_cn         Operator.basis_change (U^H M U), spinop.py:330-355
 *   msp_gemm_cn_upper,  the same basis change of sigma_mu on the integral path, blocks
 *   msp_gemm_muon_upper on or above the diagonal only, experiment.py:486-496
 *   msp_integral_obs    integrate_decaying of |O_ij|^2 (fast_evolve's C), hamiltonian.py:150-162
 *   msp_form_c          rho' * O'^T of Hamiltonian.evolve, hamiltonian.py:77-103;
 *                       |U^H sigma U|^2 of fast_evolve, hamiltonian.py:207-210
 *   msp_timesum         evolve time loop, hamiltonian.py:87-107, and the
 *                       cosine sum parallel_fast_time_evolve, cython/parallel.pyx:16-69
 *   msp_integral        Hamiltonian.integrate_decaying, hamiltonian.py:150-162
 *   msp_phase_evolve    evolve without operators (density matrices), hamiltonian.py:108-116
 *   msp_liouvillian     Lindbladian.from_hamiltonian + add_dissipative_term,
 *                       lindbladian.py:19-41; experiment.py:271-341
 *   msp_lindblad_modes  Lindbladian.evolve / integrate_decaying eigendecomposition
 *                       (numpy.linalg.eig + solve), lindbladian.py:83-97,145-160
 *   msp_mode_sum        Lindbladian.evolve expectation values, lindbladian.py:99-103
 *   msp_mode_integral   Lindbladian.integrate_decaying, lindbladian.py:163-171
 *   msp_mode_states     Lindbladian.evolve density matrices, lindbladian.py:105-109
 *   msp_adjoint_freq, msp_hadamard_real, msp_adjoint_gprime, msp_batch_sum,
 *   msp_trace_products  adjoint gradient of the fast-path polarisation for fitting; no
 *                       counterpart (FittingRunner differentiates by finite differences,
 *                       fitting.py:99-105)
 *   msp_accumulate      Re(data)*w and store_time_slice powder accumulation,
 *                       experiment.py:498; simconfig.py:347-368
 *   msp_celio_evolve    CelioHamiltonian.evolve / fast_evolve Trotter loop,
 *                       celio.py:210-293,318-476
 *   msp_graph_*, msp_profile_*  no counterpart: HIP graph capture and kernel timing
 */
#ifndef MUSPIN_HIP_H
#define MUSPIN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MSP_OK 0
#define MSP_ERR_VALUE 1       /* invalid argument      -> ValueError        */
#define MSP_ERR_NOCONV 2      /* solver not converged  -> LinAlgError       */
#define MSP_ERR_HIP 3         /* HIP runtime failure   -> RuntimeError      */
#define MSP_ERR_UNSUPPORTED 4 /* size/feature limit    -> NotImplementedError */

#define MSP_MAX_SPINS 16

/* One spin system: spins in Kronecker order (first = slowest index). */
typedef struct MspSystem {
    int32_t n_spins;
    int32_t n;            /* Hilbert dimension = prod(dims) */
    int32_t muon_index;
    int32_t term_offset;  /* first term in the term table */
    int32_t n_terms;
    int32_t pad[3];
    int32_t dims[MSP_MAX_SPINS];    /* 2I+1 per spin */
    double gammas[MSP_MAX_SPINS];   /* MHz/T */
} MspSystem;

/* One interaction term: kind 0 scalar (t[0]), 1 linear t[0..2].S_i,
 * 2 bilinear S_i.T.S_j with T = t[0..8] row-major (i == j: same-spin
 * product S_i^a S_i^b, as for quadrupolar couplings).  MHz. */
typedef struct MspTerm {
    int32_t kind;
    int32_t i;
    int32_t j;
    int32_t pad;
    double t[9];
} MspTerm;

/* Eigensolver selection for msp_eigh. */
#define MSP_EIGH_AUTO 0
#define MSP_EIGH_JACOBI 1   /* LDS-resident cyclic Jacobi, n <= 64 */
#define MSP_EIGH_DC 2       /* Householder + divide and conquer, any n */
#define MSP_EIGH_DC2 3      /* two-stage reduction (band of 16 on FP64 matrix cores,
                               then bulge chasing) + divide and conquer, 34 <= n <= 256 */

int msp_init(void);
int msp_last_error(char *h_buf, size_t len);
int msp_version(void);

/* B' = q.B + B_int and p' = q.p for each instance (q stored conjugated,
 * [w, x, y, z]); arrays are batch x 4 / batch x 3 doubles. */
int msp_rotate_fields(int batch, const double *d_q, const double *d_B, const double *d_Bint,
                      const double *d_p, double *d_Br, double *d_pr, void *stream);

/* H[b] = base[sys[b]] (if d_base != NULL) or the sum of the system's terms,
 * plus the Zeeman term sum_i gamma_i Br[b].S_i.  All instances share n.
 * zeeman = 0 skips the Zeeman part (used to build the base matrices). */
int msp_build_h(int n, int batch, const MspSystem *d_systems, const MspTerm *d_terms,
                const int32_t *d_sys_id, const double *d_Br, const double *d_base,
                int zeeman, double *d_H, void *stream);

/* msp_build_h writing only the lower triangle (column <= row) of each H[b];
 * the strict upper triangle is left as it was.  msp_eigh with MSP_EIGH_DC
 * (or AUTO resolving to it) and n <= 1024 reads only that triangle, so this
 * halves the matrix writes in front of it (same reference interfaces as
 * msp_build_h). */
int msp_build_h_lower(int n, int batch, const MspSystem *d_systems, const MspTerm *d_terms,
                      const int32_t *d_sys_id, const double *d_Br, const double *d_base,
                      int zeeman, double *d_H, void *stream);

/* Batched Hermitian eigendecomposition.  d_H is destroyed.  d_E: batch x n
 * ascending eigenvalues; d_U: batch x n x n eigenvectors (columns).
 * d_info: batch int32, 0 = converged. */
size_t msp_eigh_workspace(int n, int batch, int method);
int msp_eigh(int n, int batch, double *d_H, double *d_E, double *d_U, int32_t *d_info,
             int method, void *d_work, size_t work_bytes, void *stream);

/* First stage of the mixed-precision eigendecomposition (replaces the
 * numpy.linalg.eigh of Hermitian.diag, spinop.py:51-82, on the mixed path;
 * 2 <= n <= 192): Householder tridiagonalisation in FP32 with the whole lower
 * triangle of each H resident in LDS, then the divide and conquer and the
 * back-transformation of msp_eigh on its reflectors.  d_H (lower triangle
 * read) is preserved.  The eigenpairs carry FP32-level errors (backward error
 * of order 1e-6 ||H||); msp_refine_residual and msp_refine_correction bring
 * them to FP64 level. */
size_t msp_eigh_mixed_workspace(int n, int batch);
int msp_eigh_mixed(int n, int batch, const double *d_H, double *d_E, double *d_U, int32_t *d_info,
                   void *d_work, size_t work_bytes, void *stream);

/* msp_eigh for callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.  No
 * counterpart in the reference.  The divide-and-conquer workspace is a
 * library-owned static device array of MSP_NOALLOC_SCRATCH_BYTES, one per
 * device, held until the process ends; a call whose msp_eigh_workspace(n,
 * batch, MSP_EIGH_DC) exceeds it fails with MSP_ERR_VALUE (64 MiB: 34
 * instances at n = 192, one at n = 1024).  Calls no hipMalloc/hipFree, creates
 * no stream or event, and launches everything on `stream`.  Every call on a
 * device shares that device's array: calls that overlap on one device race,
 * so serialise them (one stream, or events).  Same results as msp_eigh. */
#define MSP_NOALLOC_SCRATCH_BYTES (64u << 20)
int msp_eigh_noalloc(int n, int batch, double *d_H, double *d_E, double *d_U, int32_t *d_info,
                     int method, void *stream);

/* Y[b] = (v[b].S_mu (x) 1) U[b] with v = d_vec (batch x 3) and the muon
 * position/dims of d_systems[d_sys_id[b]].  pauli = 1 uses sigma = 2S. */
int msp_apply_muon_op(int n, int batch, const MspSystem *d_systems, const int32_t *d_sys_id,
                      const double *d_vec, int pauli, const double *d_U, double *d_Y,
                      void *stream);

/* Y[b] = (F_0 (x) F_1 (x) ... (x) F_{k-1}) U[b] with per-instance complex
 * factors d_F (batch x sum_k dims[k]^2, row-major per factor); d_tmp is
 * scratch of the size of d_Y. */
int msp_apply_kron(int n, int batch, int n_spins, const int32_t *h_dims, const double *d_F,
                   const double *d_U, double *d_Y, double *d_tmp, void *stream);

/* C[b] = A[b]^H B[b] (complex n x n). */
int msp_gemm_cn(int n, int batch, const double *d_A, const double *d_B, double *d_C,
                void *stream);

/* C[b] = A[b]^H B[b] on the 64 x 64 blocks on and above the block diagonal
 * only (the others are left unwritten): the basis change of a Hermitian
 * observable, O = U^H (M U), for consumers that read the upper blocks
 * (Operator.basis_change, spinop.py:330-355, as msp_gemm_cn). */
int msp_gemm_cn_upper(int n, int batch, const double *d_A, const double *d_B, double *d_C,
                      void *stream);

/* O[b] = U[b]^H (f v_b.S_mu (x) 1) U[b] on the blocks on and above the block
 * diagonal (as msp_gemm_cn_upper), with the muon operator applied while the
 * second operand is loaded: the product msp_apply_muon_op then msp_gemm_cn
 * form (experiment.py:486-496 via Operator.basis_change, spinop.py:330-355)
 * in one pass for a spin-1/2 muon; f = 2 (Pauli) when pauli != 0.  An
 * instance whose system's muon has another multiplicity gets NaN blocks
 * (use msp_apply_muon_op + msp_gemm_cn_upper for it). */
int msp_gemm_muon_upper(int n, int batch, const MspSystem *d_systems, const int32_t *d_sys_id,
                        const double *d_vec, int pauli, const double *d_U, double *d_O,
                        void *stream);

/* out[b] = w[b] * out_scale * Re sum_ij C_ij / (1/tau + 2 pi i (E_i - E_j)) with
 * C_ij = alpha |O_ij|^2 + beta delta_ij Re O_ii, read from the upper blocks of a
 * Hermitian O as msp_gemm_cn_upper writes them (the blocks above the diagonal
 * count twice).  The fused form of msp_form_c + msp_integral for a real C
 * (Hamiltonian.integrate_decaying, hamiltonian.py:150-162). */
int msp_integral_obs(int n, int batch, const double *d_E, const double *d_O, double alpha,
                     double beta, double tau, double out_scale, const double *d_w, double *d_out,
                     void *stream);

/* Coefficient matrix of the time sum.
 * real_mode = 0: C_ij = rho'_ij O'_ji (complex, d_rho required).
 * real_mode = 1: C_ij = alpha |O'_ij|^2 + beta delta_ij Re O'_ii (real). */
int msp_form_c(int n, int batch, const double *d_rho, const double *d_obs, int real_mode,
               double alpha, double beta, double *d_C, void *stream);

/* P_b(t) = sum_ij C_ij exp(-2 pi i (E_i - E_j) t) for each time.
 * out_complex = 0: d_out[b*T + t] = w_b Re P_b(t) (w = 1 if d_w == NULL);
 * out_complex = 1: d_out is complex, unweighted. */
int msp_timesum(int n, int batch, int ntimes, const double *d_E, const double *d_C,
                int real_mode, const double *d_times, const double *d_w, int out_complex,
                double *d_out, void *stream);

/* Mixed-precision msp_timesum for a real symmetric C (real_mode = 1, the
 * fast_evolve and isotropic routes: Hamiltonian.fast_evolve's cosine sum,
 * hamiltonian.py:207-217 / cython/parallel.pyx:59-67, and evolve's time loop,
 * hamiltonian.py:87-107): d_out[b*T + t] = w_b sum_ij C_ij cos(2 pi (E_i - E_j) t)
 * with the phase arguments range-reduced in FP64, their sine and cosine in
 * FP32, the sum over j on FP32 matrix cores and the sum over i in FP64.
 * Errors are of order 1e-7 sum_ij |C_ij|.  d_work holds the phase table
 * (msp_timesum_mixed_workspace bytes); batch <= 65535. */
size_t msp_timesum_mixed_workspace(int n, int batch, int ntimes);
int msp_timesum_mixed(int n, int batch, int ntimes, const double *d_E, const double *d_C,
                      const double *d_times, const double *d_w, double *d_out, void *d_work,
                      size_t work_bytes, void *stream);

/* msp_timesum for a real symmetric C (real_mode = 1, the same routes and
 * reference lines as msp_timesum_mixed) in FP64 on matrix cores:
 * d_out[b*T + t] = w_b sum_ij C_ij cos(2 pi (E_i - E_j) t), the phases formed
 * exactly as msp_timesum forms them but once per (j, t), the sum over j on
 * FP64 matrix cores.  Agrees with msp_timesum to FP64 rounding.  d_work holds
 * the FP64 phase table (msp_timesum_mfma64_workspace bytes); batch <= 65535. */
size_t msp_timesum_mfma64_workspace(int n, int batch, int ntimes);
int msp_timesum_mfma64(int n, int batch, int ntimes, const double *d_E, const double *d_C,
                       const double *d_times, const double *d_w, double *d_out, void *d_work,
                       size_t work_bytes, void *stream);

/* I_b = sum_ij C_ij / (1/tau + 2 pi i (E_i - E_j)) * out_scale.
 * out_complex = 0: d_out[b] = w_b Re I_b; 1: complex, unweighted. */
int msp_integral(int n, int batch, const double *d_E, const double *d_C, int real_mode,
                 double tau, double out_scale, const double *d_w, int out_complex,
                 double *d_out, void *stream);

/* d_out[t] = exp(-2 pi i (E_i - E_j) t) rho_ij for one instance (n x n complex
 * per time, ntimes <= 65535): the eigenbasis states that Hamiltonian.evolve
 * returns when no operator is given (muspinsim/hamiltonian.py:108-116). */
int msp_phase_evolve(int n, int ntimes, const double *d_E, const double *d_rho,
                     const double *d_times, double *d_out, void *stream);

/* Liouvillian of each instance on row-major vec(rho) (N = n^2, n^2 <= 1024):
 * L = -i (H (x) 1 - 1 (x) H^T)
 *     + sum_k g_k (A_k (x) conj(A_k) - 1/2 (A_k^H A_k (x) 1 + 1 (x) (A_k^H A_k)^T)).
 * d_H: batch x n x n; d_A: batch x n_jump x n x n jump operators; d_rate:
 * batch x n_jump rates g_k; d_L: batch x N x N (complex). */
int msp_liouvillian(int n, int batch, const double *d_H, int n_jump, const double *d_A,
                    const double *d_rate, double *d_L, void *stream);

/* Workspace bytes of msp_lindblad_modes (0 when N <= 64: LDS only). */
size_t msp_lindblad_workspace(int N, int batch);

/* Eigendecomposition L = V diag(lambda) V^{-1} of each instance (d_L is
 * destroyed) and the mode amplitudes of n_ops observables:
 *   d_lam  batch x N         eigenvalues lambda_m
 *   d_amp  batch x n_ops x N amp_km = (u_k^T V)_m (V^{-1} r)_m
 *   d_c    batch x N         V^{-1} r (optional, may be NULL)
 *   d_V    batch x N x N     eigenvector columns (optional, may be NULL)
 * with d_r = vec(rho0) (batch x N) and d_u = vec(O_k^T) (batch x n_ops x N).
 * d_info[b] = 0 on success, else the index of an unconverged eigenvalue. */
int msp_lindblad_modes(int N, int batch, double *d_L, const double *d_r, int n_ops,
                       const double *d_u, double *d_lam, double *d_amp, double *d_c,
                       double *d_V, int32_t *d_info, void *d_work, size_t work_bytes,
                       void *stream);

/* P_bk(t) = sum_m amp_bkm exp(2 pi lambda_bm t).  out_complex = 1: d_out is
 * batch x n_ops x T complex; 0: d_out[(b n_ops + k) T + t] = w_b Re P. */
int msp_mode_sum(int N, int batch, int n_ops, int ntimes, const double *d_lam,
                 const double *d_amp, const double *d_times, const double *d_w,
                 int out_complex, double *d_out, void *stream);

/* I_bk = out_scale sum_m amp_bkm / (1/tau - 2 pi lambda_bm); complex or w_b Re. */
int msp_mode_integral(int N, int batch, int n_ops, const double *d_lam, const double *d_amp,
                      double tau, double out_scale, const double *d_w, int out_complex,
                      double *d_out, void *stream);

/* One instance: d_out[t] = V (c exp(2 pi lambda t)) (N complex per time, ntimes <= 65535). */
int msp_mode_states(int N, int ntimes, const double *d_lam, const double *d_c, const double *d_V,
                    const double *d_times, double *d_out, void *stream);

/* Adjoint of the fast-path time sum P_b(t) = (1/4d) sum_jk |O_jk|^2 cos(w_jk t),
 * w_jk = 2 pi (E_j - E_k), for upstream weights r_b(t) = scale w_b r[slot_b][t]:
 *   d_R[b]_jk = sum_t r_b(t) cos(w_jk t),  d_S[b]_jk = sum_t r_b(t) t sin(w_jk t)
 * (real n x n each; d_slot, d_w may be NULL for slot 0 / weight 1; batch <= 65535). */
int msp_adjoint_freq(int n, int batch, int ntimes, const double *d_E, const double *d_times,
                     const double *d_r, const int32_t *d_slot, const double *d_w, double scale,
                     double *d_R, double *d_S, void *stream);

/* d_M = d_R o d_O elementwise (R real, O and M complex). */
int msp_hadamard_real(int n, int batch, const double *d_R, const double *d_O, double *d_M,
                      void *stream);

/* Eigenbasis gradient G' (Hermitian part) of L = sum r P:
 *   G'_jj = -(pi/d) sum_k |O_jk|^2 S_jk
 *   G'_jm = (1/2d) (MO - OM)_jm / (E_j - E_m), or, when |E_j - E_m| <= degtol max|E|,
 *           (1/2d) sum_k O_jk O_km (-2 pi S_jk)  (the degenerate limit). */
int msp_adjoint_gprime(int n, int batch, const double *d_E, const double *d_O, const double *d_MO,
                       const double *d_OM, const double *d_S, double d, double degtol,
                       double *d_G, void *stream);

/* d_out[i] += sum_b d_X[b][i] over complex entries (len per instance), in batch order. */
int msp_batch_sum(int len, int batch, const double *d_X, double *d_out, void *stream);

/* d_out[p] = Re sum_ab G_ab dH_p[b][a] for nparam complex n x n matrices dH_p. */
int msp_trace_products(int n, int nparam, const double *d_G, const double *d_dH, double *d_out,
                       void *stream);

/* results[s][t] += rows[i][t] * scale for i in d_idx[d_ptr[s] .. d_ptr[s+1]),
 * in that order (deterministic powder accumulation). */
int msp_accumulate(int nslots, int width, const int32_t *d_ptr, const int32_t *d_idx,
                   const double *d_rows, double scale, double *d_results, void *stream);

/* Celio's Trotter evolution of nstates state vectors of dimension n
 * (replaces cpp/celio.cpp:23-68 with cpp/parallel.cpp:130-266, the loop of
 * CelioHamiltonian._fast_evolve_cpp, celio.py:431-476; the density-matrix
 * CelioHamiltonian.evolve, celio.py:210-293, maps onto it as a weighted sum
 * over product states).  Each state s belongs to configuration
 * c = d_state_cfg[s]; every time step writes d_out[s*ntimes + t] =
 * Re <psi|O_c|psi> and then applies k times the Trotter factors
 * d_desc[first_c .. first_c + count_c) with (first_c, count_c) =
 * d_cfg_steps[2c], d_cfg_steps[2c+1].  A descriptor is four int32
 * (m, matrix offset, fiber-table offset, element-table offset): the m x m
 * row-major complex matrix at d_mats + 2*offset acts on the fibers whose
 * base indices are d_tables[fiber .. fiber + n/m) through the m element
 * offsets d_tables[element .. element + m).  d_meas[4c..4c+3] describes O_c
 * (an operator on one spin of dimension 2..10).  init_mode 0:
 * psi = mu (x) exp(2 pi i phi)/sqrt(n/2) with the muon first, mu = the 2
 * complex at d_init_vec + 2*d_init_idx[s], phi = d_phases[s*(n/2) ..];
 * init_mode 1: product state psi[I] = prod_q v_q[digit_q(I)] with v_q at
 * d_init_vec + 2*d_init_idx[s*n_spins + q]; init_mode 2: the n complex at
 * d_init_vec + 2*d_init_idx[s].
 * n > 8192 keeps the states in d_work (msp_celio_workspace bytes); smaller
 * n keeps each state in LDS and needs no workspace.  Factor sizes m: any
 * m <= 64. */
size_t msp_celio_workspace(int n, int nstates);
int msp_celio_evolve(int n, int n_spins, const int32_t *h_dims, int nstates, int ntimes, int k,
                     const int32_t *d_state_cfg, const int32_t *d_cfg_steps,
                     const int32_t *d_desc, const int32_t *d_meas, const double *d_mats,
                     const int32_t *d_tables, int init_mode, const double *d_init_vec,
                     const int32_t *d_init_idx, const double *d_phases, double *d_out,
                     double *d_work, size_t work_bytes, void *stream);

/* Graph capture and replay of a sequence of msp_* calls on one stream (no
 * reference counterpart: the reference re-runs Python per instance).  Between
 * msp_graph_begin and msp_graph_end the calls on `stream` are recorded, not
 * run (including the side-stream work msp_eigh forks and joins); run the same
 * sequence once eagerly first, so that no kernel attribute is set during the
 * capture.  msp_graph_end instantiates the graph into *graph_exec;
 * msp_graph_launch replays it on a stream; msp_graph_destroy frees it.  The
 * buffers the calls name must stay allocated while the graph is in use. */
int msp_graph_begin(void *stream);
int msp_graph_end(void *stream, void **graph_exec);
int msp_graph_launch(void *graph_exec, void *stream);
int msp_graph_destroy(void *graph_exec);

/* Kernel timing (no reference counterpart; instrumentation for benchmarks):
 * while enabled every kernel launch is bracketed by two hipEvents on its
 * stream.  msp_profile_report waits for the recorded events, writes lines
 * "kernel_name launches total_ms" into h_buf and clears the record. */
int msp_profile_enable(int on);
int msp_profile_report(char *h_buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif
