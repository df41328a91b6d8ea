// extern "C" entry points of libmuspin_hip.so (declared in include/muspin_hip.h):
// argument validation, error reporting and kernel dispatch.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>

#include "../../include/muspin_hip.h"
#include "msp_kernels.h"

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int hip_status(hipError_t e, const char *where)
{
    if (e == hipSuccess) return MSP_OK;
    return fail(MSP_ERR_HIP, "%s: %s", where, hipGetErrorString(e));
}

inline hipStream_t S(void *s) { return (hipStream_t)s; }

constexpr int kMaxDim = 2048;
constexpr int kJacobiMax = 64;
constexpr int kMaxLiouville = 1024;      // N = n^2 of the Lindbladian path
constexpr int kSchurIts = 30;            // QR sweeps per eigenvalue

}  // namespace

// Workspace of msp_eigh_noalloc: a library-owned device array, one copy per
// device (each GPU the library runs on holds it until the process ends).  In
// host code `scratch` names a host-side stand-in, never the device memory:
// only the per-device address from hipGetSymbolAddress is passed to kernels.
static_assert(MSP_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch size");
static __device__ __align__(256) double scratch[MSP_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device with
// that device current; nullptr (and the message set) on failure.
double *scratch_ptr(const char *where)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        hip_status(e, where);
        return nullptr;
    }
    if (dev < 0 || dev >= kMaxDevices) {
        fail(MSP_ERR_UNSUPPORTED, "%s: device %d beyond %d", where, dev, kMaxDevices);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            hip_status(e == hipSuccess ? hipErrorInvalidSymbol : e, where);
            return nullptr;
        }
        g_scratch_ptr[dev] = (double *)p;
    }
    return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

int msp_version(void) { return 1; }

int msp_last_error(char *h_buf, size_t len)
{
    if (!h_buf || len == 0) return MSP_ERR_VALUE;
    strncpy(h_buf, g_err, len - 1);
    h_buf[len - 1] = '\0';
    return MSP_OK;
}

int msp_graph_begin(void *stream)
{
    if (!stream) return fail(MSP_ERR_VALUE, "msp_graph_begin: the legacy default stream cannot be captured");
    return hip_status(hipStreamBeginCapture(S(stream), hipStreamCaptureModeThreadLocal),
                      "msp_graph_begin");
}

int msp_graph_end(void *stream, void **graph_exec)
{
    if (!stream || !graph_exec) return fail(MSP_ERR_VALUE, "msp_graph_end: null stream or output");
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(S(stream), &g);
    if (e != hipSuccess) return hip_status(e, "msp_graph_end");
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return hip_status(e, "msp_graph_end");
    *graph_exec = (void *)x;
    return MSP_OK;
}

int msp_graph_launch(void *graph_exec, void *stream)
{
    if (!graph_exec) return fail(MSP_ERR_VALUE, "msp_graph_launch: null graph");
    return hip_status(hipGraphLaunch((hipGraphExec_t)graph_exec, S(stream)), "msp_graph_launch");
}

int msp_graph_destroy(void *graph_exec)
{
    if (!graph_exec) return MSP_OK;
    return hip_status(hipGraphExecDestroy((hipGraphExec_t)graph_exec), "msp_graph_destroy");
}

int msp_init(void)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_status(e, "msp_init");
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) return hip_status(e, "msp_init");
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(MSP_ERR_UNSUPPORTED, "msp_init: device is %s, library is built for gfx950",
                    prop.gcnArchName);
    return MSP_OK;
}

int msp_rotate_fields(int batch, const double *d_q, const double *d_B, const double *d_Bint,
                      const double *d_p, double *d_Br, double *d_pr, void *stream)
{
    if (batch < 0) return fail(MSP_ERR_VALUE, "msp_rotate_fields: batch < 0");
    if (batch == 0) return MSP_OK;
    if (!d_q || !d_B || !d_Bint || !d_p || !d_Br || !d_pr)
        return fail(MSP_ERR_VALUE, "msp_rotate_fields: null pointer");
    return hip_status(launch_rotate(batch, d_q, d_B, d_Bint, d_p, d_Br, d_pr, S(stream)),
                      "msp_rotate_fields");
}

int msp_build_h(int n, int batch, const MspSystem *d_systems, const MspTerm *d_terms,
                const int32_t *d_sys_id, const double *d_Br, const double *d_base, int zeeman,
                double *d_H, void *stream)
{
    if (n < 1 || n > kMaxDim) return fail(MSP_ERR_VALUE, "msp_build_h: n = %d out of range", n);
    if (batch < 0) return fail(MSP_ERR_VALUE, "msp_build_h: batch < 0");
    if (batch == 0) return MSP_OK;
    if (!d_systems || !d_H || (zeeman && !d_Br) || (!d_base && !d_terms))
        return fail(MSP_ERR_VALUE, "msp_build_h: null pointer");
    return hip_status(launch_build_h(n, batch, d_systems, d_terms, d_sys_id, d_Br, d_base, zeeman,
                                     0, d_H, S(stream)),
                      "msp_build_h");
}

int msp_build_h_lower(int n, int batch, const MspSystem *d_systems, const MspTerm *d_terms,
                      const int32_t *d_sys_id, const double *d_Br, const double *d_base,
                      int zeeman, double *d_H, void *stream)
{
    if (n < 1 || n > kMaxDim)
        return fail(MSP_ERR_VALUE, "msp_build_h_lower: n = %d out of range", n);
    if (batch < 0) return fail(MSP_ERR_VALUE, "msp_build_h_lower: batch < 0");
    if (batch == 0) return MSP_OK;
    if (!d_systems || !d_H || (zeeman && !d_Br) || (!d_base && !d_terms))
        return fail(MSP_ERR_VALUE, "msp_build_h_lower: null pointer");
    return hip_status(launch_build_h(n, batch, d_systems, d_terms, d_sys_id, d_Br, d_base, zeeman,
                                     1, d_H, S(stream)),
                      "msp_build_h_lower");
}

size_t msp_eigh_workspace(int n, int batch, int method)
{
    if (n < 1 || batch < 1) return 0;
    if (method == MSP_EIGH_AUTO) method = (n <= kJacobiMax) ? MSP_EIGH_JACOBI : MSP_EIGH_DC;
    if (method == MSP_EIGH_JACOBI) return 0;
    if (method == MSP_EIGH_DC2) return dc2_workspace_doubles(n) * sizeof(double) * (size_t)batch;
    return dc_workspace_doubles(n) * sizeof(double) * (size_t)batch;
}

int msp_eigh(int n, int batch, double *d_H, double *d_E, double *d_U, int32_t *d_info, int method,
             void *d_work, size_t work_bytes, void *stream)
{
    if (n < 1 || n > kMaxDim) return fail(MSP_ERR_VALUE, "msp_eigh: n = %d out of range", n);
    if (batch < 0) return fail(MSP_ERR_VALUE, "msp_eigh: batch < 0");
    if (batch == 0) return MSP_OK;
    if (!d_H || !d_E || !d_U || !d_info) return fail(MSP_ERR_VALUE, "msp_eigh: null pointer");
    if (method == MSP_EIGH_AUTO) method = (n <= kJacobiMax) ? MSP_EIGH_JACOBI : MSP_EIGH_DC;
    if (method == MSP_EIGH_JACOBI) {
        if (n > kJacobiMax)
            return fail(MSP_ERR_UNSUPPORTED, "msp_eigh: Jacobi path holds n <= %d in LDS (n = %d)",
                        kJacobiMax, n);
        return hip_status(launch_eigh_jacobi(n, batch, d_H, d_E, d_U, d_info, 40, S(stream)),
                          "msp_eigh(jacobi)");
    }
    if (method == MSP_EIGH_DC2) {
        if (n < kSbrMin || n > kSbrMax)
            return fail(MSP_ERR_UNSUPPORTED, "msp_eigh: two-stage path holds %d <= n <= %d (n = %d)",
                        kSbrMin, kSbrMax, n);
        const size_t need2 = msp_eigh_workspace(n, batch, MSP_EIGH_DC2);
        if (!d_work || work_bytes < need2)
            return fail(MSP_ERR_VALUE, "msp_eigh: workspace of %zu bytes needed, %zu given", need2,
                        work_bytes);
        return hip_status(launch_eigh_dc2(n, batch, d_H, d_E, d_U, d_info, (double *)d_work, S(stream)),
                          "msp_eigh(dc2)");
    }
    if (method != MSP_EIGH_DC) return fail(MSP_ERR_VALUE, "msp_eigh: unknown method %d", method);
    size_t need = msp_eigh_workspace(n, batch, MSP_EIGH_DC);
    if (!d_work || work_bytes < need)
        return fail(MSP_ERR_VALUE, "msp_eigh: workspace of %zu bytes needed, %zu given", need,
                    work_bytes);
    return hip_status(launch_eigh_dc(n, batch, d_H, d_E, d_U, d_info, (double *)d_work, S(stream)),
                      "msp_eigh(dc)");
}

size_t msp_eigh_mixed_workspace(int n, int batch)
{
    if (n < 1 || n > 192 || batch < 1) return 0;
    return (dc_workspace_doubles(n) * sizeof(double) + (size_t)n * n * 2 * sizeof(double)) *
           (size_t)batch;
}

int msp_eigh_mixed(int n, int batch, const double *d_H, double *d_E, double *d_U, int32_t *d_info,
                   void *d_work, size_t work_bytes, void *stream)
{
    if (n < 2 || n > 192)
        return fail(MSP_ERR_UNSUPPORTED, "msp_eigh_mixed: n = %d outside 2..192 (LDS-resident FP32 "
                    "reduction)", n);
    if (batch < 0) return fail(MSP_ERR_VALUE, "msp_eigh_mixed: batch < 0");
    if (batch == 0) return MSP_OK;
    if (!d_H || !d_E || !d_U || !d_info) return fail(MSP_ERR_VALUE, "msp_eigh_mixed: null pointer");
    const size_t need = msp_eigh_mixed_workspace(n, batch);
    if (!d_work || work_bytes < need)
        return fail(MSP_ERR_VALUE, "msp_eigh_mixed: workspace of %zu bytes needed, %zu given", need,
                    work_bytes);
    double *work = (double *)d_work;
    double *R = work + dc_workspace_doubles(n) * (size_t)batch;
    hipError_t e;
    {
        ProfScope ps_("k_tridiag32", S(stream));
        e = launch_tridiag32(n, batch, d_H, R, d_E, work, S(stream));
    }
    if (e != hipSuccess) return hip_status(e, "msp_eigh_mixed(tridiag32)");
    return hip_status(launch_eigh_from_tridiag(n, batch, R, d_E, d_U, d_info, work, S(stream)),
                      "msp_eigh_mixed");
}

int msp_eigh_noalloc(int n, int batch, double *d_H, double *d_E, double *d_U, int32_t *d_info,
                     int method, void *stream)
{
    if (n < 1 || n > kMaxDim) return fail(MSP_ERR_VALUE, "msp_eigh_noalloc: n = %d out of range", n);
    if (batch < 0) return fail(MSP_ERR_VALUE, "msp_eigh_noalloc: batch < 0");
    if (batch == 0) return MSP_OK;
    if (!d_H || !d_E || !d_U || !d_info) return fail(MSP_ERR_VALUE, "msp_eigh_noalloc: null pointer");
    if (method == MSP_EIGH_AUTO) method = (n <= kJacobiMax) ? MSP_EIGH_JACOBI : MSP_EIGH_DC;
    if (method == MSP_EIGH_JACOBI) {
        if (n > kJacobiMax)
            return fail(MSP_ERR_UNSUPPORTED,
                        "msp_eigh_noalloc: Jacobi path holds n <= %d in LDS (n = %d)", kJacobiMax, n);
        return hip_status(launch_eigh_jacobi(n, batch, d_H, d_E, d_U, d_info, 40, S(stream)),
                          "msp_eigh_noalloc(jacobi)");
    }
    if (method != MSP_EIGH_DC)
        return fail(MSP_ERR_VALUE, "msp_eigh_noalloc: unknown method %d", method);
    const size_t need = msp_eigh_workspace(n, batch, MSP_EIGH_DC);
    if (need > (size_t)MSP_NOALLOC_SCRATCH_BYTES)
        return fail(MSP_ERR_VALUE,
                    "msp_eigh_noalloc: n = %d x batch = %d needs %zu bytes of workspace, more than "
                    "the %zu-byte static scratch (split the batch, or use msp_eigh)",
                    n, batch, need, (size_t)MSP_NOALLOC_SCRATCH_BYTES);
    double *work = scratch_ptr("msp_eigh_noalloc");
    if (!work) return MSP_ERR_HIP;
    // every launch on `stream`: no side stream and no events
    return hip_status(launch_eigh_dc(n, batch, d_H, d_E, d_U, d_info, work, S(stream), false),
                      "msp_eigh_noalloc(dc)");
}

int msp_apply_muon_op(int n, int batch, const MspSystem *d_systems, const int32_t *d_sys_id,
                      const double *d_vec, int pauli, const double *d_U, double *d_Y,
                      void *stream)
{
    if (n < 1 || n > kMaxDim) return fail(MSP_ERR_VALUE, "msp_apply_muon_op: n out of range");
    if (batch < 0) return fail(MSP_ERR_VALUE, "msp_apply_muon_op: batch < 0");
    if (batch == 0) return MSP_OK;
    if (!d_systems || !d_vec || !d_U || !d_Y)
        return fail(MSP_ERR_VALUE, "msp_apply_muon_op: null pointer");
    return hip_status(launch_apply_muon(n, batch, d_systems, d_sys_id, d_vec, pauli, d_U, d_Y,
                                        S(stream)),
                      "msp_apply_muon_op");
}

int msp_apply_kron(int n, int batch, int n_spins, const int32_t *h_dims, const double *d_F,
                   const double *d_U, double *d_Y, double *d_tmp, void *stream)
{
    if (n < 1 || n > kMaxDim || n_spins < 1 || n_spins > MSP_MAX_SPINS || !h_dims)
        return fail(MSP_ERR_VALUE, "msp_apply_kron: bad dimensions");
    long prod = 1, fstride = 0;
    for (int k = 0; k < n_spins; ++k) {
        if (h_dims[k] < 1) return fail(MSP_ERR_VALUE, "msp_apply_kron: bad spin dimension");
        prod *= h_dims[k];
        fstride += 2L * h_dims[k] * h_dims[k];
    }
    if (prod != n) return fail(MSP_ERR_VALUE, "msp_apply_kron: prod(dims) = %ld != n = %d", prod, n);
    if (batch < 0) return fail(MSP_ERR_VALUE, "msp_apply_kron: batch < 0");
    if (batch == 0) return MSP_OK;
    if (!d_F || !d_U || !d_Y || !d_tmp) return fail(MSP_ERR_VALUE, "msp_apply_kron: null pointer");
    // ping-pong so that the last product lands in d_Y
    const double *src = d_U;
    double *bufs[2] = {d_Y, d_tmp};
    int dst = (n_spins % 2 == 1) ? 0 : 1;
    long foff = 0;
    int stride = n;
    for (int k = 0; k < n_spins; ++k) {
        stride /= h_dims[k];
        hipError_t e = launch_mode_product(n, batch, h_dims[k], stride, (int)(foff / 2),
                                           (int)(fstride / 2), d_F, src, bufs[dst], S(stream));
        if (e != hipSuccess) return hip_status(e, "msp_apply_kron");
        foff += 2L * h_dims[k] * h_dims[k];
        src = bufs[dst];
        dst ^= 1;
    }
    return MSP_OK;
}

int msp_gemm_cn(int n, int batch, const double *d_A, const double *d_B, double *d_C, void *stream)
{
    if (n < 1 || n > kMaxDim) return fail(MSP_ERR_VALUE, "msp_gemm_cn: n out of range");
    if (batch < 0) return fail(MSP_ERR_VALUE, "msp_gemm_cn: batch < 0");
    if (batch == 0) return MSP_OK;
    if (!d_A || !d_B || !d_C) return fail(MSP_ERR_VALUE, "msp_gemm_cn: null pointer");
    return hip_status(launch_gemm_cn(n, batch, d_A, d_B, d_C, S(stream)), "msp_gemm_cn");
}

int msp_gemm_cn_upper(int n, int batch, const double *d_A, const double *d_B, double *d_C,
                      void *stream)
{
    if (n < 1 || n > kMaxDim) return fail(MSP_ERR_VALUE, "msp_gemm_cn_upper: n out of range");
    if (batch < 0) return fail(MSP_ERR_VALUE, "msp_gemm_cn_upper: batch < 0");
    if (batch == 0) return MSP_OK;
    if (!d_A || !d_B || !d_C) return fail(MSP_ERR_VALUE, "msp_gemm_cn_upper: null pointer");
    return hip_status(launch_gemm_cn_upper(n, batch, d_A, d_B, d_C, S(stream)),
                      "msp_gemm_cn_upper");
}

int msp_gemm_muon_upper(int n, int batch, const MspSystem *d_systems, const int32_t *d_sys_id,
                        const double *d_vec, int pauli, const double *d_U, double *d_O,
                        void *stream)
{
    if (n < 2 || n > kMaxDim) return fail(MSP_ERR_VALUE, "msp_gemm_muon_upper: n out of range");
    if (batch < 0) return fail(MSP_ERR_VALUE, "msp_gemm_muon_upper: batch < 0");
    if (batch == 0) return MSP_OK;
    if (!d_systems || !d_vec || !d_U || !d_O)
        return fail(MSP_ERR_VALUE, "msp_gemm_muon_upper: null pointer");
    return hip_status(launch_gemm_muon_upper(n, batch, d_systems, d_sys_id, d_vec, pauli, d_U, d_O,
                                             S(stream)),
                      "msp_gemm_muon_upper");
}

int msp_integral_obs(int n, int batch, const double *d_E, const double *d_O, double alpha,
                     double beta, double tau, double out_scale, const double *d_w, double *d_out,
                     void *stream)
{
    if (n < 1 || n > kMaxDim) return fail(MSP_ERR_VALUE, "msp_integral_obs: n out of range");
    if (!(tau > 0.0)) return fail(MSP_ERR_VALUE, "msp_integral_obs: tau must be > 0");
    if (batch < 0) return fail(MSP_ERR_VALUE, "msp_integral_obs: batch < 0");
    if (batch == 0) return MSP_OK;
    if (!d_E || !d_O || !d_out) return fail(MSP_ERR_VALUE, "msp_integral_obs: null pointer");
    return hip_status(launch_integral_obs(n, batch, d_E, d_O, alpha, beta, tau, out_scale, d_w,
                                          d_out, S(stream)),
                      "msp_integral_obs");
}

int msp_form_c(int n, int batch, const double *d_rho, const double *d_obs, int real_mode,
               double alpha, double beta, double *d_C, void *stream)
{
    if (n < 1 || n > kMaxDim) return fail(MSP_ERR_VALUE, "msp_form_c: n out of range");
    if (batch < 0) return fail(MSP_ERR_VALUE, "msp_form_c: batch < 0");
    if (batch == 0) return MSP_OK;
    if (!d_obs || !d_C || (!real_mode && !d_rho)) return fail(MSP_ERR_VALUE, "msp_form_c: null pointer");
    return hip_status(launch_form_c(n, batch, d_rho, d_obs, real_mode, alpha, beta, d_C, S(stream)),
                      "msp_form_c");
}

int msp_timesum(int n, int batch, int ntimes, const double *d_E, const double *d_C, int real_mode,
                const double *d_times, const double *d_w, int out_complex, double *d_out,
                void *stream)
{
    if (n < 1 || n > kMaxDim) return fail(MSP_ERR_VALUE, "msp_timesum: n out of range");
    if (batch < 0 || ntimes < 0) return fail(MSP_ERR_VALUE, "msp_timesum: negative size");
    if (batch == 0 || ntimes == 0) return MSP_OK;
    if (!d_E || !d_C || !d_times || !d_out) return fail(MSP_ERR_VALUE, "msp_timesum: null pointer");
    if ((size_t)ntimes / 64 + 1 > 2147483647u) return fail(MSP_ERR_VALUE, "msp_timesum: too many times");
    return hip_status(launch_timesum(n, batch, ntimes, d_E, d_C, real_mode, d_times, d_w,
                                     out_complex, d_out, S(stream)),
                      "msp_timesum");
}

size_t msp_timesum_mixed_workspace(int n, int batch, int ntimes)
{
    if (n < 1 || batch < 1 || ntimes < 1) return 0;
    return timesum_mfma_workspace(n, batch, ntimes);
}

int msp_timesum_mixed(int n, int batch, int ntimes, const double *d_E, const double *d_C,
                      const double *d_times, const double *d_w, double *d_out, void *d_work,
                      size_t work_bytes, void *stream)
{
    if (n < 1 || n > kMaxDim) return fail(MSP_ERR_VALUE, "msp_timesum_mixed: n out of range");
    if (batch < 0 || ntimes < 0) return fail(MSP_ERR_VALUE, "msp_timesum_mixed: negative size");
    if (batch == 0 || ntimes == 0) return MSP_OK;
    if (batch > 65535) return fail(MSP_ERR_VALUE, "msp_timesum_mixed: batch > 65535");
    if (!d_E || !d_C || !d_times || !d_out)
        return fail(MSP_ERR_VALUE, "msp_timesum_mixed: null pointer");
    const size_t need = timesum_mfma_workspace(n, batch, ntimes);
    if (!d_work || work_bytes < need)
        return fail(MSP_ERR_VALUE, "msp_timesum_mixed: workspace of %zu bytes needed, %zu given",
                    need, work_bytes);
    return hip_status(launch_timesum_mfma(n, batch, ntimes, d_E, d_C, d_times, d_w, d_out, d_work,
                                          S(stream)),
                      "msp_timesum_mixed");
}

size_t msp_timesum_mfma64_workspace(int n, int batch, int ntimes)
{
    if (n < 1 || batch < 1 || ntimes < 1) return 0;
    return timesum_mfma64_workspace(n, batch, ntimes);
}

int msp_timesum_mfma64(int n, int batch, int ntimes, const double *d_E, const double *d_C,
                       const double *d_times, const double *d_w, double *d_out, void *d_work,
                       size_t work_bytes, void *stream)
{
    if (n < 1 || n > kMaxDim) return fail(MSP_ERR_VALUE, "msp_timesum_mfma64: n out of range");
    if (batch < 0 || ntimes < 0) return fail(MSP_ERR_VALUE, "msp_timesum_mfma64: negative size");
    if (batch == 0 || ntimes == 0) return MSP_OK;
    if (batch > 65535) return fail(MSP_ERR_VALUE, "msp_timesum_mfma64: batch > 65535");
    if (!d_E || !d_C || !d_times || !d_out)
        return fail(MSP_ERR_VALUE, "msp_timesum_mfma64: null pointer");
    const size_t need = timesum_mfma64_workspace(n, batch, ntimes);
    if (!d_work || work_bytes < need)
        return fail(MSP_ERR_VALUE, "msp_timesum_mfma64: workspace of %zu bytes needed, %zu given",
                    need, work_bytes);
    return hip_status(launch_timesum_mfma64(n, batch, ntimes, d_E, d_C, d_times, d_w, d_out, d_work,
                                            S(stream)),
                      "msp_timesum_mfma64");
}

int msp_integral(int n, int batch, const double *d_E, const double *d_C, int real_mode, double tau,
                 double out_scale, const double *d_w, int out_complex, double *d_out, void *stream)
{
    if (n < 1 || n > kMaxDim) return fail(MSP_ERR_VALUE, "msp_integral: n out of range");
    if (!(tau > 0.0)) return fail(MSP_ERR_VALUE, "msp_integral: tau must be > 0");
    if (batch < 0) return fail(MSP_ERR_VALUE, "msp_integral: batch < 0");
    if (batch == 0) return MSP_OK;
    if (!d_E || !d_C || !d_out) return fail(MSP_ERR_VALUE, "msp_integral: null pointer");
    return hip_status(launch_integral(n, batch, d_E, d_C, real_mode, tau, out_scale, d_w,
                                      out_complex, d_out, S(stream)),
                      "msp_integral");
}

int msp_phase_evolve(int n, int ntimes, const double *d_E, const double *d_rho,
                     const double *d_times, double *d_out, void *stream)
{
    if (n < 1 || n > kMaxDim) return fail(MSP_ERR_VALUE, "msp_phase_evolve: n out of range");
    if (ntimes < 0 || ntimes > 65535)
        return fail(MSP_ERR_VALUE, "msp_phase_evolve: ntimes = %d out of range", ntimes);
    if (ntimes == 0) return MSP_OK;
    if (!d_E || !d_rho || !d_times || !d_out)
        return fail(MSP_ERR_VALUE, "msp_phase_evolve: null pointer");
    return hip_status(launch_phase_evolve(n, ntimes, d_E, d_rho, d_times, d_out, S(stream)),
                      "msp_phase_evolve");
}

int msp_liouvillian(int n, int batch, const double *d_H, int n_jump, const double *d_A,
                    const double *d_rate, double *d_L, void *stream)
{
    if (n < 1 || n * n > kMaxLiouville)
        return fail(MSP_ERR_VALUE, "msp_liouvillian: n = %d out of range (n^2 <= %d)", n,
                    kMaxLiouville);
    if (batch < 0 || n_jump < 0) return fail(MSP_ERR_VALUE, "msp_liouvillian: negative size");
    if (batch == 0) return MSP_OK;
    if (!d_H || !d_L || (n_jump > 0 && (!d_A || !d_rate)))
        return fail(MSP_ERR_VALUE, "msp_liouvillian: null pointer");
    return hip_status(launch_liouvillian(n, batch, d_H, n_jump, d_A, d_rate, d_L, S(stream)),
                      "msp_liouvillian");
}

size_t msp_lindblad_workspace(int N, int batch)
{
    if (N < 1 || batch < 1) return 0;
    return lindblad_modes_workspace(N, batch);
}

int msp_lindblad_modes(int N, int batch, double *d_L, const double *d_r, int n_ops,
                       const double *d_u, double *d_lam, double *d_amp, double *d_c, double *d_V,
                       int32_t *d_info, void *d_work, size_t work_bytes, void *stream)
{
    if (N < 1 || N > kMaxLiouville)
        return fail(MSP_ERR_VALUE, "msp_lindblad_modes: N = %d out of range", N);
    if (batch < 0 || n_ops < 0) return fail(MSP_ERR_VALUE, "msp_lindblad_modes: negative size");
    if (batch == 0) return MSP_OK;
    if (!d_L || !d_r || !d_lam || !d_info || (n_ops > 0 && (!d_u || !d_amp)))
        return fail(MSP_ERR_VALUE, "msp_lindblad_modes: null pointer");
    const size_t need = lindblad_modes_workspace(N, batch);
    if (need > 0 && (!d_work || work_bytes < need))
        return fail(MSP_ERR_VALUE, "msp_lindblad_modes: workspace of %zu bytes needed", need);
    return hip_status(launch_lindblad_modes(N, batch, d_L, d_r, n_ops, d_u, d_lam, d_amp, d_c, d_V,
                                            d_info, (double *)d_work, kSchurIts, S(stream)),
                      "msp_lindblad_modes");
}

int msp_mode_sum(int N, int batch, int n_ops, int ntimes, const double *d_lam, const double *d_amp,
                 const double *d_times, const double *d_w, int out_complex, double *d_out,
                 void *stream)
{
    if (N < 1 || N > kMaxLiouville) return fail(MSP_ERR_VALUE, "msp_mode_sum: N out of range");
    if (batch < 0 || n_ops < 0 || ntimes < 0)
        return fail(MSP_ERR_VALUE, "msp_mode_sum: negative size");
    if (batch == 0 || n_ops == 0 || ntimes == 0) return MSP_OK;
    if (!d_lam || !d_amp || !d_times || !d_out)
        return fail(MSP_ERR_VALUE, "msp_mode_sum: null pointer");
    return hip_status(launch_mode_sum(N, batch, n_ops, ntimes, d_lam, d_amp, d_times, d_w,
                                      out_complex, d_out, S(stream)),
                      "msp_mode_sum");
}

int msp_mode_integral(int N, int batch, int n_ops, const double *d_lam, const double *d_amp,
                      double tau, double out_scale, const double *d_w, int out_complex,
                      double *d_out, void *stream)
{
    if (N < 1 || N > kMaxLiouville)
        return fail(MSP_ERR_VALUE, "msp_mode_integral: N out of range");
    if (!(tau > 0.0)) return fail(MSP_ERR_VALUE, "msp_mode_integral: tau must be > 0");
    if (batch < 0 || n_ops < 0) return fail(MSP_ERR_VALUE, "msp_mode_integral: negative size");
    if (batch == 0 || n_ops == 0) return MSP_OK;
    if (!d_lam || !d_amp || !d_out) return fail(MSP_ERR_VALUE, "msp_mode_integral: null pointer");
    return hip_status(launch_mode_integral(N, batch, n_ops, d_lam, d_amp, tau, out_scale, d_w,
                                           out_complex, d_out, S(stream)),
                      "msp_mode_integral");
}

int msp_mode_states(int N, int ntimes, const double *d_lam, const double *d_c, const double *d_V,
                    const double *d_times, double *d_out, void *stream)
{
    if (N < 1 || N > kMaxLiouville) return fail(MSP_ERR_VALUE, "msp_mode_states: N out of range");
    if (ntimes < 0 || ntimes > 65535)
        return fail(MSP_ERR_VALUE, "msp_mode_states: ntimes = %d out of range", ntimes);
    if (ntimes == 0) return MSP_OK;
    if (!d_lam || !d_c || !d_V || !d_times || !d_out)
        return fail(MSP_ERR_VALUE, "msp_mode_states: null pointer");
    return hip_status(launch_mode_states(N, ntimes, d_lam, d_c, d_V, d_times, d_out, S(stream)),
                      "msp_mode_states");
}

int msp_adjoint_freq(int n, int batch, int ntimes, const double *d_E, const double *d_times,
                     const double *d_r, const int32_t *d_slot, const double *d_w, double scale,
                     double *d_R, double *d_S, void *stream)
{
    if (n < 1 || n > kMaxDim) return fail(MSP_ERR_VALUE, "msp_adjoint_freq: n out of range");
    if (batch < 0 || batch > 65535 || ntimes < 0)
        return fail(MSP_ERR_VALUE, "msp_adjoint_freq: batch (<= 65535) or ntimes out of range");
    if (batch == 0) return MSP_OK;
    if (!d_E || !d_R || !d_S || (ntimes > 0 && (!d_times || !d_r)))
        return fail(MSP_ERR_VALUE, "msp_adjoint_freq: null pointer");
    return hip_status(launch_adjoint_freq(n, batch, ntimes, d_E, d_times, d_r, d_slot, d_w, scale,
                                          d_R, d_S, S(stream)),
                      "msp_adjoint_freq");
}

int msp_hadamard_real(int n, int batch, const double *d_R, const double *d_O, double *d_M,
                      void *stream)
{
    if (n < 1 || n > kMaxDim || batch < 0)
        return fail(MSP_ERR_VALUE, "msp_hadamard_real: size out of range");
    if (batch == 0) return MSP_OK;
    if (!d_R || !d_O || !d_M) return fail(MSP_ERR_VALUE, "msp_hadamard_real: null pointer");
    return hip_status(launch_hadamard_real(n, batch, d_R, d_O, d_M, S(stream)),
                      "msp_hadamard_real");
}

int msp_adjoint_gprime(int n, int batch, const double *d_E, const double *d_O, const double *d_MO,
                       const double *d_OM, const double *d_S, double d, double degtol,
                       double *d_G, void *stream)
{
    if (n < 1 || n > kMaxDim || batch < 0)
        return fail(MSP_ERR_VALUE, "msp_adjoint_gprime: size out of range");
    if (!(d > 0.0)) return fail(MSP_ERR_VALUE, "msp_adjoint_gprime: d must be > 0");
    if (batch == 0) return MSP_OK;
    if (!d_E || !d_O || !d_MO || !d_OM || !d_S || !d_G)
        return fail(MSP_ERR_VALUE, "msp_adjoint_gprime: null pointer");
    return hip_status(launch_adjoint_gprime(n, batch, d_E, d_O, d_MO, d_OM, d_S, d, degtol, d_G,
                                            S(stream)),
                      "msp_adjoint_gprime");
}

int msp_batch_sum(int len, int batch, const double *d_X, double *d_out, void *stream)
{
    if (len < 0 || batch < 0) return fail(MSP_ERR_VALUE, "msp_batch_sum: negative size");
    if (len == 0 || batch == 0) return MSP_OK;
    if (!d_X || !d_out) return fail(MSP_ERR_VALUE, "msp_batch_sum: null pointer");
    return hip_status(launch_batch_sum(len, batch, d_X, d_out, S(stream)), "msp_batch_sum");
}

int msp_trace_products(int n, int nparam, const double *d_G, const double *d_dH, double *d_out,
                       void *stream)
{
    if (n < 1 || n > kMaxDim || nparam < 0)
        return fail(MSP_ERR_VALUE, "msp_trace_products: size out of range");
    if (nparam == 0) return MSP_OK;
    if (!d_G || !d_dH || !d_out) return fail(MSP_ERR_VALUE, "msp_trace_products: null pointer");
    return hip_status(launch_trace_products(n, nparam, d_G, d_dH, d_out, S(stream)),
                      "msp_trace_products");
}

int msp_accumulate(int nslots, int width, const int32_t *d_ptr, const int32_t *d_idx,
                   const double *d_rows, double scale, double *d_results, void *stream)
{
    if (nslots < 0 || width < 0) return fail(MSP_ERR_VALUE, "msp_accumulate: negative size");
    if (nslots == 0 || width == 0) return MSP_OK;
    if (!d_ptr || !d_idx || !d_rows || !d_results)
        return fail(MSP_ERR_VALUE, "msp_accumulate: null pointer");
    return hip_status(launch_accumulate(nslots, width, d_ptr, d_idx, d_rows, scale, d_results,
                                        S(stream)),
                      "msp_accumulate");
}

size_t msp_celio_workspace(int n, int nstates)
{
    if (n < 1 || nstates < 0) return 0;
    return celio_workspace_bytes(n, nstates);
}

int msp_celio_evolve(int n, int n_spins, const int32_t *h_dims, int nstates, int ntimes, int k,
                     const int32_t *d_state_cfg, const int32_t *d_cfg_steps,
                     const int32_t *d_desc, const int32_t *d_meas, const double *d_mats,
                     const int32_t *d_tables, int init_mode, const double *d_init_vec,
                     const int32_t *d_init_idx, const double *d_phases, double *d_out,
                     double *d_work, size_t work_bytes, void *stream)
{
    if (n_spins < 1 || n_spins > MSP_MAX_SPINS || !h_dims)
        return fail(MSP_ERR_VALUE, "msp_celio_evolve: n_spins = %d out of range", n_spins);
    long long prod = 1;
    for (int q = 0; q < n_spins; ++q) {
        if (h_dims[q] < 1) return fail(MSP_ERR_VALUE, "msp_celio_evolve: bad spin dimension");
        prod *= h_dims[q];
    }
    if (n < 2 || prod != n || n > (1 << 26))
        return fail(MSP_ERR_VALUE, "msp_celio_evolve: n = %d does not match the spin dimensions", n);
    if (nstates < 0 || ntimes < 0 || k < 1)
        return fail(MSP_ERR_VALUE, "msp_celio_evolve: bad nstates / ntimes / k");
    if (nstates == 0 || ntimes == 0) return MSP_OK;
    if (init_mode < 0 || init_mode > 2)
        return fail(MSP_ERR_VALUE, "msp_celio_evolve: init_mode must be 0, 1 or 2");
    if (init_mode == 0 && (h_dims[0] != 2 || !d_phases))
        return fail(MSP_ERR_VALUE, "msp_celio_evolve: random states need the muon first and phases");
    if (!d_state_cfg || !d_cfg_steps || !d_desc || !d_meas || !d_mats || !d_tables ||
        !d_init_vec || !d_init_idx || !d_out)
        return fail(MSP_ERR_VALUE, "msp_celio_evolve: null pointer");
    const size_t need = celio_workspace_bytes(n, nstates);
    if (need > 0 && (!d_work || work_bytes < need))
        return fail(MSP_ERR_VALUE, "msp_celio_evolve: workspace of %zu bytes needed", need);
    return hip_status(launch_celio_evolve(n, n_spins, h_dims, nstates, ntimes, k, d_state_cfg,
                                          d_cfg_steps, d_desc, d_meas, d_mats, d_tables, init_mode,
                                          d_init_vec, d_init_idx, d_phases, d_out, d_work,
                                          S(stream)),
                      "msp_celio_evolve");
}

int msp_profile_enable(int on)
{
    prof_enable(on);
    return MSP_OK;
}

int msp_profile_report(char *h_buf, size_t len)
{
    if (!h_buf || len == 0) return fail(MSP_ERR_VALUE, "msp_profile_report: empty buffer");
    std::string r = prof_report();
    if (r.size() + 1 > len) return fail(MSP_ERR_VALUE, "msp_profile_report: buffer too small");
    memcpy(h_buf, r.c_str(), r.size() + 1);
    return MSP_OK;
}

}  // extern "C"
