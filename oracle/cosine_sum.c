/* C/OpenMP restatement of the reference's only native hot-path kernel,
 * parallel_fast_time_evolve (muspinsim/cython/parallel.pyx:16-69).
 *
 * TEST INFRASTRUCTURE: part of the CPU oracle and the CPU baseline.  For each
 * time t_i it forms  sum_j [ A_jj / 2 + sum_{k<j} A_jk cos(W_jk t_i) ]  in the
 * reference's loop order (j outer, k inner, lower triangle, libm cos), with
 * OpenMP threads over the time axis as the reference's prange does.  The
 * caller multiplies by 1/d.  A and W are dense row-major n x n.
 */
#include <math.h>

void oracle_cosine_sum(const double *times, long T, long n, const double *A,
                       const double *W, double *out, int threads)
{
    long i;
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (i = 0; i < T; ++i) {
        double acc = 0.0;
        const double t = times[i];
        for (long j = 0; j < n; ++j) {
            acc += 0.5 * A[j * n + j];
            for (long k = 0; k < j; ++k)
                acc += A[j * n + k] * cos(W[j * n + k] * t);
        }
        out[i] = acc;
    }
}
