/*
 * FP64 C restatement of the reference's numba kernel-sum loops, used as the
 * host-CPU baseline (timing) and as a cross-check of the numpy oracle.
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ and by bench.py's cpu_baseline
 * leg, never by the product library.
 *
 * Same loop nest as the reference: outer loop over scales, parallel loop over
 * targets, serial loop over sources, one exp per pair, float64 throughout.
 *   cpu_applyK      follows base/kernelFunctions_util.py:451-466 (+ gauss_fun :86-89)
 *   cpu_applyDiffKT follows base/kernelFunctions_util.py:773-793 (+ gauss_fun_diffT :232-236)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXK 16

void cpu_applyK(const double* y, long n, const double* x, long m, int dim, const double* a, int k,
                const double* scale, int ns, double kp, double* out) {
    double wsig = 0.0;
    for (int s = 0; s < ns; ++s) wsig += pow(scale[s], kp);
    memset(out, 0, sizeof(double) * (size_t)n * k);
    for (int s = 0; s < ns; ++s) {
        const double inv = 1.0 / scale[s];
        const double sKP = pow(scale[s], kp);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {
            double acc[MAXK] = {0};
            double yi[3] = {0, 0, 0};
            for (int d = 0; d < dim; ++d) yi[d] = y[i * dim + d] * inv;
            for (long j = 0; j < m; ++j) {
                double u = 0.0;
                for (int d = 0; d < dim; ++d) {
                    double t = yi[d] - x[j * dim + d] * inv;
                    u += t * t;
                }
                double K = exp(-u / 2);
                for (int c = 0; c < k; ++c) acc[c] += K * a[j * k + c];
            }
            for (int c = 0; c < k; ++c) out[i * k + c] += acc[c] * sKP;
        }
    }
    for (long i = 0; i < n * k; ++i) out[i] /= wsig;
}

/* mode: 0 plain, 1 sym, 2 lddmm (n == m for 1 and 2) */
void cpu_applyDiffKT(const double* y, long n, const double* x, long m, int dim, const double* p,
                     const double* a, int k, int mode, double regweight, const double* scale, int ns,
                     double kp, double* out) {
    double wsig = 0.0;
    for (int s = 0; s < ns; ++s) wsig += pow(scale[s], kp);
    memset(out, 0, sizeof(double) * (size_t)n * dim);
    for (int s = 0; s < ns; ++s) {
        const double inv = 1.0 / scale[s];
        const double sKP1 = pow(scale[s], kp - 1);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {
            double acc[3] = {0, 0, 0};
            double yi[3] = {0, 0, 0};
            for (int d = 0; d < dim; ++d) yi[d] = y[i * dim + d] * inv;
            for (long j = 0; j < m; ++j) {
                double dd[3] = {0, 0, 0};
                double u = 0.0;
                for (int d = 0; d < dim; ++d) {
                    dd[d] = yi[d] - x[j * dim + d] * inv;
                    u += dd[d] * dd[d];
                }
                double c = 0.0;
                for (int q = 0; q < k; ++q) {
                    c += p[i * k + q] * a[j * k + q];
                    if (mode >= 1) c += a[i * k + q] * p[j * k + q];
                    if (mode == 2) c -= 2 * regweight * a[i * k + q] * a[j * k + q];
                }
                double w = -exp(-u / 2) * c;
                for (int d = 0; d < dim; ++d) acc[d] += w * dd[d];
            }
            for (int d = 0; d < dim; ++d) out[i * dim + d] += acc[d] * sKP1;
        }
    }
    for (long i = 0; i < n * dim; ++i) out[i] /= wsig;
}

int cpu_threads(void) {
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}
