/* jsx_mixed.h -- mixed-precision path (placeholder until the FP32 kernels land). */
#ifndef JSX_MIXED_H
#define JSX_MIXED_H
#include "jsx_solver.h"
static inline hipError_t jsx_launch_mixed(int n_cu, unsigned long long *head, const DModel &m, const DImage &D,
                                          const float *along_f, const float *across_f, const DSolver &S, const DOut &O,
                                          int64_t npix, hipStream_t st, unsigned flags) {
    (void)n_cu; (void)head; (void)m; (void)D; (void)along_f; (void)across_f; (void)S; (void)O; (void)npix; (void)st; (void)flags;
    return hipErrorNotSupported;
}
#endif
