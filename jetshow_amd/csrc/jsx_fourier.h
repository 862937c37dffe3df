/*
 * jsx_fourier.h -- visibilities of an image: the exact direct sum
 *
 *     vis[k] = sum_j c_j exp(sgn 2 pi i (u_k x_j + v_k y_j)),
 *
 * the transform that the fitting loop applies to every image
 * (JetModelZoom.ft, jetmodel.py:114-132) and that the reference evaluates
 * approximately with finufft: FINUFFT_NUNU (fourier.py:39-94, nufft2d3,
 * scattered pixels -> scattered (u, v), eps 1e-6) and FINUFFT
 * (fourier.py:11-34, nufft2d2, a uniform pixel grid, eps 1e-8).  Both are
 * this sum with u, v in cycles per unit of x, y: the phase is in revolutions.
 *
 * The sum is VALU work (one sine and one cosine per pixel-visibility pair),
 * not a contraction: the exponent couples every pixel with every
 * visibility through u_k x_j, so there is no matrix to hand to MFMA.  Each
 * lane keeps JSX_VIS_R visibilities in registers; every lane of a block walks
 * the same pixels, so pixel coordinates and values are wave-uniform scalar
 * loads and each pixel costs its pair arithmetic only.  Long pixel lists are
 * cut into chunks (grid dimension y) so that a few thousand visibilities
 * still fill every CU; chunk partial sums are added in a fixed order by
 * jsx_k_vis_reduce, so results do not depend on timing.
 *
 * Precision.  The phase u x + v y and its reduction to [-1/2, 1/2]
 * revolutions are FP64 in both modes (phases of thousands of revolutions
 * need the bits).  FP64 mode: sincospi in FP64, FP64 sums.  Mixed mode:
 * sine and cosine in FP32, by default the hardware's (v_sin_f32 and
 * v_cos_f32 take revolutions; measured per-value error below 2e-7), or by
 * polynomials (exact quadrant reduction, Taylor series of degree 9 and 10 on
 * [-pi/4, pi/4], truncation below 2e-9; twice the instructions); FP32
 * partial sums over JSX_VIS_SUB pixels, FP64 beyond.
 */
#ifndef JSX_FOURIER_H
#define JSX_FOURIER_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef JSX_VIS_R
#define JSX_VIS_R 8        /* visibilities per lane                        */
#endif
#define JSX_VIS_BLOCK 256  /* lanes per block                              */
#define JSX_VIS_SUB 64     /* pixels per FP32 partial sum (mixed mode)     */

/* sin and cos of (pi/2) r, |r| <= 1/2: Taylor series of sin to degree 9 and of
 * cos to degree 10 in theta = (pi/2) r, written in r (coefficient of r^n is
 * (+-)(pi/2)^n / n!).  |theta| <= pi/4 bounds the truncation by
 * (pi/4)^11/11! = 1.8e-9 and (pi/4)^12/12! = 1.1e-10. */
__host__ __device__ __forceinline__ void vis_sincos_quarter(float r, float &s, float &c) {
    const float S1 = 1.5707963267948966f, S3 = -0.64596409750624625f, S5 = 0.079692626246167045f,
                S7 = -0.0046817541353186881f, S9 = 0.00016044118478735982f;
    const float C0 = 1.0f, C2 = -1.2337005501361698f, C4 = 0.25366950790104802f,
                C6 = -0.020863480763352961f, C8 = 0.00091926027483942659f, C10 = -2.5202042373060605e-05f;
    const float r2 = r * r;
    float ps = fmaf(r2, S9, S7);
    ps = fmaf(r2, ps, S5);
    ps = fmaf(r2, ps, S3);
    ps = fmaf(r2, ps, S1);
    s = r * ps;
    float pc = fmaf(r2, C10, C8);
    pc = fmaf(r2, pc, C6);
    pc = fmaf(r2, pc, C4);
    pc = fmaf(r2, pc, C2);
    c = fmaf(r2, pc, C0);
}

/* sin and cos of 2 pi ph, ph in revolutions: FP64 reduction to quarter
 * turns, t = 4 (ph - rint(ph)) in [-2, 2], q = rint(t), r = t - q in
 * [-1/2, 1/2] (every step exact: the first difference is exact for any ph,
 * 4x is a power-of-two scaling, the last difference is a Sterbenz
 * difference), so the angle is q pi/2 + (pi/2) r and only the rounding of r
 * to FP32 (2^-25 quarter turns) enters the sine and cosine. */
__host__ __device__ __forceinline__ void vis_sincos_rev_mixed(double ph, float &s, float &c) {
    const double t = 4.0 * (ph - rint(ph));
    const double q = rint(t);
    const float r = (float)(t - q);
    const int qi = (int)q & 3;
    float sr, cr;
    vis_sincos_quarter(r, sr, cr);
    const float S = (qi & 1) ? cr : sr;
    const float C = (qi & 1) ? sr : cr;
    s = (qi & 2) ? -S : S;
    c = ((qi + 1) & 2) ? -C : C;
}

#if defined(__HIPCC__)

/* The same with the hardware sine and cosine (v_sin_f32 / v_cos_f32, which
 * take their argument in revolutions): f = ph - rint(ph) in [-1/2, 1/2] in
 * FP64 (exact), rounded once to FP32 (at most 2^-26 revolutions; the
 * fractional part in [0, 1) would round to twice that). */
__device__ __forceinline__ void vis_sincos_rev_hw(double ph, float &s, float &c) {
    const float f = (float)(ph - rint(ph));
    s = __builtin_amdgcn_sinf(f);
    c = __builtin_amdgcn_cosf(f);
}

/* One block: JSX_VIS_BLOCK x JSX_VIS_R visibilities x one chunk of pixels.
 * out: the final visibilities when the launch has one chunk, else partial
 * sums [n_chunks][n_vis] (complex, interleaved re, im). */
template <bool MIXED, bool CPLX, bool HWTRIG>
__global__ __launch_bounds__(JSX_VIS_BLOCK) void jsx_k_vis(
    const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ cre,
    const double *__restrict__ cim, const double *__restrict__ u, const double *__restrict__ v,
    double2 *__restrict__ out, int64_t n_src, int64_t n_vis, int64_t chunk, double sgn) {
    const int64_t kb = (int64_t)blockIdx.x * (JSX_VIS_BLOCK * JSX_VIS_R) + threadIdx.x;
    double uk[JSX_VIS_R], vk[JSX_VIS_R];
    double A[JSX_VIS_R], B[JSX_VIS_R], Cc[JSX_VIS_R], D[JSX_VIS_R];
#pragma unroll
    for (int r = 0; r < JSX_VIS_R; ++r) {
        int64_t k = kb + (int64_t)r * JSX_VIS_BLOCK;
        if (k >= n_vis) k = n_vis - 1;   /* tail lanes compute a duplicate, never stored */
        uk[r] = u[k];
        vk[r] = v[k];
        A[r] = B[r] = Cc[r] = D[r] = 0.0;
    }
    const int64_t j0 = (int64_t)blockIdx.y * chunk;
    const int64_t j1 = j0 + chunk < n_src ? j0 + chunk : n_src;
    for (int64_t jb = j0; jb < j1; jb += JSX_VIS_SUB) {
        const int64_t je = jb + JSX_VIS_SUB < j1 ? jb + JSX_VIS_SUB : j1;
        if (MIXED) {
            float a[JSX_VIS_R], b[JSX_VIS_R], cc[JSX_VIS_R], d[JSX_VIS_R];
#pragma unroll
            for (int r = 0; r < JSX_VIS_R; ++r) a[r] = b[r] = cc[r] = d[r] = 0.0f;
            for (int64_t j = jb; j < je; ++j) {
                const double xj = x[j], yj = y[j];
                const float cr = (float)cre[j];
                const float ci = CPLX ? (float)cim[j] : 0.0f;
#pragma unroll
                for (int r = 0; r < JSX_VIS_R; ++r) {
                    const double ph = fma(uk[r], xj, vk[r] * yj);
                    float s, c;
                    if (HWTRIG)
                        vis_sincos_rev_hw(ph, s, c);
                    else
                        vis_sincos_rev_mixed(ph, s, c);
                    a[r] = fmaf(cr, c, a[r]);
                    b[r] = fmaf(cr, s, b[r]);
                    if (CPLX) {
                        cc[r] = fmaf(ci, c, cc[r]);
                        d[r] = fmaf(ci, s, d[r]);
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < JSX_VIS_R; ++r) {
                A[r] += (double)a[r];
                B[r] += (double)b[r];
                if (CPLX) {
                    Cc[r] += (double)cc[r];
                    D[r] += (double)d[r];
                }
            }
        } else {
            for (int64_t j = jb; j < je; ++j) {
                const double xj = x[j], yj = y[j];
                const double cr = cre[j];
                const double ci = CPLX ? cim[j] : 0.0;
#pragma unroll
                for (int r = 0; r < JSX_VIS_R; ++r) {
                    const double ph = fma(uk[r], xj, vk[r] * yj);
                    double s, c;
                    sincospi(2.0 * (ph - rint(ph)), &s, &c);
                    A[r] = fma(cr, c, A[r]);
                    B[r] = fma(cr, s, B[r]);
                    if (CPLX) {
                        Cc[r] = fma(ci, c, Cc[r]);
                        D[r] = fma(ci, s, D[r]);
                    }
                }
            }
        }
    }
    /* (cr + i ci)(cos + i sgn sin) */
    double2 *o = out + (size_t)blockIdx.y * (size_t)n_vis;
#pragma unroll
    for (int r = 0; r < JSX_VIS_R; ++r) {
        const int64_t k = kb + (int64_t)r * JSX_VIS_BLOCK;
        if (k < n_vis) {
            double re = A[r], im = sgn * B[r];
            if (CPLX) {
                re -= sgn * D[r];
                im += Cc[r];
            }
            o[k] = make_double2(re, im);
        }
    }
}

/* vis[k] = sum over chunks, in chunk order */
__global__ __launch_bounds__(256) void jsx_k_vis_reduce(const double2 *__restrict__ part, double2 *__restrict__ vis,
                                                        int64_t n_vis, int n_chunks) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= n_vis) return;
    double re = 0.0, im = 0.0;
    for (int c = 0; c < n_chunks; ++c) {
        const double2 p = part[(size_t)c * (size_t)n_vis + k];
        re += p.x;
        im += p.y;
    }
    vis[k] = make_double2(re, im);
}

__global__ __launch_bounds__(256) void jsx_k_vis_zero(double2 *__restrict__ vis, int64_t n_vis) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k < n_vis) vis[k] = make_double2(0.0, 0.0);
}

#endif /* __HIPCC__ */
#endif /* JSX_FOURIER_H */
