/*
 * K-SVD denoising oracle (CPU, C99 + OpenMP).  TEST INFRASTRUCTURE ONLY: the
 * checker for the HIP path and the timed CPU baseline, never the product.
 * Built by oracle/Makefile with -ffp-contract=off so that every multiply and
 * add rounds on its own, as the reference's host loops do; the fused
 * multiply-adds of the rebuild are explicit fma()/fmaf() calls.
 *
 * Two instantiations of ksvd_oracle_impl.h:
 *   orc_*_f32  REAL = float   (the reference's FP32 semantics)
 *   orc_*_f64  REAL = double  (the FP64 run of the same algorithm)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define ORC_KMAX 64          /* largest supported ompIter */
#define ORC_EIG_MAXIT 20000  /* power-iteration cap of the rank-1 solve */

static double orc_now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* Number of patches of the gather walk, cuda_svd_denoiser.cu:231-233. */
long orc_patch_count(int m, int n, int pw, int ph, int sw, int sh)
{
    if (pw > m || ph > n || pw <= 0 || ph <= 0 || sw <= 0 || sh <= 0) return 0;
    return (long)((n - ph) / sh + 1) * (long)((m - pw) / sw + 1);
}

int orc_kmax(void) { return ORC_KMAX; }

#define REAL float
#define SUFFIX _f32
#define ORC_FMA fmaf
#include "ksvd_oracle_impl.h"
#undef REAL
#undef SUFFIX
#undef ORC_FMA

#define REAL double
#define SUFFIX _f64
#define ORC_FMA fma
#include "ksvd_oracle_impl.h"
#undef REAL
#undef SUFFIX
#undef ORC_FMA
