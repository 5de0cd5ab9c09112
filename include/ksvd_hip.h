/*
 * ksvd_hip.h -- C ABI of the MI355X (gfx950) K-SVD denoising hot path.
 *
 * One opaque context per host thread (contexts on several threads may share a
 * device: their dictionary updates are ordered on the GPU, and allocations,
 * frees and graph captures take turns); every call returns an int status
 * (KSVD_OK = 0, negative on error) and keeps a per-context error string.
 * No exceptions cross this boundary; no torch or HIP types appear in it.
 * Host buffers passed in are caller-owned and copied; device memory is owned
 * by the context and freed by ksvd_destroy.
 *
 * Frame convention.  The image buffer is the reference's inputMatrix host
 * buffer after Denoiser::loadImage (denoiser.cpp:37-62): the CImg image
 * transposed, buf[x*H + y] = picture[y][x], passed with m = image width W and
 * n = image height H.  CudaKSvdDenoiser::createPatches reads it as n rows of
 * length m (cuda_svd_denoiser.cu:231-240), which for W != H cuts sheared
 * patches; KSVD_GEOM_COMPAT keeps that reading, KSVD_GEOM_TRUE (default)
 * reads the buffer as the m columns of length n it holds, and the two agree
 * for a square picture.  ksvd_rebuild returns the denoised buffer in the
 * same layout (before the final transpose of cuda_svd_denoiser.cu:489-490).
 *
 * The functions below replace, one for one, the stages of the reference's
 * CUDA_K_GESVDJ pipeline; the reference interface each one stands in for is
 * cited beside it (paths relative to the reference repository).
 */
#ifndef KSVD_HIP_H
#define KSVD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    KSVD_OK = 0,
    KSVD_ERR_ARG = -1,      /* invalid argument or parameter */
    KSVD_ERR_STATE = -2,    /* call out of order (e.g. omp before a dictionary) */
    KSVD_ERR_HIP = -3,      /* HIP runtime error (message in ksvd_last_error) */
    KSVD_ERR_ALLOC = -4,    /* device allocation failed */
    KSVD_ERR_UNSUPPORTED = -5
};

/* ---- numeric forms ------------------------------------------------------ */
enum {
    KSVD_PREC_FP32 = 0,     /* plain FP32: the drop-in for CUDA_K_GESVDJ */
    KSVD_PREC_BF16 = 1      /* BF16 correlation on MFMA, FP32 re-score/LS/SVD */
};
enum {
    KSVD_GEOM_TRUE = 0,     /* patches are blocks of the picture: the buffer is read
                               as the m columns of length n it holds, and the
                               rebuild walks the same grid */
    KSVD_GEOM_COMPAT = 1    /* the reference's walks: the buffer read as n rows of
                               length m (sheared patches when W != H), the rebuild
                               on its transposed grid for rectangular patches */
};

/* Parameters: the per-file keys of config.json (batch_denoiser.cpp:80-94,
 * Denoiser fields denoisingLib.h:44-56) plus the backend's own knobs. */
typedef struct ksvd_params {
    int patch_w;            /* patchWidthDim  (default 8)  */
    int patch_h;            /* patchHeightDim (default 8)  */
    int slide_w;            /* slidingWidth   (default 2)  */
    int slide_h;            /* slidingHeight  (default 2)  */
    int atoms;              /* atoms          (default 256) */
    int omp_iter;           /* ompIter: support-size limit (default 5) */
    int ksvd_iter;          /* ksvdIter       (default 10) */
    int min_omp_iter_batch; /* minOmpIterBatch: accepted and ignored (HBM holds
                               every patch; cublas_matrix_omp.cu:130-165) */
    int user_cap;           /* users per atom that enter the rank-1 update:
                               1024 reproduces the reference's <<<1,1024>>>
                               launches (cuda_svd_denoiser.cu:295-325); 0 = all */
    int precision;          /* KSVD_PREC_* */
    int geometry;           /* KSVD_GEOM_* */
    int device;             /* HIP device ordinal */
    int proposer;           /* 1: learned support proposer with OMP fallback
                               (BF16 form only); 0: plain OMP */
} ksvd_params;

/* Per-phase device time (HIP events on the context's stream) and counters. */
typedef struct ksvd_timings {
    double gather_ms;       /* patch gather kernel */
    double init_ms;         /* dictionary init kernel */
    double omp_ms;          /* sum over OMP kernel launches */
    double omp_kernel_ms;   /* the OMP pursuit kernel alone (roofline) */
    int64_t omp_calls;
    double update_ms;       /* sum over dictionary-update sweeps */
    int64_t update_calls;
    double rebuild_ms;      /* overlap-add + lambda + blend */
    int64_t omp_dup_stops;  /* pursuits ended by a repeated atom pick */
    int64_t omp_degenerate; /* pursuits ended by a non-positive pivot */
    int64_t eig_iters;      /* power iterations spent in rank-1 updates */
    int64_t omp_rescored;   /* BF16 form: patch-steps whose pick was re-scored in FP32 */
    int64_t omp_f64_refactors; /* BF16 form: support Grams re-factorised in FP64 */
    int64_t omp_candidates_cut; /* BF16 form: patch-steps with a third candidate in one lane group
                                   (re-scored on the slow path with every other candidate) */
    double train_ms;        /* learned proposer: training (forward, backward, AdamW) */
    int64_t proposed;       /* learned proposer: patches it proposed a support for */
    int64_t fallback;       /* learned proposer: of those, patches left to OMP */
    double train_loss;      /* learned proposer: mean loss of the last training batch */
    int64_t proposer_gated; /* learned proposer: 1 once it stood down (a pass accepted
                               under half the patches), plain OMP since */
    int64_t sweep_fallbacks; /* persistent sweeps that gave up and were undone and run
                               again with two kernels per atom */
} ksvd_timings;

typedef struct ksvd_ctx ksvd_ctx;

/* Fill p with the reference defaults (denoisingLib.h:44-53) + user_cap 1024. */
void ksvd_default_params(ksvd_params *p);

/* Create a context on p->device.  Replaces the backend construction done by
 * Denoiser::factory (denoiser.cpp:108-127) and CuBlasMatrixOps::init /
 * SvdCudaEngine::init (cublas_matrix_ops.cu:22-34, svd_cuda_engine.cu:17-44). */
int ksvd_create(const ksvd_params *p, ksvd_ctx **out);

/* Free every device buffer and the stream.  Replaces CuBlasMatrixOps::
 * finalize + SvdCudaEngine::finalize (cuda_svd_denoiser.cu:209-210). */
int ksvd_destroy(ksvd_ctx *ctx);

/* Last error message of ctx ("" if none).  With ctx NULL: the calling
 * thread's last ksvd_omp_noalloc error, else static text. */
const char *ksvd_last_error(const ksvd_ctx *ctx);

/* Upload the image buffer (host memory, m*n floats, n rows of length m).
 * Replaces Matrix::copyOnDevice of the noisy patches' source
 * (cuda_svd_denoiser.cu:246-247, matrix.cu:124-131). */
int ksvd_set_image(ksvd_ctx *ctx, const float *buf, int m, int n);

/* Same, from device memory already resident in HBM (device-to-device copy). */
int ksvd_set_image_device(ksvd_ctx *ctx, const float *dbuf, int m, int n);

/* Gather all patches into Y (dim x P, patch-contiguous) on the device.
 * Replaces CudaKSvdDenoiser::createPatches (cuda_svd_denoiser.cu:221-256). */
int ksvd_extract_patches(ksvd_ctx *ctx);

/* D = Y[:, P/2 .. P/2+K-1], each column divided by its norm.  Replaces
 * initDictionary + normalizeDictKernel (cuda_svd_denoiser.cu:52-63,261-280). */
int ksvd_init_dictionary(ksvd_ctx *ctx);

/* One batched OMP over every patch with the current dictionary.  Replaces
 * CuBlasMatrixOmp::work (cublas_matrix_omp.cu:107-402) and the pinv built by
 * CuSolverGeSvdABatch (cusolver_gesvda_batch_engine.cu:117-143). */
int ksvd_omp(ksvd_ctx *ctx);

/* Allocation-free OMP.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  The same pursuit as ksvd_omp
 * (the dictionary Gram, then the FP32 MFMA pursuit), without a context, on
 * caller-owned device memory: Y holds P patches of dim floats (patch p at
 * Y + p*dim), D holds K atoms of dim floats; outputs are the compact codes
 * of ksvd_get_codes (idx, coef: P x kmax; cnt: P), all device pointers.
 * stream is a hipStream_t (NULL: the null stream); the call only enqueues.
 * Its only scratch is a static device array of KSVD_NOALLOC_SCRATCH_BYTES
 * per GPU (the K x K Gram and counters), held until the process ends and
 * shared by every call on that GPU: calls that overlap on one GPU race.
 * Limits: dim <= 144, K <= 256, kmax <= 8 (KSVD_ERR_UNSUPPORTED beyond).
 * Errors: ksvd_last_error(NULL) gives the calling thread's last message. */
#define KSVD_NOALLOC_SCRATCH_BYTES 262272
int ksvd_omp_noalloc(const float *Y, int dim, int64_t P, const float *D, int K, int kmax,
                     int32_t *idx, float *coef, int32_t *cnt, void *stream);

/* One Gauss-Seidel sweep of rank-1 atom updates over the current codes.
 * Replaces CudaKSvdDenoiser::updateDictionary (cuda_svd_denoiser.cu:285-362)
 * with its CuBlasMatrixMult (cublas_matrix_mult.cu:41-75) and CuSolverGeSvdJ
 * (cusolver_gesvdj_engine.cu:46-69) calls. */
int ksvd_update_dictionary(ksvd_ctx *ctx);

/* iters x (ksvd_omp, ksvd_update_dictionary): the body of the loop of
 * CudaKSvdDenoiser::kSvd (cuda_svd_denoiser.cu:374-397). */
int ksvd_iterate(ksvd_ctx *ctx, int iters);

/* The whole of CudaKSvdDenoiser::kSvd (cuda_svd_denoiser.cu:368-408):
 * ksvd_iter iterations, then the final OMP. */
int ksvd_run(ksvd_ctx *ctx);

/* Several pictures' K-SVD at once.  The reference denoises a batch's files
 * one after another (the loop over BatchDenoiser::getDenoiserList in
 * main.cpp:60-68, each file's Denoiser::denoising running
 * CudaKSvdDenoiser::kSvd); here n contexts of one device
 * with the same patch size, atoms, OMP limit and user cap (plain FP32 form)
 * run their iterations together: each picture's pursuit on its own stream,
 * then one dictionary sweep launch with picture i on XCD i (n at most the
 * device's XCD count).  Each picture's result is the one ksvd_iterate gives
 * it alone, bit for bit.  Contexts that cannot share a sweep iterate one
 * after another.  ksvd_run_batch: ksvd_iter iterations (the same for all),
 * then each picture's final code. */
int ksvd_iterate_batch(ksvd_ctx *const *ctxs, int n, int iters);
int ksvd_run_batch(ksvd_ctx *const *ctxs, int n);

/* The final sparse code of kSvd (cuda_svd_denoiser.cu:400-401): ksvd_omp,
 * or, in the BF16 + learned form once the proposer is trained,
 * ksvd_proposer_omp. */
int ksvd_run_final_code(ksvd_ctx *ctx);

/* Overlap-add of D X, lambda from sigma2 (the noise variance that
 * Denoiser::loadImage estimates, denoiser.cpp:56), blend, and copy of the
 * m*n denoised buffer to host memory `out` (NULL: keep it on the device).
 * Replaces CudaKSvdDenoiser::createImage (cuda_svd_denoiser.cu:413-498) up to
 * its final transpose.  Pixels no patch covers keep their input value. */
int ksvd_rebuild(ksvd_ctx *ctx, float sigma2, float *out);

/* ksvd_rebuild of a log-domain (speckle) frame: the blend's output leaves
 * the log domain on the device, v -> 2^v, as createImage does on the host
 * when speckle is set (cuda_svd_denoiser.cu:483-484, myExp
 * denoisingLib.h:135-142); within 2 ulp of ksvd_speckle_exp on the output of
 * ksvd_rebuild.  ksvd_get_rebuild_parts still returns the log-domain sums. */
int ksvd_rebuild_speckle(ksvd_ctx *ctx, float sigma2, float *out);

/* After ksvd_rebuild: the overlap-add sums and weights (m*n each, host) and
 * lambda; any pointer may be NULL. */
int ksvd_get_rebuild_parts(ksvd_ctx *ctx, float *sum, float *weight, float *lambda, double *dsum);

/* Device pointer of the last rebuild's output (m*n floats, frame layout). */
int ksvd_rebuild_device_ptr(ksvd_ctx *ctx, const float **dptr);

/* Shapes: patch dimension, patch count, atoms, support limit. */
int ksvd_get_dims(ksvd_ctx *ctx, int *dim, int64_t *patches, int *atoms, int *kmax);

/* Copies between host memory and the context's matrices (the getters and
 * setters the reference reaches through Matrix::copyOnHost/copyOnDevice,
 * matrix.cu:124-142). */
int ksvd_get_patches(ksvd_ctx *ctx, float *Y);              /* dim x P */
int ksvd_get_dictionary(ksvd_ctx *ctx, float *D);           /* dim x K, column-major */
int ksvd_set_dictionary(ksvd_ctx *ctx, const float *D);
/* Compact codes: for patch p, cnt[p] <= kmax pairs (idx, coef) at p*kmax + i,
 * in pursuit order; dense X[idx, p] = coef. */
int ksvd_get_codes(ksvd_ctx *ctx, int32_t *idx, float *coef, int32_t *cnt);
int ksvd_set_codes(ksvd_ctx *ctx, const int32_t *idx, const float *coef, const int32_t *cnt);
/* Dense K x P column-major X, the layout of the reference's sparseCode. */
int ksvd_get_codes_dense(ksvd_ctx *ctx, float *X);

int ksvd_get_timings(ksvd_ctx *ctx, ksvd_timings *t);
int ksvd_reset_timings(ksvd_ctx *ctx);
int ksvd_synchronize(ksvd_ctx *ctx);

/* ---- learned support proposer (BF16 form, proposer = 1) ----------------
 * No counterpart in the reference: these drive the BF16 + learned form.  A
 * perceptron patch -> 512 -> 512 -> atoms (ReLU, BF16 weights, FP32
 * accumulation) proposes each patch's whole support in one shot; one FP32
 * least-squares solve gives its coefficients, and plain (BF16-form) OMP
 * takes over whenever the proposed residual exceeds 1.05 x the patch's last
 * OMP residual.  It learns from (patch, OMP support) pairs: ksvd_iterate
 * trains it on every patch after the first OMP of a dictionary and on each
 * fallback list afterwards.
 *
 * Input of a patch y with mean m: (y - m 1) / |y - m 1| (dim values), the DC
 * fraction m sqrt(dim) / |y|, and 1.  Parameter vector theta: W1 [512 x KIN]
 * | W2 [512 x 544] | W3 [KO x 544], row-major FP32; KIN = dim + 2 rounded up
 * to 32 (column dim + 1 is the bias), KO = atoms rounded up to 32, column
 * 512 of W2 and W3 is the bias. */

/* Shapes: input pitch KIN, output rows KO, parameter count. */
int ksvd_proposer_dims(ksvd_ctx *ctx, int *kin, int *ko, int64_t *nparams);
/* Seeded initialisation U(-1/sqrt(fan_in), 1/sqrt(fan_in)), biases 0. */
int ksvd_proposer_init(ksvd_ctx *ctx, uint64_t seed);
int ksvd_proposer_get_params(ksvd_ctx *ctx, float *theta);
int ksvd_proposer_set_params(ksvd_ctx *ctx, const float *theta);
/* `steps` AdamW steps (batch 16384, learning rate lr) on the current codes of
 * all patches as targets, in a seeded random order. */
int ksvd_proposer_train(ksvd_ctx *ctx, int steps, float lr, uint64_t seed);
/* One proposal pass over all patches followed by OMP on the rejected ones:
 * a replacement for ksvd_omp once the proposer is trained. */
int ksvd_proposer_omp(ksvd_ctx *ctx);
/* Mean loss and summed gradient (layout of theta) of one batch given on the
 * host: X [nb x KIN] inputs (rounded to BF16), T [nb x KO] 0/1 targets;
 * nb a multiple of 32 up to 16384.  For gradient checks. */
int ksvd_proposer_loss_grad(ksvd_ctx *ctx, const float *X, const float *T, int nb, double *loss, float *grad);

/* ---- host-side helpers (no GPU) ---------------------------------------- */

/* Seeded synthetic test image, row-major img[y*W + x]: six oriented
 * sinusoidal gratings (period 4-64 px), twenty constant rectangles, affine
 * rescale to [0,255] (clean), plus i.i.d. N(0, sigma^2) noise (noisy,
 * unclipped).  Either output may be NULL. */
int ksvd_synth_image(uint64_t seed, int W, int H, float sigma, float *clean, float *noisy);

/* CImg<float>::variance_noise() with its default method (scaled Laplacian,
 * then the MAD estimator), the estimate Denoiser::loadImage takes
 * (denoiser.cpp:56).  img is row-major W x H. */
double ksvd_noise_variance(const float *img, int W, int H);
/* The same over C planes of W x H pooled, as CImg computes it for a
 * multi-channel image (loadImage's colour case estimates it on Y, Cb, Cr). */
double ksvd_noise_variance_channels(const float *img, int W, int H, int C);

/* CImg<float>::normalize(lo, hi) in place (denoiser.cpp:75). */
void ksvd_normalize(float *img, int64_t count, float lo, float hi);

/* CImg PSNR with the default max 255 (denoiser.cpp:94-96). */
double ksvd_psnr(const float *ref, const float *img, int64_t count);

/* CImg RGBtoYCbCr(), which Denoiser::loadImage applies to a colour picture
 * before keeping channel 0 (denoiser.cpp:48-49): per pixel, 8-bit scale,
 * Y = (66 R + 129 G + 25 B + 128) / 256 + 16,
 * Cb = (-38 R - 74 G + 112 B + 128) / 256 + 128,
 * Cr = (112 R - 94 G - 18 B + 128) / 256 + 128, each clamped to [0, 255].
 * cb, cr may be NULL; ksvd_rgb_to_y is the luma alone. */
void ksvd_rgb_to_ycbcr(const float *r, const float *g, const float *b, int64_t count, float *y, float *cb,
                       float *cr);
void ksvd_rgb_to_y(const float *r, const float *g, const float *b, int64_t count, float *y);

/* The speckle transforms of Denoiser::loadImage and createImage in place
 * (myLog / myExp, denoisingLib.h:124-142; denoiser.cpp:53-54,
 * cuda_svd_denoiser.cu:483-484): v -> log2 v with 0 kept at 0, v -> 2^v. */
void ksvd_speckle_log(float *buf, int64_t count);
void ksvd_speckle_exp(float *buf, int64_t count);

/* Reference frame <-> image: buf[x*H + y] = img[y*W + x] and back. */
void ksvd_image_to_frame(const float *img, int W, int H, float *buf);
void ksvd_frame_to_image(const float *buf, int W, int H, float *img);

#ifdef __cplusplus
}
#endif
#endif /* KSVD_HIP_H */
