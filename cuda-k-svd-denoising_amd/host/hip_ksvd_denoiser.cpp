// HipKSvdDenoiser: CudaKSvdDenoiser's pipeline (src/lib/denoisingLib/
// cuda_svd_denoiser.cu:90-216) on gfx950 through the C ABI of libksvd_hip.so.
// Everything from the patch gather to the lambda blend stays on the device;
// the host sends the frame buffer and receives the denoised buffer.
#include <chrono>
#include <cmath>
#include <iostream>

#include "denoisingLib.h"
#include "ksvd_hip.h"

using namespace denoising;
using namespace baseUtl;

HipKSvdDenoiser::HipKSvdDenoiser() {}

HipKSvdDenoiser::~HipKSvdDenoiser() {}

// cuda_svd_denoiser.cu:90-102
signed char HipKSvdDenoiser::denoising()
{
    if (!loadImage()) return -1;
    if (!internalDenoising()) return -2;
    if (!saveImage()) return -3;
    return 0;
}

bool HipKSvdDenoiser::loadImage() { return Denoiser::loadImage(); }

bool HipKSvdDenoiser::saveImage() { return Denoiser::saveImage(); }

bool HipKSvdDenoiser::denoiseFrame(const float *buf, int m, int n, float noiseVar, std::vector<float> &out)
{
    ksvd_params p;
    ksvd_default_params(&p);
    p.patch_w = patchWidthDim;
    p.patch_h = patchHeightDim;
    p.slide_w = slidingWidth;
    p.slide_h = slidingHeight;
    p.atoms = atoms;
    p.omp_iter = ompIter;
    p.ksvd_iter = iter;
    p.min_omp_iter_batch = minOmpIterBatch;
    p.user_cap = userCap;
    p.geometry = geometry;
    p.device = device;
    p.precision = (type == HIP_K_SVD_BF16 || type == HIP_K_SVD_BF16_LEARNED) ? KSVD_PREC_BF16 : KSVD_PREC_FP32;
    p.proposer = type == HIP_K_SVD_BF16_LEARNED ? 1 : 0;
    ksvd_ctx *ctx = nullptr;
    int rc = ksvd_create(&p, &ctx);
    auto fail = [&](int code) {
        error = std::string("HIP backend: ") + ksvd_last_error(ctx) + " (status " + std::to_string(code) + ")";
        ksvd_destroy(ctx);
        return false;
    };
    if (rc) return fail(rc);
    if ((rc = ksvd_set_image(ctx, buf, m, n))) return fail(rc);
    if ((rc = ksvd_extract_patches(ctx))) return fail(rc);
    if ((rc = ksvd_init_dictionary(ctx))) return fail(rc);
    if ((rc = ksvd_reset_timings(ctx))) return fail(rc);
    // kSvd (cuda_svd_denoiser.cu:368-408): ksvdIter x (OMP, update), final OMP
    if ((rc = ksvd_iterate(ctx, iter))) return fail(rc);
    ksvd_timings t;
    if ((rc = ksvd_get_timings(ctx, &t))) return fail(rc);
    int dim = 0, K = 0, kmax = 0;
    int64_t P = 0;
    ksvd_get_dims(ctx, &dim, &P, &K, &kmax);
    nPatches = P;
    const double iter_ms = t.omp_ms + t.update_ms;
    patchIters = iter_ms > 0 ? (double)P * iter / (iter_ms * 1e-3) : 0;
    // the final sparse code: OMP, or the trained proposer with OMP fallback
    if ((rc = ksvd_run_final_code(ctx))) return fail(rc);
    out.assign((size_t)m * n, 0.f);
    if ((rc = ksvd_rebuild(ctx, noiseVar, out.data()))) return fail(rc);
    ksvd_destroy(ctx);
    return true;
}

bool HipKSvdDenoiser::internalDenoising()
{
    auto t0 = std::chrono::steady_clock::now();
    if (subImageWidthDim != 0 && subImageHeightDim != 0) {
        error = "sub-image tiling (subImageWidthDim / subImageHeightDim) is not built in the HIP backend; "
                "set both to 0 (the whole image fits in HBM)";
        return false;
    }
    const int W = inputMatrix->m, H = inputMatrix->n;
    std::vector<float> out;
    if (!denoiseFrame(inputMatrix->hostVector->data(), W, H, sigma, out)) return false;
    // speckle: back from the log domain (myExp, denoisingLib.h:135-142)
    if (speckle) ksvd_speckle_exp(out.data(), (int64_t)out.size());
    // back to the picture layout (the transpose of cuda_svd_denoiser.cu:489-490)
    std::vector<float> *pic = new std::vector<float>((size_t)W * H);
    ksvd_frame_to_image(out.data(), W, H, pic->data());
    outputMatrix = new Matrix(H, W, H, pic);
    timeElapsed->working =
        std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0).count();
    return true;
}
