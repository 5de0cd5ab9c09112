// HipKSvdDenoiser: CudaKSvdDenoiser's pipeline (src/lib/denoisingLib/
// cuda_svd_denoiser.cu:90-216) on gfx950 through the C ABI of libksvd_hip.so.
// Everything from the patch gather to the lambda blend stays on the device;
// the host sends the frame buffer and receives the denoised buffer.
#include <algorithm>
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
    if (grouped) return groupStatus;
    if (!loadImage()) return -1;
    if (!internalDenoising()) return -2;
    if (!saveImage()) return -3;
    return 0;
}

bool HipKSvdDenoiser::loadImage() { return Denoiser::loadImage(); }

bool HipKSvdDenoiser::saveImage() { return Denoiser::saveImage(); }

ksvd_ctx *HipKSvdDenoiser::makeContext()
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
    const int rc = ksvd_create(&p, &ctx);
    if (rc) {
        error = std::string("HIP backend: ") + ksvd_last_error(ctx) + " (status " + std::to_string(rc) + ")";
        ksvd_destroy(ctx);
        return nullptr;
    }
    return ctx;
}

bool HipKSvdDenoiser::denoiseFrame(ksvd_ctx *ctx, const float *buf, int m, int n, float noiseVar,
                                   std::vector<float> &out)
{
    int rc;
    auto fail = [&](int code) {
        error = std::string("HIP backend: ") + ksvd_last_error(ctx) + " (status " + std::to_string(code) + ")";
        return false;
    };
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
    nPatches += P;
    iterMs += t.omp_ms + t.update_ms;
    patchIterCount += (double)P * iter;
    // the final sparse code: OMP, or the trained proposer with OMP fallback
    if ((rc = ksvd_run_final_code(ctx))) return fail(rc);
    out.assign((size_t)m * n, 0.f);
    if ((rc = ksvd_rebuild(ctx, noiseVar, out.data()))) return fail(rc);
    return true;
}

bool HipKSvdDenoiser::denoiseTiles(const float *buf, int W, int H, std::vector<float> &out)
{
    // The reference cuts the tiles as it cuts patches (createPatches with the
    // patch and sliding sizes set to the sub-image sizes), in the frame as
    // the geometry reads it: rows of length fm, fn rows.
    const bool compat = geometry == KSVD_GEOM_COMPAT;
    const int fm = compat ? W : H, fn = compat ? H : W;
    const int th = subImageHeightDim, tw = subImageWidthDim;
    if (th <= 0 || tw <= 0 || th > fn || tw > fm) {
        error = "sub-image larger than the image (or not positive)";
        return false;
    }
    const int Tr = fn / th, Tc = fm / tw;
    // Pixels no tile covers: 0 in the reference's output (its rebuilt image
    // starts at 0, :505); the true geometry keeps them as they came in.
    out.assign((size_t)W * H, 0.f);
    if (!compat) {
        out.assign(buf, buf + (size_t)W * H);
        if (speckle) ksvd_speckle_exp(out.data(), (int64_t)out.size());
    }
    ksvd_ctx *ctx = makeContext();
    if (!ctx) return false;
    std::vector<float> tile((size_t)th * tw), den;
    for (int ti = 0; ti < Tr; ti++)
        for (int tj = 0; tj < Tc; tj++) {
            for (int r = 0; r < th; r++)
                std::copy(buf + (size_t)(ti * th + r) * fm + (size_t)tj * tw,
                          buf + (size_t)(ti * th + r) * fm + (size_t)tj * tw + tw, tile.data() + (size_t)r * tw);
            // Matrix(subImageHeightDim, subImageWidthDim, ...) over the tile
            // (cuda_svd_denoiser.cu:152) and its variance_noise as a
            // subImageHeightDim-wide CImg (:146,:158): the reference reads the
            // tile as rows of length th; the true geometry as the rows of
            // length tw it holds
            const float s2 = compat ? (float)ksvd_noise_variance(tile.data(), th, tw)
                                    : (float)ksvd_noise_variance(tile.data(), tw, th);
            if (!denoiseFrame(ctx, tile.data(), th, tw, s2, den)) {
                ksvd_destroy(ctx);
                return false;
            }
            if (speckle) ksvd_speckle_exp(den.data(), (int64_t)den.size());  // createImage per tile (:483-484)
            // createImageFromSubImages (:503-528): each tile back in its place
            for (int r = 0; r < th; r++)
                std::copy(den.data() + (size_t)r * tw, den.data() + (size_t)r * tw + tw,
                          out.data() + (size_t)(ti * th + r) * fm + (size_t)tj * tw);
        }
    ksvd_destroy(ctx);
    return true;
}

bool HipKSvdDenoiser::internalDenoising()
{
    auto t0 = std::chrono::steady_clock::now();
    const int W = inputMatrix->m, H = inputMatrix->n;
    std::vector<float> out;
    nPatches = 0;
    iterMs = 0;
    patchIterCount = 0;
    if (subImageWidthDim != 0 && subImageHeightDim != 0) {
        if (!denoiseTiles(inputMatrix->hostVector->data(), W, H, out)) return false;
    } else {
        ksvd_ctx *ctx = makeContext();
        if (!ctx) return false;
        const bool ok = denoiseFrame(ctx, inputMatrix->hostVector->data(), W, H, sigma, out);
        ksvd_destroy(ctx);
        if (!ok) return false;
        // speckle: back from the log domain (myExp, denoisingLib.h:135-142)
        if (speckle) ksvd_speckle_exp(out.data(), (int64_t)out.size());
    }
    patchIters = iterMs > 0 ? patchIterCount / (iterMs * 1e-3) : 0;
    // back to the picture layout (the transpose of cuda_svd_denoiser.cu:489-490)
    std::vector<float> *pic = new std::vector<float>((size_t)W * H);
    ksvd_frame_to_image(out.data(), W, H, pic->data());
    outputMatrix = new Matrix(H, W, H, pic);
    timeElapsed->working =
        std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0).count();
    return true;
}

bool HipKSvdDenoiser::groupable() const
{
    return type != HIP_K_SVD_BF16 && type != HIP_K_SVD_BF16_LEARNED && (subImageWidthDim == 0 || subImageHeightDim == 0);
}

std::string HipKSvdDenoiser::groupKey() const
{
    return std::to_string(patchWidthDim) + "," + std::to_string(patchHeightDim) + "," + std::to_string(slidingWidth) +
           "," + std::to_string(slidingHeight) + "," + std::to_string(atoms) + "," + std::to_string(ompIter) + "," +
           std::to_string(iter) + "," + std::to_string(userCap) + "," + std::to_string(geometry) + "," +
           std::to_string(device);
}

bool HipKSvdDenoiser::finishFrame(ksvd_ctx *ctx, int W, int H)
{
    int rc;
    ksvd_timings t;
    if ((rc = ksvd_get_timings(ctx, &t))) {
        error = std::string("HIP backend: ") + ksvd_last_error(ctx) + " (status " + std::to_string(rc) + ")";
        return false;
    }
    int dim = 0, K = 0, kmax = 0;
    int64_t P = 0;
    ksvd_get_dims(ctx, &dim, &P, &K, &kmax);
    nPatches = P;
    iterMs = t.omp_ms + t.update_ms;
    patchIterCount = (double)P * iter;
    patchIters = iterMs > 0 ? patchIterCount / (iterMs * 1e-3) : 0;
    std::vector<float> out((size_t)W * H, 0.f);
    if ((rc = ksvd_rebuild(ctx, sigma, out.data()))) {
        error = std::string("HIP backend: ") + ksvd_last_error(ctx) + " (status " + std::to_string(rc) + ")";
        return false;
    }
    if (speckle) ksvd_speckle_exp(out.data(), (int64_t)out.size());
    std::vector<float> *pic = new std::vector<float>((size_t)W * H);
    ksvd_frame_to_image(out.data(), W, H, pic->data());
    outputMatrix = new Matrix(H, W, H, pic);
    return true;
}

void HipKSvdDenoiser::denoiseGroups(const std::vector<Denoiser *> &list)
{
    // the groupable files by key, in file order
    std::vector<std::pair<std::string, std::vector<HipKSvdDenoiser *>>> groups;
    for (Denoiser *d : list) {
        auto *h = dynamic_cast<HipKSvdDenoiser *>(d);
        if (!h || !h->groupable()) continue;
        const std::string key = h->groupKey();
        auto it = std::find_if(groups.begin(), groups.end(), [&](const auto &g) { return g.first == key; });
        if (it == groups.end()) groups.push_back({key, {h}});
        else it->second.push_back(h);
    }
    const int per = 8;  // pictures per sweep launch: one per XCD
    for (auto &g : groups) {
        for (size_t g0 = 0; g0 + 1 < g.second.size(); g0 += per) {
            const size_t gn = std::min(g.second.size() - g0, (size_t)per);
            if (gn < 2) break;
            std::vector<HipKSvdDenoiser *> mem(g.second.begin() + g0, g.second.begin() + g0 + gn);
            std::vector<HipKSvdDenoiser *> live;
            std::vector<ksvd_ctx *> ctxs;
            for (HipKSvdDenoiser *h : mem) {
                h->grouped = true;
                if (!h->loadImage()) {
                    h->groupStatus = -1;
                    continue;
                }
                ksvd_ctx *ctx = h->makeContext();
                int rc = ctx ? 0 : -1;
                const int W = h->inputMatrix->m, H = h->inputMatrix->n;
                if (!rc && ((rc = ksvd_set_image(ctx, h->inputMatrix->hostVector->data(), W, H)) ||
                            (rc = ksvd_extract_patches(ctx)) || (rc = ksvd_init_dictionary(ctx)) ||
                            (rc = ksvd_reset_timings(ctx)))) {
                    h->error = std::string("HIP backend: ") + ksvd_last_error(ctx) + " (status " + std::to_string(rc) + ")";
                    ksvd_destroy(ctx);
                    ctx = nullptr;
                }
                if (!ctx) {
                    h->groupStatus = -2;
                    continue;
                }
                live.push_back(h);
                ctxs.push_back(ctx);
            }
            if (!ctxs.empty()) {
                const auto t1 = std::chrono::steady_clock::now();
                const int rc = ksvd_run_batch(ctxs.data(), (int)ctxs.size());
                const double ms =
                    (double)std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t1)
                        .count() * 1e-3;
                for (size_t i = 0; i < live.size(); i++) {
                    HipKSvdDenoiser *h = live[i];
                    h->groupN = (int)live.size();
                    if (rc) {
                        std::string msg;
                        for (ksvd_ctx *c : ctxs)
                            if (msg.empty()) msg = ksvd_last_error(c);
                        h->error = "HIP backend: " + msg + " (status " + std::to_string(rc) + ")";
                        h->groupStatus = -2;
                        continue;
                    }
                    const auto t2 = std::chrono::steady_clock::now();
                    if (!h->finishFrame(ctxs[i], h->inputMatrix->m, h->inputMatrix->n)) {
                        h->groupStatus = -2;
                        continue;
                    }
                    // working time: the shared K-SVD shared out evenly, plus this picture's rebuild
                    h->timeElapsed->working =
                        (int64_t)(ms / (double)live.size()) +
                        std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t2)
                            .count();
                    h->groupStatus = h->saveImage() ? 0 : -3;
                }
            }
            for (ksvd_ctx *c : ctxs) ksvd_destroy(c);
        }
    }
}
