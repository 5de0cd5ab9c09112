// denoising: the reference's pipeline interface (src/lib/denoisingLib/
// denoisingLib.h:13-122) with the gfx950 backend behind it.
//
// What stays as in the reference: the DenoiserType tag and its string map,
// Denoiser::factory(type, input, output), denoising() with status 0 / -1 load
// / -2 denoise / -3 save, getTimeElapsed(), getPsnr(), the protected
// per-image parameters with the same defaults, and BatchDenoiser::factory
// (config.json) with getDenoiserList().
// What changes: HIP_K_SVD (plain FP32) and its precision / learned variants
// are new tags; the CUDA_K_* tags run the same HIP rank-1 engine (their SVD
// choice only changes how the top singular triple is computed); an unknown
// tag is an error instead of a silent CUDA_K_GESVD.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "baseUtilityLib.h"
#include "ksvd_hip.h"

namespace img {
struct Image;
}

namespace denoising {

class Denoiser;
class HipKSvdDenoiser;
class BatchDenoiser;

enum DenoiserType {
    CUDA_K_GESVD,
    CUDA_K_GESVDJ,
    CUDA_K_GESVDA,
    HIP_K_SVD,               // plain FP32 on gfx950: the drop-in beside CUDA_K_GESVDJ
    HIP_K_SVD_BF16,          // BF16 correlation on MFMA, FP32 re-score / LS / SVD
    HIP_K_SVD_BF16_LEARNED,  // BF16 + learned support proposer, OMP fallback
    DENOISER_TYPE_UNKNOWN
};

struct DenoiserTypeMap : public std::map<std::string, DenoiserType> {
    DenoiserTypeMap()
    {
        (*this)["CUDA_K_GESVD"] = CUDA_K_GESVD;
        (*this)["CUDA_K_GESVDJ"] = CUDA_K_GESVDJ;
        (*this)["CUDA_K_GESVDA"] = CUDA_K_GESVDA;
        (*this)["HIP_K_SVD"] = HIP_K_SVD;
        (*this)["HIP_K_SVD_BF16"] = HIP_K_SVD_BF16;
        (*this)["HIP_K_SVD_BF16_LEARNED"] = HIP_K_SVD_BF16_LEARNED;
    }
    // The reference's operator[] default-inserts CUDA_K_GESVD for a typo
    // (batch_denoiser.cpp:96); lookup() reports it instead.
    DenoiserType lookup(const std::string &name) const
    {
        auto it = find(name);
        return it == end() ? DENOISER_TYPE_UNKNOWN : it->second;
    }
};

}  // namespace denoising

class denoising::Denoiser {
public:
    virtual ~Denoiser();
    static Denoiser *factory(denoising::DenoiserType, std::string inputFile, std::string outputFile);
    virtual signed char denoising() = 0;
    baseUtl::TimeElapsed *getTimeElapsed();
    std::vector<double> *getPsnr();
    // Message of the last failure ("" if none).
    const std::string &getError() const { return error; }

protected:
    int patchWidthDim = 8;
    int patchHeightDim = 8;
    int slidingWidth = 2;
    int slidingHeight = 2;
    int subImageWidthDim = 0;
    int subImageHeightDim = 0;
    int ompIter = 5;
    int atoms = 256;
    int iter = 10;
    int minOmpIterBatch = 0;
    float sigma = 0;  // noise variance estimated at load (denoiser.cpp:56)
    bool speckle = false;
    bool bw = true;
    std::string refImage;
    // backend knobs (optional config keys userCap, geometry, device)
    int userCap = 1024;
    int geometry = 0;
    int device = 0;
    baseUtl::Matrix *inputMatrix = nullptr;
    baseUtl::Matrix *outputMatrix = nullptr;
    baseUtl::TimeElapsed *timeElapsed = nullptr;
    std::vector<double> *psnr = nullptr;
    std::string error;

    Denoiser();
    virtual bool loadImage();
    virtual bool saveImage();
    virtual bool internalDenoising() = 0;

private:
    std::string inputFile, outputFile;
    img::Image *inputImage = nullptr;

    friend BatchDenoiser;
};

class denoising::HipKSvdDenoiser : public denoising::Denoiser {
public:
    ~HipKSvdDenoiser() override;
    signed char denoising() override;
    // Denoises the batch's files whose K-SVD can share dictionary-sweep
    // launches (the plain form, no sub-images, the same patch and sliding
    // sizes, atoms, OMP and K-SVD iterations, user cap, geometry and device),
    // up to eight at a time (one per XCD) through ksvd_run_batch, loading and
    // saving each as denoising() does; their later denoising() calls return
    // the stored status.  Each output is the bits denoising() gives alone.
    static void denoiseGroups(const std::vector<Denoiser *> &list);
    // Throughput of the last run: patches * ksvdIter / time of the K-SVD
    // iterations (device events), and the patch count.
    double patchItersPerSec() const { return patchIters; }
    long long patches() const { return nPatches; }
    // Pictures of the sweep launches this file's K-SVD shared (1: alone).
    int groupSize() const { return groupN; }

protected:
    bool loadImage() override;
    bool saveImage() override;
    bool internalDenoising() override;

private:
    DenoiserType type = HIP_K_SVD;
    double patchIters = 0;
    long long nPatches = 0;
    double iterMs = 0, patchIterCount = 0;

    bool grouped = false;         // denoised by denoiseGroups
    int groupN = 1;               // pictures of its group
    signed char groupStatus = 0;  // its denoising() status

    HipKSvdDenoiser();
    // Whether this file's K-SVD can share sweep launches with others, and
    // the key of the files it can share them with.
    bool groupable() const;
    std::string groupKey() const;
    // The rest of internalDenoising after the K-SVD (final code done): the
    // rebuild from ctx into the output picture.
    bool finishFrame(ksvd_ctx *ctx, int W, int H);
    // A context with this denoiser's parameters (nullptr and error on failure).
    ksvd_ctx *makeContext();
    // One frame buffer (m = W, n = H) through the whole K-SVD path on ctx;
    // out receives the denoised buffer in the same layout.
    bool denoiseFrame(ksvd_ctx *ctx, const float *buf, int m, int n, float noiseVar, std::vector<float> &out);
    // The sub-image path (cuda_svd_denoiser.cu:130-188, 503-528): the frame cut
    // into subImageHeightDim x subImageWidthDim tiles, each denoised on its own
    // with its own noise variance, and put back in place.
    bool denoiseTiles(const float *buf, int W, int H, std::vector<float> &out);

    friend Denoiser;
    friend BatchDenoiser;
};

class denoising::BatchDenoiser {
public:
    ~BatchDenoiser();
    std::vector<Denoiser *> getDenoiserList();
    // nullptr (and a message on stderr) when the file cannot be parsed, a
    // required key is missing, or a type tag is unknown.
    static BatchDenoiser *factory(std::string jsonFile);

protected:
    std::vector<Denoiser *> denoisers;

private:
    BatchDenoiser();
};
