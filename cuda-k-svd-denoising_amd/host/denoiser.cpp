// Denoiser base: image load / save / PSNR as the reference's Denoiser does
// them (src/lib/denoisingLib/denoiser.cpp:37-127), on PGM/PPM/PFM files.
#include <chrono>
#include <cmath>
#include <iostream>

#include "denoisingLib.h"
#include "image.h"
#include "ksvd_hip.h"

using namespace denoising;
using namespace baseUtl;

namespace {
int64_t ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0).count();
}

// Output path: the reference keeps the input's extension (CImg picks the
// codec); PNG/JPEG cannot be written here, so those names get ".pgm".
std::string writable_path(const std::string &p)
{
    auto dot = p.find_last_of('.');
    std::string ext = dot == std::string::npos ? "" : p.substr(dot);
    for (auto &c : ext) c = (char)tolower(c);
    if (ext == ".pgm" || ext == ".ppm" || ext == ".pfm" || ext == ".pgm16") return p;
    return (dot == std::string::npos ? p : p.substr(0, dot)) + ".pgm";
}
}  // namespace

Denoiser::Denoiser()
{
    timeElapsed = new TimeElapsed();
    psnr = new std::vector<double>(2, -1.0);
}

Denoiser::~Denoiser()
{
    delete inputImage;
    delete inputMatrix;
    delete outputMatrix;
    delete timeElapsed;
    delete psnr;
}

// denoiser.cpp:37-62: load, transpose, convert a colour picture (B&W false)
// to YCbCr in place and keep channel 0 (Y; channel 0 of the file when B&W),
// log2 for speckle (0 stays 0), and the noise variance of the whole
// (transposed, converted, un-logged) image: every channel's Laplacian
// pooled, as CImg's variance_noise does.
bool Denoiser::loadImage()
{
    auto t0 = std::chrono::steady_clock::now();
    img::Image in;
    if (!img::load(inputFile, in, error)) return false;
    img::Image t = img::transpose(in);
    if (!bw) t = img::rgb_to_ycbcr(t);
    inputImage = new img::Image(t);
    // Matrix(m = W, n = H, ld = W) over the transposed image's first plane
    inputMatrix = new Matrix(in.W, in.H, in.W, t.data.data());
    if (speckle) ksvd_speckle_log(inputMatrix->hostVector->data(), (int64_t)inputMatrix->hostVector->size());
    // the Laplacian estimator is symmetric under the transpose
    sigma = (float)ksvd_noise_variance_channels(t.data.data(), t.W, t.H, t.C);
    timeElapsed->init = ms_since(t0);
    return true;
}

// denoiser.cpp:68-99: min-max normalise to [0,255], save, PSNR of the noisy
// input and of the normalised output against the reference image.
bool Denoiser::saveImage()
{
    auto t0 = std::chrono::steady_clock::now();
    img::Image out;
    out.W = outputMatrix->n;
    out.H = outputMatrix->ld;
    out.C = 1;
    out.data = *outputMatrix->hostVector;
    ksvd_normalize(out.data.data(), (int64_t)out.data.size(), 0.f, 255.f);
    if (!img::save(writable_path(outputFile), out, error)) return false;
    timeElapsed->finalize = ms_since(t0);
    if (refImage.empty() || refImage.back() == '/') {
        (*psnr)[0] = -1;
        (*psnr)[1] = -1;
        return true;
    }
    img::Image ref, old;
    std::string e;
    if (!img::load(refImage, ref, e) || !img::load(inputFile, old, e) || ref.plane() != out.plane() ||
        old.plane() != out.plane()) {
        std::cerr << "PSNR skipped: " << (e.empty() ? "reference size differs" : e) << std::endl;
        (*psnr)[0] = -1;
        (*psnr)[1] = -1;
        return true;
    }
    (*psnr)[0] = ksvd_psnr(ref.data.data(), old.data.data(), (int64_t)out.plane());
    (*psnr)[1] = ksvd_psnr(ref.data.data(), out.data.data(), (int64_t)out.plane());
    return true;
}

// denoiser.cpp:108-127.  Every tag builds the HIP backend; unknown -> nullptr.
Denoiser *Denoiser::factory(DenoiserType type, std::string inputFile, std::string outputFile)
{
    Denoiser *denoiser = nullptr;
    switch (type) {
    case CUDA_K_GESVD:
    case CUDA_K_GESVDJ:
    case CUDA_K_GESVDA:
    case HIP_K_SVD:
    case HIP_K_SVD_BF16:
    case HIP_K_SVD_BF16_LEARNED: {
        HipKSvdDenoiser *d = new HipKSvdDenoiser();
        d->type = type;
        denoiser = d;
        break;
    }
    default:
        return nullptr;
    }
    denoiser->inputFile = inputFile;
    denoiser->outputFile = outputFile;
    return denoiser;
}

TimeElapsed *Denoiser::getTimeElapsed() { return timeElapsed; }

std::vector<double> *Denoiser::getPsnr() { return psnr; }
