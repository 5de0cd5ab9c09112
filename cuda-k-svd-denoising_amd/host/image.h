// Image files and the few CImg operations the reference's Denoiser uses
// (denoiser.cpp:37-99): load, transpose, RGB -> YCbCr, save.  Formats: PGM/PPM
// (P2, P3, P5, P6; 8 or 16 bit) and PFM (Pf grayscale, PF colour; float).
// Pixels are stored planar as CImg does: data[c*W*H + y*W + x].
#pragma once

#include <string>
#include <vector>

namespace img {

struct Image {
    int W = 0, H = 0, C = 0;
    std::vector<float> data;
    size_t plane() const { return (size_t)W * H; }
};

bool load(const std::string &path, Image &out, std::string &err);
// Extension .pfm writes floats; .pgm / .ppm write 8-bit (rounded, clamped to
// [0,255]) or, for .pgm16, 16-bit.
bool save(const std::string &path, const Image &im, std::string &err);

// CImg::transpose(): out(x, y) = in(y, x), per channel.
Image transpose(const Image &in);
// CImg::RGBtoYCbCr(): planes Y, Cb, Cr (BT.601 on the 8-bit scale, each
// clamped to [0, 255]; ksvd_rgb_to_ycbcr).  A one-channel input is taken as
// R = G = B.
Image rgb_to_ycbcr(const Image &in);

}  // namespace img
