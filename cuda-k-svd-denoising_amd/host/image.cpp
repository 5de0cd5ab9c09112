#include "image.h"
#include "ksvd_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>

namespace img {

namespace {

bool ends_with(const std::string &s, const std::string &suf)
{
    if (s.size() < suf.size()) return false;
    std::string tail = s.substr(s.size() - suf.size());
    std::transform(tail.begin(), tail.end(), tail.begin(), ::tolower);
    return tail == suf;
}

// next header token of a PNM file, skipping comments
bool token(std::istream &in, std::string &t)
{
    t.clear();
    int c;
    while ((c = in.get()) != EOF) {
        if (c == '#') {
            while ((c = in.get()) != EOF && c != '\n') {}
            continue;
        }
        if (std::isspace(c)) {
            if (!t.empty()) return true;
            continue;
        }
        t += (char)c;
    }
    return !t.empty();
}

bool load_pnm(std::istream &in, const std::string &magic, Image &out, std::string &err)
{
    std::string tw, th, tm;
    if (!token(in, tw) || !token(in, th) || !token(in, tm)) { err = "truncated PNM header"; return false; }
    const int W = std::atoi(tw.c_str()), H = std::atoi(th.c_str()), maxv = std::atoi(tm.c_str());
    if (W <= 0 || H <= 0 || maxv <= 0 || maxv > 65535) { err = "bad PNM header"; return false; }
    const int C = (magic == "P3" || magic == "P6") ? 3 : 1;
    out.W = W; out.H = H; out.C = C;
    out.data.assign((size_t)W * H * C, 0.f);
    const size_t np = (size_t)W * H;
    if (magic == "P2" || magic == "P3") {
        std::string t;
        for (size_t q = 0; q < np; q++)
            for (int c = 0; c < C; c++) {
                if (!token(in, t)) { err = "truncated PNM data"; return false; }
                out.data[c * np + q] = (float)std::atoi(t.c_str());
            }
        return true;
    }
    const int bpv = maxv > 255 ? 2 : 1;
    std::vector<unsigned char> raw(np * C * bpv);
    in.read((char *)raw.data(), (std::streamsize)raw.size());
    if ((size_t)in.gcount() != raw.size()) { err = "truncated PNM data"; return false; }
    for (size_t q = 0; q < np; q++)
        for (int c = 0; c < C; c++) {
            const size_t o = (q * C + c) * bpv;
            const unsigned v = bpv == 2 ? (unsigned)(raw[o] << 8 | raw[o + 1]) : raw[o];
            out.data[c * np + q] = (float)v;
        }
    return true;
}

bool load_pfm(std::istream &in, const std::string &magic, Image &out, std::string &err)
{
    std::string tw, th, ts;
    if (!token(in, tw) || !token(in, th) || !token(in, ts)) { err = "truncated PFM header"; return false; }
    const int W = std::atoi(tw.c_str()), H = std::atoi(th.c_str());
    const double scale = std::atof(ts.c_str());
    if (W <= 0 || H <= 0 || scale == 0) { err = "bad PFM header"; return false; }
    const int C = magic == "PF" ? 3 : 1;
    const bool little = scale < 0;
    out.W = W; out.H = H; out.C = C;
    out.data.assign((size_t)W * H * C, 0.f);
    const size_t np = (size_t)W * H;
    std::vector<float> row((size_t)W * C);
    for (int y = H - 1; y >= 0; y--) {  // PFM rows run bottom to top
        in.read((char *)row.data(), (std::streamsize)(row.size() * 4));
        if ((size_t)in.gcount() != row.size() * 4) { err = "truncated PFM data"; return false; }
        for (size_t k = 0; k < row.size(); k++) {
            uint32_t u;
            memcpy(&u, &row[k], 4);
            if (!little) u = __builtin_bswap32(u);
            float v;
            memcpy(&v, &u, 4);
            const int x = (int)(k / C), c = (int)(k % C);
            out.data[c * np + (size_t)y * W + x] = v;
        }
    }
    return true;
}

}  // namespace

bool load(const std::string &path, Image &out, std::string &err)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) { err = "cannot open " + path; return false; }
    std::string magic;
    if (!token(in, magic)) { err = "empty file " + path; return false; }
    if (magic == "P2" || magic == "P3" || magic == "P5" || magic == "P6") return load_pnm(in, magic, out, err);
    if (magic == "Pf" || magic == "PF") return load_pfm(in, magic, out, err);
    err = "unsupported image format in " + path + " (PGM/PPM/PFM are read)";
    return false;
}

bool save(const std::string &path, const Image &im, std::string &err)
{
    std::ofstream out(path, std::ios::binary);
    if (!out) { err = "cannot write " + path; return false; }
    const size_t np = im.plane();
    if (ends_with(path, ".pfm")) {
        if (im.C != 1 && im.C != 3) { err = "PFM needs 1 or 3 channels"; return false; }
        out << (im.C == 3 ? "PF" : "Pf") << "\n" << im.W << " " << im.H << "\n-1.0\n";
        std::vector<float> row((size_t)im.W * im.C);
        for (int y = im.H - 1; y >= 0; y--) {
            for (int x = 0; x < im.W; x++)
                for (int c = 0; c < im.C; c++) row[(size_t)x * im.C + c] = im.data[c * np + (size_t)y * im.W + x];
            out.write((const char *)row.data(), (std::streamsize)(row.size() * 4));
        }
        return (bool)out;
    }
    const bool sixteen = ends_with(path, ".pgm16");
    const int C = im.C >= 3 ? 3 : 1;
    out << (C == 3 ? "P6" : "P5") << "\n" << im.W << " " << im.H << "\n" << (sixteen ? 65535 : 255) << "\n";
    std::vector<unsigned char> raw(np * C * (sixteen ? 2 : 1));
    for (size_t q = 0; q < np; q++)
        for (int c = 0; c < C; c++) {
            float v = im.data[c * np + q];
            if (sixteen) {
                const unsigned u = (unsigned)std::lround(std::min(65535.f, std::max(0.f, v)));
                raw[(q * C + c) * 2] = (unsigned char)(u >> 8);
                raw[(q * C + c) * 2 + 1] = (unsigned char)(u & 255);
            } else {
                raw[q * C + c] = (unsigned char)std::lround(std::min(255.f, std::max(0.f, v)));
            }
        }
    out.write((const char *)raw.data(), (std::streamsize)raw.size());
    return (bool)out;
}

Image transpose(const Image &in)
{
    Image o;
    o.W = in.H;
    o.H = in.W;
    o.C = in.C;
    o.data.resize(in.data.size());
    const size_t np = in.plane();
    for (int c = 0; c < in.C; c++)
        for (int y = 0; y < in.H; y++)
            for (int x = 0; x < in.W; x++) o.data[c * np + (size_t)x * o.W + y] = in.data[c * np + (size_t)y * in.W + x];
    return o;
}

Image rgb_to_ycbcr(const Image &in)
{
    Image o;
    o.W = in.W;
    o.H = in.H;
    o.C = 1;
    const size_t np = in.plane();
    o.data.resize(np);
    const float *R = in.data.data();
    const float *G = in.C >= 3 ? R + np : R, *B = in.C >= 3 ? R + 2 * np : R;
    o.C = 3;
    o.data.resize(3 * np);
    ksvd_rgb_to_ycbcr(R, G, B, (int64_t)np, o.data.data(), o.data.data() + np, o.data.data() + 2 * np);
    return o;
}

}  // namespace img
