// Host-side helpers of the C ABI: seeded synthetic images, the CImg
// statistics the reference's Denoiser uses (noise variance, normalize, PSNR),
// and the image <-> frame transposes.  Plain C++17, no GPU.
#include "ksvd_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

namespace {

// splitmix64 (Steele, Lea, Flood 2014) seeds xoshiro256** (Blackman, Vigna
// 2018); both written from the published descriptions.
uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Xoshiro256 {
    uint64_t s[4];
    explicit Xoshiro256(uint64_t seed)
    {
        uint64_t sm = seed;
        for (auto &w : s) w = splitmix64(sm);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // uniform in [0, 1) with 53 random bits
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    // standard normal by the Box-Muller transform (one value per pair)
    double normal()
    {
        double u1 = 1.0 - uniform();  // (0, 1]
        double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

}  // namespace

extern "C" int ksvd_synth_image(uint64_t seed, int W, int H, float sigma, float *clean, float *noisy)
{
    if (W <= 0 || H <= 0 || sigma < 0) return KSVD_ERR_ARG;
    Xoshiro256 rng(seed);
    const size_t N = (size_t)W * H;
    std::vector<double> f(N, 0.0);
    // six oriented sinusoidal gratings: angle U[0,pi), period U[4,64], phase U[0,2pi)
    for (int g = 0; g < 6; g++) {
        const double th = rng.uniform() * M_PI;
        const double period = 4.0 + rng.uniform() * 60.0;
        const double ph = rng.uniform() * 2.0 * M_PI;
        const double kx = std::cos(th) * 2.0 * M_PI / period;
        const double ky = std::sin(th) * 2.0 * M_PI / period;
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) f[(size_t)y * W + x] += std::sin(kx * x + ky * y + ph);
    }
    double fmin = f[0], fmax = f[0];
    for (double v : f) { fmin = std::min(fmin, v); fmax = std::max(fmax, v); }
    // twenty constant rectangles, sides U[dim/16, dim/4], value U[fmin, fmax]
    for (int r = 0; r < 20; r++) {
        const int rw = std::max(1, (int)(W / 16.0 + rng.uniform() * (W / 4.0 - W / 16.0)));
        const int rh = std::max(1, (int)(H / 16.0 + rng.uniform() * (H / 4.0 - H / 16.0)));
        const int x0 = (int)(rng.uniform() * (W - rw + 1));
        const int y0 = (int)(rng.uniform() * (H - rh + 1));
        const double v = fmin + rng.uniform() * (fmax - fmin);
        for (int y = y0; y < std::min(H, y0 + rh); y++)
            for (int x = x0; x < std::min(W, x0 + rw); x++) f[(size_t)y * W + x] = v;
    }
    double lo = f[0], hi = f[0];
    for (double v : f) { lo = std::min(lo, v); hi = std::max(hi, v); }
    const double scale = hi > lo ? 255.0 / (hi - lo) : 0.0;
    for (size_t q = 0; q < N; q++) {
        const double c = (f[q] - lo) * scale;
        if (clean) clean[q] = (float)c;
        const double z = rng.normal();
        if (noisy) noisy[q] = (float)(c + (double)sigma * z);
    }
    return KSVD_OK;
}

// CImg's variance_noise() default (variance_method = 2): each channel is
// replaced by its 4-neighbour Laplacian scaled by 1/sqrt(20) (so that white
// noise keeps its variance), borders replicated, and the variance is the
// squared MAD estimate (1.4828 * median |L - median(L)|)^2 over the values
// of every channel together, medians taken as element siz/2 of the sorted
// values.  img holds C planes of W x H, row-major.
extern "C" double ksvd_noise_variance_channels(const float *img, int W, int H, int C)
{
    if (!img || W <= 0 || H <= 0 || C <= 0) return 0.0;
    const size_t np = (size_t)W * H, N = np * (size_t)C;
    std::vector<double> L(N);
    const double cste = 1.0 / std::sqrt(20.0);
    for (int c = 0; c < C; c++) {
        const float *p = img + (size_t)c * np;
        double *l = L.data() + (size_t)c * np;
        for (int y = 0; y < H; y++) {
            const int yp = y > 0 ? y - 1 : 0, yn = y < H - 1 ? y + 1 : H - 1;
            for (int x = 0; x < W; x++) {
                const int xp = x > 0 ? x - 1 : 0, xn = x < W - 1 ? x + 1 : W - 1;
                const double v = p[(size_t)y * W + x];
                l[(size_t)y * W + x] = cste * ((double)p[(size_t)y * W + xn] + (double)p[(size_t)y * W + xp] +
                                               (double)p[(size_t)yn * W + x] + (double)p[(size_t)yp * W + x] - 4 * v);
            }
        }
    }
    const size_t mid = N >> 1;
    std::nth_element(L.begin(), L.begin() + mid, L.end());
    const double med = L[mid];
    for (auto &v : L) v = std::fabs(v - med);
    std::nth_element(L.begin(), L.begin() + mid, L.end());
    const double sig = 1.4828 * L[mid];
    const double var = sig * sig;
    return var > 0 ? var : 0;
}

extern "C" double ksvd_noise_variance(const float *img, int W, int H)
{
    return ksvd_noise_variance_channels(img, W, H, 1);
}

// CImg normalize(a, b): min-max stretch in float; a constant image is filled
// with the lower bound.
extern "C" void ksvd_normalize(float *img, int64_t count, float lo, float hi)
{
    if (!img || count <= 0) return;
    const float a = lo < hi ? lo : hi, b = lo < hi ? hi : lo;
    float m = img[0], M = img[0];
    for (int64_t q = 0; q < count; q++) { m = std::min(m, img[q]); M = std::max(M, img[q]); }
    if (m == M) {
        for (int64_t q = 0; q < count; q++) img[q] = a;
        return;
    }
    if (m == a && M == b) return;
    for (int64_t q = 0; q < count; q++) img[q] = (img[q] - m) / (M - m) * (b - a) + a;
}

extern "C" double ksvd_psnr(const float *ref, const float *img, int64_t count)
{
    if (!ref || !img || count <= 0) return -1.0;
    double s = 0;
    for (int64_t q = 0; q < count; q++) {
        const double d = (double)ref[q] - (double)img[q];
        s += d * d;
    }
    const double mse = s / (double)count;
    if (mse == 0) return INFINITY;
    return 10.0 * std::log10(255.0 * 255.0 / mse);
}

// CImg RGBtoYCbCr() on 8-bit-scale planes (BT.601, studio swing), each
// component clamped to [0, 255]; cb and cr may be NULL.
extern "C" void ksvd_rgb_to_ycbcr(const float *r, const float *g, const float *b, int64_t count, float *y, float *cb,
                                  float *cr)
{
    if (!r || !g || !b || !y || count <= 0) return;
    auto cut = [](float v) { return std::min(255.f, std::max(0.f, v)); };
    for (int64_t q = 0; q < count; q++) {
        const float R = r[q], G = g[q], B = b[q];
        y[q] = cut((66 * R + 129 * G + 25 * B + 128) / 256 + 16);
        if (cb) cb[q] = cut((-38 * R - 74 * G + 112 * B + 128) / 256 + 128);
        if (cr) cr[q] = cut((112 * R - 94 * G - 18 * B + 128) / 256 + 128);
    }
}

extern "C" void ksvd_rgb_to_y(const float *r, const float *g, const float *b, int64_t count, float *y)
{
    ksvd_rgb_to_ycbcr(r, g, b, count, y, nullptr, nullptr);
}

namespace {

// The speckle transforms of the reference's host path (myLog / myExp,
// denoisingLib.h:124-142): log2 with 0 kept at 0, and 2^v.  Each element is
// computed alone by the same libm call, so splitting a large buffer over
// host threads (up to 16, a million elements or more each) gives the same
// bits as one thread.
template <class F>
void each_element(float *buf, int64_t count, F f)
{
    const int64_t chunk = 1 << 20;
    int64_t nt = count / chunk;
    const int64_t hw = std::max(1u, std::thread::hardware_concurrency());
    nt = std::min<int64_t>(std::min<int64_t>(nt, hw), 16);
    if (nt <= 1) {
        for (int64_t q = 0; q < count; q++) buf[q] = f(buf[q]);
        return;
    }
    std::vector<std::thread> th;
    const int64_t per = (count + nt - 1) / nt;
    for (int64_t t = 0; t < nt; t++) {
        const int64_t a = t * per, b = std::min(count, a + per);
        th.emplace_back([=] {
            for (int64_t q = a; q < b; q++) buf[q] = f(buf[q]);
        });
    }
    for (auto &x : th) x.join();
}

}  // namespace

extern "C" void ksvd_speckle_log(float *buf, int64_t count)
{
    if (!buf || count <= 0) return;
    each_element(buf, count, [](float v) { return v == 0.f ? 0.f : log2f(v); });
}

extern "C" void ksvd_speckle_exp(float *buf, int64_t count)
{
    if (!buf || count <= 0) return;
    each_element(buf, count, [](float v) { return powf(2.f, v); });
}

extern "C" void ksvd_image_to_frame(const float *img, int W, int H, float *buf)
{
    for (int x = 0; x < W; x++)
        for (int y = 0; y < H; y++) buf[(size_t)x * H + y] = img[(size_t)y * W + x];
}

extern "C" void ksvd_frame_to_image(const float *buf, int W, int H, float *img)
{
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) img[(size_t)y * W + x] = buf[(size_t)x * H + y];
}
