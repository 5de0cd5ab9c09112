// Patch gather, dictionary init and dictionary Gram kernels (gfx950).
//
// Gather: one workgroup per (patch row, run of TP patches along that row).
// The ph image rows the run touches are staged in LDS with coalesced loads,
// then the run's TP*dim outputs -- one contiguous range of Y -- are written
// with coalesced stores.  Pure copies: bit-exact against the oracle.
#include "ksvd_internal.hpp"

namespace {

__global__ __launch_bounds__(256) void gather_patches_kernel(const float *__restrict__ img, int m,
                                                             int pw, int ph, int sw, int sh, int Pc,
                                                             int TP, int TW, float *__restrict__ Y)
{
    extern __shared__ __attribute__((aligned(16))) float tile[];  // ph x TW
    const int pr = blockIdx.y;
    const int pc0 = blockIdx.x * TP;
    const int np = min(TP, Pc - pc0);
    const int a0 = pr * sh, b0 = pc0 * sw;
    const int tw = (np - 1) * sw + pw;  // columns this run touches
    for (int q = threadIdx.x; q < ph * tw; q += blockDim.x) {
        const int r = q / tw, c = q - r * tw;
        tile[r * TW + c] = img[(size_t)(a0 + r) * m + b0 + c];
    }
    __syncthreads();
    const int dim = pw * ph;
    float *out = Y + ((size_t)pr * Pc + pc0) * dim;
    for (int o = threadIdx.x; o < np * dim; o += blockDim.x) {
        const int qp = o / dim, e = o - qp * dim;
        const int r = e / pw, c = e - r * pw;
        out[o] = tile[r * TW + qp * sw + c];
    }
}

// One thread per atom for its norm, then the block's atoms divided and
// stored together.  The squared norm is accumulated the way the reference's
// in-kernel thrust::transform_reduce(square, 0, plus<float>) does it
// (cuda_svd_denoiser.cu:59): the int init value makes the running sum an int,
// so each float partial sum is truncated toward zero.  The elements are
// loaded 16 at a time ahead of that chain; the division pass walks the
// block's rows (consecutive patches, so one contiguous run of Y and of D)
// row by row, each row's elements across the threads.
__global__ __launch_bounds__(256) void init_dict_kernel(const float *__restrict__ Y, int dim, int64_t off, int K,
                                                        float *__restrict__ D)
{
    __shared__ float snorm[256];
    const int k0 = blockIdx.x * 256;
    const int k = k0 + threadIdx.x;
    if (k < K) {
        const float *y = Y + (off + k) * (int64_t)dim;
        constexpr int GB = 16;
        int acc = 0;
        int e = 0;
        for (; e + GB <= dim; e += GB) {
            float v[GB];
#pragma unroll
            for (int i = 0; i < GB; i++) v[i] = y[e + i];
#pragma unroll
            for (int i = 0; i < GB; i++) {
                const float sq = v[i] * v[i];
                acc = (int)((float)acc + sq);
            }
        }
        for (; e < dim; e++) {
            const float sq = y[e] * y[e];
            acc = (int)((float)acc + sq);
        }
        snorm[threadIdx.x] = sqrtf((float)acc);
    }
    __syncthreads();
    const int na = min(256, K - k0);
    const float *yb = Y + (off + k0) * (int64_t)dim;
    float *db = D + (int64_t)k0 * dim;
#pragma unroll 4
    for (int a = 0; a < na; a++) {
        const float nr = snorm[a];
        for (int e = threadIdx.x; e < dim; e += 256) db[(int64_t)a * dim + e] = yb[(int64_t)a * dim + e] / nr;
    }
}

// G = D^T D (K x K).  One thread per Gram entry; element order, fused
// multiply-adds.  The elements are loaded 16 at a time ahead of their
// products (VEC: as 16-byte loads, for rows that stay 16-byte aligned), so
// that a thread waits for one round of loads per 16 products instead of one
// per product; the sum is the same chain, element by element.
template <bool VEC>
__global__ void dict_gram_kernel(const float *__restrict__ D, int dim, int K, float *__restrict__ G)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)K * K) return;
    const int a = (int)(q / K), b = (int)(q - (int64_t)a * K);
    const float *da = D + (int64_t)a * dim, *db = D + (int64_t)b * dim;
    // accumulated in FP64 and rounded once: the pursuit's support Grams have
    // condition numbers near 1e3 on a dictionary cut from overlapping
    // patches, where an FP32 running sum's error would reach the codes
    constexpr int GB = 16;
    double s = 0.0;
    int e = 0;
    for (; e + GB <= dim; e += GB) {
        float xa[GB], xb[GB];
        if (VEC) {
#pragma unroll
            for (int i = 0; i < GB; i += 4) {
                const float4 va = *(const float4 *)(da + e + i), vb = *(const float4 *)(db + e + i);
                xa[i] = va.x, xa[i + 1] = va.y, xa[i + 2] = va.z, xa[i + 3] = va.w;
                xb[i] = vb.x, xb[i + 1] = vb.y, xb[i + 2] = vb.z, xb[i + 3] = vb.w;
            }
        } else {
#pragma unroll
            for (int i = 0; i < GB; i++) {
                xa[i] = da[e + i];
                xb[i] = db[e + i];
            }
        }
#pragma unroll
        for (int i = 0; i < GB; i++) s = fma((double)xa[i], (double)xb[i], s);
    }
    for (; e < dim; e++) s = fma((double)da[e], (double)db[e], s);
    G[q] = (float)s;
}

}  // namespace

int launch_gather(ksvd_ctx *ctx)
{
    const int pw = ctx->prm.patch_w, ph = ctx->prm.patch_h;
    const int sw = ctx->prm.slide_w, sh = ctx->prm.slide_h;
    // runs of TP patches, the staged tile kept under 32 KB
    int TP = ((8192 / ph) - pw) / sw + 1;
    TP = TP < 1 ? 1 : (TP > 64 ? 64 : TP);
    const int TW = (TP - 1) * sw + pw;
    const size_t lds = sizeof(float) * (size_t)ph * TW;
    if (lds > 64 * 1024) return ksvd_fail(ctx, KSVD_ERR_UNSUPPORTED, "patch too large for the gather tile");
    dim3 grid((ctx->Pc + TP - 1) / TP, ctx->Pr);
    hipLaunchKernelGGL(gather_patches_kernel, grid, dim3(256), lds, ctx->stream, (const float *)ctx->img.ptr,
                       ctx->m, pw, ph, sw, sh, ctx->Pc, TP, TW, (float *)ctx->Y.ptr);
    KSVD_HIP_CHECK(ctx, hipGetLastError());
    return KSVD_OK;
}

int launch_init_dict(ksvd_ctx *ctx)
{
    hipLaunchKernelGGL(init_dict_kernel, dim3((ctx->K + 255) / 256), dim3(256), 0, ctx->stream,
                       (const float *)ctx->Y.ptr, ctx->dim, ctx->P / 2, ctx->K, (float *)ctx->D.ptr);
    KSVD_HIP_CHECK(ctx, hipGetLastError());
    return KSVD_OK;
}

// 16-byte loads of the Gram's rows: dim a multiple of 4 on a 16-byte aligned D.
static bool gram_vec(const float *D, int dim) { return dim % 4 == 0 && (uintptr_t)D % 16 == 0; }

int launch_dict_gram(ksvd_ctx *ctx)
{
    const int64_t kk = (int64_t)ctx->K * ctx->K;
    auto kern = gram_vec((const float *)ctx->D.ptr, ctx->dim) ? dict_gram_kernel<true> : dict_gram_kernel<false>;
    hipLaunchKernelGGL(kern, dim3((unsigned)((kk + 255) / 256)), dim3(256), 0, ctx->stream,
                       (const float *)ctx->D.ptr, ctx->dim, ctx->K, (float *)ctx->G.ptr);
    KSVD_HIP_CHECK(ctx, hipGetLastError());
    ctx->gram_valid = true;
    return KSVD_OK;
}

// G = D^T D of a dictionary given by pointer (K atoms of dim floats), on
// `stream`; the allocation-free pursuit's Gram (omp.hip).
hipError_t launch_dict_gram_raw(const float *D, int dim, int K, float *G, hipStream_t stream)
{
    const int64_t kk = (int64_t)K * K;
    auto kern = gram_vec(D, dim) ? dict_gram_kernel<true> : dict_gram_kernel<false>;
    hipLaunchKernelGGL(kern, dim3((unsigned)((kk + 255) / 256)), dim3(256), 0, stream, D, dim, K, G);
    return hipGetLastError();
}
