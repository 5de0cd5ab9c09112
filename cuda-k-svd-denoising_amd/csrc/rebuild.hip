// Deterministic overlap-add rebuild, lambda and blend on gfx950
// (CudaKSvdDenoiser::createImage, cuda_svd_denoiser.cu:413-498).
//
// rebuild_pixels: one thread per pixel gathers the patches that cover it in
//   ascending patch order (no atomics).  Each covering patch's value at that
//   pixel is D x restricted to the support, summed in ascending atom order
//   with fused multiply-adds.  Writes the sum and the weight (count) and a
//   per-workgroup double partial of (sum/w - y)^2 over covered pixels.
// rebuild_lambda: one workgroup folds the partials in a fixed order and forms
//   lambda = |sqrtf(d / (sigma2 * N)) - 1| in float as the reference does.
// rebuild_blend: out = (y * lambda + sum) / (lambda + w); uncovered pixels
//   (w = 0) keep y.
#include "ksvd_internal.hpp"

namespace {

__global__ __launch_bounds__(256) void rebuild_pixels_kernel(
    const int *__restrict__ idx, const float *__restrict__ coef, const int *__restrict__ cnt,
    const float *__restrict__ D, int dim, int kmax, int m, int n, int Pr, int Pc, int rext, int cext, int rstep,
    int cstep, const float *__restrict__ img, float *__restrict__ sum, float *__restrict__ wgt,
    double *__restrict__ part)
{
    __shared__ double red[256];
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double contrib = 0.0;
    if (q < (int64_t)m * n) {
        const int a = (int)(q / m), b = (int)(q - (int64_t)a * m);
        const int pr_lo = a >= rext ? (a - rext) / rstep + 1 : 0;
        const int pr_hi = min(a / rstep, Pr - 1);
        const int pc_lo = b >= cext ? (b - cext) / cstep + 1 : 0;
        const int pc_hi = min(b / cstep, Pc - 1);
        float s = 0.f, w = 0.f;
        for (int pr = pr_lo; pr <= pr_hi; pr++) {
            const int r = a - pr * rstep;
            for (int pc = pc_lo; pc <= pc_hi; pc++) {
                const int c = b - pc * cstep;
                const int64_t p = (int64_t)pr * Pc + pc;
                const int e = r * cext + c;
                // support sorted by atom index (insertion, at most kmax entries)
                int ai[KSVD_MAX_SUPPORT];
                float ac[KSVD_MAX_SUPPORT];
                const int np = cnt[p];
                for (int i = 0; i < np; i++) {
                    const int at = idx[p * kmax + i];
                    const float cf = coef[p * kmax + i];
                    int z = i;
                    while (z > 0 && ai[z - 1] > at) {
                        ai[z] = ai[z - 1];
                        ac[z] = ac[z - 1];
                        z--;
                    }
                    ai[z] = at;
                    ac[z] = cf;
                }
                float v = 0.f;
                for (int i = 0; i < np; i++) v = fmaf(D[(int64_t)ai[i] * dim + e], ac[i], v);
                s = s + v;
                w = w + 1.f;
            }
        }
        sum[q] = s;
        wgt[q] = w;
        if (w > 0.f) {
            const float first = s / w;
            const float diff = first - img[q];
            const float sq = diff * diff;
            contrib = (double)sq;
        }
    }
    red[threadIdx.x] = contrib;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void rebuild_lambda_kernel(const double *__restrict__ part, int nparts,
                                                             float sigma2, int64_t N, float *__restrict__ lam_out,
                                                             double *__restrict__ dsum_out)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += 256) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float d = (float)red[0];
        d = d / (sigma2 * (float)N);
        lam_out[0] = fabsf(sqrtf(d) - 1.f);
        dsum_out[0] = red[0];
    }
}

__global__ __launch_bounds__(256) void rebuild_blend_kernel(const float *__restrict__ img,
                                                            const float *__restrict__ sum,
                                                            const float *__restrict__ wgt,
                                                            const float *__restrict__ lam_p, int64_t N,
                                                            float *__restrict__ out)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N) return;
    const float lam = lam_p[0];
    const float w = wgt[q];
    if (w == 0.f) {
        out[q] = img[q];
        return;
    }
    float num = img[q] * lam;
    num = num + sum[q];
    out[q] = num / (lam + w);
}

}  // namespace

int launch_rebuild(ksvd_ctx *ctx, float sigma2)
{
    const int pw = ctx->prm.patch_w, ph = ctx->prm.patch_h;
    const int sw = ctx->prm.slide_w, sh = ctx->prm.slide_h;
    int rext = ph, cext = pw, rstep = sh, cstep = sw;
    if (ctx->prm.geometry == KSVD_GEOM_COMPAT) {
        rext = pw;
        cext = ph;
        rstep = sw;
        cstep = sh;
    }
    const int Pr = (ctx->n - rext) / rstep + 1, Pc = (ctx->m - cext) / cstep + 1;
    if ((int64_t)Pr * Pc != ctx->P)
        return ksvd_fail(ctx, KSVD_ERR_UNSUPPORTED,
                         "compat geometry: the rebuild walk does not visit as many patches as the gather");
    const int64_t N = (int64_t)ctx->m * ctx->n;
    const int nblk = (int)((N + 255) / 256);
    hipLaunchKernelGGL(rebuild_pixels_kernel, dim3(nblk), dim3(256), 0, ctx->stream, (const int *)ctx->idx.ptr,
                       (const float *)ctx->coef.ptr, (const int *)ctx->cnt.ptr, (const float *)ctx->D.ptr, ctx->dim,
                       ctx->kmax, ctx->m, ctx->n, Pr, Pc, rext, cext, rstep, cstep, (const float *)ctx->img.ptr,
                       (float *)ctx->rsum.ptr, (float *)ctx->rweight.ptr, (double *)ctx->rpart.ptr);
    hipLaunchKernelGGL(rebuild_lambda_kernel, dim3(1), dim3(256), 0, ctx->stream, (const double *)ctx->rpart.ptr,
                       nblk, sigma2, N, (float *)ctx->rlam.ptr, (double *)((float *)ctx->rlam.ptr + 2));
    hipLaunchKernelGGL(rebuild_blend_kernel, dim3(nblk), dim3(256), 0, ctx->stream, (const float *)ctx->img.ptr,
                       (const float *)ctx->rsum.ptr, (const float *)ctx->rweight.ptr, (const float *)ctx->rlam.ptr, N,
                       (float *)ctx->rout.ptr);
    KSVD_HIP_CHECK(ctx, hipGetLastError());
    return KSVD_OK;
}
