// Deterministic overlap-add rebuild, lambda and blend on gfx950
// (CudaKSvdDenoiser::createImage, cuda_svd_denoiser.cu:413-498).
//
// rebuild_tiles: one thread per pixel gathers the patches that cover it in
//   ascending patch order (no atomics).  Each covering patch's value at that
//   pixel is D x restricted to the support, summed in ascending atom order
//   with fused multiply-adds.  Writes the sum and the weight (count) and a
//   per-workgroup double partial of (sum/w - y)^2 over covered pixels.  The
//   workgroups are persistent (one per CU) and hold the whole dictionary in
//   LDS; per 16 x 64-pixel tile the supports of the covering patches are
//   sorted once into LDS, so a patch's codes are read once per tile instead
//   of once per pixel it covers.
// rebuild_pixels: the same sums from global memory, one pixel per thread, for
//   dictionaries that do not fit in LDS.
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

constexpr int RB_TA = 16, RB_TB = 64;  // tile: 16 rows x 64 consecutive pixels, one per thread

// Covering-patch slots a tile can need along one axis: the patches whose
// window [p*step, p*step + ext) meets [a0, a0 + t).
__host__ __device__ inline int rb_span(int t, int ext, int step) { return (t - 1 + ext) / step + 1; }

struct RbTile {
    int a0, b0, PR0, PC0, nc, ns;
};

__device__ inline RbTile rb_tile(int t, int ntb, int m, int n, int Pr, int Pc, int rext, int cext, int rstep,
                                 int cstep)
{
    RbTile g;
    g.a0 = (t / ntb) * RB_TA;
    g.b0 = (t - (t / ntb) * ntb) * RB_TB;
    const int a1 = min(g.a0 + RB_TA, n) - 1, b1 = min(g.b0 + RB_TB, m) - 1;
    g.PR0 = g.a0 >= rext ? (g.a0 - rext) / rstep + 1 : 0;
    g.PC0 = g.b0 >= cext ? (g.b0 - cext) / cstep + 1 : 0;
    const int nr = min(a1 / rstep, Pr - 1) - g.PR0 + 1;
    g.nc = min(b1 / cstep, Pc - 1) - g.PC0 + 1;
    g.ns = (nr > 0 && g.nc > 0) ? nr * g.nc : 0;
    return g;
}

// KM: bound on the support size (kmax <= KM), and the stride of a slot's
// support in LDS.  nslot: bound on a tile's
// covering patches, at most RB_TA * RB_TB (one per thread, checked at launch).
// DL: the dictionary in LDS; without it (dictionaries above the LDS, 16 x 12
// and 22 x 22 patches) the dictionary is read from global memory, where it
// stays in L2, and only the tile's supports are in LDS.
template <int KM, bool DL>
__global__ __launch_bounds__(RB_TA * RB_TB) void rebuild_tiles_kernel(
    const int *__restrict__ idx, const float *__restrict__ coef, const int *__restrict__ cnt,
    const float *__restrict__ D, int dim, int K, int kmax, int m, int n, int Pr, int Pc, int rext, int cext,
    int rstep, int cstep, int ntb, int ntiles, int nslot, const float *__restrict__ img, float *__restrict__ sum,
    float *__restrict__ wgt, double *__restrict__ part)
{
    extern __shared__ __attribute__((aligned(16))) float Ds[];  // K x dim, then the tile's supports
    __shared__ double red[RB_TA * RB_TB / 64];
    const int tid = threadIdx.x;
    const int kd = DL ? K * dim : 0;
    float2 *sup = (float2 *)(Ds + ((kd + 3) & ~3));  // [slot][KM]: (atom * dim as bits, coefficient)
    int *sn = (int *)(sup + (size_t)nslot * KM);  // [slot]: support size
    // this thread's covering patch of the next tile: its codes, loaded a tile ahead
    int np = 0, at[KM];
    float cf[KM];
    auto fetch = [&](int t) {
        np = 0;
        if (t >= ntiles) return;
        const RbTile g = rb_tile(t, ntb, m, n, Pr, Pc, rext, cext, rstep, cstep);
        if (tid >= g.ns) return;
        const int64_t p = (int64_t)(g.PR0 + tid / g.nc) * Pc + (g.PC0 + tid % g.nc);
        np = cnt[p];
#pragma unroll
        for (int i = 0; i < KM; i++) {
            if (i < kmax) {
                at[i] = idx[p * kmax + i];
                cf[i] = coef[p * kmax + i];
            }
        }
    };
    fetch(blockIdx.x);
    if (!DL) {
    } else if ((kd & 3) == 0) {
        for (int i = tid; i < kd / 4; i += RB_TA * RB_TB) ((float4 *)Ds)[i] = ((const float4 *)D)[i];
    } else {
        for (int i = tid; i < kd; i += RB_TA * RB_TB) Ds[i] = D[i];
    }
    double acc = 0.0;
    const int ta = tid / RB_TB, tb = tid - ta * RB_TB;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const RbTile g = rb_tile(t, ntb, m, n, Pr, Pc, rext, cext, rstep, cstep);
        const int a0 = g.a0, b0 = g.b0, PR0 = g.PR0, PC0 = g.PC0, nc = g.nc;
        __syncthreads();  // the previous tile's readers (and the dictionary's writers) are done
        if (tid < g.ns) {
            // support sorted by atom index, stably (the order an insertion sort
            // gives); entries np..KM-1 padded with (atom 0, coefficient 0)
            float2 *sj = sup + (size_t)tid * KM;
#pragma unroll
            for (int i = 0; i < KM; i++) {
                int rank = i;
                if (i < np) {
                    rank = 0;
#pragma unroll
                    for (int j = 0; j < KM; j++)
                        rank += (j < np && (at[j] < at[i] || (at[j] == at[i] && j < i))) ? 1 : 0;
                }
                sj[rank] = i < np ? make_float2(__int_as_float(at[i] * dim), cf[i]) : make_float2(0.f, 0.f);
            }
            sn[tid] = np;
        }
        __syncthreads();
        fetch(t + gridDim.x);
        const int a = a0 + ta, b = b0 + tb;
        if (a < n && b < m) {
            const int pr_lo = a >= rext ? (a - rext) / rstep + 1 : 0;
            const int pr_hi = min(a / rstep, Pr - 1);
            const int pc_lo = b >= cext ? (b - cext) / cstep + 1 : 0;
            const int pc_hi = min(b / cstep, Pc - 1);
            float s = 0.f, w = 0.f;
            for (int pr = pr_lo; pr <= pr_hi; pr++) {
                const int r = a - pr * rstep;
                for (int pc = pc_lo; pc <= pc_hi; pc++) {
                    const int e = r * cext + (b - pc * cstep);
                    const int j = (pr - PR0) * nc + (pc - PC0);
                    const float2 *sj = sup + (size_t)j * KM;
                    const int np = sn[j];
                    // every load of the support first, then the products in
                    // ascending atom order (the padding is never summed)
                    float2 u[KM];
                    float dv[KM];
#pragma unroll
                    for (int i = 0; i < KM; i++) u[i] = sj[i];
#pragma unroll
                    for (int i = 0; i < KM; i++) dv[i] = DL ? Ds[__float_as_int(u[i].x) + e] : D[__float_as_int(u[i].x) + e];
                    float v = 0.f;
#pragma unroll
                    for (int i = 0; i < KM; i++)
                        if (i < np) v = fmaf(dv[i], u[i].y, v);
                    s = s + v;
                    w = w + 1.f;
                }
            }
            const int64_t q = (int64_t)a * m + b;
            sum[q] = s;
            wgt[q] = w;
            if (w > 0.f) {
                const float first = s / w;
                const float diff = first - img[q];
                const float sq = diff * diff;
                acc += (double)sq;
            }
        }
    }
    // fixed-order fold: lanes by a butterfly, then the wavefronts in order
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((tid & 63) == 0) red[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) {
        double tsum = 0.0;
        for (int i = 0; i < RB_TA * RB_TB / 64; i++) tsum += red[i];
        part[blockIdx.x] = tsum;
    }
}

// LDS bytes of rebuild_tiles_kernel: the dictionary (when it is held there)
// and nslot supports of stride km.
size_t rebuild_tiles_lds(int K, int dim, int km, int64_t nslot, bool dl)
{
    const size_t slots = (size_t)nslot;
    const size_t kd = dl ? (size_t)K * dim : 0;
    return sizeof(float) * ((kd + 3) & ~(size_t)3) + slots * km * sizeof(float2) + slots * sizeof(int);
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

// SPECKLE: the output leaves the log domain, v -> 2^v (createImage's
// speckle branch, cuda_svd_denoiser.cu:483-484, which the reference runs on
// the host with powf(2, v)); exp2f is within 2 ulp of it.
template <bool SPECKLE>
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
    float v;
    if (w == 0.f) {
        v = img[q];
    } else {
        float num = img[q] * lam;
        num = num + sum[q];
        v = num / (lam + w);
    }
    out[q] = SPECKLE ? exp2f(v) : v;
}

}  // namespace

int launch_rebuild(ksvd_ctx *ctx, float sigma2, bool speckle)
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
    const int nblk = (int)((N + 255) / 256);  // blend grid, and the capacity of the partials
    const int64_t slots = (int64_t)rb_span(RB_TA, rext, rstep) * rb_span(RB_TB, cext, cstep);
    const int km = ctx->kmax <= 4 ? 4 : ctx->kmax <= 5 ? 5 : ctx->kmax <= 8 ? 8 : 16;
    const size_t lds_cap = 160 * 1024 - sizeof(double) * (RB_TA * RB_TB / 64);
    const bool dl = rebuild_tiles_lds(ctx->K, ctx->dim, km, slots, true) <= lds_cap;
    const size_t lds = rebuild_tiles_lds(ctx->K, ctx->dim, km, slots, dl);
    int nparts = nblk;
    if (ctx->kmax <= 16 && slots <= RB_TA * RB_TB && lds <= lds_cap) {
        const int ntb = (ctx->m + RB_TB - 1) / RB_TB;
        const int64_t nt64 = (int64_t)((ctx->n + RB_TA - 1) / RB_TA) * ntb;
        if (nt64 > INT32_MAX) return ksvd_fail(ctx, KSVD_ERR_UNSUPPORTED, "rebuild: too many tiles");
        const int ntiles = (int)nt64;
        // with the dictionary in L2, two workgroups a CU hide its loads
        int grid = dl ? ctx->num_cus : 2 * ctx->num_cus;
        if (grid > ntiles) grid = ntiles;
        if (grid > nblk) grid = nblk;
        nparts = grid;
        typedef void (*RbFn)(const int *, const float *, const int *, const float *, int, int, int, int, int, int,
                             int, int, int, int, int, int, int, int, const float *, float *, float *, double *);
        RbFn kern;
        if (dl)
            kern = km == 4 ? rebuild_tiles_kernel<4, true>
                 : km == 5 ? rebuild_tiles_kernel<5, true>
                 : km == 8 ? rebuild_tiles_kernel<8, true>
                           : rebuild_tiles_kernel<16, true>;
        else
            kern = km == 4 ? rebuild_tiles_kernel<4, false>
                 : km == 5 ? rebuild_tiles_kernel<5, false>
                 : km == 8 ? rebuild_tiles_kernel<8, false>
                           : rebuild_tiles_kernel<16, false>;
        const size_t key = (size_t)km * 2 + (dl ? 1 : 0);
        if (ctx->rebuild_lds_set != key) {  // one host call per context and variant, not per rebuild
            KSVD_HIP_CHECK(ctx, allow_max_lds((const void *)kern));
            ctx->rebuild_lds_set = key;
        }
        hipLaunchKernelGGL(kern, dim3(grid), dim3(RB_TA * RB_TB), lds, ctx->stream,
                           (const int *)ctx->idx.ptr, (const float *)ctx->coef.ptr, (const int *)ctx->cnt.ptr,
                           (const float *)ctx->D.ptr, ctx->dim, ctx->K, ctx->kmax, ctx->m, ctx->n, Pr, Pc, rext,
                           cext, rstep, cstep, ntb, ntiles, (int)slots, (const float *)ctx->img.ptr, (float *)ctx->rsum.ptr,
                           (float *)ctx->rweight.ptr, (double *)ctx->rpart.ptr);
    } else {
        hipLaunchKernelGGL(rebuild_pixels_kernel, dim3(nblk), dim3(256), 0, ctx->stream, (const int *)ctx->idx.ptr,
                           (const float *)ctx->coef.ptr, (const int *)ctx->cnt.ptr, (const float *)ctx->D.ptr,
                           ctx->dim, ctx->kmax, ctx->m, ctx->n, Pr, Pc, rext, cext, rstep, cstep,
                           (const float *)ctx->img.ptr, (float *)ctx->rsum.ptr, (float *)ctx->rweight.ptr,
                           (double *)ctx->rpart.ptr);
    }
    hipLaunchKernelGGL(rebuild_lambda_kernel, dim3(1), dim3(256), 0, ctx->stream, (const double *)ctx->rpart.ptr,
                       nparts, sigma2, N, (float *)ctx->rlam.ptr, (double *)((float *)ctx->rlam.ptr + 2));
    auto blend = speckle ? rebuild_blend_kernel<true> : rebuild_blend_kernel<false>;
    hipLaunchKernelGGL(blend, dim3(nblk), dim3(256), 0, ctx->stream, (const float *)ctx->img.ptr,
                       (const float *)ctx->rsum.ptr, (const float *)ctx->rweight.ptr, (const float *)ctx->rlam.ptr, N,
                       (float *)ctx->rout.ptr);
    KSVD_HIP_CHECK(ctx, hipGetLastError());
    return KSVD_OK;
}
