// Dictionary update (one Gauss-Seidel sweep of rank-1 atom updates) on gfx950.
//
// Users of every atom (CSR, ascending patch index, deterministic):
//   users_count   one workgroup per 256 patches; an LDS bitmask per atom marks
//                 the block's patches that use it (coef != 0); popcounts give
//                 per-block counts;
//   users_scan    per atom, exclusive scan of the block counts;
//   users_offsets exclusive scan over atoms;
//   users_scatter rank of a patch among the block's users of an atom = popcount
//                 of the mask below it; writes the patch and its support slot.
// Per atom k, in order (cuda_svd_denoiser.cu:309-361):
//   upd_build     E[:,t] = y_p - sum_{j != k} d_j x_{j,p} for the first
//                 min(c_k, cap) users (sparse: only the patch's support);
//   upd_gram      G = E E^T, one 16x16 tile per workgroup, fixed order;
//   upd_eig       top eigenvector of G by power iteration in LDS, started
//                 from the previous atom; written to D[:,k];
//   upd_coef      x_{k,p} = e_t . d_k (= s v_t) for t < cap, 0 beyond.
#include <cstdlib>

#include "ksvd_internal.hpp"

namespace {

constexpr int UB = 256;           // patches per users_* workgroup
constexpr int UW = UB / 64;       // 64-bit mask words per atom per workgroup
constexpr int BCOLS = 16;         // E columns per upd_build workgroup
constexpr int BUILD_BLOCKS = 64;  // covers a 1024-user cap in one pass

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ void build_masks(unsigned long long *msk, const int *__restrict__ idx, const float *__restrict__ coef,
                            const int *__restrict__ cnt, int64_t P, int K, int kmax)
{
    for (int q = threadIdx.x; q < K * UW; q += UB) msk[q] = 0ull;
    __syncthreads();
    const int64_t p = (int64_t)blockIdx.x * UB + threadIdx.x;
    if (p < P) {
        const int c = cnt[p];
        for (int i = 0; i < c; i++)
            if (coef[p * kmax + i] != 0.f)
                atomicOr(&msk[idx[p * kmax + i] * UW + (threadIdx.x >> 6)], 1ull << (threadIdx.x & 63));
    }
    __syncthreads();
}

__global__ __launch_bounds__(UB) void users_count_kernel(const int *__restrict__ idx, const float *__restrict__ coef,
                                                         const int *__restrict__ cnt, int64_t P, int K, int kmax,
                                                         int *__restrict__ blockcnt)
{
    extern __shared__ unsigned long long msk[];
    build_masks(msk, idx, coef, cnt, P, K, kmax);
    for (int a = threadIdx.x; a < K; a += UB) {
        int c = 0;
#pragma unroll
        for (int w = 0; w < UW; w++) c += __popcll(msk[a * UW + w]);
        blockcnt[(int64_t)blockIdx.x * K + a] = c;
    }
}

__global__ __launch_bounds__(256) void users_scan_kernel(int *__restrict__ blockcnt, int nb, int K,
                                                         int *__restrict__ ucount)
{
    __shared__ int sh[256];
    const int a = blockIdx.x;
    int carry = 0;
    for (int b0 = 0; b0 < nb; b0 += 256) {
        const int b = b0 + threadIdx.x;
        const int v = b < nb ? blockcnt[(int64_t)b * K + a] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (b < nb) blockcnt[(int64_t)b * K + a] = carry + sh[threadIdx.x] - v;
        carry += sh[255];
        __syncthreads();
    }
    if (threadIdx.x == 0) ucount[a] = carry;
}

__global__ __launch_bounds__(1024) void users_offsets_kernel(const int *__restrict__ ucount, int K,
                                                             int *__restrict__ uoff)
{
    __shared__ int sh[1024];
    const int a = threadIdx.x;
    const int v = a < K ? ucount[a] : 0;
    sh[a] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int t = a >= off ? sh[a - off] : 0;
        __syncthreads();
        sh[a] += t;
        __syncthreads();
    }
    if (a < K) uoff[a] = sh[a] - v;
    if (a == 0) uoff[K] = sh[1023];
}

__global__ __launch_bounds__(UB) void users_scatter_kernel(const int *__restrict__ idx, const float *__restrict__ coef,
                                                           const int *__restrict__ cnt, int64_t P, int K, int kmax,
                                                           const int *__restrict__ blockoff,
                                                           const int *__restrict__ uoff, int *__restrict__ users,
                                                           int *__restrict__ uslot)
{
    extern __shared__ unsigned long long msk[];
    build_masks(msk, idx, coef, cnt, P, K, kmax);
    const int64_t p = (int64_t)blockIdx.x * UB + threadIdx.x;
    if (p >= P) return;
    const int w = threadIdx.x >> 6, bit = threadIdx.x & 63;
    const int c = cnt[p];
    for (int i = 0; i < c; i++) {
        if (coef[p * kmax + i] == 0.f) continue;
        const int a = idx[p * kmax + i];
        int rank = __popcll(msk[a * UW + w] & ((1ull << bit) - 1ull));
        for (int ww = 0; ww < w; ww++) rank += __popcll(msk[a * UW + ww]);
        const int pos = uoff[a] + blockoff[(int64_t)blockIdx.x * K + a] + rank;
        users[pos] = (int)p;
        uslot[pos] = i;
    }
}

__device__ __forceinline__ int atom_cols(const int *uoff, int k, int cap, int *c_out)
{
    const int c = uoff[k + 1] - uoff[k];
    *c_out = c;
    return (cap > 0 && c > cap) ? cap : c;
}

__global__ __launch_bounds__(256) void upd_build_kernel(int k, const float *__restrict__ Y, int dim,
                                                        const float *__restrict__ D, int kmax,
                                                        const int *__restrict__ idx, const float *__restrict__ coef,
                                                        const int *__restrict__ cnt, const int *__restrict__ uoff,
                                                        const int *__restrict__ users, const int *__restrict__ uslot,
                                                        int cap, float *__restrict__ E)
{
    __shared__ int sp[BCOLS], sn[BCOLS];
    __shared__ int sa[BCOLS][KSVD_MAX_SUPPORT];
    __shared__ float sc[BCOLS][KSVD_MAX_SUPPORT];
    int c;
    const int mcols = atom_cols(uoff, k, cap, &c);
    const int base = uoff[k];
    for (int t0 = blockIdx.x * BCOLS; t0 < mcols; t0 += gridDim.x * BCOLS) {
        const int nc = min(BCOLS, mcols - t0);
        if (threadIdx.x < nc) {
            const int tc = threadIdx.x;
            const int pos = base + t0 + tc;
            const int p = users[pos], my = uslot[pos];
            const int cp = cnt[p];
            int na = 0;
            for (int i = 0; i < cp; i++)
                if (i != my) {
                    sa[tc][na] = idx[(int64_t)p * kmax + i];
                    sc[tc][na] = coef[(int64_t)p * kmax + i];
                    na++;
                }
            sp[tc] = p;
            sn[tc] = na;
        }
        __syncthreads();
        for (int q = threadIdx.x; q < nc * dim; q += blockDim.x) {
            const int tc = q / dim, e = q - tc * dim;
            float v = Y[(int64_t)sp[tc] * dim + e];
            const int na = sn[tc];
            for (int i = 0; i < na; i++) v = fmaf(-D[(int64_t)sa[tc][i] * dim + e], sc[tc][i], v);
            E[(int64_t)(t0 + tc) * dim + e] = v;
        }
        __syncthreads();
    }
}

// G = E E^T, tile (ba, bb) of 16 x 16 per workgroup; t ascending, FMA chain.
__global__ __launch_bounds__(256) void upd_gram_kernel(int k, int dim, const int *__restrict__ uoff, int cap,
                                                       const float *__restrict__ E, float *__restrict__ G)
{
    __shared__ float ea[64][17], eb[64][17];
    int c;
    const int mcols = atom_cols(uoff, k, cap, &c);
    const int nbt = (dim + 15) / 16;
    const int ba = blockIdx.x / nbt, bb = blockIdx.x - ba * nbt;
    const int ta = threadIdx.x >> 4, tb = threadIdx.x & 15;
    const int a = ba * 16 + ta, b = bb * 16 + tb;
    float s = 0.f;
    for (int t0 = 0; t0 < mcols; t0 += 64) {
        for (int q = threadIdx.x; q < 64 * 16; q += 256) {
            const int tt = q >> 4, ee = q & 15;
            const int t = t0 + tt;
            const int ia = ba * 16 + ee, ib = bb * 16 + ee;
            ea[tt][ee] = (t < mcols && ia < dim) ? E[(int64_t)t * dim + ia] : 0.f;
            eb[tt][ee] = (t < mcols && ib < dim) ? E[(int64_t)t * dim + ib] : 0.f;
        }
        __syncthreads();
        const int nt = min(64, mcols - t0);
        for (int tt = 0; tt < nt; tt++) s = fmaf(ea[tt][ta], eb[tt][tb], s);
        __syncthreads();
    }
    if (a < dim && b < dim) G[a * dim + b] = s;
}

// Block-wide sum over 1024 threads; every thread returns the same value.
__device__ float block_sum_1024(float v, float *red)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    float s = 0.f;
#pragma unroll
    for (int w = 0; w < 16; w++) s += red[w];
    __syncthreads();
    return s;
}

// Top eigenvector of G (dim x dim in LDS) by power iteration from d_k.
// Stops when two consecutive normalised iterates differ by less than
// sqrt(tol2) (never before the second), or after maxit iterations.
__global__ __launch_bounds__(1024) void upd_eig_kernel(int k, int dim, float *__restrict__ D,
                                                       const int *__restrict__ uoff, const float *__restrict__ G,
                                                       int maxit, float tol2,
                                                       unsigned long long *__restrict__ counters)
{
    extern __shared__ float sm[];
    float *Gs = sm;                   // dim * dim
    float *u = Gs + dim * dim;        // dim
    float *part = u + dim;            // 1024
    __shared__ float red[16];
    if (uoff[k + 1] - uoff[k] <= 0) return;
    const int tid = threadIdx.x;
    for (int q = tid; q < dim * dim; q += 1024) Gs[q] = G[q];
    for (int e = tid; e < dim; e += 1024) u[e] = D[(int64_t)k * dim + e];
    __syncthreads();
    const int S = 1024 / dim;         // threads per row
    const int row = tid / S, seg = tid - row * S;
    int it = 0;
    for (;;) {
        // z = G u: row `row` split over S threads, segments combined in order
        float ps = 0.f;
        if (row < dim)
            for (int q = seg; q < dim; q += S) ps = fmaf(Gs[row * dim + q], u[q], ps);
        part[tid] = ps;
        __syncthreads();
        float z = 0.f;
        if (tid < dim)
            for (int q = 0; q < S; q++) z += part[tid * S + q];
        const float nz2 = block_sum_1024(tid < dim ? z * z : 0.f, red);
        const float nz = sqrtf(nz2);
        const float un = (tid < dim && nz > 0.f) ? z / nz : 0.f;
        const float du = tid < dim ? un - u[tid] : 0.f;
        const float d2 = block_sum_1024(du * du, red);
        it++;
        if (nz == 0.f) break;
        if (tid < dim) u[tid] = un;
        __syncthreads();
        if ((it >= 2 && d2 < tol2) || it >= maxit) break;
    }
    if (tid < dim) D[(int64_t)k * dim + tid] = u[tid];
    if (tid == 0) atomicAdd(&counters[CNT_EIG_IT], (unsigned long long)it);
}

// x_{k,p} = e_t . d_k for the first min(c_k, cap) users, 0 for the rest.
__global__ __launch_bounds__(256) void upd_coef_kernel(int k, int dim, const float *__restrict__ D, int kmax,
                                                       float *__restrict__ coef, const int *__restrict__ uoff,
                                                       const int *__restrict__ users, const int *__restrict__ uslot,
                                                       int cap, const float *__restrict__ E)
{
    __shared__ float u[KSVD_MAX_DIM];
    int c;
    const int mcols = atom_cols(uoff, k, cap, &c);
    if (c <= 0) return;
    for (int e = threadIdx.x; e < dim; e += blockDim.x) u[e] = D[(int64_t)k * dim + e];
    __syncthreads();
    const int base = uoff[k];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = gridDim.x * (blockDim.x >> 6);
    for (int t = blockIdx.x * (blockDim.x >> 6) + wave; t < mcols; t += nw) {
        const float *et = E + (int64_t)t * dim;
        float s = 0.f;
        for (int e = lane; e < dim; e += 64) s = fmaf(et[e], u[e], s);
        s = wave_sum(s);
        if (lane == 0) {
            const int pos = base + t;
            coef[(int64_t)users[pos] * kmax + uslot[pos]] = s;
        }
    }
    const int nthreads = gridDim.x * blockDim.x;
    for (int t = mcols + blockIdx.x * blockDim.x + threadIdx.x; t < c; t += nthreads) {
        const int pos = base + t;
        coef[(int64_t)users[pos] * kmax + uslot[pos]] = 0.f;
    }
}

}  // namespace

// Enqueue one sweep: user lists, then per atom build -> Gram -> eigenvector ->
// coefficients.  Everything is stream-ordered; no host synchronisation.
static int enqueue_update(ksvd_ctx *ctx)
{
    const int K = ctx->K, kmax = ctx->kmax, dim = ctx->dim;
    const int64_t P = ctx->P;
    const int nb = (int)((P + UB - 1) / UB);
    const int cap = ctx->prm.user_cap;
    const size_t mlds = sizeof(unsigned long long) * (size_t)K * UW;
    int *blockcnt = (int *)ctx->blockcnt.ptr;
    hipLaunchKernelGGL(users_count_kernel, dim3(nb), dim3(UB), mlds, ctx->stream, (const int *)ctx->idx.ptr,
                       (const float *)ctx->coef.ptr, (const int *)ctx->cnt.ptr, P, K, kmax, blockcnt);
    hipLaunchKernelGGL(users_scan_kernel, dim3(K), dim3(256), 0, ctx->stream, blockcnt, nb, K, (int *)ctx->ucount.ptr);
    hipLaunchKernelGGL(users_offsets_kernel, dim3(1), dim3(1024), 0, ctx->stream, (const int *)ctx->ucount.ptr, K,
                       (int *)ctx->uoff.ptr);
    hipLaunchKernelGGL(users_scatter_kernel, dim3(nb), dim3(UB), mlds, ctx->stream, (const int *)ctx->idx.ptr,
                       (const float *)ctx->coef.ptr, (const int *)ctx->cnt.ptr, P, K, kmax,
                       (const int *)blockcnt, (const int *)ctx->uoff.ptr, (int *)ctx->users.ptr,
                       (int *)ctx->uslot.ptr);
    KSVD_HIP_CHECK(ctx, hipGetLastError());

    const size_t eig_lds = sizeof(float) * ((size_t)dim * dim + dim + 1024);
    const int nbt = (dim + 15) / 16;
    float *E = (float *)ctx->E.ptr;
    float *Gk = (float *)ctx->wbuf.ptr;
    const float tol = 1e-6f;
    for (int k = 0; k < K; k++) {
        hipLaunchKernelGGL(upd_build_kernel, dim3(BUILD_BLOCKS), dim3(256), 0, ctx->stream, k,
                           (const float *)ctx->Y.ptr, dim, (const float *)ctx->D.ptr, kmax, (const int *)ctx->idx.ptr,
                           (const float *)ctx->coef.ptr, (const int *)ctx->cnt.ptr, (const int *)ctx->uoff.ptr,
                           (const int *)ctx->users.ptr, (const int *)ctx->uslot.ptr, cap, E);
        hipLaunchKernelGGL(upd_gram_kernel, dim3(nbt * nbt), dim3(256), 0, ctx->stream, k, dim,
                           (const int *)ctx->uoff.ptr, cap, (const float *)E, Gk);
        hipLaunchKernelGGL(upd_eig_kernel, dim3(1), dim3(1024), eig_lds, ctx->stream, k, dim, (float *)ctx->D.ptr,
                           (const int *)ctx->uoff.ptr, (const float *)Gk, 300, tol * tol,
                           (unsigned long long *)ctx->counters.ptr);
        hipLaunchKernelGGL(upd_coef_kernel, dim3(BUILD_BLOCKS), dim3(256), 0, ctx->stream, k, dim,
                           (const float *)ctx->D.ptr, kmax, (float *)ctx->coef.ptr, (const int *)ctx->uoff.ptr,
                           (const int *)ctx->users.ptr, (const int *)ctx->uslot.ptr, cap, (const float *)E);
    }
    KSVD_HIP_CHECK(ctx, hipGetLastError());
    return KSVD_OK;
}

void drop_update_graph(ksvd_ctx *ctx)
{
    if (ctx->upd_exec) hipGraphExecDestroy((hipGraphExec_t)ctx->upd_exec);
    if (ctx->upd_graph) hipGraphDestroy((hipGraph_t)ctx->upd_graph);
    ctx->upd_exec = nullptr;
    ctx->upd_graph = nullptr;
}

// The sweep's 4K + 4 launches are captured once per buffer layout into a
// HIP graph and replayed (set KSVD_NO_GRAPH=1 to launch them one by one).
int launch_update(ksvd_ctx *ctx)
{
    const size_t eig_lds = sizeof(float) * ((size_t)ctx->dim * ctx->dim + ctx->dim + 1024);
    KSVD_HIP_CHECK(ctx, hipFuncSetAttribute((const void *)upd_eig_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                            (int)eig_lds));
    const char *ng = getenv("KSVD_NO_GRAPH");
    const bool no_graph = ng && ng[0] && ng[0] != '0';
    if (no_graph) {
        int rc = enqueue_update(ctx);
        if (rc) return rc;
    } else {
        if (!ctx->upd_exec) {
            hipGraph_t g = nullptr;
            hipGraphExec_t ge = nullptr;
            KSVD_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            int rc = enqueue_update(ctx);
            hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (ec != hipSuccess) return ksvd_fail(ctx, KSVD_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
            KSVD_HIP_CHECK(ctx, hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            ctx->upd_graph = g;
            ctx->upd_exec = ge;
        }
        KSVD_HIP_CHECK(ctx, hipGraphLaunch((hipGraphExec_t)ctx->upd_exec, ctx->stream));
    }
    ctx->gram_valid = false;
    return KSVD_OK;
}
