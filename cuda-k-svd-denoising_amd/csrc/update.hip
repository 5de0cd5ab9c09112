// Dictionary update (one Gauss-Seidel sweep of rank-1 atom updates) on gfx950:
// the user lists and the column table the sweep kernels (sweep.hip) read.
//
// Users of every atom (CSR, ascending patch index, deterministic):
//   users_count   one workgroup per 256 patches; an LDS bitmask per atom marks
//                 the block's patches that use it (coef != 0); popcounts give
//                 per-block counts;
//   users_scan    per atom, exclusive scan of the block counts;
//   users_offsets exclusive scans over atoms: the users lists' offsets and
//                 the capped columns' offsets in the column table;
//   users_scatter rank of a patch among the block's users of an atom = popcount
//                 of the mask below it; writes the patch, its support slot and,
//                 per support entry, the patch's rank in that atom's list.
// The users of atom k are the patches whose code has a nonzero coefficient on
// k when the sweep starts (findRelevantIndicesKernel,
// cuda_svd_denoiser.cu:36-50, run once per sweep at :298).
#include <cstdlib>
#include <mutex>

#include "ksvd_internal.hpp"

namespace {

constexpr int UB = 256;           // patches per users_* workgroup
constexpr int UW = UB / 64;       // 64-bit mask words per atom per workgroup

// Patch block of this workgroup.  Workgroups go to the 8 XCDs round robin;
// XCD x takes a contiguous range of blocks, so the runs that neighbouring
// blocks write into an atom's list share that XCD's L2 and leave it as whole
// lines (a bijection on [0, nb)).
__device__ inline int users_block(int nb)
{
    const int b = blockIdx.x, x = b & 7, q = nb >> 3, r = nb & 7;
    return x * q + min(x, r) + (b >> 3);
}

__device__ void build_masks(unsigned long long *msk, const int *__restrict__ idx, const float *__restrict__ coef,
                            const int *__restrict__ cnt, int64_t P, int K, int kmax, int blk)
{
    for (int q = threadIdx.x; q < K * UW; q += UB) msk[q] = 0ull;
    __syncthreads();
    const int64_t p = (int64_t)blk * UB + threadIdx.x;
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
    const int blk = users_block(gridDim.x);
    build_masks(msk, idx, coef, cnt, P, K, kmax, blk);
    for (int a = threadIdx.x; a < K; a += UB) {
        int c = 0;
#pragma unroll
        for (int w = 0; w < UW; w++) c += __popcll(msk[a * UW + w]);
        blockcnt[(int64_t)blk * K + a] = c;
    }
}

// Per atom (one workgroup each), exclusive scan of its block counts in
// place; ucount[a] = the total.  Thread t takes UQ consecutive blocks per
// chunk of 256 UQ (all loads in flight at once), then the 256 thread sums are
// scanned across the wavefront by shuffles and across the 4 wavefronts in LDS.
constexpr int UQ = 8;
__global__ __launch_bounds__(256) void users_scan_kernel(int *__restrict__ blockcnt, int nb, int K,
                                                         int *__restrict__ ucount)
{
    __shared__ int wsum[4];
    const int a = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int carry = 0;
    for (int b0 = 0; b0 < nb; b0 += 256 * UQ) {
        const int bt = b0 + tid * UQ;
        int v[UQ];
#pragma unroll
        for (int q = 0; q < UQ; q++) v[q] = bt + q < nb ? blockcnt[(int64_t)(bt + q) * K + a] : 0;
        int tsum = 0;
#pragma unroll
        for (int q = 0; q < UQ; q++) tsum += v[q];
        // inclusive scan of the thread sums across the wavefront
        int incl = tsum;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int o = __shfl_up(incl, off, 64);
            if (lane >= off) incl += o;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        int before = carry;
        for (int w = 0; w < wave; w++) before += wsum[w];
        const int chunk = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        int run = before + incl - tsum;
#pragma unroll
        for (int q = 0; q < UQ; q++) {
            if (bt + q < nb) blockcnt[(int64_t)(bt + q) * K + a] = run;
            run += v[q];
        }
        carry += chunk;
        __syncthreads();  // wsum is rewritten by the next chunk
    }
    if (tid == 0) ucount[a] = carry;
}

// Exclusive scan over the 1024 threads of a workgroup (K <= 1024 atoms):
// shuffles within each wavefront, the 16 wavefront sums in LDS.  Returns the
// exclusive prefix of v; *total gets the sum.
__device__ int block_scan_1024(int v, int *wsum, int *total)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(incl, off, 64);
        if (lane >= off) incl += o;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int before = 0, all = 0;
    for (int w = 0; w < 16; w++) {
        before += w < wave ? wsum[w] : 0;
        all += wsum[w];
    }
    __syncthreads();  // wsum is reused by the next scan
    *total = all;
    return before + incl - v;
}

// Offsets of every atom's users list (uoff) and of its capped columns in the
// column table (moff, counts capped at cap when cap > 0), in one launch.
__global__ __launch_bounds__(1024) void users_offsets_kernel(const int *__restrict__ ucount, int K, int cap,
                                                             int *__restrict__ uoff, int *__restrict__ moff)
{
    __shared__ int wsum[16];
    const int a = threadIdx.x;
    const int v = a < K ? ucount[a] : 0;
    int tot;
    const int ex = block_scan_1024(v, wsum, &tot);
    if (a < K) uoff[a] = ex;
    if (a == 0) uoff[K] = tot;
    const int vc = (cap > 0 && v > cap) ? cap : v;
    const int exc = block_scan_1024(vc, wsum, &tot);
    if (a < K) moff[a] = exc;
    if (a == 0) moff[K] = tot;
}

__global__ __launch_bounds__(UB) void users_scatter_kernel(const int *__restrict__ idx, const float *__restrict__ coef,
                                                           const int *__restrict__ cnt, int64_t P, int K, int kmax,
                                                           const int *__restrict__ blockoff,
                                                           const int *__restrict__ uoff, int *__restrict__ users,
                                                           int *__restrict__ uslot, int *__restrict__ srank)
{
    extern __shared__ unsigned long long msk[];
    const int blk = users_block(gridDim.x);
    build_masks(msk, idx, coef, cnt, P, K, kmax, blk);
    const int64_t p = (int64_t)blk * UB + threadIdx.x;
    if (p >= P) return;
    const int w = threadIdx.x >> 6, bit = threadIdx.x & 63;
    const int c = cnt[p];
    for (int i = 0; i < kmax; i++) srank[p * kmax + i] = -1;
    for (int i = 0; i < c; i++) {
        if (coef[p * kmax + i] == 0.f) continue;
        const int a = idx[p * kmax + i];
        int rank = __popcll(msk[a * UW + w] & ((1ull << bit) - 1ull));
        for (int ww = 0; ww < w; ww++) rank += __popcll(msk[a * UW + ww]);
        const int pos = uoff[a] + blockoff[(int64_t)blk * K + a] + rank;
        users[pos] = (int)p;
        uslot[pos] = i;
        srank[p * kmax + i] = pos - uoff[a];  // rank among atom a's users
    }
}

// ---------------------------------------------------------------------------
// Column table of one sweep, built in parallel before it: for atom k and its
// t-th capped user (t < m_k = min(c_k, cap)), the patch, the coefficient slot
// of atom k, and the (atom, coefficient slot) pairs of the patch's other
// support entries.  Only coefficient VALUES change during the sweep, so the
// table holds for all of it and the sweep's loads stop depending on each
// other.  moff[k] = sum_{j<k} m_j.

__global__ __launch_bounds__(256) void cols_table_kernel(int K, int kmax, int cap, const int *__restrict__ idx,
                                                         const int *__restrict__ cnt, const int *__restrict__ uoff,
                                                         const int *__restrict__ moff, const int *__restrict__ users,
                                                         const int *__restrict__ uslot, const int *__restrict__ srank,
                                                         int *__restrict__ tab_p, int *__restrict__ tab_self,
                                                         int4 *__restrict__ tab_oth)
{
    const int k = blockIdx.y;
    const int c = uoff[k + 1] - uoff[k];
    const int m = (cap > 0 && c > cap) ? cap : c;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < m; t += gridDim.x * blockDim.x) {
        const int pos = uoff[k] + t;
        const int p = users[pos], my = uslot[pos];
        const int64_t col = moff[k] + t;
        tab_p[col] = p;
        tab_self[col] = p * kmax + my;
        const int cp = cnt[p];
        const int no = kmax > 1 ? kmax - 1 : 1;  // entries per column (one padding entry when kmax = 1)
        int na = 0;
        for (int i = 0; i < cp; i++)
            if (i != my)
            {
                // .w = 1: atom k-1's coefficient, which the sweep hands over
                // instead of reading it
                const int at = idx[(int64_t)p * kmax + i], rk = srank[(int64_t)p * kmax + i];
                tab_oth[col * no + na++] = make_int4(at, p * kmax + i, rk, (at == k - 1 && rk >= 0) ? 1 : 0);
            }
        for (; na < no; na++) tab_oth[col * no + na] = make_int4(-1, 0, -1, 0);
    }
}

}  // namespace

int launch_sweep(ksvd_ctx *ctx, float tol);

// The power iteration's tolerance on |u_i - u_{i-1}| (power_stop also stops
// on its a-posteriori error estimate); a much smaller value would sit below
// the rounding floor of the FP32 products and never trigger.
constexpr float SWEEP_TOL = 1e-5f;

// The user lists and the column table of a sweep, on the context's stream.
static int enqueue_users(ksvd_ctx *ctx)
{
    const int K = ctx->K, kmax = ctx->kmax;
    const int64_t P = ctx->P;
    const int nb = (int)((P + UB - 1) / UB);
    const int cap = ctx->prm.user_cap;
    const size_t mlds = sizeof(unsigned long long) * (size_t)K * UW;
    int *blockcnt = (int *)ctx->blockcnt.ptr;
    hipLaunchKernelGGL(users_count_kernel, dim3(nb), dim3(UB), mlds, ctx->stream, (const int *)ctx->idx.ptr,
                       (const float *)ctx->coef.ptr, (const int *)ctx->cnt.ptr, P, K, kmax, blockcnt);
    hipLaunchKernelGGL(users_scan_kernel, dim3(K), dim3(256), 0, ctx->stream, blockcnt, nb, K, (int *)ctx->ucount.ptr);
    int *moff = (int *)ctx->moff.ptr;
    hipLaunchKernelGGL(users_offsets_kernel, dim3(1), dim3(1024), 0, ctx->stream, (const int *)ctx->ucount.ptr, K,
                       cap, (int *)ctx->uoff.ptr, moff);
    hipLaunchKernelGGL(users_scatter_kernel, dim3(nb), dim3(UB), mlds, ctx->stream, (const int *)ctx->idx.ptr,
                       (const float *)ctx->coef.ptr, (const int *)ctx->cnt.ptr, P, K, kmax,
                       (const int *)blockcnt, (const int *)ctx->uoff.ptr, (int *)ctx->users.ptr,
                       (int *)ctx->uslot.ptr, (int *)ctx->srank.ptr);
    hipLaunchKernelGGL(cols_table_kernel, dim3(4, K), dim3(256), 0, ctx->stream, K, kmax, cap,
                       (const int *)ctx->idx.ptr, (const int *)ctx->cnt.ptr, (const int *)ctx->uoff.ptr,
                       (const int *)moff, (const int *)ctx->users.ptr, (const int *)ctx->uslot.ptr,
                       (const int *)ctx->srank.ptr, (int *)ctx->tab_p.ptr, (int *)ctx->tab_self.ptr,
                       (int4 *)ctx->tab_oth.ptr);
    KSVD_HIP_CHECK(ctx, hipGetLastError());
    return KSVD_OK;
}

// Enqueue one sweep: user lists, column table, then the sweep's kernels.
// Everything is stream-ordered; no host synchronisation.
static int enqueue_update(ksvd_ctx *ctx)
{
    int rc = enqueue_users(ctx);
    if (rc) return rc;
    return launch_sweep(ctx, SWEEP_TOL);
}

void drop_update_graph(ksvd_ctx *ctx)
{
    if (ctx->upd_exec) (void)hipGraphExecDestroy((hipGraphExec_t)ctx->upd_exec);
    if (ctx->upd_graph) (void)hipGraphDestroy((hipGraph_t)ctx->upd_graph);
    ctx->upd_exec = nullptr;
    ctx->upd_graph = nullptr;
}

static bool graphs_off()
{
    const char *ng = getenv("KSVD_NO_GRAPH");
    return ng && ng[0] && ng[0] != '0';
}

// Persistent sweeps of different contexts on one device must not run at the
// same time: each needs all the CUs of one XCD, and two sharing them would
// each wait for workgroups the other holds until their waits give up.  So
// every update on a device waits (on the GPU, not the host) for the last
// update enqueued on that device by any context; the wait, the launch and
// the record are ordered under one lock.  The lock also covers every graph
// capture below: the last update's event may have been recorded on a stream
// that another thread is capturing, and HIP refuses a wait on such an event
// from a stream that is not capturing.  A context that goes away takes the
// device's event with it if it recorded it last, so that no wait names an
// event of a destroyed stream.
namespace {
constexpr int UPD_MAX_DEVICES = 64;
std::mutex upd_mu;
hipEvent_t upd_last[UPD_MAX_DEVICES];
const ksvd_ctx *upd_owner[UPD_MAX_DEVICES];
}  // namespace

// HIP frees (and may allocate) device memory only after synchronising every
// stream of the device, which breaks a capture in progress on another
// thread's stream; so allocations, frees and captures take turns under the
// same lock.
std::mutex &device_op_mutex() { return upd_mu; }

void update_release(ksvd_ctx *ctx)
{
    if (ctx->device < 0 || ctx->device >= UPD_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lock(upd_mu);
    const int d = ctx->device;
    if (upd_owner[d] == ctx) {
        // the context's stream is synchronised and about to go: nothing
        // enqueued later waits on this event any more
        (void)hipEventDestroy(upd_last[d]);
        upd_last[d] = nullptr;
        upd_owner[d] = nullptr;
    }
}

// The update's launches depend only on the context's buffers and sizes, so
// they are captured once per buffer layout into a HIP graph (KSVD_NO_GRAPH=1
// launches them one by one).  ksvd_set_image prepares the graph right after
// laying the buffers out, so that no K-SVD iteration pays the capture.
// Called with upd_mu held.
static int prepare_update_graph_locked(ksvd_ctx *ctx)
{
    if (graphs_off() || ctx->upd_exec) return KSVD_OK;
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
    return KSVD_OK;
}

int prepare_update_graph(ksvd_ctx *ctx)
{
    std::lock_guard<std::mutex> lock(upd_mu);
    return prepare_update_graph_locked(ctx);
}

int launch_update(ksvd_ctx *ctx)
{
    // the users lists read the codes: the last sweep's clearing first
    int rc = flush_cap_zero(ctx);
    if (rc) return rc;
    if (ctx->device < 0 || ctx->device >= UPD_MAX_DEVICES)
        return ksvd_fail(ctx, KSVD_ERR_UNSUPPORTED, "device ordinal beyond 63");
    std::lock_guard<std::mutex> lock(upd_mu);
    hipEvent_t &last = upd_last[ctx->device];
    if (!last) KSVD_HIP_CHECK(ctx, hipEventCreateWithFlags(&last, hipEventDisableTiming));
    else KSVD_HIP_CHECK(ctx, hipStreamWaitEvent(ctx->stream, last, 0));
    if (graphs_off() || ctx->force_two_kernel) {
        rc = enqueue_update(ctx);
        if (rc) return rc;
    } else {
        rc = prepare_update_graph_locked(ctx);
        if (rc) return rc;
        KSVD_HIP_CHECK(ctx, hipGraphLaunch((hipGraphExec_t)ctx->upd_exec, ctx->stream));
    }
    KSVD_HIP_CHECK(ctx, hipEventRecord(last, ctx->stream));
    upd_owner[ctx->device] = ctx;
    ctx->cap_zero_pending = ctx->sweep_defers_zero;
    ctx->gram_valid = false;
    return KSVD_OK;
}

int sweep_multi_tol_setup(ksvd_ctx *const *ctxs, int n) { return sweep_multi_setup(ctxs, n, SWEEP_TOL); }

// One update of several pictures: each picture's user lists and column table
// on its own stream, then one multi-picture sweep on ctxs[0]'s stream once
// the other pictures' streams and the device's last update are done; the
// other streams then wait for it.  Ordered with single updates under the same
// lock.
int launch_update_batch(ksvd_ctx *const *ctxs, int n)
{
    int rc;
    for (int i = 0; i < n; i++) {
        if ((rc = flush_cap_zero(ctxs[i]))) return rc;
        if ((rc = enqueue_users(ctxs[i]))) return rc;
    }
    if ((rc = sweep_multi_reset(ctxs, n))) return rc;
    ksvd_ctx *c0 = ctxs[0];
    if (c0->device < 0 || c0->device >= UPD_MAX_DEVICES)
        return ksvd_fail(c0, KSVD_ERR_UNSUPPORTED, "device ordinal beyond 63");
    std::lock_guard<std::mutex> lock(upd_mu);
    for (int i = 1; i < n; i++) {
        ksvd_ctx *c = ctxs[i];
        if (!c->batch_ev) KSVD_HIP_CHECK(c, hipEventCreateWithFlags((hipEvent_t *)&c->batch_ev, hipEventDisableTiming));
        KSVD_HIP_CHECK(c, hipEventRecord((hipEvent_t)c->batch_ev, c->stream));
        KSVD_HIP_CHECK(c0, hipStreamWaitEvent(c0->stream, (hipEvent_t)c->batch_ev, 0));
    }
    hipEvent_t &last = upd_last[c0->device];
    if (!last) KSVD_HIP_CHECK(c0, hipEventCreateWithFlags(&last, hipEventDisableTiming));
    else KSVD_HIP_CHECK(c0, hipStreamWaitEvent(c0->stream, last, 0));
    if ((rc = launch_sweep_multi(ctxs, n))) return rc;
    KSVD_HIP_CHECK(c0, hipEventRecord(last, c0->stream));
    upd_owner[c0->device] = c0;
    for (int i = 1; i < n; i++) KSVD_HIP_CHECK(ctxs[i], hipStreamWaitEvent(ctxs[i]->stream, last, 0));
    for (int i = 0; i < n; i++) {
        ctxs[i]->cap_zero_pending = ctxs[i]->sweep_defers_zero;
        ctxs[i]->gram_valid = false;
    }
    return KSVD_OK;
}
