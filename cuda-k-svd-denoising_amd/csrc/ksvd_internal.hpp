// Internal declarations of the gfx950 K-SVD library: the context behind the
// opaque ksvd_ctx handle and the launchers of every kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "ksvd_hip.h"

// Learned support proposer: hidden width, padded activation pitch (the
// constant-1 column at MLP_H), training batch, split-K partials of the weight
// gradients, proposal chunk (rows whose activations stay in the last-level
// cache).
constexpr int MLP_H = 512;
constexpr int MLP_HP = 544;
constexpr int MLP_BATCH = 16384;
constexpr int MLP_SPLIT = 16;
constexpr int MLP_CHUNK = 65536;

// Largest supported ompIter of the device path.
constexpr int KSVD_MAX_SUPPORT = 16;
// Largest patch dimension (pw * ph) of the device path.
constexpr int KSVD_MAX_DIM = 1024;

// Device counters, one slot each.
// Slot 3 counts sweep hand-over lists that overflowed (sweep.hip); slots 4-6
// belong to the BF16 pursuit: patch-steps re-scored in FP32, FP64
// re-factorisations, patch-steps whose candidate list was cut at 8.
// Slots 7-8: patches offered to the learned proposer, and those it left to OMP.
enum { CNT_DUP = 0, CNT_DEGEN = 1, CNT_EIG_IT = 2, CNT_HAND_OVF = 3, CNT_RESCORE = 4, CNT_F64 = 5,
       CNT_CAND_CUT = 6, CNT_PROPOSED = 7, CNT_FALLBACK = 8, CNT_SLOTS = 16 };

struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
};

struct ksvd_ctx {
    ksvd_params prm{};
    int device = 0;
    int num_cus = 256;
    int n_xcd = 8;  // XCDs of the device
    hipStream_t stream = nullptr;
    hipEvent_t ev_a = nullptr, ev_b = nullptr, ev_c = nullptr;

    // frame and patch grid: the buffer is n rows of length m
    int m = 0, n = 0;
    int dim = 0;
    int Pr = 0, Pc = 0;
    int64_t P = 0;
    int K = 0, kmax = 0;

    // state
    bool has_image = false, has_patches = false, has_dict = false, has_codes = false;
    bool has_rebuild = false;
    bool gram_valid = false;

    // device buffers
    DevBuf img, Y, D, G, idx, coef, cnt, counters;
    DevBuf c0;  // first-step correlations of a chunk of patches (pursuit of large patches or dictionaries)
    DevBuf blockcnt, ucount, uoff, users, uslot, E, wbuf;
    DevBuf moff, tab_p, tab_self, tab_oth;  // column table of the sweep
    DevBuf srank, handx, E2, gpart, ticket; // sweep kernels (sweep.hip); ticket: flags of the persistent sweep
    DevBuf rsum, rweight, rout, rpart, rlam;
    DevBuf Dbf, Dq, bfscal;  // BF16 form: LDS image of D, padded FP32 atoms, largest atom norm
    DevBuf rbound;           // per patch: residual norm of its last OMP (the proposer's bound)

    // learned support proposer (mlp.hip)
    DevBuf mlp_theta, mlp_m, mlp_v, mlp_wb, mlp_grad;        // weights (FP32 master, Adam moments, BF16), split-K grads
    DevBuf mlp_xb, mlp_h1b, mlp_h2b, mlp_zb, mlp_dz, mlp_dh2, mlp_dh1, mlp_lpart;  // one training batch
    DevBuf mlp_tr;  // transposed operands of the backward products
    DevBuf mlp_x, mlp_perm;                                  // normalised patches (all), training order
    uint64_t mlp_perm_seed = 0;   // mlp_perm holds the seeded permutation of mlp_perm_n patches
    int64_t mlp_perm_n = -1;
    DevBuf mlp_h1c, mlp_h2c, mlp_zc, mlp_fb, mlp_cnt;        // proposal chunk, fallback list and its length
    DevBuf mlp_rej, mlp_bcnt;                                // per-patch fallback flags, per-block counts
    DevBuf multi;                                            // a batch's pictures for the multi-picture sweep
    int multi_mode = 0;                                      // of that batch: 1 two-team sweeps, 2 large patches, 3 eight-team sweeps, 4 one-launch large patches
    bool mlp_ready = false, mlp_x_valid = false, mlp_trained = false;
    bool mlp_gated = false;  // the proposer stood down for this dictionary (acceptance below break-even)
    // The two-team sweep leaves the coefficients of the users beyond the cap
    // to one clearing pass; the pass is run only before something reads the
    // codes (they are rewritten by the next OMP otherwise)
    bool sweep_defers_zero = false, cap_zero_pending = false;
    // A persistent sweep that gave up (check_sweep sets sweep_failed) is
    // undone from the snapshots and run again with the two-kernel sweep
    // (force_two_kernel, launched eagerly, not from the captured graph)
    bool sweep_failed = false, force_two_kernel = false;
    DevBuf snap_D, snap_idx, snap_coef, snap_cnt;
    // rebuild: the support stride whose tiled kernel this context has given
    // its LDS ceiling (0: none yet)
    size_t rebuild_lds_set = 0;
    // the last proposal pass's fallback count, copied to pinned host memory
    // behind gate_ev and read at the next learned-form iteration
    unsigned *gate_host = nullptr;
    void *gate_ev = nullptr;  // hipEvent_t
    void *batch_ev = nullptr;  // hipEvent_t: this context's work before a multi-picture sweep
    bool gate_pending = false, gate_probe = false;
    int64_t gate_n = 0;
    int64_t mlp_chunk_rows = 0;
    int64_t mlp_step = 0;
    int64_t iters_done = 0;  // K-SVD iterations run on this image (learned form schedule)

    ksvd_timings tm{};
    std::string err;
    void *timeline = nullptr;  // event bookkeeping owned by capi.hip
    void *upd_graph = nullptr;  // captured dictionary sweep (hipGraph_t)
    void *upd_exec = nullptr;   // its executable (hipGraphExec_t)
};

// Set ctx->err and return code.
int ksvd_fail(ksvd_ctx *ctx, int code, const std::string &msg);

// Let kernel fn use as much dynamic LDS as a workgroup can hold beside its
// static LDS (160 KB in all on gfx950).  The attribute belongs to the kernel,
// which every context on the device shares, so it is never set to one
// context's size, which another context's launch could find too small.
inline hipError_t allow_max_lds(const void *fn)
{
    hipFuncAttributes fa;
    const hipError_t e = hipFuncGetAttributes(&fa, fn);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - (int)fa.sharedSizeBytes);
}
// Grow a device buffer to at least `bytes` (contents not preserved).
int ksvd_reserve(ksvd_ctx *ctx, DevBuf &b, size_t bytes);

// Blocking copies and fills on the context's own stream: the legacy null
// stream that hipMemcpy / hipMemset use may not wait on a stream that another
// thread is capturing, and HIP refuses such a call.
inline hipError_t ksvd_copy(ksvd_ctx *ctx, void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, ctx->stream);
    return e != hipSuccess ? e : hipStreamSynchronize(ctx->stream);
}

inline hipError_t ksvd_set(ksvd_ctx *ctx, void *dst, int value, size_t bytes)
{
    const hipError_t e = hipMemsetAsync(dst, value, bytes, ctx->stream);
    return e != hipSuccess ? e : hipStreamSynchronize(ctx->stream);
}

#define KSVD_HIP_CHECK(ctx, call)                                                                    \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            return ksvd_fail((ctx), KSVD_ERR_HIP,                                                    \
                             std::string(#call) + ": " + hipGetErrorString(e_));                     \
    } while (0)

// ---- launchers (each enqueues on ctx->stream and returns KSVD_OK or fails) --
int launch_gather(ksvd_ctx *ctx);
int launch_init_dict(ksvd_ctx *ctx);
int launch_dict_gram(ksvd_ctx *ctx);
hipError_t launch_dict_gram_raw(const float *D, int dim, int K, float *G, hipStream_t stream);
int launch_omp(ksvd_ctx *ctx);
int launch_omp_bf16(ksvd_ctx *ctx);
// BF16 OMP over the patches listed in `plist` (length *plist_n on the device).
int launch_omp_bf16_list(ksvd_ctx *ctx, const int *plist, const unsigned *plist_n);
// The sparse-coding kernel of a context: the FP32 pursuit for the plain form
// and for the learned form (its fallback and the judge of its proposals),
// the BF16-correlation pursuit for the BF16 form alone.
bool omp_uses_fp32(const ksvd_ctx *ctx);
int launch_omp_fp32_list(ksvd_ctx *ctx, const int *plist, const unsigned *plist_n);
// Runs the clearing pass the last sweep left pending, if any (sweep.hip).
int flush_cap_zero(ksvd_ctx *ctx);
// learned support proposer (mlp.hip)
int mlp_dims(const ksvd_ctx *ctx, int *kin, int *ko);
int64_t mlp_param_count(const ksvd_ctx *ctx);
int mlp_init(ksvd_ctx *ctx, uint64_t seed);
int mlp_set_params(ksvd_ctx *ctx, const float *theta);
int mlp_get_params(ksvd_ctx *ctx, float *theta);
int mlp_loss_grad_host(ksvd_ctx *ctx, const float *X, const float *T, int nb, double *loss, float *grad);
int mlp_train_all(ksvd_ctx *ctx, int steps, float lr, uint64_t seed);
int mlp_prepare(ksvd_ctx *ctx, uint64_t init_seed, uint64_t perm_seed);
int mlp_train_fallback(ksvd_ctx *ctx, int steps, float lr);
// proposals for patches [0, n); write = false only counts the rejected ones
int mlp_propose(ksvd_ctx *ctx, float accept, int64_t n, bool write);
size_t bf16_dict_bytes(int dim);
size_t bf16_dictq_bytes(int dim);
int launch_update(ksvd_ctx *ctx);
int launch_update_batch(ksvd_ctx *const *ctxs, int n);  // one update of several pictures (multi-picture sweep)
int sweep_multi_setup(ksvd_ctx *const *ctxs, int n, float tol);
int sweep_multi_tol_setup(ksvd_ctx *const *ctxs, int n);  // sweep_multi_setup at the update's tolerance
int sweep_multi_reset(ksvd_ctx *const *ctxs, int n);
int launch_sweep_multi(ksvd_ctx *const *ctxs, int n);
void update_release(ksvd_ctx *ctx);  // at destroy: the device's last-update event, if this context's
std::mutex &device_op_mutex();       // held over allocations, frees and graph captures
int prepare_update_graph(ksvd_ctx *ctx);  // capture the update's launches for the current buffers
void drop_update_graph(ksvd_ctx *ctx);
size_t sweep_gpart_bytes(int dim);
// Synchronises the stream; fails if a persistent sweep gave up waiting.
int check_sweep(ksvd_ctx *ctx);
int launch_rebuild(ksvd_ctx *ctx, float sigma2, bool speckle);
// The calling thread's last ksvd_omp_noalloc error, or nullptr (omp.hip).
const char *noalloc_last_error();
