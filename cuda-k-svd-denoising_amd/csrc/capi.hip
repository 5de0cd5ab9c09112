// C ABI of the gfx950 K-SVD library (declarations and reference citations in
// include/ksvd_hip.h).  Host orchestration only: buffer sizing, state checks,
// launches on the context's stream, HIP-event timing of every phase.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ksvd_internal.hpp"

namespace {

enum SpanKind { SPAN_GATHER, SPAN_INIT, SPAN_OMP, SPAN_OMP_KERNEL, SPAN_UPDATE, SPAN_REBUILD, SPAN_TRAIN, SPAN_KINDS };

struct Span {
    hipEvent_t a, b;
    int kind;
};

// Per-context event bookkeeping, kept beside the context so that the public
// struct stays plain.
struct Timeline {
    std::vector<hipEvent_t> pool;
    std::vector<Span> pending;
};

Timeline &timeline(ksvd_ctx *ctx)
{
    if (!ctx->timeline) ctx->timeline = new Timeline();
    return *(Timeline *)ctx->timeline;
}

void drop_timeline(ksvd_ctx *ctx)
{
    if (!ctx->timeline) return;
    Timeline *tl = (Timeline *)ctx->timeline;
    for (auto &s : tl->pending) {
        (void)hipEventDestroy(s.a);
        (void)hipEventDestroy(s.b);
    }
    for (auto e : tl->pool) (void)hipEventDestroy(e);
    delete tl;
    ctx->timeline = nullptr;
}

int get_event(ksvd_ctx *ctx, hipEvent_t *e)
{
    Timeline &tl = timeline(ctx);
    if (!tl.pool.empty()) {
        *e = tl.pool.back();
        tl.pool.pop_back();
        return KSVD_OK;
    }
    KSVD_HIP_CHECK(ctx, hipEventCreate(e));
    return KSVD_OK;
}

int span_begin(ksvd_ctx *ctx, Span *s, int kind)
{
    s->kind = kind;
    int rc = get_event(ctx, &s->a);
    if (rc) return rc;
    rc = get_event(ctx, &s->b);
    if (rc) return rc;
    KSVD_HIP_CHECK(ctx, hipEventRecord(s->a, ctx->stream));
    return KSVD_OK;
}

int span_end(ksvd_ctx *ctx, Span *s)
{
    KSVD_HIP_CHECK(ctx, hipEventRecord(s->b, ctx->stream));
    timeline(ctx).pending.push_back(*s);
    return KSVD_OK;
}

// Wait for the stream and fold every finished span into ctx->tm.
int resolve(ksvd_ctx *ctx)
{
    KSVD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    Timeline &tl = timeline(ctx);
    for (auto &s : tl.pending) {
        float ms = 0.f;
        KSVD_HIP_CHECK(ctx, hipEventElapsedTime(&ms, s.a, s.b));
        switch (s.kind) {
        case SPAN_GATHER: ctx->tm.gather_ms += ms; break;
        case SPAN_INIT: ctx->tm.init_ms += ms; break;
        case SPAN_OMP: ctx->tm.omp_ms += ms; ctx->tm.omp_calls++; break;
        case SPAN_OMP_KERNEL: ctx->tm.omp_kernel_ms += ms; break;
        case SPAN_UPDATE: ctx->tm.update_ms += ms; ctx->tm.update_calls++; break;
        case SPAN_REBUILD: ctx->tm.rebuild_ms += ms; break;
        case SPAN_TRAIN: ctx->tm.train_ms += ms; break;
        default: break;
        }
        tl.pool.push_back(s.a);
        tl.pool.push_back(s.b);
    }
    tl.pending.clear();
    return KSVD_OK;
}

int set_device(ksvd_ctx *ctx)
{
    KSVD_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    return KSVD_OK;
}

int check_params(ksvd_ctx *ctx, const ksvd_params *p)
{
    if (p->patch_w <= 0 || p->patch_h <= 0) return ksvd_fail(ctx, KSVD_ERR_ARG, "patch dimensions must be positive");
    if (p->slide_w <= 0 || p->slide_h <= 0) return ksvd_fail(ctx, KSVD_ERR_ARG, "sliding steps must be positive");
    if (p->atoms <= 0 || p->atoms > 1024) return ksvd_fail(ctx, KSVD_ERR_ARG, "atoms must be in [1, 1024]");
    if (p->omp_iter <= 0 || p->omp_iter > KSVD_MAX_SUPPORT)
        return ksvd_fail(ctx, KSVD_ERR_ARG, "ompIter must be in [1, 16]");

    if (p->ksvd_iter < 0) return ksvd_fail(ctx, KSVD_ERR_ARG, "ksvdIter must be >= 0");
    if (p->user_cap < 0) return ksvd_fail(ctx, KSVD_ERR_ARG, "user_cap must be >= 0");
    const int dim = p->patch_w * p->patch_h;
    if (dim > KSVD_MAX_DIM) return ksvd_fail(ctx, KSVD_ERR_UNSUPPORTED, "patch dimension above 1024");
    if (p->precision != KSVD_PREC_FP32 && p->precision != KSVD_PREC_BF16)
        return ksvd_fail(ctx, KSVD_ERR_ARG, "unknown precision");
    if (p->precision == KSVD_PREC_BF16 && (dim > 144 || p->atoms > 256 || p->omp_iter > 8))
        return ksvd_fail(ctx, KSVD_ERR_UNSUPPORTED, "the BF16 form needs patch dim <= 144, atoms <= 256, ompIter <= 8");
    if (p->geometry != KSVD_GEOM_TRUE && p->geometry != KSVD_GEOM_COMPAT)
        return ksvd_fail(ctx, KSVD_ERR_ARG, "unknown geometry");
    if (p->proposer != 0 && p->proposer != 1) return ksvd_fail(ctx, KSVD_ERR_ARG, "proposer must be 0 or 1");
    if (p->proposer && p->precision != KSVD_PREC_BF16)
        return ksvd_fail(ctx, KSVD_ERR_UNSUPPORTED, "the learned proposer belongs to the BF16 form");
    return KSVD_OK;
}

int sync_copy_d2h(ksvd_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    int rc = resolve(ctx);
    if (rc) return rc;
    KSVD_HIP_CHECK(ctx, ksvd_copy(ctx, dst, src, bytes, hipMemcpyDeviceToHost));
    return KSVD_OK;
}

}  // namespace

int ksvd_fail(ksvd_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

int ksvd_reserve(ksvd_ctx *ctx, DevBuf &b, size_t bytes)
{
    if (bytes == 0) bytes = 256;
    if (b.bytes >= bytes) return KSVD_OK;
    if (b.ptr) {
        KSVD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        hipError_t ef;
        {
            std::lock_guard<std::mutex> lock(device_op_mutex());
            ef = hipFree(b.ptr);
        }
        KSVD_HIP_CHECK(ctx, ef);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    hipError_t e;
    {
        std::lock_guard<std::mutex> lock(device_op_mutex());
        e = hipMalloc(&b.ptr, bytes);
    }
    if (e != hipSuccess) {
        b.ptr = nullptr;
        return ksvd_fail(ctx, KSVD_ERR_ALLOC, std::string("hipMalloc of ") + std::to_string(bytes) +
                                                  " bytes failed: " + hipGetErrorString(e));
    }
    b.bytes = bytes;
    return KSVD_OK;
}

extern "C" {

void ksvd_default_params(ksvd_params *p)
{
    if (!p) return;
    p->patch_w = 8;
    p->patch_h = 8;
    p->slide_w = 2;
    p->slide_h = 2;
    p->atoms = 256;
    p->omp_iter = 5;
    p->ksvd_iter = 10;
    p->min_omp_iter_batch = 0;
    p->user_cap = 1024;
    p->precision = KSVD_PREC_FP32;
    p->geometry = KSVD_GEOM_TRUE;
    p->device = 0;
    p->proposer = 0;
}

const char *ksvd_last_error(const ksvd_ctx *ctx)
{
    if (!ctx) {
        const char *m = noalloc_last_error();
        return m ? m : "null context";
    }
    return ctx->err.c_str();
}

int ksvd_create(const ksvd_params *p, ksvd_ctx **out)
{
    if (!p || !out) return KSVD_ERR_ARG;
    *out = nullptr;
    ksvd_ctx *ctx = new ksvd_ctx();
    int rc = check_params(ctx, p);
    if (rc) {
        // keep the context so the caller can read the message, then destroy it
        *out = ctx;
        return rc;
    }
    ctx->prm = *p;
    ctx->device = p->device;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        *out = ctx;
        return ksvd_fail(ctx, KSVD_ERR_HIP, "no HIP device available");
    }
    if (p->device < 0 || p->device >= ndev) {
        *out = ctx;
        return ksvd_fail(ctx, KSVD_ERR_ARG, "device ordinal out of range");
    }
    *out = ctx;
    rc = set_device(ctx);
    if (rc) return rc;
    hipDeviceProp_t prop;
    KSVD_HIP_CHECK(ctx, hipGetDeviceProperties(&prop, ctx->device));
    ctx->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    int nx = 0;
    if (hipDeviceGetAttribute(&nx, hipDeviceAttributeNumberOfXccs, ctx->device) == hipSuccess && nx > 0)
        ctx->n_xcd = nx;
    KSVD_HIP_CHECK(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    ctx->dim = p->patch_w * p->patch_h;
    ctx->K = p->atoms;
    ctx->kmax = p->omp_iter;
    rc = ksvd_reserve(ctx, ctx->counters, sizeof(unsigned long long) * CNT_SLOTS);
    if (rc) return rc;
    KSVD_HIP_CHECK(ctx, hipMemsetAsync(ctx->counters.ptr, 0, ctx->counters.bytes, ctx->stream));
    rc = ksvd_reserve(ctx, ctx->D, sizeof(float) * (size_t)ctx->dim * ctx->K);
    if (rc) return rc;
    // the dictionary Gram, with 256 floats of slack after it: the pursuit
    // reads whole 16-byte groups of a row and zeroes the entries past K
    rc = ksvd_reserve(ctx, ctx->G, sizeof(float) * ((size_t)ctx->K * ctx->K + 256));
    if (rc) return rc;
    rc = ksvd_reserve(ctx, ctx->wbuf, sizeof(float) * (size_t)ctx->dim * ctx->dim);
    if (rc) return rc;
    rc = ksvd_reserve(ctx, ctx->rlam, 16);
    if (rc) return rc;
    if (p->precision == KSVD_PREC_BF16) {
        if ((rc = ksvd_reserve(ctx, ctx->Dbf, bf16_dict_bytes(ctx->dim)))) return rc;
        if ((rc = ksvd_reserve(ctx, ctx->Dq, bf16_dictq_bytes(ctx->dim)))) return rc;
        if ((rc = ksvd_reserve(ctx, ctx->bfscal, 256))) return rc;
    }

    KSVD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return KSVD_OK;
}

int ksvd_destroy(ksvd_ctx *ctx)
{
    if (!ctx) return KSVD_OK;
    if (ctx->stream) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
    }
    drop_timeline(ctx);
    drop_update_graph(ctx);
    update_release(ctx);
    std::lock_guard<std::mutex> lock(device_op_mutex());
    if (ctx->gate_ev) (void)hipEventDestroy((hipEvent_t)ctx->gate_ev);
    if (ctx->batch_ev) (void)hipEventDestroy((hipEvent_t)ctx->batch_ev);
    if (ctx->gate_host) (void)hipHostFree(ctx->gate_host);
    DevBuf *bufs[] = {&ctx->img,  &ctx->Y,      &ctx->D,     &ctx->G,       &ctx->idx,
                      &ctx->coef, &ctx->cnt,    &ctx->counters, &ctx->blockcnt, &ctx->ucount, &ctx->uoff,
                      &ctx->users, &ctx->uslot, &ctx->E,     &ctx->wbuf, &ctx->moff, &ctx->tab_p, &ctx->tab_self, &ctx->tab_oth, &ctx->srank, &ctx->handx, &ctx->E2, &ctx->gpart, &ctx->ticket,    &ctx->rsum,    &ctx->rweight,
                      &ctx->rout, &ctx->rpart,  &ctx->rlam, &ctx->Dbf, &ctx->Dq, &ctx->bfscal, &ctx->rbound,
                      &ctx->mlp_theta, &ctx->mlp_m, &ctx->mlp_v, &ctx->mlp_wb, &ctx->mlp_grad, &ctx->mlp_xb,
                      &ctx->mlp_h1b, &ctx->mlp_h2b, &ctx->mlp_zb, &ctx->mlp_dz, &ctx->mlp_dh2, &ctx->mlp_dh1,
                      &ctx->mlp_lpart, &ctx->mlp_tr, &ctx->mlp_x, &ctx->mlp_perm, &ctx->mlp_h1c, &ctx->mlp_h2c, &ctx->mlp_zc,
                      &ctx->mlp_fb, &ctx->mlp_cnt, &ctx->mlp_rej, &ctx->mlp_bcnt, &ctx->multi, &ctx->snap_D, &ctx->snap_idx, &ctx->snap_coef, &ctx->snap_cnt, &ctx->c0};
    for (DevBuf *b : bufs)
        if (b->ptr) (void)hipFree(b->ptr);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return KSVD_OK;
}

static int set_image_common(ksvd_ctx *ctx, const float *buf, int m, int n, hipMemcpyKind kind)
{
    if (!ctx) return KSVD_ERR_ARG;
    if (!ctx->stream) return ksvd_fail(ctx, KSVD_ERR_STATE, "context was not created successfully");
    if (!buf || m <= 0 || n <= 0) return ksvd_fail(ctx, KSVD_ERR_ARG, "image buffer and sizes must be valid");
    const ksvd_params &p = ctx->prm;
    // The caller's buffer holds the picture column by column (buf[x*H + y],
    // m = W, n = H), which the reference reads as n rows of length m.  The
    // true geometry reads it as what it holds, m rows of length n, so that a
    // patch is a block of the picture even when W != H; the compat geometry
    // keeps the reference's reading (the same for a square picture).
    const bool true_geom = p.geometry == KSVD_GEOM_TRUE;
    const int fm = true_geom ? n : m, fn = true_geom ? m : n;
    if (p.patch_w > fm || p.patch_h > fn)
        return ksvd_fail(ctx, KSVD_ERR_ARG, "patch larger than the image");
    int rc = set_device(ctx);
    if (rc) return rc;
    rc = resolve(ctx);
    if (rc) return rc;
    const int64_t oldP = ctx->P;
    ctx->m = fm;
    ctx->n = fn;
    m = fm;
    n = fn;
    ctx->Pr = (n - p.patch_h) / p.slide_h + 1;
    ctx->Pc = (m - p.patch_w) / p.slide_w + 1;
    ctx->P = (int64_t)ctx->Pr * ctx->Pc;
    const int64_t P = ctx->P, N = (int64_t)m * n;
    const int dim = ctx->dim, K = ctx->K, kmax = ctx->kmax;
    const int64_t nb = (P + 255) / 256;
    const int64_t ecols = p.user_cap > 0 ? p.user_cap : P;
    // columns of the sweep's table: sum over atoms of min(users, cap) <= both
    // P * kmax (every support entry) and K * cap
    const int64_t tcols = p.user_cap > 0 ? std::min<int64_t>(P * kmax, (int64_t)K * p.user_cap) : P * kmax;
    struct {
        DevBuf *b;
        size_t bytes;
    } plan[] = {
        {&ctx->img, sizeof(float) * (size_t)N},
        {&ctx->Y, sizeof(float) * (size_t)dim * P},
        {&ctx->idx, sizeof(int) * (size_t)P * kmax},
        {&ctx->coef, sizeof(float) * (size_t)P * kmax},
        {&ctx->cnt, sizeof(int) * (size_t)P},
        {&ctx->blockcnt, sizeof(int) * (size_t)nb * K},
        {&ctx->ucount, sizeof(int) * (size_t)K},
        {&ctx->uoff, sizeof(int) * (size_t)(K + 1)},
        {&ctx->users, sizeof(int) * (size_t)P * kmax},
        {&ctx->uslot, sizeof(int) * (size_t)P * kmax},
        // E columns: pitch up to 192 floats (the tile sweep pads columns to 4Q)
        {&ctx->E, sizeof(float) * (size_t)std::max(dim, 192) * ecols},
        {&ctx->moff, sizeof(int) * (size_t)(K + 1)},
        {&ctx->tab_p, sizeof(int) * (size_t)tcols},
        {&ctx->tab_self, sizeof(int) * (size_t)tcols},
        {&ctx->tab_oth, sizeof(int) * 4 * (size_t)tcols * (kmax > 1 ? kmax - 1 : 1)},
        {&ctx->srank, sizeof(int) * (size_t)P * kmax},
        {&ctx->handx, sizeof(float) * (size_t)tcols * (kmax > 1 ? kmax - 1 : 1)},
        {&ctx->E2, sizeof(float) * (size_t)std::max(dim, 192) * ecols},
        {&ctx->gpart, sweep_gpart_bytes(dim)},
        {&ctx->ticket, 1024},  // flags of the persistent sweep
        {&ctx->rsum, sizeof(float) * (size_t)N},
        {&ctx->rweight, sizeof(float) * (size_t)N},
        {&ctx->rout, sizeof(float) * (size_t)N},
        {&ctx->rpart, sizeof(double) * (size_t)((N + 255) / 256)},
        {&ctx->rbound, sizeof(float) * (size_t)P},
    };
    // the update's graph holds these buffers' addresses and P: it is kept
    // only when neither changes
    bool same = ctx->P == oldP;
    for (auto &e : plan) {
        void *before = e.b->ptr;
        rc = ksvd_reserve(ctx, *e.b, e.bytes);
        if (rc) {
            drop_update_graph(ctx);
            return rc;
        }
        same = same && e.b->ptr == before;
    }
    if (!same) drop_update_graph(ctx);
    // flags of the persistent sweep (its reset kernel keeps the sticky abort word)
    KSVD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    KSVD_HIP_CHECK(ctx, ksvd_set(ctx, ctx->ticket.ptr, 0, ctx->ticket.bytes));
    KSVD_HIP_CHECK(ctx, ksvd_copy(ctx, ctx->img.ptr, buf, sizeof(float) * (size_t)N, kind));
    if ((rc = prepare_update_graph(ctx))) return rc;
    ctx->has_image = true;
    ctx->has_patches = ctx->has_codes = ctx->has_rebuild = false;
    ctx->cap_zero_pending = false;
    ctx->mlp_x_valid = false;
    return KSVD_OK;
}

int ksvd_set_image(ksvd_ctx *ctx, const float *buf, int m, int n)
{
    return set_image_common(ctx, buf, m, n, hipMemcpyHostToDevice);
}

int ksvd_set_image_device(ksvd_ctx *ctx, const float *dbuf, int m, int n)
{
    return set_image_common(ctx, dbuf, m, n, hipMemcpyDeviceToDevice);
}

// learned form schedule constants (see proposer_step)
constexpr int PROP_EPOCHS0 = 1, PROP_MIN_STEPS0 = 16, PROP_MAX_STEPS0 = 32, PROP_STEPS1 = 4, PROP_PROBE_STEPS = 8;
constexpr float PROP_LR = 2e-3f, PROP_ACCEPT = 1.05f, PROP_MIN_ACCEPT = 0.5f;
constexpr uint64_t PROP_SEED = 0x5EEDull;

// The pinned word and the event through which a proposal pass's fallback
// count reaches the host (created once per context).
static int gate_setup(ksvd_ctx *ctx)
{
    if (!ctx->gate_host) {
        std::lock_guard<std::mutex> lock(device_op_mutex());
        KSVD_HIP_CHECK(ctx, hipHostMalloc((void **)&ctx->gate_host, sizeof(unsigned), hipHostMallocDefault));
    }
    if (!ctx->gate_ev) {
        hipEvent_t e;
        KSVD_HIP_CHECK(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        ctx->gate_ev = e;
    }
    return KSVD_OK;
}

int ksvd_extract_patches(ksvd_ctx *ctx)
{
    if (!ctx) return KSVD_ERR_ARG;
    if (!ctx->has_image) return ksvd_fail(ctx, KSVD_ERR_STATE, "no image set");
    int rc = set_device(ctx);
    if (rc) return rc;
    Span s;
    if ((rc = span_begin(ctx, &s, SPAN_GATHER))) return rc;
    if ((rc = launch_gather(ctx))) return rc;
    if ((rc = span_end(ctx, &s))) return rc;
    ctx->has_patches = true;
    ctx->has_codes = ctx->has_rebuild = false;
    ctx->cap_zero_pending = false;
    ctx->mlp_x_valid = false;
    // the learned form's network, inputs, buffers and first training order,
    // and the pinned word its gate reads
    if (ctx->prm.proposer && ctx->P >= MLP_BATCH) {
        if ((rc = mlp_prepare(ctx, PROP_SEED, PROP_SEED))) return rc;
        if ((rc = gate_setup(ctx))) return rc;
    }
    return KSVD_OK;
}

int ksvd_init_dictionary(ksvd_ctx *ctx)
{
    if (!ctx) return KSVD_ERR_ARG;
    if (!ctx->has_patches) return ksvd_fail(ctx, KSVD_ERR_STATE, "patches not extracted");
    if (ctx->P / 2 + ctx->K > ctx->P)
        return ksvd_fail(ctx, KSVD_ERR_ARG, "fewer patches than P/2 + atoms: the dictionary cannot be initialised");
    int rc = set_device(ctx);
    if (rc) return rc;
    Span s;
    if ((rc = span_begin(ctx, &s, SPAN_INIT))) return rc;
    if ((rc = launch_init_dict(ctx))) return rc;
    if ((rc = span_end(ctx, &s))) return rc;
    ctx->has_dict = true;
    ctx->gram_valid = false;
    ctx->mlp_trained = false;  // a new dictionary lineage: its atoms are not the ones the proposer learned
    ctx->mlp_gated = false;
    ctx->gate_pending = false;
    ctx->iters_done = 0;
    return KSVD_OK;
}

int ksvd_omp(ksvd_ctx *ctx)
{
    if (!ctx) return KSVD_ERR_ARG;
    if (!ctx->has_patches || !ctx->has_dict) return ksvd_fail(ctx, KSVD_ERR_STATE, "patches and dictionary required");
    int rc = set_device(ctx);
    if (rc) return rc;
    Span outer, inner;
    if ((rc = span_begin(ctx, &outer, SPAN_OMP))) return rc;
    if (!ctx->gram_valid && (rc = launch_dict_gram(ctx))) return rc;
    if ((rc = span_begin(ctx, &inner, SPAN_OMP_KERNEL))) return rc;
    if ((rc = launch_omp(ctx))) return rc;
    if ((rc = span_end(ctx, &inner))) return rc;
    if ((rc = span_end(ctx, &outer))) return rc;
    ctx->has_codes = true;
    ctx->has_rebuild = false;
    ctx->cap_zero_pending = false;  // every code was rewritten
    return KSVD_OK;
}

// Enqueues one sweep; a sweep that gave up is reported by check_sweep.
static int update_dictionary_async(ksvd_ctx *ctx)
{
    if (!ctx->has_codes || !ctx->has_dict) return ksvd_fail(ctx, KSVD_ERR_STATE, "codes and dictionary required");
    if (ctx->kmax > 8) return ksvd_fail(ctx, KSVD_ERR_UNSUPPORTED, "the dictionary sweep handles ompIter <= 8");
    int rc = set_device(ctx);
    if (rc) return rc;
    Span s;
    if ((rc = span_begin(ctx, &s, SPAN_UPDATE))) return rc;
    if ((rc = launch_update(ctx))) return rc;
    if ((rc = span_end(ctx, &s))) return rc;
    ctx->has_rebuild = false;
    return KSVD_OK;
}

// The state a persistent sweep that gives up must be undone to: the
// dictionary and (codes = true) the compact codes, copied on the stream.
static int take_snapshot(ksvd_ctx *ctx, bool codes)
{
    const size_t nd = sizeof(float) * (size_t)ctx->K * ctx->dim;
    int rc = ksvd_reserve(ctx, ctx->snap_D, nd);
    if (rc) return rc;
    KSVD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->snap_D.ptr, ctx->D.ptr, nd, hipMemcpyDeviceToDevice, ctx->stream));
    if (codes) {
        const size_t ni = sizeof(int) * (size_t)ctx->P * ctx->kmax, nc = sizeof(int) * (size_t)ctx->P;
        if ((rc = ksvd_reserve(ctx, ctx->snap_idx, ni))) return rc;
        if ((rc = ksvd_reserve(ctx, ctx->snap_coef, ni))) return rc;
        if ((rc = ksvd_reserve(ctx, ctx->snap_cnt, nc))) return rc;
        KSVD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->snap_idx.ptr, ctx->idx.ptr, ni, hipMemcpyDeviceToDevice, ctx->stream));
        KSVD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->snap_coef.ptr, ctx->coef.ptr, ni, hipMemcpyDeviceToDevice, ctx->stream));
        KSVD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->snap_cnt.ptr, ctx->cnt.ptr, nc, hipMemcpyDeviceToDevice, ctx->stream));
    }
    return KSVD_OK;
}

static int restore_snapshot(ksvd_ctx *ctx, bool codes)
{
    const size_t nd = sizeof(float) * (size_t)ctx->K * ctx->dim;
    KSVD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->D.ptr, ctx->snap_D.ptr, nd, hipMemcpyDeviceToDevice, ctx->stream));
    if (codes) {
        const size_t ni = sizeof(int) * (size_t)ctx->P * ctx->kmax, nc = sizeof(int) * (size_t)ctx->P;
        KSVD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->idx.ptr, ctx->snap_idx.ptr, ni, hipMemcpyDeviceToDevice, ctx->stream));
        KSVD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->coef.ptr, ctx->snap_coef.ptr, ni, hipMemcpyDeviceToDevice, ctx->stream));
        KSVD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->cnt.ptr, ctx->snap_cnt.ptr, nc, hipMemcpyDeviceToDevice, ctx->stream));
    }
    ctx->gram_valid = false;
    ctx->cap_zero_pending = false;
    return KSVD_OK;
}

int ksvd_update_dictionary(ksvd_ctx *ctx)
{
    if (!ctx) return KSVD_ERR_ARG;
    if (!ctx->has_codes || !ctx->has_dict) return ksvd_fail(ctx, KSVD_ERR_STATE, "codes and dictionary required");
    int rc = set_device(ctx);
    if (rc) return rc;
    // the codes as a reader sees them (the last sweep's clearing done), then
    // the state a failed persistent sweep is undone to
    if ((rc = flush_cap_zero(ctx))) return rc;
    if ((rc = take_snapshot(ctx, true))) return rc;
    if ((rc = update_dictionary_async(ctx))) return rc;
    rc = check_sweep(ctx);
    if (rc && ctx->sweep_failed) {
        // a persistent sweep gave up: undo it and sweep again with two
        // kernels per atom, which wait on nothing
        ctx->sweep_failed = false;
        ctx->err.clear();
        if ((rc = restore_snapshot(ctx, true))) return rc;
        ctx->force_two_kernel = true;
        rc = update_dictionary_async(ctx);
        ctx->force_two_kernel = false;
        if (rc) return rc;
        ctx->tm.sweep_fallbacks++;
        return check_sweep(ctx);
    }
    return rc;
}

// Learned form schedule: the first iteration on a dictionary runs OMP on
// every patch and trains the proposer on those pairs for PROP_PROBE_STEPS
// steps, then probes: proposals for the first MLP_CHUNK patches, counted
// against the residual bound but not kept (the codes stay OMP's, so a
// proposer that stands down leaves the BF16 form's codes).  Later iterations
// propose, run OMP on the rejected patches,
// and train on them (up to PROP_STEPS1 steps); the one after a probe that
// passed first trains the rest of the first-iteration steps.  A proposal pass
// costs about as much as an OMP pass, so when the probe or a pass accepts
// less than PROP_MIN_ACCEPT of its patches the proposer stands down for the
// rest of this dictionary's iterations and OMP runs instead.  Each pass's
// fallback count reaches the host behind an event and is read at the next
// iteration, so no iteration waits for the GPU.  The network, its inputs,
// its buffers and the first training order are prepared with the patches
// (ksvd_extract_patches), so that the first iteration does not pay for them.

static int proposer_pass(ksvd_ctx *ctx, int64_t n, bool probe);

// KSVD_PROPOSER_MIN_ACCEPT (tests): the acceptance below which the proposer
// stands down; 0 keeps it proposing whatever it accepts
static double min_accept()
{
    const char *v = getenv("KSVD_PROPOSER_MIN_ACCEPT");
    return v && v[0] ? atof(v) : (double)PROP_MIN_ACCEPT;
}

// Queues the copy of the last pass's fallback count (n patches) to the host.
static int post_gate(ksvd_ctx *ctx, int64_t n, bool probe)
{
    int rc = gate_setup(ctx);
    if (rc) return rc;
    KSVD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->gate_host, ctx->mlp_cnt.ptr, sizeof(unsigned), hipMemcpyDeviceToHost,
                                       ctx->stream));
    KSVD_HIP_CHECK(ctx, hipEventRecord((hipEvent_t)ctx->gate_ev, ctx->stream));
    ctx->gate_pending = true;
    ctx->gate_n = n;
    ctx->gate_probe = probe;
    return KSVD_OK;
}

static int proposer_step(ksvd_ctx *ctx)
{
    int rc;
    Span s;
    const int steps0 = (int)std::min<int64_t>(PROP_MAX_STEPS0, std::max<int64_t>(PROP_MIN_STEPS0,
                                                                                PROP_EPOCHS0 * ctx->P / MLP_BATCH));
    if (ctx->gate_pending) {
        // the last pass's count (its copy was queued before at least one
        // dictionary sweep, so this normally does not wait)
        KSVD_HIP_CHECK(ctx, hipEventSynchronize((hipEvent_t)ctx->gate_ev));
        ctx->gate_pending = false;
        const int64_t n = ctx->gate_n;
        if ((double)(n - (int64_t)*ctx->gate_host) < min_accept() * (double)n) {
            ctx->mlp_gated = true;
        } else if (ctx->gate_probe) {
            if ((rc = flush_cap_zero(ctx))) return rc;
            if ((rc = span_begin(ctx, &s, SPAN_TRAIN))) return rc;
            if ((rc = mlp_train_all(ctx, steps0 - PROP_PROBE_STEPS, PROP_LR, PROP_SEED + 1 + ctx->iters_done))) return rc;
            if ((rc = span_end(ctx, &s))) return rc;
        }
    }
    if (ctx->mlp_gated) return ksvd_omp(ctx);
    if (!ctx->mlp_trained) {
        if ((rc = ksvd_omp(ctx))) return rc;
        if (ctx->P < MLP_BATCH) return KSVD_OK;  // too few patches to train on: plain OMP
        if (!ctx->mlp_ready && (rc = mlp_init(ctx, PROP_SEED))) return rc;
        if ((rc = span_begin(ctx, &s, SPAN_TRAIN))) return rc;
        if ((rc = mlp_train_all(ctx, PROP_PROBE_STEPS, PROP_LR, PROP_SEED + ctx->iters_done))) return rc;
        if ((rc = span_end(ctx, &s))) return rc;
        ctx->mlp_trained = true;
        const int64_t n = std::min<int64_t>(ctx->P, MLP_CHUNK);
        if ((rc = proposer_pass(ctx, n, true))) return rc;
        return post_gate(ctx, n, true);
    }
    if ((rc = ksvd_proposer_omp(ctx))) return rc;
    if ((rc = post_gate(ctx, ctx->P, false))) return rc;
    if ((rc = span_begin(ctx, &s, SPAN_TRAIN))) return rc;
    if ((rc = mlp_train_fallback(ctx, PROP_STEPS1, PROP_LR))) return rc;
    return span_end(ctx, &s);
}

static int iterate_async(ksvd_ctx *ctx, int iters)
{
    for (int i = 0; i < iters; i++) {
        int rc = ctx->prm.proposer ? proposer_step(ctx) : ksvd_omp(ctx);
        if (rc) return rc;
        rc = update_dictionary_async(ctx);
        if (rc) return rc;
        ctx->iters_done++;
    }
    return KSVD_OK;
}

int ksvd_iterate(ksvd_ctx *ctx, int iters)
{
    if (!ctx) return KSVD_ERR_ARG;
    if (iters < 0) return ksvd_fail(ctx, KSVD_ERR_ARG, "iters must be >= 0");
    if (iters == 0) return KSVD_OK;
    if (!ctx->has_dict) return ksvd_fail(ctx, KSVD_ERR_STATE, "dictionary required");
    int rc = set_device(ctx);
    if (rc) return rc;
    // the iterations' first OMP rewrites the codes: the dictionary is the
    // state a failed persistent sweep is undone to (the sweeps are checked
    // once, after the last iteration, so that the host never waits between
    // them; a failure anywhere replays them all with the two-kernel sweep)
    if ((rc = take_snapshot(ctx, false))) return rc;
    const int64_t done0 = ctx->iters_done;
    if ((rc = iterate_async(ctx, iters))) return rc;
    rc = check_sweep(ctx);
    if (rc && ctx->sweep_failed) {
        ctx->sweep_failed = false;
        ctx->err.clear();
        if ((rc = restore_snapshot(ctx, false))) return rc;
        ctx->iters_done = done0;
        ctx->force_two_kernel = true;
        rc = iterate_async(ctx, iters);
        ctx->force_two_kernel = false;
        if (rc) return rc;
        ctx->tm.sweep_fallbacks++;
        return check_sweep(ctx);
    }
    return rc;
}

// Several pictures' iterations at once (ksvd_iterate_batch): each picture's
// pursuit on its own stream, then one multi-picture sweep with picture i on
// XCD i.  Contexts that cannot share a sweep (other sizes or forms, more
// pictures than XCDs) iterate one after another instead.
int ksvd_iterate_batch(ksvd_ctx *const *ctxs, int n, int iters)
{
    if (!ctxs || n <= 0) return KSVD_ERR_ARG;
    for (int i = 0; i < n; i++) {
        if (!ctxs[i]) return KSVD_ERR_ARG;
        for (int j = 0; j < i; j++)
            if (ctxs[j] == ctxs[i]) return ksvd_fail(ctxs[i], KSVD_ERR_ARG, "a context appears twice in the batch");
    }
    if (iters < 0) return ksvd_fail(ctxs[0], KSVD_ERR_ARG, "iters must be >= 0");
    if (iters == 0) return KSVD_OK;
    bool shared = n > 1;
    for (int i = 0; i < n; i++) {
        ksvd_ctx *c = ctxs[i];
        if (!c->has_dict) return ksvd_fail(c, KSVD_ERR_STATE, "dictionary required");
        if (c->prm.precision != KSVD_PREC_FP32 || c->prm.proposer || c->device != ctxs[0]->device) shared = false;
    }
    int rc;
    if (shared) {
        if ((rc = set_device(ctxs[0]))) return rc;
        rc = sweep_multi_tol_setup(ctxs, n);
        if (rc < 0) return rc;
        shared = rc == 1;
    }
    if (!shared) {
        for (int i = 0; i < n; i++)
            if ((rc = ksvd_iterate(ctxs[i], iters))) return rc;
        return KSVD_OK;
    }
    std::vector<int64_t> done0(n);
    for (int i = 0; i < n; i++) {
        if ((rc = take_snapshot(ctxs[i], false))) return rc;
        done0[i] = ctxs[i]->iters_done;
    }
    std::vector<Span> sp(n);
    for (int it = 0; it < iters; it++) {
        for (int i = 0; i < n; i++)
            if ((rc = ksvd_omp(ctxs[i]))) return rc;
        for (int i = 0; i < n; i++) {
            if (ctxs[i]->kmax > 8)
                return ksvd_fail(ctxs[i], KSVD_ERR_UNSUPPORTED, "the dictionary sweep handles ompIter <= 8");
            if ((rc = span_begin(ctxs[i], &sp[i], SPAN_UPDATE))) return rc;
        }
        if ((rc = launch_update_batch(ctxs, n))) return rc;
        for (int i = 0; i < n; i++) {
            if ((rc = span_end(ctxs[i], &sp[i]))) return rc;
            ctxs[i]->has_rebuild = false;
            ctxs[i]->iters_done++;
        }
    }
    // a picture whose sweep gave up is undone and iterated again alone with
    // the two-kernel sweep
    for (int i = 0; i < n; i++) {
        ksvd_ctx *c = ctxs[i];
        rc = check_sweep(c);
        if (rc && c->sweep_failed) {
            c->sweep_failed = false;
            c->err.clear();
            if ((rc = restore_snapshot(c, false))) return rc;
            c->iters_done = done0[i];
            c->force_two_kernel = true;
            rc = iterate_async(c, iters);
            c->force_two_kernel = false;
            if (rc) return rc;
            c->tm.sweep_fallbacks++;
            rc = check_sweep(c);
        }
        if (rc) return rc;
    }
    return KSVD_OK;
}

int ksvd_run_batch(ksvd_ctx *const *ctxs, int n)
{
    if (!ctxs || n <= 0) return KSVD_ERR_ARG;
    for (int i = 0; i < n; i++)
        if (!ctxs[i]) return KSVD_ERR_ARG;
    // the pictures share the batch's iteration count (ctxs[0]'s ksvd_iter)
    for (int i = 1; i < n; i++)
        if (ctxs[i]->prm.ksvd_iter != ctxs[0]->prm.ksvd_iter)
            return ksvd_fail(ctxs[i], KSVD_ERR_ARG, "the batch's contexts differ in ksvd_iter");
    int rc = ksvd_iterate_batch(ctxs, n, ctxs[0]->prm.ksvd_iter);
    if (rc) return rc;
    for (int i = 0; i < n; i++)
        if ((rc = ksvd_run_final_code(ctxs[i]))) return rc;
    return KSVD_OK;
}

int ksvd_run_final_code(ksvd_ctx *ctx)
{
    if (!ctx) return KSVD_ERR_ARG;
    if (ctx->prm.proposer && ctx->mlp_trained && !ctx->mlp_gated) return ksvd_proposer_omp(ctx);
    return ksvd_omp(ctx);
}

int ksvd_run(ksvd_ctx *ctx)
{
    if (!ctx) return KSVD_ERR_ARG;
    int rc = ksvd_iterate(ctx, ctx->prm.ksvd_iter);
    if (rc) return rc;
    return ksvd_run_final_code(ctx);
}

static int rebuild_common(ksvd_ctx *ctx, float sigma2, float *out, bool speckle)
{
    if (!ctx) return KSVD_ERR_ARG;
    if (!ctx->has_codes || !ctx->has_dict || !ctx->has_image)
        return ksvd_fail(ctx, KSVD_ERR_STATE, "image, dictionary and codes required");
    if (!(sigma2 > 0.f) || !std::isfinite(sigma2))
        return ksvd_fail(ctx, KSVD_ERR_ARG, "sigma2 must be a positive finite noise variance");
    int rc = set_device(ctx);
    if (rc) return rc;
    if ((rc = flush_cap_zero(ctx))) return rc;
    Span s;
    if ((rc = span_begin(ctx, &s, SPAN_REBUILD))) return rc;
    if ((rc = launch_rebuild(ctx, sigma2, speckle))) return rc;
    if ((rc = span_end(ctx, &s))) return rc;
    ctx->has_rebuild = true;
    if (out) return sync_copy_d2h(ctx, out, ctx->rout.ptr, sizeof(float) * (size_t)ctx->m * ctx->n);
    return KSVD_OK;
}

int ksvd_rebuild(ksvd_ctx *ctx, float sigma2, float *out)
{
    return rebuild_common(ctx, sigma2, out, false);
}

int ksvd_rebuild_speckle(ksvd_ctx *ctx, float sigma2, float *out)
{
    return rebuild_common(ctx, sigma2, out, true);
}

int ksvd_get_rebuild_parts(ksvd_ctx *ctx, float *sum, float *weight, float *lambda, double *dsum)
{
    if (!ctx) return KSVD_ERR_ARG;
    if (!ctx->has_rebuild) return ksvd_fail(ctx, KSVD_ERR_STATE, "no rebuild yet");
    const size_t N = (size_t)ctx->m * ctx->n;
    int rc;
    if (sum && (rc = sync_copy_d2h(ctx, sum, ctx->rsum.ptr, sizeof(float) * N))) return rc;
    if (weight && (rc = sync_copy_d2h(ctx, weight, ctx->rweight.ptr, sizeof(float) * N))) return rc;
    float lam[4];
    if ((rc = sync_copy_d2h(ctx, lam, ctx->rlam.ptr, sizeof(lam)))) return rc;
    if (lambda) *lambda = lam[0];
    if (dsum) memcpy(dsum, &lam[2], sizeof(double));
    return KSVD_OK;
}

int ksvd_rebuild_device_ptr(ksvd_ctx *ctx, const float **dptr)
{
    if (!ctx || !dptr) return KSVD_ERR_ARG;
    if (!ctx->has_rebuild) return ksvd_fail(ctx, KSVD_ERR_STATE, "no rebuild yet");
    *dptr = (const float *)ctx->rout.ptr;
    return KSVD_OK;
}

int ksvd_get_dims(ksvd_ctx *ctx, int *dim, int64_t *patches, int *atoms, int *kmax)
{
    if (!ctx) return KSVD_ERR_ARG;
    if (dim) *dim = ctx->dim;
    if (patches) *patches = ctx->P;
    if (atoms) *atoms = ctx->K;
    if (kmax) *kmax = ctx->kmax;
    return KSVD_OK;
}

int ksvd_get_patches(ksvd_ctx *ctx, float *Y)
{
    if (!ctx || !Y) return KSVD_ERR_ARG;
    if (!ctx->has_patches) return ksvd_fail(ctx, KSVD_ERR_STATE, "patches not extracted");
    return sync_copy_d2h(ctx, Y, ctx->Y.ptr, sizeof(float) * (size_t)ctx->dim * ctx->P);
}

int ksvd_get_dictionary(ksvd_ctx *ctx, float *D)
{
    if (!ctx || !D) return KSVD_ERR_ARG;
    if (!ctx->has_dict) return ksvd_fail(ctx, KSVD_ERR_STATE, "no dictionary");
    return sync_copy_d2h(ctx, D, ctx->D.ptr, sizeof(float) * (size_t)ctx->dim * ctx->K);
}

int ksvd_set_dictionary(ksvd_ctx *ctx, const float *D)
{
    if (!ctx || !D) return KSVD_ERR_ARG;
    if (!ctx->stream) return ksvd_fail(ctx, KSVD_ERR_STATE, "context was not created successfully");
    int rc = set_device(ctx);
    if (rc) return rc;
    if ((rc = resolve(ctx))) return rc;
    KSVD_HIP_CHECK(ctx, ksvd_copy(ctx, ctx->D.ptr, D, sizeof(float) * (size_t)ctx->dim * ctx->K, hipMemcpyHostToDevice));
    ctx->has_dict = true;
    ctx->gram_valid = false;
    ctx->has_rebuild = false;
    ctx->mlp_trained = false;
    ctx->mlp_gated = false;
    ctx->gate_pending = false;
    ctx->iters_done = 0;
    return KSVD_OK;
}

int ksvd_get_codes(ksvd_ctx *ctx, int32_t *idx, float *coef, int32_t *cnt)
{
    if (!ctx) return KSVD_ERR_ARG;
    if (!ctx->has_codes) return ksvd_fail(ctx, KSVD_ERR_STATE, "no codes");
    const size_t pk = (size_t)ctx->P * ctx->kmax;
    int rc = set_device(ctx);
    if (rc) return rc;
    if ((rc = flush_cap_zero(ctx))) return rc;
    if (idx && (rc = sync_copy_d2h(ctx, idx, ctx->idx.ptr, sizeof(int) * pk))) return rc;
    if (coef && (rc = sync_copy_d2h(ctx, coef, ctx->coef.ptr, sizeof(float) * pk))) return rc;
    if (cnt && (rc = sync_copy_d2h(ctx, cnt, ctx->cnt.ptr, sizeof(int) * (size_t)ctx->P))) return rc;
    return KSVD_OK;
}

int ksvd_set_codes(ksvd_ctx *ctx, const int32_t *idx, const float *coef, const int32_t *cnt)
{
    if (!ctx || !idx || !coef || !cnt) return KSVD_ERR_ARG;
    if (!ctx->has_image) return ksvd_fail(ctx, KSVD_ERR_STATE, "no image set (patch count unknown)");
    const int64_t P = ctx->P;
    const int kmax = ctx->kmax, K = ctx->K;
    for (int64_t p = 0; p < P; p++) {
        if (cnt[p] < 0 || cnt[p] > kmax) return ksvd_fail(ctx, KSVD_ERR_ARG, "code count out of range");
        for (int i = 0; i < cnt[p]; i++) {
            const int a = idx[p * kmax + i];
            if (a < 0 || a >= K) return ksvd_fail(ctx, KSVD_ERR_ARG, "atom index out of range");
            for (int q = 0; q < i; q++)
                if (idx[p * kmax + q] == a) return ksvd_fail(ctx, KSVD_ERR_ARG, "repeated atom in a support");
        }
    }
    int rc = set_device(ctx);
    if (rc) return rc;
    if ((rc = resolve(ctx))) return rc;
    const size_t pk = (size_t)P * kmax;
    KSVD_HIP_CHECK(ctx, ksvd_copy(ctx, ctx->idx.ptr, idx, sizeof(int) * pk, hipMemcpyHostToDevice));
    KSVD_HIP_CHECK(ctx, ksvd_copy(ctx, ctx->coef.ptr, coef, sizeof(float) * pk, hipMemcpyHostToDevice));
    KSVD_HIP_CHECK(ctx, ksvd_copy(ctx, ctx->cnt.ptr, cnt, sizeof(int) * (size_t)P, hipMemcpyHostToDevice));
    ctx->has_codes = true;
    ctx->has_rebuild = false;
    ctx->cap_zero_pending = false;
    return KSVD_OK;
}

int ksvd_get_codes_dense(ksvd_ctx *ctx, float *X)
{
    if (!ctx || !X) return KSVD_ERR_ARG;
    if (!ctx->has_codes) return ksvd_fail(ctx, KSVD_ERR_STATE, "no codes");
    const int64_t P = ctx->P;
    const int kmax = ctx->kmax, K = ctx->K;
    std::vector<int> idx((size_t)P * kmax), cnt((size_t)P);
    std::vector<float> coef((size_t)P * kmax);
    int rc = ksvd_get_codes(ctx, idx.data(), coef.data(), cnt.data());
    if (rc) return rc;
    memset(X, 0, sizeof(float) * (size_t)K * P);
    for (int64_t p = 0; p < P; p++)
        for (int i = 0; i < cnt[p]; i++) X[(size_t)p * K + idx[(size_t)p * kmax + i]] = coef[(size_t)p * kmax + i];
    return KSVD_OK;
}

// ---- learned support proposer ------------------------------------------------

__global__ void proposer_count_kernel(const unsigned *n, int64_t P, unsigned long long *counters)
{
    counters[CNT_PROPOSED] += (unsigned long long)P;
    counters[CNT_FALLBACK] += (unsigned long long)*n;
}

static int proposer_ok(ksvd_ctx *ctx)
{
    if (!ctx) return KSVD_ERR_ARG;
    if (!ctx->stream) return ksvd_fail(ctx, KSVD_ERR_STATE, "context was not created successfully");
    if (ctx->prm.precision != KSVD_PREC_BF16)
        return ksvd_fail(ctx, KSVD_ERR_UNSUPPORTED, "the learned proposer belongs to the BF16 form");
    return set_device(ctx);
}

int ksvd_proposer_dims(ksvd_ctx *ctx, int *kin, int *ko, int64_t *nparams)
{
    int rc = proposer_ok(ctx);
    if (rc) return rc;
    int a, b;
    mlp_dims(ctx, &a, &b);
    if (kin) *kin = a;
    if (ko) *ko = b;
    if (nparams) *nparams = mlp_param_count(ctx);
    return KSVD_OK;
}

int ksvd_proposer_init(ksvd_ctx *ctx, uint64_t seed)
{
    int rc = proposer_ok(ctx);
    if (rc) return rc;
    ctx->mlp_trained = false;
    ctx->gate_pending = false;
    return mlp_init(ctx, seed);
}

int ksvd_proposer_get_params(ksvd_ctx *ctx, float *theta)
{
    int rc = proposer_ok(ctx);
    if (rc) return rc;
    if (!theta) return ksvd_fail(ctx, KSVD_ERR_ARG, "null parameter buffer");
    return mlp_get_params(ctx, theta);
}

int ksvd_proposer_set_params(ksvd_ctx *ctx, const float *theta)
{
    int rc = proposer_ok(ctx);
    if (rc) return rc;
    if (!theta) return ksvd_fail(ctx, KSVD_ERR_ARG, "null parameter buffer");
    return mlp_set_params(ctx, theta);
}

int ksvd_proposer_train(ksvd_ctx *ctx, int steps, float lr, uint64_t seed)
{
    int rc = proposer_ok(ctx);
    if (rc) return rc;
    if (!ctx->has_codes || !ctx->has_patches) return ksvd_fail(ctx, KSVD_ERR_STATE, "codes required as targets");
    if (steps < 0 || !(lr > 0.f)) return ksvd_fail(ctx, KSVD_ERR_ARG, "steps >= 0 and lr > 0 required");
    if (ctx->P < MLP_BATCH) return ksvd_fail(ctx, KSVD_ERR_ARG, "fewer patches than one training batch (16384)");
    if (!ctx->mlp_ready && (rc = mlp_init(ctx, PROP_SEED))) return rc;
    if ((rc = flush_cap_zero(ctx))) return rc;
    Span s;
    if ((rc = span_begin(ctx, &s, SPAN_TRAIN))) return rc;
    if ((rc = mlp_train_all(ctx, steps, lr, seed))) return rc;
    if ((rc = span_end(ctx, &s))) return rc;
    ctx->mlp_trained = true;
    return KSVD_OK;
}

// Proposals for patches [0, n), OMP for the ones rejected.  A probe counts
// the rejections only: the codes (OMP's, just computed) stay as they are.
static int proposer_pass(ksvd_ctx *ctx, int64_t n, bool probe)
{
    int rc;
    Span outer;
    if ((rc = span_begin(ctx, &outer, SPAN_OMP))) return rc;
    if (!ctx->gram_valid && (rc = launch_dict_gram(ctx))) return rc;
    if ((rc = mlp_propose(ctx, PROP_ACCEPT, n, !probe))) return rc;
    if (!probe &&
        (rc = omp_uses_fp32(ctx)
                  ? launch_omp_fp32_list(ctx, (const int *)ctx->mlp_fb.ptr, (const unsigned *)ctx->mlp_cnt.ptr)
                  : launch_omp_bf16_list(ctx, (const int *)ctx->mlp_fb.ptr, (const unsigned *)ctx->mlp_cnt.ptr)))
        return rc;
    hipLaunchKernelGGL(proposer_count_kernel, dim3(1), dim3(1), 0, ctx->stream, (const unsigned *)ctx->mlp_cnt.ptr,
                       n, (unsigned long long *)ctx->counters.ptr);
    KSVD_HIP_CHECK(ctx, hipGetLastError());
    if ((rc = span_end(ctx, &outer))) return rc;
    ctx->has_codes = true;
    ctx->has_rebuild = false;
    if (!probe) ctx->cap_zero_pending = false;  // proposals and fallback rewrote every code
    return KSVD_OK;
}

int ksvd_proposer_omp(ksvd_ctx *ctx)
{
    int rc = proposer_ok(ctx);
    if (rc) return rc;
    if (!ctx->has_patches || !ctx->has_dict) return ksvd_fail(ctx, KSVD_ERR_STATE, "patches and dictionary required");
    if (!ctx->mlp_ready) return ksvd_fail(ctx, KSVD_ERR_STATE, "proposer not initialised");
    if (!ctx->has_codes) return ksvd_fail(ctx, KSVD_ERR_STATE, "a first OMP (the residual bounds) is required");
    return proposer_pass(ctx, ctx->P, false);
}

int ksvd_proposer_loss_grad(ksvd_ctx *ctx, const float *X, const float *T, int nb, double *loss, float *grad)
{
    int rc = proposer_ok(ctx);
    if (rc) return rc;
    if (!X || !T || !loss || !grad) return ksvd_fail(ctx, KSVD_ERR_ARG, "null buffer");
    if (!ctx->mlp_ready) return ksvd_fail(ctx, KSVD_ERR_STATE, "proposer not initialised");
    if ((rc = resolve(ctx))) return rc;
    return mlp_loss_grad_host(ctx, X, T, nb, loss, grad);
}

int ksvd_get_timings(ksvd_ctx *ctx, ksvd_timings *t)
{
    if (!ctx || !t) return KSVD_ERR_ARG;
    if (!ctx->stream) return ksvd_fail(ctx, KSVD_ERR_STATE, "context was not created successfully");
    int rc = resolve(ctx);
    if (rc) return rc;
    unsigned long long c[CNT_SLOTS];
    KSVD_HIP_CHECK(ctx, ksvd_copy(ctx, c, ctx->counters.ptr, sizeof(c), hipMemcpyDeviceToHost));
    ctx->tm.omp_dup_stops = (int64_t)c[CNT_DUP];
    ctx->tm.omp_degenerate = (int64_t)c[CNT_DEGEN];
    ctx->tm.eig_iters = (int64_t)c[CNT_EIG_IT];
    ctx->tm.omp_rescored = (int64_t)c[CNT_RESCORE];
    ctx->tm.omp_f64_refactors = (int64_t)c[CNT_F64];
    ctx->tm.omp_candidates_cut = (int64_t)c[CNT_CAND_CUT];
    ctx->tm.proposer_gated = ctx->mlp_gated ? 1 : 0;
    ctx->tm.proposed = (int64_t)c[CNT_PROPOSED];
    ctx->tm.fallback = (int64_t)c[CNT_FALLBACK];
    if (ctx->mlp_ready) {
        float l = 0.f;
        KSVD_HIP_CHECK(ctx, ksvd_copy(ctx, &l, (const float *)ctx->mlp_cnt.ptr + 1, sizeof(float), hipMemcpyDeviceToHost));
        ctx->tm.train_loss = l;
    }
    *t = ctx->tm;
    return KSVD_OK;
}

int ksvd_reset_timings(ksvd_ctx *ctx)
{
    if (!ctx) return KSVD_ERR_ARG;
    int rc = resolve(ctx);
    if (rc) return rc;
    ctx->tm = ksvd_timings{};
    KSVD_HIP_CHECK(ctx, ksvd_set(ctx, ctx->counters.ptr, 0, ctx->counters.bytes));
    return KSVD_OK;
}

int ksvd_synchronize(ksvd_ctx *ctx)
{
    if (!ctx) return KSVD_ERR_ARG;
    if (!ctx->stream) return ksvd_fail(ctx, KSVD_ERR_STATE, "context was not created successfully");
    return resolve(ctx);
}

}  // extern "C"
