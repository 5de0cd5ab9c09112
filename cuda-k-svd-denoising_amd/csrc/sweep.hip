// The Gauss-Seidel dictionary sweep (cuda_svd_denoiser.cu:309-361) as two
// kernels per atom, spread over CUs, with no grid-wide barrier (a kernel
// boundary is the only cross-workgroup synchronisation).
//
// K1(k), B1 workgroups x 1024 threads:
//   (a) the new coefficients of atom k-1 for its users: x_t = e_t . d_{k-1}
//       for its first min(c, cap) users, 0 for the rest;
//   (b) the restricted residual E_k of atom k, a slice of columns per
//       workgroup: e_t = y_p - sum_{j != k} d_j x_{j,p}.  Atom k-1's
//       coefficient for p is not read from memory (part (a) of this same
//       launch is writing it) but recomputed as e_{k-1,r} . d_{k-1} from the
//       stored E_{k-1} (r = rank of p among atom k-1's users), with the same
//       code as (a), so both agree bit for bit;
//   (c) the partial Gram E_slice E_slice^T of the slice on MFMA
//       (v_mfma_f32_16x16x4_f32), upper 16x16 tiles only.
// K2(k), one workgroup x 1024 threads: sums the B1 partial Grams in a fixed
//   order straight into LDS (G = E_k E_k^T) and takes its top eigenvector by
//   power iteration from d_k; that vector becomes the new atom d_k.
// A final K1(K) writes atom K-1's coefficients.  E_k alternates between two
// buffers so that K1(k+1) reads E_k while writing E_{k+1}.
#include <algorithm>

#include "ksvd_internal.hpp"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int B1 = 8;       // K1 workgroups
constexpr int T1 = 1024;    // K1 threads
constexpr int CHUNK = 64;   // E columns per K1 chunk (held in LDS)
constexpr int T2 = 1024;    // K2 threads (4 lanes per row of the Gram)
constexpr int HAND_MAX = 4096;  // handoff entries per atom held in LDS
constexpr int MAX_TPW = 5;  // Gram tiles per wavefront in K1 (16 waves -> 80 tiles, dpad <= 192)

struct SweepArgs {
    const float *Y;
    float *D;
    float *coef;
    const int *uoff;
    const int *users;
    const int *uslot;
    const int *moff;
    const int *tab_p;
    const int *tab_self;
    const int4 *tab_oth;
    float *E0, *E1;
    float *Gpart;
    float *handx;  // atom k-1's new coefficient for the table entries of atom k's columns
    float *G;
    unsigned long long *counters;
    int dim, dpad, K, kmax, cap;
    int nbt, ntiles;
    int maxit;
    float tol2;
};

__device__ __forceinline__ int atom_m(const int *uoff, int k, int cap, int *c_out)
{
    const int c = uoff[k + 1] - uoff[k];
    *c_out = c;
    return (cap > 0 && c > cap) ? cap : c;
}

// e . u by one wavefront: lane l sums elements l, l+64, l+128, ... in order,
// then a fixed butterfly.  Used for the stored coefficients and for the
// recomputed ones, which therefore agree exactly.
__device__ __forceinline__ float wave_dot(const float *__restrict__ e, const float *__restrict__ u, int dim)
{
    const int lane = threadIdx.x & 63;
    float v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int q = lane + 64 * j;
        v[j] = q < dim ? e[q] : 0.f;
    }
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int q = lane + 64 * j;
        if (q < dim) s = fmaf(v[j], u[q], s);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    return s;
}

// upper tile index -> (ba, bb), ba <= bb, row-major over the upper triangle
__device__ __forceinline__ void tile_coords(int ti, int nbt, int *ba, int *bb)
{
    int a = 0, rem = ti;
    while (rem >= nbt - a) {
        rem -= nbt - a;
        a++;
    }
    *ba = a;
    *bb = a + rem;
}

template <int TPW, int NO, int BU>
__global__ __launch_bounds__(T1) void k1_kernel(SweepArgs a, int k)
{
    extern __shared__ float sm[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int dim = a.dim, dpad = a.dpad, kmax = a.kmax;
    const int no = kmax > 1 ? kmax - 1 : 1;
    // ---- (a) coefficients of atom k-1 --------------------------------------
    int cprev = 0, mprev = 0;
    const float *Eprev = nullptr, *uprev = nullptr;
    (void)no;
    if (k >= 1) {
        mprev = atom_m(a.uoff, k - 1, a.cap, &cprev);
        Eprev = ((k - 1) & 1) ? a.E1 : a.E0;
        uprev = a.D + (int64_t)(k - 1) * dim;
        if (cprev > 0) {
            const int64_t cbp = a.moff[k - 1];
            const int basep = a.uoff[k - 1];
            // wavefront w of workgroup b takes columns t = g*W + w for its
            // share of g; the columns' loads are issued before any use
            const int W = gridDim.x * (T1 / 64);
            const int w0 = blockIdx.x * (T1 / 64) + wave;
            constexpr int CPW = 4;
            for (int g0 = 0; g0 * W < mprev; g0 += CPW) {
                float ev[CPW][4], uv[4];
                int slot[CPW];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int q = lane + 64 * j;
                    uv[j] = q < dim ? uprev[q] : 0.f;
                }
#pragma unroll
                for (int g = 0; g < CPW; g++) {
                    const int t = (g0 + g) * W + w0;
                    slot[g] = t < mprev ? a.tab_self[cbp + t] : -1;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const int q = lane + 64 * j;
                        ev[g][j] = (t < mprev && q < dim) ? Eprev[(int64_t)t * dpad + q] : 0.f;
                    }
                }
#pragma unroll
                for (int g = 0; g < CPW; g++) {
                    float sdot = 0.f;
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        if (lane + 64 * j < dim) sdot = fmaf(ev[g][j], uv[j], sdot);
#pragma unroll
                    for (int off = 32; off > 0; off >>= 1) sdot += __shfl_xor(sdot, off);
                    if (lane == 0 && slot[g] >= 0) a.coef[slot[g]] = sdot;
                }
            }
            for (int t = mprev + blockIdx.x * T1 + tid; t < cprev; t += gridDim.x * T1) {
                const int pos = basep + t;
                a.coef[(int64_t)a.users[pos] * kmax + a.uslot[pos]] = 0.f;
            }
        }
    }
    if (k >= a.K) return;
    int c;
    const int m = atom_m(a.uoff, k, a.cap, &c);
    if (m == 0) return;
    const int cw = (m + B1 - 1) / B1;
    const int t_begin = blockIdx.x * cw;
    const int t_end = min(m, t_begin + cw);
    if (t_begin >= t_end) return;
    float *Ecur = (k & 1) ? a.E1 : a.E0;
    const int64_t cb = a.moff[k];

    float *Es = sm;                                   // CHUNK x dpad

    // MFMA accumulators of this wavefront's Gram tiles, kept across chunks
    f4 acc[TPW];
#pragma unroll
    for (int q = 0; q < TPW; q++) acc[q] = f4{0.f, 0.f, 0.f, 0.f};
    const int cpt = T1 / dim;
    const int tc = tid / dim, te = tid - tc * dim;

    for (int c0 = t_begin; c0 < t_end; c0 += CHUNK) {
        const int nc = min(CHUNK, t_end - c0);
        for (int q = tid; q < CHUNK * dpad; q += T1) Es[q] = 0.f;
        __syncthreads();
        // element-parallel build: thread (tc, te) takes element te of BU
        // columns at once; each thread reads its columns' table entries
        // directly (the same addresses across a column: broadcast), then the
        // patch, the atoms and the coefficients, all in flight together.
        // Atom k-1's coefficient comes from the handoff values that K2(k-1)
        // computed for exactly these entries.
        if (tc < cpt) {
            for (int t = tc; t < nc; t += BU * cpt) {
                int pp[BU];
                int4 oo[BU][NO];
#pragma unroll
                for (int b = 0; b < BU; b++) {
                    const int tt = t + b * cpt;
                    const int64_t col = cb + c0 + (tt < nc ? tt : 0);
                    pp[b] = a.tab_p[col];
#pragma unroll
                    for (int i = 0; i < NO; i++)
                        oo[b][i] = i < no ? a.tab_oth[col * no + i] : make_int4(-1, 0, -1, 0);
                }
                float v[BU];
#pragma unroll
                for (int b = 0; b < BU; b++) v[b] = a.Y[(int64_t)pp[b] * dim + te];
#pragma unroll
                for (int i = 0; i < NO; i++) {
                    float dv[BU], xv[BU];
#pragma unroll
                    for (int b = 0; b < BU; b++) {
                        const int tt = t + b * cpt;
                        const int4 o = oo[b][i];
                        const int64_t ent = (cb + c0 + tt) * no + i;
                        dv[b] = o.x >= 0 ? a.D[(int64_t)o.x * dim + te] : 0.f;
                        xv[b] = o.x < 0 ? 0.f : o.w ? a.handx[ent] : a.coef[o.y];
                    }
#pragma unroll
                    for (int b = 0; b < BU; b++) v[b] = fmaf(-dv[b], xv[b], v[b]);
                }
#pragma unroll
                for (int b = 0; b < BU; b++) {
                    const int tt = t + b * cpt;
                    if (tt < nc) {
                        Es[tt * dpad + te] = v[b];
                        Ecur[(int64_t)(c0 + tt) * dpad + te] = v[b];
                    }
                }
            }
        }
        __syncthreads();
        // partial Gram of the chunk: tile (ba, bb) += E_chunk[ba] E_chunk[bb]^T
        const int col16 = lane & 15, grp = lane >> 4;
#pragma unroll
        for (int q = 0; q < TPW; q++) {
            const int ti = wave + 16 * q;
            if (ti < a.ntiles) {
                int ba, bb;
                tile_coords(ti, a.nbt, &ba, &bb);
                for (int kk = 0; kk < CHUNK / 4; kk++) {
                    const int t = 4 * kk + grp;
                    const float av = Es[t * dpad + ba * 16 + col16];
                    const float bv = Es[t * dpad + bb * 16 + col16];
                    acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[q], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // write this workgroup's partial tiles
    float *gp = a.Gpart + (size_t)blockIdx.x * a.ntiles * 256;
#pragma unroll
    for (int q = 0; q < TPW; q++) {
        const int ti = wave + 16 * q;
        if (ti < a.ntiles)
            *(f4 *)(gp + (size_t)ti * 256 + lane * 4) = acc[q];
    }
}

__device__ __forceinline__ float block_sum_t2(float v, float *red)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float s = 0.f;
#pragma unroll
    for (int w = 0; w < T2 / 64; w++) s += red[w];
    __syncthreads();
    return s;
}

__global__ __launch_bounds__(T2) void k2_kernel(SweepArgs a, int k)
{
    extern __shared__ float sm[];
    __shared__ float red[T2 / 64];
    const int tid = threadIdx.x;
    const int dim = a.dim;
    int c;
    const int m = atom_m(a.uoff, k, a.cap, &c);
    if (m == 0) return;
    const int cw = (m + B1 - 1) / B1;
    const int nb_used = (m + cw - 1) / cw;
    float *Gs = sm;              // dim x dim (symmetric: column t read as row t)
    float *u = Gs + dim * dim;   // dim
    // ---- G = sum of the partial Grams (workgroup order), straight into LDS ----
    // float4 f of tile ti holds lane l = f % 64 of the MFMA output: rows
    // 4*(l/16) .. +3, column l % 16 of the tile
    const int nf4 = a.ntiles * 64;
    for (int f0 = 0; f0 < nf4; f0 += 2 * T2) {
        float4 acc2[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int f = f0 + h * T2 + tid;
            float4 pv[B1];
#pragma unroll
            for (int b = 0; b < B1; b++)
                pv[b] = (f < nf4 && b < nb_used) ? ((const float4 *)a.Gpart)[(size_t)b * nf4 + f]
                                                 : make_float4(0.f, 0.f, 0.f, 0.f);
            float4 sacc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int b = 0; b < B1; b++) {
                sacc.x += pv[b].x;
                sacc.y += pv[b].y;
                sacc.z += pv[b].z;
                sacc.w += pv[b].w;
            }
            acc2[h] = sacc;
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int f = f0 + h * T2 + tid;
            if (f >= nf4) continue;
            const int ti = f >> 6, l = f & 63;
            int ba, bb;
            tile_coords(ti, a.nbt, &ba, &bb);
            const float vals[4] = {acc2[h].x, acc2[h].y, acc2[h].z, acc2[h].w};
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int ga = ba * 16 + 4 * (l >> 4) + i, gb = bb * 16 + (l & 15);
                if (ga < dim && gb < dim) {
                    Gs[ga * dim + gb] = vals[i];
                    Gs[gb * dim + ga] = vals[i];
                }
            }
        }
    }
    for (int e = tid; e < dim; e += T2) u[e] = a.D[(int64_t)k * dim + e];
    __syncthreads();
    // ---- top eigenvector by power iteration: 4 lanes per row -----------------
    const int row = tid >> 2, part = tid & 3;
    const int qlen = (dim + 3) / 4;
    const int q0 = part * qlen, q1 = min(dim, q0 + qlen);
    int it = 0;
    for (;;) {
        float z = 0.f;
        if (row < dim)
            for (int q = q0; q < q1; q++) z = fmaf(Gs[q * dim + row], u[q], z);
        z += __shfl_xor(z, 1);
        z += __shfl_xor(z, 2);
        const bool owner = row < dim && part == 0;
        const float nz = sqrtf(block_sum_t2(owner ? z * z : 0.f, red));
        if (nz == 0.f) break;  // E = 0: keep d_k (its coefficients become 0)
        const float un = owner ? z / nz : 0.f;
        const float du = owner ? un - u[row] : 0.f;
        const float d2 = block_sum_t2(du * du, red);
        if (owner) u[row] = un;
        __syncthreads();
        it++;
        if ((it >= 2 && d2 < a.tol2) || it >= a.maxit) break;
    }
    for (int e = tid; e < dim; e += T2) a.D[(int64_t)k * dim + e] = u[e];
    if (tid == 0) atomicAdd(&a.counters[CNT_EIG_IT], (unsigned long long)it);
    // ---- handoff: atom k's new coefficient for every table entry of atom
    // k+1's columns that names atom k (x = e_r . d_k for rank r < m, else 0),
    // computed with the same wave-level dot as K1's coefficient pass ----------
    if (k + 1 >= a.K) return;
    int c1;
    const int m1 = atom_m(a.uoff, k + 1, a.cap, &c1);
    if (m1 == 0) return;
    const int no = a.kmax > 1 ? a.kmax - 1 : 1;
    const int64_t cb1 = a.moff[k + 1];
    const int nent = m1 * no;
    int *hlist = (int *)(u + dim);      // LDS list of entries to compute
    __shared__ int hcount;
    if (tid == 0) hcount = 0;
    __syncthreads();
    const float *Ek = (k & 1) ? a.E1 : a.E0;
    for (int q = tid; q < nent; q += T2) {
        const int4 o = a.tab_oth[cb1 * no + q];
        if (o.x == k && o.z >= 0) {
            if (o.z < m) {
                const int slot = atomicAdd(&hcount, 1);
                if (slot < HAND_MAX) hlist[slot] = q;
            } else {
                a.handx[cb1 * no + q] = 0.f;
            }
        }
    }
    __syncthreads();
    const int nh = min(hcount, HAND_MAX);
    const int lane = tid & 63, wave = tid >> 6;
    for (int h = wave; h < nh; h += T2 / 64) {
        const int q = hlist[h];
        const int r = a.tab_oth[cb1 * no + q].z;
        const float x = wave_dot(Ek + (int64_t)r * a.dpad, u, dim);
        if (lane == 0) a.handx[cb1 * no + q] = x;
    }
    if (hcount > HAND_MAX && tid == 0) atomicAdd(&a.counters[CNT_DEGEN + 2], 1ull);
}

}  // namespace

int launch_sweep(ksvd_ctx *ctx, float tol)
{
    const int dim = ctx->dim, K = ctx->K, kmax = ctx->kmax;
    const int dpad = (dim + 15) / 16 * 16;
    const int nbt = dpad / 16;
    const int ntiles = nbt * (nbt + 1) / 2;
    if (ntiles > 16 * MAX_TPW) return ksvd_fail(ctx, KSVD_ERR_UNSUPPORTED, "patch dimension too large for the sweep");
    SweepArgs a;
    a.Y = (const float *)ctx->Y.ptr;
    a.D = (float *)ctx->D.ptr;
    a.coef = (float *)ctx->coef.ptr;
    a.uoff = (const int *)ctx->uoff.ptr;
    a.users = (const int *)ctx->users.ptr;
    a.uslot = (const int *)ctx->uslot.ptr;
    a.moff = (const int *)ctx->moff.ptr;
    a.tab_p = (const int *)ctx->tab_p.ptr;
    a.tab_self = (const int *)ctx->tab_self.ptr;
    a.tab_oth = (const int4 *)ctx->tab_oth.ptr;
    a.E0 = (float *)ctx->E.ptr;
    a.E1 = (float *)ctx->E2.ptr;
    a.Gpart = (float *)ctx->gpart.ptr;
    a.handx = (float *)ctx->handx.ptr;
    a.G = (float *)ctx->wbuf.ptr;
    a.counters = (unsigned long long *)ctx->counters.ptr;
    a.dim = dim;
    a.dpad = dpad;
    a.K = K;
    a.kmax = kmax;
    a.cap = ctx->prm.user_cap;
    a.nbt = nbt;
    a.ntiles = ntiles;
    a.maxit = 300;
    a.tol2 = tol * tol;
    const int no = kmax > 1 ? kmax - 1 : 1;
    const size_t lds1 = sizeof(float) * (size_t)CHUNK * dpad;
    (void)no;
    const size_t lds2 = sizeof(float) * ((size_t)dim * dim + dim) + sizeof(int) * HAND_MAX;
    // K1 variants: Gram tiles per wavefront x other support entries per column
    const int tpw = (ntiles + 15) / 16;
    const bool wide = kmax > 5;
    typedef void (*K1Fn)(SweepArgs, int);
    K1Fn k1 = nullptr;
    if (!wide) k1 = tpw <= 1 ? k1_kernel<1, 4, 4> : tpw <= 2 ? k1_kernel<2, 4, 3> : tpw <= 3 ? k1_kernel<3, 4, 3>
                                                                                             : k1_kernel<5, 4, 2>;
    else k1 = tpw <= 1 ? k1_kernel<1, 7, 2> : tpw <= 2 ? k1_kernel<2, 7, 2> : tpw <= 3 ? k1_kernel<3, 7, 2>
                                                                                       : k1_kernel<5, 7, 2>;
    KSVD_HIP_CHECK(ctx, hipFuncSetAttribute((const void *)k1, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds1));
    KSVD_HIP_CHECK(ctx, hipFuncSetAttribute((const void *)k2_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2));
    for (int k = 0; k <= K; k++) {
        hipLaunchKernelGGL(k1, dim3(B1), dim3(T1), lds1, ctx->stream, a, k);
        if (k < K) hipLaunchKernelGGL(k2_kernel, dim3(1), dim3(T2), lds2, ctx->stream, a, k);
    }
    KSVD_HIP_CHECK(ctx, hipGetLastError());
    return KSVD_OK;
}

size_t sweep_gpart_bytes(int dim)
{
    const int nbt = (dim + 15) / 16;
    return sizeof(float) * (size_t)B1 * (nbt * (nbt + 1) / 2) * 256;
}
