// The Gauss-Seidel dictionary sweep (cuda_svd_denoiser.cu:309-361).  Atom k
// depends on atom k-1's new vector and coefficients, so the sweep is a chain
// of 256 short dependent steps; three implementations, chosen in launch_sweep:
//
//  * two-team persistent sweep (pt_sweep_kernel, the default for user caps up
//    to 1024): two teams of 16 workgroups on one XCD take alternate atoms; a
//    team runs its atom's power iteration as all-to-all exchanges of
//    E_h (E_h^T u) vectors while the other team prepares the next atom;
//  * master/helper persistent sweep (ps_sweep_kernel, user caps up to 1984):
//    helpers form partial Gram matrices on MFMA, a master workgroup runs the
//    power iteration on G = E E^T;
//  * two kernels per atom (k1_kernel, k2_kernel: any cap, including none).
//
// The two-kernel form, K1(k), B1 workgroups x 1024 threads:
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
#include <cstdlib>

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
    float tol2;  // stop when |u_i - u_{i-1}|^2 falls below it
    float est2;  // or when the estimated |u_i - limit|^2 does (power_stop)
    int spec;    // eight-team sweep: speculate on the first exchange (0: the plain path for every atom)
};

// Stop rule of every sweep's power iteration, after iterate i with d2 =
// |u_i - u_{i-1}|^2 and d2p = |u_{i-1} - u_{i-2}|^2: never before i = 2; stop
// when the step is below the tolerance (d2 < tol2), or when the a-posteriori
// error estimate is: the steps of a power iteration shrink by r = lambda2 /
// lambda1 of E E^T each time, r ~ sqrt(d2 / d2p), and u_i lies about
// |u_i - u_{i-1}| r / (1 - r) from the limit -- used when r <= 1/2, stop when
// its square is below est2.  Always at maxit.
__device__ __forceinline__ bool power_stop(int i, float d2, float d2p, float tol2, float est2, int maxit)
{
    if (i >= maxit) return true;
    if (i < 2) return false;
    if (d2 < tol2) return true;
    if (!(d2p > 0.f) || d2 > 0.25f * d2p) return false;
    const float r = sqrtf(d2 / d2p);
    const float e = r / (1.f - r);
    return d2 * e * e < est2;
}

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

// DPP lane moves (VALU, no LDS round trip).  Rows whose bit is clear in
// row_mask receive 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_f(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xF, false));
}

// Sum of the 4 lanes of each quad, in every lane of the quad (fixed order).
__device__ __forceinline__ float quad_sum(float v)
{
    v += dpp_f<0xB1, 0xF>(v);  // quad_perm [1,0,3,2]
    v += dpp_f<0x4E, 0xF>(v);  // quad_perm [2,3,0,1]
    return v;
}

// Sum over the 64 lanes of a wavefront in a fixed order, returned to every
// lane: quads, half rows, rows (mirrors), then rows 0+1 and 2+3 and the total
// in lane 63 (row broadcasts), read back as a scalar.
__device__ __forceinline__ float wave_sum_dpp(float v)
{
    v = quad_sum(v);
    v += dpp_f<0x141, 0xF>(v);  // row_half_mirror
    v += dpp_f<0x140, 0xF>(v);  // row_mirror
    v += dpp_f<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_f<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

template <int TPW, int NO, int BU>
__global__ __launch_bounds__(T1) void k1_kernel(SweepArgs a, int k)
{
    extern __shared__ __attribute__((aligned(16))) float sm[];
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

// Top eigenvector of the symmetric dim x dim matrix Gs (LDS) by power
// iteration from u (LDS, overwritten with the result), T2 threads, 4 lanes per
// row.  Stops by power_stop; a zero product keeps u.
// Returns the number of steps.
// Sums of a and of b over the T2 threads of a workgroup (fixed order),
// returned to every thread; red2 holds 2 x (T2 / 64) float2 and `phase`
// alternates between calls so that one barrier suffices.
__device__ __forceinline__ float2 block_sum2(float a, float b, float2 *red2, int &phase, bool act = true)
{
    if (act) {
        a = wave_sum_dpp(a);
        b = wave_sum_dpp(b);
    } else {
        a = b = 0.f;
    }
    float2 *r = red2 + phase * (T2 / 64);
    phase ^= 1;
    if ((threadIdx.x & 63) == 0) r[threadIdx.x >> 6] = make_float2(a, b);
    __syncthreads();
    float2 s = make_float2(0.f, 0.f);
#pragma unroll
    for (int w = 0; w < T2 / 64; w++) {
        const float2 v = r[w];
        s.x += v.x;
        s.y += v.y;
    }
    return s;
}

// Power iteration u <- G u / |G u| from u (LDS) on the symmetric Gs (LDS,
// pitch ldg), T2 threads, 4 per row, each holding its quarter of a row of G
// in registers (quarters of ql = 4 * ceil(dim / 16) entries, ql <= QL; u must
// hold 4 * QL floats, zero past dim; zb is scratch for 2 x 4 * QL floats).
// Stops at the first iterate u_i that power_stop accepts; a zero product
// keeps u_i.  One barrier per pass: pass i
// reads z_i = G u_{i-1} unnormalised and its norm n_i (from the previous
// pass's reduction), forms u_i = z_i / n_i on the fly, writes z_{i+1} =
// G u_i to the other buffer, and one block reduction gives |z_{i+1}|^2 and
// |u_i - u_{i-1}|^2 together.  Returns i; u holds u_i.
template <int QL>
__device__ int power_iteration_ps(const float *Gs, int ldg, float *u, float *zb, int dim, float2 *red2, float tol2,
                                  float est2, int maxit)
{
    const int tid = threadIdx.x;
    const int row = tid >> 2, part = tid & 3;
    const int ql = (dim + 15) / 16 * 4;
    const int q0 = part * ql;
    // only the wavefronts that hold rows compute; the others take part in
    // the barriers with zero partial sums
    const bool act = (tid >> 6) < (4 * dim + 63) / 64;
    float g[QL];
#pragma unroll
    for (int i = 0; i < QL; i++) {
        const int q = q0 + i;
        g[i] = (act && row < dim && i < ql && q < dim) ? Gs[q * ldg + row] : 0.f;
    }
    const bool owner = row < dim && part == 0;
    // the iterate buffers past dim stay 0 (g is 0 there, but 0 * NaN is not);
    // the first pass's barrier orders these stores before any read
    for (int e = dim + tid; e < 4 * QL; e += blockDim.x) {
        zb[e] = 0.f;
        zb[4 * QL + e] = 0.f;
    }
    const float *src = u;  // u_i = src * scale
    float scale = 1.f, uprev = 0.f, ui = 0.f, d2p = 0.f;
    int phase = 0;
    int it = 0;
    for (;;) {
        // terms past the thread's range have g = 0 and add exactly nothing;
        // four partial sums shorten the dependent chain
        float z = 0.f;
        if (act) {
            float z4[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int c = 0; c < QL / 4; c++) {
                const f4 uu = *(const f4 *)(src + q0 + 4 * c);
                z4[0] = fmaf(g[4 * c], uu.x, z4[0]);
                z4[1] = fmaf(g[4 * c + 1], uu.y, z4[1]);
                z4[2] = fmaf(g[4 * c + 2], uu.z, z4[2]);
                z4[3] = fmaf(g[4 * c + 3], uu.w, z4[3]);
            }
            z = quad_sum((z4[0] + z4[1]) + (z4[2] + z4[3])) * scale;
        }
        ui = owner ? src[row] * scale : 0.f;
        const float du = ui - uprev;
        float *dst = zb + ((it + 1) & 1) * 4 * QL;
        if (owner) dst[row] = z;
        const float2 sz = block_sum2(owner ? z * z : 0.f, owner ? du * du : 0.f, red2, phase, act);
        if (power_stop(it, sz.y, d2p, tol2, est2, maxit)) break;
        d2p = sz.y;
        const float nz = sqrtf(sz.x);
        if (nz == 0.f) break;  // E = 0: keep u_i (d_k at the start; its coefficients become 0)
        uprev = ui;
        src = dst;
        scale = 1.f / nz;
        it++;
    }
    if (owner) u[row] = ui;
    __syncthreads();
    return it;
}

__device__ int power_iteration(const float *Gs, float *u, int dim, float *red, float tol2, float est2, int maxit)
{
    const int tid = threadIdx.x;
    const int row = tid >> 2, part = tid & 3;
    const int qlen = (dim + 3) / 4;
    const int q0 = part * qlen, q1 = min(dim, q0 + qlen);
    int it = 0;
    float d2p = 0.f;
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
        if (power_stop(it, d2, d2p, tol2, est2, maxit)) break;
        d2p = d2;
    }
    return it;
}

__global__ __launch_bounds__(T2) void k2_kernel(SweepArgs a, int k)
{
    extern __shared__ __attribute__((aligned(16))) float sm[];
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
    const int it = power_iteration(Gs, u, dim, red, a.tol2, a.est2, a.maxit);
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
    if (hcount > HAND_MAX && tid == 0) atomicAdd(&a.counters[CNT_HAND_OVF], 1ull);
}

// ===========================================================================
// Persistent sweep on one XCD: the whole Gauss-Seidel sweep in one launch,
// used when every atom's capped column count fits the helpers' LDS slices
// (cap <= H * PS_CWMAX).
//
// The workgroups read the XCD they run on (hardware register XCC_ID); the
// first workgroup to arrive chooses its XCD, the first H + 1 to arrive there
// take roles, one master and H helpers, and the others leave at once.  Every
// role that runs to its end counts itself, so the host also finds a sweep in
// which roles never ran.  All parties then share that XCD's L2, so data
// passes between them without an L2 write-back: the writer stores normally,
// waits for its stores (vmcnt(0)) and sets a flag with an L1-bypassing
// (sc1) store; the reader polls the flag and reads the data with sc1 loads.
// Data that nobody changes during the sweep (patches, column table, users)
// is read with ordinary loads.
//
// Phase k (atom k); helper h owns columns [h*cw, (h+1)*cw) of atom k's m
// capped users, cw = ceil(m / H), and keeps them in LDS buffer k & 1:
//   P1 (before atom k-1 is final) builds its columns of E_k from every
//      support entry except atom k-1's, once every helper has written atom
//      k-2's final coefficients (cdone = k);
//   W  waits until the master has published d_{k-1} (aflag = k);
//   P2 for each column whose patch is among atom k-1's first m_{k-1} users
//      (rank r), x = e_{k-1,r} . d_{k-1} (column r from global memory) and
//      e -= d_{k-1} x;
//   P3 partial Gram of its columns on MFMA -> Gpart[h]; done[h] = k + 1;
//      then its E_k columns -> global (atom k+1's hand-over reads them after
//      aflag = k + 1, which follows every rdone[*] = k + 1 and the store
//      wait that precedes each rdone);
//   RS once every done[*] = k + 1: sums the partials of the Gram tiles
//      h, h + H, ... in helper order -> Gred; rdone[h] = k + 1;
//   P4 atom k-1's final coefficients for its own columns of atom k-1 (its
//      other LDS buffer) -- the same arithmetic on the same values as P2, so
//      both agree bit for bit -- and 0 for its share of the users beyond the
//      cap; cdone[h] = k + 1.
// Master, phase k: once every rdone[*] = k + 1, reads Gred into LDS, runs the
// power iteration from d_k, writes d_k and sets aflag = k + 1.  After phase
// K - 1 the helpers run P4 for atom K - 1.
// Every wait is bounded: one that runs out sets an abort word (and a sticky
// copy the host reads) which every other wait also polls, so all workgroups
// leave; the host then reports the sweep as failed.
constexpr int PS_T = T2;                // threads per workgroup (the power iteration's block size)
constexpr int PS_W = PS_T / 64;         // wavefronts per workgroup
constexpr int PS_HMAX = 31;             // helpers at most (one master + 31 helpers fill one XCD's 32 CUs)
constexpr int PS_CWMAX = 64;            // columns per helper slice
constexpr int PS_CPW = PS_CWMAX / PS_W; // columns per wavefront in the dot passes
constexpr int PS_NOMAX = 7;             // other support entries per column (ompIter <= 8)
constexpr long PS_SPIN_MAX = 1L << 21;  // polls before a wait gives up (default)
// flag words: per-XCD registration; per-helper done, rdone, cdone; the
// master's aflag; the abort word and its sticky copy
enum {
    PSF_REG = 0,
    PSF_DONE = 32,
    PSF_RDONE = 64,
    PSF_CDONE = 96,
    PSF_AFLAG = 128,
    PSF_ABORT = 129,
    PSF_STICKY = 130,
    PSF_XSEL = 131,      // the XCD the sweep runs on, plus 1 (0: none chosen yet)
    PSF_STARTED = 132,   // roles launched, summed over sweeps (kept across resets)
    PSF_FINISHED = 133,  // roles that completed, summed over sweeps (kept across resets)
    PSF_SPIN = 134,      // polls before a wait gives up (set by the reset kernel)
    PSF_WORDS = 160
};

// The first workgroup to arrive chooses its own XCD for the sweep; the others
// take a role (slot < nroles) only on that XCD.  Returns the slot, or -1.
__device__ __forceinline__ int ps_register(unsigned *flags, int nroles)
{
    const unsigned xcc = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15u;  // XCC_ID[3:0]
    const unsigned old = atomicCAS(&flags[PSF_XSEL], 0u, xcc + 1u);
    if ((old == 0u ? xcc + 1u : old) != xcc + 1u) return -1;
    const unsigned slot = atomicAdd(&flags[PSF_REG + xcc], 1u);
    return slot < (unsigned)nroles ? (int)slot : -1;
}

__device__ __forceinline__ unsigned ld_flag(const unsigned *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The polls a wait may spend before it gives up (PS_SPIN_MAX unless a test
// asks for fewer).
__device__ __forceinline__ long spin_limit(const unsigned *flags) { return (long)ld_flag(flags + PSF_SPIN); }

// Flag and exchange stores are workgroup-scope atomics: single stores that
// reach the XCD's L2 because gfx950's L1 is write-through, where the
// persistent sweeps' workgroups, all on the sweep's XCD, read them with
// L1-bypassing loads.  Readers in other workgroups see them through that
// write-through L1, which the HIP memory model does not promise for
// workgroup scope: this relies on the hardware.  Agent scope would also write
// them through to memory for the other XCDs, which none of the readers is on;
// that made each pair store's round trip twice as long and the 2048 x 2048
// update 10 % slower (2.90 -> 2.62 ms).
__device__ __forceinline__ void st_flag(unsigned *p, unsigned v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// L1-bypassing (sc1) load of a value another workgroup wrote
__device__ __forceinline__ float ld_c(const float *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// sc1 16-byte load, `off` bytes into a buffer of `bytes` bytes (0 past its end)
__device__ __forceinline__ f4 ld4_c(const float *base, unsigned bytes, unsigned off)
{
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(base), 0, bytes, 0x00020000);
    return __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 16));
}

// Drops this CU's L1 lines, so that plain loads after it read what other
// workgroups stored to the XCD's L2 before the flags this workgroup waited
// for.  The invalidation is agent-scope (sc1): a workgroup-scope one (sc0)
// leaves lines that other CUs' stores have made stale in place, since a
// workgroup's own waves share one L1, and plain loads after it then read
// stale atoms and coefficients -- results that changed with the buffers'
// addresses from one context to the next.  Wavefront 0 issues the
// invalidation; the barrier orders every wavefront's later loads after it.
// (The persistent sweep reads the atoms and coefficients of the other atoms'
// shares this way: with L1-bypassing sc1 loads instead, the 2048 x 2048
// sweep ran 20 % slower in about one context in four, depending on where the
// dictionary and coefficient buffers sat in memory.)
__device__ __forceinline__ void l1_drop()
{
    if ((threadIdx.x >> 6) == 0) asm volatile("buffer_inv sc1\n\ts_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
}

// This workgroup's global stores reach the XCD's L2, then *flag = v.
__device__ __forceinline__ void ps_publish(unsigned *flag, unsigned v)
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0), expcnt and lgkmcnt untouched
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __syncthreads();
    if (threadIdx.x == 0) st_flag(flag, v);
}

// Threads 0 .. count-1 each wait for flags[first + t] >= v.  Returns false
// (for the whole workgroup) after an abort or a wait that ran out.  Up to
// three groups of flags at once: thread t < n0 waits for flags[f0 + t] >= v0,
// the next n1 threads for flags[f1 + ..] >= v1, the next n2 for f2, v2.
__device__ bool ps_wait3(unsigned *flags, int f0, int n0, unsigned v0, int f1, int n1, unsigned v1, int f2, int n2,
                         unsigned v2, int *bail)
{
    const int t = (int)threadIdx.x;
    // bail[1] holds the launch's spin limit; 0 (tests only) gives up at the
    // first wait, so that an abort is certain
    if (bail[1] == 0 && t == 0) {
        st_flag(flags + PSF_STICKY, 1u);
        st_flag(flags + PSF_ABORT, 1u);
        *bail = 1;
    }
    if (t < n0 + n1 + n2 && bail[1] != 0) {
        const unsigned *f = t < n0 ? flags + f0 + t : t < n0 + n1 ? flags + f1 + (t - n0) : flags + f2 + (t - n0 - n1);
        const unsigned v = t < n0 ? v0 : t < n0 + n1 ? v1 : v2;
        long spins = 0;
        while (ld_flag(f) < v) {
            __builtin_amdgcn_s_sleep(2);
            if ((++spins & 127) == 0) {
                if (ld_flag(flags + PSF_ABORT)) {
                    *bail = 1;
                    break;
                }
                if (spins > spin_limit(flags)) {
                    st_flag(flags + PSF_STICKY, 1u);
                    st_flag(flags + PSF_ABORT, 1u);
                    *bail = 1;
                    break;
                }
            }
        }
    }
    __syncthreads();
    const bool ok = *bail == 0;
    __syncthreads();
    return ok;
}

__device__ __forceinline__ bool ps_wait(unsigned *flags, int first, int count, unsigned v, int *bail)
{
    return ps_wait3(flags, first, count, v, 0, 0, 0u, 0, 0, 0u, bail);
}

// v + (v at lane ^ 16) and v + (v at lane ^ 32) on the VALU: the permlane
// swaps of a value with itself return, in their two results, every lane's
// value and its partner's (rows 0 and 2 against rows 1 and 3; the lower half
// against the upper), so the sum is the same in both lanes of a pair
__device__ __forceinline__ float add_xor16(float v)
{
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

__device__ __forceinline__ float add_xor32(float v)
{
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// Sums over the 64 lanes of four values a[0..3] at once, returned to every
// lane in t[0..3] (fixed order): exchanges over lane bits 0 and 1 leave each
// lane one column's quad sum (column 2 (lane & 1) + ((lane >> 1) & 1)), then
// rotations by 8 and 4 within rows and exchanges over lane bits 4 and 5 sum
// the 16 lanes of that column, read back from lanes 0, 2, 1, 3.
__device__ __forceinline__ void reduce4_cols(const float *a, float *t)
{
    const int lane = threadIdx.x & 63;
    const bool odd = (lane & 1) != 0, hi = (lane & 2) != 0;
    const float x = (odd ? a[2] : a[0]) + dpp_f<0xB1, 0xF>(odd ? a[0] : a[2]);
    const float y = (odd ? a[3] : a[1]) + dpp_f<0xB1, 0xF>(odd ? a[1] : a[3]);
    float v = (hi ? y : x) + dpp_f<0x4E, 0xF>(hi ? x : y);
    v += dpp_f<0x128, 0xF>(v);  // row_ror:8
    v += dpp_f<0x124, 0xF>(v);  // row_ror:4
    v = add_xor32(add_xor16(v));
    t[0] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    t[1] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 2));
    t[2] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 1));
    t[3] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 3));
}

// e . u for one column held by a wavefront as ev[j] = e[lane + 64j], with u
// held the same way (0 past dim): lane partials in j order with fmaf, then
// wave_sum_dpp.  The hand-over (P2) and the stored coefficient (P4) both use
// it on the same values, so they agree bit for bit.
__device__ __forceinline__ float ps_dot(const float *ev, const float *uv, int dim)
{
    const int lane = threadIdx.x & 63;
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < 3; j++)
        if (lane + 64 * j < dim) s = fmaf(ev[j], uv[j], s);
    return wave_sum_dpp(s);
}

// P2 for the hand-over list hl[0 .. nh): wavefront w takes entries w + PS_W*g,
// g < CPW.  Column tl = hl[i] of E_k (LDS) loses d_{k-1} x, x = e . d_{k-1}
// with e column hrank[tl] of E_{k-1} (global; all loads issued first).
template <int CPW>
__device__ __forceinline__ void ps_handover(float *Es, int dpad, int dim, const int *hl, const int *hrank, int nh,
                                            const float *Eprev, const float *uv)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float ev[CPW][3];
    int tls[CPW];
#pragma unroll
    for (int g = 0; g < CPW; g++) {
        const int i = wave + PS_W * g;
        const int tl = i < nh ? hl[i] : -1;
        tls[g] = tl;
        const float *col = Eprev + (int64_t)(tl >= 0 ? hrank[tl] : 0) * dpad;
#pragma unroll
        for (int j = 0; j < 3; j++) ev[g][j] = ld_c(col + min(lane + 64 * j, dim - 1));
    }
#pragma unroll
    for (int g = 0; g < CPW; g++) {
        const int tl = tls[g];
        if (tl < 0) continue;
        const float x = ps_dot(ev[g], uv, dim);
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int e = lane + 64 * j;
            if (e < dim) Es[tl * dpad + e] = fmaf(-uv[j], x, Es[tl * dpad + e]);
        }
    }
}

template <int QL>
__device__ void ps_master(const SweepArgs &a, unsigned *flags, int H, const float *Gred, float *sm, int *bail)
{
    __shared__ float2 red2[2 * (PS_T / 64)];
    const int tid = threadIdx.x, dim = a.dim, dpad = a.dpad;
    float *Gs = sm;               // dpad x dpad (rows and columns past dim are 0)
    float *u = Gs + dpad * dpad;  // 4 * QL, zero past dim
    float *zb = u + 4 * QL;       // 2 x 4 * QL: the power iteration's unnormalised iterates
    const int nf4 = a.ntiles * 64;
    const unsigned rbytes = (unsigned)(nf4 * 16);
    unsigned long long iters = 0;
    for (int k = 0; k < a.K; k++) {
        int c;
        const int m = atom_m(a.uoff, k, a.cap, &c);
        if (m > 0) {
            if (!ps_wait(flags, PSF_RDONE, H, (unsigned)(k + 1), bail)) return;
            // the reduced upper tiles, both triangles; tiles cover dpad x dpad
            // and the padding rows of E are 0, so the padding of Gs is 0
            for (int f0 = tid; f0 < nf4; f0 += 4 * PS_T) {
                f4 v[4];
#pragma unroll
                for (int i = 0; i < 4; i++) v[i] = ld4_c(Gred, rbytes, (unsigned)((f0 + PS_T * i) * 16));
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int f = f0 + PS_T * i;
                    if (f >= nf4) continue;
                    const int ti = f >> 6, l = f & 63;
                    int ba, bb;
                    tile_coords(ti, a.nbt, &ba, &bb);
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const int ga = ba * 16 + 4 * (l >> 4) + j, gb = bb * 16 + (l & 15);
                        Gs[ga * dpad + gb] = v[i][j];
                        Gs[gb * dpad + ga] = v[i][j];
                    }
                }
            }
            for (int e = tid; e < 4 * QL; e += PS_T) u[e] = e < dim ? ld_c(a.D + (int64_t)k * dim + e) : 0.f;
            __syncthreads();
            iters += (unsigned long long)power_iteration_ps<QL>(Gs, dpad, u, zb, dim, red2, a.tol2, a.est2, a.maxit);
            for (int e = tid; e < dim; e += PS_T) a.D[(int64_t)k * dim + e] = u[e];
        }
        ps_publish(flags + PSF_AFLAG, (unsigned)(k + 1));
    }
    if (tid == 0) atomicAdd(&a.counters[CNT_EIG_IT], iters);
}

template <int TPW, int NO, int BU>
__device__ void ps_helper(const SweepArgs &a, unsigned *flags, int H, int h, float *Gred, float *sm, int *bail)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int dim = a.dim, dpad = a.dpad, kmax = a.kmax, K = a.K;
    const int no = kmax > 1 ? kmax - 1 : 1;
    const int nf4 = a.ntiles * 64;
    const unsigned gbytes = (unsigned)(H * nf4 * 16);
    float *Esb = sm;                                   // 2 x PS_CWMAX x dpad: columns of E_k in buffer k & 1
    float *sx = Esb + 2 * PS_CWMAX * dpad;             // PS_CWMAX x PS_NOMAX coefficients
    int *sat = (int *)(sx + PS_CWMAX * PS_NOMAX);      // PS_CWMAX x PS_NOMAX atoms (-1: none)
    int *sp = sat + PS_CWMAX * PS_NOMAX;               // PS_CWMAX patches
    int *hrank = sp + PS_CWMAX;                        // PS_CWMAX: rank among atom k-1's users, -1: none
    int *hl = hrank + PS_CWMAX;                        // PS_CWMAX: columns with a hand-over (any order)
    int *nhp = hl + PS_CWMAX;                          // their number
    for (int k = 0; k <= K; k++) {
        float *Es = Esb + (k & 1) * PS_CWMAX * dpad;
        const float *Eo = Esb + ((k + 1) & 1) * PS_CWMAX * dpad;  // this helper's columns of E_{k-1}
        // ---- P1: E_k without atom k-1 ---------------------------------------
        int m = 0, nc = 0, tb = 0;
        int cprev = 0, mprev = 0;
        if (k >= 1) mprev = atom_m(a.uoff, k - 1, a.cap, &cprev);
        if (tid == 0) *nhp = 0;
        if (k < K) {
            int c;
            m = atom_m(a.uoff, k, a.cap, &c);
            if (k >= 1 && !ps_wait(flags, PSF_CDONE, H, (unsigned)k, bail)) return;
            if (m > 0) {
                const int cw = (m + H - 1) / H;
                tb = h * cw;
                nc = max(0, min(m, tb + cw) - tb);
            }
            if (nc > 0) {
                const int64_t cb = a.moff[k] + tb;
                for (int tl = tid; tl < nc; tl += PS_T) {
                    const int64_t col = cb + tl;
                    sp[tl] = a.tab_p[col];
                    // all loads unconditional (entries past `no` reread the
                    // last one and are then dropped), so none waits on a branch
                    int4 o[NO];
#pragma unroll
                    for (int i = 0; i < NO; i++) o[i] = a.tab_oth[col * no + min(i, no - 1)];
                    float xv[NO];
#pragma unroll
                    for (int i = 0; i < NO; i++) xv[i] = ld_c(a.coef + (o[i].x >= 0 ? o[i].y : 0));
                    int hr = -1;
#pragma unroll
                    for (int i = 0; i < NO; i++) {
                        int at = i < no ? o[i].x : -1;
                        if (at == k - 1) {
                            if (o[i].z >= 0 && o[i].z < mprev) hr = o[i].z;
                            at = -1;
                        }
                        sat[tl * PS_NOMAX + i] = at;
                        sx[tl * PS_NOMAX + i] = at >= 0 ? xv[i] : 0.f;
                    }
                    hrank[tl] = hr;
                    if (hr >= 0) hl[atomicAdd(nhp, 1)] = tl;
                }
                const int nc4 = (nc + 3) & ~3;
                for (int q = tid; q < nc4 * dpad; q += PS_T) {
                    const int tl = q / dpad, e = q - tl * dpad;
                    if (tl >= nc || e >= dim) Es[q] = 0.f;
                }
                __syncthreads();
                const int cpt = PS_T / dim;
                const int tc = tid / dim, te = tid - tc * dim;
                if (tc < cpt) {
                    for (int t = tc; t < nc; t += BU * cpt) {
                        float v[BU], dv[BU][NO];
                        // unconditional loads: a column past nc rereads column
                        // nc-1, an empty entry reads atom 0 (its coefficient
                        // is 0, so it adds exactly nothing)
#pragma unroll
                        for (int b = 0; b < BU; b++) {
                            const int tt = min(t + b * cpt, nc - 1);
                            v[b] = a.Y[(int64_t)sp[tt] * dim + te];
#pragma unroll
                            for (int i = 0; i < NO; i++)
                                dv[b][i] = ld_c(a.D + (int64_t)max(sat[tt * PS_NOMAX + i], 0) * dim + te);
                        }
#pragma unroll
                        for (int b = 0; b < BU; b++) {
                            const int tt = t + b * cpt;
                            if (tt < nc) {
#pragma unroll
                                for (int i = 0; i < NO; i++) v[b] = fmaf(-dv[b][i], sx[tt * PS_NOMAX + i], v[b]);
                                Es[tt * dpad + te] = v[b];
                            }
                        }
                    }
                }
            }
        }
        // ---- W: d_{k-1} -------------------------------------------------------
        // every wavefront keeps d_{k-1} in registers: uv[j] = d_{k-1}[lane + 64j]
        float uv[3] = {0.f, 0.f, 0.f};
        if (k >= 1) {
            if (!ps_wait(flags, PSF_AFLAG, 1, (unsigned)k, bail)) return;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const float v = ld_c(a.D + (int64_t)(k - 1) * dim + min(lane + 64 * j, dim - 1));
                uv[j] = lane + 64 * j < dim ? v : 0.f;
            }
        }
        // ---- P2: atom k-1's share of E_k --------------------------------------
        if (k >= 1 && nc > 0) {
            const float *Eprev = ((k - 1) & 1) ? a.E1 : a.E0;
            const int nh = *nhp;
            if (nh > 2 * PS_W) ps_handover<4>(Es, dpad, dim, hl, hrank, nh, Eprev, uv);
            else if (nh > PS_W) ps_handover<2>(Es, dpad, dim, hl, hrank, nh, Eprev, uv);
            else if (nh > 0) ps_handover<1>(Es, dpad, dim, hl, hrank, nh, Eprev, uv);
        }
        __syncthreads();
        // ---- P3: partial Gram of this helper's columns --------------------------
        if (k < K) {
            if (nc > 0) {
                f4 acc[TPW];
#pragma unroll
                for (int q = 0; q < TPW; q++) acc[q] = f4{0.f, 0.f, 0.f, 0.f};
                const int col16 = lane & 15, grp = lane >> 4;
                const int nk = (nc + 3) >> 2;
                // the wavefront's tiles advance together: UK k-steps of
                // operands are read first, then the MFMA chains of the tiles
                // interleave (independent accumulators)
                int oa[TPW], ob[TPW];
                bool tok[TPW];
#pragma unroll
                for (int q = 0; q < TPW; q++) {
                    const int ti = wave + PS_W * q;
                    tok[q] = ti < a.ntiles;
                    int ba = 0, bb = 0;
                    if (tok[q]) tile_coords(ti, a.nbt, &ba, &bb);
                    oa[q] = ba * 16 + col16;
                    ob[q] = bb * 16 + col16;
                }
                const float *Eg = Es + grp * dpad;
                constexpr int UK = TPW <= 3 ? 4 : 2;  // k-steps per batch of reads (register budget)
                for (int kk = 0; kk < nk; kk += UK) {
                    float av[UK][TPW], bv[UK][TPW];
#pragma unroll
                    for (int u = 0; u < UK; u++) {
                        // columns past nc4 hold stale values: kept out
                        const bool in = kk + u < nk;
                        const float *row = Eg + 4 * (kk + u) * dpad;
#pragma unroll
                        for (int q = 0; q < TPW; q++) {
                            av[u][q] = in ? row[oa[q]] : 0.f;
                            bv[u][q] = in ? row[ob[q]] : 0.f;
                        }
                    }
#pragma unroll
                    for (int u = 0; u < UK; u++)
#pragma unroll
                        for (int q = 0; q < TPW; q++)
                            if (tok[q]) acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[u][q], bv[u][q], acc[q], 0, 0, 0);
                }
                float *gp = a.Gpart + (size_t)h * nf4 * 4;
#pragma unroll
                for (int q = 0; q < TPW; q++) {
                    const int ti = wave + PS_W * q;
                    if (ti < a.ntiles) *(f4 *)(gp + (size_t)ti * 256 + lane * 4) = acc[q];
                }
            }
            ps_publish(flags + PSF_DONE + h, (unsigned)(k + 1));
            // E_k's columns to global memory for atom k+1's hand-over: issued
            // after the flag (they are read only after atom k+1's vector is
            // out, which follows this helper's next publish and its wait)
            if (nc > 0) {
                float *Ecur = (k & 1) ? a.E1 : a.E0;
                for (int q = tid; q < nc * dim; q += PS_T) {
                    const int tl = q / dim, e = q - tl * dim;
                    Ecur[(int64_t)(tb + tl) * dpad + e] = Es[tl * dpad + e];
                }
            }
            // ---- RS: this helper's tiles of the Gram, summed over the partials
            if (m > 0) {
                if (!ps_wait(flags, PSF_DONE, H, (unsigned)(k + 1), bail)) return;
                const int cw = (m + H - 1) / H;
                const int nb = (m + cw - 1) / cw;
                for (int ti = h + H * wave; ti < a.ntiles; ti += H * PS_W) {
                    const int f = ti * 64 + lane;
                    f4 sum = f4{0.f, 0.f, 0.f, 0.f};
                    for (int b0 = 0; b0 < nb; b0 += 8) {
                        f4 pv[8];
#pragma unroll
                        for (int b = 0; b < 8; b++) pv[b] = ld4_c(a.Gpart, gbytes, (unsigned)(((b0 + b) * nf4 + f) * 16));
#pragma unroll
                        for (int b = 0; b < 8; b++)
                            if (b0 + b < nb) sum += pv[b];
                    }
                    *(f4 *)(Gred + (size_t)f * 4) = sum;
                }
            }
            ps_publish(flags + PSF_RDONE + h, (unsigned)(k + 1));
        }
        // ---- P4: atom k-1's final coefficients, from this helper's LDS copy ---
        if (k >= 1 && cprev > 0) {
            const int cwp = (mprev + H - 1) / H;
            const int tbp = h * cwp;
            const int ncp = max(0, min(mprev, tbp + cwp) - tbp);
            const int64_t cbp = a.moff[k - 1] + tbp;
#pragma unroll
            for (int g = 0; g < PS_CPW; g++) {
                const int tl = wave + PS_W * g;
                if (tl >= ncp) continue;
                float ev[3];
#pragma unroll
                for (int jj = 0; jj < 3; jj++) ev[jj] = Eo[tl * dpad + min(lane + 64 * jj, dim - 1)];
                const float x = ps_dot(ev, uv, dim);
                if (lane == 0) a.coef[a.tab_self[cbp + tl]] = x;
            }
            const int basep = a.uoff[k - 1];
            for (int t = mprev + h * PS_T + tid; t < cprev; t += H * PS_T) {
                const int pos = basep + t;
                a.coef[(int64_t)a.users[pos] * kmax + a.uslot[pos]] = 0.f;
            }
        }
        if (k < K) ps_publish(flags + PSF_CDONE + h, (unsigned)(k + 1));
    }
}

template <int TPW, int NO, int BU>
__global__ __launch_bounds__(PS_T) void ps_sweep_kernel(SweepArgs a, unsigned *flags, int H, float *Gred)
{
    extern __shared__ __attribute__((aligned(16))) float sm[];
    __shared__ int s_role, s_bail[2];  // bail word, spin limit
    if (threadIdx.x == 0) {
        s_role = ps_register(flags, H + 1);
        s_bail[0] = 0;
        s_bail[1] = (int)ld_flag(flags + PSF_SPIN);
    }
    __syncthreads();
    const int role = s_role;
    if (role < 0) return;
    // Gram entries per master thread: 4 * QL >= dpad (TPW <= 3: dpad <= 144)
    if (role == 0) ps_master<TPW <= 3 ? 36 : 48>(a, flags, H, Gred, sm, s_bail);
    else ps_helper<TPW, NO, BU>(a, flags, H, role - 1, Gred, sm, s_bail);
    // a role that ran to its end (a bail makes every thread leave early)
    if (threadIdx.x == 0 && s_bail[0] == 0) atomicAdd(&flags[PSF_FINISHED], 1u);
}

// ===========================================================================
// Two-team persistent sweep: no master and no Gram matrix.  2 x PT_HT
// workgroups of PT_T threads on one XCD form two teams; team t takes the atoms
// k = t, t + 2, ...  The PT_HT helpers of a team hold atom k's capped columns
// (helper h: [h*cw, (h+1)*cw), cw = ceil(m / PT_HT)) in LDS and run the power
// iteration together.  Exchange i: each forms w_h = E_h (E_h^T u_{i-1}) from
// its own columns and stores it as (sequence number, value) pairs, one 64-bit
// store each; every team member reads the team's PT_HT vectors, rereading a
// pair until it carries the exchange's number (no flag and no store wait on
// this path), and sums them in a fixed order: z_i = E_k E_k^T u_{i-1},
// u_i = z_i / |z_i|.  All members run the same code on the same values, so
// they hold the same u_i and take the same stop decision bit for bit.  The
// pairs alternate between two buffers per team by sequence parity: a member
// reads exchange s's pairs before it stores its pairs of s + 1, so nobody
// stores s + 2 while a member may still read s.  The buffers are cleared
// before each launch, so no pair from an earlier sweep carries a current
// number.
//
// While one team iterates on atom k, the other prepares its next atom k+1:
// the final coefficients of its atom k-1, then atom k+1's columns without
// atom k (atom k's share is handed over when d_k is out).  Atom k, team t,
// helper h:
//   start  three thread groups at once: team 1-t's d_{k-1} pairs (tagged
//       k); the columns of E_{k-1} that the hand-over needs, each read once
//       the team 1-t member that holds it has certified it (ecert >= k);
//       atom k+2's column table and starting vector (static until that
//       atom's own update) into the other of two table sets;
//   P2  for each column whose patch is among atom k-1's first m_{k-1} users
//       (rank r): x = e_{k-1,r} . d_{k-1} and e -= d_{k-1} x;
//   PI  exchanges i = 1, 2, ... until power_stop accepts u_i, or z_i = 0
//       (then u_{i-1} stays).  After exchange 1's store: the columns of E_k
//       that atom k+1's hand-over reads, to global memory; ecert[h] = k + 1 at
//       exchange 2's barrier (after a store wait, which by then finds the
//       columns out), or after the iteration when it took one exchange.
//       PT_PW wavefronts form w_h; every wavefront takes part in the
//       exchange's reads and forms u_i and the stop decision;
//   out helper 0 stores d_k as pairs tagged k + 1 for team 1-t (no store
//       wait on this path), then into the dictionary, and sets done[t] =
//       k + 1 (after a store wait);
//   P4  atom k's final coefficients for its columns (x = e . d_k, the same
//       arithmetic on the same values as team 1-t's P2, so both agree bit for
//       bit); ccert[h] = k + 1 (after a store wait);
//   P1  once every ccert of team t is k + 1, every ccert of team 1-t is k and
//       done[t] = k + 1: atom k+2's patches, and every other atom's share
//       subtracted (atoms <= k final, the others as they were).
// Users beyond the cap get coefficient 0 after the sweep (pa_zero_kernel);
// until then the column tables drop their entries.
constexpr int PT_T = 1024;               // threads per helper workgroup
constexpr int PT_W = PT_T / 64;          // wavefronts per helper
constexpr int PT_PW = 8;                 // wavefronts that run the power iteration's arithmetic
constexpr int PT_HT = 16;                // helpers per team (the two teams fill one XCD's 32 CUs)
constexpr int PT_GRP = 4;                // groups reading the exchange, PT_HT / PT_GRP pairs each (4 groups: 3.0 ms per
                                         // update at 2048x2048; 7 groups: 3.3 ms, 1 group: 5.1 ms)
constexpr int PT_XH = PT_HT / PT_GRP;    // exchange pairs read per thread
// flag words: per-XCD registration; per-helper ecert and ccert (index
// t * PT_HT + h); per-team done; the abort word and its sticky copy
constexpr int PTF_ECERT = 32, PTF_CCERT = 64, PTF_DONE = 96;

__device__ __forceinline__ unsigned long long ld64_c(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st64_c(unsigned long long *p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ int cols_of(int m, int H, int h, int *tb)
{
    const int cw = (m + H - 1) / H;
    *tb = h * cw;
    return max(0, min(m, *tb + cw) - *tb);
}

// Capped user count of atom k from the LDS copy of uoff.
__device__ __forceinline__ int pa_m(const int *suo, int k, int cap)
{
    const int c = suo[k + 1] - suo[k];
    return c > cap ? cap : c;
}

// Column table of this helper's columns [tb, tb + nc) of atom kb: the patch
// (sp), the column's own coefficient slot (ss), every support entry except
// atom kb-1's whose coefficient is not 0 at this point of the sweep, as atom
// and coefficient index (sat, sci; -1 for none), and the rank of the patch
// among atom kb-1's first mx users where it is one (hrank, listed in hl in any
// order).  An entry of an atom j < kb-1 (updated already) whose patch is not
// among atom j's first m_j users has coefficient 0 from atom j's update on;
// its coefficient word is cleared only after the sweep, so it is dropped here.
// sx: 1 for a column whose patch is also among atom kb+1's first m_{kb+1}
// users (the only columns of E_kb that atom kb+1's hand-over reads).
// Threads t0 .. t0 + tn - 1 of the workgroup build it.
template <int NO>
__device__ void pa_table(const SweepArgs &a, const int *suo, const int *smo, int kb, int tb, int nc, int mx, int *sp,
                         int *ss, int *sat, int *sci, int *hrank, int *hl, int *nhp, int *sx, int t0, int tn)
{
    const int no = a.kmax > 1 ? a.kmax - 1 : 1;
    const int mnext = kb + 1 < a.K ? pa_m(suo, kb + 1, a.cap) : 0;
    const int64_t cb = (int64_t)smo[kb] + tb;
    for (int tl = (int)threadIdx.x - t0; tl >= 0 && tl < nc; tl += tn) {
        const int64_t col = cb + tl;
        sp[tl] = a.tab_p[col];
        ss[tl] = a.tab_self[col];
        // all loads unconditional (entries past `no` reread the last one and
        // are then dropped), so none waits on a branch
        int4 o[NO];
#pragma unroll
        for (int i = 0; i < NO; i++) o[i] = a.tab_oth[col * no + min(i, no - 1)];
        int hr = -1, ex = 0;
#pragma unroll
        for (int i = 0; i < NO; i++) {
            int at = i < no ? o[i].x : -1;
            const int rk = o[i].z;
            if (at == kb + 1 && rk >= 0 && rk < mnext) ex = 1;
            if (at >= 0 && at == kb - 1) {
                if (rk >= 0 && rk < mx) hr = rk;
                at = -1;
            } else if (at >= 0 && (rk < 0 || (at < kb - 1 && rk >= pa_m(suo, at, a.cap)))) {
                at = -1;
            }
            sat[tl * PS_NOMAX + i] = at;
            sci[tl * PS_NOMAX + i] = at >= 0 ? o[i].y : -1;
        }
        hrank[tl] = hr;
        sx[tl] = ex;
        if (hr >= 0) hl[atomicAdd(nhp, 1)] = tl;
    }
}

// The columns e_t = y_p - sum_i d_{sat_i} x_i (x_i read at sci_i) of the
// table into Es (pitch dpad), element-parallel: thread (tc, te) takes element
// te of BU columns at once, with every load in flight before the first use.
// The atoms and coefficients are read with L1-bypassing loads (SC1), or with
// plain loads after the caller has dropped its L1 lines (l1_drop) behind the
// waits that certify them.
template <int NO, int BU, bool SC1 = true>
__device__ void pa_build(const SweepArgs &a, int nc, float *Es, const int *sp, const int *sat, const int *sci)
{
    const int tid = threadIdx.x, dim = a.dim, dpad = a.dpad;
    const int cpt = PT_T / dim;
    const int tc = tid / dim, te = tid - tc * dim;
    if (tc >= cpt) return;
    for (int t = tc; t < nc; t += BU * cpt) {
        float v[BU], dv[BU][NO], xv[BU][NO];
        // a column past nc rereads column nc-1, an empty entry reads atom 0
        // and coefficient 0 and is then dropped
#pragma unroll
        for (int b = 0; b < BU; b++) {
            const int tt = min(t + b * cpt, nc - 1);
            v[b] = a.Y[(int64_t)sp[tt] * dim + te];
#pragma unroll
            for (int i = 0; i < NO; i++) {
                const float *dp = a.D + (int64_t)max(sat[tt * PS_NOMAX + i], 0) * dim + te;
                const float *cp = a.coef + max(sci[tt * PS_NOMAX + i], 0);
                dv[b][i] = SC1 ? ld_c(dp) : *dp;
                xv[b][i] = SC1 ? ld_c(cp) : *cp;
            }
        }
#pragma unroll
        for (int b = 0; b < BU; b++) {
            const int tt = t + b * cpt;
            if (tt < nc) {
#pragma unroll
                for (int i = 0; i < NO; i++)
                    if (sci[tt * PS_NOMAX + i] >= 0) v[b] = fmaf(-dv[b][i], xv[b][i], v[b]);
                Es[tt * dpad + te] = v[b];
            }
        }
    }
}

// After the sweep: coefficient 0 for every user of atom k beyond its first
// min(c_k, cap) (workgroups (k, 0 .. gridDim.y) share atom k's users).
__global__ void pa_zero_kernel(SweepArgs a)
{
    const int k = blockIdx.x;
    const int base = a.uoff[k], c = a.uoff[k + 1] - base;
    const int step = gridDim.y * blockDim.x;
    for (int t = a.cap + blockIdx.y * blockDim.x + threadIdx.x; t < c; t += step) {
        const int pos = base + t;
        a.coef[(int64_t)a.users[pos] * a.kmax + a.uslot[pos]] = 0.f;
    }
}

// A team member's table of one atom's columns (pa_table) and the atom's
// starting vector; two sets, so that atom k+2's is loaded while atom k runs.
constexpr int PT_TABW = 2 * PS_CWMAX * PS_NOMAX + 5 * PS_CWMAX + 1;  // ints of one set
constexpr int PT_MULTI_MAX = 16;  // pictures of one multi-picture sweep at most
constexpr int PT_EH0 = 192;  // threads [PT_EH0, PT_TB0): the hand-over columns at an atom's start
constexpr int PT_TB0 = 512;  // threads [PT_TB0, PT_T): the next atom's table at an atom's start

struct PtTab {
    int *sat, *sci, *sp, *hrank, *hl, *ss, *sx, *nhp;
    float *d0;
};

template <int NO, int BU>
__device__ void pt_helper(const SweepArgs &a, unsigned *flags, int team, int h, unsigned long long *Wx, float *sm,
                          int *bail)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int K = a.K, cap = a.cap, H = PT_HT;
    int dim = a.dim, dpad = a.dpad;
    // LDS layout (pitch dpad)
    float *Es = sm;                           // PS_CWMAX x dpad: this helper's columns of the team's atom
    float *Eh = Es + PS_CWMAX * dpad;         // PS_CWMAX x dpad: columns of E_{k-1} for the hand-over (hl order)
    float *dprev = Eh + PS_CWMAX * dpad;      // dpad: d_{k-1}, then d_k for P4
    float *wpw = dprev + dpad;                // PT_W x dpad: per-wavefront partial sums of E v
    float *wx = wpw + PT_W * dpad;            // PT_GRP x dpad: partial sums of the exchange
    float *d0s = wx + PT_GRP * dpad;          // 2 x dpad: the starting vectors of the two table sets
    int *tabs = (int *)(d0s + 2 * dpad);      // 2 x PT_TABW: the two table sets
    int *suo = tabs + 2 * PT_TABW;            // K + 1: uoff
    int *smo = suo + K + 1;                   // K + 1: moff
    auto tab = [&](int set) {
        PtTab t;
        t.sat = tabs + set * PT_TABW;         // PS_CWMAX x PS_NOMAX atoms (-1: none)
        t.sci = t.sat + PS_CWMAX * PS_NOMAX;  // PS_CWMAX x PS_NOMAX coefficient indices (-1: none)
        t.sp = t.sci + PS_CWMAX * PS_NOMAX;   // PS_CWMAX patches
        t.hrank = t.sp + PS_CWMAX;            // PS_CWMAX: rank among atom k-1's users, -1: none
        t.hl = t.hrank + PS_CWMAX;            // PS_CWMAX: columns with a hand-over
        t.ss = t.hl + PS_CWMAX;               // PS_CWMAX: own coefficient slots of the columns
        t.sx = t.ss + PS_CWMAX;               // PS_CWMAX: 1 for a column atom k+1's hand-over reads
        t.nhp = t.sx + PS_CWMAX;              // the number of hand-over columns
        t.d0 = d0s + set * dpad;              // dpad: the atom as it stands (the starting vector)
        return t;
    };
    unsigned *ecert = flags + PTF_ECERT, *ccert = flags + PTF_CCERT;
    const int me = team * PT_HT + h, other = (1 - team) * PT_HT;
    unsigned long long *Wt = Wx + (size_t)team * 2 * PT_HT * dpad;  // this team's two pair buffers
    // d_k pairs of each team (after both teams' exchange buffers)
    unsigned long long *Dp = Wx + (size_t)4 * PT_HT * dpad;
    unsigned seq = 0;
    unsigned long long iters = 0;

    for (int i = tid; i <= K; i += PT_T) {
        suo[i] = a.uoff[i];
        smo[i] = a.moff[i];
    }
    if (tid == 0) {
        *tab(0).nhp = 0;
        *tab(1).nhp = 0;
    }
    __syncthreads();
    // the team's first atom k = team: its columns without atom k-1 (nothing
    // is final yet) and atom k's vector
    int k = team;
    if (k < K) {
        const PtTab T0 = tab(0);
        int tb0;
        const int nc0 = cols_of(pa_m(suo, k, cap), H, h, &tb0);
        pa_table<NO>(a, suo, smo, k, tb0, nc0, k >= 1 ? pa_m(suo, k - 1, cap) : 0, T0.sp, T0.ss, T0.sat, T0.sci,
                     T0.hrank, T0.hl, T0.nhp, T0.sx, 0, PT_T);
        if (tid < dim) T0.d0[tid] = ld_c(a.D + (int64_t)k * dim + tid);
        __syncthreads();
        pa_build<NO, BU>(a, nc0, Es, T0.sp, T0.sat, T0.sci);
        __syncthreads();
    }
    for (; k < K; k += 2) {
        // dim and dpad pass through an opaque copy each atom, so that offsets
        // derived from them are recomputed where they are used instead of
        // being held in registers across the loop
        asm volatile("" : "+s"(dim), "+s"(dpad));
        const int cs = (k >> 1) & 1;  // this atom's table set; atom k+2's is the other
        const PtTab T = tab(cs), Tn = tab(cs ^ 1);
        int tb;
        const int m = pa_m(suo, k, cap);
        const int nc = cols_of(m, H, h, &tb);
        float *Ecur = (k & 1) ? a.E1 : a.E0;
        const float *Eprev = (k & 1) ? a.E0 : a.E1;
        const int kn = k + 2;
        // ---- atom start: three thread groups at once --------------------------
        //   threads < dim: team 1-t's d_{k-1} pairs (tagged k);
        //   [PT_EH0, PT_TB0): the hand-over columns of E_{k-1}, each once its
        //     owner (the member of team 1-t that holds that rank) has certified
        //     its columns (ecert >= k), read with L1-bypassing loads;
        //   [PT_TB0, PT_T): atom k+2's table and starting vector (static until
        //     that atom's own update).
        if (k >= 1 && tid < dim) {
            const unsigned long long *pp = Dp + (size_t)(1 - team) * dpad + tid;
            unsigned long long pv = ld64_c(pp);
            long spins = 0;
            while ((unsigned)(pv >> 32) != (unsigned)k) {
                __builtin_amdgcn_s_sleep(2);
                if ((++spins & 15) == 0) {
                    if (ld_flag(flags + PSF_ABORT)) {
                        *bail = 1;
                        break;
                    }
                    if (spins > spin_limit(flags) / 2) {
                        st_flag(flags + PSF_STICKY, 1u);
                        st_flag(flags + PSF_ABORT, 1u);
                        *bail = 1;
                        break;
                    }
                }
                pv = ld64_c(pp);
            }
            dprev[tid] = __uint_as_float((unsigned)pv);
        } else if (k >= 1 && tid >= PT_EH0 && tid < PT_TB0) {
            const int nh = *T.nhp;
            const int cwp = (pa_m(suo, k - 1, cap) + H - 1) / H;  // team 1-t's columns per member for atom k-1
            int polled = -1;
            bool gaveup = false;
            for (int q = tid - PT_EH0; q < nh * dpad && !gaveup; q += PT_TB0 - PT_EH0) {
                const int i = q / dpad, c = q - i * dpad;
                const int r = T.hrank[T.hl[i]];
                const int owner = r / cwp;
                if (owner != polled) {
                    long spins = 0;
                    while (ld_flag(ecert + other + owner) < (unsigned)k) {
                        __builtin_amdgcn_s_sleep(2);
                        if ((++spins & 15) == 0) {
                            if (ld_flag(flags + PSF_ABORT)) {
                                *bail = 1;
                                gaveup = true;
                                break;
                            }
                            if (spins > spin_limit(flags)) {
                                st_flag(flags + PSF_STICKY, 1u);
                                st_flag(flags + PSF_ABORT, 1u);
                                *bail = 1;
                                gaveup = true;
                                break;
                            }
                        }
                    }
                    polled = owner;
                }
                if (!gaveup) Eh[i * dpad + c] = ld_c(Eprev + (size_t)r * dpad + c);
            }
        } else if (kn < K && tid >= PT_TB0) {
            int tbn;
            const int ncn = cols_of(pa_m(suo, kn, cap), H, h, &tbn);
            pa_table<NO>(a, suo, smo, kn, tbn, ncn, pa_m(suo, k + 1, cap), Tn.sp, Tn.ss, Tn.sat, Tn.sci, Tn.hrank,
                         Tn.hl, Tn.nhp, Tn.sx, PT_TB0, PT_T - PT_TB0);
            if (tid - PT_TB0 < dim) Tn.d0[tid - PT_TB0] = ld_c(a.D + (int64_t)kn * dim + (tid - PT_TB0));
        }
        __syncthreads();
        if (*bail) return;
        // ---- P2: atom k-1's share --------------------------------------------
        if (k >= 1) {
            const int nh = *T.nhp;
            float dv[3];
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int e = lane + 64 * j;
                dv[j] = e < dim ? dprev[e] : 0.f;
            }
            for (int i = wave; i < nh; i += PT_W) {
                const int tl = T.hl[i];
                float ev[3];
#pragma unroll
                for (int j = 0; j < 3; j++) ev[j] = Eh[i * dpad + min(lane + 64 * j, dim - 1)];
                const float x = ps_dot(ev, dv, dim);
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const int e = lane + 64 * j;
                    if (e < dim) Es[tl * dpad + e] = fmaf(-dv[j], x, Es[tl * dpad + e]);
                }
            }
            __syncthreads();
        }
        // ---- PI ------------------------------------------------------------------
        float uj[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int e = lane + 64 * j;
            uj[j] = e < dim ? T.d0[e] : 0.f;
        }
        // E_k's columns that team 1-t's hand-over of atom k+1 reads out, and ecert
        auto store_e = [&](bool publish) {
            const int d4 = dpad / 4;
            for (int q = tid; q < nc * d4; q += PT_T) {
                const int tl = q / d4, c4 = q - tl * d4;
                if (T.sx[tl])
                    *(f4 *)(Ecur + (int64_t)(tb + tl) * dpad + 4 * c4) = *(const f4 *)(Es + tl * dpad + 4 * c4);
            }
            if (publish) ps_publish(ecert + me, (unsigned)(k + 1));
        };
        // issue: wavefront w < PT_PW takes columns w, w + PT_PW, ...: v_t =
        // e_t . u_{i-1} and its partial sum of e_t v_t; w_h = the PT_PW
        // partials in wavefront order, stored as pairs
        const bool piw = wave < PT_PW;
        // cert: this is exchange 2, and the E_k columns stored after exchange
        // 1's vectors are certified at its barrier (ecert), long before team
        // 1-t reads them at atom k+1's start
        auto issue = [&](unsigned s, bool cert) {
            unsigned long long *Wb = Wt + (size_t)(s & 1) * PT_HT * dpad;
            if (piw) {
                float pw[3] = {0.f, 0.f, 0.f};
                for (int g0 = 0; wave + PT_PW * g0 < nc; g0 += 4) {
                    float ev[4][3];
#pragma unroll
                    for (int b = 0; b < 4; b++) {
                        const int t = min(wave + PT_PW * (g0 + b), nc - 1);
#pragma unroll
                        for (int j = 0; j < 3; j++) ev[b][j] = Es[t * dpad + min(lane + 64 * j, dim - 1)];
                    }
                    // the four columns' dots (u is 0 past dim) reduced together
                    float a4[4];
#pragma unroll
                    for (int b = 0; b < 4; b++) {
                        a4[b] = 0.f;
#pragma unroll
                        for (int j = 0; j < 3; j++) a4[b] = fmaf(ev[b][j], uj[j], a4[b]);
                    }
                    float vt[4];
                    reduce4_cols(a4, vt);
#pragma unroll
                    for (int b = 0; b < 4; b++)
                        if (wave + PT_PW * (g0 + b) < nc)
#pragma unroll
                            for (int j = 0; j < 3; j++) pw[j] = fmaf(ev[b][j], vt[b], pw[j]);
                }
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const int e = lane + 64 * j;
                    if (e < dim) wpw[wave * dpad + e] = pw[j];
                }
            }
            if (cert) {
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the E_k columns are out
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
            }
            __syncthreads();
            if (cert && tid == 0) st_flag(ecert + me, (unsigned)(k + 1));
            if (tid < dim) {
                float w = 0.f;
#pragma unroll
                for (int q = 0; q < PT_PW; q++) w += wpw[q * dpad + tid];
                st64_c(Wb + (size_t)h * dpad + tid, ((unsigned long long)s << 32) | __float_as_uint(w));
            }
        };
        // collect: z_i = the team's vectors of exchange s (group gq sums
        // helpers gq, gq + ngrp, ..., then the groups in order), then u_i;
        // returns -1 to leave (abort), 1 when the iteration stops, else 0
        int it = 0;
        float d2prev = 0.f;
        auto collect = [&](unsigned s) -> int {
            const unsigned long long *Wb = Wt + (size_t)(s & 1) * PT_HT * dpad;
            const int ngrp = min(PT_T / dim, PT_GRP);
            const int gq = tid / dim, ge = tid - gq * dim;
            if (gq < ngrp) {
                unsigned long long pv[PT_XH];
#pragma unroll
                for (int j = 0; j < PT_XH; j++) {
                    const int hh = gq + ngrp * j;
                    pv[j] = hh < H ? ld64_c(Wb + (size_t)hh * dpad + ge) : ((unsigned long long)s << 32);
                }
                // a thread still missing pairs rereads all its missing ones
                // together (a few 8-byte loads per round)
                long spins = 0;
                for (;;) {
                    bool ok = true;
#pragma unroll
                    for (int j = 0; j < PT_XH; j++) ok = ok && (unsigned)(pv[j] >> 32) == s;
                    if (ok) break;
                    __builtin_amdgcn_s_sleep(1);
                    if ((++spins & 31) == 0) {
                        if (ld_flag(flags + PSF_ABORT)) {
                            *bail = 1;
                            break;
                        }
                        if (spins > spin_limit(flags)) {
                            st_flag(flags + PSF_STICKY, 1u);
                            st_flag(flags + PSF_ABORT, 1u);
                            *bail = 1;
                            break;
                        }
                    }
#pragma unroll
                    for (int j = 0; j < PT_XH; j++) {
                        const int hh = gq + ngrp * j;
                        if (hh < H && (unsigned)(pv[j] >> 32) != s) pv[j] = ld64_c(Wb + (size_t)hh * dpad + ge);
                    }
                }
                float z = __uint_as_float((unsigned)pv[0]);
#pragma unroll
                for (int j = 1; j < PT_XH; j++)
                    if (gq + ngrp * j < H) z += __uint_as_float((unsigned)pv[j]);
                wx[gq * dpad + ge] = z;
            }
            __syncthreads();
            if (*bail) return -1;
            // every wavefront forms z_i, |z_i|^2, u_i and |u_i - u_{i-1}|^2
            // (each the same, bit for bit) and takes the decision itself, so
            // no barrier passes it on
            int dec;
            {
                float zj[3], n2 = 0.f;
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const int e = lane + 64 * j;
                    float z = 0.f;
                    if (e < dim)
                        for (int q = 0; q < ngrp; q++) z += wx[q * dpad + e];
                    zj[j] = z;
                    n2 = fmaf(z, z, n2);
                }
                n2 = wave_sum_dpp(n2);
                if (n2 == 0.f) {
                    dec = 1;  // E = 0: keep u_{i-1} (d_k at the start; its coefficients become 0)
                } else {
                    const float inv = 1.f / sqrtf(n2);
                    float d2 = 0.f;
#pragma unroll
                    for (int j = 0; j < 3; j++) {
                        const float un = zj[j] * inv;
                        const float du = un - uj[j];
                        d2 = fmaf(du, du, d2);
                        uj[j] = un;
                    }
                    d2 = wave_sum_dpp(d2);
                    dec = power_stop(it + 1, d2, d2prev, a.tol2, a.est2, a.maxit) ? 2 : 0;
                    d2prev = d2;
                }
            }
            if (dec != 1) it++;
            return dec != 0 ? 1 : 0;
        };
        if (m > 0) {
            issue(++seq, false);
            store_e(false);
            int r = collect(seq);
            bool certified = false;
            while (r == 0) {
                issue(++seq, !certified);
                certified = true;
                r = collect(seq);
            }
            if (r < 0) return;
            // one exchange only: certify the columns now
            if (!certified) ps_publish(ecert + me, (unsigned)(k + 1));
            iters += (unsigned long long)it;
        } else {
            store_e(true);
        }
        // ---- d_k out ---------------------------------------------------------
        if (wave == 0) {
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int e = lane + 64 * j;
                if (e < dim) {
                    if (h == 0)
                        st64_c(Dp + (size_t)team * dpad + e, ((unsigned long long)(k + 1) << 32) | __float_as_uint(uj[j]));
                    dprev[e] = uj[j];
                    if (h == 0) a.D[(int64_t)k * dim + e] = uj[j];
                }
            }
        }
        if (h == 0) ps_publish(flags + PTF_DONE + team, (unsigned)(k + 1));
        else __syncthreads();
        // ---- P4: atom k's final coefficients ---------------------------------
        {
            float dv[3];
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int e = lane + 64 * j;
                dv[j] = e < dim ? dprev[e] : 0.f;
            }
#pragma unroll
            for (int g = 0; g < PS_CWMAX / PT_W; g++) {
                const int tl = wave + PT_W * g;
                if (tl >= nc) continue;
                float ev[3];
#pragma unroll
                for (int j = 0; j < 3; j++) ev[j] = Es[tl * dpad + min(lane + 64 * j, dim - 1)];
                const float x = ps_dot(ev, dv, dim);
                if (lane == 0) a.coef[T.ss[tl]] = x;
            }
        }
        ps_publish(ccert + me, (unsigned)(k + 1));
        // ---- P1: atom k+2's columns without atom k+1 ------------------------------
        // (its table and vector were loaded at this atom's start; once the
        // other atoms' coefficients are certified, the patches and the other
        // atoms' shares in one pass)
        if (kn >= K) break;
        // this atom's set is refilled at atom k+2's start, for atom k+4 (the
        // barriers of the wait below order this before those threads' use)
        if (tid == 0) *T.nhp = 0;
        int tbn;
        const int ncn = cols_of(pa_m(suo, kn, cap), H, h, &tbn);
        if (!ps_wait3(flags, PTF_CCERT + team * PT_HT, H, (unsigned)(k + 1), PTF_CCERT + other, k >= 1 ? H : 0,
                      (unsigned)k, PTF_DONE + team, 1, (unsigned)(k + 1), bail))
            return;
        l1_drop();
        pa_build<NO, BU, false>(a, ncn, Es, Tn.sp, Tn.sat, Tn.sci);
        __syncthreads();
    }
    if (h == 0 && tid == 0) atomicAdd(&a.counters[CNT_EIG_IT], iters);
}

template <int NO, int BU>
__global__ __launch_bounds__(PT_T) void pt_sweep_kernel(SweepArgs a, unsigned *flags, unsigned long long *Wx)
{
    extern __shared__ __attribute__((aligned(16))) float sm[];
    __shared__ int s_role, s_bail[2];  // bail word, spin limit
    if (threadIdx.x == 0) {
        s_role = ps_register(flags, 2 * PT_HT);
        s_bail[0] = 0;
        s_bail[1] = (int)ld_flag(flags + PSF_SPIN);
    }
    __syncthreads();
    const int role = s_role;
    if (role < 0) return;
    pt_helper<NO, BU>(a, flags, role & 1, role >> 1, Wx, sm, s_bail);
    if (threadIdx.x == 0 && s_bail[0] == 0) atomicAdd(&flags[PSF_FINISHED], 1u);
}

// Several pictures' two-team sweeps in one launch, picture i on XCD i: the
// launch has n_xcd x 2 PT_HT workgroups, the XCDs take them round robin, and
// the 2 PT_HT on XCD i < n take the roles of picture i's sweep (its own flags,
// pairs and buffers); the rest leave.  Each picture's sweep is the two-team
// sweep above, bit for bit.
struct PtPicture {
    SweepArgs a;
    unsigned *flags;
    unsigned long long *wx;
};

template <int NO, int BU>
__global__ __launch_bounds__(PT_T) void pt_multi_kernel(const PtPicture *__restrict__ pics, int n)
{
    extern __shared__ __attribute__((aligned(16))) float sm[];
    __shared__ int s_role, s_bail[2];  // bail word, spin limit
    const unsigned xcc = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15u;  // XCC_ID[3:0]
    if (xcc >= (unsigned)n) return;
    unsigned *flags = pics[xcc].flags;
    if (threadIdx.x == 0) {
        (void)atomicCAS(&flags[PSF_XSEL], 0u, xcc + 1u);
        const unsigned slot = atomicAdd(&flags[PSF_REG + xcc], 1u);
        s_role = slot < (unsigned)(2 * PT_HT) ? (int)slot : -1;
        s_bail[0] = 0;
        s_bail[1] = (int)ld_flag(flags + PSF_SPIN);
    }
    __syncthreads();
    const int role = s_role;
    if (role < 0) return;
    const SweepArgs a = pics[xcc].a;
    pt_helper<NO, BU>(a, flags, role & 1, role >> 1, pics[xcc].wx, sm, s_bail);
    if (threadIdx.x == 0 && s_bail[0] == 0) atomicAdd(&flags[PSF_FINISHED], 1u);
}

// ===========================================================================
// Eight-team sweep (et_sweep_kernel, the default for user caps up to 1024).
// The 32 workgroups of one XCD form 8 teams of 4 members; atom k belongs to
// team k mod 8.  A member holds its quarter of atom k's capped columns (at
// most 256) in registers: thread (row lane rb = lane & 15, column lane
// q = 4 wave + lane / 16) holds rows rb + 16 i (i < NR) of columns
// tl = 32 j + q (j < ET_CB).  An exchange of the power iteration forms
// E_h (E_h^T u) from registers, and four members exchange their vectors.
// A team has seven other atoms' time to prepare its next atom k:
//   * build, once atoms up to k - 8 are certified: the columns without the
//     shares of atoms k-7 .. k-1 (old values for the atoms after k);
//   * window: the shares of atoms k-7 .. k-3, each once its team has
//     certified its vector and coefficients;
//   * speculation: the first exchange, from d_k, on the columns still without
//     atoms k-2's and k-1's shares.  Each member publishes those of its
//     columns that the two shares change (the list of the atom), so that
//     every member can correct the speculative sum exactly once d_{k-1} is in:
//     + e (e . d_k) - e' (e' . d_k) for each listed column, e' as published,
//     e = e' - d_{k-2} x_{k-2} - d_{k-1} x_{k-1};
//   * critical path: d_{k-1} in, atom k-1's coefficient of each listed
//     column as e . d_{k-1} from the column of E_{k-1} that atom k-1's team
//     stored (the same dot, bit for bit, as the stored coefficient), the
//     correction, then the remaining exchanges.
// Then d_k out as tagged pairs to team k+1, and the coefficients of atom
// k's columns.  A team list longer than the stage (or an atom without users)
// takes the plain path: atom k-2's share in the window, atom k-1's share,
// then every exchange after it.

constexpr int ET_T = 1024;                 // threads per member
constexpr int ET_W = ET_T / 64;            // wavefronts per member
constexpr int ET_QL = 4 * ET_W;            // column lanes (4 lane rows of 16 lanes per wavefront)
constexpr int ET_HT = 4;                   // members per team
constexpr int ET_NT = 8;                   // teams (ET_HT x ET_NT = one XCD's 32 CUs)
constexpr int ET_CB = 4;                   // column slots per thread
constexpr int ET_COLS = ET_QL * ET_CB;     // columns per member (a cap of up to 1024 per atom)
constexpr int ET_WIN = ET_NT - 1;          // atoms whose shares are deferred past the build
constexpr int ET_LMAX = 48;                // listed columns per team (global layout)
constexpr int ET_LE = 1 + PS_NOMAX;        // ints per list entry (global layout)
constexpr int ET_LM = 1 + ET_LMAX * ET_LE; // ints of a member's list: its count, then the entries

// Listed columns a team stages in LDS (the speculative path's limit; a
// longer list takes the plain path).
__host__ __device__ constexpr int et_hch(int) { return ET_LMAX; }

// Sum over the 16 lanes of a lane row, returned to each of them (rotations
// by 8, 4, 2, 1: every lane ends with the same bits).
__device__ __forceinline__ float row_sum16(float v)
{
    v += dpp_f<0x128, 0xF>(v);  // row_ror:8
    v += dpp_f<0x124, 0xF>(v);  // row_ror:4
    v += dpp_f<0x122, 0xF>(v);  // row_ror:2
    v += dpp_f<0x121, 0xF>(v);  // row_ror:1
    return v;
}

// e . d for a column held by the 16 lanes of a lane row as e[i] = row
// rb + 16 i: fmaf over i in order, then row_sum16.  The coefficients (P4),
// the hand-over (P2) and the speculative correction use it on the same
// values.
template <int NR>
__device__ __forceinline__ float et_dot(const float *e, const float *d)
{
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < NR; i++) s = fmaf(e[i], d[i], s);
    return row_sum16(s);
}

// A member's table of one atom's columns, in LDS.
struct EtTab {
    int *sp;    // ET_COLS: patch
    int *ss;    // ET_COLS: the column's own coefficient slot
    int *sx;    // ET_COLS: 1 when the patch is also among atom k+1's first m_{k+1} users
    int *hidx;  // ET_COLS: position in the hand-over list, -1 for none
    int *hr;    // ET_COLS: rank among atom k-1's first m_{k-1} users (hand-over columns), -1 for none
    int *hl;    // ET_COLS: the hand-over columns
    int *s2;    // ET_COLS: 1 when a support entry of atom k-2 is deferred
    int *sidx;  // ET_COLS: position in the list (hand-over or atom k-2 entry), -1 for none
    int *sl;    // ET_COLS: the listed columns
    int *eat;   // ET_COLS x NO: atom of a support entry, -1 for none
    int *eci;   // ET_COLS x NO: its coefficient index
    int *ek;    // ET_COLS x NO: 1 subtracted at the build, 2 deferred to the window
    int *nh;    // the hand-over list's length
    int *ns;    // the list's length
};

// Column table of this member's columns [tb, tb + nc) of atom kb, one thread
// per column (threads 0 .. ET_COLS-1).  Entries of atoms updated before kb
// (j < kb) whose patch is beyond atom j's first m_j users have coefficient 0
// from atom j's update on and are dropped; atom kb-1's entry becomes the
// hand-over; atoms kb-7 .. kb-2 are deferred; the rest are subtracted at the
// build (atoms up to kb-8 with their new values, atoms after kb with their
// old ones).  The lists follow (et_lists).
template <int NO>
__device__ __forceinline__ void et_table(const SweepArgs &a, const int *suo, const int *smo, int kb, int tb, int nc, const EtTab &T)
{
    const int no = a.kmax > 1 ? a.kmax - 1 : 1;
    const int mprev = kb >= 1 ? pa_m(suo, kb - 1, a.cap) : 0;
    const int mnext = kb + 1 < a.K ? pa_m(suo, kb + 1, a.cap) : 0;
    const int64_t cb = (int64_t)smo[kb] + tb;
    const int tl = (int)threadIdx.x;
    if (tl >= nc) return;
    const int64_t col = cb + tl;
    T.sp[tl] = a.tab_p[col];
    T.ss[tl] = a.tab_self[col];
    int4 o[NO];
#pragma unroll
    for (int i = 0; i < NO; i++) o[i] = a.tab_oth[col * no + min(i, no - 1)];
    int hr = -1, ex = 0, s2 = 0;
#pragma unroll
    for (int i = 0; i < NO; i++) {
        const int at = i < no ? o[i].x : -1;
        const int rk = o[i].z;
        int kind = 0;
        if (at == kb + 1 && rk >= 0 && rk < mnext) ex = 1;
        if (at >= 0 && rk >= 0) {
            if (at == kb - 1) {
                if (rk < mprev) hr = rk;
            } else if (at < kb) {
                if (rk < pa_m(suo, at, a.cap)) kind = at >= kb - ET_WIN ? 2 : 1;
            } else {
                kind = 1;
            }
        }
        if (kind == 2 && at == kb - 2) s2 = 1;
        T.eat[tl * NO + i] = kind ? at : -1;
        T.eci[tl * NO + i] = kind ? o[i].y : 0;
        T.ek[tl * NO + i] = kind;
    }
    T.hr[tl] = hr;
    T.sx[tl] = ex;
    T.s2[tl] = s2;
}

// The hand-over list (columns with atom kb-1's entry) and the list (atom
// kb-1's or kb-2's), both in ascending column order (the speculative
// correction sums the listed columns in list order, so the order is fixed):
// flags counted per wavefront by ballot, then a prefix over the four
// wavefronts of the ET_COLS columns.  Every thread calls it after the table
// is written.
__device__ void et_lists(const EtTab &T, int nc, int *wcnt)
{
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const bool fh = tid < nc && T.hr[tid] >= 0;
    const bool fs = tid < nc && (T.hr[tid] >= 0 || T.s2[tid] != 0);
    const unsigned long long bh = __ballot(fh), bs = __ballot(fs);
    if (tid < ET_COLS && lane == 0) {
        wcnt[w] = __popcll(bh);
        wcnt[4 + w] = __popcll(bs);
    }
    __syncthreads();
    if (tid < ET_COLS) {
        const unsigned long long lo = (1ull << lane) - 1ull;
        int ph = 0, ps = 0;
        for (int x = 0; x < w; x++) {
            ph += wcnt[x];
            ps += wcnt[4 + x];
        }
        ph += __popcll(bh & lo);
        ps += __popcll(bs & lo);
        T.hidx[tid] = fh ? ph : -1;
        T.sidx[tid] = fs ? ps : -1;
        if (fh) T.hl[ph] = tid;
        if (fs) T.sl[ps] = tid;
        if (tid == 0) {
            *T.nh = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
            *T.ns = wcnt[4] + wcnt[5] + wcnt[6] + wcnt[7];
        }
    }
    __syncthreads();
}

// The build: E[i][j] = y_p - sum of the shares marked 1 (plain loads: the
// caller has dropped its L1 lines behind the certificates of the atoms whose
// new values it reads).
template <int NR, int NO>
__device__ __forceinline__ void et_build(const SweepArgs &a, float (&E)[NR][ET_CB], int nc, const EtTab &T, int q,
                                         int rb)
{
    const int dim = a.dim;
    // the patches, then one support entry at a time for the four column
    // slots together (every load of a step in flight; a bounded register
    // set); an entry that does not apply subtracts 0 x d
#pragma unroll
    for (int j = 0; j < ET_CB; j++) {
        const int tl = ET_QL * j + q;
        const int p = tl < nc ? T.sp[tl] : 0;
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const int r = rb + 16 * i;
            const float y = a.Y[(int64_t)p * dim + min(r, dim - 1)];
            E[i][j] = tl < nc && r < dim ? y : 0.f;
        }
    }
#pragma unroll 1
    for (int e = 0; e < NO; e++) {
        // two column slots at a time (their loads in flight together)
#pragma unroll
        for (int j0 = 0; j0 < ET_CB; j0 += 2) {
            float xv[2], dv[2][NR];
#pragma unroll
            for (int jj = 0; jj < 2; jj++) {
                const int tl = ET_QL * (j0 + jj) + q;
                const int tc = min(tl, max(nc - 1, 0));
                const bool k1 = tl < nc && T.ek[tc * NO + e] == 1;
                const float x = a.coef[k1 ? T.eci[tc * NO + e] : 0];
                xv[jj] = k1 ? x : 0.f;
                const float *dr = a.D + (int64_t)(k1 ? T.eat[tc * NO + e] : 0) * dim;
#pragma unroll
                for (int i = 0; i < NR; i++) {
                    const float d = dr[min(rb + 16 * i, dim - 1)];
                    dv[jj][i] = rb + 16 * i < dim ? d : 0.f;
                }
            }
#pragma unroll
            for (int jj = 0; jj < 2; jj++)
#pragma unroll
                for (int i = 0; i < NR; i++) E[i][j0 + jj] = fmaf(-dv[jj][i], xv[jj], E[i][j0 + jj]);
        }
    }
}

// The deferred shares of atom j (certified; plain loads after an L1 drop).
template <int NR, int NO>
__device__ __forceinline__ void et_window(const SweepArgs &a, float (&E)[NR][ET_CB], int nc, const EtTab &T, int q,
                                          int rb, int j)
{
    const int dim = a.dim;
    float dv[NR];
#pragma unroll
    for (int i = 0; i < NR; i++) {
        const float d = a.D[(int64_t)j * dim + min(rb + 16 * i, dim - 1)];
        dv[i] = rb + 16 * i < dim ? d : 0.f;
    }
#pragma unroll
    for (int s = 0; s < ET_CB; s++) {
        const int tl = ET_QL * s + q;
        const int tc = min(tl, max(nc - 1, 0));
#pragma unroll
        for (int e = 0; e < NO; e++) {
            const bool mt = tl < nc && T.ek[tc * NO + e] == 2 && T.eat[tc * NO + e] == j;
            if (!__any(mt)) continue;
            const float c = a.coef[mt ? T.eci[tc * NO + e] : 0];
            const float x = mt ? c : 0.f;
#pragma unroll
            for (int i = 0; i < NR; i++) E[i][s] = fmaf(-dv[i], x, E[i][s]);
        }
    }
}

// Waits until flags[f0 + t] >= v for the threads t < n (ps_wait3); false
// after an abort or a wait that ran out.
__device__ bool et_wait(unsigned *flags, int f0, int n, unsigned v, int *bail)
{
    return ps_wait3(flags, f0, n, v, 0, 0, 0u, 0, 0, 0u, bail);
}

// Waits until flags[f0 + i] >= v for i < n (the calling thread polls all n);
// false after an abort or a wait that ran out.
__device__ bool et_poll(unsigned *flags, int f0, int n, unsigned v, int *bail)
{
    for (int i = 0; i < n; i++) {
        long spins = 0;
        while (ld_flag(flags + f0 + i) < v) {
            __builtin_amdgcn_s_sleep(2);
            if ((++spins & 15) == 0) {
                if (ld_flag(flags + PSF_ABORT)) {
                    *bail = 1;
                    return false;
                }
                if (spins > spin_limit(flags)) {
                    st_flag(flags + PSF_STICKY, 1u);
                    st_flag(flags + PSF_ABORT, 1u);
                    *bail = 1;
                    return false;
                }
            }
        }
    }
    return true;
}

__device__ __forceinline__ int ld_i(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Flag words of the eight-team sweep besides the certificates: per member,
// its list and listed columns are out (atom + 1).
constexpr int ETF_SCERT = 96;

// Global scratch of the eight-team sweep in the partial-Gram buffer: the
// exchange pairs (8 teams x 2 buffers x 4 members x dpad) and the d_k pairs
// (8 x dpad), then each member's listed columns (ET_LMAX x dpad floats) and
// its list (ET_LM ints).
__host__ __device__ constexpr size_t et_pairs_n(int dpad) { return (size_t)(ET_NT * 2 * ET_HT + ET_NT) * dpad; }

// (Inlined into its kernels: a call would take the kernel argument's address
// and copy it to private memory, and every load of a buffer pointer would go
// through it.)
template <int NR, int NO>
__device__ __forceinline__ void et_member(const SweepArgs &a, unsigned *flags, int team, int h,
                                          unsigned long long *Wx, float *sm, int *bail)
{
    constexpr int RP = 16 * NR;      // row pitch of the LDS vectors
    constexpr int HCH = et_hch(NR);  // listed columns staged (the plain path's hand-over chunk too)
    int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    int rb = lane & 15, q = wave * 4 + (lane >> 4);
    const int K = a.K, cap = a.cap;
    int dim = a.dim, dpad = a.dpad;
    // LDS: every offset a compile-time constant except the two K-sized
    // arrays at the end (fewer live registers)
    constexpr int F_ZS = ET_W * RP, F_ZSP = F_ZS + RP, F_DPREV = F_ZSP + RP, F_D0 = F_DPREV + RP,
                  F_DK2 = F_D0 + RP, F_SS = F_DK2 + RP, F_SH = F_SS + HCH * RP, F_CP = F_SH + HCH * RP,
                  F_X2 = F_CP + HCH * RP, F_END = F_X2 + HCH * NO;
    float *part = sm;            // ET_W x RP: per-wavefront partial sums of E (E^T u)
    float *zs = sm + F_ZS;       // RP: the summed exchange, rows past dim 0
    float *zsp = sm + F_ZSP;     // RP: the speculative exchange's sum
    float *dprev = sm + F_DPREV; // RP: d_{k-1}
    float *d0 = sm + F_D0;       // RP: the atom's starting vector
    float *dk2 = sm + F_DK2;     // RP: d_{k-2}
    float *stageS = sm + F_SS;   // HCH x RP: the listed columns as published
    float *stageH = sm + F_SH;   // HCH x RP: their columns of E_{k-1} (hand-over; plain path: own)
    float *cpart = sm + F_CP;    // HCH x RP: each listed column's correction of the speculative sum
    float *x2 = sm + F_X2;       // HCH: atom k-2's coefficient of each listed column (entries summed)
    int *ti = (int *)(sm + F_END);
    constexpr int I_SS = ET_COLS, I_SX = 2 * ET_COLS, I_HIDX = 3 * ET_COLS, I_HR = 4 * ET_COLS,
                  I_HL = 5 * ET_COLS, I_S2 = 6 * ET_COLS, I_SIDX = 7 * ET_COLS, I_SL = 8 * ET_COLS,
                  I_EAT = 9 * ET_COLS, I_ECI = I_EAT + ET_COLS * NO, I_EK = I_ECI + ET_COLS * NO,
                  I_NH = I_EK + ET_COLS * NO, I_NS = I_NH + 1, I_CNT = I_NS + 1, I_SRK = I_CNT + ET_HT,
                  I_SCI = I_SRK + HCH, I_WC = I_SCI + HCH * NO, I_END = I_WC + 8;
    EtTab T;
    T.sp = ti;
    T.ss = ti + I_SS;
    T.sx = ti + I_SX;
    T.hidx = ti + I_HIDX;
    T.hr = ti + I_HR;
    T.hl = ti + I_HL;
    T.s2 = ti + I_S2;
    T.sidx = ti + I_SIDX;
    T.sl = ti + I_SL;
    T.eat = ti + I_EAT;
    T.eci = ti + I_ECI;
    T.ek = ti + I_EK;
    T.nh = ti + I_NH;
    T.ns = ti + I_NS;
    int *cntl = ti + I_CNT;  // ET_HT: the members' list lengths
    int *srk = ti + I_SRK;   // HCH: the listed columns' ranks among atom k-1's users (-1: none)
    int *sci = ti + I_SCI;   // HCH x NO: coefficient index of each atom k-2 entry (-1: none)
    int *wcnt = ti + I_WC;   // 8: per-wavefront counts of the two lists
    int *suo = ti + I_END;   // K + 1: uoff
    int *smo = suo + K + 1;  // K + 1: moff

    unsigned *ecert = flags + PTF_ECERT, *ccert = flags + PTF_CCERT;
    const int me = team * ET_HT + h;
    unsigned long long *Wt = Wx + (size_t)team * 2 * ET_HT * dpad;   // this team's two pair buffers
    unsigned long long *Dp = Wx + (size_t)ET_NT * 2 * ET_HT * dpad;  // each team's d_k pairs
    float *Esp = (float *)(Wx + et_pairs_n(dpad));
    int *Lst = (int *)(Esp + (size_t)ET_NT * ET_HT * ET_LMAX * dpad);
    unsigned seq = 0;
    unsigned long long iters = 0;

    for (int i = tid; i <= K; i += ET_T) {
        suo[i] = a.uoff[i];
        smo[i] = a.moff[i];
    }
    for (int i = tid; i < (ET_W + 6) * RP; i += ET_T) sm[i] = 0.f;
    __syncthreads();

    float E[NR][ET_CB];
    float u[NR];
    int k = team;
    int tb = 0, nc = 0;
    if (k < K) {
        nc = cols_of(pa_m(suo, k, cap), ET_HT, h, &tb);
        et_table<NO>(a, suo, smo, k, tb, nc, T);
        if (tid < dim) d0[tid] = a.D[(int64_t)k * dim + tid];
        __syncthreads();
        et_lists(T, nc, wcnt);
        et_build<NR, NO>(a, E, nc, T, q, rb);
    }

    // One exchange of the power iteration: zs = the team's sum of E_h (E_h^T
    // v), in a fixed order (every member gets the same bits).  The stores
    // issued before it are certified at its barrier: the hand-over columns
    // (ecert) or the list and listed columns (scert).
    auto exchange = [&](const float *v, bool certE, bool certS, int kk) -> bool {
        const unsigned s = ++seq;
        unsigned long long *Wb = Wt + (size_t)(s & 1) * ET_HT * dpad;
        float t[ET_CB];
#pragma unroll
        for (int j = 0; j < ET_CB; j++) {
            float p = 0.f;
#pragma unroll
            for (int i = 0; i < NR; i++) p = fmaf(E[i][j], v[i], p);
            t[j] = row_sum16(p);
        }
        float w[NR];
#pragma unroll
        for (int i = 0; i < NR; i++) {
            float x = 0.f;
#pragma unroll
            for (int j = 0; j < ET_CB; j++) x = fmaf(E[i][j], t[j], x);
            w[i] = add_xor32(add_xor16(x));
        }
        if ((lane >> 4) == 0) {
#pragma unroll
            for (int i = 0; i < NR; i++) part[wave * RP + rb + 16 * i] = w[i];
        }
        if (certE || certS) {
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the certified stores are out
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
        }
        __syncthreads();
        if (tid == 0) {
            if (certE) st_flag(ecert + me, (unsigned)(kk + 1));
            if (certS) st_flag(flags + ETF_SCERT + me, (unsigned)(kk + 1));
        }
        if (tid < dim) {
            float x = 0.f;
#pragma unroll
            for (int wv = 0; wv < ET_W; wv++) x += part[wv * RP + tid];
            st64_c(Wb + (size_t)h * dpad + tid, ((unsigned long long)s << 32) | __float_as_uint(x));
        }
        // lanes 4r .. 4r+3 read member 0..3's pair of row r
        for (int idx = tid; idx < ET_HT * dim; idx += ET_T) {
            const int rr = idx >> 2, mm = idx & 3;
            const unsigned long long *pp = Wb + (size_t)mm * dpad + rr;
            unsigned long long pv = ld64_c(pp);
            long spins = 0;
            while ((unsigned)(pv >> 32) != s) {
                __builtin_amdgcn_s_sleep(1);
                if ((++spins & 31) == 0) {
                    if (ld_flag(flags + PSF_ABORT)) {
                        *bail = 1;
                        break;
                    }
                    if (spins > spin_limit(flags)) {
                        st_flag(flags + PSF_STICKY, 1u);
                        st_flag(flags + PSF_ABORT, 1u);
                        *bail = 1;
                        break;
                    }
                }
                pv = ld64_c(pp);
            }
            const float z = quad_sum(__uint_as_float((unsigned)pv));
            if (mm == 0) zs[rr] = z;
        }
        __syncthreads();
        return *bail == 0;
    };
    // u_i = zs / |zs| and the stop rule (every wavefront takes the same
    // decision); true when the iteration stops (E = 0 keeps u)
    int it = 0;
    float d2prev = 0.f;
    auto step = [&]() -> bool {
        float z[NR], n2 = 0.f;
#pragma unroll
        for (int i = 0; i < NR; i++) {
            z[i] = zs[rb + 16 * i];
            n2 = fmaf(z[i], z[i], n2);
        }
        n2 = row_sum16(n2);
        if (n2 == 0.f) return true;
        const float inv = 1.f / sqrtf(n2);
        float d2 = 0.f;
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const float un = z[i] * inv;
            const float du = un - u[i];
            d2 = fmaf(du, du, d2);
            u[i] = un;
        }
        d2 = row_sum16(d2);
        it++;
        if (power_stop(it, d2, d2prev, a.tol2, a.est2, a.maxit)) return true;
        d2prev = d2;
        return false;
    };

    for (; k < K; k += ET_NT) {
        // the lane coordinates and sizes pass through an opaque copy each
        // atom, so that the addresses derived from them are formed where they
        // are used instead of being held (and spilled) across the loop
        asm volatile("" : "+s"(dim), "+s"(dpad));
        asm volatile("" : "+v"(rb), "+v"(q), "+v"(tid), "+v"(lane));
        asm volatile("" : "+s"(wave));
        const int m = pa_m(suo, k, cap);
        float *Ecur = (k & 1) ? a.E1 : a.E0;
        const float *Eprev = (k & 1) ? a.E0 : a.E1;
        const int cwp = k >= 1 ? (pa_m(suo, k - 1, cap) + ET_HT - 1) / ET_HT : 1;
        const int tprev = (k + ET_NT - 1) % ET_NT;
        auto owner_wait = [&](int rk, int &polled) -> bool {
            const int owner = rk / cwp;
            if (owner == polled) return true;
            polled = owner;
            return et_poll(flags, PTF_ECERT + tprev * ET_HT + owner, 1, (unsigned)k, bail);
        };
        auto window = [&](int j) -> bool {
            if (!et_wait(flags, PTF_CCERT + (j % ET_NT) * ET_HT, ET_HT, (unsigned)(j + 1), bail)) return false;
            l1_drop();
            et_window<NR, NO>(a, E, nc, T, q, rb, j);
            return true;
        };
        // ---- window: atoms k-7 .. k-3, each once certified -----------------
        for (int j = max(0, k - ET_WIN); j <= k - 3; j++)
            if (!window(j)) return;
        const int nh = *T.nh, ns = *T.ns;
        // ---- speculation ------------------------------------------------------
        const bool spec = a.spec && k >= 1 && m > 0;
        bool sp = false;
        int total = 0, goff = 0;
        if (spec) {
            // this member's list: the listed columns as they stand (without
            // atoms k-2 and k-1), each column's rank among atom k-1's users
            // and its atom k-2 entries' coefficient indices
            int *lm = Lst + (size_t)me * ET_LM;
            float *em = Esp + (size_t)me * ET_LMAX * dpad;
#pragma unroll
            for (int s = 0; s < ET_CB; s++) {
                const int tl = ET_QL * s + q;
                const int si = tl < nc ? T.sidx[tl] : -1;
                if (si >= 0 && si < HCH) {
#pragma unroll
                    for (int i = 0; i < NR; i++)
                        if (rb + 16 * i < dim) em[(size_t)si * dpad + rb + 16 * i] = E[i][s];
                    if (rb == 0) {
                        int *le = lm + 1 + si * ET_LE;
                        le[0] = T.hr[tl];
#pragma unroll
                        for (int e = 0; e < NO; e++)
                            le[1 + e] = T.ek[tl * NO + e] == 2 && T.eat[tl * NO + e] == k - 2 ? T.eci[tl * NO + e] : -1;
                    }
                }
            }
            if (tid == 0) lm[0] = ns;
            // the speculative exchange from d_k; its barrier certifies the list
#pragma unroll
            for (int i = 0; i < NR; i++) u[i] = d0[rb + 16 * i];
            if (!exchange(u, false, true, k)) return;
            if (tid < dim) zsp[tid] = zs[tid];
            // the team's list
            if (tid < ET_HT) {
                const bool ok = et_poll(flags, ETF_SCERT + team * ET_HT + tid, 1, (unsigned)(k + 1), bail);
                cntl[tid] = ok ? ld_i(Lst + (size_t)(team * ET_HT + tid) * ET_LM) : 0;
            }
            __syncthreads();
            if (*bail) return;
            int off[ET_HT + 1];
            off[0] = 0;
#pragma unroll
            for (int mm = 0; mm < ET_HT; mm++) off[mm + 1] = off[mm] + cntl[mm];
            total = off[ET_HT];
            goff = off[h];
            sp = total <= HCH;
            if (sp) {
                auto member_of = [&](int g, int &i) {
                    int mm = 0;
#pragma unroll
                    for (int x = 1; x < ET_HT; x++) mm += g >= off[x] ? 1 : 0;
                    i = g - off[mm];
                    return mm;
                };
                for (int idx = tid; idx < total * (1 + NO); idx += ET_T) {
                    const int g = idx / (1 + NO), f = idx - g * (1 + NO);
                    int i;
                    const int mm = member_of(g, i);
                    const int v = ld_i(Lst + (size_t)(team * ET_HT + mm) * ET_LM + 1 + i * ET_LE + f);
                    if (f == 0) srk[g] = v;
                    else sci[g * NO + f - 1] = v;
                }
                for (int qq = tid; qq < total * dim; qq += ET_T) {
                    const int g = qq / dim, r = qq - g * dim;
                    int i;
                    const int mm = member_of(g, i);
                    stageS[g * RP + r] = ld_c(Esp + ((size_t)(team * ET_HT + mm) * ET_LMAX + i) * dpad + r);
                }
                __syncthreads();
                // atom k-2's vector and the listed entries' coefficients, once
                // certified
                if (k >= 2) {
                    if (tid < ET_HT) et_poll(flags, PTF_CCERT + ((k - 2) % ET_NT) * ET_HT + tid, 1, (unsigned)(k - 1), bail);
                    __syncthreads();
                    if (*bail) return;
                    if (tid < dim) dk2[tid] = ld_c(a.D + (int64_t)(k - 2) * dim + tid);
                    // atom k-2's coefficient of each listed column: its entries'
                    // coefficients summed in entry order
                    if (tid < total) {
                        float c[NO];
#pragma unroll
                        for (int e = 0; e < NO; e++) c[e] = ld_c(a.coef + max(sci[tid * NO + e], 0));
                        float xs = 0.f;
#pragma unroll
                        for (int e = 0; e < NO; e++) xs += sci[tid * NO + e] >= 0 ? c[e] : 0.f;
                        x2[tid] = xs;
                    }
                } else if (tid < total) {
                    x2[tid] = 0.f;
                }
                // the hand-over columns of E_{k-1}, each once its owner
                // certified its columns
                int polled = -1;
                for (int qq = tid; qq < total * dim; qq += ET_T) {
                    const int g = qq / dim, r = qq - g * dim;
                    const int rk = srk[g];
                    if (rk < 0) continue;
                    if (!owner_wait(rk, polled)) break;
                    stageH[g * RP + r] = ld_c(Eprev + (size_t)rk * dpad + r);
                }
            }
        }
        if (!sp && k >= 2 && !window(k - 2)) return;
        // ---- atom start: d_{k-1} (threads < dim); on the plain path, this
        //      member's own hand-over columns (from thread 256 on) ----------------
        auto stage_own = [&](int c0, int t0, int tn) {
            const int n = min(nh - c0, HCH);
            int polled = -1;
            for (int qq = (int)tid - t0; qq >= 0 && qq < n * dim; qq += tn) {
                const int i = qq / dim, r = qq - i * dim;
                const int rk = T.hr[T.hl[c0 + i]];
                if (!owner_wait(rk, polled)) break;
                stageH[i * RP + r] = ld_c(Eprev + (size_t)rk * dpad + r);
            }
        };
        if (k >= 1 && tid < dim) {
            const unsigned long long *pp = Dp + (size_t)tprev * dpad + tid;
            unsigned long long pv = ld64_c(pp);
            long spins = 0;
            while ((unsigned)(pv >> 32) != (unsigned)k) {
                __builtin_amdgcn_s_sleep(1);
                if ((++spins & 15) == 0) {
                    if (ld_flag(flags + PSF_ABORT)) {
                        *bail = 1;
                        break;
                    }
                    if (spins > spin_limit(flags)) {
                        st_flag(flags + PSF_STICKY, 1u);
                        st_flag(flags + PSF_ABORT, 1u);
                        *bail = 1;
                        break;
                    }
                }
                pv = ld64_c(pp);
            }
            dprev[tid] = __uint_as_float((unsigned)pv);
        } else if (k >= 1 && !sp && tid >= 256) {
            stage_own(0, 256, ET_T - 256);
        }
        __syncthreads();
        if (*bail) return;
        // ---- atom k-1's share (and, speculating, atom k-2's): x = e . d_{k-1}
        //      from the hand-over column (its rows past dim masked), then
        //      e -= d_{k-1} x --------------------------------------------------------
        if (sp) {
            float dv[NR], d2v[NR];
#pragma unroll
            for (int i = 0; i < NR; i++) {
                dv[i] = dprev[rb + 16 * i];
                d2v[i] = dk2[rb + 16 * i];
            }
#pragma unroll
            for (int s = 0; s < ET_CB; s++) {
                const int tl = ET_QL * s + q;
                const int si = tl < nc ? T.sidx[tl] : -1;
                const bool mine = si >= 0;
                if (!__any(mine)) continue;
                const int g = goff + max(si, 0);
                const float x2g = mine ? x2[g] : 0.f;
                const bool hand = mine && srk[g] >= 0;
                float ev[NR];
#pragma unroll
                for (int i = 0; i < NR; i++) {
                    E[i][s] = fmaf(-d2v[i], x2g, E[i][s]);
                    ev[i] = hand && rb + 16 * i < dim ? stageH[g * RP + rb + 16 * i] : 0.f;
                }
                const float xd = et_dot<NR>(ev, dv);
                const float x = hand ? xd : 0.f;
#pragma unroll
                for (int i = 0; i < NR; i++) E[i][s] = fmaf(-dv[i], x, E[i][s]);
            }
        } else if (k >= 1 && nh > 0) {
            float dv[NR];
#pragma unroll
            for (int i = 0; i < NR; i++) dv[i] = dprev[rb + 16 * i];
            for (int c0 = 0; c0 < nh; c0 += HCH) {
                if (c0 > 0) {
                    __syncthreads();  // the previous pass's reads of the stage are done
                    stage_own(c0, 0, ET_T);
                    __syncthreads();
                    if (*bail) return;
                }
#pragma unroll
                for (int s = 0; s < ET_CB; s++) {
                    const int tl = ET_QL * s + q;
                    const int hi = tl < nc ? T.hidx[tl] - c0 : -1;
                    const bool mine = hi >= 0 && hi < HCH;
                    if (!__any(mine)) continue;
                    float ev[NR];
#pragma unroll
                    for (int i = 0; i < NR; i++)
                        ev[i] = mine && rb + 16 * i < dim ? stageH[max(hi, 0) * RP + rb + 16 * i] : 0.f;
                    const float xd = et_dot<NR>(ev, dv);
                    const float x = mine ? xd : 0.f;
#pragma unroll
                    for (int i = 0; i < NR; i++) E[i][s] = fmaf(-dv[i], x, E[i][s]);
                }
            }
        }
        // ---- the columns atom k+1's hand-over reads --------------------------
#pragma unroll
        for (int s = 0; s < ET_CB; s++) {
            const int tl = ET_QL * s + q;
            if (tl < nc && T.sx[tl]) {
                float *col = Ecur + (int64_t)(tb + tl) * dpad;
#pragma unroll
                for (int i = 0; i < NR; i++)
                    if (rb + 16 * i < dim) col[rb + 16 * i] = E[i][s];
            }
        }
        // ---- power iteration from d_k ------------------------------------------
#pragma unroll
        for (int i = 0; i < NR; i++) u[i] = d0[rb + 16 * i];
        it = 0;
        d2prev = 0.f;
        bool stop = m == 0;
        bool certified = false;
        if (sp) {
            // the speculative sum corrected for each listed column (every
            // member the same bits), lane row q taking columns q, q + 32
            for (int g0 = 0; g0 < total; g0 += ET_QL) {
                if (!__any(g0 + q < total)) continue;
                const int g = min(g0 + q, total - 1);
                const bool hand = srk[g] >= 0;
                const float x2g = x2[g];
                float en[NR], ev[NR], dv[NR];
                // atom k-1's coefficient of the column, then e = e' - shares
                // (the owner's operations on the same values)
#pragma unroll
                for (int i = 0; i < NR; i++) {
                    ev[i] = hand && rb + 16 * i < dim ? stageH[g * RP + rb + 16 * i] : 0.f;
                    dv[i] = dprev[rb + 16 * i];
                }
                const float xd = et_dot<NR>(ev, dv);
                const float xh = hand ? xd : 0.f;
#pragma unroll
                for (int i = 0; i < NR; i++) {
                    const float es = rb + 16 * i < dim ? stageS[g * RP + rb + 16 * i] : 0.f;
                    en[i] = fmaf(-dv[i], xh, fmaf(-dk2[rb + 16 * i], x2g, es));
                }
                const float an = et_dot<NR>(en, u);
#pragma unroll
                for (int i = 0; i < NR; i++) ev[i] = rb + 16 * i < dim ? stageS[g * RP + rb + 16 * i] : 0.f;
                const float as = et_dot<NR>(ev, u);
                if (g0 + q < total) {
#pragma unroll
                    for (int i = 0; i < NR; i++) cpart[g * RP + rb + 16 * i] = en[i] * an - ev[i] * as;
                }
            }
            __syncthreads();
            if (tid < dim) {
                // four partial sums over the listed columns (g mod 4), then
                // (d0 + d1) + (d2 + d3): a fixed order
                float dz[4] = {0.f, 0.f, 0.f, 0.f};
                for (int g0 = 0; g0 < total; g0 += 4) {
                    float c[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) c[j] = g0 + j < total ? cpart[(g0 + j) * RP + tid] : 0.f;
#pragma unroll
                    for (int j = 0; j < 4; j++) dz[j] += c[j];
                }
                zs[tid] = zsp[tid] + ((dz[0] + dz[1]) + (dz[2] + dz[3]));
            }
            __syncthreads();
            stop = step();
        }
        while (!stop) {
            if (!exchange(u, !certified, false, k)) return;
            certified = true;
            stop = step();
        }
        if (!certified) ps_publish(ecert + me, (unsigned)(k + 1));
        iters += (unsigned long long)it;
        // ---- d_k out ---------------------------------------------------------
        if (h == 0 && tid < 16) {
#pragma unroll
            for (int i = 0; i < NR; i++) {
                const int rr = rb + 16 * i;
                if (rr < dim) {
                    st64_c(Dp + (size_t)team * dpad + rr, ((unsigned long long)(k + 1) << 32) | __float_as_uint(u[i]));
                    a.D[(int64_t)k * dim + rr] = u[i];
                }
            }
        }
        // ---- P4: atom k's final coefficients ---------------------------------
#pragma unroll
        for (int s = 0; s < ET_CB; s++) {
            const int tl = ET_QL * s + q;
            if (!__any(tl < nc)) continue;
            float ev[NR];
#pragma unroll
            for (int i = 0; i < NR; i++) ev[i] = E[i][s];
            const float x = et_dot<NR>(ev, u);
            if (tl < nc && rb == 0) a.coef[T.ss[tl]] = x;
        }
        ps_publish(ccert + me, (unsigned)(k + 1));
        // ---- the team's next atom kn = k + 8: table, starting vector, and
        //      once atoms k-7 .. k are certified, its columns ---------------------
        const int kn = k + ET_NT;
        if (kn >= K) break;
        nc = cols_of(pa_m(suo, kn, cap), ET_HT, h, &tb);
        __syncthreads();
        et_table<NO>(a, suo, smo, kn, tb, nc, T);
        if (tid < dim) d0[tid] = a.D[(int64_t)kn * dim + tid];
        __syncthreads();
        et_lists(T, nc, wcnt);
        // thread t < 32 waits for member t % 4 of the team of atom k - 7 + t / 4
        {
            const int j = k - ET_WIN + tid / ET_HT;
            const bool w = tid < ET_NT * ET_HT && j >= 0;
            if (w && bail[1] == 0 && tid == 0) {
                st_flag(flags + PSF_STICKY, 1u);
                st_flag(flags + PSF_ABORT, 1u);
                *bail = 1;
            }
            if (w && bail[1] != 0)
                et_poll(flags, PTF_CCERT + (j % ET_NT) * ET_HT + (tid % ET_HT), 1, (unsigned)(j + 1), bail);
        }
        __syncthreads();
        if (*bail) return;
        l1_drop();
        et_build<NR, NO>(a, E, nc, T, q, rb);
    }
    if (h == 0 && tid == 0) atomicAdd(&a.counters[CNT_EIG_IT], iters);
}

template <int NR, int NO>
__global__ __launch_bounds__(ET_T) void et_sweep_kernel(SweepArgs a, unsigned *flags, unsigned long long *Wx)
{
    extern __shared__ __attribute__((aligned(16))) float sm[];
    __shared__ int s_role, s_bail[2];  // bail word, spin limit
    if (threadIdx.x == 0) {
        s_role = ps_register(flags, ET_NT * ET_HT);
        s_bail[0] = 0;
        s_bail[1] = (int)ld_flag(flags + PSF_SPIN);
    }
    __syncthreads();
    const int role = s_role;
    if (role < 0) return;
    et_member<NR, NO>(a, flags, role % ET_NT, role / ET_NT, Wx, sm, s_bail);
    if (threadIdx.x == 0 && s_bail[0] == 0) atomicAdd(&flags[PSF_FINISHED], 1u);
}

// Several pictures' eight-team sweeps in one launch, picture i on XCD i (as
// pt_multi_kernel).
template <int NR, int NO>
__global__ __launch_bounds__(ET_T) void et_multi_kernel(const PtPicture *__restrict__ pics, int n)
{
    extern __shared__ __attribute__((aligned(16))) float sm[];
    __shared__ int s_role, s_bail[2];
    const unsigned xcc = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15u;  // XCC_ID[3:0]
    if (xcc >= (unsigned)n) return;
    unsigned *flags = pics[xcc].flags;
    if (threadIdx.x == 0) {
        (void)atomicCAS(&flags[PSF_XSEL], 0u, xcc + 1u);
        const unsigned slot = atomicAdd(&flags[PSF_REG + xcc], 1u);
        s_role = slot < (unsigned)(ET_NT * ET_HT) ? (int)slot : -1;
        s_bail[0] = 0;
        s_bail[1] = (int)ld_flag(flags + PSF_SPIN);
    }
    __syncthreads();
    const int role = s_role;
    if (role < 0) return;
    const SweepArgs a = pics[xcc].a;
    et_member<NR, NO>(a, flags, role % ET_NT, role / ET_NT, pics[xcc].wx, sm, s_bail);
    if (threadIdx.x == 0 && s_bail[0] == 0) atomicAdd(&flags[PSF_FINISHED], 1u);
}

// Clears the sweep's flags (all but the sticky abort word and the role
// counts) and the nw exchange pairs before a launch, and adds the launch's
// nroles roles to the started count.
__global__ void ps_reset_kernel(unsigned *flags, unsigned long long *wx, int nw, unsigned nroles, unsigned spin)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < PSF_WORDS && i != PSF_STICKY && i != PSF_STARTED && i != PSF_FINISHED) flags[i] = i == PSF_SPIN ? spin : 0u;
    if (i == PSF_STARTED) flags[i] += nroles;
    if (i < nw) wx[i] = 0ull;
}

// ===========================================================================
// Large patches (padded dim above 192, up to 1024): three launches per atom,
// E_k and its Gram in global memory, for the patch sizes the LDS-resident
// sweeps cannot hold where the one-launch large-patch sweep below does not
// apply (dim above 512, caps above 1024 or none, KSVD_SWEEP_KL=1).
//   kl_build_kernel(k), one pass of many workgroups: E_k's capped columns,
//     e_t = y_p - sum_{j != k} d_j x_{j,p} in the column table's entry order,
//     every coefficient read from memory (atom k-1's were written, final, by
//     the previous atom's last launch; users beyond an updated atom's cap were
//     zeroed there too);
//   kl_gram_kernel(k): G = E_k E_k^T on MFMA, one workgroup per 64 x 64 tile
//     of the upper triangle (mirrored) and column range (KG_SPLIT partial
//     Grams, each summed over its columns in order);
//   kq_power_kernel(k), the workgroups of one XCD: the partial Grams added in
//     order into each role's rows in LDS, the power iteration from d_k with
//     the stop rule of the other sweeps by tagged-pair exchanges, d_k = u,
//     then x_t = e_t . d_k for the first m users and 0 for the others.
// The fallback after a persistent sweep gave up runs the power iteration in
// one workgroup (kl_power_kernel, the single Gram) and the coefficients in
// kl_coef_kernel, which wait on nothing.
constexpr int KL_T = 1024;
constexpr int KG_TB = 64, KG_KC = 64;
constexpr int KG_SPLIT = 4;  // column ranges of the large-patch Gram when the persistent power iteration reads it

__device__ __forceinline__ void kl_build_body(const SweepArgs &a, float *E, int k)
{
    int c;
    const int m = atom_m(a.uoff, k, a.cap, &c);
    const int dim = a.dim, kmax = a.kmax;
    const int no = kmax > 1 ? kmax - 1 : 1;
    const int cpt = KL_T / dim;  // columns per pass (dim <= 1024)
    const int tc = threadIdx.x / dim, te = threadIdx.x - tc * dim;
    if (tc >= cpt) return;
    const int64_t cb = a.moff[k];
    for (int t = blockIdx.x * cpt + tc; t < m; t += gridDim.x * cpt) {
        const int64_t col = cb + t;
        const int p = a.tab_p[col];
        float v = a.Y[(int64_t)p * dim + te];
        for (int i = 0; i < no; i++) {
            const int4 o = a.tab_oth[col * no + i];
            if (o.x >= 0) v = fmaf(-a.D[(int64_t)o.x * dim + te], a.coef[o.y], v);
        }
        E[(int64_t)t * dim + te] = v;
    }
}

__global__ __launch_bounds__(KL_T) void kl_build_kernel(SweepArgs a, float *E, int k) { kl_build_body(a, E, k); }

// A batch's picture in the large-patch kernels: its sweep arguments, E_k, its
// (partial) Grams and the power iteration's flags, pairs and per-atom words.
struct KlPicture {
    SweepArgs a;
    float *E, *G;
    unsigned *flags;
    unsigned long long *wx;
    unsigned *kw;
};

// picture blockIdx.z of the batch
__global__ __launch_bounds__(KL_T) void kl_build_multi(const KlPicture *__restrict__ pics, int k)
{
    const KlPicture *p = pics + blockIdx.z;
    kl_build_body(p->a, p->E, k);
}

// Tile (bi, bj), bi <= bj, of G = E E^T (dim x dim, pitch gp), four
// wavefronts of 32 x 32; the columns of E in chunks of KG_KC staged in LDS,
// the next chunk's loads in flight (registers) while the current one is on
// MFMA.  With gridDim.y = S > 1, workgroup (tile, s) sums the s-th of S
// column ranges into partial Gram s (at G + s dim gp); the reader adds the
// partials in order.
constexpr int KG_LD = (KG_KC * KG_TB) / 256;  // values of each operand a thread stages per chunk

__device__ __forceinline__ void kl_gram_body(const SweepArgs &a, const float *__restrict__ E, float *__restrict__ G,
                                             int gp, int k)
{
    __shared__ float Ea[KG_KC][KG_TB + 4];
    __shared__ float Eb[KG_KC][KG_TB + 4];
    int c;
    const int m = atom_m(a.uoff, k, a.cap, &c);
    const int dim = a.dim;
    // this workgroup's column range [tb, te)
    const int S = (int)gridDim.y, sp = (int)blockIdx.y;
    const int ck = ((m + S - 1) / S + KG_KC - 1) / KG_KC * KG_KC;
    const int tb = min(m, sp * ck), te = min(m, tb + ck);
    G += (size_t)sp * dim * gp;
    const int nb = (dim + KG_TB - 1) / KG_TB;
    // tile index -> (bi, bj) of the upper triangle, row by row
    int ti = blockIdx.x, bi = 0;
    while (ti >= nb - bi) { ti -= nb - bi; bi++; }
    const int bj = bi + ti;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 15, g = lane >> 4;
    const int wr = wave >> 1, wc = wave & 1;
    const int e0 = bi * KG_TB, f0 = bj * KG_TB;
    f4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) acc[x][y] = f4{0.f, 0.f, 0.f, 0.f};
    // thread tid stages entries q = tid + 256 i: row q / KG_TB, column q % KG_TB
    float ra[KG_LD], rb[KG_LD];
    auto fetch = [&](int t0) {
#pragma unroll
        for (int i = 0; i < KG_LD; i++) {
            const int q = tid + 256 * i, r = q / KG_TB, cc = q - r * KG_TB, t = t0 + r;
            ra[i] = (t < te && e0 + cc < dim) ? E[(int64_t)t * dim + e0 + cc] : 0.f;
            rb[i] = (t < te && f0 + cc < dim) ? E[(int64_t)t * dim + f0 + cc] : 0.f;
        }
    };
    if (te > tb) fetch(tb);
    for (int t0 = tb; t0 < te; t0 += KG_KC) {
#pragma unroll
        for (int i = 0; i < KG_LD; i++) {
            const int q = tid + 256 * i, r = q / KG_TB, cc = q - r * KG_TB;
            Ea[r][cc] = ra[i];
            Eb[r][cc] = rb[i];
        }
        __syncthreads();
        if (t0 + KG_KC < te) fetch(t0 + KG_KC);
#pragma unroll
        for (int kk = 0; kk < KG_KC; kk += 4) {
            float av[2], bv[2];
#pragma unroll
            for (int x = 0; x < 2; x++) av[x] = Ea[kk + g][wr * 32 + x * 16 + col];
#pragma unroll
            for (int y = 0; y < 2; y++) bv[y] = Eb[kk + g][wc * 32 + y * 16 + col];
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 2; y++)
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[x], bv[y], acc[x][y], 0, 0, 0);
        }
        __syncthreads();
    }
    // lane (col, g) holds G[row 4g + i][column col] of each 16 x 16 block
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int e = e0 + wr * 32 + x * 16 + 4 * g + i, f = f0 + wc * 32 + y * 16 + col;
                if (e < dim && f < dim) {
                    G[(int64_t)e * gp + f] = acc[x][y][i];
                    G[(int64_t)f * gp + e] = acc[x][y][i];
                }
            }
}

__global__ __launch_bounds__(256) void kl_gram_kernel(SweepArgs a, const float *__restrict__ E, float *__restrict__ G,
                                                      int gp, int k)
{
    kl_gram_body(a, E, G, gp, k);
}

__global__ __launch_bounds__(256) void kl_gram_multi(const KlPicture *__restrict__ pics, int gp, int k)
{
    const KlPicture *p = pics + blockIdx.z;
    kl_gram_body(p->a, p->E, p->G, gp, k);
}

// Power iteration on G (global, symmetric, row pitch gp = dim rounded up to
// 4): z = G u with thread (sl, c) summing four columns 4c .. 4c+3 over the
// rows f of slice sl by 16-byte loads (coalesced across c, KP_U rows in
// flight), the slices' partials added in order.
constexpr int KP_U = 8;

__global__ __launch_bounds__(KL_T) void kl_power_kernel(SweepArgs a, const float *__restrict__ G, int gp, int k)
{
    __shared__ float u[KSVD_MAX_DIM];
    __shared__ __attribute__((aligned(16))) float zp[4 * KL_T];  // slices x gp <= 4 KL_T
    __shared__ float red[KL_T / 64];
    int c;
    const int m = atom_m(a.uoff, k, a.cap, &c);
    if (m == 0) return;
    const int dim = a.dim, tid = threadIdx.x;
    const int ncg = gp / 4;
    const int ns = KL_T / ncg;  // slices (gp <= 1024)
    const int sl = tid / ncg, cg = tid - sl * ncg;
    const int fl = (dim + ns - 1) / ns;
    const int fb = sl * fl, fe = min(dim, fb + fl);
    if (tid < dim) u[tid] = a.D[(int64_t)k * dim + tid];
    __syncthreads();
    int it = 0;
    float d2p = 0.f;
    for (;;) {
        if (sl < ns) {
            f4 s4 = f4{0.f, 0.f, 0.f, 0.f};
            const float *gc = G + 4 * cg;
            int f = fb;
            for (; f + KP_U <= fe; f += KP_U) {
                f4 gv[KP_U];
#pragma unroll
                for (int q = 0; q < KP_U; q++) gv[q] = *(const f4 *)(gc + (int64_t)(f + q) * gp);
#pragma unroll
                for (int q = 0; q < KP_U; q++) {
                    const float uf = u[f + q];
                    s4.x = fmaf(gv[q].x, uf, s4.x);
                    s4.y = fmaf(gv[q].y, uf, s4.y);
                    s4.z = fmaf(gv[q].z, uf, s4.z);
                    s4.w = fmaf(gv[q].w, uf, s4.w);
                }
            }
            for (; f < fe; f++) {
                const f4 g4 = *(const f4 *)(gc + (int64_t)f * gp);
                const float uf = u[f];
                s4.x = fmaf(g4.x, uf, s4.x);
                s4.y = fmaf(g4.y, uf, s4.y);
                s4.z = fmaf(g4.z, uf, s4.z);
                s4.w = fmaf(g4.w, uf, s4.w);
            }
            *(f4 *)(zp + sl * gp + 4 * cg) = s4;
        }
        __syncthreads();
        float zv = 0.f;
        if (tid < dim)
            for (int q = 0; q < ns; q++) zv += zp[q * gp + tid];
        const float n2 = block_sum_t2(zv * zv, red);
        if (n2 == 0.f) break;  // E = 0: d_k stays (its coefficients become 0)
        const float un = zv / sqrtf(n2);
        const float du = tid < dim ? un - u[tid] : 0.f;
        const float d2 = block_sum_t2(du * du, red);
        it++;
        if (tid < dim) u[tid] = un;
        __syncthreads();
        if (power_stop(it, d2, d2p, a.tol2, a.est2, a.maxit)) break;
        d2p = d2;
    }
    if (tid == 0) atomicAdd(&a.counters[CNT_EIG_IT], (unsigned long long)it);
    if (tid < dim) a.D[(int64_t)k * dim + tid] = u[tid];
}

// e_t . u by one wavefront (lane-strided over dim, fixed butterfly); every
// load in flight before the first product (dim <= KSVD_MAX_DIM).
__device__ __forceinline__ float kl_dot(const float *__restrict__ e, const float *__restrict__ u, int dim)
{
    constexpr int J = KSVD_MAX_DIM / 64;
    const int lane = threadIdx.x & 63;
    float ev[J];
#pragma unroll
    for (int j = 0; j < J; j++) ev[j] = lane + 64 * j < dim ? e[lane + 64 * j] : 0.f;
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < J; j++)
        if (lane + 64 * j < dim) s = fmaf(ev[j], u[lane + 64 * j], s);
    return wave_sum_dpp(s);
}

// Power iteration on G by the KQ_R workgroups of one XCD (one launch per
// atom): the first workgroup to arrive chooses its XCD for this atom (word
// kw[k]), the others take roles only there (counter kw[K + k]) and the rest
// leave.  Role r holds rows [r rpr, (r + 1) rpr) of G in LDS.  Each iteration
// every role forms its rows of z = G u (a wavefront per row, lane-strided,
// fixed butterfly) and stores them as (tag, value) pairs, tag = (k << 12) |
// iteration, into one of two buffers by the iteration's parity; every role then
// reads all dim pairs, rereading one until it carries the tag, and forms |z|,
// u and the stop decision from the same values in the same order, so all roles
// hold the same u bit for bit.  A role reads iteration i's pairs before it
// stores those of i + 1, so nobody stores i + 2 while i may still be read.  The
// pairs and the per-atom words are cleared before each sweep.  The waits give
// up as the persistent sweeps' do (sticky abort; started and finished counts).
constexpr int KQ_R = 32, KQ_T = 1024;

// Sum over the workgroup (fixed order: DPP within each wavefront, then the
// wavefronts in order), returned to every thread.
__device__ __forceinline__ float kq_block_sum(float v, float *red)
{
    v = wave_sum_dpp(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float s = 0.f;
#pragma unroll
    for (int w = 0; w < KQ_T / 64; w++) s += red[w];
    __syncthreads();
    return s;
}

// The power iteration and coefficients of one role (see kq_power_kernel).
__device__ __forceinline__ void kq_power_body(const SweepArgs &a, const float *__restrict__ G, int gp, int nsplit,
                                              const float *__restrict__ E, int k, unsigned *flags,
                                              unsigned long long *wx, int role, int *s_bail, float *gs, float *u,
                                              float *red)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int c;
    const int m = atom_m(a.uoff, k, a.cap, &c);
    const int dim = a.dim;
    const int rpr = (dim + KQ_R - 1) / KQ_R;
    const int r0 = min(dim, role * rpr), r1 = min(dim, r0 + rpr), nr = r1 - r0;
    int it = 0;
    if (m > 0) {
        // this role's rows, the partial Grams added in order
        const int g4 = gp / 4;
        const size_t pstride = (size_t)dim * gp;
        for (int q = tid; q < nr * g4; q += KQ_T) {
            const int r = q / g4, c4 = q - r * g4;
            const float *src = G + (int64_t)(r0 + r) * gp + 4 * c4;
            f4 v = *(const f4 *)src;
            for (int sp = 1; sp < nsplit; sp++) {
                const f4 w = *(const f4 *)(src + sp * pstride);
                v.x += w.x;
                v.y += w.y;
                v.z += w.z;
                v.w += w.w;
            }
            *(f4 *)(gs + r * gp + 4 * c4) = v;
        }
        if (tid < dim) u[tid] = a.D[(int64_t)k * dim + tid];
        __syncthreads();
        float d2p = 0.f;
        for (;;) {
            const unsigned tag = ((unsigned)k << 12) | (unsigned)(it + 1);
            unsigned long long *W = wx + (size_t)(it & 1) * gp;
            // this role's rows of z = G u
            for (int r = wave; r < nr; r += KQ_T / 64) {
                float z = 0.f;
                for (int f = lane; f < dim; f += 64) z = fmaf(gs[r * gp + f], u[f], z);
                z = wave_sum_dpp(z);
                if (lane == 0) st64_c(W + r0 + r, ((unsigned long long)tag << 32) | __float_as_uint(z));
            }
            // every role's rows (a spin limit of 0, tests only, gives up at once)
            float zv = 0.f;
            if (s_bail[1] == 0 && tid == 0) {
                st_flag(flags + PSF_STICKY, 1u);
                st_flag(flags + PSF_ABORT, 1u);
                s_bail[0] = 1;
            }
            if (tid < dim && s_bail[1] != 0) {
                unsigned long long pv = ld64_c(W + tid);
                long spins = 0;
                while ((unsigned)(pv >> 32) != tag) {
                    __builtin_amdgcn_s_sleep(1);
                    if ((++spins & 31) == 0) {
                        if (ld_flag(flags + PSF_ABORT)) {
                            s_bail[0] = 1;
                            break;
                        }
                        if (spins > (long)s_bail[1]) {
                            st_flag(flags + PSF_STICKY, 1u);
                            st_flag(flags + PSF_ABORT, 1u);
                            s_bail[0] = 1;
                            break;
                        }
                    }
                    pv = ld64_c(W + tid);
                }
                zv = __uint_as_float((unsigned)pv);
            }
            const float n2 = kq_block_sum(zv * zv, red);
            if (s_bail[0]) return;
            if (n2 == 0.f) break;  // E = 0: d_k stays (its coefficients become 0)
            const float un = zv / sqrtf(n2);
            const float du = tid < dim ? un - u[tid] : 0.f;
            const float d2 = kq_block_sum(du * du, red);
            it++;
            if (tid < dim) u[tid] = un;
            __syncthreads();
            if (power_stop(it, d2, d2p, a.tol2, a.est2, a.maxit)) break;
            d2p = d2;
        }
        if (role == 0 && tid < dim) a.D[(int64_t)k * dim + tid] = u[tid];
        // the new coefficients x_t = e_t . d_k of the first m users, a
        // wavefront per column over all roles
        const int64_t cb = a.moff[k];
        for (int t = role * (KQ_T / 64) + wave; t < m; t += KQ_R * (KQ_T / 64)) {
            const float x = kl_dot(E + (int64_t)t * dim, u, dim);
            if (lane == 0) a.coef[a.tab_self[cb + t]] = x;
        }
    }
    // users beyond the cap: coefficient 0
    const int base = a.uoff[k];
    for (int t = m + role * KQ_T + tid; t < c; t += KQ_R * KQ_T)
        a.coef[(int64_t)a.users[base + t] * a.kmax + a.uslot[base + t]] = 0.f;
    if (tid == 0) {
        if (role == 0) atomicAdd(&a.counters[CNT_EIG_IT], (unsigned long long)it);
        atomicAdd(&flags[PSF_FINISHED], 1u);
    }
}

__global__ __launch_bounds__(KQ_T) void kq_power_kernel(SweepArgs a, const float *__restrict__ G, int gp, int nsplit,
                                                        const float *__restrict__ E, int k, unsigned *flags,
                                                        unsigned long long *wx, unsigned *kw)
{
    extern __shared__ __attribute__((aligned(16))) float gs[];  // rpr x gp
    __shared__ float u[KSVD_MAX_DIM];
    __shared__ float red[KQ_T / 64];
    __shared__ int s_role, s_bail[2];
    if (threadIdx.x == 0) {
        const unsigned xcc = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15u;  // XCC_ID[3:0]
        const unsigned old = atomicCAS(&kw[k], 0u, xcc + 1u);
        int role = -1;
        if ((old == 0u ? xcc + 1u : old) == xcc + 1u) {
            const unsigned r = atomicAdd(&kw[a.K + k], 1u);
            if (r < (unsigned)KQ_R) role = (int)r;
        }
        s_role = role;
        s_bail[0] = 0;
        s_bail[1] = (int)ld_flag(flags + PSF_SPIN);
    }
    __syncthreads();
    if (s_role < 0) return;
    kq_power_body(a, G, gp, nsplit, E, k, flags, wx, s_role, s_bail, gs, u, red);
}

// Several pictures' power iterations in one launch, picture i on XCD i (its
// roles counted in its own per-atom word).
__global__ __launch_bounds__(KQ_T) void kq_power_multi(const KlPicture *__restrict__ pics, int n, int gp, int nsplit,
                                                       int k)
{
    extern __shared__ __attribute__((aligned(16))) float gs[];  // rpr x gp
    __shared__ float u[KSVD_MAX_DIM];
    __shared__ float red[KQ_T / 64];
    __shared__ int s_role, s_bail[2];
    const unsigned xcc = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15u;  // XCC_ID[3:0]
    if (xcc >= (unsigned)n) return;
    const KlPicture *p = pics + xcc;
    if (threadIdx.x == 0) {
        const unsigned r = atomicAdd(&p->kw[p->a.K + k], 1u);
        s_role = r < (unsigned)KQ_R ? (int)r : -1;
        s_bail[0] = 0;
        s_bail[1] = (int)ld_flag(p->flags + PSF_SPIN);
    }
    __syncthreads();
    if (s_role < 0) return;
    kq_power_body(p->a, p->G, gp, nsplit, p->E, k, p->flags, p->wx, s_role, s_bail, gs, u, red);
}

// ===========================================================================
// Large patches in one launch (padded dim above 192, up to 512; caps up to
// 1024): the KX_R workgroups of one XCD form one team that takes the atoms in
// order.  Role r holds its share of atom k's capped columns (at most KX_CW) in
// LDS, built from the column table as kl_build_kernel builds them (same
// arithmetic, same bits), and the power iteration runs on E_k itself, with no
// Gram: each iteration every role forms its partial z_r = E_r (E_r^T u) (the
// column dots lane-strided with a fixed butterfly, then each row summed over
// the role's columns in order) and stores it as tagged pairs, tag = (k << 12)
// | iteration, into one of two buffers by the iteration's parity; every role
// reads all KX_R partials of each row and adds them in role order, so every
// role holds the same z, u and stop decision bit for bit.  A role reads
// iteration i's pairs before it stores those of i + 1, and the roles meet
// between atoms, so no pair is overwritten while it may still be read.  When
// the iteration stops, role 0 writes d_k = u, every role writes its columns'
// coefficients x_t = e_t . d_k (kl_dot's arithmetic) and zeroes its share of
// the users beyond the cap, then publishes its done flag; atom k+1's columns
// are built once every role is done with atom k (they read atom k's vector
// and coefficients).  Waits give up as the other persistent sweeps' do.
constexpr int KX_R = 32;     // roles: one XCD's CUs
constexpr int KX_T = 1024;   // threads per role
constexpr int KX_CW = 32;    // columns per role at most (caps up to KX_R * KX_CW)
constexpr int KX_NO = 8;     // other support entries per column staged in LDS (ompIter <= 9)
constexpr int KX_B = 8;      // columns whose loads are in flight together in the build

// LDS of one role: its columns (pitch ep = dim rounded up to 4), then the
// column table entries and coefficients of the build.
__host__ __device__ constexpr int kx_ep(int dim) { return (dim + 3) / 4 * 4; }
__host__ __device__ constexpr size_t kx_lds_bytes(int dim)
{
    return sizeof(float) * (size_t)KX_CW * kx_ep(dim) + sizeof(int) * KX_CW * (2 + KX_NO) + sizeof(float) * KX_CW * KX_NO;
}

__device__ void kx_team(const SweepArgs &a, unsigned *flags, unsigned long long *wx, int role, float *sm, int *bail)
{
    int tid = threadIdx.x, lane = tid & 63;
    int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int K = a.K, cap = a.cap, kmax = a.kmax;
    int dim = a.dim, dpad = a.dpad;
    const int no = kmax > 1 ? kmax - 1 : 1;
    int ep = kx_ep(dim);
    float *Es = sm;                                // KX_CW x ep: this role's columns
    int *sp = (int *)(Es + (size_t)KX_CW * ep);    // KX_CW: patch of each column
    int *sself = sp + KX_CW;                       // KX_CW: the column's own coefficient slot
    int *sat = sself + KX_CW;                      // KX_CW x KX_NO: other atoms of the column (-1: none)
    float *sx = (float *)(sat + KX_CW * KX_NO);    // KX_CW x KX_NO: their coefficients
    __shared__ float u[KSVD_MAX_DIM];
    __shared__ float vc[KX_CW];
    __shared__ float red[KX_T / 64];
    // the build: column group g of cg takes rows r of columns g + cg j
    int cg = KX_T / dim;
    int g = tid / dim, r = tid - g * dim;
    unsigned long long iters = 0;
    for (int k = 0; k < K; k++) {
        // the thread coordinates and sizes pass through an opaque copy each
        // atom, so that the addresses derived from them are formed where they
        // are used instead of being held (and spilled) across the loop
        asm volatile("" : "+s"(dim), "+s"(dpad), "+s"(ep), "+s"(cg), "+s"(wave));
        asm volatile("" : "+v"(tid), "+v"(lane), "+v"(g), "+v"(r));
        int c;
        const int m = atom_m(a.uoff, k, cap, &c);
        const int cw = (m + KX_R - 1) / KX_R;
        const int tb = min(m, role * cw);
        const int nc = min(m, tb + cw) - tb;
        const int64_t cb = a.moff[k] + tb;
        // ---- this role's columns of E_k: e = y - sum_j d_j x_j in the column
        //      table's entry order (kl_build_kernel's sums) -------------------------
        if (tid < nc) {
            sp[tid] = a.tab_p[cb + tid];
            sself[tid] = a.tab_self[cb + tid];
        }
        for (int q = tid; q < nc * KX_NO; q += KX_T) {
            const int t = q / KX_NO, i = q - t * KX_NO;
            int at = -1;
            float xv = 0.f;
            if (i < no) {
                const int4 o = a.tab_oth[(cb + t) * no + i];
                if (o.x >= 0) {
                    at = o.x;
                    xv = a.coef[o.y];
                }
            }
            sat[q] = at;
            sx[q] = xv;
        }
        if (tid < dim) u[tid] = a.D[(int64_t)k * dim + tid];
        __syncthreads();
        if (g < cg) {
            for (int t0 = g; t0 < nc; t0 += cg * KX_B) {
                float v[KX_B];
#pragma unroll
                for (int j = 0; j < KX_B; j++) v[j] = a.Y[(int64_t)sp[min(t0 + j * cg, nc - 1)] * dim + r];
                // every entry's atom row loaded without a branch (row 0 for
                // empty entries, whose term is then not taken), so that all
                // the loads are in flight together
#pragma unroll
                for (int i = 0; i < KX_NO; i++) {
                    if (i >= no) break;
                    float dv[KX_B];
#pragma unroll
                    for (int j = 0; j < KX_B; j++) {
                        const int t = min(t0 + j * cg, nc - 1);  // a column of this atom's table
                        const int at = t0 + j * cg < nc ? sat[t * KX_NO + i] : -1;
                        dv[j] = a.D[(int64_t)max(at, 0) * dim + r];
                    }
#pragma unroll
                    for (int j = 0; j < KX_B; j++) {
                        const int t = min(t0 + j * cg, nc - 1);
                        const int at = t0 + j * cg < nc ? sat[t * KX_NO + i] : -1;
                        const float nv = fmaf(-dv[j], sx[t * KX_NO + i], v[j]);
                        v[j] = at >= 0 ? nv : v[j];
                    }
                }
#pragma unroll
                for (int j = 0; j < KX_B; j++) {
                    const int t = t0 + j * cg;
                    if (t < nc) Es[t * ep + r] = v[j];
                }
            }
        }
        __syncthreads();
        // ---- power iteration from d_k -------------------------------------------
        int it = 0;
        float d2p = 0.f;
        if (m > 0) {
            for (;;) {
                // the column dots e_t . u, a wavefront per column
                for (int t = wave; t < nc; t += KX_T / 64) {
                    float s = 0.f;
                    for (int f = lane; f < dim; f += 64) s = fmaf(Es[t * ep + f], u[f], s);
                    s = wave_sum_dpp(s);
                    if (lane == 0) vc[t] = s;
                }
                __syncthreads();
                const unsigned tag = ((unsigned)k << 12) | (unsigned)(it + 1);
                unsigned long long *W = wx + (size_t)(it & 1) * KX_R * dpad;      // stage 1: partials
                unsigned long long *Z = wx + (size_t)2 * KX_R * dpad + (size_t)(it & 1) * dpad;  // stage 2: rows
                float zv = 0.f;
                if (tid < dim) {
                    float z = 0.f;
                    for (int t = 0; t < nc; t++) z = fmaf(Es[t * ep + tid], vc[t], z);
                    st64_c(W + (size_t)role * dpad + tid, ((unsigned long long)tag << 32) | __float_as_uint(z));
                }
                if (bail[1] == 0 && tid == 0) {
                    st_flag(flags + PSF_STICKY, 1u);
                    st_flag(flags + PSF_ABORT, 1u);
                    bail[0] = 1;
                }
                // stage 1: this role's rows [role rpr, role rpr + rpr), lane
                // group (row, source role) of 32 lanes each; the 32 partials
                // added by a fixed butterfly, the sum stored as a tagged pair
                const int rpr = (dim + KX_R - 1) / KX_R;
                const int row = role * rpr + (tid >> 5), src = tid & 31;
                if (tid < KX_R * rpr && bail[1] != 0) {
                    float pz = 0.f;
                    if (row < dim) {
                        const unsigned long long *pp = W + (size_t)src * dpad + row;
                        unsigned long long pv = ld64_c(pp);
                        long spins = 0;
                        while ((unsigned)(pv >> 32) != tag) {
                            __builtin_amdgcn_s_sleep(1);
                            if ((++spins & 31) == 0) {
                                if (ld_flag(flags + PSF_ABORT)) {
                                    bail[0] = 1;
                                    break;
                                }
                                if (spins > spin_limit(flags)) {
                                    st_flag(flags + PSF_STICKY, 1u);
                                    st_flag(flags + PSF_ABORT, 1u);
                                    bail[0] = 1;
                                    break;
                                }
                            }
                            pv = ld64_c(pp);
                        }
                        pz = __uint_as_float((unsigned)pv);
                    }
#pragma unroll
                    for (int off = 1; off < 32; off <<= 1) pz += __shfl_xor(pz, off);
                    if (src == 0 && row < dim) st64_c(Z + row, ((unsigned long long)tag << 32) | __float_as_uint(pz));
                }
                // stage 2: every row's sum
                if (tid < dim && bail[1] != 0) {
                    const unsigned long long *pp = Z + tid;
                    unsigned long long pv = ld64_c(pp);
                    long spins = 0;
                    while ((unsigned)(pv >> 32) != tag) {
                        __builtin_amdgcn_s_sleep(1);
                        if ((++spins & 31) == 0) {
                            if (ld_flag(flags + PSF_ABORT)) {
                                bail[0] = 1;
                                break;
                            }
                            if (spins > spin_limit(flags)) {
                                st_flag(flags + PSF_STICKY, 1u);
                                st_flag(flags + PSF_ABORT, 1u);
                                bail[0] = 1;
                                break;
                            }
                        }
                        pv = ld64_c(pp);
                    }
                    zv = __uint_as_float((unsigned)pv);
                }
                const float n2 = kq_block_sum(zv * zv, red);
                if (bail[0]) return;
                if (n2 == 0.f) break;  // E = 0: d_k stays (its coefficients become 0)
                const float un = zv / sqrtf(n2);
                const float du = tid < dim ? un - u[tid] : 0.f;
                const float d2 = kq_block_sum(du * du, red);
                it++;
                if (tid < dim) u[tid] = un;
                __syncthreads();
                if (power_stop(it, d2, d2p, a.tol2, a.est2, a.maxit)) break;
                d2p = d2;
            }
            if (role == 0 && tid < dim) a.D[(int64_t)k * dim + tid] = u[tid];
            // the new coefficients of this role's columns
            for (int t = wave; t < nc; t += KX_T / 64) {
                float s = 0.f;
                for (int f = lane; f < dim; f += 64) s = fmaf(Es[t * ep + f], u[f], s);
                s = wave_sum_dpp(s);
                if (lane == 0) a.coef[sself[t]] = s;
            }
        }
        iters += (unsigned long long)it;
        // users beyond the cap: coefficient 0
        const int base = a.uoff[k];
        for (int t = m + role * KX_T + tid; t < c; t += KX_R * KX_T)
            a.coef[(int64_t)a.users[base + t] * kmax + a.uslot[base + t]] = 0.f;
        // ---- every role done with atom k before atom k+1's columns ------------
        ps_publish(flags + PSF_DONE + role, (unsigned)(k + 1));
        if (k + 1 < K) {
            if (!ps_wait(flags, PSF_DONE, KX_R, (unsigned)(k + 1), bail)) return;
            l1_drop();
        }
    }
    if (role == 0 && tid == 0) atomicAdd(&a.counters[CNT_EIG_IT], iters);
}

__global__ __launch_bounds__(KX_T) void kx_sweep_kernel(SweepArgs a, unsigned *flags, unsigned long long *wx)
{
    extern __shared__ __attribute__((aligned(16))) float sm[];
    __shared__ int s_role, s_bail[2];  // bail word, spin limit
    if (threadIdx.x == 0) {
        s_role = ps_register(flags, KX_R);
        s_bail[0] = 0;
        s_bail[1] = (int)ld_flag(flags + PSF_SPIN);
    }
    __syncthreads();
    const int role = s_role;
    if (role < 0) return;
    kx_team(a, flags, wx, role, sm, s_bail);
    if (threadIdx.x == 0 && s_bail[0] == 0) atomicAdd(&flags[PSF_FINISHED], 1u);
}

// Several pictures' large-patch sweeps in one launch, picture i on XCD i (its
// own flags and pairs), each the sweep above bit for bit.
__global__ __launch_bounds__(KX_T) void kx_multi_kernel(const PtPicture *__restrict__ pics, int n)
{
    extern __shared__ __attribute__((aligned(16))) float sm[];
    __shared__ int s_role, s_bail[2];  // bail word, spin limit
    const unsigned xcc = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15u;  // XCC_ID[3:0]
    if (xcc >= (unsigned)n) return;
    unsigned *flags = pics[xcc].flags;
    if (threadIdx.x == 0) {
        (void)atomicCAS(&flags[PSF_XSEL], 0u, xcc + 1u);
        const unsigned slot = atomicAdd(&flags[PSF_REG + xcc], 1u);
        s_role = slot < (unsigned)KX_R ? (int)slot : -1;
        s_bail[0] = 0;
        s_bail[1] = (int)ld_flag(flags + PSF_SPIN);
    }
    __syncthreads();
    const int role = s_role;
    if (role < 0) return;
    const SweepArgs a = pics[xcc].a;
    kx_team(a, flags, pics[xcc].wx, role, sm, s_bail);
    if (threadIdx.x == 0 && s_bail[0] == 0) atomicAdd(&flags[PSF_FINISHED], 1u);
}

__global__ __launch_bounds__(256) void kl_coef_kernel(SweepArgs a, const float *__restrict__ E, int k)
{
    __shared__ float u[KSVD_MAX_DIM];
    int c;
    const int m = atom_m(a.uoff, k, a.cap, &c);
    const int dim = a.dim, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int e = threadIdx.x; e < dim; e += 256) u[e] = a.D[(int64_t)k * dim + e];
    __syncthreads();
    const int64_t cb = a.moff[k];
    for (int t = blockIdx.x * 4 + wave; t < m; t += gridDim.x * 4) {
        const float x = kl_dot(E + (int64_t)t * dim, u, dim);
        if (lane == 0) a.coef[a.tab_self[cb + t]] = x;
    }
    // users beyond the cap: coefficient 0
    const int base = a.uoff[k];
    for (int t = m + blockIdx.x * 256 + threadIdx.x; t < c; t += gridDim.x * 256)
        a.coef[(int64_t)a.users[base + t] * a.kmax + a.uslot[base + t]] = 0.f;
}

}  // namespace

static bool env_flag(const char *name)
{
    const char *v = getenv(name);
    return v && v[0] && v[0] != '0';
}

static bool env_flag_default(const char *name, bool dflt)
{
    const char *v = getenv(name);
    return v && v[0] ? v[0] != '0' : dflt;
}

int flush_cap_zero(ksvd_ctx *ctx)
{
    if (!ctx->cap_zero_pending) return KSVD_OK;
    SweepArgs a = {};
    a.coef = (float *)ctx->coef.ptr;
    a.uoff = (const int *)ctx->uoff.ptr;
    a.users = (const int *)ctx->users.ptr;
    a.uslot = (const int *)ctx->uslot.ptr;
    a.kmax = ctx->kmax;
    a.cap = ctx->prm.user_cap;
    hipLaunchKernelGGL(pa_zero_kernel, dim3(ctx->K, 16), dim3(256), 0, ctx->stream, a);
    KSVD_HIP_CHECK(ctx, hipGetLastError());
    ctx->cap_zero_pending = false;
    return KSVD_OK;
}

// The sweep's view of a context's buffers.
static SweepArgs sweep_args(ksvd_ctx *ctx, float tol)
{
    const int dim = ctx->dim;
    const int dpad = (dim + 15) / 16 * 16;
    const int nbt = dpad / 16;
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
    a.K = ctx->K;
    a.kmax = ctx->kmax;
    a.cap = ctx->prm.user_cap;
    a.nbt = nbt;
    a.ntiles = nbt * (nbt + 1) / 2;
    a.maxit = 300;
    a.tol2 = tol * tol;
    a.est2 = 1e-12f;  // estimated error of the kept atom below 1e-6
    // KSVD_SWEEP_SPEC=0 (tests): the eight-team sweep's plain path for every atom
    a.spec = env_flag_default("KSVD_SWEEP_SPEC", true) ? 1 : 0;
    return a;
}

// LDS bytes of the two-team sweep for this context's sizes
static size_t pt_lds_bytes(const ksvd_ctx *ctx)
{
    const size_t dpad = (size_t)(ctx->dim + 15) / 16 * 16;
    return sizeof(float) * (2 * (size_t)PS_CWMAX * dpad + (3 + PT_W + PT_GRP) * dpad) +
           sizeof(int) * (2 * (size_t)PT_TABW + 2 * ((size_t)ctx->K + 1));
}

// Rows per thread of the eight-team sweep (16 NR >= padded dim): 4 or 9.
static int et_nr(int dim) { return dim <= 64 ? 4 : 9; }

// Bytes of the eight-team sweep's global scratch (et_pairs_n and the
// members' lists and listed columns).
static size_t et_scratch_bytes(int dim)
{
    const int dpad = (dim + 15) / 16 * 16;
    return sizeof(unsigned long long) * et_pairs_n(dpad) + sizeof(float) * (size_t)ET_NT * ET_HT * ET_LMAX * dpad +
           sizeof(int) * (size_t)ET_NT * ET_HT * ET_LM;
}

// LDS bytes of the eight-team sweep for this context's sizes
static size_t et_lds_bytes(const ksvd_ctx *ctx)
{
    const int nr = et_nr(ctx->dim);
    const size_t rp = 16 * (size_t)nr, hch = (size_t)et_hch(nr);
    const size_t no = ctx->kmax > 5 ? 7 : 4;
    return sizeof(float) * ((ET_W + 6 + 3 * hch) * rp + hch * no) +
           sizeof(int) * (2 * ((size_t)ctx->K + 1) + (9 + 3 * no) * (size_t)ET_COLS + 2 + ET_HT + hch * (1 + no) + 8);
}

// Whether the eight-team sweep applies: KSVD_SWEEP_TEAMS=2 selects the
// two-team sweep instead; cap up to 1024 (256 columns per member), OMP limit
// up to 8, dim up to 144 (at 16 x 12 patches its register-held columns spill
// and the two-team sweep is faster), its LDS within 150 KB.
static bool et_applies(const ksvd_ctx *ctx, bool two_kernel)
{
    const char *tv = getenv("KSVD_SWEEP_TEAMS");
    if (tv && tv[0] && atoi(tv) == 2) return false;
    const int cap = ctx->prm.user_cap;
    return cap > 0 && cap <= ET_HT * ET_COLS && ctx->kmax <= PS_NOMAX + 1 && ctx->dim <= 144 && !two_kernel &&
           et_lds_bytes(ctx) <= 150 * 1024 && et_scratch_bytes(ctx->dim) <= ctx->gpart.bytes;
}

typedef void (*EtFn)(SweepArgs, unsigned *, unsigned long long *);
typedef void (*EmFn)(const PtPicture *, int);

static EtFn et_kernel(const ksvd_ctx *ctx)
{
    const bool wide = ctx->kmax > 5;
    if (et_nr(ctx->dim) == 4) return wide ? et_sweep_kernel<4, 7> : et_sweep_kernel<4, 4>;
    return wide ? et_sweep_kernel<9, 7> : et_sweep_kernel<9, 4>;
}

static EmFn et_multi(const ksvd_ctx *ctx)
{
    const bool wide = ctx->kmax > 5;
    if (et_nr(ctx->dim) == 4) return wide ? et_multi_kernel<4, 7> : et_multi_kernel<4, 4>;
    return wide ? et_multi_kernel<9, 7> : et_multi_kernel<9, 4>;
}

// Whether the two-team sweep applies to this context (cap up to 1024, OMP
// limit up to 8, padded dim up to 192, its LDS within 160 KB).
static bool pt_applies(const ksvd_ctx *ctx, bool two_kernel)
{
    const int cap = ctx->prm.user_cap, nbt = (ctx->dim + 15) / 16;
    return cap > 0 && cap <= PT_HT * PS_CWMAX && ctx->kmax <= PS_NOMAX + 1 && nbt * (nbt + 1) / 2 <= 16 * MAX_TPW &&
           !two_kernel && pt_lds_bytes(ctx) <= 160 * 1024;
}

// Where the large-patch kernels keep their data in a context's partial-Gram
// buffer: the Gram (nsplit partials, row pitch gp), then the power
// iteration's pairs (2 gp) and per-atom words (K + K), cleared as nw 64-bit
// words; and the power iteration's LDS.  False when it does not fit.
struct KlLayout {
    int gp, nsplit, nw;
    size_t lds_q;
    float *G;
    unsigned long long *wx;
    unsigned *kw;
};

static bool kl_layout(const ksvd_ctx *ctx, int nsplit, KlLayout *L)
{
    const int dim = ctx->dim;
    L->gp = (dim + 3) / 4 * 4;
    L->nsplit = nsplit;
    L->G = (float *)ctx->gpart.ptr;
    unsigned long long *wx = (unsigned long long *)(L->G + (size_t)nsplit * dim * L->gp + 4);
    L->wx = (unsigned long long *)(((uintptr_t)wx + 15) & ~(uintptr_t)15);
    L->kw = (unsigned *)(L->wx + 2 * (size_t)L->gp);
    L->nw = 2 * L->gp + ctx->K;
    L->lds_q = sizeof(float) * (size_t)((dim + KQ_R - 1) / KQ_R) * L->gp;
    return dim <= KSVD_MAX_DIM && (char *)(L->wx + L->nw) <= (char *)ctx->gpart.ptr + ctx->gpart.bytes &&
           L->lds_q <= 150 * 1024;
}

// Whether the one-launch large-patch sweep applies (padded dim above 192 up
// to 512, caps up to 1024, its support entries staged, its pairs within the
// partial-Gram buffer); KSVD_SWEEP_KL=1 selects the three launches per atom.
static bool kx_applies(const ksvd_ctx *ctx, bool two_kernel)
{
    const int dpad = (ctx->dim + 15) / 16 * 16;
    return !two_kernel && !env_flag("KSVD_SWEEP_KL") && ctx->prm.user_cap > 0 &&
           ctx->prm.user_cap <= KX_R * KX_CW && ctx->dim <= KX_T / 2 && ctx->kmax - 1 <= KX_NO &&
           sizeof(unsigned long long) * 2 * (KX_R + 1) * (size_t)dpad <= ctx->gpart.bytes &&
           kx_lds_bytes(ctx->dim) <= 120 * 1024;
}

int launch_sweep(ksvd_ctx *ctx, float tol)
{
    ctx->sweep_defers_zero = false;
    const int dim = ctx->dim, K = ctx->K, kmax = ctx->kmax;
    SweepArgs a = sweep_args(ctx, tol);
    const int dpad = a.dpad;
    const int ntiles = a.ntiles;
    const bool large = ntiles > 16 * MAX_TPW;  // dpad > 192: the LDS-resident sweeps cannot hold it
    const int tpw = (ntiles + 15) / 16;
    const bool wide = kmax > 5;

    // KSVD_SWEEP_SPIN (tests): fewer polls before a persistent sweep's wait
    // gives up, so that the abort and the two-kernel fallback can be exercised
    unsigned spin = (unsigned)PS_SPIN_MAX;
    if (const char *sv = getenv("KSVD_SWEEP_SPIN")) spin = (unsigned)std::max(0, atoi(sv));
    const bool two_kernel = ctx->force_two_kernel || env_flag("KSVD_SWEEP_TWO_KERNEL");

    // ---- large patches: E in global memory, four kernels per atom ------------
    auto launch_large = [&]() -> int {
        if (dim > KSVD_MAX_DIM) return ksvd_fail(ctx, KSVD_ERR_UNSUPPORTED, "patch dimension above 1024");
        if (kx_applies(ctx, two_kernel)) {
            // one launch: the columns in the LDS of one XCD's workgroups
            unsigned *flags = (unsigned *)ctx->ticket.ptr;
            unsigned long long *wx = (unsigned long long *)a.Gpart;
            const int nw = 2 * (KX_R + 1) * dpad;
            KSVD_HIP_CHECK(ctx, allow_max_lds((const void *)kx_sweep_kernel));
            hipLaunchKernelGGL(ps_reset_kernel, dim3((std::max(nw, (int)PSF_WORDS) + 255) / 256), dim3(256), 0,
                               ctx->stream, flags, wx, nw, (unsigned)KX_R, spin);
            hipLaunchKernelGGL(kx_sweep_kernel, dim3(ctx->n_xcd * KX_R), dim3(KX_T), kx_lds_bytes(dim), ctx->stream, a,
                               flags, wx);
            KSVD_HIP_CHECK(ctx, hipGetLastError());
            return KSVD_OK;
        }
        float *E = (float *)ctx->E.ptr;  // m x dim, m <= min(users, cap): the E buffer's columns
        const int nb = (dim + KG_TB - 1) / KG_TB;
        const int ntg = nb * (nb + 1) / 2;
        // enough workgroups for one pass over the most columns an atom can
        // have (its capped users; every patch without a cap)
        const int64_t ecols = a.cap > 0 ? a.cap : ctx->P;
        const int cpt = KL_T / dim;
        const unsigned gb = (unsigned)std::min<int64_t>(4096, std::max<int64_t>(1, (ecols + cpt - 1) / cpt));
        const unsigned gc = (unsigned)std::min<int64_t>(4096, std::max<int64_t>(1, (ecols + 3) / 4));
        // the power iteration on one XCD's workgroups (their pairs and per-atom
        // words after G), or, for the fallback, on one workgroup
        const bool persistent = !two_kernel;
        // the persistent power iteration reads the Gram once per atom, so the
        // Gram is split over KG_SPLIT column ranges (more workgroups on MFMA)
        KlLayout L;
        const bool fits = kl_layout(ctx, persistent ? KG_SPLIT : 1, &L);
        const int gp = L.gp, nsplit = L.nsplit, nw = L.nw;
        const size_t lds_q = L.lds_q;
        float *G = L.G;
        unsigned *flags = (unsigned *)ctx->ticket.ptr;
        unsigned long long *wx = L.wx;
        unsigned *kw = L.kw;
        if (!persistent && sizeof(float) * (size_t)dim * gp > ctx->gpart.bytes)
            return ksvd_fail(ctx, KSVD_ERR_UNSUPPORTED, "sweep: Gram buffer too small");
        if (persistent) {
            if (!fits) return ksvd_fail(ctx, KSVD_ERR_UNSUPPORTED, "sweep: large-patch power iteration does not fit");
            KSVD_HIP_CHECK(ctx, allow_max_lds((const void *)kq_power_kernel));
            hipLaunchKernelGGL(ps_reset_kernel, dim3((std::max(nw, (int)PSF_WORDS) + 255) / 256), dim3(256), 0,
                               ctx->stream, flags, wx, nw, (unsigned)(K * KQ_R), spin);
        }
        for (int k = 0; k < K; k++) {
            hipLaunchKernelGGL(kl_build_kernel, dim3(gb), dim3(KL_T), 0, ctx->stream, a, E, k);
            hipLaunchKernelGGL(kl_gram_kernel, dim3(ntg, nsplit), dim3(256), 0, ctx->stream, a, (const float *)E, G, gp,
                               k);
            if (persistent) {
                // power iteration and the new coefficients
                hipLaunchKernelGGL(kq_power_kernel, dim3(ctx->n_xcd * KQ_R), dim3(KQ_T), lds_q, ctx->stream, a,
                                   (const float *)G, gp, nsplit, (const float *)E, k, flags, wx, kw);
            } else {
                hipLaunchKernelGGL(kl_power_kernel, dim3(1), dim3(KL_T), 0, ctx->stream, a, (const float *)G, gp, k);
                hipLaunchKernelGGL(kl_coef_kernel, dim3(gc), dim3(256), 0, ctx->stream, a, (const float *)E, k);
            }
        }
        KSVD_HIP_CHECK(ctx, hipGetLastError());
        return KSVD_OK;
    };
    if (large) return launch_large();
    // ---- persistent one-XCD sweep -------------------------------------------
    const int cap = a.cap;
    if (et_applies(ctx, two_kernel)) {
        const EtFn et = et_kernel(ctx);
        KSVD_HIP_CHECK(ctx, allow_max_lds((const void *)et));
        unsigned *flags = (unsigned *)ctx->ticket.ptr;
        unsigned long long *wx = (unsigned long long *)a.Gpart;
        const int nw = (int)et_pairs_n(dpad);
        hipLaunchKernelGGL(ps_reset_kernel, dim3((std::max(nw, (int)PSF_WORDS) + 255) / 256), dim3(256), 0,
                           ctx->stream, flags, wx, nw, (unsigned)(ET_NT * ET_HT), spin);
        // the XCDs take workgroups round robin: 32 of the n_xcd x 32 land on each
        hipLaunchKernelGGL(et, dim3(ctx->n_xcd * ET_NT * ET_HT), dim3(ET_T), et_lds_bytes(ctx), ctx->stream, a, flags,
                           wx);
        KSVD_HIP_CHECK(ctx, hipGetLastError());
        ctx->sweep_defers_zero = true;
        return KSVD_OK;
    }
    if (pt_applies(ctx, two_kernel)) {
        const size_t lds = pt_lds_bytes(ctx);
        {
            typedef void (*PtFn)(SweepArgs, unsigned *, unsigned long long *);
            // BU = 5: a member's columns (at most 64) in two passes of 35 with every
            // load in flight (2.33 -> 2.28 ms per update at 2048 x 2048 against BU = 4)
            PtFn pt = wide ? pt_sweep_kernel<7, 2> : pt_sweep_kernel<4, 5>;
            KSVD_HIP_CHECK(ctx, allow_max_lds((const void *)pt));
            unsigned *flags = (unsigned *)ctx->ticket.ptr;
            // the exchange pairs (two teams x two buffers of PT_HT x dpad) and
            // the d_k pairs (two teams x dpad) sit at the start of the
            // partial-Gram buffer
            unsigned long long *wx = (unsigned long long *)a.Gpart;
            const int nw = 2 * 2 * PT_HT * dpad + 2 * dpad;  // and each team's d_k pairs
            hipLaunchKernelGGL(ps_reset_kernel, dim3((std::max(nw, (int)PSF_WORDS) + 255) / 256), dim3(256), 0,
                               ctx->stream, flags, wx, nw, (unsigned)(2 * PT_HT), spin);
            // the XCDs take workgroups round robin: 2 PT_HT of the n_xcd x 2 PT_HT land on each
            hipLaunchKernelGGL(pt, dim3(ctx->n_xcd * 2 * PT_HT), dim3(PT_T), lds, ctx->stream, a, flags, wx);
            KSVD_HIP_CHECK(ctx, hipGetLastError());
            // the users beyond the cap get coefficient 0 in flush_cap_zero
            ctx->sweep_defers_zero = true;
            return KSVD_OK;
        }
    }
    int H = 24;
    if (const char *hv = getenv("KSVD_SWEEP_HELPERS")) H = std::max(1, std::min(PS_HMAX, atoi(hv)));
    if (cap > 0) H = std::max(H, (cap + PS_CWMAX - 1) / PS_CWMAX);
    const bool persistent = cap > 0 && H <= PS_HMAX && kmax <= PS_NOMAX + 1 && !two_kernel;
    if (persistent) {
        const size_t lds_helper = sizeof(float) * (2 * (size_t)PS_CWMAX * dpad + (size_t)PS_CWMAX * PS_NOMAX) +
                                  sizeof(int) * ((size_t)PS_CWMAX * PS_NOMAX + 3 * PS_CWMAX + 1);
        const size_t lds_master = sizeof(float) * ((size_t)dpad * dpad + 3 * 4 * 48);
        const size_t lds = std::max(lds_helper, lds_master);
        typedef void (*PsFn)(SweepArgs, unsigned *, int, float *);
        PsFn ps = nullptr;
        if (!wide) ps = tpw <= 1 ? ps_sweep_kernel<1, 4, 4> : tpw <= 2 ? ps_sweep_kernel<2, 4, 4>
                      : tpw <= 3 ? ps_sweep_kernel<3, 4, 4> : ps_sweep_kernel<5, 4, 4>;
        else ps = tpw <= 1 ? ps_sweep_kernel<1, 7, 2> : tpw <= 2 ? ps_sweep_kernel<2, 7, 2>
                 : tpw <= 3 ? ps_sweep_kernel<3, 7, 2> : ps_sweep_kernel<5, 7, 2>;
        KSVD_HIP_CHECK(ctx, allow_max_lds((const void *)ps));
        unsigned *flags = (unsigned *)ctx->ticket.ptr;
        // the reduced Gram follows the partials in the same buffer
        float *Gred = a.Gpart + (size_t)std::max(B1, PS_HMAX) * ntiles * 256;
        hipLaunchKernelGGL(ps_reset_kernel, dim3(1), dim3(256), 0, ctx->stream, flags, (unsigned long long *)nullptr, 0,
                           (unsigned)(H + 1), spin);
        // the XCDs take workgroups round robin: H + 1 of the n_xcd (H + 1) land on each
        hipLaunchKernelGGL(ps, dim3(ctx->n_xcd * (H + 1)), dim3(PS_T), lds, ctx->stream, a, flags, H, Gred);
        KSVD_HIP_CHECK(ctx, hipGetLastError());
        return KSVD_OK;
    }

    // ---- two kernels per atom (any cap) --------------------------------------
    const size_t lds1 = sizeof(float) * (size_t)CHUNK * dpad;
    const size_t lds2 = sizeof(float) * ((size_t)dim * dim + dim) + sizeof(int) * HAND_MAX;
    // the Gram of E must fit K2's LDS; beyond that, the Gram-free pair
    if (lds2 > 160 * 1024) return launch_large();
    typedef void (*K1Fn)(SweepArgs, int);
    K1Fn k1 = nullptr;
    if (!wide) k1 = tpw <= 1 ? k1_kernel<1, 4, 4> : tpw <= 2 ? k1_kernel<2, 4, 3> : tpw <= 3 ? k1_kernel<3, 4, 3>
                                                                                             : k1_kernel<5, 4, 2>;
    else k1 = tpw <= 1 ? k1_kernel<1, 7, 2> : tpw <= 2 ? k1_kernel<2, 7, 2> : tpw <= 3 ? k1_kernel<3, 7, 2>
                                                                                       : k1_kernel<5, 7, 2>;
    KSVD_HIP_CHECK(ctx, allow_max_lds((const void *)k1));
    KSVD_HIP_CHECK(ctx, allow_max_lds((const void *)k2_kernel));
    for (int k = 0; k <= K; k++) {
        hipLaunchKernelGGL(k1, dim3(B1), dim3(T1), lds1, ctx->stream, a, k);
        if (k < K) hipLaunchKernelGGL(k2_kernel, dim3(1), dim3(T2), lds2, ctx->stream, a, k);
    }
    KSVD_HIP_CHECK(ctx, hipGetLastError());
    return KSVD_OK;
}

// Reads and clears the persistent sweep's sticky abort word and role counts
// (synchronises the context's stream).  Returns KSVD_OK, or KSVD_ERR_HIP if a
// sweep gave up or left roles unfinished (no workgroup took them).
// Sets up the multi-picture sweep for a batch call: whether it applies to
// these contexts (same sizes, the two-team sweep for each, at most one picture
// per XCD), and the pictures' sweep arguments in ctxs[0]'s device buffer
// (copied once, waiting for ctxs[0]'s stream).  Returns 1 when it applies, 0
// when not, or a negative status.
int sweep_multi_setup(ksvd_ctx *const *ctxs, int n, float tol)
{
    ksvd_ctx *c0 = ctxs[0];
    const bool two_kernel = env_flag("KSVD_SWEEP_TWO_KERNEL");
    c0->multi_mode = 0;
    if (n < 2 || n > c0->n_xcd || n > PT_MULTI_MAX || two_kernel) return 0;
    const int nbt = (c0->dim + 15) / 16;
    const bool large = nbt * (nbt + 1) / 2 > 16 * MAX_TPW;
    KlLayout L;
    for (int i = 0; i < n; i++) {
        const ksvd_ctx *c = ctxs[i];
        if (c->device != c0->device || c->dim != c0->dim || c->K != c0->K || c->kmax != c0->kmax ||
            c->prm.user_cap != c0->prm.user_cap || c->force_two_kernel)
            return 0;
        if (large ? !kl_layout(c, KG_SPLIT, &L) : !(et_applies(c, two_kernel) || pt_applies(c, two_kernel))) return 0;
        if (!large && et_applies(c, two_kernel) != et_applies(c0, two_kernel)) return 0;
        if (large && kx_applies(c, two_kernel) != kx_applies(c0, two_kernel)) return 0;
    }
    const bool kx = large && kx_applies(c0, two_kernel);
    const size_t bytes = std::max(sizeof(PtPicture), sizeof(KlPicture)) * PT_MULTI_MAX;
    int rc = ksvd_reserve(c0, c0->multi, bytes);
    if (rc) return rc;
    if (!large || kx) {
        PtPicture h[PT_MULTI_MAX];
        for (int i = 0; i < n; i++) {
            h[i].a = sweep_args(ctxs[i], tol);
            h[i].flags = (unsigned *)ctxs[i]->ticket.ptr;
            h[i].wx = (unsigned long long *)h[i].a.Gpart;
        }
        KSVD_HIP_CHECK(c0, ksvd_copy(c0, c0->multi.ptr, h, sizeof(PtPicture) * n, hipMemcpyHostToDevice));
    } else {
        KlPicture h[PT_MULTI_MAX];
        for (int i = 0; i < n; i++) {
            kl_layout(ctxs[i], KG_SPLIT, &L);
            h[i].a = sweep_args(ctxs[i], tol);
            h[i].E = (float *)ctxs[i]->E.ptr;
            h[i].G = L.G;
            h[i].flags = (unsigned *)ctxs[i]->ticket.ptr;
            h[i].wx = L.wx;
            h[i].kw = L.kw;
        }
        KSVD_HIP_CHECK(c0, ksvd_copy(c0, c0->multi.ptr, h, sizeof(KlPicture) * n, hipMemcpyHostToDevice));
    }
    c0->multi_mode = kx ? 4 : large ? 2 : et_applies(c0, two_kernel) ? 3 : 1;
    return 1;
}

// Each picture's flag and pair reset on its own stream, before an update.
int sweep_multi_reset(ksvd_ctx *const *ctxs, int n)
{
    unsigned spin = (unsigned)PS_SPIN_MAX;
    if (const char *sv = getenv("KSVD_SWEEP_SPIN")) spin = (unsigned)std::max(0, atoi(sv));
    const bool large = ctxs[0]->multi_mode == 2;
    const int dpad = (ctxs[0]->dim + 15) / 16 * 16;
    for (int i = 0; i < n; i++) {
        ksvd_ctx *c = ctxs[i];
        c->sweep_defers_zero = false;
        unsigned long long *wx = (unsigned long long *)c->gpart.ptr;
        int nw = 2 * 2 * PT_HT * dpad + 2 * dpad;
        unsigned nroles = 2 * PT_HT;
        if (ctxs[0]->multi_mode == 3) {
            nw = (int)et_pairs_n(dpad);
            nroles = ET_NT * ET_HT;
        }
        if (ctxs[0]->multi_mode == 4) {
            nw = 2 * (KX_R + 1) * dpad;
            nroles = KX_R;
        }
        if (large) {
            KlLayout L;
            kl_layout(c, KG_SPLIT, &L);
            wx = L.wx;
            nw = L.nw;
            nroles = (unsigned)(c->K * KQ_R);
        }
        hipLaunchKernelGGL(ps_reset_kernel, dim3((std::max(nw, (int)PSF_WORDS) + 255) / 256), dim3(256), 0, c->stream,
                           (unsigned *)c->ticket.ptr, wx, nw, nroles, spin);
        KSVD_HIP_CHECK(c, hipGetLastError());
    }
    return KSVD_OK;
}

// The multi-picture sweep on ctxs[0]'s stream (after sweep_multi_setup and
// sweep_multi_reset, and after that stream waits for the other pictures'
// streams).
int launch_sweep_multi(ksvd_ctx *const *ctxs, int n)
{
    ksvd_ctx *c0 = ctxs[0];
    if (c0->multi_mode == 2) {
        // large patches: each atom's three launches cover every picture
        const int dim = c0->dim, K = c0->K;
        KlLayout L;
        kl_layout(c0, KG_SPLIT, &L);
        const int nb = (dim + KG_TB - 1) / KG_TB;
        const int ntg = nb * (nb + 1) / 2;
        const int cpt = KL_T / dim;
        int64_t ecols = 0;
        for (int i = 0; i < n; i++)
            ecols = std::max<int64_t>(ecols, ctxs[i]->prm.user_cap > 0 ? ctxs[i]->prm.user_cap : ctxs[i]->P);
        const unsigned gb = (unsigned)std::min<int64_t>(4096, std::max<int64_t>(1, (ecols + cpt - 1) / cpt));
        const KlPicture *pics = (const KlPicture *)c0->multi.ptr;
        KSVD_HIP_CHECK(c0, allow_max_lds((const void *)kq_power_multi));
        for (int k = 0; k < K; k++) {
            hipLaunchKernelGGL(kl_build_multi, dim3(gb, 1, n), dim3(KL_T), 0, c0->stream, pics, k);
            hipLaunchKernelGGL(kl_gram_multi, dim3(ntg, L.nsplit, n), dim3(256), 0, c0->stream, pics, L.gp, k);
            hipLaunchKernelGGL(kq_power_multi, dim3(c0->n_xcd * KQ_R), dim3(KQ_T), L.lds_q, c0->stream, pics, n, L.gp,
                               L.nsplit, k);
        }
        KSVD_HIP_CHECK(c0, hipGetLastError());
        return KSVD_OK;
    }
    if (c0->multi_mode == 4) {
        KSVD_HIP_CHECK(c0, allow_max_lds((const void *)kx_multi_kernel));
        hipLaunchKernelGGL(kx_multi_kernel, dim3(c0->n_xcd * KX_R), dim3(KX_T), kx_lds_bytes(c0->dim), c0->stream,
                           (const PtPicture *)c0->multi.ptr, n);
        KSVD_HIP_CHECK(c0, hipGetLastError());
        return KSVD_OK;
    }
    if (c0->multi_mode == 3) {
        const EmFn em = et_multi(c0);
        KSVD_HIP_CHECK(c0, allow_max_lds((const void *)em));
        hipLaunchKernelGGL(em, dim3(c0->n_xcd * ET_NT * ET_HT), dim3(ET_T), et_lds_bytes(c0), c0->stream,
                           (const PtPicture *)c0->multi.ptr, n);
        KSVD_HIP_CHECK(c0, hipGetLastError());
        for (int i = 0; i < n; i++) ctxs[i]->sweep_defers_zero = true;
        return KSVD_OK;
    }
    const bool wide = c0->kmax > 5;
    typedef void (*PmFn)(const PtPicture *, int);
    PmFn pm = wide ? pt_multi_kernel<7, 2> : pt_multi_kernel<4, 5>;
    KSVD_HIP_CHECK(c0, allow_max_lds((const void *)pm));
    hipLaunchKernelGGL(pm, dim3(c0->n_xcd * 2 * PT_HT), dim3(PT_T), pt_lds_bytes(c0), c0->stream,
                       (const PtPicture *)c0->multi.ptr, n);
    KSVD_HIP_CHECK(c0, hipGetLastError());
    for (int i = 0; i < n; i++) ctxs[i]->sweep_defers_zero = true;
    return KSVD_OK;
}

int check_sweep(ksvd_ctx *ctx)
{
    unsigned v[4] = {0, 0, 0, 0};  // sticky, XCD, started, finished
    unsigned *flags = (unsigned *)ctx->ticket.ptr;
    if (!flags) return KSVD_OK;
    KSVD_HIP_CHECK(ctx, hipMemcpyAsync(v, flags + PSF_STICKY, sizeof(v), hipMemcpyDeviceToHost, ctx->stream));
    KSVD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (v[0] || v[2] != v[3]) {
        ctx->sweep_failed = true;
        KSVD_HIP_CHECK(ctx, hipMemsetAsync(flags + PSF_STICKY, 0, sizeof(unsigned), ctx->stream));
        KSVD_HIP_CHECK(ctx, hipMemsetAsync(flags + PSF_STARTED, 0, 2 * sizeof(unsigned), ctx->stream));
        KSVD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        if (v[0])
            return ksvd_fail(ctx, KSVD_ERR_HIP,
                             "dictionary sweep: its workgroups did not all run on one XCD (a wait ran out); "
                             "set KSVD_SWEEP_TWO_KERNEL=1 to use the two-kernel sweep");
        return ksvd_fail(ctx, KSVD_ERR_HIP,
                         "dictionary sweep: " + std::to_string(v[3]) + " of " + std::to_string(v[2]) +
                             " workgroup roles completed; set KSVD_SWEEP_TWO_KERNEL=1 to use the two-kernel sweep");
    }
    if (v[2]) {
        KSVD_HIP_CHECK(ctx, hipMemsetAsync(flags + PSF_STARTED, 0, 2 * sizeof(unsigned), ctx->stream));
    }
    return KSVD_OK;
}

size_t sweep_gpart_bytes(int dim)
{
    const int nbt = (dim + 15) / 16;
    // B1 (or PS_HMAX) partial Grams, then the reduced one; the eight-team
    // sweep's pairs and speculation inputs
    return std::max(sizeof(float) * (size_t)(std::max(B1, PS_HMAX) + 1) * (nbt * (nbt + 1) / 2) * 256,
                    et_scratch_bytes(std::min(dim, 192)));
}
