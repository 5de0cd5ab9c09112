// Learned support proposer of the BF16 form: a 3-layer perceptron
// patch -> 512 -> 512 -> atoms (ReLU hidden layers, one logit per atom)
// trained on the GPU on (patch, OMP support) pairs, and its use as a
// one-shot support proposer with OMP as fallback.
//
// Layout.  Every activation matrix carries a constant-1 column, so each
// layer's bias is one more column of its weight matrix:
//   X  [rows x KIN]  bf16: the zero-mean part of the patch over its norm in
//                    columns 0..dim-1, the DC fraction in column dim, 1 in
//                    column dim + 1
//   H1 [rows x HP]   bf16, ReLU(X W1^T) in columns 0..511, 1 in column 512
//   H2 [rows x HP]   bf16, ReLU(H1 W2^T), 1 in column 512
//   Z  [rows x KO]   logits H2 W3^T (bf16 for proposals, FP32 in training)
//   W1 [512 x KIN], W2 [512 x HP], W3 [KO x HP]: FP32 master weights, BF16
//   copies for the GEMMs; KIN = dim + 2 rounded up to 32, HP = 544,
//   KO = atoms rounded up to 32; padding rows and columns stay 0.
//
// One GEMM kernel (C = A B^T on v_mfma_f32_16x16x32_bf16, FP32
// accumulation, 128 x 128 tiles, operands staged through LDS) serves all
// eight products of a training step, every operand read K-major (stored
// rows x K): the backward products take tiled transposes of the activations
// and weights:
//   dZ  = (sigmoid(Z) - T) / B                      (loss: mean over the batch
//                                                    of the summed BCE)
//   dW3 = dZ^T H2,  dH2 = (dZ W3) * [H2 > 0]
//   dW2 = dH2^T H1, dH1 = (dH2 W2) * [H1 > 0]
//   dW1 = dH1^T X
// Weight gradients reduce over the batch in 16 split-K partials that the
// AdamW kernel sums in a fixed order (no float atomics anywhere).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ksvd_internal.hpp"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int TM = 128, TN = 128, TK = 32;  // GEMM tile
constexpr int LDT = TK + 8;                 // LDS pitch (bf16): 80 bytes, odd multiple of 16

enum { EPI_F32 = 0, EPI_RELU_BF16 = 1, EPI_MASK_BF16 = 2, EPI_BF16 = 3 };

// Global -> register staging of one 128 x 32 operand tile (two 16-byte
// chunks per thread).  K-major: rows x K storage, chunk c covers row c/4,
// k 8*(c%4)..+7.  Rows-major: K x rows storage, chunk c covers k c/16,
// rows 8*(c%16)..+7.  Outside [0, rows) x [k0, kend) the values are 0.
template <bool KMAJ>
__device__ __forceinline__ void stage_load(int4 *v, const __bf16 *__restrict__ X, int ld, int rows, int r0, int k0,
                                           int kend)
{
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int c = threadIdx.x + 256 * h;
        int4 t = make_int4(0, 0, 0, 0);
        if (KMAJ) {
            const int row = c >> 2, kc = (c & 3) * 8;
            if (r0 + row < rows && k0 + kc < kend) t = *(const int4 *)(X + (int64_t)(r0 + row) * ld + k0 + kc);
        } else {
            const int kk = c >> 4, mc = (c & 15) * 8;
            if (k0 + kk < kend && r0 + mc < rows) t = *(const int4 *)(X + (int64_t)(k0 + kk) * ld + r0 + mc);
        }
        v[h] = t;
    }
}

template <bool KMAJ>
__device__ __forceinline__ void stage_store(__bf16 *S, const int4 *v)
{
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int c = threadIdx.x + 256 * h;
        if (KMAJ) {
            const int row = c >> 2, kc = (c & 3) * 8;
            *(int4 *)(S + row * LDT + kc) = v[h];
        } else {
            const int kk = c >> 4, mc = (c & 15) * 8;
            const __bf16 *e = (const __bf16 *)&v[h];
#pragma unroll
            for (int i = 0; i < 8; i++) S[(mc + i) * LDT + kk] = e[i];
        }
    }
}

// C[m, n] = sum_k A(m, k) B(n, k) over k in this split's range
// [z * kchunk, min(Kred, (z+1) * kchunk)).  256 threads, 4 wavefronts in 2 x 2,
// each 64 x 64 as 4 x 4 MFMA tiles.  EPI_F32 writes Cf + z * cstride;
// EPI_RELU_BF16 writes bf16(max(c, 0)); EPI_MASK_BF16 writes bf16(c) where
// mask[m * ldm + n] > 0, else 0; EPI_BF16 writes bf16(c).
template <bool AK, bool BK, int EPI>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(const __bf16 *__restrict__ A, int lda,
                                                        const __bf16 *__restrict__ B, int ldb, int M, int N, int Kred,
                                                        int kchunk, float *__restrict__ Cf, __bf16 *__restrict__ Cb,
                                                        int ldc, int64_t cstride, const __bf16 *__restrict__ mask,
                                                        int ldm)
{
    __shared__ __attribute__((aligned(16))) __bf16 As[2][TM * LDT];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[2][TN * LDT];
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int kbeg = blockIdx.z * kchunk, kend = min(Kred, kbeg + kchunk);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, g = lane >> 4;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;

    f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

    int4 va[2], vb[2];
    stage_load<AK>(va, A, lda, M, m0, kbeg, kend);
    stage_load<BK>(vb, B, ldb, N, n0, kbeg, kend);
    stage_store<AK>(As[0], va);
    stage_store<BK>(Bs[0], vb);
    __syncthreads();
    int buf = 0;
    for (int k0 = kbeg; k0 < kend; k0 += TK) {
        const bool more = k0 + TK < kend;
        if (more) {
            stage_load<AK>(va, A, lda, M, m0, k0 + TK, kend);
            stage_load<BK>(vb, B, ldb, N, n0, k0 + TK, kend);
        }
        bf8 a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; i++) a[i] = *(const bf8 *)(As[buf] + (wm + 16 * i + col) * LDT + 8 * g);
#pragma unroll
        for (int j = 0; j < 4; j++) b[j] = *(const bf8 *)(Bs[buf] + (wn + 16 * j + col) * LDT + 8 * g);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        if (more) {
            stage_store<AK>(As[buf ^ 1], va);
            stage_store<BK>(Bs[buf ^ 1], vb);
        }
        __syncthreads();
        buf ^= 1;
    }
    // lane holds C[wm + 16i + 4g + r][wn + 16j + col]
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + 16 * i + 4 * g + r, n = n0 + wn + 16 * j + col;
                if (m >= M || n >= N) continue;
                const float c = acc[i][j][r];
                const int64_t o = (int64_t)m * ldc + n;
                if (EPI == EPI_F32) Cf[blockIdx.z * cstride + o] = c;
                else if (EPI == EPI_RELU_BF16) Cb[o] = (__bf16)fmaxf(c, 0.f);
                else if (EPI == EPI_MASK_BF16) Cb[o] = (float)mask[(int64_t)m * ldm + n] > 0.f ? (__bf16)c : (__bf16)0.f;
                else Cb[o] = (__bf16)c;
            }
}

// Scale-free features of a patch (OMP's support does not change with the
// patch's scale): columns 0..dim-1 the zero-mean part (y - m 1) / |y - m 1|,
// column dim the DC fraction m sqrt(dim) / |y| (in [-1, 1]), column dim + 1
// the constant 1, 0 after; a zero part gives 0.  One wavefront per patch.
__global__ __launch_bounds__(256) void mlp_input_kernel(const float *__restrict__ Y, int dim, int64_t P, int kin,
                                                        __bf16 *__restrict__ X)
{
    const int lane = threadIdx.x & 63;
    const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (p >= P) return;
    float sum = 0.f, ss = 0.f;
    for (int e = lane; e < dim; e += 64) {
        const float v = Y[p * dim + e];
        sum += v;
        ss = fmaf(v, v, ss);
    }
    for (int off = 32; off > 0; off >>= 1) {
        sum += __shfl_xor(sum, off);
        ss += __shfl_xor(ss, off);
    }
    const float m = sum / (float)dim;
    float sa = 0.f;
    for (int e = lane; e < dim; e += 64) {
        const float v = Y[p * dim + e] - m;
        sa = fmaf(v, v, sa);
    }
    for (int off = 32; off > 0; off >>= 1) sa += __shfl_xor(sa, off);
    const float inva = sa > 0.f ? 1.f / sqrtf(sa) : 0.f;
    const float dc = ss > 0.f ? m * sqrtf((float)dim) / sqrtf(ss) : 0.f;
    for (int e = lane; e < kin; e += 64) {
        const float v = e < dim ? (Y[p * dim + e] - m) * inva : (e == dim ? dc : (e == dim + 1 ? 1.f : 0.f));
        X[p * kin + e] = (__bf16)v;
    }
}

// Mean loss of the last training batch: thread t sums partials t, t + 256,
// ... in order, then a fixed tree over the 256 threads.
__global__ __launch_bounds__(256) void loss_sum_kernel(const float *__restrict__ lpart, int n, int nb,
                                                       float *__restrict__ out)
{
    __shared__ float red[4];
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += 256) s += lpart[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) *out = ((red[0] + red[1]) + (red[2] + red[3])) / (float)nb;
}

// out [cols x rows] = in [rows x cols]^T (bf16, leading dimensions ldi / ldo),
// 64 x 64 tiles through LDS.
__global__ __launch_bounds__(256) void transpose_bf16_kernel(const __bf16 *__restrict__ in, int rows, int cols,
                                                             int ldi, __bf16 *__restrict__ out, int ldo)
{
    __shared__ __bf16 t[64][66];
    const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4)
        t[r][tx] = (r0 + r < rows && c0 + tx < cols) ? in[(int64_t)(r0 + r) * ldi + c0 + tx] : (__bf16)0.f;
    __syncthreads();
    for (int c = ty; c < 64; c += 4)
        if (c0 + c < cols && r0 + tx < rows) out[(int64_t)(c0 + c) * ldo + r0 + tx] = t[tx][c];
}

// The same transpose with 8 x 8 blocks in registers (no LDS): a thread
// loads 8 rows of 8 elements (16 bytes each) and stores 8 rows of 8, so
// every access is a 16-byte vector and 8 threads cover a 128-byte segment.
// rows, cols, ldi and ldo multiples of 8 (checked by the caller).  256
// threads cover a 128 x 128 tile.
__global__ __launch_bounds__(256) void transpose8_bf16_kernel(const __bf16 *__restrict__ in, int rows, int cols,
                                                              int ldi, __bf16 *__restrict__ out, int ldo)
{
    const int bc = threadIdx.x & 15, br = threadIdx.x >> 4;
    const int r = blockIdx.y * 128 + 8 * br, c = blockIdx.x * 128 + 8 * bc;
    if (r >= rows || c >= cols) return;
    uint4 v[8];
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = *(const uint4 *)(in + (int64_t)(r + i) * ldi + c);
    auto word = [](const uint4 &q, int k) { return k == 0 ? q.x : k == 1 ? q.y : k == 2 ? q.z : q.w; };
#pragma unroll
    for (int j = 0; j < 8; j++) {
        unsigned w[4];
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const unsigned a = word(v[2 * p], j >> 1), b = word(v[2 * p + 1], j >> 1);
            w[p] = (j & 1) ? ((a >> 16) | (b & 0xffff0000u)) : ((a & 0xffffu) | (b << 16));
        }
        *(uint4 *)(out + (int64_t)(c + j) * ldo + r) = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// rows [r0, r0 + n) of a [rows x ld] bf16 matrix: column c1 = 1, other
// columns from c1 on = 0 (the constant-1 column and its padding)
__global__ void ones_column_kernel(__bf16 *__restrict__ H, int ld, int c1, int64_t rows)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    for (int c = c1; c < ld; c++) H[r * ld + c] = (__bf16)(c == c1 ? 1.f : 0.f);
}

// Training batch: rows perm[b0 + b] of X.
__global__ __launch_bounds__(256) void gather_rows_kernel(const __bf16 *__restrict__ X, int kin,
                                                          const int *__restrict__ perm, int b0, int nb,
                                                          __bf16 *__restrict__ Xb)
{
    const int c16 = kin / 8;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)nb * c16) return;
    const int b = (int)(q / c16), c = (int)(q - (int64_t)b * c16);
    const int p = perm[b0 + b];
    ((int4 *)(Xb + (int64_t)b * kin))[c] = ((const int4 *)(X + (int64_t)p * kin))[c];
}

// dZ = (sigmoid(z) - t) / nb for atoms < K (0 for the padding), t = 1 when
// the atom is in the support of patch perm[b0 + b] (codes idx/coef/cnt, a
// nonzero coefficient); lpart[b] = the row's loss sum over the atoms of
// softplus(z) - t z.  One wavefront per row, 4 rows a block (ko <= 256):
// lane L holds atoms L + 64 q, q = 0..3, each group of 64 atoms is summed by
// an xor butterfly and the group sums of lane 0 as (g0 + g1) + (g2 + g3).
__global__ __launch_bounds__(256) void bce_grad_kernel(const float *__restrict__ Z, int ko, int K, int nb,
                                                       const int *__restrict__ perm, int b0,
                                                       const int *__restrict__ idx, const float *__restrict__ coef,
                                                       const int *__restrict__ cnt, int kmax,
                                                       __bf16 *__restrict__ dZ, float *__restrict__ lpart)
{
    const int lane = threadIdx.x & 63;
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (b >= nb) return;  // the whole wavefront
    const int64_t p = perm[b0 + b];
    const int ns = min(cnt[p], kmax);
    bool t[4] = {false, false, false, false};
    for (int i = 0; i < ns; i++) {
        const int a = idx[p * kmax + i];
        if (coef[p * kmax + i] != 0.f) {
#pragma unroll
            for (int q = 0; q < 4; q++) t[q] |= a == lane + 64 * q;
        }
    }
    float l[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int k = lane + 64 * q;
        l[q] = 0.f;
        if (k < ko) {
            float d = 0.f;
            if (k < K) {
                const float tv = t[q] ? 1.f : 0.f;
                const float z = Z[(int64_t)b * ko + k];
                const float sg = 1.f / (1.f + expf(-z));
                d = (sg - tv) / (float)nb;
                l[q] = fmaxf(z, 0.f) + log1pf(expf(-fabsf(z))) - tv * z;
            }
            dZ[(int64_t)b * ko + k] = (__bf16)d;
        }
    }
#pragma unroll
    for (int q = 0; q < 4; q++)
        for (int off = 32; off > 0; off >>= 1) l[q] += __shfl_xor(l[q], off);
    if (lane == 0) lpart[b] = (l[0] + l[1]) + (l[2] + l[3]);
}

// AdamW over the three weight matrices (contiguous FP32 master copy):
// gradient = sum of the nsplit partials in split order; decoupled weight
// decay on weights (not on the bias column); padding entries held at 0;
// BF16 copies written alongside.
struct AdamArgs {
    float *theta, *m, *v;
    __bf16 *wb;
    const float *g1, *g2, *g3;
    int64_t n1, n2, n3;     // entries of W1, W2, W3
    int c1, cb1, cmax1;     // W1: pitch, bias column, last used column + 1 (dim + 2)
    int c2, cb2;            // W2 / W3: pitch HP, bias column 512
    int rows3;              // W3 rows that are atoms (K)
    int nsplit;
    float lr, b1, b2, eps, wd, bc1, bc2;
};

__global__ __launch_bounds__(256) void adamw_kernel(AdamArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = a.n1 + a.n2 + a.n3;
    if (i >= n) return;
    const float *gp;
    int64_t li, stride;
    int pitch, bias_col, used_cols, used_rows;
    if (i < a.n1) {
        gp = a.g1; li = i; stride = a.n1; pitch = a.c1; bias_col = a.cb1; used_cols = a.cmax1; used_rows = 1 << 30;
    } else if (i < a.n1 + a.n2) {
        gp = a.g2; li = i - a.n1; stride = a.n2; pitch = a.c2; bias_col = a.cb2; used_cols = a.cb2 + 1;
        used_rows = 1 << 30;
    } else {
        gp = a.g3; li = i - a.n1 - a.n2; stride = a.n3; pitch = a.c2; bias_col = a.cb2; used_cols = a.cb2 + 1;
        used_rows = a.rows3;
    }
    const int row = (int)(li / pitch), c = (int)(li - (int64_t)row * pitch);
    float th = a.theta[i];
    if (c >= used_cols || row >= used_rows) {
        th = 0.f;
    } else {
        float gr = 0.f;
        for (int s = 0; s < a.nsplit; s++) gr += gp[s * stride + li];
        const float m = a.b1 * a.m[i] + (1.f - a.b1) * gr;
        const float v = a.b2 * a.v[i] + (1.f - a.b2) * gr * gr;
        a.m[i] = m;
        a.v[i] = v;
        const float step = (m / a.bc1) / (sqrtf(v / a.bc2) + a.eps);
        if (c != bias_col) th -= a.lr * a.wd * th;
        th -= a.lr * step;
    }
    a.theta[i] = th;
    a.wb[i] = (__bf16)th;
}

__global__ void f32_to_bf16_kernel(const float *__restrict__ s, __bf16 *__restrict__ d, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (__bf16)s[i];
}

// ---- proposals -----------------------------------------------------------
// 16 lanes per patch (4 patches per wavefront): the kmax largest logits
// (lowest atom on ties) form the support; least squares on the original
// patch with the FP32 atoms (Cholesky of the support Gram from the FP32
// dictionary Gram, two triangular solves); accepted when the residual is at
// most accept x the patch's last OMP residual and every pivot exceeds
// 1e-6 of its Gram diagonal, else the patch is flagged for the fallback
// list (rej[p] = 1), which fallback_compact builds in ascending patch order.
template <int KM, int RJ>
__global__ __launch_bounds__(256) void propose_kernel(const __bf16 *__restrict__ Z, int ko, int K,
                                                      const float *__restrict__ Y, int dim, int64_t p0, int64_t np,
                                                      const float *__restrict__ D, const float *__restrict__ G,
                                                      int kmax, const float *__restrict__ rbound, float accept,
                                                      int write, int *__restrict__ idx_out, float *__restrict__ coef_out,
                                                      int *__restrict__ cnt_out, unsigned char *__restrict__ rej)
{
    const int lane = threadIdx.x & 63;
    const int sub = lane & 15;
    const int64_t q = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
    const bool valid = q < np;
    const int64_t p = p0 + (valid ? q : 0);
    // logits: lane sub holds atoms sub + 16 i
    float z[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int a = sub + 16 * i;
        z[i] = (valid && a < K) ? (float)Z[q * ko + a] : -INFINITY;
    }
    int S[KM];
#pragma unroll
    for (int t = 0; t < KM; t++) {
        float bvl = -INFINITY;
        int bi = 0x7fffffff;
#pragma unroll
        for (int i = 0; i < 16; i++)
            if (z[i] > bvl) { bvl = z[i]; bi = sub + 16 * i; }
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) {
            const float ov = __shfl_xor(bvl, off);
            const int oi = __shfl_xor(bi, off);
            if (ov > bvl || (ov == bvl && oi < bi)) { bvl = ov; bi = oi; }
        }
        S[t] = (t < kmax && bi < K) ? bi : -1;
#pragma unroll
        for (int i = 0; i < 16; i++)
            if (sub + 16 * i == bi) z[i] = -INFINITY;
    }
    float y[RJ];
#pragma unroll
    for (int j = 0; j < RJ; j++) {
        const int e = sub + 16 * j;
        y[j] = (valid && e < dim) ? Y[p * dim + e] : 0.f;
    }
    // the support's atoms (this lane's elements), read once
    float dS[KM][RJ];
#pragma unroll
    for (int i = 0; i < KM; i++)
#pragma unroll
        for (int j = 0; j < RJ; j++) {
            const int e = sub + 16 * j;
            dS[i][j] = (i < kmax && S[i] >= 0 && e < dim) ? D[(int64_t)S[i] * dim + e] : 0.f;
        }
    // b = D_S^T y and the Cholesky factor of the support Gram
    float bvec[KM], L[KM * KM], x[KM];
    bool ok = valid;
#pragma unroll
    for (int i = 0; i < KM; i++) {
        float s = 0.f;
#pragma unroll
        for (int j = 0; j < RJ; j++) s = fmaf(dS[i][j], y[j], s);
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) s += __shfl_xor(s, off);
        bvec[i] = s;
        ok = ok && (i >= kmax || S[i] >= 0);
    }
#pragma unroll
    for (int i = 0; i < KM; i++) {
        x[i] = 0.f;
#pragma unroll
        for (int q2 = 0; q2 < KM; q2++) L[i * KM + q2] = 0.f;
    }
#pragma unroll
    for (int i = 0; i < KM; i++) {
        if (i < kmax && ok) {
            const float gii = G[(int64_t)S[i] * K + S[i]];
            float piv = gii;
#pragma unroll
            for (int q2 = 0; q2 < i; q2++) {
                float v = G[(int64_t)S[i] * K + S[q2]];
#pragma unroll
                for (int m = 0; m < q2; m++) v = fmaf(-L[i * KM + m], L[q2 * KM + m], v);
                v = v / L[q2 * KM + q2];
                L[i * KM + q2] = v;
                piv = fmaf(-v, v, piv);
            }
            if (piv > 1e-6f * gii) L[i * KM + i] = sqrtf(piv);
            else ok = false;
        }
    }
    if (ok) {
        float zz[KM];
#pragma unroll
        for (int i = 0; i < KM; i++) {
            zz[i] = 0.f;
            if (i < kmax) {
                float v = bvec[i];
#pragma unroll
                for (int q2 = 0; q2 < i; q2++) v = fmaf(-L[i * KM + q2], zz[q2], v);
                zz[i] = v / L[i * KM + i];
            }
        }
#pragma unroll
        for (int i = KM - 1; i >= 0; i--) {
            if (i < kmax) {
                float v = zz[i];
#pragma unroll
                for (int q2 = i + 1; q2 < KM; q2++)
                    if (q2 < kmax) v = fmaf(-L[q2 * KM + i], x[q2], v);
                x[i] = v / L[i * KM + i];
            }
        }
    }
    // residual norm
    float nr = 0.f;
#pragma unroll
    for (int j = 0; j < RJ; j++) {
        const int e = sub + 16 * j;
        float a = 0.f;
#pragma unroll
        for (int i = 0; i < KM; i++)
            if (i < kmax && ok && e < dim) a = fmaf(dS[i][j], x[i], a);
        const float r = y[j] - a;
        nr = fmaf(r, r, nr);
    }
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) nr += __shfl_xor(nr, off);
    const bool accept_it = ok && sqrtf(nr) <= accept * rbound[p];
    if (valid && sub == 0) {
        rej[p] = accept_it ? 0 : 1;
        if (accept_it && write) {
            cnt_out[p] = kmax;
#pragma unroll
            for (int i = 0; i < KM; i++)
                if (i < kmax) {
                    idx_out[p * kmax + i] = S[i];
                    coef_out[p * kmax + i] = x[i];
                }
        }
    }
}

// ---- fallback list: the flagged patches in ascending order ----------------
// Three passes over n flags: counts per block of 1024, an exclusive scan of
// the counts (one workgroup, carried over tiles of 1024; the total goes to
// *total), and the stores at each block's offset by wavefront ballots.  The
// list is the same whatever order the proposals ran in, so the training
// batches cut from it are too.
constexpr int FB_B = 1024;

__global__ __launch_bounds__(FB_B) void fb_count_kernel(const unsigned char *__restrict__ f, int64_t n,
                                                        int *__restrict__ bcnt)
{
    __shared__ int ws[FB_B / 64];
    const int64_t i = (int64_t)blockIdx.x * FB_B + threadIdx.x;
    const unsigned long long b = __ballot(i < n && f[i] != 0);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = __popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < FB_B / 64; w++) t += ws[w];
        bcnt[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(FB_B) void fb_scan_kernel(int *__restrict__ bcnt, int nb, unsigned *__restrict__ total)
{
    __shared__ int sh[FB_B];
    int carry = 0;
    for (int b0 = 0; b0 < nb; b0 += FB_B) {
        const int i = b0 + (int)threadIdx.x;
        const int v = i < nb ? bcnt[i] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        // Hillis-Steele inclusive scan in LDS
        for (int off = 1; off < FB_B; off <<= 1) {
            const int a = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0;
            __syncthreads();
            sh[threadIdx.x] += a;
            __syncthreads();
        }
        if (i < nb) bcnt[i] = carry + sh[threadIdx.x] - v;
        carry += sh[FB_B - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = (unsigned)carry;
}

__global__ __launch_bounds__(FB_B) void fb_store_kernel(const unsigned char *__restrict__ f, int64_t n,
                                                        const int *__restrict__ boff, int *__restrict__ list)
{
    __shared__ int ws[FB_B / 64];
    const int64_t i = (int64_t)blockIdx.x * FB_B + threadIdx.x;
    const bool on = i < n && f[i] != 0;
    const unsigned long long b = __ballot(on);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) ws[w] = __popcll(b);
    __syncthreads();
    int off = boff[blockIdx.x];
    for (int v = 0; v < w; v++) off += ws[v];
    if (on) list[off + __popcll(b & ((1ull << lane) - 1ull))] = (int)i;
}

// Host-side helpers ----------------------------------------------------------

template <bool AK, bool BK, int EPI>
int gemm(ksvd_ctx *ctx, const __bf16 *A, int lda, const __bf16 *B, int ldb, int M, int N, int Kred, int nsplit,
         float *Cf, __bf16 *Cb, int ldc, int64_t cstride, const __bf16 *mask, int ldm)
{
    const int kchunk = ((Kred + nsplit - 1) / nsplit + TK - 1) / TK * TK;
    dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM, nsplit);
    hipLaunchKernelGGL((gemm_bf16_kernel<AK, BK, EPI>), grid, dim3(256), 0, ctx->stream, A, lda, B, ldb, M, N, Kred,
                       kchunk, Cf, Cb, ldc, cstride, mask, ldm);
    KSVD_HIP_CHECK(ctx, hipGetLastError());
    return KSVD_OK;
}

}  // namespace

// ---- the proposer's state ---------------------------------------------------
int mlp_dims(const ksvd_ctx *ctx, int *kin, int *ko)
{
    *kin = (ctx->dim + 2 + 31) / 32 * 32;
    *ko = (ctx->K + 31) / 32 * 32;
    return KSVD_OK;
}

static int64_t mlp_nparams(const ksvd_ctx *ctx)
{
    int kin, ko;
    mlp_dims(ctx, &kin, &ko);
    return (int64_t)MLP_H * kin + (int64_t)MLP_H * MLP_HP + (int64_t)ko * MLP_HP;
}

// Allocate the weights and the training batch buffers (once per context).
static int mlp_alloc(ksvd_ctx *ctx)
{
    if (ctx->mlp_ready) return KSVD_OK;
    int kin, ko;
    mlp_dims(ctx, &kin, &ko);
    const int64_t np = mlp_nparams(ctx);
    const int64_t B = MLP_BATCH;
    const int64_t n1 = (int64_t)MLP_H * kin, n2 = (int64_t)MLP_H * MLP_HP, n3 = (int64_t)ko * MLP_HP;
    struct {
        DevBuf *b;
        size_t bytes;
    } plan[] = {
        {&ctx->mlp_theta, sizeof(float) * np},
        {&ctx->mlp_m, sizeof(float) * np},
        {&ctx->mlp_v, sizeof(float) * np},
        {&ctx->mlp_wb, sizeof(__bf16) * np},
        {&ctx->mlp_grad, sizeof(float) * MLP_SPLIT * (n1 + n2 + n3)},
        {&ctx->mlp_xb, sizeof(__bf16) * B * kin},
        {&ctx->mlp_h1b, sizeof(__bf16) * B * MLP_HP},
        {&ctx->mlp_h2b, sizeof(__bf16) * B * MLP_HP},
        {&ctx->mlp_zb, sizeof(float) * B * ko},
        {&ctx->mlp_dz, sizeof(__bf16) * B * ko},
        {&ctx->mlp_dh2, sizeof(__bf16) * B * MLP_H},
        {&ctx->mlp_dh1, sizeof(__bf16) * B * MLP_H},
        {&ctx->mlp_lpart, sizeof(float) * B},
        {&ctx->mlp_tr, sizeof(__bf16) * ((size_t)(MLP_HP + MLP_H + kin) * B + (size_t)MLP_H * ko + (size_t)MLP_H * MLP_H)},
        {&ctx->mlp_cnt, 256},
    };
    for (auto &e : plan) {
        int rc = ksvd_reserve(ctx, *e.b, e.bytes);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(ones_column_kernel, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, ctx->stream,
                       (__bf16 *)ctx->mlp_h1b.ptr, MLP_HP, MLP_H, B);
    hipLaunchKernelGGL(ones_column_kernel, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, ctx->stream,
                       (__bf16 *)ctx->mlp_h2b.ptr, MLP_HP, MLP_H, B);
    KSVD_HIP_CHECK(ctx, hipGetLastError());
    ctx->mlp_ready = true;
    return KSVD_OK;
}

// splitmix64 for the weight initialisation
static uint64_t sm64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static int upload_theta(ksvd_ctx *ctx, const std::vector<float> &th)
{
    const int64_t np = (int64_t)th.size();
    KSVD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    KSVD_HIP_CHECK(ctx, ksvd_copy(ctx, ctx->mlp_theta.ptr, th.data(), sizeof(float) * np, hipMemcpyHostToDevice));
    KSVD_HIP_CHECK(ctx, ksvd_set(ctx, ctx->mlp_m.ptr, 0, sizeof(float) * np));
    KSVD_HIP_CHECK(ctx, ksvd_set(ctx, ctx->mlp_v.ptr, 0, sizeof(float) * np));
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, ctx->stream,
                       (const float *)ctx->mlp_theta.ptr, (__bf16 *)ctx->mlp_wb.ptr, np);
    KSVD_HIP_CHECK(ctx, hipGetLastError());
    ctx->mlp_step = 0;
    return KSVD_OK;
}

int mlp_init(ksvd_ctx *ctx, uint64_t seed)
{
    int rc = mlp_alloc(ctx);
    if (rc) return rc;
    int kin, ko;
    mlp_dims(ctx, &kin, &ko);
    std::vector<float> th((size_t)mlp_nparams(ctx), 0.f);
    uint64_t s = seed;
    // U(-1/sqrt(fan_in), 1/sqrt(fan_in)) on the used weights, biases 0
    auto fill = [&](float *W, int rows, int pitch, int fan_in, int used_rows) {
        const float a = 1.f / std::sqrt((float)fan_in);
        for (int r = 0; r < used_rows; r++)
            for (int c = 0; c < fan_in; c++) {
                const double u = (double)(sm64(s) >> 11) * (1.0 / 9007199254740992.0);
                W[(size_t)r * pitch + c] = (float)((2.0 * u - 1.0) * a);
            }
        (void)rows;
    };
    float *W1 = th.data(), *W2 = W1 + (size_t)MLP_H * kin, *W3 = W2 + (size_t)MLP_H * MLP_HP;
    fill(W1, MLP_H, kin, ctx->dim + 1, MLP_H);
    fill(W2, MLP_H, MLP_HP, MLP_H, MLP_H);
    fill(W3, ko, MLP_HP, MLP_H, ctx->K);
    return upload_theta(ctx, th);
}

int mlp_set_params(ksvd_ctx *ctx, const float *theta)
{
    int rc = mlp_alloc(ctx);
    if (rc) return rc;
    std::vector<float> th(theta, theta + mlp_nparams(ctx));
    return upload_theta(ctx, th);
}

int mlp_get_params(ksvd_ctx *ctx, float *theta)
{
    if (!ctx->mlp_ready) return ksvd_fail(ctx, KSVD_ERR_STATE, "proposer not initialised");
    KSVD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    KSVD_HIP_CHECK(ctx, ksvd_copy(ctx, theta, ctx->mlp_theta.ptr, sizeof(float) * mlp_nparams(ctx), hipMemcpyDeviceToHost));
    return KSVD_OK;
}

int64_t mlp_param_count(const ksvd_ctx *ctx) { return mlp_nparams(ctx); }

// The normalised patches X [P x KIN] (bf16), computed once per image.
static int mlp_inputs(ksvd_ctx *ctx)
{
    if (ctx->mlp_x_valid) return KSVD_OK;
    int kin, ko;
    mlp_dims(ctx, &kin, &ko);
    int rc = ksvd_reserve(ctx, ctx->mlp_x, sizeof(__bf16) * (size_t)ctx->P * kin);
    if (rc) return rc;
    hipLaunchKernelGGL(mlp_input_kernel, dim3((unsigned)((ctx->P + 3) / 4)), dim3(256), 0, ctx->stream,
                       (const float *)ctx->Y.ptr, ctx->dim, ctx->P, kin, (__bf16 *)ctx->mlp_x.ptr);
    KSVD_HIP_CHECK(ctx, hipGetLastError());
    ctx->mlp_x_valid = true;
    return KSVD_OK;
}

// Forward of nb rows of Xb through the three layers: H1b, H2b (bf16) and Zb
// (FP32 logits).
static int mlp_forward_batch(ksvd_ctx *ctx, const __bf16 *Xb, int nb)
{
    int kin, ko;
    mlp_dims(ctx, &kin, &ko);
    const __bf16 *W1 = (const __bf16 *)ctx->mlp_wb.ptr, *W2 = W1 + (size_t)MLP_H * kin,
                 *W3 = W2 + (size_t)MLP_H * MLP_HP;
    __bf16 *H1 = (__bf16 *)ctx->mlp_h1b.ptr, *H2 = (__bf16 *)ctx->mlp_h2b.ptr;
    int rc;
    if ((rc = gemm<true, true, EPI_RELU_BF16>(ctx, Xb, kin, W1, kin, nb, MLP_H, kin, 1, nullptr, H1, MLP_HP, 0,
                                              nullptr, 0)))
        return rc;
    if ((rc = gemm<true, true, EPI_RELU_BF16>(ctx, H1, MLP_HP, W2, MLP_HP, nb, MLP_H, MLP_HP, 1, nullptr, H2, MLP_HP,
                                              0, nullptr, 0)))
        return rc;
    return gemm<true, true, EPI_F32>(ctx, H2, MLP_HP, W3, MLP_HP, nb, ko, MLP_HP, 1, (float *)ctx->mlp_zb.ptr, nullptr,
                                     ko, 0, nullptr, 0);
}

static int transpose(ksvd_ctx *ctx, const __bf16 *in, int rows, int cols, int ldi, __bf16 *out, int ldo)
{
    const bool v16 = ((rows | cols | ldi | ldo) & 7) == 0 && ((uintptr_t)in & 15) == 0 && ((uintptr_t)out & 15) == 0;
    if (v16) {
        dim3 grid((cols + 127) / 128, (rows + 127) / 128);
        hipLaunchKernelGGL(transpose8_bf16_kernel, grid, dim3(256), 0, ctx->stream, in, rows, cols, ldi, out, ldo);
        KSVD_HIP_CHECK(ctx, hipGetLastError());
        return KSVD_OK;
    }
    dim3 grid((cols + 63) / 64, (rows + 63) / 64);
    hipLaunchKernelGGL(transpose_bf16_kernel, grid, dim3(256), 0, ctx->stream, in, rows, cols, ldi, out, ldo);
    KSVD_HIP_CHECK(ctx, hipGetLastError());
    return KSVD_OK;
}

// Backward of the batch whose logits are in Zb and whose input is Xb: dZ is
// already in mlp_dz.  Every product runs K-major: the activations and the
// weights are transposed first (tiled, through LDS).  Leaves the split-K
// weight-gradient partials in mlp_grad.
static int mlp_backward_batch(ksvd_ctx *ctx, const __bf16 *Xb, int nb)
{
    int kin, ko;
    mlp_dims(ctx, &kin, &ko);
    const __bf16 *W1 = (const __bf16 *)ctx->mlp_wb.ptr, *W2 = W1 + (size_t)MLP_H * kin,
                 *W3 = W2 + (size_t)MLP_H * MLP_HP;
    const __bf16 *H1 = (const __bf16 *)ctx->mlp_h1b.ptr, *H2 = (const __bf16 *)ctx->mlp_h2b.ptr;
    const __bf16 *dZ = (const __bf16 *)ctx->mlp_dz.ptr;
    __bf16 *dH2 = (__bf16 *)ctx->mlp_dh2.ptr, *dH1 = (__bf16 *)ctx->mlp_dh1.ptr;
    // transposed operands
    __bf16 *T0 = (__bf16 *)ctx->mlp_tr.ptr;                  // [HP x nb]: H2^T, then H1^T
    __bf16 *T1 = T0 + (size_t)MLP_HP * MLP_BATCH;             // [512 x nb]: dZ^T, dH2^T, dH1^T
    __bf16 *T2 = T1 + (size_t)MLP_H * MLP_BATCH;              // [kin x nb]: X^T
    __bf16 *W3t = T2 + (size_t)kin * MLP_BATCH;               // [512 x ko]
    __bf16 *W2t = W3t + (size_t)MLP_H * ko;                   // [512 x 512]
    const int64_t n1 = (int64_t)MLP_H * kin, n2 = (int64_t)MLP_H * MLP_HP, n3 = (int64_t)ko * MLP_HP;
    float *g1 = (float *)ctx->mlp_grad.ptr, *g2 = g1 + MLP_SPLIT * n1, *g3 = g2 + MLP_SPLIT * n2;
    int rc;
    // dW3 [ko x HP] = dZ^T H2
    if ((rc = transpose(ctx, dZ, nb, ko, ko, T1, nb))) return rc;
    if ((rc = transpose(ctx, H2, nb, MLP_HP, MLP_HP, T0, nb))) return rc;
    if ((rc = gemm<true, true, EPI_F32>(ctx, T1, nb, T0, nb, ko, MLP_HP, nb, MLP_SPLIT, g3, nullptr, MLP_HP, n3,
                                        nullptr, 0)))
        return rc;
    // dH2 [nb x 512] = (dZ W3) * [H2 > 0]
    if ((rc = transpose(ctx, W3, ko, MLP_H, MLP_HP, W3t, ko))) return rc;
    if ((rc = gemm<true, true, EPI_MASK_BF16>(ctx, dZ, ko, W3t, ko, nb, MLP_H, ko, 1, nullptr, dH2, MLP_H, 0, H2,
                                              MLP_HP)))
        return rc;
    // dW2 [512 x HP] = dH2^T H1
    if ((rc = transpose(ctx, dH2, nb, MLP_H, MLP_H, T1, nb))) return rc;
    if ((rc = transpose(ctx, H1, nb, MLP_HP, MLP_HP, T0, nb))) return rc;
    if ((rc = gemm<true, true, EPI_F32>(ctx, T1, nb, T0, nb, MLP_H, MLP_HP, nb, MLP_SPLIT, g2, nullptr, MLP_HP, n2,
                                        nullptr, 0)))
        return rc;
    // dH1 [nb x 512] = (dH2 W2) * [H1 > 0]
    if ((rc = transpose(ctx, W2, MLP_H, MLP_H, MLP_HP, W2t, MLP_H))) return rc;
    if ((rc = gemm<true, true, EPI_MASK_BF16>(ctx, dH2, MLP_H, W2t, MLP_H, nb, MLP_H, MLP_H, 1, nullptr, dH1, MLP_H,
                                              0, H1, MLP_HP)))
        return rc;
    // dW1 [512 x kin] = dH1^T X
    if ((rc = transpose(ctx, dH1, nb, MLP_H, MLP_H, T1, nb))) return rc;
    if ((rc = transpose(ctx, Xb, nb, kin, kin, T2, nb))) return rc;
    return gemm<true, true, EPI_F32>(ctx, T1, nb, T2, nb, MLP_H, kin, nb, MLP_SPLIT, g1, nullptr, kin, n1, nullptr,
                                     0);
}

static int mlp_adamw(ksvd_ctx *ctx, float lr)
{
    int kin, ko;
    mlp_dims(ctx, &kin, &ko);
    ctx->mlp_step++;
    AdamArgs a;
    a.theta = (float *)ctx->mlp_theta.ptr;
    a.m = (float *)ctx->mlp_m.ptr;
    a.v = (float *)ctx->mlp_v.ptr;
    a.wb = (__bf16 *)ctx->mlp_wb.ptr;
    a.n1 = (int64_t)MLP_H * kin;
    a.n2 = (int64_t)MLP_H * MLP_HP;
    a.n3 = (int64_t)ko * MLP_HP;
    a.g1 = (const float *)ctx->mlp_grad.ptr;
    a.g2 = a.g1 + MLP_SPLIT * a.n1;
    a.g3 = a.g2 + MLP_SPLIT * a.n2;
    a.c1 = kin;
    a.cb1 = ctx->dim + 1;
    a.cmax1 = ctx->dim + 2;
    a.c2 = MLP_HP;
    a.cb2 = MLP_H;
    a.rows3 = ctx->K;
    a.nsplit = MLP_SPLIT;
    a.lr = lr;
    a.b1 = 0.9f;
    a.b2 = 0.999f;
    a.eps = 1e-8f;
    a.wd = 1e-4f;
    a.bc1 = 1.f - std::pow(0.9f, (float)ctx->mlp_step);
    a.bc2 = 1.f - std::pow(0.999f, (float)ctx->mlp_step);
    const int64_t n = a.n1 + a.n2 + a.n3;
    hipLaunchKernelGGL(adamw_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, a);
    KSVD_HIP_CHECK(ctx, hipGetLastError());
    return KSVD_OK;
}

// Loss and gradients of one batch given on the host (test hook): X [nb x kin]
// and targets T [nb x ko] (0/1), both FP32 (rounded to bf16 / used as is).
int mlp_loss_grad_host(ksvd_ctx *ctx, const float *X, const float *T, int nb, double *loss, float *grad)
{
    if (!ctx->mlp_ready) return ksvd_fail(ctx, KSVD_ERR_STATE, "proposer not initialised");
    if (nb <= 0 || nb > MLP_BATCH || nb % 32 != 0)
        return ksvd_fail(ctx, KSVD_ERR_ARG, "batch must be a positive multiple of 32 up to the training batch");
    int kin, ko;
    mlp_dims(ctx, &kin, &ko);
    // targets as a fake code table: patch b has its target atoms as support
    std::vector<__bf16> xb((size_t)nb * kin);
    for (size_t i = 0; i < xb.size(); i++) xb[i] = (__bf16)X[i];
    int kmx = 0;
    for (int b = 0; b < nb; b++) {
        int c = 0;
        for (int k = 0; k < ctx->K; k++) c += T[(size_t)b * ko + k] != 0.f;
        kmx = std::max(kmx, c);
    }
    kmx = std::max(kmx, 1);
    std::vector<int> idx((size_t)nb * kmx, -1), cnt(nb, 0), perm(nb);
    std::vector<float> coef((size_t)nb * kmx, 0.f);
    for (int b = 0; b < nb; b++) {
        perm[b] = b;
        for (int k = 0; k < ctx->K; k++)
            if (T[(size_t)b * ko + k] != 0.f) {
                idx[(size_t)b * kmx + cnt[b]] = k;
                coef[(size_t)b * kmx + cnt[b]] = 1.f;
                cnt[b]++;
            }
    }
    DevBuf didx, dcoef, dcnt, dperm;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(ctx->stream);
        std::lock_guard<std::mutex> lock(device_op_mutex());
        for (DevBuf *b : {&didx, &dcoef, &dcnt, &dperm})
            if (b->ptr) (void)hipFree(b->ptr);
    };
    int rc = 0;
    if ((rc = ksvd_reserve(ctx, didx, sizeof(int) * idx.size())) || (rc = ksvd_reserve(ctx, dcoef, sizeof(float) * coef.size())) ||
        (rc = ksvd_reserve(ctx, dcnt, sizeof(int) * nb)) || (rc = ksvd_reserve(ctx, dperm, sizeof(int) * nb))) {
        cleanup();
        return rc;
    }
    hipError_t e = hipSuccess;
    e = e ? e : ksvd_copy(ctx, ctx->mlp_xb.ptr, xb.data(), sizeof(__bf16) * xb.size(), hipMemcpyHostToDevice);
    e = e ? e : ksvd_copy(ctx, didx.ptr, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice);
    e = e ? e : ksvd_copy(ctx, dcoef.ptr, coef.data(), sizeof(float) * coef.size(), hipMemcpyHostToDevice);
    e = e ? e : ksvd_copy(ctx, dcnt.ptr, cnt.data(), sizeof(int) * nb, hipMemcpyHostToDevice);
    e = e ? e : ksvd_copy(ctx, dperm.ptr, perm.data(), sizeof(int) * nb, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        cleanup();
        return ksvd_fail(ctx, KSVD_ERR_HIP, std::string("upload: ") + hipGetErrorString(e));
    }
    const __bf16 *Xb = (const __bf16 *)ctx->mlp_xb.ptr;
    if ((rc = mlp_forward_batch(ctx, Xb, nb))) {
        cleanup();
        return rc;
    }
    const int nblk = nb;
    hipLaunchKernelGGL(bce_grad_kernel, dim3((nb + 3) / 4), dim3(256), 0, ctx->stream, (const float *)ctx->mlp_zb.ptr, ko,
                       ctx->K, nb, (const int *)dperm.ptr, 0, (const int *)didx.ptr, (const float *)dcoef.ptr,
                       (const int *)dcnt.ptr, kmx, (__bf16 *)ctx->mlp_dz.ptr, (float *)ctx->mlp_lpart.ptr);
    if ((rc = mlp_backward_batch(ctx, Xb, nb))) {
        cleanup();
        return rc;
    }
    e = hipStreamSynchronize(ctx->stream);
    std::vector<float> lp(nblk);
    const int64_t n1 = (int64_t)MLP_H * kin, n2 = (int64_t)MLP_H * MLP_HP, n3 = (int64_t)ko * MLP_HP;
    std::vector<float> gp((size_t)MLP_SPLIT * (n1 + n2 + n3));
    e = e ? e : ksvd_copy(ctx, lp.data(), ctx->mlp_lpart.ptr, sizeof(float) * nblk, hipMemcpyDeviceToHost);
    e = e ? e : ksvd_copy(ctx, gp.data(), ctx->mlp_grad.ptr, sizeof(float) * gp.size(), hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) return ksvd_fail(ctx, KSVD_ERR_HIP, std::string("gradients: ") + hipGetErrorString(e));
    double l = 0.0;
    for (float v : lp) l += v;
    *loss = l / nb;
    const int64_t off[3] = {0, MLP_SPLIT * n1, MLP_SPLIT * (n1 + n2)}, len[3] = {n1, n2, n3};
    int64_t o = 0;
    for (int L = 0; L < 3; L++) {
        for (int64_t i = 0; i < len[L]; i++) {
            float s = 0.f;
            for (int sp = 0; sp < MLP_SPLIT; sp++) s += gp[off[L] + sp * len[L] + i];
            grad[o + i] = s;
        }
        o += len[L];
    }
    return KSVD_OK;
}

// Train `steps` AdamW steps on batches of the patches perm[0 .. n), taken in
// order (wrapping), with the current codes as targets.
static int mlp_train_list(ksvd_ctx *ctx, const int *perm, int64_t n, int steps, float lr)
{
    int kin, ko;
    mlp_dims(ctx, &kin, &ko);
    int rc = mlp_inputs(ctx);
    if (rc) return rc;
    const int B = MLP_BATCH;
    if (n < B) return KSVD_OK;
    for (int st = 0; st < steps; st++) {
        const int64_t b0 = ((int64_t)st * B) % (n - B + 1);
        const int c16 = kin / 8;
        hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)(((int64_t)B * c16 + 255) / 256)), dim3(256), 0,
                           ctx->stream, (const __bf16 *)ctx->mlp_x.ptr, kin, perm, (int)b0, B,
                           (__bf16 *)ctx->mlp_xb.ptr);
        const __bf16 *Xb = (const __bf16 *)ctx->mlp_xb.ptr;
        if ((rc = mlp_forward_batch(ctx, Xb, B))) return rc;
        const int nblk = B;
        hipLaunchKernelGGL(bce_grad_kernel, dim3((B + 3) / 4), dim3(256), 0, ctx->stream, (const float *)ctx->mlp_zb.ptr,
                           ko, ctx->K, B, perm, (int)b0, (const int *)ctx->idx.ptr, (const float *)ctx->coef.ptr,
                           (const int *)ctx->cnt.ptr, ctx->kmax, (__bf16 *)ctx->mlp_dz.ptr,
                           (float *)ctx->mlp_lpart.ptr);
        if ((rc = mlp_backward_batch(ctx, Xb, B))) return rc;
        if ((rc = mlp_adamw(ctx, lr))) return rc;
        if (st == steps - 1)
            hipLaunchKernelGGL(loss_sum_kernel, dim3(1), dim3(256), 0, ctx->stream, (const float *)ctx->mlp_lpart.ptr,
                               nblk, B, (float *)ctx->mlp_cnt.ptr + 1);
    }
    KSVD_HIP_CHECK(ctx, hipGetLastError());
    return KSVD_OK;
}

// A seeded permutation of all patches (host Fisher-Yates), for training on
// the current codes.
// the training order of seed `seed` into mlp_perm, unless it is there already
static int mlp_permutation(ksvd_ctx *ctx, uint64_t seed)
{
    const int64_t P = ctx->P;
    if (ctx->mlp_perm_n == P && ctx->mlp_perm_seed == seed) return KSVD_OK;
    std::vector<int> perm((size_t)P);
    for (int64_t i = 0; i < P; i++) perm[i] = (int)i;
    uint64_t s = seed;
    for (int64_t i = P - 1; i > 0; i--) {
        const int64_t j = (int64_t)(sm64(s) % (uint64_t)(i + 1));
        std::swap(perm[i], perm[j]);
    }
    ctx->mlp_perm_n = -1;
    int rc = ksvd_reserve(ctx, ctx->mlp_perm, sizeof(int) * (size_t)P);
    if (rc) return rc;
    KSVD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    KSVD_HIP_CHECK(ctx, ksvd_copy(ctx, ctx->mlp_perm.ptr, perm.data(), sizeof(int) * (size_t)P, hipMemcpyHostToDevice));
    ctx->mlp_perm_n = P;
    ctx->mlp_perm_seed = seed;
    return KSVD_OK;
}

int mlp_train_all(ksvd_ctx *ctx, int steps, float lr, uint64_t seed)
{
    int rc = mlp_alloc(ctx);
    if (rc) return rc;
    if ((rc = mlp_permutation(ctx, seed))) return rc;
    return mlp_train_list(ctx, (const int *)ctx->mlp_perm.ptr, ctx->P, steps, lr);
}

// Everything the learned form needs for an image that does not depend on the
// codes: the network's buffers and initial weights (seed init_seed, once per
// context), the normalised patches, the proposal passes' buffers and the
// first training order (seed perm_seed).  Run when the patches are
// extracted, so that the K-SVD iterations neither allocate nor wait for it.
int mlp_prepare(ksvd_ctx *ctx, uint64_t init_seed, uint64_t perm_seed)
{
    int rc = 0;
    if (!ctx->mlp_ready && (rc = mlp_init(ctx, init_seed))) return rc;
    if ((rc = mlp_inputs(ctx))) return rc;
    // the proposal passes' chunk buffers and fallback list, at their largest
    int kin, ko;
    mlp_dims(ctx, &kin, &ko);
    const int64_t CH = std::min<int64_t>(MLP_CHUNK, (ctx->P + 127) / 128 * 128);
    if ((rc = ksvd_reserve(ctx, ctx->mlp_h1c, sizeof(__bf16) * CH * MLP_HP))) return rc;
    if ((rc = ksvd_reserve(ctx, ctx->mlp_h2c, sizeof(__bf16) * CH * MLP_HP))) return rc;
    if ((rc = ksvd_reserve(ctx, ctx->mlp_zc, sizeof(__bf16) * CH * ko))) return rc;
    if ((rc = ksvd_reserve(ctx, ctx->mlp_fb, sizeof(int) * (size_t)ctx->P))) return rc;
    if ((rc = ksvd_reserve(ctx, ctx->mlp_rej, (size_t)ctx->P))) return rc;
    if ((rc = ksvd_reserve(ctx, ctx->mlp_bcnt, sizeof(int) * (size_t)((ctx->P + FB_B - 1) / FB_B)))) return rc;
    return mlp_permutation(ctx, perm_seed);
}

// Continual training on the last fallback list (patches whose codes plain
// OMP has just produced): up to `steps` steps; needs the list's length on
// the host (one synchronisation).
int mlp_train_fallback(ksvd_ctx *ctx, int steps, float lr)
{
    unsigned n = 0;
    KSVD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    KSVD_HIP_CHECK(ctx, ksvd_copy(ctx, &n, ctx->mlp_cnt.ptr, sizeof(unsigned), hipMemcpyDeviceToHost));
    const int usable = (int)std::min<int64_t>(steps, (int64_t)n / MLP_BATCH);
    if (usable <= 0) return KSVD_OK;
    return mlp_train_list(ctx, (const int *)ctx->mlp_fb.ptr, n, usable, lr);
}

// Proposals for every patch, in chunks whose activations stay in the
// last-level cache; rejected patches are listed in mlp_fb (count in
// mlp_cnt[0]).  The caller then runs OMP over the list.
int mlp_propose(ksvd_ctx *ctx, float accept, int64_t n, bool write)
{
    int kin, ko;
    mlp_dims(ctx, &kin, &ko);
    int rc = mlp_inputs(ctx);
    if (rc) return rc;
    const int64_t P = std::min<int64_t>(n, ctx->P);
    const int64_t CH = std::min<int64_t>(MLP_CHUNK, (P + 127) / 128 * 128);
    if ((rc = ksvd_reserve(ctx, ctx->mlp_h1c, sizeof(__bf16) * CH * MLP_HP))) return rc;
    if ((rc = ksvd_reserve(ctx, ctx->mlp_h2c, sizeof(__bf16) * CH * MLP_HP))) return rc;
    if ((rc = ksvd_reserve(ctx, ctx->mlp_zc, sizeof(__bf16) * CH * ko))) return rc;
    if ((rc = ksvd_reserve(ctx, ctx->mlp_fb, sizeof(int) * (size_t)P))) return rc;
    if ((rc = ksvd_reserve(ctx, ctx->mlp_rej, (size_t)P))) return rc;
    const int nfb = (int)((P + FB_B - 1) / FB_B);
    if ((rc = ksvd_reserve(ctx, ctx->mlp_bcnt, sizeof(int) * (size_t)nfb))) return rc;
    if (ctx->mlp_chunk_rows < CH) {
        hipLaunchKernelGGL(ones_column_kernel, dim3((unsigned)((CH + 255) / 256)), dim3(256), 0, ctx->stream,
                           (__bf16 *)ctx->mlp_h1c.ptr, MLP_HP, MLP_H, CH);
        hipLaunchKernelGGL(ones_column_kernel, dim3((unsigned)((CH + 255) / 256)), dim3(256), 0, ctx->stream,
                           (__bf16 *)ctx->mlp_h2c.ptr, MLP_HP, MLP_H, CH);
        ctx->mlp_chunk_rows = CH;
    }
    const __bf16 *W1 = (const __bf16 *)ctx->mlp_wb.ptr, *W2 = W1 + (size_t)MLP_H * kin,
                 *W3 = W2 + (size_t)MLP_H * MLP_HP;
    __bf16 *H1 = (__bf16 *)ctx->mlp_h1c.ptr, *H2 = (__bf16 *)ctx->mlp_h2c.ptr, *Z = (__bf16 *)ctx->mlp_zc.ptr;
    for (int64_t p0 = 0; p0 < P; p0 += CH) {
        const int nb = (int)std::min<int64_t>(CH, P - p0);
        const __bf16 *X = (const __bf16 *)ctx->mlp_x.ptr + p0 * kin;
        if ((rc = gemm<true, true, EPI_RELU_BF16>(ctx, X, kin, W1, kin, nb, MLP_H, kin, 1, nullptr, H1, MLP_HP, 0,
                                                  nullptr, 0)))
            return rc;
        if ((rc = gemm<true, true, EPI_RELU_BF16>(ctx, H1, MLP_HP, W2, MLP_HP, nb, MLP_H, MLP_HP, 1, nullptr, H2,
                                                  MLP_HP, 0, nullptr, 0)))
            return rc;
        if ((rc = gemm<true, true, EPI_BF16>(ctx, H2, MLP_HP, W3, MLP_HP, nb, ko, MLP_HP, 1, nullptr, Z, ko, 0,
                                             nullptr, 0)))
            return rc;
        const dim3 grid((unsigned)((nb + 15) / 16));
        auto launch = [&](auto kern) {
            hipLaunchKernelGGL(kern, grid, dim3(256), 0, ctx->stream, (const __bf16 *)Z, ko, ctx->K,
                               (const float *)ctx->Y.ptr, ctx->dim, p0, (int64_t)nb, (const float *)ctx->D.ptr,
                               (const float *)ctx->G.ptr, ctx->kmax, (const float *)ctx->rbound.ptr, accept,
                               write ? 1 : 0, (int *)ctx->idx.ptr, (float *)ctx->coef.ptr, (int *)ctx->cnt.ptr,
                               (unsigned char *)ctx->mlp_rej.ptr);
        };
        const int rj = (ctx->dim + 15) / 16;
        if (ctx->kmax <= 5) {
            if (rj <= 4) launch(propose_kernel<5, 4>);
            else launch(propose_kernel<5, 9>);
        } else {
            if (rj <= 4) launch(propose_kernel<8, 4>);
            else launch(propose_kernel<8, 9>);
        }
        KSVD_HIP_CHECK(ctx, hipGetLastError());
    }
    // the fallback count (and, when the codes are kept, the list)
    const unsigned char *rej = (const unsigned char *)ctx->mlp_rej.ptr;
    int *bcnt = (int *)ctx->mlp_bcnt.ptr;
    if (nfb > 0)
        hipLaunchKernelGGL(fb_count_kernel, dim3((unsigned)nfb), dim3(FB_B), 0, ctx->stream, rej, P, bcnt);
    hipLaunchKernelGGL(fb_scan_kernel, dim3(1), dim3(FB_B), 0, ctx->stream, bcnt, nfb, (unsigned *)ctx->mlp_cnt.ptr);
    if (write && nfb > 0)
        hipLaunchKernelGGL(fb_store_kernel, dim3((unsigned)nfb), dim3(FB_B), 0, ctx->stream, rej, P,
                           (const int *)bcnt, (int *)ctx->mlp_fb.ptr);
    KSVD_HIP_CHECK(ctx, hipGetLastError());
    return KSVD_OK;
}
