// Batched Orthogonal Matching Pursuit on gfx950.
//
// omp_mfma_kernel: persistent, one workgroup per CU, eight wavefronts (two
// per SIMD; four with OMP limits above 5).  The whole dictionary (K x dim
// FP32, atom-major, row pitch LD = 4Q + 4 floats so that the 16-byte LDS
// reads of one wavefront hit distinct banks) sits in LDS for the kernel's
// lifetime.  Each wavefront owns a tile of 16 patches at a time:
//   * lane (col, g), col = lane % 16, g = lane / 16, holds quarter g of patch
//     col: elements g*Q .. g*Q+Q-1 of y and of the residual r;
//   * step 0's correlation C0 (K x 16) = D^T Y is v_mfma_f32_16x16x4_f32
//     with the reduction index permuted so that MFMA number kk takes element
//     g*Q + kk from lane group g (A from LDS, B straight from y's registers);
//     later steps take D^T r = C0 - G[:, S] x from the dictionary Gram
//     (Batch-OMP), the rows of the support atoms read from L2 a few blocks
//     of atoms at a time, each block folded into the argmax at once;
//   * |C| argmax per patch: in-lane scan in ascending atom order, then two
//     cross-lane steps with the lowest index winning ties (cublasIsamax);
//   * the support Gram grows by one Cholesky row per step (entries from the
//     dictionary Gram G = D^T D), then two triangular solves give x; all four
//     lanes of a patch carry identical copies of S, L, x;
//   * ||r||^2 = ||y||^2 - b_S . x ends the step; where that is too small to
//     rule out ||r|| < 1e-3 (the retirement test), the explicit
//     r = y - D_S x decides.
// Semantics follow CuBlasMatrixOmp::work (cublas_matrix_omp.cu:107-402).
#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <mutex>

#include "ksvd_internal.hpp"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

// Wavefronts per workgroup (one workgroup per CU, the dictionary in its LDS):
// two per SIMD where the support's registers leave room (OMP limit <= 5), so
// one wavefront's MFMA correlation and L2 waits overlap the other's pursuit
// steps; one per SIMD with the 8 x 8 Cholesky factor.
constexpr int omp_waves(int km) { return km <= 5 ? 8 : 4; }

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

// Cross-lane sum over the four lanes that hold one patch (lanes col, col+16,
// col+32, col+48).  Every lane ends with the same bits.
__device__ __forceinline__ float patch_sum(float v)
{
    return add_xor32(add_xor16(v));
}

// Folds four correlations c of atoms a0 .. a0+3 into a running first-index
// argmax of |c| (ascending atom order within the lane).
__device__ __forceinline__ void amax_fold(const f4 &c, int a0, float &best, int &bi)
{
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float v = fabsf(c[i]);
        if (v > best) { best = v; bi = a0 + i; }
    }
}

// Step t > 0 of the pursuit, for supports of at most NR atoms: D^T r = C0 -
// G[:, S] x (Batch-OMP; the lane's atoms tt*16 + 4g .. +3, rows S_q of the
// symmetric G, support entries in order), two blocks of atoms at a time:
// every row's entries of the blocks are loaded together, corrected and
// folded into the argmax, so no full copy of the correlations is held beside
// C0.  All NR rows are read without branches; rows past a lane's support
// repeat its first row with weight 0, which leaves the sums as they are.
// FULL: K is the whole 16 KT.  Otherwise the reads run past row ends (into
// the next row, or the 256 floats of slack after G) and the entries of atoms
// >= K are replaced by 0, so their correlations stay 0 (zero rows of the LDS
// image in step 0).
template <int KT, int NR, bool FULL, int KMX>
__device__ __forceinline__ void gram_argmax(const f4 (&c0)[KT], const float *__restrict__ G, int K, const int (&S)[KMX],
                                            const float (&x)[KMX], int s, int g, float &best, int &bi)
{
    constexpr int CH = 2;
    float xq[NR];
    const float *gr[NR];
    const int s0 = s > 0 ? S[0] : 0;
#pragma unroll
    for (int q = 0; q < NR; q++) {
        xq[q] = q < s ? x[q] : 0.f;
        gr[q] = G + (int64_t)(q < s ? S[q] : s0) * K + 4 * g;
    }
#pragma unroll
    for (int tc = 0; tc < KT; tc += CH) {
        f4 a[CH];
#pragma unroll
        for (int c = 0; c < CH; c++) a[c] = c0[tc + c];
        f4 gv[NR][CH];
#pragma unroll
        for (int q = 0; q < NR; q++)
#pragma unroll
            for (int c = 0; c < CH; c++) {
                gv[q][c] = *(const f4 *)(gr[q] + 16 * (tc + c));
                if (!FULL) {
                    const int a0 = (tc + c) * 16 + 4 * g;
#pragma unroll
                    for (int i = 0; i < 4; i++) gv[q][c][i] = a0 + i < K ? gv[q][c][i] : 0.f;
                }
            }
#pragma unroll
        for (int q = 0; q < NR; q++)
#pragma unroll
            for (int c = 0; c < CH; c++)
#pragma unroll
                for (int i = 0; i < 4; i++) a[c][i] = fmaf(-gv[q][c][i], xq[q], a[c][i]);
#pragma unroll
        for (int c = 0; c < CH; c++) amax_fold(a[c], (tc + c) * 16 + 4 * g, best, bi);
    }
}

// Start of row i of a packed lower triangle.
__host__ __device__ constexpr int tri(int i) { return i * (i + 1) / 2; }

// Quarter-row of atom j (elements e0 .. e0 + 4Q - 1 of its row, 0 past dim)
// from global memory, 16 bytes at a time where they lie inside the row.
template <int Q>
__device__ __forceinline__ f4 drow4(const float *__restrict__ D, int dim, int j, int e)
{
    if (e + 3 < dim) return *(const f4 *)(D + (int64_t)j * dim + e);
    f4 v;
#pragma unroll
    for (int c = 0; c < 4; c++) v[c] = e + c < dim ? D[(int64_t)j * dim + e + c] : 0.f;
    return v;
}

// GC (global correlation): step 0's correlations come from C0 (row p of the
// launch's patches at C0 + p K, the GEMM of corr_gemm_kernel) instead of the
// MFMA, and the atoms from global memory instead of LDS, for patch
// dimensions whose dictionary does not fit the LDS (up to 4Q).
template <int Q, int KT, int KM, int W, bool FULL, bool GC = false>
__global__ __launch_bounds__(64 * W, 1) void omp_mfma_kernel(
    const float *__restrict__ Y, int dim, int64_t P, const float *__restrict__ D, int K,
    const float *__restrict__ G, int kmax, int ntiles, const int *__restrict__ plist,
    const unsigned *__restrict__ plist_n, int *__restrict__ idx_out, float *__restrict__ coef_out,
    int *__restrict__ cnt_out, float *__restrict__ rnorm_out, unsigned long long *__restrict__ counters,
    const float *__restrict__ C0 = nullptr)
{
    constexpr int LD = 4 * Q + 4;
    constexpr int KP = 16 * KT;
    constexpr int AH = W > 4 ? 4 : KT;  // blocks of A read from LDS at a time
    static_assert(W % 4 == 0, "the tile order takes whole sets of four wavefronts");
    extern __shared__ __attribute__((aligned(16))) float Ds[];  // KP x LD, zero beyond (K, dim)
    if (!GC) {
        for (int q = threadIdx.x; q < KP * LD; q += blockDim.x) {
            const int a = q / LD, e = q - a * LD;
            Ds[q] = (a < K && e < dim) ? D[(int64_t)a * dim + e] : 0.f;
        }
        __syncthreads();
    }

    unsigned long long ndup = 0, ndeg = 0;

    // list mode: the tiles cover plist[0 .. *plist_n)
    const int64_t nl = plist ? (int64_t)*plist_n : 0;
    const int ntl = plist ? (int)((nl + 15) / 16) : ntiles;
    // Tiles go round robin over the launch's wavefronts in an order that
    // takes one wavefront of every SIMD (wavefront w of a workgroup runs on
    // SIMD w % 4) before the second wavefront of any: the last, partial
    // round then adds at most one tile to any SIMD.
    const int wave = threadIdx.x >> 6;
    const int slot0 = (wave / 4) * (gridDim.x * 4) + blockIdx.x * 4 + (wave & 3);
    for (int tile = slot0; tile < ntl; tile += gridDim.x * W) {
        // The lane's coordinates pass through an opaque copy per tile, so
        // the LDS addresses and bounds masks derived from them are formed
        // where they are used instead of being hoisted out of the tile loop
        // and held (or spilled) across it.
        int lane = threadIdx.x & 63;
        asm volatile("" : "+v"(lane));
        const int col = lane & 15, g = lane >> 4;
        const int e0 = g * Q;
        const int64_t slot = (int64_t)tile * 16 + col;
        const int64_t p = plist ? (slot < nl ? (int64_t)plist[slot] : P) : slot;
        const bool valid = p < P;
        float y[Q];
#pragma unroll
        for (int kb = 0; kb < Q; kb += 4) {
            if (valid && e0 + kb + 3 < dim) {
                const f4 v = *(const f4 *)(Y + p * dim + e0 + kb);
#pragma unroll
                for (int c = 0; c < 4; c++) y[kb + c] = v[c];
            } else {
#pragma unroll
                for (int c = 0; c < 4; c++) y[kb + c] = (valid && e0 + kb + c < dim) ? Y[p * dim + e0 + kb + c] : 0.f;
            }
        }
        // ||y||^2, and the band below which the retirement test needs the
        // explicit residual (see the residual norm below)
        float yy = 0.f;
        if (Q > 64) {
            f4 y4 = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kk = 0; kk < Q; kk += 4)
#pragma unroll
                for (int c = 0; c < 4; c++) y4[c] = fmaf(y[kk + c], y[kk + c], y4[c]);
            yy = (y4[0] + y4[1]) + (y4[2] + y4[3]);
        } else {
#pragma unroll
            for (int kk = 0; kk < Q; kk++) yy = fmaf(y[kk], y[kk], yy);
        }
        yy = patch_sum(yy);
        const float band = 1e-6f + yy * 0x1p-10f + sqrtf(yy) * 0x1p-8f;
        // the Cholesky factor's lower triangle, row by row: entry (i, q),
        // q <= i, at tri(i) + q
        int S[KM];
        float x[KM], bv[KM], L[KM * (KM + 1) / 2];
#pragma unroll
        for (int i = 0; i < KM; i++) { S[i] = -1; x[i] = 0.f; bv[i] = 0.f; }
#pragma unroll
        for (int i = 0; i < KM * (KM + 1) / 2; i++) L[i] = 0.f;
        int s = 0;
        bool active = valid;
        float rlast = yy;  // squared residual norm of the current support

        f4 c0[KT];  // D^T y (step 0's correlation)
        for (int t = 0; t < kmax; t++) {
            if (__ballot(active) == 0) break;
            // again per step: nothing derived from the coordinates stays
            // live across the steps
            int lane = threadIdx.x & 63;
            asm volatile("" : "+v"(lane));
            const int col = lane & 15, g = lane >> 4;
            const int e0 = g * Q;
            float best = -1.f;
            int bi = 0x7fffffff;
            if (t == 0 && GC) {
                // ---- correlation C = D^T Y from the GEMM ----------------------
                const int64_t pc = valid ? p : 0;
#pragma unroll
                for (int tt = 0; tt < KT; tt++) {
                    const int a0 = tt * 16 + 4 * g;
                    if (FULL) {
                        c0[tt] = *(const f4 *)(C0 + pc * K + a0);
                    } else {
#pragma unroll
                        for (int i = 0; i < 4; i++) c0[tt][i] = a0 + i < K ? C0[pc * K + a0 + i] : 0.f;
                    }
                    if (!valid) c0[tt] = f4{0.f, 0.f, 0.f, 0.f};
                }
#pragma unroll
                for (int tt = 0; tt < KT; tt++) amax_fold(c0[tt], tt * 16 + 4 * g, best, bi);
            } else if (t == 0) {
                // ---- correlation C = D^T Y on MFMA ----------------------------
#pragma unroll
                for (int tt = 0; tt < KT; tt++) c0[tt] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kb = 0; kb < Q; kb += 4)
#pragma unroll
                    for (int th = 0; th < KT; th += AH) {
                        f4 av[AH];
#pragma unroll
                        for (int tt = 0; tt < AH; tt++)
                            av[tt] = *(const f4 *)(Ds + ((th + tt) * 16 + col) * LD + e0 + kb);
#pragma unroll
                        for (int sub = 0; sub < 4; sub++)
#pragma unroll
                            for (int tt = 0; tt < AH; tt++)
                                c0[th + tt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[tt][sub], y[kb + sub], c0[th + tt], 0, 0, 0);
                    }
#pragma unroll
                for (int tt = 0; tt < KT; tt++) amax_fold(c0[tt], tt * 16 + 4 * g, best, bi);
            } else {
                // ---- correlation from the Gram: D^T r = D^T y - G[:, S] x ------
                // A step t < kmax <= KM corrects by at most KM - 1 rows: two
                // code paths, one for the first two steps (at most two rows)
                // and one for the rest.
                if (t <= 2) gram_argmax<KT, 2, FULL>(c0, G, K, S, x, s, g, best, bi);
                else gram_argmax<KT, KM - 1, FULL>(c0, G, K, S, x, s, g, best, bi);
            }
            // ---- first index of max |c| across the patch's four lanes --------
#pragma unroll
            for (int off = 16; off <= 32; off <<= 1) {
                const float ob = __shfl_xor(best, off);
                const int oi = __shfl_xor(bi, off);
                if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
            }
            const int j = bi;
            // ---- b_j = d_j . y ------------------------------------------------
            float bp = 0.f;
            if (Q > 64) {
                // long quarters: four partial sums (elements 4i + c), added
                // in a fixed order, keep the rounding of a 4Q-term dot low
                f4 b4 = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kb = 0; kb < Q; kb += 4) {
                    const f4 dv = drow4<Q>(D, dim, j < K ? j : 0, e0 + kb);
#pragma unroll
                    for (int c = 0; c < 4; c++) b4[c] = fmaf(dv[c], y[kb + c], b4[c]);
                }
                bp = (b4[0] + b4[1]) + (b4[2] + b4[3]);
            } else {
                const float *dj = Ds + (j < KP ? j : 0) * LD + e0;
#pragma unroll
                for (int kb = 0; kb < Q; kb += 4) {
                    const f4 dv = GC ? drow4<Q>(D, dim, j < K ? j : 0, e0 + kb) : *(const f4 *)(dj + kb);
                    bp = fmaf(dv[0], y[kb + 0], bp);
                    bp = fmaf(dv[1], y[kb + 1], bp);
                    bp = fmaf(dv[2], y[kb + 2], bp);
                    bp = fmaf(dv[3], y[kb + 3], bp);
                }
            }
            bp = patch_sum(bp);
            // ---- support bookkeeping and Cholesky row ------------------------
            bool dup = false;
#pragma unroll
            for (int i = 0; i < KM; i++) dup |= (i < s) && (S[i] == j);
            if (active && dup) {
                active = false;
                if (g == 0) ndup++;
            }
            // a non-finite residual leaves no atom with a finite |c|: retire
            if (active && j >= K) {
                active = false;
                if (g == 0) ndeg++;
            }
            if (active) {
                // every Gram entry of the new Cholesky row, loaded together
                // (one round trip to the L2, not one per support entry)
                float gS[KM];
#pragma unroll
                for (int i = 0; i < KM; i++) gS[i] = G[(int64_t)(i < s ? S[i] : j) * K + j];
                float w[KM];
                float piv = G[(int64_t)j * K + j];
#pragma unroll
                for (int i = 0; i < KM; i++) {
                    w[i] = 0.f;
                    if (i < s) {
                        float gi = gS[i];
#pragma unroll
                        for (int q = 0; q < i; q++) gi = fmaf(-L[tri(i) + q], w[q], gi);
                        w[i] = gi / L[tri(i) + i];
                        piv = fmaf(-w[i], w[i], piv);
                    }
                }
                if (!(piv > 0.f)) {
                    active = false;
                    if (g == 0) ndeg++;
                } else {
                    const float lss = sqrtf(piv);
#pragma unroll
                    for (int i = 0; i < KM; i++)
                        if (i == s) {
#pragma unroll
                            for (int q = 0; q < i; q++) L[tri(i) + q] = w[q];
                            L[tri(i) + i] = lss;
                            S[i] = j;
                            bv[i] = bp;
                        }
                    s++;
                    float z[KM];
#pragma unroll
                    for (int i = 0; i < KM; i++) {
                        z[i] = 0.f;
                        if (i < s) {
                            float v = bv[i];
#pragma unroll
                            for (int q = 0; q < i; q++) v = fmaf(-L[tri(i) + q], z[q], v);
                            z[i] = v / L[tri(i) + i];
                        }
                    }
#pragma unroll
                    for (int i = KM - 1; i >= 0; i--) {
                        if (i < s) {
                            float v = z[i];
#pragma unroll
                            for (int q = i + 1; q < KM; q++)
                                if (q < s) v = fmaf(-L[tri(q) + i], x[q], v);
                            x[i] = v / L[tri(i) + i];
                        }
                    }
                }
            }
            // ---- residual norm -------------------------------------------------
            // For the least-squares x, ||r||^2 = ||y||^2 - b_S . x.  If x is
            // off by dx, the explicit ||y - D_S x||^2 exceeds the exact one by
            // dx^T G_S dx and the formula errs by at most ||y|| |G_S^1/2 dx|,
            // so an explicit norm below 1e-6 (the retirement test ||r|| <
            // 1e-3) always leaves the formula inside `band`.  Only there is
            // r = y - D_S x formed, and its norm decides as before.
            float nr = yy;
#pragma unroll
            for (int i = 0; i < KM; i++)
                if (i < s) nr = fmaf(-bv[i], x[i], nr);
            const bool near = active && nr < band;
            if (__ballot(near) != 0ull) {
                float ne = 0.f;
                if (near) {
#pragma unroll
                    for (int kb = 0; kb < Q; kb += 4) {
                        f4 a4 = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                        for (int i = 0; i < KM; i++)
                            if (i < s) {
                                const f4 dv = GC ? drow4<Q>(D, dim, S[i], e0 + kb) : *(const f4 *)(Ds + S[i] * LD + e0 + kb);
                                a4[0] = fmaf(dv[0], x[i], a4[0]);
                                a4[1] = fmaf(dv[1], x[i], a4[1]);
                                a4[2] = fmaf(dv[2], x[i], a4[2]);
                                a4[3] = fmaf(dv[3], x[i], a4[3]);
                            }
#pragma unroll
                        for (int c4 = 0; c4 < 4; c4++) {
                            const float rv = y[kb + c4] - a4[c4];
                            ne = fmaf(rv, rv, ne);
                        }
                    }
                }
                ne = patch_sum(ne);
                if (near) nr = ne;
            }
            if (active) rlast = nr;
            if (active && sqrtf(nr) < 1e-3f) active = false;
        }
        if (valid && g == 0) {
            if (rnorm_out) rnorm_out[p] = sqrtf(fmaxf(rlast, 0.f));
            cnt_out[p] = s;
#pragma unroll
            for (int i = 0; i < KM; i++)
                if (i < kmax) {
                    idx_out[p * kmax + i] = i < s ? S[i] : -1;
                    coef_out[p * kmax + i] = i < s ? x[i] : 0.f;
                }
        }
    }
    if (ndup) atomicAdd(&counters[CNT_DUP], ndup);
    if (ndeg) atomicAdd(&counters[CNT_DEGEN], ndeg);
}

// Pursuit for the shapes the MFMA kernel does not cover (dim > 144, K > 256
// or ompIter > 8): the first step's correlations C0 = D^T Y come from an
// MFMA GEMM over a chunk of patches, and a pursuit kernel, one wavefront per
// patch, takes every later step's D^T r = C0 - G[:, S] x from the dictionary
// Gram (Batch-OMP, as the MFMA kernel does).  Same semantics as
// omp_mfma_kernel: first index of max |c|, Cholesky rows from G, two
// triangular solves, ||r||^2 = ||y||^2 - b.x with the explicit residual only
// near zero, duplicate and degenerate stops.

// C0[p][a] = d_a . y_p on v_mfma_f32_16x16x4_f32.  Workgroup tile: 256
// atoms x 64 patches, so every patch is read from HBM once (D, at most 512 KB,
// stays in L2); wavefront w takes atoms [64w, 64w + 64) of the tile, sixteen
// 16 x 16 blocks.  The reduction runs over the patch dimension in chunks of
// 16 staged in LDS (zero past dim, K and P), double-buffered: the next
// chunk's loads are in flight while the MFMAs consume this one, and one
// barrier per chunk separates them.  Lane (col, g) holds elements
// [4g, 4g + 4) of a chunk row, read with one 16-byte LDS load; MFMA step s
// takes element 4g + s of A and B alike, so the four steps sum the chunk.
// VEC: dim and K are multiples of 4 (16-byte global loads and stores).
// TP: patches per tile, 64 or 128 (twice the MFMA work per LDS chunk of D,
// at two thirds of the occupancy); both give the same bits, since every
// correlation is summed over the same chunks and steps in the same order,
// and the first call for a shape times both and keeps the faster.
constexpr int CG_TA = 256, CG_KC = 16, CG_LD = CG_KC + 4;

template <bool VEC>
__device__ __forceinline__ f4 cg_load4(const float *__restrict__ row, int e, int dim, bool ok)
{
    f4 v = f4{0.f, 0.f, 0.f, 0.f};
    if (VEC) {
        if (ok && e < dim) v = *(const f4 *)(row + e);
    } else if (ok) {
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (e + i < dim) v[i] = row[e + i];
    }
    return v;
}

template <bool VEC, int TP>
__global__ __launch_bounds__(256) void corr_gemm_kernel(const float *__restrict__ Y, int dim, int64_t P,
                                                        const float *__restrict__ D, int K, float *__restrict__ C0)
{
    constexpr int NB = TP / 64;  // patch rows this thread loads per chunk
    constexpr int PB = TP / 16;  // 16-patch blocks of each wavefront
    __shared__ __attribute__((aligned(16))) float As[2][CG_TA][CG_LD];
    __shared__ __attribute__((aligned(16))) float Bs[2][TP][CG_LD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 15, g = lane >> 4;
    const int a0 = blockIdx.y * CG_TA;
    const int64_t p0 = (int64_t)blockIdx.x * TP;
    // this thread's loads: rows tid / 4 + 64 j of the atom tile and of the
    // patch tile, elements [4 (tid & 3), +4) of each chunk
    const int lr = tid >> 2, lc = 4 * (tid & 3);
    const float *arow[4];
    bool aok[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int at = a0 + lr + 64 * j;
        aok[j] = at < K;
        arow[j] = D + (int64_t)(aok[j] ? at : 0) * dim;
    }
    const float *brow[NB];
    bool bok[NB];
#pragma unroll
    for (int j = 0; j < NB; j++) {
        bok[j] = p0 + lr + 64 * j < P;
        brow[j] = Y + (bok[j] ? p0 + lr + 64 * j : 0) * dim;
    }
    f4 acc[4][PB];
#pragma unroll
    for (int ab = 0; ab < 4; ab++)
#pragma unroll
        for (int pb = 0; pb < PB; pb++) acc[ab][pb] = f4{0.f, 0.f, 0.f, 0.f};
    f4 pa[4], pbv[NB];
#pragma unroll
    for (int j = 0; j < 4; j++) pa[j] = cg_load4<VEC>(arow[j], lc, dim, aok[j]);
#pragma unroll
    for (int j = 0; j < NB; j++) pbv[j] = cg_load4<VEC>(brow[j], lc, dim, bok[j]);
#pragma unroll
    for (int j = 0; j < 4; j++) *(f4 *)&As[0][lr + 64 * j][lc] = pa[j];
#pragma unroll
    for (int j = 0; j < NB; j++) *(f4 *)&Bs[0][lr + 64 * j][lc] = pbv[j];
    __syncthreads();
    const int nch = (dim + CG_KC - 1) / CG_KC;
    for (int c = 0; c < nch; c++) {
        const int cur = c & 1;
        const bool more = c + 1 < nch;
        if (more) {
            const int e = (c + 1) * CG_KC + lc;
#pragma unroll
            for (int j = 0; j < 4; j++) pa[j] = cg_load4<VEC>(arow[j], e, dim, aok[j]);
#pragma unroll
            for (int j = 0; j < NB; j++) pbv[j] = cg_load4<VEC>(brow[j], e, dim, bok[j]);
        }
        f4 av[4], bv[PB];
#pragma unroll
        for (int ab = 0; ab < 4; ab++) av[ab] = *(const f4 *)&As[cur][wave * 64 + ab * 16 + col][4 * g];
#pragma unroll
        for (int pb = 0; pb < PB; pb++) bv[pb] = *(const f4 *)&Bs[cur][pb * 16 + col][4 * g];
#pragma unroll
        for (int st = 0; st < 4; st++)
#pragma unroll
            for (int ab = 0; ab < 4; ab++)
#pragma unroll
                for (int pb = 0; pb < PB; pb++)
                    acc[ab][pb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[ab][st], bv[pb][st], acc[ab][pb], 0, 0, 0);
        if (more) {
#pragma unroll
            for (int j = 0; j < 4; j++) *(f4 *)&As[cur ^ 1][lr + 64 * j][lc] = pa[j];
#pragma unroll
            for (int j = 0; j < NB; j++) *(f4 *)&Bs[cur ^ 1][lr + 64 * j][lc] = pbv[j];
        }
        __syncthreads();
    }
    // lane (col, g) holds C[atom 4g + i][patch col] of each block
#pragma unroll
    for (int ab = 0; ab < 4; ab++)
#pragma unroll
        for (int pb = 0; pb < PB; pb++) {
            const int at = a0 + wave * 64 + ab * 16 + 4 * g;
            const int64_t pp = p0 + pb * 16 + col;
            if (pp >= P) continue;
            if (VEC && at + 3 < K) {
                *(f4 *)(C0 + pp * K + at) = acc[ab][pb];
            } else {
#pragma unroll
                for (int i = 0; i < 4; i++)
                    if (at + i < K) C0[pp * K + at + i] = acc[ab][pb][i];
            }
        }
}

// Fixed butterfly over the wavefront: every lane ends with the same bits.
__device__ __forceinline__ float wave_sum_xor(float v)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) v += __shfl_xor(v, off);
    return v;
}

// One wavefront per patch: lane l holds atoms l, l + 64, ... (CPL of them) of
// the correlations and elements l, l + 64, ... (RJ of them) of the patch.
template <int CPL, int RJ, int KM>
__global__ __launch_bounds__(256) void omp_batch_kernel(
    const float *__restrict__ Y, int dim, int64_t P, const float *__restrict__ D, int K,
    const float *__restrict__ G, const float *__restrict__ C0, int kmax, int *__restrict__ idx_out,
    float *__restrict__ coef_out, int *__restrict__ cnt_out, unsigned long long *__restrict__ counters)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned long long ndup = 0, ndeg = 0;
    for (int64_t p = (int64_t)blockIdx.x * 4 + wave; p < P; p += (int64_t)gridDim.x * 4) {
        float y[RJ];
#pragma unroll
        for (int jj = 0; jj < RJ; jj++) {
            const int e = lane + 64 * jj;
            y[jj] = e < dim ? Y[p * dim + e] : 0.f;
        }
        float yy = 0.f;
#pragma unroll
        for (int jj = 0; jj < RJ; jj++) yy = fmaf(y[jj], y[jj], yy);
        yy = wave_sum_xor(yy);
        const float band = 1e-6f + yy * 0x1p-10f + sqrtf(yy) * 0x1p-8f;
        float c0[CPL];
#pragma unroll
        for (int i = 0; i < CPL; i++) {
            const int at = lane + 64 * i;
            c0[i] = at < K ? C0[p * K + at] : 0.f;
        }
        int S[KM];
        float x[KM], bv[KM], L[KM * KM];
#pragma unroll
        for (int i = 0; i < KM; i++) { S[i] = -1; x[i] = 0.f; bv[i] = 0.f; }
#pragma unroll
        for (int i = 0; i < KM * KM; i++) L[i] = 0.f;
        int s = 0;
        for (int t = 0; t < kmax; t++) {
            // D^T r = C0 - G[:, S] x, support entries in order
            float c[CPL];
#pragma unroll
            for (int i = 0; i < CPL; i++) c[i] = c0[i];
#pragma unroll
            for (int q = 0; q < KM; q++)
                if (q < s) {
                    const float *gr = G + (int64_t)S[q] * K;
#pragma unroll
                    for (int i = 0; i < CPL; i++) {
                        const int at = lane + 64 * i;
                        if (at < K) c[i] = fmaf(-gr[at], x[q], c[i]);
                    }
                }
            // first index of max |c| (cublasIsamax)
            float best = -1.f;
            int bi = 0x7fffffff;
#pragma unroll
            for (int i = 0; i < CPL; i++) {
                const int at = lane + 64 * i;
                if (at < K) {
                    const float v = fabsf(c[i]);
                    if (v > best) { best = v; bi = at; }
                }
            }
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const float ob = __shfl_xor(best, off);
                const int oi = __shfl_xor(bi, off);
                if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
            }
            const int j = bi;
            bool dup = false;
#pragma unroll
            for (int i = 0; i < KM; i++) dup |= (i < s) && (S[i] == j);
            if (dup) { if (lane == 0) ndup++; break; }
            // a non-finite residual leaves no atom with a finite |c|: retire
            if (j >= K) { if (lane == 0) ndeg++; break; }
            // b_j = d_j . y
            float bp = 0.f;
#pragma unroll
            for (int jj = 0; jj < RJ; jj++) {
                const int e = lane + 64 * jj;
                if (e < dim) bp = fmaf(D[(int64_t)j * dim + e], y[jj], bp);
            }
            bp = wave_sum_xor(bp);
            // Cholesky row of the support Gram from the dictionary Gram
            float gS[KM];
#pragma unroll
            for (int i = 0; i < KM; i++) gS[i] = G[(int64_t)(i < s ? S[i] : j) * K + j];
            float w[KM];
            float piv = G[(int64_t)j * K + j];
#pragma unroll
            for (int i = 0; i < KM; i++) {
                w[i] = 0.f;
                if (i < s) {
                    float gi = gS[i];
#pragma unroll
                    for (int q = 0; q < i; q++) gi = fmaf(-L[i * KM + q], w[q], gi);
                    w[i] = gi / L[i * KM + i];
                    piv = fmaf(-w[i], w[i], piv);
                }
            }
            if (!(piv > 0.f)) { if (lane == 0) ndeg++; break; }
            const float lss = sqrtf(piv);
#pragma unroll
            for (int i = 0; i < KM; i++)
                if (i == s) {
#pragma unroll
                    for (int q = 0; q < i; q++) L[i * KM + q] = w[q];
                    L[i * KM + i] = lss;
                    S[i] = j;
                    bv[i] = bp;
                }
            s++;
            float z[KM];
#pragma unroll
            for (int i = 0; i < KM; i++) {
                z[i] = 0.f;
                if (i < s) {
                    float v = bv[i];
#pragma unroll
                    for (int q = 0; q < i; q++) v = fmaf(-L[i * KM + q], z[q], v);
                    z[i] = v / L[i * KM + i];
                }
            }
#pragma unroll
            for (int i = KM - 1; i >= 0; i--)
                if (i < s) {
                    float v = z[i];
#pragma unroll
                    for (int q = i + 1; q < KM; q++)
                        if (q < s) v = fmaf(-L[q * KM + i], x[q], v);
                    x[i] = v / L[i * KM + i];
                }
            // ||r||^2 = ||y||^2 - b.x; the explicit residual where that is
            // too small to rule out ||r|| < 1e-3
            float nr = yy;
#pragma unroll
            for (int i = 0; i < KM; i++)
                if (i < s) nr = fmaf(-bv[i], x[i], nr);
            if (nr < band) {
                float ne = 0.f;
#pragma unroll
                for (int jj = 0; jj < RJ; jj++) {
                    const int e = lane + 64 * jj;
                    float a = 0.f;
#pragma unroll
                    for (int i = 0; i < KM; i++)
                        if (i < s && e < dim) a = fmaf(D[(int64_t)S[i] * dim + e], x[i], a);
                    const float rv = e < dim ? y[jj] - a : 0.f;
                    ne = fmaf(rv, rv, ne);
                }
                nr = wave_sum_xor(ne);
            }
            if (sqrtf(nr) < 1e-3f) break;
        }
        if (lane < kmax) {
            int si = -1;
            float xi = 0.f;
#pragma unroll
            for (int q = 0; q < KM; q++)
                if (q == lane && q < s) { si = S[q]; xi = x[q]; }
            idx_out[p * kmax + lane] = si;
            coef_out[p * kmax + lane] = xi;
        }
        if (lane == 0) cnt_out[p] = s;
    }
    if (ndup) atomicAdd(&counters[CNT_DUP], ndup);
    if (ndeg) atomicAdd(&counters[CNT_DEGEN], ndeg);
}

template <int Q, int KM>
int launch_mfma(ksvd_ctx *ctx, const int *plist, const unsigned *plist_n, float *rnorm)
{
    constexpr int KT = 16;
    constexpr int LD = 4 * Q + 4;
    const size_t lds = sizeof(float) * (size_t)(16 * KT) * LD;
    constexpr int W = omp_waves(KM);
    // whole Gram rows (no per-lane bounds) when K fills the 256-atom tile
    auto kern = ctx->K == 16 * KT ? omp_mfma_kernel<Q, KT, KM, W, true> : omp_mfma_kernel<Q, KT, KM, W, false>;
    KSVD_HIP_CHECK(ctx, allow_max_lds((const void *)kern));
    const int ntiles = (int)((ctx->P + 15) / 16);
    int grid = (ntiles + W - 1) / W;
    if (grid > ctx->num_cus) grid = ctx->num_cus;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * W), lds, ctx->stream, (const float *)ctx->Y.ptr,
                       ctx->dim, ctx->P, (const float *)ctx->D.ptr, ctx->K, (const float *)ctx->G.ptr, ctx->kmax,
                       ntiles, plist, plist_n, (int *)ctx->idx.ptr, (float *)ctx->coef.ptr, (int *)ctx->cnt.ptr,
                       rnorm, (unsigned long long *)ctx->counters.ptr, (const float *)nullptr);
    KSVD_HIP_CHECK(ctx, hipGetLastError());
    return KSVD_OK;
}

template <int CPL, int RJ, int KM>
int launch_batch_t(ksvd_ctx *ctx, int64_t p0, int64_t np, float *C0)
{
    const int64_t waves = np;
    int64_t grid = (waves + 3) / 4;
    if (grid > 8 * (int64_t)ctx->num_cus) grid = 8 * ctx->num_cus;
    const int kmax = ctx->kmax, dim = ctx->dim;
    hipLaunchKernelGGL((omp_batch_kernel<CPL, RJ, KM>), dim3((unsigned)grid), dim3(256), 0, ctx->stream,
                       (const float *)ctx->Y.ptr + p0 * dim, dim, np, (const float *)ctx->D.ptr, ctx->K,
                       (const float *)ctx->G.ptr, (const float *)C0, kmax, (int *)ctx->idx.ptr + p0 * kmax,
                       (float *)ctx->coef.ptr + p0 * kmax, (int *)ctx->cnt.ptr + p0,
                       (unsigned long long *)ctx->counters.ptr);
    KSVD_HIP_CHECK(ctx, hipGetLastError());
    return KSVD_OK;
}

template <int CPL, int RJ>
int launch_batch_k(ksvd_ctx *ctx, int64_t p0, int64_t np, float *C0)
{
    return ctx->kmax <= 8 ? launch_batch_t<CPL, RJ, 8>(ctx, p0, np, C0)
                          : launch_batch_t<CPL, RJ, KSVD_MAX_SUPPORT>(ctx, p0, np, C0);
}

template <int CPL>
int launch_batch_c(ksvd_ctx *ctx, int64_t p0, int64_t np, float *C0)
{
    const int dim = ctx->dim;
    if (dim <= 256) return launch_batch_k<CPL, 4>(ctx, p0, np, C0);
    if (dim <= 512) return launch_batch_k<CPL, 8>(ctx, p0, np, C0);
    return launch_batch_k<CPL, 16>(ctx, p0, np, C0);
}

// The tile pursuit (16 patches a wavefront) on a chunk of patches whose
// step-0 correlations are in C0: omp_mfma_kernel's steps, the atoms from
// global memory.
template <int Q, int KM>
int launch_tiles_gc(ksvd_ctx *ctx, int64_t p0, int64_t np, const float *C0)
{
    constexpr int KT = 16;
    constexpr int W = Q > 48 ? 4 : omp_waves(KM);  // a quarter above 48 elements needs a SIMD's registers
    const int kmax = ctx->kmax, dim = ctx->dim;
    auto kern = ctx->K == 16 * KT ? omp_mfma_kernel<Q, KT, KM, W, true, true>
                                  : omp_mfma_kernel<Q, KT, KM, W, false, true>;
    const int ntiles = (int)((np + 15) / 16);
    int grid = (ntiles + W - 1) / W;
    if (grid > ctx->num_cus) grid = ctx->num_cus;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * W), 0, ctx->stream, (const float *)ctx->Y.ptr + p0 * dim, dim, np,
                       (const float *)ctx->D.ptr, ctx->K, (const float *)ctx->G.ptr, kmax, ntiles, (const int *)nullptr,
                       (const unsigned *)nullptr, (int *)ctx->idx.ptr + p0 * kmax, (float *)ctx->coef.ptr + p0 * kmax,
                       (int *)ctx->cnt.ptr + p0, (float *)nullptr, (unsigned long long *)ctx->counters.ptr, C0);
    KSVD_HIP_CHECK(ctx, hipGetLastError());
    return KSVD_OK;
}

// Where the tile pursuit takes the GEMM's correlations: patch dims 145-496
// (a quarter of up to 124 elements a lane), atoms up to 256, OMP limit up to
// 8; KSVD_OMP_WAVE_PER_PATCH=1 keeps the one-wavefront-per-patch pursuit.
static bool tiles_gc_ok(const ksvd_ctx *ctx)
{
    const char *v = getenv("KSVD_OMP_WAVE_PER_PATCH");
    return ctx->dim <= 496 && ctx->K <= 256 && ctx->kmax <= 8 && !(v && v[0] == '1');
}

static int launch_tiles_gc_any(ksvd_ctx *ctx, int64_t p0, int64_t np, const float *C0)
{
    const int dim = ctx->dim;
    if (ctx->kmax <= 5)
        return dim <= 192   ? launch_tiles_gc<48, 5>(ctx, p0, np, C0)
               : dim <= 256 ? launch_tiles_gc<64, 5>(ctx, p0, np, C0)
                            : launch_tiles_gc<124, 5>(ctx, p0, np, C0);
    return dim <= 192   ? launch_tiles_gc<48, 8>(ctx, p0, np, C0)
           : dim <= 256 ? launch_tiles_gc<64, 8>(ctx, p0, np, C0)
                        : launch_tiles_gc<124, 8>(ctx, p0, np, C0);
}

// The step-0 GEMM over np patches at Yc into C0.  16-byte loads and stores
// when the rows of Y, D and C0 stay 16-byte aligned.  The patch tile (64 or
// 128, the same bits) is chosen per (dim, K, vec) by timing both on the
// first call for that shape in the process (KSVD_GEMM_TP=64 or 128 forces
// one); a stream under capture takes 64 without timing.
std::mutex cg_mu;
int cg_pick[2][KSVD_MAX_DIM + 1][2];  // [vec][dim][K <= 256]: 0 unknown, else the tile

int launch_corr_gemm(ksvd_ctx *ctx, const float *Yc, int64_t np, float *C0)
{
    const int K = ctx->K, dim = ctx->dim;
    const bool vec = dim % 4 == 0 && K % 4 == 0;
    auto grid = [&](int tp) { return dim3((unsigned)((np + tp - 1) / tp), (unsigned)((K + CG_TA - 1) / CG_TA)); };
    auto run = [&](int tp) {
        auto kern = tp == 128 ? (vec ? corr_gemm_kernel<true, 128> : corr_gemm_kernel<false, 128>)
                              : (vec ? corr_gemm_kernel<true, 64> : corr_gemm_kernel<false, 64>);
        hipLaunchKernelGGL(kern, grid(tp), dim3(256), 0, ctx->stream, Yc, dim, np, (const float *)ctx->D.ptr, K, C0);
    };
    int tp = 0;
    const char *ev = getenv("KSVD_GEMM_TP");
    if (ev && (atoi(ev) == 64 || atoi(ev) == 128)) tp = atoi(ev);
    const bool known_shape = dim <= KSVD_MAX_DIM;
    if (!tp && known_shape) {
        std::lock_guard<std::mutex> lock(cg_mu);
        tp = cg_pick[vec][dim][K > 256];
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    KSVD_HIP_CHECK(ctx, hipStreamIsCapturing(ctx->stream, &cs));
    if (!tp && (cs != hipStreamCaptureStatusNone || !known_shape || np < 16 * 1024)) tp = 64;
    if (tp) {
        run(tp);
        KSVD_HIP_CHECK(ctx, hipGetLastError());
        return KSVD_OK;
    }
    // first call for this shape: both tiles, timed with events (each run
    // writes the same C0), the faster kept
    hipEvent_t ev3[3];
    for (auto &e : ev3) KSVD_HIP_CHECK(ctx, hipEventCreate(&e));
    float ms[2] = {0.f, 0.f};
    hipError_t he = hipEventRecord(ev3[0], ctx->stream);
    run(64);
    if (he == hipSuccess) he = hipEventRecord(ev3[1], ctx->stream);
    run(128);
    if (he == hipSuccess) he = hipEventRecord(ev3[2], ctx->stream);
    if (he == hipSuccess) he = hipGetLastError();
    if (he == hipSuccess) he = hipEventSynchronize(ev3[2]);
    if (he == hipSuccess) he = hipEventElapsedTime(&ms[0], ev3[0], ev3[1]);
    if (he == hipSuccess) he = hipEventElapsedTime(&ms[1], ev3[1], ev3[2]);
    for (auto &e : ev3) (void)hipEventDestroy(e);
    KSVD_HIP_CHECK(ctx, he);
    std::lock_guard<std::mutex> lock(cg_mu);
    cg_pick[vec][dim][K > 256] = ms[1] < ms[0] ? 128 : 64;
    return KSVD_OK;
}

// The GEMM-correlation pursuit over every patch, in chunks whose C0 stays
// under 128 MB.
int launch_batch(ksvd_ctx *ctx)
{
    const int K = ctx->K;
    const int64_t P = ctx->P;
    int64_t chunk = ((int64_t)32 << 20) / K;
    chunk = std::max<int64_t>(128, chunk / 128 * 128);  // whole tiles of either shape
    if (chunk > P) chunk = P;
    int rc = ksvd_reserve(ctx, ctx->c0, sizeof(float) * (size_t)chunk * K);
    if (rc) return rc;
    float *C0 = (float *)ctx->c0.ptr;
    for (int64_t p0 = 0; p0 < P; p0 += chunk) {
        const int64_t np = std::min(chunk, P - p0);
        if ((rc = launch_corr_gemm(ctx, (const float *)ctx->Y.ptr + p0 * ctx->dim, np, C0))) return rc;
        if (tiles_gc_ok(ctx))
            rc = launch_tiles_gc_any(ctx, p0, np, C0);
        else
            rc = K <= 256 ? launch_batch_c<4>(ctx, p0, np, C0) : launch_batch_c<16>(ctx, p0, np, C0);
        if (rc) return rc;
    }
    return KSVD_OK;
}

}  // namespace

// True when the MFMA pursuit kernel covers the context's shapes.
static bool mfma_ok(const ksvd_ctx *ctx) { return ctx->dim <= 144 && ctx->K <= 256 && ctx->kmax <= 8; }

// The FP32 pursuit over every patch, or over the patches plist[0 ..
// *plist_n); each patch's final residual norm into rnorm when it is given.
static int launch_mfma_any(ksvd_ctx *ctx, const int *plist, const unsigned *plist_n, float *rnorm)
{
    // the support's registers (S, x, the Cholesky factor) are sized by KM:
    // 5 covers the usual OMP limit with a 5 x 5 factor instead of 8 x 8
    if (ctx->kmax <= 5)
        return ctx->dim <= 64 ? launch_mfma<16, 5>(ctx, plist, plist_n, rnorm)
                              : launch_mfma<36, 5>(ctx, plist, plist_n, rnorm);
    return ctx->dim <= 64 ? launch_mfma<16, 8>(ctx, plist, plist_n, rnorm)
                          : launch_mfma<36, 8>(ctx, plist, plist_n, rnorm);
}

bool omp_uses_fp32(const ksvd_ctx *ctx) { return ctx->prm.precision == KSVD_PREC_FP32 || ctx->prm.proposer; }

int launch_omp_fp32_list(ksvd_ctx *ctx, const int *plist, const unsigned *plist_n)
{
    if (!mfma_ok(ctx)) return ksvd_fail(ctx, KSVD_ERR_UNSUPPORTED, "the listed pursuit needs dim <= 144, atoms <= 256, ompIter <= 8");
    return launch_mfma_any(ctx, plist, plist_n, (float *)ctx->rbound.ptr);
}

int launch_omp(ksvd_ctx *ctx)
{
    if (!ctx->gram_valid) {
        int rc = launch_dict_gram(ctx);
        if (rc) return rc;
    }
    if (!omp_uses_fp32(ctx)) return launch_omp_bf16(ctx);
    // the learned form keeps each patch's residual norm: its proposals are
    // accepted against it
    if (mfma_ok(ctx)) return launch_mfma_any(ctx, nullptr, nullptr, ctx->prm.proposer ? (float *)ctx->rbound.ptr : nullptr);
    return launch_batch(ctx);
}

// ---- allocation-free pursuit (ksvd_omp_noalloc) ---------------------------
// For callers that cannot allocate device memory (code under HIP graph
// capture, serving loops).  The only device memory beyond the caller's inputs
// and outputs is `scratch`: the dictionary Gram (K x K floats, K <= 256) and
// the pursuit's counters.  Each GPU holds its own copy until the process
// ends; every call on that GPU reuses it, so calls that overlap on one GPU
// race.  Host code never passes the array's name to a kernel: it takes the
// device address hipGetSymbolAddress returns, once per device.
namespace {

constexpr size_t NOALLOC_GRAM_BYTES = sizeof(float) * 256 * 256;
constexpr size_t NOALLOC_SCRATCH_BYTES = NOALLOC_GRAM_BYTES + sizeof(unsigned long long) * CNT_SLOTS;
static_assert(NOALLOC_SCRATCH_BYTES == KSVD_NOALLOC_SCRATCH_BYTES, "scratch size differs from the header's");

__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mu;
unsigned char *noalloc_ptr[NOALLOC_MAX_DEVICES];  // device address of `scratch` per device
int noalloc_cus[NOALLOC_MAX_DEVICES];
thread_local char noalloc_err[256];

__global__ void noalloc_clear_kernel(unsigned long long *c, int n)
{
    if ((int)threadIdx.x < n) c[threadIdx.x] = 0ull;
}

int noalloc_fail(int code, const char *msg, hipError_t e = hipSuccess)
{
    if (e != hipSuccess) snprintf(noalloc_err, sizeof(noalloc_err), "ksvd_omp_noalloc: %s: %s", msg, hipGetErrorString(e));
    else snprintf(noalloc_err, sizeof(noalloc_err), "ksvd_omp_noalloc: %s", msg);
    return code;
}

template <int Q, int KM>
hipError_t noalloc_launch(const float *Y, int dim, int64_t P, const float *D, int K, const float *G, int kmax,
                          int *idx, float *coef, int *cnt, unsigned long long *counters, int ncu, hipStream_t st)
{
    constexpr int KT = 16;
    constexpr int LD = 4 * Q + 4;
    const size_t lds = sizeof(float) * (size_t)(16 * KT) * LD;
    const int ntiles = (int)((P + 15) / 16);
    constexpr int W = omp_waves(KM);
    int grid = (ntiles + W - 1) / W;
    if (grid > ncu) grid = ncu;
    auto kern = K == 16 * KT ? omp_mfma_kernel<Q, KT, KM, W, true> : omp_mfma_kernel<Q, KT, KM, W, false>;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * W), lds, st, Y, dim, P, D, K, G, kmax, ntiles,
                       (const int *)nullptr, (const unsigned *)nullptr, idx, coef, cnt, (float *)nullptr, counters,
                       (const float *)nullptr);
    return hipGetLastError();
}

// Once per device, with that device current: the address of its copy of
// `scratch`, its CU count, and the pursuit kernels' LDS ceiling.
int noalloc_setup(int dev, unsigned char **ptr, int *ncu)
{
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (!noalloc_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return noalloc_fail(KSVD_ERR_HIP, "hipGetSymbolAddress(scratch)", e);
        int cus = 0;
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return noalloc_fail(KSVD_ERR_HIP, "hipDeviceGetAttribute", e);
        const void *kerns[] = {(const void *)omp_mfma_kernel<16, 16, 5, omp_waves(5), true>,
                               (const void *)omp_mfma_kernel<36, 16, 5, omp_waves(5), true>,
                               (const void *)omp_mfma_kernel<16, 16, 8, omp_waves(8), true>,
                               (const void *)omp_mfma_kernel<36, 16, 8, omp_waves(8), true>,
                               (const void *)omp_mfma_kernel<16, 16, 5, omp_waves(5), false>,
                               (const void *)omp_mfma_kernel<36, 16, 5, omp_waves(5), false>,
                               (const void *)omp_mfma_kernel<16, 16, 8, omp_waves(8), false>,
                               (const void *)omp_mfma_kernel<36, 16, 8, omp_waves(8), false>};
        for (const void *k : kerns)
            if ((e = allow_max_lds(k)) != hipSuccess) return noalloc_fail(KSVD_ERR_HIP, "hipFuncSetAttribute", e);
        noalloc_cus[dev] = cus > 0 ? cus : 256;
        noalloc_ptr[dev] = (unsigned char *)p;
    }
    *ptr = noalloc_ptr[dev];
    *ncu = noalloc_cus[dev];
    return KSVD_OK;
}

}  // namespace

const char *noalloc_last_error() { return noalloc_err[0] ? noalloc_err : nullptr; }

extern "C" int ksvd_omp_noalloc(const float *Y, int dim, int64_t P, const float *D, int K, int kmax, int32_t *idx,
                                float *coef, int32_t *cnt, void *stream)
{
    noalloc_err[0] = 0;
    if (!Y || !D || !idx || !coef || !cnt) return noalloc_fail(KSVD_ERR_ARG, "null pointer argument");
    if (dim < 1 || K < 1 || kmax < 1 || P < 0) return noalloc_fail(KSVD_ERR_ARG, "dim, atoms, ompIter must be >= 1 and patches >= 0");
    if (dim > 144 || K > 256 || kmax > 8)
        return noalloc_fail(KSVD_ERR_UNSUPPORTED,
                            "sizes beyond the scratch and kernel maximum (dim <= 144, atoms <= 256, ompIter <= 8)");
    if ((P + 15) / 16 > (int64_t)INT32_MAX) return noalloc_fail(KSVD_ERR_ARG, "too many patches");
    if (P == 0) return KSVD_OK;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return noalloc_fail(KSVD_ERR_HIP, "hipGetDevice", e);
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return noalloc_fail(KSVD_ERR_UNSUPPORTED, "device ordinal beyond 63");
    unsigned char *ptr = nullptr;
    int ncu = 0;
    int rc = noalloc_setup(dev, &ptr, &ncu);
    if (rc) return rc;
    float *G = (float *)ptr;
    unsigned long long *counters = (unsigned long long *)(ptr + NOALLOC_GRAM_BYTES);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(noalloc_clear_kernel, dim3(1), dim3(64), 0, st, counters, (int)CNT_SLOTS);
    if ((e = hipGetLastError()) != hipSuccess) return noalloc_fail(KSVD_ERR_HIP, "counter clear launch", e);
    if ((e = launch_dict_gram_raw(D, dim, K, G, st)) != hipSuccess) return noalloc_fail(KSVD_ERR_HIP, "Gram launch", e);
    if (kmax <= 5)
        e = dim <= 64 ? noalloc_launch<16, 5>(Y, dim, P, D, K, G, kmax, idx, coef, cnt, counters, ncu, st)
                      : noalloc_launch<36, 5>(Y, dim, P, D, K, G, kmax, idx, coef, cnt, counters, ncu, st);
    else
        e = dim <= 64 ? noalloc_launch<16, 8>(Y, dim, P, D, K, G, kmax, idx, coef, cnt, counters, ncu, st)
                      : noalloc_launch<36, 8>(Y, dim, P, D, K, G, kmax, idx, coef, cnt, counters, ncu, st);
    if (e != hipSuccess) return noalloc_fail(KSVD_ERR_HIP, "pursuit launch", e);
    return KSVD_OK;
}
