// Batched OMP, BF16 precision form, on gfx950.
//
// Same pursuit as omp_mfma_kernel (CuBlasMatrixOmp::work semantics,
// cublas_matrix_omp.cu:107-402) with the correlation C = D^T R computed by
// BF16 MFMA and the atom choice won back in FP32:
//
//   * the dictionary sits in LDS as two BF16 images, D_hi = bf16(D) and
//     D_lo = bf16(D - D_hi) (2 x 77.8 KB at dim 144), shared by the 4
//     wavefronts of a workgroup (one per SIMD); each wavefront owns a tile of
//     16 patches, lane (col, g) holding quarter g (Q elements) of patch col's
//     y and residual r in FP32;
//   * C = D_hi^T r_hi + D_hi^T r_lo + D_lo^T r_hi (r_hi = bf16(r), r_lo =
//     bf16(r - r_hi)): three v_mfma_f32_16x16x32_bf16 per 32-element slice
//     (16x16x16 for the last 4 elements of a 36-element quarter) with FP32
//     accumulation, the reduction index permuted as in the FP32 kernel so that
//     the B operands come straight from r's registers;
//   * each score is then within
//         delta = |d|max |r| (3 * 2^-16 + N 2^-23),  N = 3 * 4Q products,
//     of its exact value (hi/lo splits leave 2^-16 relative per operand,
//     products are exact in FP32, at most N FP32 additions), so the FP32
//     argmax is among the atoms whose score is within 2 delta of the best.
//     Those candidates -- the best two keys of each of the patch's four
//     lanes, and, on a slow path taken when a lane holds a third, every other
//     candidate as well -- are re-scored in FP32 (d_a . r with the FP32 atom,
//     lowest index on ties) when there are two or more; a single one is the
//     argmax already;
//   * least squares stay in FP32 (incremental Cholesky of the support Gram
//     from the FP32 dictionary Gram, two triangular solves); when a new pivot
//     falls below 1e-6 of its diagonal Gram entry the support Gram and D_S^T y
//     are formed again in FP64 from the FP32 atoms and factorised and solved
//     in FP64 (x rounded to FP32), and the patch stays in FP64 from then on;
//   * the residual r = y - D_S x uses the FP32 atoms (padded FP32 copy of D
//     in global memory, read through L1/L2).
#include "ksvd_internal.hpp"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));

constexpr int BW = 4;     // wavefronts per workgroup (one per SIMD: 512 registers each)
constexpr int KPAD = 256; // atom rows in LDS (zero past K)

// LDS row of one atom: quarter g's first QM elements at g*QM (16-byte
// aligned chunks for the 16x16x32 MFMA), its last QT (0 or 4) elements at
// 4*QM + g*QT; the pitch LDB is a multiple of 8 elements with LDB/8 odd, so
// the 16 rows one MFMA reads start in distinct 16-byte bank groups.
template <int Q>
struct BfRow {
    static constexpr int QM = Q / 8 * 8;
    static constexpr int QT = Q - QM;
    static constexpr int USED = 4 * Q;
    static constexpr int LDB = ((USED + 7) / 8 % 2 == 0) ? (USED + 7) / 8 * 8 + 8 : (USED + 7) / 8 * 8;
};

__device__ __forceinline__ float patch_sum(float v)
{
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}

__device__ __forceinline__ unsigned umax(unsigned a, unsigned b) { return a > b ? a : b; }
__device__ __forceinline__ unsigned umin(unsigned a, unsigned b) { return a < b ? a : b; }
// median of three (v_med3_u32)
__device__ __forceinline__ unsigned umed3(unsigned a, unsigned b, unsigned c)
{
    unsigned d;
    asm("v_med3_u32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}

// Score key of atom a: |c| with the 8 low significand bits replaced by
// 255 - a, so an unsigned max picks the largest |c| and, among equal
// truncated scores, the lowest atom.  Inf/NaN give keys >= 0x7F800000.
__device__ __forceinline__ unsigned score_key(float c, int a)
{
    return (__float_as_uint(c) & 0x7FFFFF00u) | (unsigned)(255 - a);
}

__device__ __forceinline__ int key_atom(unsigned k) { return 255 - (int)(k & 255u); }

// One-time layout of the dictionary for the BF16 kernel: the BF16 LDS
// images D_hi (KPAD x LDB, zero past K and dim) and D_lo right after it, the
// FP32 atoms with each quarter padded to Q elements (KPAD x 4Q), and the
// largest atom norm (float bits).
template <int Q>
__global__ __launch_bounds__(256) void bf16_prep_kernel(const float *__restrict__ D, int dim, int K,
                                                        __bf16 *__restrict__ Dbf, float *__restrict__ Dq,
                                                        unsigned *__restrict__ bits)
{
    using R = BfRow<Q>;
    const int a = blockIdx.x;
    __shared__ float red[4];
    float ss = 0.f;
    for (int q = threadIdx.x; q < R::LDB; q += blockDim.x) {
        // LDS element q of row a -> quarter g, element e of that quarter
        float v = 0.f;
        if (q < R::USED) {
            int g, e;
            if (q < 4 * R::QM) {
                g = q / R::QM;
                e = q - g * R::QM;
            } else {
                g = (q - 4 * R::QM) / (R::QT > 0 ? R::QT : 1);
                e = R::QM + (q - 4 * R::QM) - g * R::QT;
            }
            const int el = g * Q + e;
            v = (a < K && el < dim) ? D[(int64_t)a * dim + el] : 0.f;
        }
        const __bf16 hi = (__bf16)v;
        Dbf[(int64_t)a * R::LDB + q] = hi;
        Dbf[(int64_t)(KPAD + a) * R::LDB + q] = (__bf16)(v - (float)hi);
    }
    for (int q = threadIdx.x; q < 4 * Q; q += blockDim.x) {
        const float v = (a < K && q < dim) ? D[(int64_t)a * dim + q] : 0.f;
        Dq[(int64_t)a * 4 * Q + q] = v;
        ss = fmaf(v, v, ss);
    }
    for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ss;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(&bits[0], __float_as_uint(sqrtf(red[0] + red[1] + red[2] + red[3])));
}

__global__ void zero_u32_kernel(unsigned *p) { p[threadIdx.x] = 0u; }

__device__ __forceinline__ double patch_sum_d(double v)
{
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}

// FP64 least squares on the support [S_0 .. S_{s-1}, j]: the support Gram and
// b = D_S^T y are formed in FP64 from the FP32 atoms (quarter rows of Dq, the
// lane's quarter of the patch yp) and summed over the patch's four lanes, the Gram
// is Cholesky-factorised and L L^T x = b solved in FP64; x is returned rounded
// to FP32.  Returns false if a pivot is not positive.  All four lanes of a
// patch call it together.  Rolled loops over local arrays keep this rarely
// taken path out of the registers.
template <int Q, int KM>
__device__ __noinline__ bool f64_solve(const float *__restrict__ Dq, int e0, const float *__restrict__ yp, int dim,
                                       const int *S, int j, int s, float *xout)
{
    int Sx[KM];
    double b[KM], Ld[KM * KM], z[KM];
    const int n = s + 1;
#pragma unroll 1
    for (int i = 0; i < n; i++) Sx[i] = i < s ? S[i] : j;
#pragma unroll 1
    for (int i = 0; i < n; i++) {
        const float *di = Dq + (int64_t)Sx[i] * 4 * Q + e0;
        double v = 0.0;
#pragma unroll 4
        for (int e = 0; e < Q; e++)
            if (e0 + e < dim) v += (double)di[e] * (double)yp[e0 + e];
        b[i] = patch_sum_d(v);
#pragma unroll 1
        for (int q = 0; q <= i; q++) {
            const float *dq = Dq + (int64_t)Sx[q] * 4 * Q + e0;
            double gv = 0.0;
#pragma unroll 4
            for (int e = 0; e < Q; e++) gv += (double)di[e] * (double)dq[e];
            gv = patch_sum_d(gv);
#pragma unroll 1
            for (int m = 0; m < q; m++) gv -= Ld[i * KM + m] * Ld[q * KM + m];
            if (q < i) {
                Ld[i * KM + q] = gv / Ld[q * KM + q];
            } else {
                if (!(gv > 0.0)) return false;
                Ld[i * KM + i] = sqrt(gv);
            }
        }
    }
#pragma unroll 1
    for (int i = 0; i < n; i++) {
        double v = b[i];
#pragma unroll 1
        for (int q = 0; q < i; q++) v -= Ld[i * KM + q] * z[q];
        z[i] = v / Ld[i * KM + i];
    }
#pragma unroll 1
    for (int i = n - 1; i >= 0; i--) {
        double v = z[i];
#pragma unroll 1
        for (int q = i + 1; q < n; q++) v -= Ld[q * KM + i] * z[q];
        z[i] = v / Ld[i * KM + i];  // z holds x from row i on
    }
#pragma unroll 1
    for (int i = 0; i < n; i++) xout[i] = (float)z[i];
    return true;
}

// Quarter g (elements e0 .. e0+Q-1, zero past dim) of patch p into v[].
template <int Q>
__device__ __forceinline__ void load_quarter(float *v, const float *__restrict__ Y, int64_t p, int dim, int e0,
                                             bool valid)
{
    if ((dim & 3) == 0) {
        // 16-byte chunks lie wholly inside or wholly past dim
        const f4 *src = (const f4 *)(Y + p * dim + e0);
#pragma unroll
        for (int c4 = 0; c4 < Q / 4; c4++) {
            const f4 t = (valid && e0 + 4 * c4 < dim) ? src[c4] : f4{0.f, 0.f, 0.f, 0.f};
            v[4 * c4] = t[0];
            v[4 * c4 + 1] = t[1];
            v[4 * c4 + 2] = t[2];
            v[4 * c4 + 3] = t[3];
        }
    } else {
#pragma unroll
        for (int e = 0; e < Q; e++) v[e] = (valid && e0 + e < dim) ? Y[p * dim + e0 + e] : 0.f;
    }
}

template <int Q, int KM>
__global__ __launch_bounds__(64 * BW, 1) void omp_bf16_kernel(
    const float *__restrict__ Y, int dim, int64_t P, const __bf16 *__restrict__ Dbf, const float *__restrict__ Dq,
    const unsigned *__restrict__ dmax_bits, int K, const float *__restrict__ G, int kmax, int ntiles,
    const int *__restrict__ plist, const unsigned *__restrict__ plist_n, int *__restrict__ idx_out,
    float *__restrict__ coef_out, int *__restrict__ cnt_out, float *__restrict__ rnorm_out,
    unsigned long long *__restrict__ counters)
{
    using R = BfRow<Q>;
    constexpr int LDB = R::LDB, QM = R::QM, QT = R::QT, NKK = QM / 8;
    constexpr int KT = KPAD / 16;
    extern __shared__ __attribute__((aligned(16))) __bf16 Ds[];  // D_hi (KPAD x LDB), then D_lo
    const __bf16 *Dlo = Ds + KPAD * LDB;
    {
        const int4 *src = (const int4 *)Dbf;
        int4 *dst = (int4 *)Ds;
        constexpr int n16 = 2 * KPAD * LDB * 2 / 16;
        for (int q = threadIdx.x; q < n16; q += blockDim.x) dst[q] = src[q];
    }
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, g = lane >> 4;
    const int e0 = g * Q;
    // this lane's A-operand rows: atom tile t adds t * 16 * LDB
    const __bf16 *bh = Ds + col * LDB + g * QM, *bl = Dlo + col * LDB + g * QM;
    const __bf16 *bth = Ds + col * LDB + 4 * QM + 4 * g, *btl = Dlo + col * LDB + 4 * QM + 4 * g;
    // score error bound per unit |r|: 3 * 2^-16 for the hi/lo splits, 2^-23
    // per FP32 addition of the 3 * 4Q products, 0.1 % slack
    const float delta_u = __uint_as_float(dmax_bits[0]) * (3.f * 1.5259e-5f + 3.f * 4.f * Q * 1.1921e-7f) * 1.001f;
    unsigned long long ndup = 0, ndeg = 0, nresc = 0, nf64 = 0, ncut = 0;

    // list mode: the tiles cover plist[0 .. *plist_n)
    const int64_t nl = plist ? (int64_t)*plist_n : 0;
    const int ntl = plist ? (int)((nl + 15) / 16) : ntiles;
    for (int tile = blockIdx.x * BW + wave; tile < ntl; tile += gridDim.x * BW) {
        const int64_t slot = (int64_t)tile * 16 + col;
        const int64_t p = plist ? (slot < nl ? (int64_t)plist[slot] : P) : slot;
        const bool valid = p < P;
        float y[Q], r[Q];
        load_quarter<Q>(y, Y, p, dim, e0, valid);
        float ny = 0.f;
#pragma unroll
        for (int e = 0; e < Q; e++) {
            r[e] = y[e];
            ny = fmaf(y[e], y[e], ny);
        }
        float rn = sqrtf(patch_sum(ny));
        int S[KM];
        float x[KM], bv[KM], L[KM * KM];
#pragma unroll
        for (int i = 0; i < KM; i++) { S[i] = -1; x[i] = 0.f; bv[i] = 0.f; }
#pragma unroll
        for (int i = 0; i < KM * KM; i++) L[i] = 0.f;
        int s = 0;
        bool active = valid;
        bool f64mode = false;  // the support Gram went below the pivot bound: FP64 from here on

        for (int t = 0; t < kmax; t++) {
            if (__ballot(active) == 0) break;
            // ---- correlation: three BF16 MFMAs per slice ---------------------
            // (the first slice starts from a zero accumulator constant)
            f4 acc[KT];
#pragma unroll
            for (int kk = 0; kk < NKK; kk++) {
                bf8 rh, rl;
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const float v = r[8 * kk + i];
                    rh[i] = (__bf16)v;
                    rl[i] = (__bf16)(v - (float)rh[i]);
                }
#pragma unroll
                for (int t0 = 0; t0 < KT; t0 += 4) {
                    bf8 ah[4], al[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        ah[u] = *(const bf8 *)(bh + (t0 + u) * 16 * LDB + 8 * kk);
                        al[u] = *(const bf8 *)(bl + (t0 + u) * 16 * LDB + 8 * kk);
                    }
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const f4 c0 = kk == 0 ? f4{0.f, 0.f, 0.f, 0.f} : acc[t0 + u];
                        acc[t0 + u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[u], rh, c0, 0, 0, 0);
                        acc[t0 + u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[u], rl, acc[t0 + u], 0, 0, 0);
                        acc[t0 + u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[u], rh, acc[t0 + u], 0, 0, 0);
                    }
                }
            }
            if constexpr (QT == 4) {
                bf4 rh, rl;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const float v = r[QM + i];
                    rh[i] = (__bf16)v;
                    rl[i] = (__bf16)(v - (float)rh[i]);
                }
#pragma unroll
                for (int t0 = 0; t0 < KT; t0 += 4) {
                    bf4 ah[4], al[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        ah[u] = *(const bf4 *)(bth + (t0 + u) * 16 * LDB);
                        al[u] = *(const bf4 *)(btl + (t0 + u) * 16 * LDB);
                    }
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        acc[t0 + u] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[u], rh, acc[t0 + u], 0, 0, 0);
                        acc[t0 + u] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[u], rl, acc[t0 + u], 0, 0, 0);
                        acc[t0 + u] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al[u], rh, acc[t0 + u], 0, 0, 0);
                    }
                }
            }
            // ---- candidates: scores within 2 delta of the best ----------------
            // lane (col, g) holds atoms tt*16 + 4g + i; keys replace the scores
            // per-lane best two keys: m2 = med3(m1, m2, k) keeps the second
            // largest (m2 <= m1 throughout), in four independent chains
            unsigned c1[4], c2[4];
#pragma unroll
            for (int c = 0; c < 4; c++) c1[c] = c2[c] = 0u;
#pragma unroll
            for (int tt = 0; tt < KT; tt++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const unsigned k = score_key(acc[tt][i], tt * 16 + 4 * g + i);
                    acc[tt][i] = __uint_as_float(k);
                    c2[i] = umed3(c1[i], c2[i], k);
                    c1[i] = umax(c1[i], k);
                }
            unsigned m1 = c1[0], m2 = c2[0];
#pragma unroll
            for (int c = 1; c < 4; c++) {
                m2 = umax(umin(m1, c1[c]), umax(m2, c2[c]));
                m1 = umax(m1, c1[c]);
            }
            unsigned M = umax(m1, (unsigned)__shfl_xor((int)m1, 16));
            M = umax(M, (unsigned)__shfl_xor((int)M, 32));
            const float Ms = __uint_as_float(M & 0x7FFFFF00u);
            // keys drop 8 significand bits: 2^-15 of the best score more
            const float tau = Ms - 2.f * delta_u * rn - Ms * 6.11e-5f;
            const unsigned tk = tau > 0.f ? (__float_as_uint(tau) & 0x7FFFFF00u) : 0u;
            const bool finite = M < 0x7F800000u && rn == rn;
            // candidates: the two best keys of each of the patch's four lanes
            // that reach tk (at most 8), in lane-group order
            unsigned ck[8];
            int nc = 0;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const unsigned kq = (unsigned)__shfl((int)(q < 4 ? m1 : m2), col + 16 * (q & 3));
                ck[q] = 0u;
                if (kq >= tk) {
#pragma unroll
                    for (int z = 0; z < 8; z++)
                        if (z == nc) ck[z] = kq;
                    nc++;
                }
            }
            // a lane with two candidates may hold a third: such a patch takes
            // the slow path below, which re-scores every candidate beyond the
            // lanes' best two
            bool cut = false;
            if (__ballot(active && m2 >= tk) != 0ull) {
                unsigned m3 = 0u;
#pragma unroll
                for (int tt = 0; tt < KT; tt++)
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const unsigned k = __float_as_uint(acc[tt][i]);
                        m3 = umax(m3, k < m2 ? k : 0u);
                    }
                m3 = umax(m3, (unsigned)__shfl_xor((int)m3, 16));
                m3 = umax(m3, (unsigned)__shfl_xor((int)m3, 32));
                cut = active && finite && m3 >= tk && m3 > 0u;
                if (cut && g == 0) ncut++;
            }
            if (!finite) nc = 0;
            if (active && g == 0 && nc >= 2) nresc++;
            // ---- FP32 re-score of the candidates; b_j = d_j . y --------------
            float best32 = -1.f, bp = 0.f;
            int j = 0x7fffffff;
#pragma unroll 1
            for (int e = 0; e < 8; e++) {
                const bool mine = active && e < nc;
                if (__ballot(mine) == 0ull) break;
                unsigned ke = ck[0];
#pragma unroll
                for (int z = 1; z < 8; z++)
                    if (z == e) ke = ck[z];
                const int a = mine ? key_atom(ke) : 0;
                const f4 *row = (const f4 *)(Dq + (int64_t)a * 4 * Q + e0);
                float cr = 0.f;
#pragma unroll
                for (int c4 = 0; c4 < Q / 4; c4++) {
                    const f4 dv = row[c4];
#pragma unroll
                    for (int i = 0; i < 4; i++) cr = fmaf(dv[i], r[4 * c4 + i], cr);
                }
                cr = patch_sum(cr);
                const float v = fabsf(cr);
                if (mine && (v > best32 || (v == best32 && a < j))) {
                    best32 = v;
                    j = a;
                }
            }
            // slow path: every further key >= tk, lane by lane in descending
            // order (round q takes each lane's next key below the last one it
            // gave), re-scored in FP32 against the best so far
            if (__ballot(cut) != 0ull) {
                unsigned prev = m2;
#pragma unroll 1
                for (;;) {
                    unsigned mn = 0u;
#pragma unroll
                    for (int tt = 0; tt < KT; tt++)
#pragma unroll
                        for (int i = 0; i < 4; i++) {
                            const unsigned k = __float_as_uint(acc[tt][i]);
                            mn = umax(mn, k < prev ? k : 0u);
                        }
                    const bool more = cut && mn >= tk && mn > 0u;
                    if (__ballot(more) == 0ull) break;
#pragma unroll 1
                    for (int q = 0; q < 4; q++) {
                        const unsigned kq = (unsigned)__shfl((int)mn, col + 16 * q);
                        const bool mine = cut && kq >= tk && kq > 0u;
                        if (__ballot(mine) == 0ull) continue;
                        const int a = mine ? key_atom(kq) : 0;
                        const f4 *row = (const f4 *)(Dq + (int64_t)a * 4 * Q + e0);
                        float cr = 0.f;
#pragma unroll
                        for (int c4 = 0; c4 < Q / 4; c4++) {
                            const f4 dv = row[c4];
#pragma unroll
                            for (int i = 0; i < 4; i++) cr = fmaf(dv[i], r[4 * c4 + i], cr);
                        }
                        cr = patch_sum(cr);
                        const float v = fabsf(cr);
                        if (mine && (v > best32 || (v == best32 && a < j))) {
                            best32 = v;
                            j = a;
                        }
                    }
                    prev = more ? mn : 0u;
                }
            }
            // b_j = d_j . y
            if (__ballot(active && j < K) != 0ull) {
                const f4 *row = (const f4 *)(Dq + (int64_t)(j < K ? j : 0) * 4 * Q + e0);
#pragma unroll
                for (int c4 = 0; c4 < Q / 4; c4++) {
                    const f4 dv = row[c4];
#pragma unroll
                    for (int i = 0; i < 4; i++) bp = fmaf(dv[i], y[4 * c4 + i], bp);
                }
                bp = patch_sum(bp);
            }
            // ---- support bookkeeping and Cholesky row ------------------------
            bool dup = false;
#pragma unroll
            for (int i = 0; i < KM; i++) dup |= (i < s) && (S[i] == j);
            if (active && dup) {
                active = false;
                if (g == 0) ndup++;
            }
            // no finite candidate (non-finite residual): retire
            if (active && j >= K) {
                active = false;
                if (g == 0) ndeg++;
            }
            bool need64 = active && f64mode;
            if (active && !f64mode) {
                float w[KM];
                const float gjj = G[(int64_t)j * K + j];
                float piv = gjj;
#pragma unroll
                for (int i = 0; i < KM; i++) {
                    w[i] = 0.f;
                    if (i < s) {
                        float gi = G[(int64_t)S[i] * K + j];
#pragma unroll
                        for (int q = 0; q < i; q++) gi = fmaf(-L[i * KM + q], w[q], gi);
                        w[i] = gi / L[i * KM + i];
                        piv = fmaf(-w[i], w[i], piv);
                    }
                }
                if (piv > 1e-6f * gjj) {
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
                    for (int i = KM - 1; i >= 0; i--) {
                        if (i < s) {
                            float v = z[i];
#pragma unroll
                            for (int q = i + 1; q < KM; q++)
                                if (q < s) v = fmaf(-L[q * KM + i], x[q], v);
                            x[i] = v / L[i * KM + i];
                        }
                    }
                } else {
                    need64 = true;
                    f64mode = true;
                    if (g == 0) nf64++;
                }
            }
            if (need64) {
                // small pivot (now or at an earlier step): least squares in FP64
                float xs[KM];
                if (!f64_solve<Q, KM>(Dq, e0, Y + p * dim, dim, S, j, s, xs)) {
                    active = false;
                    if (g == 0) ndeg++;
                } else {
#pragma unroll
                    for (int i = 0; i < KM; i++)
                        if (i == s) {
                            S[i] = j;
                            bv[i] = bp;
                        }
                    s++;
#pragma unroll
                    for (int i = 0; i < KM; i++)
                        if (i < s) x[i] = xs[i];
                }
            }
            // ---- residual r = y - D_S x with the FP32 atoms, and its norm ----
            // (D_S x accumulates in r's registers, then r = y - D_S x)
            float nr = 0.f;
            if (active) {
#pragma unroll
                for (int e = 0; e < Q; e++) r[e] = 0.f;
#pragma unroll
                for (int i = 0; i < KM; i++)
                    if (i < s) {
                        const f4 *row = (const f4 *)(Dq + (int64_t)S[i] * 4 * Q + e0);
#pragma unroll
                        for (int c4 = 0; c4 < Q / 4; c4++) {
                            const f4 dv = row[c4];
#pragma unroll
                            for (int c = 0; c < 4; c++) r[4 * c4 + c] = fmaf(dv[c], x[i], r[4 * c4 + c]);
                        }
                    }
#pragma unroll
                for (int e = 0; e < Q; e++) {
                    const float rv = y[e] - r[e];
                    r[e] = rv;
                    nr = fmaf(rv, rv, nr);
                }
            }
            nr = patch_sum(nr);
            rn = sqrtf(nr);
            if (active && rn < 1e-3f) active = false;
        }
        if (valid && g == 0) {
            if (rnorm_out) rnorm_out[p] = rn;
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
    if (nresc) atomicAdd(&counters[CNT_RESCORE], nresc);
    if (nf64) atomicAdd(&counters[CNT_F64], nf64);
    if (ncut) atomicAdd(&counters[CNT_CAND_CUT], ncut);
}

template <int Q>
int launch_bf16_q(ksvd_ctx *ctx, const int *plist, const unsigned *plist_n)
{
    using R = BfRow<Q>;
    unsigned *dmax = (unsigned *)ctx->bfscal.ptr;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, ctx->stream, dmax);
    hipLaunchKernelGGL(bf16_prep_kernel<Q>, dim3(KPAD), dim3(256), 0, ctx->stream, (const float *)ctx->D.ptr,
                       ctx->dim, ctx->K, (__bf16 *)ctx->Dbf.ptr, (float *)ctx->Dq.ptr, dmax);
    const size_t lds = sizeof(__bf16) * 2 * (size_t)KPAD * R::LDB;
    const int ntiles = (int)((ctx->P + 15) / 16);
    int grid = (ntiles + BW - 1) / BW;
    if (grid > ctx->num_cus) grid = ctx->num_cus;
    auto run = [&](auto kern) -> int {
        KSVD_HIP_CHECK(ctx, allow_max_lds((const void *)kern));
        hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * BW), lds, ctx->stream, (const float *)ctx->Y.ptr, ctx->dim,
                           ctx->P, (const __bf16 *)ctx->Dbf.ptr, (const float *)ctx->Dq.ptr,
                           (const unsigned *)dmax, ctx->K, (const float *)ctx->G.ptr, ctx->kmax, ntiles, plist,
                           plist_n, (int *)ctx->idx.ptr, (float *)ctx->coef.ptr, (int *)ctx->cnt.ptr,
                           (float *)ctx->rbound.ptr, (unsigned long long *)ctx->counters.ptr);
        KSVD_HIP_CHECK(ctx, hipGetLastError());
        return KSVD_OK;
    };
    if (ctx->kmax <= 5) return run(omp_bf16_kernel<Q, 5>);
    return run(omp_bf16_kernel<Q, 8>);
}

}  // namespace

size_t bf16_dict_bytes(int dim)
{
    const int Q = dim <= 64 ? 16 : 36;
    const int ldb = Q == 16 ? BfRow<16>::LDB : BfRow<36>::LDB;
    return sizeof(__bf16) * 2 * (size_t)KPAD * ldb;
}

size_t bf16_dictq_bytes(int dim)
{
    const int Q = dim <= 64 ? 16 : 36;
    return sizeof(float) * (size_t)KPAD * 4 * Q;
}

int launch_omp_bf16(ksvd_ctx *ctx)
{
    if (ctx->dim > 144 || ctx->K > KPAD || ctx->kmax > 8)
        return ksvd_fail(ctx, KSVD_ERR_UNSUPPORTED, "the BF16 form needs patch dim <= 144, atoms <= 256, ompIter <= 8");
    return ctx->dim <= 64 ? launch_bf16_q<16>(ctx, nullptr, nullptr) : launch_bf16_q<36>(ctx, nullptr, nullptr);
}

int launch_omp_bf16_list(ksvd_ctx *ctx, const int *plist, const unsigned *plist_n)
{
    if (ctx->dim > 144 || ctx->K > KPAD || ctx->kmax > 8)
        return ksvd_fail(ctx, KSVD_ERR_UNSUPPORTED, "the BF16 form needs patch dim <= 144, atoms <= 256, ompIter <= 8");
    return ctx->dim <= 64 ? launch_bf16_q<16>(ctx, plist, plist_n) : launch_bf16_q<36>(ctx, plist, plist_n);
}
