/*
 * K-SVD denoising oracle: CPU restatement of the reference algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, not the product: only
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load
 * it.  The product path (libksvd_hip.so) never links or calls it.
 *
 * Included twice by ksvd_oracle.c: once with REAL=float (suffix _f32, the
 * reference's FP32 semantics) and once with REAL=double (suffix _f64, the FP64
 * run of the same algorithm used to state every result's error).
 *
 * Reference files restated (paths relative to the reference repository):
 *   patch gather      src/lib/denoisingLib/cuda_svd_denoiser.cu:221-256
 *   dictionary init   src/lib/denoisingLib/cuda_svd_denoiser.cu:52-63,261-280
 *   batched OMP       src/lib/matUtilityLib/cublas_matrix_omp.cu:107-402
 *   least squares     src/lib/svdLib/cusolver_gesvda_batch_engine.cu:117-143
 *                     + cublas_matrix_omp.cu:270-338 (x = pinv(A_S) y)
 *   dictionary update src/lib/denoisingLib/cuda_svd_denoiser.cu:11-50,285-362
 *   rank-1 SVD        src/lib/svdLib/cusolver_gesvdj_engine.cu:46-69
 *   K-SVD driver      src/lib/denoisingLib/cuda_svd_denoiser.cu:368-408
 *   rebuild + blend   src/lib/denoisingLib/cuda_svd_denoiser.cu:413-498
 *
 * Pinning (which independent check fixes each function) is listed in the
 * header comment of every function below; tests/test_oracle.py runs them.
 */

#define ORC_CAT_(a, b) a##b
#define ORC_CAT(a, b) ORC_CAT_(a, b)
#define FN(name) ORC_CAT(orc_##name, SUFFIX)

/* ------------------------------------------------------------------------ */
/* Patch gather.  cuda_svd_denoiser.cu:231-240: the image buffer is walked as
 * n rows of length m (row stride m); the patch at (row i, col j) is the ph x pw
 * block, stored row after row (element e = r*pw + c).  Patches are numbered
 * with the row origin outer and the column origin inner.
 * Pinned by: closed-form patch count and a brute-force index-set check
 * (tests/test_oracle.py::test_extract_*). */
void FN(extract)(const float *buf, int m, int n, int pw, int ph, int sw, int sh, REAL *Y)
{
    long p = 0;
    const int dim = pw * ph;
    for (int i = 0; i + ph <= n; i += sh)
        for (int j = 0; j + pw <= m; j += sw) {
            REAL *y = Y + p * (long)dim;
            int e = 0;
            for (int r = 0; r < ph; r++)
                for (int c = 0; c < pw; c++)
                    y[e++] = (REAL)buf[(long)(i + r) * m + j + c];
            p++;
        }
}

/* ------------------------------------------------------------------------ */
/* Dictionary init.  cuda_svd_denoiser.cu:266-268 copies the K patches that
 * start at column floor(P/2); normalizeDictKernel (:52-63) divides each by
 * sqrtf of thrust::transform_reduce(square, init=0, plus<float>).  The init
 * value 0 is an int, so the reduction's accumulator is an int: every partial
 * sum is computed in float and truncated toward zero before the next term.
 * Pinned by: integer-valued patches (truncation exact -> unit norm against
 * numpy), a hand-computed fractional case, and the column pick. */
void FN(init_dict)(const REAL *Y, int dim, long P, int K, REAL *D)
{
    const long off = P / 2;
    for (int k = 0; k < K; k++) {
        const REAL *y = Y + (off + k) * (long)dim;
        REAL *d = D + (long)k * dim;
        int acc = 0;
        for (int e = 0; e < dim; e++) {
            REAL sq = y[e] * y[e];
            acc = (int)((REAL)acc + sq);
        }
        REAL norm = (REAL)sqrt((double)(REAL)acc);
        if (sizeof(REAL) == sizeof(float)) norm = (REAL)sqrtf((float)acc);
        for (int e = 0; e < dim; e++) d[e] = y[e] / norm;
    }
}

/* ------------------------------------------------------------------------ */
/* Batched OMP, cublas_matrix_omp.cu:107-402, per patch:
 *   norm = 1 on entry (:116), X = 0 (:79), r = y (:118);
 *   for t < kmax, while norm >= 1e-3 (:218):
 *     c = D^T r (:193-206);  j = first argmax |c| (cublasIsamax, :225-233);
 *     S += {j};  x = pinv(D_S) y on the ORIGINAL patch (:327-338);
 *     X[S_i] = x_i (:342-350);  approx = D_S x (:357-368);
 *     r = y - approx (:376-380);  norm = ||r||_2 (:386-390).
 * pinv(D_S) y is the least-squares solution; it is computed here by the
 * Cholesky factor of the support Gram D_S^T D_S, grown one row per step.
 * ADR (documented): a pick that repeats an atom already in S (only possible
 * when r is numerically orthogonal to every atom) or a non-positive Cholesky
 * pivot ends the patch's pursuit with its current code; stats counts both.
 * Output is compact: idx[p*kmax+i], coef[p*kmax+i] for i < cnt[p].
 * Dt is the dictionary transposed (dim x K row-major) so that the K dot
 * products of one correlation vectorise across atoms; each dot product is
 * still accumulated in element order.
 * Pinned by: brute-force OMP built on numpy lstsq (tests/test_oracle.py),
 * the orthogonality invariant r ⟂ span(D_S), first-index tie-break cases. */
void FN(omp)(const REAL *Y, int dim, long P, const REAL *D, int K, int kmax,
             int *idx, REAL *coef, int *cnt, long *stats)
{
    REAL *Dt = (REAL *)malloc(sizeof(REAL) * (size_t)dim * K);
    for (int k = 0; k < K; k++)
        for (int e = 0; e < dim; e++) Dt[(long)e * K + k] = D[(long)k * dim + e];
    long dup = 0, degen = 0, early = 0;

#pragma omp parallel reduction(+ : dup, degen, early)
    {
        REAL *r = (REAL *)malloc(sizeof(REAL) * dim);
        REAL *c = (REAL *)malloc(sizeof(REAL) * K);
        REAL L[ORC_KMAX * ORC_KMAX], b[ORC_KMAX], x[ORC_KMAX], w[ORC_KMAX];
        int S[ORC_KMAX];
#pragma omp for schedule(dynamic, 64)
        for (long p = 0; p < P; p++) {
            const REAL *y = Y + p * (long)dim;
            int s = 0;
            for (int e = 0; e < dim; e++) r[e] = y[e];
            for (int t = 0; t < kmax; t++) {
                /* correlation c = D^T r, one sequential dot product per atom */
                for (int k = 0; k < K; k++) c[k] = 0;
                for (int e = 0; e < dim; e++) {
                    const REAL re = r[e];
                    const REAL *row = Dt + (long)e * K;
                    for (int k = 0; k < K; k++) c[k] = c[k] + row[k] * re;
                }
                /* first index of the largest |c| (Isamax semantics) */
                int j = 0;
                REAL best = (REAL)fabs((double)c[0]);
                for (int k = 1; k < K; k++) {
                    REAL a = (REAL)fabs((double)c[k]);
                    if (a > best) { best = a; j = k; }
                }
                int is_dup = 0;
                for (int i = 0; i < s; i++)
                    if (S[i] == j) is_dup = 1;
                if (is_dup) { dup++; break; }
                /* grow the Cholesky factor of the support Gram by one row */
                const REAL *dj = D + (long)j * dim;
                for (int i = 0; i < s; i++) {
                    const REAL *di = D + (long)S[i] * dim;
                    REAL g = 0;
                    for (int e = 0; e < dim; e++) g = g + di[e] * dj[e];
                    for (int q = 0; q < i; q++) g = g - L[i * ORC_KMAX + q] * w[q];
                    w[i] = g / L[i * ORC_KMAX + i];
                }
                REAL gjj = 0;
                for (int e = 0; e < dim; e++) gjj = gjj + dj[e] * dj[e];
                REAL piv = gjj;
                for (int q = 0; q < s; q++) piv = piv - w[q] * w[q];
                if (!(piv > 0)) { degen++; break; }
                for (int q = 0; q < s; q++) L[s * ORC_KMAX + q] = w[q];
                L[s * ORC_KMAX + s] = (REAL)sqrt((double)piv);
                if (sizeof(REAL) == sizeof(float)) L[s * ORC_KMAX + s] = (REAL)sqrtf((float)piv);
                REAL bj = 0;
                for (int e = 0; e < dim; e++) bj = bj + dj[e] * y[e];
                b[s] = bj;
                S[s] = j;
                s++;
                /* solve L L^T x = b */
                for (int i = 0; i < s; i++) {
                    REAL v = b[i];
                    for (int q = 0; q < i; q++) v = v - L[i * ORC_KMAX + q] * w[q];
                    w[i] = v / L[i * ORC_KMAX + i];
                }
                for (int i = s - 1; i >= 0; i--) {
                    REAL v = w[i];
                    for (int q = i + 1; q < s; q++) v = v - L[q * ORC_KMAX + i] * x[q];
                    x[i] = v / L[i * ORC_KMAX + i];
                }
                /* residual against the original patch and its norm */
                REAL nrm2 = 0;
                for (int e = 0; e < dim; e++) {
                    REAL a = 0;
                    for (int i = 0; i < s; i++) a = a + D[(long)S[i] * dim + e] * x[i];
                    r[e] = y[e] - a;
                    nrm2 = nrm2 + r[e] * r[e];
                }
                REAL nrm = (REAL)sqrt((double)nrm2);
                if (sizeof(REAL) == sizeof(float)) nrm = (REAL)sqrtf((float)nrm2);
                if (nrm < (REAL)0.001) { if (t + 1 < kmax) early++; break; }
            }
            cnt[p] = s;
            for (int i = 0; i < kmax; i++) {
                idx[p * kmax + i] = i < s ? S[i] : -1;
                coef[p * kmax + i] = i < s ? x[i] : 0;
            }
        }
        free(r);
        free(c);
    }
    free(Dt);
    if (stats) { stats[0] = dup; stats[1] = degen; stats[2] = early; }
}

/* ------------------------------------------------------------------------ */
/* Users of every atom: omega_k = ascending {p : X[k,p] != 0}, computed once
 * from the codes at entry (findRelevantIndicesKernel, cuda_svd_denoiser.cu:
 * 36-50, launched once per sweep at :298).  Output CSR: off[K+1], users[]. */
long FN(users)(const int *idx, const REAL *coef, const int *cnt, long P, int K, int kmax,
               long *off, int *users)
{
    for (int k = 0; k <= K; k++) off[k] = 0;
    for (long p = 0; p < P; p++)
        for (int i = 0; i < cnt[p]; i++)
            if (coef[p * kmax + i] != 0) off[idx[p * kmax + i] + 1]++;
    for (int k = 0; k < K; k++) off[k + 1] += off[k];
    long *pos = (long *)malloc(sizeof(long) * K);
    for (int k = 0; k < K; k++) pos[k] = off[k];
    for (long p = 0; p < P; p++)
        for (int i = 0; i < cnt[p]; i++)
            if (coef[p * kmax + i] != 0) users[pos[idx[p * kmax + i]]++] = (int)p;
    free(pos);
    return off[K];
}

/* Top eigenvector of the symmetric n x n matrix G (double), by power
 * iteration started from u (overwritten).  Returns the eigenvalue estimate;
 * *iters receives the iteration count. */
static double ORC_CAT(orc_top_eig, SUFFIX)(const double *G, int n, double *u, double *z, int *iters)
{
    double nu = 0;
    for (int e = 0; e < n; e++) nu += u[e] * u[e];
    nu = sqrt(nu);
    if (nu == 0) { u[0] = 1; nu = 1; }
    for (int e = 0; e < n; e++) u[e] /= nu;
    double lam = 0;
    int it;
    for (it = 0; it < ORC_EIG_MAXIT; it++) {
        for (int a = 0; a < n; a++) {
            double s = 0;
            const double *g = G + (long)a * n;
            for (int q = 0; q < n; q++) s += g[q] * u[q];
            z[a] = s;
        }
        double nz = 0;
        for (int a = 0; a < n; a++) nz += z[a] * z[a];
        nz = sqrt(nz);
        if (nz == 0) { lam = 0; break; }
        double diff = 0;
        for (int a = 0; a < n; a++) {
            double v = z[a] / nz;
            diff += (v - u[a]) * (v - u[a]);
            u[a] = v;
        }
        lam = nz;
        if (diff < 1e-26) break;
    }
    if (iters) *iters = it;
    return lam;
}

/* ------------------------------------------------------------------------ */
/* Dictionary update, cuda_svd_denoiser.cu:285-362 (Gauss-Seidel over atoms):
 *   for k with |omega_k| >= 1 (:311), T = first min(|omega_k|, cap) users
 *   (the <<<1,1024>>> launches at :315,:324 only fill 1024 columns of E; the
 *   other columns stay 0, :306);
 *   E[:,t] = y_{T_t} - D (x_{T_t} with entry k zeroed)  (:315-325);
 *   (u, s, v) = top singular triple of E (gesvdj, :328-332);
 *   d_k = u (:343-346);  X[k, omega_k[t]] = s v_t (:349-358), which is the
 *   column projection e_t^T u, and 0 for t >= cap.
 * D x is summed over the support in ascending atom order (dense D·x with the
 * zero entries dropped).  The top singular vector is the top eigenvector of
 * E E^T formed in double (power iteration from the previous atom, so the sign
 * follows d_k's previous direction).  cap <= 0 means uncapped.
 * gap (optional, K entries) receives lambda2/lambda1 of E E^T per atom (-1 if
 * skipped).
 * Pinned by: numpy/LAPACK SVD of the same E, sign-fixed (tests), hand-built
 * cap cases, and a Gauss-Seidel ordering case. */
void FN(update)(const REAL *Y, int dim, long P, REAL *D, int K, int kmax,
                const int *idx, REAL *coef, const int *cnt, int cap, double *gap, long *eig_iters)
{
    long *off = (long *)malloc(sizeof(long) * (K + 1));
    long nnz = 0;
    for (long p = 0; p < P; p++) nnz += cnt[p];
    int *users = (int *)malloc(sizeof(int) * (nnz + 1));
    FN(users)(idx, coef, cnt, P, K, kmax, off, users);
    double *G = (double *)malloc(sizeof(double) * dim * dim);
    double *u = (double *)malloc(sizeof(double) * dim);
    double *z = (double *)malloc(sizeof(double) * dim);
    long tot_it = 0;

    for (int k = 0; k < K; k++) {
        const long c = off[k + 1] - off[k];
        if (gap) gap[k] = -1;
        if (c < 1) continue;
        const long mcols = (cap > 0 && c > cap) ? cap : c;
        const int *T = users + off[k];
        REAL *E = (REAL *)malloc(sizeof(REAL) * (size_t)dim * mcols);
#pragma omp parallel for schedule(static)
        for (long t = 0; t < mcols; t++) {
            const long p = T[t];
            int ai[ORC_KMAX];
            REAL ac[ORC_KMAX];
            int na = 0;
            for (int i = 0; i < cnt[p]; i++) {
                int a = idx[p * kmax + i];
                if (a == k) continue;
                /* insertion by ascending atom index */
                int q = na++;
                while (q > 0 && ai[q - 1] > a) { ai[q] = ai[q - 1]; ac[q] = ac[q - 1]; q--; }
                ai[q] = a;
                ac[q] = coef[p * kmax + i];
            }
            const REAL *y = Y + p * (long)dim;
            REAL *et = E + t * (long)dim;
            for (int e = 0; e < dim; e++) {
                REAL a = 0;
                for (int q = 0; q < na; q++) a = a + D[(long)ai[q] * dim + e] * ac[q];
                et[e] = y[e] - a;
            }
        }
#pragma omp parallel for schedule(static)
        for (int a = 0; a < dim; a++)
            for (int b2 = 0; b2 <= a; b2++) {
                double s = 0;
                for (long t = 0; t < mcols; t++)
                    s += (double)E[t * dim + a] * (double)E[t * dim + b2];
                G[(long)a * dim + b2] = s;
                G[(long)b2 * dim + a] = s;
            }
        for (int e = 0; e < dim; e++) u[e] = (double)D[(long)k * dim + e];
        int it = 0;
        double lam1 = ORC_CAT(orc_top_eig, SUFFIX)(G, dim, u, z, &it);
        tot_it += it;
        if (gap) {
            /* second eigenvalue by deflation, for reporting only */
            double *G2 = (double *)malloc(sizeof(double) * dim * dim);
            double *u2 = (double *)malloc(sizeof(double) * dim);
            for (int a = 0; a < dim; a++)
                for (int b2 = 0; b2 < dim; b2++)
                    G2[(long)a * dim + b2] = G[(long)a * dim + b2] - lam1 * u[a] * u[b2];
            for (int e = 0; e < dim; e++) u2[e] = (double)((e * 7 + 3) % 11) - 5.0;
            int it2 = 0;
            double lam2 = ORC_CAT(orc_top_eig, SUFFIX)(G2, dim, u2, z, &it2);
            gap[k] = lam1 > 0 ? lam2 / lam1 : -1;
            free(G2);
            free(u2);
        }
        if (lam1 > 0) {
            for (int e = 0; e < dim; e++) D[(long)k * dim + e] = (REAL)u[e];
        }
        /* new coefficients: s v_t = e_t^T u for t < cap, 0 beyond */
#pragma omp parallel for schedule(static)
        for (long t = 0; t < c; t++) {
            const long p = T[t];
            double v = 0;
            if (t < mcols && lam1 > 0) {
                const REAL *et = E + t * (long)dim;
                for (int e = 0; e < dim; e++) v += (double)et[e] * u[e];
            }
            for (int i = 0; i < cnt[p]; i++)
                if (idx[p * kmax + i] == k) coef[p * kmax + i] = (REAL)v;
        }
        free(E);
    }
    if (eig_iters) *eig_iters = tot_it;
    free(off);
    free(users);
    free(G);
    free(u);
    free(z);
}

/* ------------------------------------------------------------------------ */
/* Rebuild, cuda_svd_denoiser.cu:421-481:
 *   patch values Yhat = D X (:421-423), summed over the support in ascending
 *   atom order with fused multiply-adds;
 *   overlap-add sum and count (:431-456) in patch order; the count is the
 *   number of covering patches;
 *   first = sum / w;  d = sum (first - buf)^2 over covered pixels (float
 *   difference and square, accumulated in double; the reference sums in float
 *   sequentially) ; d /= sigma2 * N (float);  lambda = |sqrtf(d) - 1|;
 *   out = (lambda * buf + sum) / (lambda + w) (float, no contraction).
 * ADR: a pixel with w = 0 adds nothing to d and its output is buf itself.
 * Geometry: the patch frame of the gather (rows of length m).  compat != 0
 * walks the reference's rebuild frame (:432-453), which swaps the patch
 * extents and strides; identical when pw == ph and sw == sh.
 * Outputs: out, sum, weight (each m*n), and dl[0] = d sum, dl[1] = lambda.
 * Pinned by: brute-force numpy overlap-add on small cases; closed-form lambda. */
void FN(rebuild)(const int *idx, const REAL *coef, const int *cnt, const REAL *D, int dim,
                 long P, int kmax, const float *buf, int m, int n, int pw, int ph, int sw,
                 int sh, int compat, float sigma2, REAL *out, REAL *sum, REAL *weight,
                 double *dl)
{
    const long N = (long)m * n;
    for (long q = 0; q < N; q++) { sum[q] = 0; weight[q] = 0; }
    /* frame of the walk */
    int rext = ph, cext = pw, rstep = sh, cstep = sw;
    if (compat) { rext = pw; cext = ph; rstep = sw; cstep = sh; }
    REAL *yh = (REAL *)malloc(sizeof(REAL) * dim);
    long p = 0;
    for (int i = 0; i + rext <= n; i += rstep)
        for (int j = 0; j + cext <= m; j += cstep) {
            if (p >= P) break;
            int ai[ORC_KMAX];
            REAL ac[ORC_KMAX];
            int na = 0;
            for (int q = 0; q < cnt[p]; q++) {
                int a = idx[p * kmax + q];
                int s = na++;
                while (s > 0 && ai[s - 1] > a) { ai[s] = ai[s - 1]; ac[s] = ac[s - 1]; s--; }
                ai[s] = a;
                ac[s] = coef[p * kmax + q];
            }
            for (int e = 0; e < dim; e++) {
                REAL acc = 0;
                for (int q = 0; q < na; q++) acc = ORC_FMA(D[(long)ai[q] * dim + e], ac[q], acc);
                yh[e] = acc;
            }
            int e = 0;
            for (int r = 0; r < rext; r++)
                for (int c = 0; c < cext; c++) {
                    long q = (long)(i + r) * m + j + c;
                    sum[q] = sum[q] + yh[e++];
                    weight[q] = weight[q] + 1;
                }
            p++;
        }
    free(yh);
    double dsum = 0;
    for (long q = 0; q < N; q++) {
        if (weight[q] == 0) continue;
        REAL first = sum[q] / weight[q];
        REAL diff = first - (REAL)buf[q];
        REAL sq = diff * diff;
        dsum += (double)sq;
    }
    REAL d = (REAL)dsum;
    d = d / ((REAL)sigma2 * (REAL)N);
    REAL lam = (REAL)fabs(sqrt((double)d) - 1.0);
    if (sizeof(REAL) == sizeof(float)) lam = (REAL)fabsf(sqrtf((float)d) - 1.f);
    for (long q = 0; q < N; q++) {
        if (weight[q] == 0) { out[q] = (REAL)buf[q]; continue; }
        REAL num = (REAL)buf[q] * lam;
        num = num + sum[q];
        out[q] = num / (lam + weight[q]);
    }
    dl[0] = dsum;
    dl[1] = (double)lam;
}

/* ------------------------------------------------------------------------ */
/* The K-SVD loop, cuda_svd_denoiser.cu:368-408: iters x (OMP, update), then a
 * final OMP.  Used as the CPU baseline; t_ms receives [omp ms, update ms]. */
void FN(ksvd)(const REAL *Y, int dim, long P, REAL *D, int K, int kmax, int iters, int cap,
              int final_omp, int *idx, REAL *coef, int *cnt, double *t_ms)
{
    double t_omp = 0, t_upd = 0;
    for (int it = 0; it < iters; it++) {
        double t0 = orc_now_ms();
        FN(omp)(Y, dim, P, D, K, kmax, idx, coef, cnt, NULL);
        double t1 = orc_now_ms();
        FN(update)(Y, dim, P, D, K, kmax, idx, coef, cnt, cap, NULL, NULL);
        double t2 = orc_now_ms();
        t_omp += t1 - t0;
        t_upd += t2 - t1;
    }
    if (final_omp) FN(omp)(Y, dim, P, D, K, kmax, idx, coef, cnt, NULL);
    if (t_ms) { t_ms[0] = t_omp; t_ms[1] = t_upd; }
}

#undef FN
