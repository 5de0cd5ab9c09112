"""Writes the small golden fixtures of tests/golden/ from independent numpy /
LAPACK restatements of the reference algorithm (no code shared with the C
oracle):

  omp_golden.npz      OMP on a tiny seeded problem; least squares by
                      numpy.linalg.lstsq (LAPACK gelsd), selection by
                      numpy.argmax of |D^T r| (first index wins a tie);
                      semantics of cublas_matrix_omp.cu:107-402.
  update_golden.npz   one Gauss-Seidel dictionary sweep; rank-1 factor by
                      numpy.linalg.svd (LAPACK gesdd) of the restricted
                      residual E, with the user cap of
                      cuda_svd_denoiser.cu:295-325; semantics of
                      cuda_svd_denoiser.cu:285-362.
  rebuild_golden.npz  overlap-add and lambda blend by explicit loops
                      (cuda_svd_denoiser.cu:421-481), in float64.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def ref_omp(Y, D, kmax):
    """Y (P, n), D (K, n) float64 -> idx, coef, cnt (compact, pursuit order)."""
    P, n = Y.shape
    idx = -np.ones((P, kmax), np.int32)
    coef = np.zeros((P, kmax))
    cnt = np.zeros(P, np.int32)
    for p in range(P):
        y = Y[p]
        r = y.copy()
        S = []
        x = np.zeros(0)
        for t in range(kmax):
            c = D @ r
            j = int(np.argmax(np.abs(c)))
            if j in S:
                break
            S.append(j)
            A = D[S].T
            x = np.linalg.lstsq(A, y, rcond=None)[0]
            r = y - A @ x
            if np.linalg.norm(r) < 1e-3:
                break
        cnt[p] = len(S)
        idx[p, :len(S)] = S
        coef[p, :len(S)] = x
    return idx, coef, cnt


def ref_update(Y, D, idx, coef, cnt, cap):
    """One sweep; returns new D (K, n) and coef (P, kmax), float64."""
    D = D.copy()
    coef = coef.copy()
    P, n = Y.shape
    K = D.shape[0]
    kmax = idx.shape[1]
    users = [[] for _ in range(K)]
    for p in range(P):
        for i in range(cnt[p]):
            if coef[p, i] != 0:
                users[idx[p, i]].append((p, i))
    for k in range(K):
        us = users[k]
        if not us:
            continue
        m = min(len(us), cap) if cap > 0 else len(us)
        E = np.zeros((n, len(us)))
        for t in range(m):
            p, _ = us[t]
            e = Y[p].copy()
            for i in range(cnt[p]):
                if idx[p, i] != k:
                    e -= D[idx[p, i]] * coef[p, i]
            E[:, t] = e
        U, s, Vt = np.linalg.svd(E, full_matrices=False)
        u, v = U[:, 0], Vt[0]
        if u @ D[k] < 0:  # the oracle's power iteration keeps d_k's direction
            u, v = -u, -v
        D[k] = u
        for t, (p, i) in enumerate(us):
            coef[p, i] = s[0] * v[t]
    return D, coef


def ref_rebuild(idx, coef, cnt, D, buf, m, n, pw, ph, sw, sh, sigma2):
    """Frame buffer (n, m).  Returns out, sum, weight (n, m) and lambda."""
    img = buf.reshape(n, m).astype(np.float64)
    s = np.zeros((n, m))
    w = np.zeros((n, m))
    p = 0
    for i in range(0, n - ph + 1, sh):
        for j in range(0, m - pw + 1, sw):
            v = np.zeros(pw * ph)
            for q in range(cnt[p]):
                v += D[idx[p, q]] * coef[p, q]
            s[i:i + ph, j:j + pw] += v.reshape(ph, pw)
            w[i:i + ph, j:j + pw] += 1
            p += 1
    cov = w > 0
    first = np.where(cov, s / np.where(cov, w, 1), 0)
    d = np.sum(((first - img) ** 2)[cov]) / (sigma2 * m * n)
    lam = abs(np.sqrt(d) - 1)
    out = np.where(cov, (img * lam + s) / (lam + w), img)
    return out, s, w, lam


def unit_rows(A):
    return A / np.linalg.norm(A, axis=1, keepdims=True)


def main():
    rng = np.random.default_rng(20260101)
    # --- OMP: n = 16, K = 32, P = 64, k = 4 --------------------------------
    n, K, P, k = 16, 32, 64, 4
    D = unit_rows(rng.standard_normal((K, n)))
    X = np.zeros((P, K))
    for p in range(P):
        sup = rng.choice(K, 3, replace=False)
        X[p, sup] = rng.uniform(1, 4, 3) * rng.choice([-1, 1], 3)
    Y = X @ D + 0.05 * rng.standard_normal((P, n))
    Y[5] = 0.0                      # zero patch: one pick, coefficient 0
    Y[9] = 2.5 * D[7]               # exactly one atom: retires after one step
    idx, coef, cnt = ref_omp(Y, D, k)
    np.savez(os.path.join(HERE, "omp_golden.npz"), Y=Y, D=D, kmax=k, idx=idx, coef=coef, cnt=cnt)

    # --- update: n = 12, K = 10, P = 1500, k = 3, cap 1024 (atom 0 capped) --
    n, K, P, k = 12, 10, 1500, 3
    D = unit_rows(rng.standard_normal((K, n)))
    idx = np.zeros((P, k), np.int32)
    coef = np.zeros((P, k))
    cnt = np.full(P, k, np.int32)
    for p in range(P):
        others = rng.choice(np.arange(1, K), k - 1, replace=False)
        sup = np.concatenate([[0], others])
        rng.shuffle(sup)
        idx[p] = sup
        coef[p] = rng.uniform(0.5, 3, k) * rng.choice([-1, 1], k)
    Xd = np.zeros((P, K))
    for p in range(P):
        Xd[p, idx[p]] = coef[p]
    Y = Xd @ D + 0.1 * rng.standard_normal((P, n))
    D1, coef1 = ref_update(Y, D, idx, coef, cnt, 1024)
    D0, coef0 = ref_update(Y, D, idx, coef, cnt, 0)
    np.savez(os.path.join(HERE, "update_golden.npz"), Y=Y, D=D, idx=idx, coef=coef, cnt=cnt, D_cap=D1,
             coef_cap=coef1, D_nocap=D0, coef_nocap=coef0)

    # --- rebuild: frame 23 x 19 (m x n), 5 x 5 patches, stride 3 (uncovered
    # border: 23 - 5 = 18 = 6*3 -> covered; 19 - 5 = 14 -> last row uncovered)
    m, n, pw, ph, sw, sh, K, k = 23, 19, 5, 5, 3, 3, 12, 3
    buf = rng.uniform(0, 255, (n, m)).astype(np.float32)
    D = unit_rows(rng.standard_normal((K, pw * ph)))
    P = ((n - ph) // sh + 1) * ((m - pw) // sw + 1)
    idx = np.zeros((P, k), np.int32)
    coef = np.zeros((P, k))
    cnt = rng.integers(1, k + 1, P).astype(np.int32)
    for p in range(P):
        idx[p] = rng.choice(K, k, replace=False)
        coef[p] = rng.uniform(-200, 200, k)
    sigma2 = 400.0
    out, s, w, lam = ref_rebuild(idx, coef, cnt, D, buf, m, n, pw, ph, sw, sh, sigma2)
    np.savez(os.path.join(HERE, "rebuild_golden.npz"), buf=buf, D=D, idx=idx, coef=coef, cnt=cnt,
             geom=np.array([m, n, pw, ph, sw, sh]), sigma2=sigma2, out=out, sum=s, weight=w, lam=lam)


if __name__ == "__main__":
    main()
