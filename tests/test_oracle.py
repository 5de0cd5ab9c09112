"""Pins the CPU oracle (oracle/) against the golden fixtures, closed forms,
brute-force restatements and hand-made cases.  CPU only.

Each test names the reference behaviour it pins (paths relative to the
reference repository).
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def brute_patches(frame, pw, ph, sw, sh):
    """Independent index-set enumeration of cuda_svd_denoiser.cu:231-240."""
    n, m = frame.shape
    out = []
    for i in range(0, n - ph + 1, sh):
        for j in range(0, m - pw + 1, sw):
            out.append(frame[i:i + ph, j:j + pw].ravel())
    return np.array(out)


# ---- patch gather ---------------------------------------------------------
@pytest.mark.parametrize("m,n,pw,ph,sw,sh", [
    (16, 16, 8, 8, 2, 2), (128, 128, 8, 8, 2, 2), (512, 512, 12, 12, 3, 3), (2048, 2048, 12, 12, 3, 3),
    (37, 23, 5, 7, 3, 2), (12, 12, 12, 12, 3, 3), (10, 30, 4, 6, 5, 1)])
def test_patch_count_closed_form(orc, m, n, pw, ph, sw, sh):
    expect = ((n - ph) // sh + 1) * ((m - pw) // sw + 1)
    assert orc.patch_count(m, n, pw, ph, sw, sh) == expect
    if m * n <= 4096:
        assert len(brute_patches(np.zeros((n, m)), pw, ph, sw, sh)) == expect


def test_patch_count_known_sizes(orc):
    # sizes stated for the three benchmark shapes (128^2 8x8/2, 512^2 and 2048^2 12x12/3)
    assert orc.patch_count(128, 128, 8, 8, 2, 2) == 3721
    assert orc.patch_count(512, 512, 12, 12, 3, 3) == 27889
    assert orc.patch_count(2048, 2048, 12, 12, 3, 3) == 461041


@pytest.mark.parametrize("m,n,pw,ph,sw,sh", [(16, 16, 8, 8, 2, 2), (37, 23, 5, 7, 3, 2), (9, 9, 3, 3, 1, 1)])
def test_extract_index_sets(orc, m, n, pw, ph, sw, sh):
    # the reference's walk of its frame (compat): n rows of length m
    frame = np.arange(m * n, dtype=np.float32).reshape(n, m)
    Y = orc.extract(frame, m, n, pw, ph, sw, sh, compat=True)
    np.testing.assert_array_equal(Y, brute_patches(frame, pw, ph, sw, sh))
    Y64 = orc.extract(frame, m, n, pw, ph, sw, sh, real="f64", compat=True)
    np.testing.assert_array_equal(Y64, Y.astype(np.float64))


def picture_blocks(pic, pw, ph, sw, sh):
    """True geometry from the picture itself (H x W): patch (a, b) is the
    block of columns x in [a, a + ph) and rows y in [b, b + pw), a stepping by
    sh and b by sw, elements in (x, y) order -- for a square picture the
    reference's patch of its transposed frame (cuda_svd_denoiser.cu:231-240,
    denoiser.cpp:46)."""
    H, W = pic.shape
    out = []
    for a in range(0, W - ph + 1, sh):
        for b in range(0, H - pw + 1, sw):
            out.append(pic[b:b + pw, a:a + ph].T.ravel())
    return np.array(out)


@pytest.mark.parametrize("W,H,pw,ph,sw,sh", [(16, 16, 8, 8, 2, 2), (37, 23, 5, 7, 3, 2), (96, 64, 12, 8, 3, 2),
                                             (10, 30, 4, 6, 5, 1), (23, 37, 7, 5, 2, 3)])
def test_extract_true_geometry_is_picture_blocks(orc, W, H, pw, ph, sw, sh):
    pic = np.arange(W * H, dtype=np.float32).reshape(H, W)
    frame = np.ascontiguousarray(pic.T).ravel()  # buf[x*H + y] = pic[y][x] (denoiser.cpp:46)
    Y = orc.extract(frame, W, H, pw, ph, sw, sh)
    np.testing.assert_array_equal(Y, picture_blocks(pic, pw, ph, sw, sh))
    if W == H:
        np.testing.assert_array_equal(Y, orc.extract(frame, W, H, pw, ph, sw, sh, compat=True))
    elif pw * ph > 1:
        assert not np.array_equal(Y[:1], orc.extract(frame, W, H, pw, ph, sw, sh, compat=True)[:1])


def test_rebuild_true_geometry_overlap_add(orc):
    """Overlap-add of the true geometry from the picture: every patch's values
    go back onto its own block; sums and weights in picture coordinates."""
    rng = np.random.default_rng(3)
    W, H, pw, ph, sw, sh, K, k = 29, 17, 6, 4, 3, 2, 10, 3
    pic = rng.uniform(0, 255, (H, W)).astype(np.float32)
    frame = np.ascontiguousarray(pic.T).ravel()
    P = orc.patch_count(H, W, pw, ph, sw, sh)
    D = rng.standard_normal((K, pw * ph)).astype(np.float32)
    idx = np.stack([rng.choice(K, k, replace=False) for _ in range(P)]).astype(np.int32)
    coef = rng.uniform(-9, 9, (P, k)).astype(np.float32)
    cnt = np.full(P, k, np.int32)
    out, s, w, dsum, lam = orc.rebuild(idx, coef, cnt, D.astype(np.float64), frame, W, H, pw, ph, sw, sh, 50.0,
                                       real="f64")
    S = np.zeros((H, W))
    Wt = np.zeros((H, W))
    p = 0
    for a in range(0, W - ph + 1, sh):
        for b in range(0, H - pw + 1, sw):
            v = sum(coef[p, i].astype(np.float64) * D[idx[p, i]].astype(np.float64) for i in range(k))
            S[b:b + pw, a:a + ph] += v.reshape(ph, pw).T
            Wt[b:b + pw, a:a + ph] += 1
            p += 1
    assert p == P
    # the oracle's arrays are in the buffer's layout (row x of length H)
    np.testing.assert_array_equal(w.reshape(W, H).T, Wt)
    np.testing.assert_allclose(s.reshape(W, H).T, S, rtol=1e-12, atol=1e-9)


# ---- dictionary init --------------------------------------------------------
def test_init_dict_integer_patches_unit_norm(orc):
    # integer pixels: every partial sum of squares is an integer, so the int
    # accumulator of cuda_svd_denoiser.cu:59 loses nothing -> exact unit norm
    rng = np.random.default_rng(1)
    Y = rng.integers(0, 256, (600, 64)).astype(np.float32)
    D = orc.init_dict(Y, 256)
    np.testing.assert_array_equal(D[0] * 0 + Y[300] / np.float32(np.sqrt(np.sum(Y[300].astype(np.int64) ** 2))),
                                  D[0])
    np.testing.assert_allclose(np.linalg.norm(D.astype(np.float64), axis=1), 1.0, rtol=2e-7)
    # the pick starts at column floor(P/2) (cuda_svd_denoiser.cu:267)
    for k in (0, 17, 255):
        assert np.allclose(D[k] * np.linalg.norm(Y[300 + k]), Y[300 + k], rtol=1e-6)


def test_init_dict_truncating_accumulator(orc):
    # patch [1.5]*4: squares 2.25; the int running sum goes 2, 4, 6, 8, so the
    # norm is sqrt(8) (the exact norm would be 3)
    Y = np.full((3, 4), 1.5, np.float32)
    D = orc.init_dict(Y, 2)
    assert D[0, 0] == np.float32(1.5) / np.float32(np.sqrt(np.float32(8)))
    assert abs(D[0, 0] - 0.5303300858899106) < 1e-7


# ---- OMP ----------------------------------------------------------------------
def test_omp_matches_lstsq_golden(orc):
    g = np.load(os.path.join(GOLD, "omp_golden.npz"))
    idx, coef, cnt, st = orc.omp(g["Y"], g["D"], int(g["kmax"]), real="f64")
    np.testing.assert_array_equal(cnt, g["cnt"])
    for p in range(len(cnt)):
        np.testing.assert_array_equal(idx[p, :cnt[p]], g["idx"][p, :cnt[p]])
    np.testing.assert_allclose(coef, g["coef"], rtol=1e-9, atol=1e-9)
    # FP32 run of the same problem: same supports, coefficients close
    idx32, coef32, cnt32, _ = orc.omp(g["Y"], g["D"], int(g["kmax"]), real="f32")
    np.testing.assert_array_equal(cnt32, g["cnt"])
    np.testing.assert_array_equal(idx32, idx)
    np.testing.assert_allclose(coef32, g["coef"], rtol=2e-4, atol=2e-4)


def test_omp_special_patches(orc):
    g = np.load(os.path.join(GOLD, "omp_golden.npz"))
    idx, coef, cnt, st = orc.omp(g["Y"], g["D"], int(g["kmax"]), real="f64")
    # zero patch: first pick is atom 0 (all |c| tie), coefficient 0, retired
    assert cnt[5] == 1 and idx[5, 0] == 0 and coef[5, 0] == 0.0
    # exact multiple of atom 7 retires after one step
    assert cnt[9] == 1 and idx[9, 0] == 7 and abs(coef[9, 0] - 2.5) < 1e-12


def test_omp_residual_orthogonal_to_support(orc):
    rng = np.random.default_rng(3)
    D = rng.standard_normal((64, 20))
    D /= np.linalg.norm(D, axis=1, keepdims=True)
    Y = rng.standard_normal((200, 20)) * 50
    idx, coef, cnt, _ = orc.omp(Y, D, 6, real="f64")
    for p in range(200):
        S = idx[p, :cnt[p]]
        r = Y[p] - coef[p, :cnt[p]] @ D[S]
        assert np.max(np.abs(D[S] @ r)) < 1e-9 * np.linalg.norm(Y[p])
        assert cnt[p] == 6


def test_omp_first_index_tie_break(orc):
    # atoms 3 and 7 identical: |c_3| == |c_7| exactly, Isamax returns the first
    rng = np.random.default_rng(4)
    D = rng.standard_normal((12, 8))
    D /= np.linalg.norm(D, axis=1, keepdims=True)
    D[7] = D[3]
    Y = 3.0 * D[3][None, :] + np.zeros((1, 8))
    for real in ("f32", "f64"):
        idx, coef, cnt, _ = orc.omp(Y, D, 3, real=real)
        assert idx[0, 0] == 3 and cnt[0] == 1


def test_omp_duplicate_pick_stops(orc):
    # D spans only 2 of 3 dimensions: after two picks r is orthogonal to every
    # atom, all |c| are ~0 and the third pick would repeat an atom -> stop.
    D = np.array([[1, 0, 0], [0, 1, 0], [0.6, 0.8, 0]], dtype=np.float64)
    Y = np.array([[3.0, 2.0, 5.0]])
    idx, coef, cnt, st = orc.omp(Y, D, 3, real="f64")
    assert cnt[0] == 2
    assert st[0] + st[1] == 1


# ---- dictionary update -----------------------------------------------------
def _signfix(D, coef, idx, cnt):
    D = D.copy()
    coef = coef.copy()
    for k in range(D.shape[0]):
        if D[k, np.argmax(np.abs(D[k]))] < 0:
            D[k] = -D[k]
            sel = (idx == k) & (np.arange(idx.shape[1])[None, :] < cnt[:, None])
            coef[sel] = -coef[sel]
    return D, coef


@pytest.mark.parametrize("cap,key", [(1024, "cap"), (0, "nocap")])
def test_update_matches_svd_golden(orc, cap, key):
    g = np.load(os.path.join(GOLD, "update_golden.npz"))
    D, coef, gap, it = orc.update(g["Y"], g["D"], g["idx"], g["coef"], g["cnt"], cap=cap, real="f64",
                                  want_gap=True)
    Dg, cg = _signfix(g["D_" + key], g["coef_" + key], g["idx"], g["cnt"])
    Do, co = _signfix(D, coef, g["idx"], g["cnt"])
    np.testing.assert_allclose(Do, Dg, atol=1e-9)
    np.testing.assert_allclose(co, cg, atol=1e-8)


def test_update_cap_zeroes_late_users(orc):
    g = np.load(os.path.join(GOLD, "update_golden.npz"))
    D, coef, _, _ = orc.update(g["Y"], g["D"], g["idx"], g["coef"], g["cnt"], cap=1024, real="f64")
    idx = g["idx"]
    users0 = [p for p in range(len(idx)) if 0 in idx[p]]
    assert len(users0) == 1500
    slot = [list(idx[p]).index(0) for p in users0]
    late = np.array([coef[p, s] for p, s in zip(users0[1024:], slot[1024:])])
    early = np.array([coef[p, s] for p, s in zip(users0[:1024], slot[:1024])])
    assert np.all(late == 0)
    assert np.all(early != 0)
    # the cap changes the answer: the uncapped sweep differs on atom 0
    D0, _, _, _ = orc.update(g["Y"], g["D"], g["idx"], g["coef"], g["cnt"], cap=0, real="f64")
    assert np.max(np.abs(np.abs(D0[0]) - np.abs(D[0]))) > 1e-6


def test_update_is_gauss_seidel(orc):
    # atom 9 is updated after atoms 0..8 changed: its restricted residual must
    # use their new atoms and coefficients.  The golden (sequential) agrees;
    # a Jacobi-style update of atom 9 from the entry state does not.
    g = np.load(os.path.join(GOLD, "update_golden.npz"))
    Y, idx, coef, cnt = g["Y"], g["idx"], g["coef"], g["cnt"]
    D, _, _, _ = orc.update(Y, g["D"], idx, coef, cnt, cap=1024, real="f64")
    keep = np.where((idx == 9).any(axis=1))[0]
    e_entry = []
    for p in keep[:1024]:
        e = Y[p].copy()
        for i in range(cnt[p]):
            if idx[p, i] != 9:
                e -= g["D"][idx[p, i]] * coef[p, i]
        e_entry.append(e)
    u_j = np.linalg.svd(np.array(e_entry).T, full_matrices=False)[0][:, 0]
    assert np.max(np.abs(np.abs(u_j) - np.abs(D[9]))) > 1e-6
    np.testing.assert_allclose(np.abs(D[9]), np.abs(g["D_cap"][9]), atol=1e-9)


def test_update_fp32_close_to_fp64(orc):
    g = np.load(os.path.join(GOLD, "update_golden.npz"))
    D64, c64, _, _ = orc.update(g["Y"], g["D"], g["idx"], g["coef"], g["cnt"], cap=1024, real="f64")
    D32, c32, _, _ = orc.update(g["Y"], g["D"], g["idx"], g["coef"], g["cnt"], cap=1024, real="f32")
    np.testing.assert_allclose(D32, D64, atol=2e-6)
    np.testing.assert_allclose(c32, c64, rtol=1e-5, atol=1e-5)


# ---- rebuild --------------------------------------------------------------------
def test_rebuild_matches_loop_golden(orc):
    g = np.load(os.path.join(GOLD, "rebuild_golden.npz"))
    m, n, pw, ph, sw, sh = [int(v) for v in g["geom"]]
    out, s, w, dsum, lam = orc.rebuild(g["idx"], g["coef"], g["cnt"], g["D"], g["buf"], m, n, pw, ph, sw, sh,
                                       float(g["sigma2"]), compat=True, real="f64")
    np.testing.assert_array_equal(w, g["weight"])
    np.testing.assert_allclose(s, g["sum"], rtol=1e-12, atol=1e-9)
    assert abs(lam - float(g["lam"])) < 1e-6 * max(1.0, float(g["lam"]))
    np.testing.assert_allclose(out, g["out"], rtol=1e-6, atol=1e-6)
    # uncovered last row keeps the input (lambda-mask decision)
    assert np.all(g["weight"][-1] == 0)
    np.testing.assert_array_equal(out[-1], g["buf"][-1].astype(np.float64))


def test_rebuild_f32_sums_close(orc):
    g = np.load(os.path.join(GOLD, "rebuild_golden.npz"))
    m, n, pw, ph, sw, sh = [int(v) for v in g["geom"]]
    out, s, w, dsum, lam = orc.rebuild(g["idx"], g["coef"], g["cnt"], g["D"], g["buf"], m, n, pw, ph, sw, sh,
                                       float(g["sigma2"]), compat=True, real="f32")
    np.testing.assert_array_equal(w, g["weight"])
    np.testing.assert_allclose(s, g["sum"], rtol=1e-5, atol=1e-3)
    np.testing.assert_allclose(out, g["out"], rtol=1e-4, atol=1e-3)


def test_rebuild_lambda_closed_form(orc):
    # codes that reproduce the image exactly: first == y, d = 0, lambda = 1,
    # out = (y + w*y) / (1 + w) = y
    m = n = 8
    pw = ph = 4
    sw = sh = 2
    frame = np.random.default_rng(5).uniform(0, 100, (n, m)).astype(np.float32)
    Y = orc.extract(frame, m, n, pw, ph, sw, sh, real="f64")
    P = len(Y)
    D = Y / np.linalg.norm(Y, axis=1, keepdims=True)
    idx = np.arange(P, dtype=np.int32)[:, None]
    coef = np.linalg.norm(Y, axis=1)[:, None]
    cnt = np.ones(P, np.int32)
    out, s, w, dsum, lam = orc.rebuild(idx, coef, cnt, D, frame, m, n, pw, ph, sw, sh, 25.0, real="f64")
    assert dsum < 1e-20 and lam == 1.0
    np.testing.assert_allclose(out, frame, rtol=1e-12)


def test_rebuild_compat_equals_true_for_square(orc):
    g = np.load(os.path.join(GOLD, "rebuild_golden.npz"))
    rng = np.random.default_rng(6)
    m = n = 17
    pw = ph = 5
    sw = sh = 3
    frame = rng.uniform(0, 255, (n, m)).astype(np.float32)
    P = orc.patch_count(m, n, pw, ph, sw, sh)
    D = g["D"]
    idx = np.stack([rng.choice(D.shape[0], 3, replace=False) for _ in range(P)]).astype(np.int32)
    coef = rng.uniform(-50, 50, (P, 3))
    cnt = np.full(P, 3, np.int32)
    a = orc.rebuild(idx, coef, cnt, D, frame, m, n, pw, ph, sw, sh, 100.0, compat=False, real="f32")
    b = orc.rebuild(idx, coef, cnt, D, frame, m, n, pw, ph, sw, sh, 100.0, compat=True, real="f32")
    for x, y in zip(a[:3], b[:3]):
        np.testing.assert_array_equal(x, y)


# ---- whole loop -------------------------------------------------------------------
def test_ksvd_small_denoises(orc, kl):
    clean, noisy = kl.synth_image(7, 64, 64, 20.0)
    frame = kl.image_to_frame(noisy)
    Y = orc.extract(frame, 64, 64, 8, 8, 2, 2)
    D0 = orc.init_dict(Y, 128)
    D, idx, coef, cnt, t = orc.ksvd(Y, D0, 5, 3)
    s2 = kl.noise_variance(noisy)
    out, *_ = orc.rebuild(idx, coef, cnt, D, frame, 64, 64, 8, 8, 2, 2, s2)
    img = kl.frame_to_image(out.astype(np.float32), 64, 64)
    assert kl.psnr(clean, img) > kl.psnr(clean, noisy) + 3.0
