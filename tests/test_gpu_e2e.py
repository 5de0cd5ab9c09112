"""End-to-end parity of the plain FP32 form at the 512x512 and 2048x2048
workloads, and checks of the dictionary sweep that hold whatever an atom's
spectral gap.

Bars (the plain form's contract; cuda_svd_denoiser.cu:368-408, :413-481):
  * full K-SVD (10 iterations of OMP + Gauss-Seidel sweep, the final OMP and
    the rebuild) on 24 seeds at 512x512 and six at 2048x2048: the mean of
    (gpu - FP64 oracle) raw PSNR within 0.02 dB plus twice its standard error
    under FP32 rounding (the FP32 oracle's spread), and the rms of it no more
    than 0.01 dB above the rms of (FP32 oracle - FP64 oracle), the spread that
    FP32 rounding alone causes over ten chaotic iterations;
  * each iteration of the GPU's trajectory restarted in the oracle: raw PSNR
    after one step within 0.02 dB;
  * the first OMP at 2048x2048: identical supports on >= 99.9 % of patches,
    coefficients within 1e-4 relative of the FP64 oracle's;
  * one sweep from the oracle's codes: sign-fixed atoms and coefficients
    within 1e-4 where the gap lambda2/lambda1 < 0.9, and for every atom the
    GPU's d_k is a top singular vector of the restricted residual E_k the GPU
    itself saw: ||E_k^T d_k||^2 >= (1 - 1e-6) sigma_1(E_k)^2, with the new
    coefficients equal to E_k^T d_k (cuda_svd_denoiser.cu:343-358).
Every result also states its error against the FP64 oracle.
"""
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CFG2 = dict(W=512, H=512, sigma=25.0, patch=(12, 12), slide=(3, 3), atoms=256, iters=10, k=5, seed=2)
CFG3 = dict(W=2048, H=2048, sigma=20.0, patch=(12, 12), slide=(3, 3), atoms=256, iters=10, k=5, seed=3)


def _say(msg):
    print(msg, flush=True)


def sorted_support(idx, coef, cnt):
    """Supports sorted by atom (padding -1 / 0) with their coefficients."""
    P, kmax = idx.shape
    live = np.arange(kmax)[None, :] < cnt[:, None]
    key = np.where(live, idx, np.iinfo(np.int32).max)
    o = np.argsort(key, axis=1, kind="stable")
    si = np.take_along_axis(np.where(live, idx, -1), o, axis=1)
    sc = np.take_along_axis(np.where(live, coef, 0).astype(np.float64), o, axis=1)
    return si, sc


def compare_codes(a, b):
    """(fraction of patches with identical supports, max relative coefficient
    error over those patches) between two compact codes."""
    ia, ca = sorted_support(*a)
    ib, cb = sorted_support(*b)
    same = np.all(ia == ib, axis=1) & (a[2] == b[2])
    nb = np.maximum(np.linalg.norm(cb, axis=1), 1e-30)
    err = np.max(np.abs(ca - cb), axis=1) / nb
    return same.mean(), (err[same].max() if same.any() else np.inf)


def signfix(D):
    s = np.sign(D[np.arange(D.shape[0]), np.argmax(np.abs(D), axis=1)])
    s[s == 0] = 1
    return D * s[:, None], s


def sweep_self_check(Y, D0, idx, c0, cnt, D1, c1, cap):
    """Per atom k: rebuild the restricted residual E_k the sweep worked on (the
    first min(c_k, cap) users of atom k in the codes at entry, ascending patch;
    atoms before k at their new values, atoms after k at their old ones), and
    return (1 - ||E_k^T d_k||^2 / sigma_1^2, max |coef - E_k^T d_k| / max |E_k^T d_k|,
    | ||d_k|| - 1 |) for every atom that has users."""
    P, kmax = idx.shape
    live = (np.arange(kmax)[None, :] < cnt[:, None]) & (c0 != 0)
    K, dim = D0.shape
    out = []
    for k in range(K):
        pk, sk = np.nonzero(live & (idx == k))  # row-major: ascending patch
        if len(pk) == 0:
            continue
        m = min(len(pk), cap) if cap > 0 else len(pk)
        T, slot = pk[:m], sk[:m]
        ii = idx[T]
        new = ii < k
        cc = np.where(new, c1[T], c0[T]).astype(np.float64)
        keep = (np.arange(kmax)[None, :] < cnt[T][:, None]) & (ii != k)
        cc = np.where(keep, cc, 0.0)
        Dn = np.where(new[..., None], D1[ii], D0[ii]).astype(np.float64)
        E = Y[T].astype(np.float64) - np.einsum("tj,tjd->td", cc, Dn)
        s1 = np.linalg.svd(E, compute_uv=False)[0]
        d = D1[k].astype(np.float64)
        proj = E @ d
        deficit = 1.0 - float(proj @ proj) / max(s1 * s1, 1e-300)
        cerr = np.max(np.abs(c1[T, slot] - proj)) / max(np.max(np.abs(proj)), 1e-30)
        out.append((deficit, cerr, abs(np.linalg.norm(d) - 1.0)))
    return np.array(out)


def run_gpu(kl, cfg, **kw):
    clean, noisy = kl.synth_image(cfg["seed"], cfg["W"], cfg["H"], cfg["sigma"])
    ks = kl.KSvd(patch=cfg["patch"], slide=cfg["slide"], atoms=cfg["atoms"], omp_iter=cfg["k"],
                 ksvd_iter=cfg["iters"], **kw)
    ks.set_picture(noisy)
    ks.extract()
    ks.init_dictionary()
    ks.run()
    img = ks.rebuild_picture(kl.noise_variance(noisy))
    t = ks.timings()
    ks.close()
    return clean, noisy, img, t


def run_oracle(kl, orc, cfg, real, perturb=False):
    """The oracle's whole K-SVD and rebuild.  perturb: the initial dictionary's
    first entry moved by one FP32 ulp (what one rounding does to the result)."""
    clean, noisy = kl.synth_image(cfg["seed"], cfg["W"], cfg["H"], cfg["sigma"])
    W, H = cfg["W"], cfg["H"]
    (pw, ph), (sw, sh) = cfg["patch"], cfg["slide"]
    fr = kl.image_to_frame(noisy)
    Y = orc.extract(fr, W, H, pw, ph, sw, sh, real=real)
    D0 = orc.init_dict(Y, cfg["atoms"], real=real)
    if perturb:
        D0[0, 0] = np.nextafter(D0[0, 0], D0.dtype.type(10))
    D, idx, coef, cnt, _ = orc.ksvd(Y, D0, cfg["k"], cfg["iters"], real=real)
    out, *_ = orc.rebuild(idx, coef, cnt, D, fr, W, H, pw, ph, sw, sh, kl.noise_variance(noisy), real=real)
    return kl.frame_to_image(out.astype(np.float32), W, H)


def end_to_end(kl, orc, cfg, perturbed=True):
    """Raw PSNR of the GPU, FP32-oracle, FP64-oracle and (perturbed) one-ulp
    FP32-oracle runs of one seeded input, the noisy PSNR, the outputs and the
    clean picture."""
    clean, noisy, img, _ = run_gpu(kl, cfg)
    outs = [img, run_oracle(kl, orc, cfg, "f32"), run_oracle(kl, orc, cfg, "f64")]
    if perturbed:
        outs.append(run_oracle(kl, orc, cfg, "f32", perturb=True))
    return [kl.psnr(clean, x) for x in outs], kl.psnr(clean, noisy), outs, clean


def rms(v):
    return float(np.sqrt(np.mean(np.square(v))))


@pytest.mark.parametrize("cfg,seeds,perturbed", [(CFG2, range(2, 26), True), (CFG3, range(3, 9), False)],
                         ids=["512sq_24seeds", "2048sq_6seeds"])
def test_end_to_end_psnr(kl, orc, cfg, seeds, perturbed):
    """Full K-SVD, final OMP and rebuild on several seeds.  Ten K-SVD
    iterations amplify any rounding: the FP64 oracle, or the FP32 oracle from
    a dictionary one ulp away, ends a few hundredths of a dB (up to 0.09 dB at
    2048^2) from the FP32 oracle, whose own sums are one arbitrary FP32 order
    among many, like the reference's CUDA sums.  So the FP64 oracle stands for
    the algorithm and the 0.02 dB bar holds (a) for the mean over seeds of
    (gpu - oracle64), the systematic part, up to twice its standard error
    under FP32 rounding alone (2 rms(oracle32 - oracle64) / sqrt(seeds):
    0.01 dB on 24 seeds at 512^2, 0.04 dB on 6 at 2048^2), while (b) the
    GPU's spread around the FP64 oracle is no wider than the FP32 oracle's:
    rms(gpu - oracle64) <= rms(oracle32 - oracle64) + 0.01 dB.  gpu - oracle32
    is reported.  The
    normalized PSNR (the reference's report) hinges on the picture's extreme
    pixels and is reported only; each seed must denoise."""
    rows = []
    for seed in seeds:
        t0 = time.time()
        c = dict(cfg, seed=seed)
        r, rn, outs, clean = end_to_end(kl, orc, c, perturbed)
        rg, r32, r64 = r[:3]
        pg, p32, p64 = (kl.psnr(clean, kl.normalize(x)) for x in outs[:3])
        ulp = (r[3] - r32) if perturbed else np.nan
        rows.append((rg - r32, r64 - r32, rg - r64, ulp, pg - p32, p64 - p32))
        _say(f"  {cfg['W']}^2 seed {seed} ({time.time() - t0:.0f} s): raw PSNR gpu {rg:.4f} oracle32 {r32:.4f} "
             f"oracle64 {r64:.4f} (noisy {rn:.4f}); gpu - oracle32 {rg - r32:+.4f}, oracle64 - oracle32 "
             f"{r64 - r32:+.4f}, gpu - oracle64 {rg - r64:+.4f}" + (f", one-ulp oracle32 - oracle32 {ulp:+.4f}" if perturbed else "")
             + f" dB; normalized gpu {pg:.4f} oracle32 {p32:.4f} oracle64 {p64:.4f}")
        assert rg > rn + 3
    a = np.array(rows)
    _say(f"{cfg['W']}^2 over {len(a)} seeds: mean gpu - oracle32 {a[:, 0].mean():+.4f} dB, oracle64 - oracle32 "
         f"{a[:, 1].mean():+.4f}, gpu - oracle64 {a[:, 2].mean():+.4f}; rms gpu - oracle32 {rms(a[:, 0]):.4f}, "
         f"rms oracle64 - oracle32 {rms(a[:, 1]):.4f}, rms gpu - oracle64 {rms(a[:, 2]):.4f}"
         + (f", rms one-ulp move {rms(a[:, 3]):.4f}" if perturbed else "")
         + f"; normalized PSNR (min-max stretch, set by single extreme pixels): rms gpu - oracle32 {rms(a[:, 4]):.3f}, "
         f"oracle64 - oracle32 {rms(a[:, 5]):.3f}")
    se = rms(a[:, 1]) / np.sqrt(len(a))
    _say(f"bars: |mean gpu - oracle64| {abs(a[:, 2].mean()):.4f} <= 0.02 + 2 x {se:.4f}; rms gpu - oracle64 "
         f"{rms(a[:, 2]):.4f} <= {rms(a[:, 1]):.4f} + 0.01")
    assert abs(a[:, 2].mean()) <= 0.02 + 2 * se
    assert rms(a[:, 2]) <= rms(a[:, 1]) + 0.01


@pytest.mark.parametrize("cfg", [CFG2, CFG3], ids=["512sq", "2048sq"])
def test_one_step_psnr_along_trajectory(kl, orc, cfg):
    """Each of the ten iterations of the GPU's run, restarted in the oracle
    from the same dictionary: OMP, the sweep, the final OMP and the rebuild of
    the GPU and of the FP32 oracle give raw PSNR within 0.02 dB (no
    amplification over iterations enters this comparison)."""
    clean, noisy = kl.synth_image(cfg["seed"], cfg["W"], cfg["H"], cfg["sigma"])
    W, H = cfg["W"], cfg["H"]
    (pw, ph), (sw, sh) = cfg["patch"], cfg["slide"]
    s2 = kl.noise_variance(noisy)
    fr = kl.image_to_frame(noisy)
    ks = kl.KSvd(patch=cfg["patch"], slide=cfg["slide"], atoms=cfg["atoms"], omp_iter=cfg["k"], ksvd_iter=1)
    ks.set_picture(noisy)
    ks.extract()
    ks.init_dictionary()
    Y = ks.patches()
    D = ks.dictionary()
    worst = 0.0
    for it in range(cfg["iters"]):
        ks.set_dictionary(D)
        ks.omp()
        ks.update()
        Dg = ks.dictionary()
        ks.omp()
        rg = kl.psnr(clean, ks.rebuild_picture(s2))
        io, co, no, _ = orc.omp(Y, D, cfg["k"], real="f32")
        Do, _, _, _ = orc.update(Y, D, io, co, no, cap=1024, real="f32")
        io, co, no, _ = orc.omp(Y, Do, cfg["k"], real="f32")
        out, *_ = orc.rebuild(io, co, no, Do, fr, W, H, pw, ph, sw, sh, s2, real="f32")
        ro = kl.psnr(clean, kl.frame_to_image(out.astype(np.float32), W, H))
        worst = max(worst, abs(rg - ro))
        _say(f"  {W}^2 iteration {it + 1}: raw PSNR gpu {rg:.4f} oracle32 {ro:.4f} ({rg - ro:+.5f} dB)")
        assert abs(rg - ro) <= 0.02
        D = Dg
    ks.close()
    _say(f"{W}^2 one-step parity over {cfg['iters']} iterations: max |gpu - oracle32| {worst:.5f} dB")


@pytest.fixture(scope="module")
def big(kl, orc):
    """The 2048x2048 workload's patches, initial dictionary and the oracle's
    first OMP (FP32 and FP64)."""
    cfg = CFG3
    clean, noisy = kl.synth_image(cfg["seed"], cfg["W"], cfg["H"], cfg["sigma"])
    ks = kl.KSvd(patch=cfg["patch"], slide=cfg["slide"], atoms=cfg["atoms"], omp_iter=cfg["k"],
                 ksvd_iter=cfg["iters"])
    ks.set_picture(noisy)
    ks.extract()
    ks.init_dictionary()
    Y = ks.patches()
    D0 = ks.dictionary()
    o32 = orc.omp(Y, D0, cfg["k"], real="f32")[:3]
    o64 = orc.omp(Y, D0.astype(np.float64), cfg["k"], real="f64")[:3]
    yield ks, Y, D0, o32, o64
    ks.close()


def test_first_omp_2048(big):
    """The first OMP of the 2048x2048 workload.  Its dictionary is cut from
    overlapping patches, so the support Grams have condition numbers near 1e3
    and FP32 least squares is good to about 1e-4 there: the FP32 oracle's
    sequential sums themselves stray up to 2e-4 from the FP64 oracle.  So the
    supports are held to the FP32 oracle, the coefficients to the FP64 oracle
    (the exact algorithm) within 1e-4, and no worse than the FP32 oracle's."""
    ks, Y, D0, o32, o64 = big
    ks.set_dictionary(D0)
    ks.omp()
    g = ks.codes()
    frac, err = compare_codes(g, o32)
    frac64, err64 = compare_codes(g, o64)
    frac3264, err3264 = compare_codes(o32, o64)
    _say(f"\nfirst OMP 2048^2 ({len(g[2])} patches): supports equal {frac * 100:.4f}% of the FP32 oracle's "
         f"(coef err {err:.2e}); vs FP64 oracle {frac64 * 100:.4f}% (coef err {err64:.2e}); FP32 oracle vs FP64 "
         f"oracle {frac3264 * 100:.4f}% (coef err {err3264:.2e})")
    assert frac >= 0.999
    assert err64 <= 1e-4
    assert err64 <= err3264


def test_one_sweep_2048(big, orc):
    ks, Y, D0, o32, o64 = big
    io, co, no = o32
    ks.set_dictionary(D0)
    ks.set_codes(io, co, no)
    ks.update()
    Dg = ks.dictionary()
    cg = ks.codes()[1]
    Do, coo, gap, _ = orc.update(Y, D0, io, co, no, cap=1024, real="f32", want_gap=True)
    D64, _, _, _ = orc.update(Y, D0.astype(np.float64), io, co, no.copy(), cap=1024, real="f64")
    good = (gap >= 0) & (gap < 0.9)
    Dgs, sg = signfix(Dg)
    Dos, so = signfix(Do)
    D64s, _ = signfix(D64)
    aerr = np.max(np.abs(Dgs - Dos), axis=1)
    aerr64 = np.max(np.abs(Dgs - D64s), axis=1)
    P, kmax = io.shape
    live = np.arange(kmax)[None, :] < no[:, None]
    cerr = np.zeros(D0.shape[0])
    for k in np.nonzero(good)[0]:
        sel = (io == k) & live
        cerr[k] = np.max(np.abs(cg[sel] * sg[k] - coo[sel] * so[k])) / max(np.max(np.abs(coo[sel])), 1e-30)
    chk = sweep_self_check(Y, D0, io, co, no, Dg, cg, 1024)
    _say(f"\none sweep 2048^2: {good.sum()} of {(gap >= 0).sum()} atoms with gap < 0.9: atom err {aerr[good].max():.2e} "
         f"(vs FP64 oracle {aerr64[good].max():.2e}), coef err {cerr[good].max():.2e}; all atoms: "
         f"Rayleigh deficit max {chk[:, 0].max():.2e}, coef vs E^T d {chk[:, 1].max():.2e}, "
         f"|1 - norm| {chk[:, 2].max():.2e}; worst-gap atom err {aerr[~good].max() if (~good).any() else 0:.2e}")
    assert aerr[good].max() <= 1e-4
    assert cerr[good].max() <= 1e-4
    assert chk[:, 0].max() <= 1e-6
    assert chk[:, 1].max() <= 1e-4
    assert chk[:, 2].max() <= 1e-5


@pytest.mark.parametrize("k,cap,sweep", [(5, 1024, "persistent"), (5, 0, "two_kernel"), (8, 1536, "persistent"),
                                         (5, 64, "persistent"), (5, 1024, "two_team"), (5, 1024, "plain")],
                         ids=["k5_cap1024", "k5_nocap", "k8_cap1536_master", "k5_cap64", "k5_cap1024_2team",
                              "k5_cap1024_plain"])
def test_sweep_every_atom_512(kl, orc, k, cap, sweep, monkeypatch):
    """The gap-independent check at 512x512 on the sweeps that the cap and
    OMP limit select (eight-team, master/helper, two-kernel), the eight-team
    sweep's plain path (no speculation) and the two-team sweep."""
    monkeypatch.setenv("KSVD_SWEEP_TWO_KERNEL", "1" if sweep == "two_kernel" else "0")
    monkeypatch.setenv("KSVD_SWEEP_TEAMS", "2" if sweep == "two_team" else "8")
    monkeypatch.setenv("KSVD_SWEEP_SPEC", "0" if sweep == "plain" else "1")
    cfg = dict(CFG2, k=k)
    clean, noisy = kl.synth_image(cfg["seed"], cfg["W"], cfg["H"], cfg["sigma"])
    ks = kl.KSvd(patch=cfg["patch"], slide=cfg["slide"], atoms=cfg["atoms"], omp_iter=k, user_cap=cap)
    ks.set_picture(noisy)
    ks.extract()
    ks.init_dictionary()
    Y = ks.patches()
    D0 = ks.dictionary()
    io, co, no, _ = orc.omp(Y, D0, k, real="f32")
    ks.set_codes(io, co, no)
    ks.update()
    Dg = ks.dictionary()
    cg = ks.codes()[1]
    ks.close()
    chk = sweep_self_check(Y, D0, io, co, no, Dg, cg, cap)
    _say(f"\nsweep 512^2 k={k} cap={cap} {sweep}: {len(chk)} atoms: Rayleigh deficit max {chk[:, 0].max():.2e}, "
         f"coef vs E^T d {chk[:, 1].max():.2e}, |1 - norm| {chk[:, 2].max():.2e}")
    assert chk[:, 0].max() <= 1e-6
    assert chk[:, 1].max() <= 1e-4
    assert chk[:, 2].max() <= 1e-5
