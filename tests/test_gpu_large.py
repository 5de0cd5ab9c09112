"""Large and rectangular patches: 16 x 12 (dim 192, the LDS-resident
two-team sweep at its widest) and 22 x 22 (dim 484: the generic pursuit and
the one-launch large-patch sweep with each workgroup's columns in LDS, or,
with KSVD_SWEEP_KL=1, three launches per atom with E and its Gram in global
memory), the patch sizes of the batched workload's files and per-file
overrides.

Bars, against the oracle on the same seeded inputs:
  * OMP: identical supports on >= 99.9 % of patches, coefficients within
    1e-4 of the FP64 oracle's;
  * one sweep from the oracle's codes: sign-fixed atoms within 1e-4 where
    lambda2/lambda1 < 0.9, and for every atom the gap-independent check of
    test_gpu_e2e (||E_k^T d_k||^2 >= (1 - 1e-6) sigma_1^2, coefficients
    equal to E_k^T d_k);
  * three K-SVD iterations end to end: raw PSNR within 0.05 dB of the FP32
    oracle (the FP64 oracle's distance printed beside it).
"""
import numpy as np
import pytest

from test_gpu_e2e import compare_codes, signfix, sweep_self_check

pytestmark = pytest.mark.gpu

CASES = [dict(W=256, H=192, patch=(16, 12), slide=(3, 3), atoms=256, k=5, seed=41, sigma=20.0),
         dict(W=256, H=192, patch=(22, 22), slide=(3, 3), atoms=256, k=5, seed=42, sigma=25.0)]
IDS = ["16x12_dim192", "22x22_dim484"]


def context(kl, cfg, **kw):
    clean, noisy = kl.synth_image(cfg["seed"], cfg["W"], cfg["H"], cfg["sigma"])
    ks = kl.KSvd(patch=cfg["patch"], slide=cfg["slide"], atoms=cfg["atoms"], omp_iter=cfg["k"], ksvd_iter=3, **kw)
    ks.set_picture(noisy)
    ks.extract()
    ks.init_dictionary()
    return clean, noisy, ks


OMP_CASES = CASES + [dict(CASES[0], atoms=200, seed=43), dict(CASES[0], seed=44, wave_per_patch=True)]
OMP_IDS = IDS + ["16x12_200atoms", "16x12_wave_per_patch"]


@pytest.mark.parametrize("cfg", OMP_CASES, ids=OMP_IDS)
def test_large_patch_omp(kl, orc, cfg, monkeypatch):
    """Both sizes run the 16-patch tile pursuit with the GEMM's correlations
    (whole Gram rows at 256 atoms, bounded reads at 200; 22 x 22 quarters of
    121 elements at one wavefront per SIMD); KSVD_OMP_WAVE_PER_PATCH=1 runs one
    wavefront per patch."""
    if cfg.get("wave_per_patch"):
        monkeypatch.setenv("KSVD_OMP_WAVE_PER_PATCH", "1")
    clean, noisy, ks = context(kl, cfg)
    ks.omp()
    g = ks.codes()
    Y, D = ks.patches(), ks.dictionary()
    ks.close()
    o32 = orc.omp(Y, D, cfg["k"], real="f32")[:3]
    o64 = orc.omp(Y, D.astype(np.float64), cfg["k"], real="f64")[:3]
    frac, err = compare_codes(g, o32)
    frac64, err64 = compare_codes(g, o64)
    print(f"\nOMP {cfg['patch']} (dim {Y.shape[1]}, {len(Y)} patches): supports equal {frac * 100:.4f}% "
          f"(coef err {err:.2e}); FP64 oracle {frac64 * 100:.4f}% (coef err {err64:.2e})")
    assert frac >= 0.999
    assert err64 <= 1e-4


@pytest.mark.parametrize("cfg", CASES + [dict(CASES[0], atoms=97, seed=45)], ids=IDS + ["16x12_97atoms"])
def test_gemm_tiles_same_bits(kl, cfg, monkeypatch):
    """The step-0 GEMM's two patch tiles (64 and 128 patches, forced with
    KSVD_GEMM_TP; 97 atoms takes the element-wise loads) give the same codes
    bit for bit, so the first-use choice between them never changes a
    result."""
    res = []
    for tp in ("64", "128"):
        monkeypatch.setenv("KSVD_GEMM_TP", tp)
        clean, noisy, ks = context(kl, cfg)
        ks.omp()
        res.append(ks.codes())
        ks.close()
    for a, b in zip(*res):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("path", ["default", "kl"])
@pytest.mark.parametrize("cfg", CASES, ids=IDS)
def test_large_patch_sweep(kl, orc, cfg, path, monkeypatch):
    """path "kl": KSVD_SWEEP_KL=1 keeps 22 x 22 patches on the three launches
    per atom (build, Gram, persistent power iteration) instead of the
    one-launch sweep with the columns in LDS."""
    if path == "kl":
        monkeypatch.setenv("KSVD_SWEEP_KL", "1")
    clean, noisy, ks = context(kl, cfg)
    Y, D0 = ks.patches(), ks.dictionary()
    io, co, no, _ = orc.omp(Y, D0, cfg["k"], real="f32")
    ks.set_codes(io, co, no)
    ks.update()
    Dg = ks.dictionary()
    cg = ks.codes()[1]
    ks.close()
    Do, coo, gap, _ = orc.update(Y, D0, io, co, no, cap=1024, real="f32", want_gap=True)
    good = (gap >= 0) & (gap < 0.9)
    Dgs, _ = signfix(Dg)
    Dos, _ = signfix(Do)
    aerr = np.max(np.abs(Dgs - Dos), axis=1)
    chk = sweep_self_check(Y, D0, io, co, no, Dg, cg, 1024)
    print(f"\nsweep {cfg['patch']}: {good.sum()} of {(gap >= 0).sum()} atoms with gap < 0.9, atom err "
          f"{aerr[good].max():.2e}; all atoms: Rayleigh deficit {chk[:, 0].max():.2e}, coef vs E^T d "
          f"{chk[:, 1].max():.2e}, |1 - norm| {chk[:, 2].max():.2e}")
    assert aerr[good].max() <= 1e-4
    assert chk[:, 0].max() <= 1e-6
    assert chk[:, 1].max() <= 1e-4
    assert chk[:, 2].max() <= 1e-5


@pytest.mark.parametrize("cfg", CASES, ids=IDS)
def test_large_patch_end_to_end(kl, orc, cfg):
    clean, noisy, ks = context(kl, cfg)
    ks.run()
    s2 = kl.noise_variance(noisy)
    img = ks.rebuild_picture(s2)
    ks.close()
    W, H = cfg["W"], cfg["H"]
    (pw, ph), (sw, sh) = cfg["patch"], cfg["slide"]
    fr = kl.image_to_frame(noisy)
    po = {}
    for real in ("f32", "f64"):
        Y = orc.extract(fr, W, H, pw, ph, sw, sh, real=real)
        D, idx, coef, cnt, _ = orc.ksvd(Y, orc.init_dict(Y, cfg["atoms"], real=real), cfg["k"], 3, real=real)
        out, *_ = orc.rebuild(idx, coef, cnt, D, fr, W, H, pw, ph, sw, sh, s2, real=real)
        po[real] = kl.psnr(clean, kl.frame_to_image(out.astype(np.float32), W, H))
    pg = kl.psnr(clean, img)
    print(f"\nend to end {cfg['patch']}: raw PSNR gpu {pg:.4f}, FP32 oracle {po['f32']:.4f}, FP64 oracle "
          f"{po['f64']:.4f} (noisy {kl.psnr(clean, noisy):.4f})")
    assert abs(pg - po["f32"]) <= 0.05
    assert pg > kl.psnr(clean, noisy) + 2


def test_large_patch_sweep_that_gives_up_is_redone(kl, monkeypatch):
    """22 x 22: a persistent power iteration that gives up (a spin limit of 0)
    is undone and the iterations run again with the one-workgroup power
    iteration: the result equals that path's, bit for bit, and the fallback is
    counted."""
    cfg = dict(CASES[1], W=128, H=96)

    def run():
        clean, noisy, ks = context(kl, cfg)
        ks.iterate(2)
        out = ks.dictionary(), ks.codes()[1], ks.timings()["sweep_fallbacks"]
        ks.close()
        return out

    monkeypatch.setenv("KSVD_SWEEP_TWO_KERNEL", "1")
    d2, c2, f2 = run()
    monkeypatch.setenv("KSVD_SWEEP_TWO_KERNEL", "0")
    monkeypatch.setenv("KSVD_SWEEP_SPIN", "0")
    d, c, f = run()
    print(f"\nlarge-patch sweep fallbacks: {f} (two-kernel run: {f2})")
    assert f2 == 0 and f == 1
    np.testing.assert_array_equal(d, d2)
    np.testing.assert_array_equal(c, c2)
