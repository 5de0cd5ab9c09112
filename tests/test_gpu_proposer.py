"""Learned support proposer of the BF16 form, through the C ABI.

  * the GPU's loss and weight gradients (hand-written BF16 MFMA GEMMs,
    split-K reductions) against the CPU oracle of the same network with the
    same BF16 rounding points (oracle/mlp_oracle.py, itself pinned by finite
    differences);
  * training on OMP's (patch, support) pairs at the 2048 x 2048 workload:
    training lowers the loss, and the trained network's proposals pass the
    1.05 x OMP-residual test more often than the untrained network's (what
    the training buys); every code, accepted proposal or OMP fallback, meets
    that bound;
  * the fallback list through the FP32 pursuit gives the full pass's codes
    bit for bit;
  * end to end, BF16 + learned: PSNR no more than 0.05 dB below the plain FP32
    form, fallback rate reported;
  * determinism: two runs give the same bits, on the stand-down path and on
    the trained-proposal path (forced by a zero acceptance gate).
"""
import numpy as np
import pytest

import mlp_oracle as M
from test_gpu_parity import CFG2, make

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def test_proposer_gradients_match_oracle(kl):
    ks = kl.KSvd(patch=(12, 12), slide=(3, 3), atoms=256, omp_iter=5, precision=kl.PREC_BF16)
    ks.proposer_init(11)
    kin, ko, n = ks.proposer_dims()
    W1, W2, W3 = ks.proposer_params()
    rng = np.random.default_rng(5)
    nb = 512
    X = np.zeros((nb, kin), np.float32)
    y = rng.standard_normal((nb, 144)) + 2.0
    yc = y - y.mean(axis=1, keepdims=True)
    X[:, :144] = yc / np.linalg.norm(yc, axis=1, keepdims=True)
    X[:, 144] = y.mean(axis=1) * 12.0 / np.linalg.norm(y, axis=1)
    X[:, 145] = 1.0
    T = np.zeros((nb, ko), np.float32)
    for b in range(nb):
        T[b, rng.choice(256, 5, replace=False)] = 1.0
    loss, (g1, g2, g3) = ks.proposer_loss_grad(X, T)
    lo, o1, o2, o3 = M.forward_backward(W1, W2, W3, X, T, 256, rounding="bf16")
    e = [rel(g1, o1), rel(g2, o2), rel(g3, o3)]
    print(f"\nloss gpu {loss:.6f} oracle {lo:.6f}; gradient rel err W1 {e[0]:.2e} W2 {e[1]:.2e} W3 {e[2]:.2e}")
    assert abs(loss - lo) <= 1e-4 * abs(lo)
    assert max(e) <= 2e-2
    # the bias columns carry the summed output gradients
    assert rel(g3[:256, 512], o3[:256, 512]) <= 2e-2
    # a plain float64 network is close too (BF16 rounding is the only difference)
    _, f1, f2, f3 = M.forward_backward(W1, W2, W3, X, T, 256, rounding=None)
    print(f"vs float64 network: W1 {rel(g1, f1):.2e} W2 {rel(g2, f2):.2e} W3 {rel(g3, f3):.2e}")
    ks.close()


def acceptance(ks, kl):
    ks.reset_timings()
    ks.proposer_omp()
    t = ks.timings()
    return 1.0 - t["fallback"] / max(t["proposed"], 1)


def test_training_gains_and_bound(kl):
    side = 2048
    clean, noisy = kl.synth_image(3, side, side, 20.0)
    ks = kl.KSvd(patch=(12, 12), slide=(3, 3), atoms=256, omp_iter=5, precision=kl.PREC_BF16)
    ks.set_picture(noisy)
    ks.extract()
    ks.init_dictionary()
    ks.omp()                                   # plain OMP: targets and residual bounds
    idx0, coef0, cnt0 = ks.codes()
    ks.proposer_init(3)
    acc_untrained = acceptance(ks, kl)
    ks.omp()                                   # OMP codes again as targets
    ks.proposer_init(3)
    ks.reset_timings()
    ks.proposer_train(1, lr=1e-3, seed=9)
    loss0 = ks.timings()["train_loss"]
    ks.proposer_init(3)
    ks.reset_timings()
    ks.proposer_train(96, lr=3e-3, seed=9)
    t = ks.timings()
    t_train, loss = t["train_ms"], t["train_loss"]
    acc_trained = acceptance(ks, kl)
    idx, coef, cnt = ks.codes()
    Y = ks.patches()
    D = ks.dictionary().astype(np.float64)
    rng = np.random.default_rng(0)
    worst = 0.0
    for p in rng.choice(len(cnt), 4000, replace=False):
        y = Y[p].astype(np.float64)
        r_new = np.linalg.norm(y - coef[p, :cnt[p]] @ D[idx[p, :cnt[p]]])
        r_omp = np.linalg.norm(y - coef0[p, :cnt0[p]] @ D[idx0[p, :cnt0[p]]])
        worst = max(worst, r_new / max(r_omp, 1e-12))
    print(f"\nloss after 1 step {loss0:.3f}, after 96 steps {loss:.3f} ({t_train:.1f} ms); proposals accepted: "
          f"untrained {acc_untrained * 100:.2f}%, trained {acc_trained * 100:.2f}%; worst residual / OMP residual "
          f"{worst:.4f}")
    assert loss < 0.25 * loss0
    assert acc_trained >= acc_untrained + 0.02
    assert worst <= 1.05 * (1 + 1e-4)
    ks.close()


def test_learned_fallback_list_matches_full_pursuit(kl):
    """The learned form codes with the FP32 pursuit, and its fallback runs
    that kernel over a list of patches.  With an untrained network almost
    every proposal fails the 1.05 x bound, so almost every patch goes through
    the list: its codes must be the full pass's bit for bit, and every
    accepted proposal must meet the bound."""
    cfg = dict(CFG2, W=256, H=256)
    clean, noisy, ks = make(kl, cfg, precision=kl.PREC_BF16, proposer=1)
    ks.extract()
    ks.init_dictionary()
    ks.omp()                                    # the FP32 pass: codes and residual bounds
    idx0, coef0, cnt0 = ks.codes()
    ks.proposer_init(5)
    ks.reset_timings()
    ks.proposer_omp()
    idx, coef, cnt = ks.codes()
    t = ks.timings()
    same = (cnt == cnt0) & np.all(idx == idx0, axis=1) & np.all(coef == coef0, axis=1)
    Y = ks.patches().astype(np.float64)
    D = ks.dictionary().astype(np.float64)
    worst = 0.0
    for p in np.nonzero(~same)[0]:
        r_new = np.linalg.norm(Y[p] - coef[p, :cnt[p]] @ D[idx[p, :cnt[p]]])
        r_omp = np.linalg.norm(Y[p] - coef0[p, :cnt0[p]] @ D[idx0[p, :cnt0[p]]])
        worst = max(worst, r_new / max(r_omp, 1e-12))
    print(f"\nfallback {t['fallback']} of {t['proposed']}; codes equal to the full pass on {same.mean() * 100:.2f}%; "
          f"worst accepted residual / OMP residual {worst:.4f}")
    assert t["fallback"] >= 0.9 * t["proposed"]
    assert same.sum() >= t["fallback"]
    assert worst <= 1.05 * (1 + 1e-4)
    ks.close()


def test_learned_run_deterministic(kl):
    """Two runs of the learned form on the same image (training, probe,
    stand-down decision, iterations, rebuild) give the same bits."""
    cfg = CFG2
    res = []
    for _ in range(2):
        clean, noisy, ks = make(kl, cfg, precision=kl.PREC_BF16, proposer=1)
        ks.extract()
        ks.init_dictionary()
        ks.run()
        t = ks.timings()
        res.append((ks.dictionary(), ks.codes(), ks.rebuild_picture(kl.noise_variance(noisy)), t["proposed"],
                    t["fallback"], t["train_loss"]))
        ks.close()
    a, b = res
    print(f"\nlearned run twice: proposed {a[3]} / {b[3]}, fallback {a[4]} / {b[4]}, loss {a[5]:.5f} / {b[5]:.5f}")
    assert a[3] > 0
    np.testing.assert_array_equal(a[0], b[0])
    for x, y in zip(a[1], b[1]):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(a[2], b[2])
    assert a[3:] == b[3:]


def test_learned_trained_path_deterministic(kl, monkeypatch):
    """The trained-proposal path (the proposer kept whatever it accepts, so
    every iteration after the probe proposes, codes its fallback list with
    OMP and trains on that list) twice on the same image gives the same bits:
    the fallback list is built in ascending patch order, whatever order the
    proposals ran in, so the training batches cut from it are the same."""
    monkeypatch.setenv("KSVD_PROPOSER_MIN_ACCEPT", "0")
    cfg = CFG2
    res = []
    for _ in range(2):
        clean, noisy, ks = make(kl, cfg, precision=kl.PREC_BF16, proposer=1)
        ks.extract()
        ks.init_dictionary()
        ks.run()
        t = ks.timings()
        res.append((ks.dictionary(), ks.codes(), ks.rebuild_picture(kl.noise_variance(noisy)), t["proposed"],
                    t["fallback"], t["train_loss"], t["proposer_gated"]))
        ks.close()
    a, b = res
    P = a[1][2].shape[0]
    print(f"\ntrained path twice: proposed {a[3]} / {b[3]} ({P} patches), fallback {a[4]} / {b[4]}, "
          f"loss {a[5]:.5f} / {b[5]:.5f}, gated {a[6]} / {b[6]}; PSNR raw {kl.psnr(clean, a[2]):.4f}")
    assert a[6] == 0
    # the probe, then a full pass per later iteration and the final code
    assert a[3] > 2 * P
    assert 0 < a[4] < a[3]
    np.testing.assert_array_equal(a[0], b[0])
    for x, y in zip(a[1], b[1]):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(a[2], b[2])
    assert a[3:] == b[3:]
    assert kl.psnr(clean, a[2]) > kl.psnr(clean, noisy) + 3


def test_bf16_learned_end_to_end_psnr(kl):
    cfg = CFG2
    out, tm = {}, {}
    for form, kw in (("plain", dict(precision=kl.PREC_FP32)),
                     ("learned", dict(precision=kl.PREC_BF16, proposer=1))):
        clean, noisy, ks = make(kl, cfg, **kw)
        ks.extract()
        ks.init_dictionary()
        ks.reset_timings()
        ks.run()
        tm[form] = ks.timings()
        out[form] = ks.rebuild_picture(kl.noise_variance(noisy))
        ks.close()
    r32, rl = kl.psnr(clean, out["plain"]), kl.psnr(clean, out["learned"])
    t = tm["learned"]
    fb = t["fallback"] / max(t["proposed"], 1)
    print(f"\nPSNR raw 512^2: plain FP32 {r32:.4f}, BF16 + learned {rl:.4f} (delta {rl - r32:+.4f} dB); "
          f"fallback rate {fb * 100:.2f}% of {t['proposed']} proposals; training {t['train_ms']:.2f} ms; "
          f"proposer stood down: {bool(t['proposer_gated'])}")
    # the bar bounds the quality lost against the plain form; accepted
    # proposals (residual within 1.05x of OMP's) change the codes and so the
    # dictionary's path, which can land above the plain form's PSNR as well
    assert rl >= r32 - 0.05
    assert t["proposed"] > 0
