"""K-SVD denoising benchmark on one MI355X.

A step is one K-SVD iteration (sparse coding of every patch + one
Gauss-Seidel sweep of rank-1 atom updates) over one synthetic image.  The
default workload is the 2048 x 2048 grayscale synthetic image (seeded
gratings + rectangles, Gaussian sigma 20), 12 x 12 patches sliding 3
(461,041 patches of dim 144), 256 atoms, OMP limit 5, 10 iterations, run in
three forms on the same input:

  * plain FP32 (the drop-in for the reference's CUDA_K_GESVDJ path);
  * BF16 (correlation on BF16 MFMA, FP32 re-score, FP32 least squares);
  * BF16 + learned (the BF16 form whose sparse coding is proposed by the
    trained perceptron, OMP as fallback; training is inside the timed
    iterations).

`value` is the BF16 + learned form's patch-iterations per second (patches x
iterations / wall time of the iterations, inputs resident in HBM);
`plain_fp32` and `bf16` give the other forms' rates, `speedup_vs_plain` the
ratio, `delta_psnr_db` the denoised PSNR difference to the plain form (one
seed, the workload's own 10 iterations; the statistics over seeds are in
tests/test_gpu_bf16.py and tests/test_gpu_e2e.py).  `headline_form` and
`dtype` name the arithmetic the headline's timed iterations ran.
`fp64_oracle` compares the plain form's first OMP with the FP64 oracle on a
sample of patches.

`roofline` is for the kernel that dominates the step: the persistent
dictionary sweep (et_sweep_kernel, the eight-team sweep, one launch per
iteration).  Its
algorithmic bytes are those of the reference's update for the capped users
of every atom (re-read each user's patch and code, write its new
coefficient); its duration is the library's HIP-event span of the update on
its stream.  `roofline_omp` is the same record for the plain form's OMP
kernel (FP32 MFMA correlation in the first step, Gram updates after it),
priced by the flops that kernel performs.  `cpu_baseline` times the CPU oracle (FP32, OpenMP) on
one K-SVD iteration of the same workload on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "cuda-k-svd-denoising_amd")
ORACLE = os.path.join(ROOT, "oracle")
sys.path.insert(0, PKG)

# The three square workloads (seeded synthetic images); 2048 is the
# benchmark line (the configuration the headline metric is quoted on), the
# others are selectable with --config.
WORKLOADS = {
    "128": dict(W=128, H=128, sigma=20.0, patch=(8, 8), slide=(2, 2), atoms=256, omp_iter=5, seed=1, iters=5),
    "512": dict(W=512, H=512, sigma=25.0, patch=(12, 12), slide=(3, 3), atoms=256, omp_iter=5, seed=2, iters=10),
    "2048": dict(W=2048, H=2048, sigma=20.0, patch=(12, 12), slide=(3, 3), atoms=256, omp_iter=5, seed=3, iters=10),
}
PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0

# HBM bytes per dispatch of the dominant kernels, measured with rocprofv3
# FETCH_SIZE / WRITE_SIZE (KiB) and scaled by what the same counters read for
# a device copy of known size (FETCH x 2, WRITE x 1):
# profiles/hbm_traffic_ksvd_2048sq.md.
MEASURED_TRAFFIC = {
    "2048": {"sweep": (112745.7 * 2 + 42044.3) * 1024, "omp": (131621.6 * 2 + 20223.8) * 1024},
}

def omp_flops_per_patch(n, K, k):
    """Algorithmic flops of one patch's pursuit (residual-correlation OMP):
    sum over steps s of 2nK (correlation) + 2n(s-1) (Gram column) + 2n (rhs)
    + 2ns (approximation) + 3n (residual and norm) + 3s^2 (Cholesky)."""
    return sum(2 * n * K + 2 * n * (s - 1) + 2 * n + 2 * n * s + 3 * n + 3 * s * s for s in range(1, k + 1))


def batch_omp_flops_per_patch(n, K, k):
    """Flops of one patch's pursuit in the FP32 kernel (Batch-OMP): |y|^2
    once (2n); the first step correlates D^T y (2nK); step s > 1 updates it
    from the dictionary Gram, c = D^T y - G[:, S] x (2K(s-1)); every step
    forms b_j = d_j . y (2n), the Cholesky row and solves (3s^2) and the
    residual norm |y|^2 - b.x (2s).  The support Gram entries are read from
    G; the explicit residual is formed only for near-zero norms."""
    return 2 * n + sum((2 * n * K if s == 1 else 2 * K * (s - 1)) + 2 * n + 3 * s * s + 2 * s
                       for s in range(1, k + 1))


def update_flops_per_patch(n, k):
    return k * (2 * n * (k - 1) + 4 * n)


def bytes_per_patch(n, k):
    return 4 * n + 8 * k + k * (4 * n + 8 * k + 4)


def update_bytes_per_user(n, k):
    """Algorithmic HBM bytes of one (atom, user) column of the rank-1 update:
    re-read the user's patch (4n) and code (8k), write its new coefficient."""
    return 4 * n + 8 * k + 4


def capped_users(idx, coef, cnt, K, cap):
    """Sum over atoms of min(c_k, cap), c_k = patches whose code for atom k is
    non-zero: the columns one dictionary sweep processes."""
    mask = (np.arange(idx.shape[1])[None, :] < cnt[:, None]) & (coef != 0)
    c = np.bincount(idx[mask], minlength=K)
    return int(np.minimum(c, cap).sum())


def calibrate_copy(mib):
    """Device-to-device copies of `mib` MiB (the runtime's copy kernel): the
    known-size reference that scales rocprofv3's FETCH_SIZE / WRITE_SIZE."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"calibration": "d2d copy", "bytes_per_copy": nbytes, "copies": 3}))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def code_agreement(a, b):
    """(fraction of patches with identical supports, max coefficient error
    relative to the patch's coefficient norm on those) of two compact codes."""
    def srt(idx, coef, cnt):
        live = np.arange(idx.shape[1])[None, :] < cnt[:, None]
        o = np.argsort(np.where(live, idx, np.iinfo(np.int32).max), axis=1, kind="stable")
        return (np.take_along_axis(np.where(live, idx, -1), o, 1),
                np.take_along_axis(np.where(live, coef, 0).astype(np.float64), o, 1))
    ia, ca = srt(*a)
    ib, cb = srt(*b)
    same = np.all(ia == ib, axis=1) & (a[2] == b[2])
    err = np.max(np.abs(ca - cb), axis=1) / np.maximum(np.linalg.norm(cb, axis=1), 1e-30)
    return float(same.mean()), float(err[same].max() if same.any() else np.inf)


def run_form(K, wl, frame, steps, warmup, **kw):
    """Time `steps` K-SVD iterations of one form on a fresh context (after
    `warmup` untimed iterations on another context); returns the rate, the
    library's timings and the context's dims."""
    def make():
        ks = K.KSvd(patch=wl["patch"], slide=wl["slide"], atoms=wl["atoms"], omp_iter=wl["omp_iter"],
                    ksvd_iter=steps, **kw)
        ks.set_frame(frame, wl["W"], wl["H"])
        ks.extract()
        ks.init_dictionary()
        ks.synchronize()
        return ks

    if warmup > 0:
        ks = make()
        ks.iterate(warmup)
        ks.synchronize()
        ks.close()
    ks = make()
    ks.reset_timings()
    t0 = time.perf_counter()
    ks.iterate(steps)
    ks.synchronize()
    T = time.perf_counter() - t0
    tm = ks.timings()
    P = ks.P
    ks.close()
    return P * steps / T, T, tm, P


def denoise(K, wl, frame, clean, steps, **kw):
    """Whole image (gather -> iterations -> final code -> rebuild): images/s and
    PSNR of the denoised picture (raw and min-max normalised)."""
    s2 = K.noise_variance(K.frame_to_image(frame, wl["W"], wl["H"]))
    ks = K.KSvd(patch=wl["patch"], slide=wl["slide"], atoms=wl["atoms"], omp_iter=wl["omp_iter"], ksvd_iter=steps,
                **kw)
    ks.set_frame(frame, wl["W"], wl["H"])
    ks.synchronize()
    t1 = time.perf_counter()
    ks.extract()
    ks.init_dictionary()
    ks.run()
    ks.rebuild(s2, copy=False)
    ks.synchronize()
    T_img = time.perf_counter() - t1
    out = ks.rebuild_picture(s2)
    ks.close()
    return 1.0 / T_img, K.psnr(clean, out), K.psnr(clean, K.normalize(out))


def batch_specs(n=32, seed=7):
    """The uneven batch: n seeded pictures sized 256 x 256 to 2048 x 1536 at
    mixed aspect ratios; the even ones grayscale with Gaussian sigma in
    [10, 50], the odd ones colour with multiplicative Gamma speckle (L = 4,
    denoised in the log domain); 16 x 12 patches, every fourth file with a
    22 x 22 override; sliding 3, 256 atoms, 10 K-SVD iterations."""
    sizes = [(256, 256), (384, 256), (512, 384), (640, 480), (768, 512), (1024, 768), (1280, 720), (1536, 1024),
             (2048, 1536), (1024, 1024), (512, 512), (2048, 1024), (800, 600), (1920, 1080), (320, 480), (1536, 1536)]
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        W, H = sizes[i % len(sizes)]
        out.append(dict(W=W, H=H, rgb=i % 2 == 1, sigma=float(rng.uniform(10, 50)),
                        patch=(22, 22) if i % 4 == 3 else (16, 12), seed=1000 + i))
    return out


def batch_inputs(K, specs):
    """Frame buffers and noise variances as the front end prepares them."""
    rng = np.random.default_rng(11)
    prepared = []
    for sp in specs:
        W, H = sp["W"], sp["H"]
        if sp["rgb"]:
            clean = np.stack([K.synth_image(sp["seed"] + 7919 * c, W, H, 0.0)[0] for c in range(3)], axis=-1) + 10.0
            noisy = (clean * rng.gamma(4.0, 0.25, size=clean.shape)).astype(np.float32)
            planes = K.rgb_to_ycbcr(noisy)
            frame = K.speckle_log(K.image_to_frame(planes[0]))
        else:
            _, noisy = K.synth_image(sp["seed"], W, H, sp["sigma"])
            planes = noisy[None]
            frame = K.image_to_frame(noisy)
        prepared.append((frame, np.float32(K.noise_variance_channels(planes))))
    return prepared


def run_batch(K, specs, inputs, iters, group=8, resident=None):
    """Every picture through the whole path (gather, iterations, final code,
    rebuild, speckle's exp2).  Pictures of one patch size go in groups of up
    to `group` (one per XCD): each picture on its own context, their
    iterations together (ksvd_run_batch: each picture's pursuit on its own
    stream, one dictionary sweep launch for the group where a persistent
    one-XCD sweep applies, one picture after another otherwise).  Contexts
    are reused from group to group.

    resident: each picture's frame already in HBM (device pointers); the
    output stays there too (ksvd_rebuild_speckle for the speckle pictures).
    None: frames copied in from host memory, outputs copied back and
    speckle's exp2 on the host threads (ksvd_speckle_exp), as the front end
    does.  Returns (seconds, patch-iterations)."""
    pools = {}
    pi = 0
    order = {}
    for i, sp in enumerate(specs):
        order.setdefault(sp["patch"], []).append(i)
    t0 = time.perf_counter()
    for key, idxs in order.items():
        pool = pools.setdefault(key, [])
        for g0 in range(0, len(idxs), group):
            members = idxs[g0:g0 + group]
            while len(pool) < len(members):
                pool.append(K.KSvd(patch=key, slide=(3, 3), atoms=256, omp_iter=5, ksvd_iter=iters))
            ctxs = pool[:len(members)]
            for ks, i in zip(ctxs, members):
                sp, (frame, s2) = specs[i], inputs[i]
                if resident is None:
                    ks.set_frame(frame, sp["W"], sp["H"])
                else:
                    ks.set_frame_device(resident[i], sp["W"], sp["H"])
                ks.extract()
                ks.init_dictionary()
            K.run_batch(ctxs)
            for ks, i in zip(ctxs, members):
                if resident is None:
                    out = ks.rebuild(inputs[i][1])
                    if specs[i]["rgb"]:
                        K.speckle_exp(out)
                else:
                    ks.rebuild(inputs[i][1], copy=False, speckle=specs[i]["rgb"])
                pi += ks.P * iters
    for pool in pools.values():
        for ks in pool:
            ks.synchronize()
    T = time.perf_counter() - t0
    for pool in pools.values():
        for ks in pool:
            ks.close()
    return T, pi


def main_batch(args):
    import torch  # device memory for the resident frames
    import pyksvd as K
    specs = batch_specs()
    inputs = batch_inputs(K, specs)
    iters = args.steps or 10
    dev = [torch.from_numpy(np.ascontiguousarray(f, dtype=np.float32).ravel()).to("cuda:0") for f, _ in inputs]
    torch.cuda.synchronize()
    ptrs = [t.data_ptr() for t in dev]
    run_batch(K, specs[:2], inputs[:2], 1, resident=ptrs[:2])  # warm-up (libraries, first launches)
    T, pi = run_batch(K, specs, inputs, iters, resident=ptrs)
    Th, _ = run_batch(K, specs, inputs, iters)
    n16 = sum(1 for s in specs if s["patch"] == (16, 12))
    rec = {"metric": "images/s, uneven batch (plain FP32 form)", "value": round(len(specs) / T, 3), "unit": "images/s",
           "n_gpus": 1, "steps": iters, "warmup": 1, "ms_per_image": round(T * 1e3 / len(specs), 2),
           "higher_is_better": True, "dtype": "f32", "data": "synthetic (seeded gratings + rectangles)",
           "config": {"workload": f"{len(specs)} pictures 256x256 to 2048x1536, half colour with Gamma speckle (L=4), "
                                  f"{n16} at 16x12 patches and {len(specs) - n16} at 22x22, sliding 3, 256 atoms, "
                                  f"{iters} K-SVD iterations", "pictures_per_sweep_launch": 8,
                      "frames": "resident in HBM before the timed region; outputs left in HBM (speckle exp2 on the device)"},
           "patch_iters_per_s": round(pi / T, 1),
           "host_inclusive_images_per_s": round(len(specs) / Th, 3),
           "host_inclusive": "frames copied in from host memory, outputs copied back, speckle exp2 on the host threads"}
    print(json.dumps(rec))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--config", choices=sorted(WORKLOADS) + ["batch"], default="2048",
                    help="image side of the workload (2048: the benchmark line; batch: the uneven batch, images/s)")
    ap.add_argument("--steps", type=int, default=None, help="K-SVD iterations timed (default: the config's)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--plain-only", action="store_true", help="time the plain FP32 form only (profiling runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run device copies of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the multi-GPU sharding is a "
              "design note, not a measured path", file=sys.stderr)
        sys.exit(2)

    if args.config == "batch":
        main_batch(args)
        return
    import pyksvd as K
    wl = WORKLOADS[args.config]
    if args.steps is None:
        args.steps = wl["iters"]
    clean, noisy = K.synth_image(wl["seed"], wl["W"], wl["H"], wl["sigma"])
    frame = K.image_to_frame(noisy)
    n = wl["patch"][0] * wl["patch"][1]
    kk = wl["omp_iter"]

    forms = {"plain": dict(precision=K.PREC_FP32)}
    if not args.plain_only:
        forms["bf16"] = dict(precision=K.PREC_BF16)
        forms["learned"] = dict(precision=K.PREC_BF16, proposer=1)
    res = {}
    for name, kw in forms.items():
        rate, T, tm, P = run_form(K, wl, frame, args.steps, args.warmup, **kw)
        # images/s and PSNR at the workload's own iteration count, whatever --steps is
        ips, praw, pnorm = denoise(K, wl, frame, clean, wl["iters"], **kw)
        res[name] = dict(rate=rate, T=T, tm=tm, ips=ips, psnr_raw=praw, psnr_norm=pnorm)
        print(json.dumps({"form": name, "patch_iters_per_s": round(rate, 1), "ms_per_step": round(T * 1e3 / args.steps, 4),
                          "omp_ms_per_step": round(tm["omp_ms"] / args.steps, 4),
                          "update_ms_per_step": round(tm["update_ms"] / args.steps, 4),
                          "sweep_fallbacks": tm["sweep_fallbacks"],
                          "psnr_raw": round(praw, 4)}), file=sys.stderr, flush=True)
    plain = res["plain"]
    head = res.get("learned", plain)
    tm = plain["tm"]

    # the plain form's OMP kernel on FP32 MFMA
    omp_avg_ms = tm["omp_kernel_ms"] / max(tm["omp_calls"], 1)
    flops_launch = P * batch_omp_flops_per_patch(n, wl["atoms"], kk)
    achieved = flops_launch / (omp_avg_ms * 1e-3) / 1e12

    # Columns swept per iteration: replay the plain form's (deterministic)
    # iterations step by step, outside any timed region, and count each
    # sweep's capped users.
    ks3 = K.KSvd(patch=wl["patch"], slide=wl["slide"], atoms=wl["atoms"], omp_iter=kk, ksvd_iter=args.steps)
    ks3.set_frame(frame, wl["W"], wl["H"])
    ks3.extract()
    ks3.init_dictionary()
    cols = 0
    for _ in range(args.steps):
        ks3.omp()
        cols += capped_users(*ks3.codes(), wl["atoms"], 1024)
        ks3.update()
    ks3.close()
    sweep_ms = tm["update_ms"] / max(tm["update_calls"], 1)
    sweep_bytes = cols / args.steps * update_bytes_per_user(n, kk)
    sweep_gbs = sweep_bytes / (sweep_ms * 1e-3) / 1e9
    traffic = MEASURED_TRAFFIC.get(args.config, {})

    cpu = None
    fp64 = None
    if not args.no_cpu:
        sys.path.insert(0, ORACLE)
        import pyoracle as O
        Y = O.extract(frame, wl["W"], wl["H"], *wl["patch"], *wl["slide"])
        D0 = O.init_dict(Y, wl["atoms"])
        O.ksvd(Y, D0, kk, 1, final_omp=False)  # untimed warm run (threads, pages)
        c0 = time.perf_counter()
        O.ksvd(Y, D0, kk, 1, final_omp=False)
        Tc = time.perf_counter() - c0
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": round(P / Tc, 1), "unit": "patch-iters/s", "cores": cores, "kind": "port",
               "model": cpu_model(),
               "sample": f"one K-SVD iteration (OMP + dictionary sweep) of the same {wl['W']}x{wl['H']} workload, "
                         f"{P} patches, FP32 oracle with OpenMP, {Tc:.3f} s"}
        # the plain form's first OMP against the FP64 oracle on a sample of patches
        ks4 = K.KSvd(patch=wl["patch"], slide=wl["slide"], atoms=wl["atoms"], omp_iter=kk)
        ks4.set_frame(frame, wl["W"], wl["H"])
        ks4.extract()
        ks4.init_dictionary()
        ks4.omp()
        gi, gc, gn = ks4.codes()
        Dg = ks4.dictionary()
        ks4.close()
        sel = np.linspace(0, P - 1, min(P, 20000)).astype(np.int64)
        oi, oc, on, _ = O.omp(Y[sel].astype(np.float64), Dg.astype(np.float64), kk, real="f64")
        agree, cerr = code_agreement((gi[sel], gc[sel], gn[sel]), (oi, oc, on))
        fp64 = {"first_omp_supports_equal": round(agree, 6), "first_omp_coef_rel_err_max": float(f"{cerr:.3g}"),
                "sample": f"{len(sel)} patches evenly spaced over the {P}",
                "end_to_end": "raw PSNR against the FP64 oracle over seeds: tests/test_gpu_e2e.py::test_end_to_end_psnr"}

    lt = head["tm"]
    # the arithmetic the headline's timed iterations ran: with the proposer
    # stood down after its probe, the learned form codes with the FP32 pursuit
    if "learned" not in res:
        head_dtype, head_form = "f32", "plain FP32"
    elif lt["proposer_gated"]:
        head_dtype, head_form = "f32", ("BF16 + learned, proposer stood down after its probe: FP32 pursuit, "
                                        "BF16 training and probe")
    else:
        head_dtype, head_form = "bf16", "BF16 + learned"
    rec = {
        "metric": "K-SVD patch-iters/sec, 1 GPU: BF16+learned vs plain FP32 at \u0394PSNR\u22640.05dB",
        "value": round(head["rate"], 1),
        "unit": "patch-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["T"] * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": head_dtype,
        "data": "synthetic (seeded gratings + rectangles + Gaussian noise)",
        "config": {"workload": f"K-SVD denoising, {wl['W']}x{wl['H']} sigma {wl['sigma']:g}, "
                               f"{wl['patch'][0]}x{wl['patch'][1]} patches sliding {wl['slide'][0]}, "
                               f"{wl['atoms']} atoms, OMP limit {kk}, user cap 1024",
                   "patches": P, "patch_dim": n, "atoms": wl["atoms"], "omp_iter": kk, "ksvd_iter": args.steps},
        "plain_fp32": {"value": round(plain["rate"], 1), "ms_per_step": round(plain["T"] * 1e3 / args.steps, 4),
                       "images_per_s": round(plain["ips"], 3), "psnr_raw": round(plain["psnr_raw"], 4),
                       "psnr_normalized": round(plain["psnr_norm"], 4)},
        "roofline": {"bound": "hbm", "kernel": "dictionary update of the plain form: et_sweep_kernel (eight-team persistent "
                               "sweep) + its users / column-table / cap-zeroing kernels",
                     "achieved": round(sweep_gbs, 3), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(sweep_gbs / PEAK_HBM_GBS, 6),
                     "traffic": round(traffic["sweep"]) if "sweep" in traffic else None,
                     "avg_launch_ms": round(sweep_ms, 5), "bytes_per_launch": round(sweep_bytes),
                     "users_per_atom": round(cols / args.steps / wl["atoms"], 1)},
        "roofline_omp": {"bound": "mfma", "kernel": "omp_mfma_kernel (plain form)", "achieved": round(achieved, 3),
                         "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 5),
                         "traffic": round(traffic["omp"]) if "omp" in traffic else None,
                         "avg_launch_ms": round(omp_avg_ms, 5), "flops_per_launch": flops_launch},
        "power_iters_per_atom": round(tm["eig_iters"] / max(tm["update_calls"], 1) / wl["atoms"], 2),
        "psnr_noisy": round(K.psnr(clean, noisy), 4),
        "headline_form": head_form,
        "psnr_iterations": wl["iters"],
        "fp64_oracle": fp64,
        "cpu_baseline": cpu,
    }
    if "learned" in res:
        b = res["bf16"]
        bt = b["tm"]
        bomp_ms = bt["omp_kernel_ms"] / max(bt["omp_calls"], 1)
        bflops = P * omp_flops_per_patch(n, wl["atoms"], kk)
        bach = bflops / (bomp_ms * 1e-3) / 1e12
        rec["speedup_vs_plain"] = round(head["rate"] / plain["rate"], 4)
        rec["delta_psnr_db"] = round(head["psnr_raw"] - plain["psnr_raw"], 4)
        rec["images_per_s"] = round(head["ips"], 3)
        rec["learned"] = {"fallback_rate": round(lt["fallback"] / max(lt["proposed"], 1), 4),
                          "proposed": lt["proposed"], "train_ms": round(lt["train_ms"], 3),
                          "train_loss": round(lt["train_loss"], 4), "stood_down": bool(lt["proposer_gated"]),
                          "psnr_raw": round(head["psnr_raw"], 4)}
        rec["bf16"] = {"value": round(b["rate"], 1), "ms_per_step": round(b["T"] * 1e3 / args.steps, 4),
                       "speedup_vs_plain": round(b["rate"] / plain["rate"], 4),
                       "delta_psnr_db": round(b["psnr_raw"] - plain["psnr_raw"], 4),
                       "omp_kernel": {"kernel": "omp_bf16_kernel", "avg_launch_ms": round(bomp_ms, 5),
                                      "achieved": round(bach, 3), "unit": "TFLOP/s (algorithmic)",
                                      "peak_bf16": PEAK_BF16_TFLOPS, "frac": round(bach / PEAK_BF16_TFLOPS, 5),
                                      "rescored_steps": bt["omp_rescored"], "f64_refactors": bt["omp_f64_refactors"]}}
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
