"""K-SVD denoising benchmark on one MI355X.

A step is one K-SVD iteration (batched OMP over every patch + one
Gauss-Seidel sweep of rank-1 atom updates) over the synthetic image.  The
workload is the plain FP32 form on the 512 x 512 grayscale synthetic image,
sigma 25, 12 x 12 patches sliding 3 (27,889 patches of dim 144), 256 atoms,
OMP limit 5.  Inputs are resident in HBM before the timed region; the timed
region is exactly --steps iterations from the reference's dictionary init.

Prints one JSON line.  `roofline` is for the kernel pair that dominates the
step: one atom of the Gauss-Seidel dictionary sweep (k1_kernel builds the
restricted residual and partial Grams, k2_kernel reduces them and takes the
top eigenvector).  Its algorithmic bytes are those of the reference's update
for the atom's m users (re-read each user's patch and code, write its new
coefficient); its duration is the library's own HIP-event span of the
dictionary update on the library's stream, divided by the atoms swept.
`roofline_omp` is the same record for the OMP pursuit kernel (correlation on
MFMA FP32).  `cpu_baseline` times the CPU oracle (FP32, OpenMP) on one K-SVD
iteration of the same workload on this host.
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

# The three square workloads (seeded synthetic images; the 512^2 one is the
# default benchmark line, the others are selectable with --config).
WORKLOADS = {
    "128": dict(W=128, H=128, sigma=20.0, patch=(8, 8), slide=(2, 2), atoms=256, omp_iter=5, seed=1, iters=5),
    "512": dict(W=512, H=512, sigma=25.0, patch=(12, 12), slide=(3, 3), atoms=256, omp_iter=5, seed=2, iters=10),
    "2048": dict(W=2048, H=2048, sigma=20.0, patch=(12, 12), slide=(3, 3), atoms=256, omp_iter=5, seed=3, iters=10),
}
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0

# HBM bytes per dispatch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on
# the 512 x 512 workload (profiles/hbm_traffic_ksvd_512sq.md), scaled by what
# the same counters read for a device copy of known size (FETCH x 2, WRITE x 1).
MEASURED_TRAFFIC = {
    "512": {"omp": (9367.4 * 2 + 1198.5) * 1024,
            "sweep_atom": (485.4 * 2 + 578.9 + 196.1 * 2 + 0.9) * 1024},
}


def omp_flops_per_patch(n, K, k):
    """Algorithmic flops of one patch's pursuit (residual-correlation OMP):
    sum over steps s of 2nK (correlation) + 2n(s-1) (Gram column) + 2n (rhs)
    + 2ns (approximation) + 3n (residual and norm) + 3s^2 (Cholesky)."""
    return sum(2 * n * K + 2 * n * (s - 1) + 2 * n + 2 * n * s + 3 * n + 3 * s * s for s in range(1, k + 1))


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--config", choices=sorted(WORKLOADS), default="512",
                    help="image side of the workload (512: the benchmark line)")
    ap.add_argument("--steps", type=int, default=None, help="K-SVD iterations timed (default: the config's)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
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

    import pyksvd as K
    wl = WORKLOADS[args.config]
    if args.steps is None:
        args.steps = wl["iters"]
    clean, noisy = K.synth_image(wl["seed"], wl["W"], wl["H"], wl["sigma"])
    frame = K.image_to_frame(noisy)

    def make():
        ks = K.KSvd(patch=wl["patch"], slide=wl["slide"], atoms=wl["atoms"], omp_iter=wl["omp_iter"],
                    ksvd_iter=args.steps)
        ks.set_frame(frame, wl["W"], wl["H"])
        ks.extract()
        ks.init_dictionary()
        ks.synchronize()
        return ks

    # warmup on its own context: module load, first-touch, clocks
    if args.warmup > 0:
        ks = make()
        ks.iterate(args.warmup)
        ks.synchronize()
        ks.close()

    ks = make()
    ks.reset_timings()
    t0 = time.perf_counter()
    ks.iterate(args.steps)
    ks.synchronize()
    T = time.perf_counter() - t0
    tm = ks.timings()
    P = ks.P
    n = wl["patch"][0] * wl["patch"][1]
    kk = wl["omp_iter"]
    value = P * args.steps / T

    # whole image (gather -> iterations -> final OMP -> rebuild), for images/s
    s2 = K.noise_variance(noisy)
    ks2 = K.KSvd(patch=wl["patch"], slide=wl["slide"], atoms=wl["atoms"], omp_iter=kk, ksvd_iter=args.steps)
    ks2.set_frame(frame, wl["W"], wl["H"])
    ks2.synchronize()
    t1 = time.perf_counter()
    ks2.extract()
    ks2.init_dictionary()
    ks2.run()
    ks2.rebuild(s2, copy=False)
    ks2.synchronize()
    T_img = time.perf_counter() - t1
    out = ks2.rebuild_picture(s2)
    psnr_noisy = K.psnr(clean, noisy)
    psnr_out = K.psnr(clean, K.normalize(out))
    psnr_raw = K.psnr(clean, out)
    ks2.close()

    omp_avg_ms = tm["omp_kernel_ms"] / max(tm["omp_calls"], 1)
    flops_launch = P * omp_flops_per_patch(n, wl["atoms"], kk)
    achieved = flops_launch / (omp_avg_ms * 1e-3) / 1e12
    upd_avg_ms = tm["update_ms"] / max(tm["update_calls"], 1)
    ks.close()

    # Columns swept per iteration: replay the same (deterministic) iterations
    # step by step, outside any timed region, and count each sweep's users.
    ks3 = make()
    cols = 0
    for _ in range(args.steps):
        ks3.omp()
        cols += capped_users(*ks3.codes(), wl["atoms"], 1024)
        ks3.update()
    ks3.close()
    atoms_swept = args.steps * wl["atoms"]
    atom_ms = tm["update_ms"] / atoms_swept
    atom_bytes = cols / atoms_swept * update_bytes_per_user(n, kk)
    atom_gbs = atom_bytes / (atom_ms * 1e-3) / 1e9
    traffic = MEASURED_TRAFFIC.get(args.config, {})

    cpu = None
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
               "sample": f"one K-SVD iteration (OMP + dictionary sweep) of the same {wl['W']}x{wl['H']} workload, "
                         f"{P} patches, FP32 oracle with OpenMP, {Tc:.3f} s"}

    rec = {
        "metric": "K-SVD patch-iters/sec, 1 GPU: plain FP32 form",
        "value": round(value, 1),
        "unit": "patch-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(T * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded gratings + rectangles + Gaussian noise)",
        "config": {"workload": f"plain FP32 K-SVD, {wl['W']}x{wl['H']} sigma {wl['sigma']:g}, "
                               f"{wl['patch'][0]}x{wl['patch'][1]} patches sliding {wl['slide'][0]}, "
                               f"{wl['atoms']} atoms, OMP limit {kk}, user cap 1024",
                   "patches": P, "patch_dim": n, "atoms": wl["atoms"], "omp_iter": kk,
                   "ksvd_iter": args.steps},
        "roofline": {"bound": "hbm", "kernel": "k1_kernel + k2_kernel (one atom of the dictionary sweep)",
                     "achieved": round(atom_gbs, 3), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(atom_gbs / PEAK_HBM_GBS, 6),
                     "traffic": round(traffic["sweep_atom"]) if "sweep_atom" in traffic else None,
                     "avg_launch_ms": round(atom_ms, 5), "bytes_per_launch": round(atom_bytes),
                     "users_per_atom": round(cols / atoms_swept, 1)},
        "roofline_omp": {"bound": "mfma", "kernel": "omp_mfma_kernel", "achieved": round(achieved, 3),
                         "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(achieved / PEAK_FP32_TFLOPS, 5),
                         "traffic": round(traffic["omp"]) if "omp" in traffic else None,
                         "avg_launch_ms": round(omp_avg_ms, 5), "flops_per_launch": flops_launch},
        "phase_ms_per_step": {"omp": round(tm["omp_ms"] / args.steps, 4),
                              "update": round(upd_avg_ms, 4)},
        "power_iters_per_atom": round(tm["eig_iters"] / max(tm["update_calls"], 1) / wl["atoms"], 2),
        "images_per_s": round(1.0 / T_img, 3),
        "psnr": {"noisy": round(psnr_noisy, 4), "denoised_normalized": round(psnr_out, 4),
                 "denoised_raw": round(psnr_raw, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
