"""The `denoising config.json` front end: config schema and error handling on
the CPU, a full run on the GPU.  The schema is the reference's
(batch_denoiser.cpp:36-117): thirteen required globalParams keys, per-file
overrides, output <outputFolder>/<pw>_<ph>_<name>."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "cuda-k-svd-denoising_amd", "denoising")

GLOBAL = {"patchWidthDim": 8, "patchHeightDim": 8, "slidingWidth": 2, "slidingHeight": 2, "atoms": 256,
          "ompIter": 5, "ksvdIter": 2, "B&W": True, "speckle": False, "type": "HIP_K_SVD",
          "minOmpIterBatch": 0, "subImageWidthDim": 0, "subImageHeightDim": 0}


@pytest.fixture(scope="module")
def denoising_bin(kl):
    if not os.path.exists(BIN):
        subprocess.check_call(["make", "-s", "-C", os.path.dirname(BIN), "denoising"])
    return BIN


def write_pgm(path, img):
    img = np.clip(np.rint(img), 0, 255).astype(np.uint8)
    with open(path, "wb") as f:
        f.write(b"P5\n%d %d\n255\n" % (img.shape[1], img.shape[0]))
        f.write(img.tobytes())


def write_pfm(path, img):
    img = np.asarray(img, np.float32)
    with open(path, "wb") as f:
        f.write(b"Pf\n%d %d\n-1.0\n" % (img.shape[1], img.shape[0]))
        f.write(img[::-1].astype("<f4").tobytes())


def read_pgm(path):
    with open(path, "rb") as f:
        data = f.read()
    parts = data.split(maxsplit=4)
    assert parts[0] == b"P5"
    w, h = int(parts[1]), int(parts[2])
    return np.frombuffer(parts[4][: w * h], np.uint8).reshape(h, w)


def run(binary, cfg, tmp_path):
    p = tmp_path / "config.json"
    p.write_text(json.dumps(cfg) if isinstance(cfg, dict) else cfg)
    return subprocess.run([binary, str(p)], capture_output=True, text=True, timeout=600)


def test_missing_config_file(denoising_bin, tmp_path):
    r = subprocess.run([denoising_bin, str(tmp_path / "nope.json")], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open" in r.stderr


def test_malformed_json(denoising_bin, tmp_path):
    r = run(denoising_bin, '{"inputFolder": "x", ', tmp_path)
    assert r.returncode == 1 and "config" in r.stderr


@pytest.mark.parametrize("key", sorted(GLOBAL))
def test_every_global_key_is_required(denoising_bin, tmp_path, key):
    g = dict(GLOBAL)
    del g[key]
    cfg = {"inputFolder": str(tmp_path), "outputFolder": str(tmp_path), "globalParams": g,
           "files": [{"name": "a.pgm"}]}
    r = run(denoising_bin, cfg, tmp_path)
    assert r.returncode == 1
    assert f'missing required key "{key}"' in r.stderr


def test_file_name_is_required(denoising_bin, tmp_path):
    cfg = {"inputFolder": str(tmp_path), "outputFolder": str(tmp_path), "globalParams": GLOBAL,
           "files": [{"ref": "a.pgm"}]}
    r = run(denoising_bin, cfg, tmp_path)
    assert r.returncode == 1 and '"name"' in r.stderr


def test_unknown_type_is_an_error(denoising_bin, tmp_path):
    # the reference silently falls back to CUDA_K_GESVD (map operator[]); here it is reported
    cfg = {"inputFolder": str(tmp_path), "outputFolder": str(tmp_path), "globalParams": dict(GLOBAL, type="FOO"),
           "files": [{"name": "a.pgm"}]}
    r = run(denoising_bin, cfg, tmp_path)
    assert r.returncode == 1 and 'unknown denoiser type "FOO"' in r.stderr


def test_missing_image_reports_load_failure(denoising_bin, tmp_path):
    cfg = {"inputFolder": str(tmp_path), "outputFolder": str(tmp_path), "globalParams": GLOBAL,
           "files": [{"name": "absent.pgm"}]}
    r = run(denoising_bin, cfg, tmp_path)
    assert r.returncode == 2 and "Image Loading Failed" in r.stdout


def tiles_cut(frame, fm, fn, th, tw):
    """Tiles of a frame (fn rows of length fm), th rows x tw columns each,
    rows of tiles outer (createPatches with patch = slide = the sub-image
    size, cuda_svd_denoiser.cu:130-140,231-240)."""
    f = np.asarray(frame, np.float32).reshape(fn, fm)
    return [f[ti * th:(ti + 1) * th, tj * tw:(tj + 1) * tw].copy()
            for ti in range(fn // th) for tj in range(fm // tw)]


def tiles_join_reference(tiles, fm, fn, th, tw):
    """createImageFromSubImages (cuda_svd_denoiser.cu:503-528) written out as
    the reference's loops: stride = patchIdx * tile size, stride2 = tiles
    per row * tw, rows of the output of length stride2."""
    img = np.zeros(fm * fn, np.float32)
    flat = np.concatenate([t.ravel() for t in tiles])
    per_row = fm // tw
    tsz = th * tw
    for patch_idx in range(0, len(tiles), per_row):
        stride = patch_idx * tsz
        for i in range(per_row):
            stride2 = per_row * tw
            for j in range(th):
                start = tsz * (i + patch_idx) + j * tw
                o = stride + stride2 * j + i * tw
                img[o:o + tw] = flat[start:start + tw]
    return img


def test_tile_join_is_the_reference_loop():
    """The block placement the front end uses equals the reference's loop
    where the frame's sides are multiples of the tile's (the case its loop
    lays out rows of the frame's own length)."""
    rng = np.random.default_rng(2)
    fm, fn, th, tw = 48, 36, 12, 16
    tiles = [rng.standard_normal((th, tw)).astype(np.float32) for _ in range((fn // th) * (fm // tw))]
    out = np.zeros((fn, fm), np.float32)
    k = 0
    for ti in range(fn // th):
        for tj in range(fm // tw):
            out[ti * th:(ti + 1) * th, tj * tw:(tj + 1) * tw] = tiles[k]
            k += 1
    np.testing.assert_array_equal(out.ravel(), tiles_join_reference(tiles, fm, fn, th, tw))
    frame = rng.standard_normal(fm * fn).astype(np.float32)
    back = tiles_join_reference(tiles_cut(frame, fm, fn, th, tw), fm, fn, th, tw)
    np.testing.assert_array_equal(back, frame)


@pytest.mark.gpu
def test_denoising_cli_end_to_end(denoising_bin, tmp_path, kl):
    clean, noisy = kl.synth_image(9, 128, 128, 20.0)
    write_pfm(tmp_path / "noisy.pfm", noisy)
    write_pfm(tmp_path / "clean.pfm", clean)
    cfg = {"inputFolder": str(tmp_path), "outputFolder": str(tmp_path), "globalParams": GLOBAL,
           "files": [{"name": "noisy.pfm", "ref": "clean.pfm", "ksvdIter": 3},
                     {"name": "noisy.pfm", "ref": "clean.pfm", "type": "CUDA_K_GESVDJ", "patchWidthDim": 6,
                      "patchHeightDim": 6}]}
    r = run(denoising_bin, cfg, tmp_path)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("PSNR after:") == 2
    before = [float(l.split(":")[1]) for l in r.stdout.splitlines() if "PSNR before:" in l]
    after = [float(l.split(":")[1]) for l in r.stdout.splitlines() if "PSNR after:" in l]
    assert abs(before[0] - kl.psnr(clean, noisy)) < 1e-3  # printed with 6 significant digits
    assert all(a > 0 for a in after)
    assert os.path.exists(tmp_path / "8_8_noisy.pfm") and os.path.exists(tmp_path / "6_6_noisy.pfm")


def read_pfm(path):
    with open(path, "rb") as f:
        data = f.read()
    parts = data.split(maxsplit=4)
    assert parts[0] in (b"Pf", b"PF")
    C = 3 if parts[0] == b"PF" else 1
    w, h, scale = int(parts[1]), int(parts[2]), float(parts[3])
    a = np.frombuffer(parts[4][: w * h * C * 4], "<f4" if scale < 0 else ">f4").reshape(h, w, C)[::-1]
    return a[..., 0] if C == 1 else a


def write_pfm_rgb(path, rgb):
    rgb = np.asarray(rgb, np.float32)
    with open(path, "wb") as f:
        f.write(b"PF\n%d %d\n-1.0\n" % (rgb.shape[1], rgb.shape[0]))
        f.write(rgb[::-1].astype("<f4").tobytes())


def printed(stdout, key):
    return [float(l.split(":")[1]) for l in stdout.splitlines() if key in l]


@pytest.mark.gpu
def test_cli_on_reference_pictures(denoising_bin, tmp_path, kl):
    """`denoising config.json` with the reference's settings (8x8 patches,
    sliding 2, 256 atoms, CUDA_K_GESVDJ) on its three test pictures at noise
    variance 0.05 with the published iteration counts: the printed PSNRs are
    those of the noisy file and of the min-max stretched output that the
    C ABI gives for the same picture."""
    import json as js
    gold = os.path.join(ROOT, "tests", "golden")
    meta = js.load(open(os.path.join(gold, "reference_psnr.json")))
    arr = np.load(os.path.join(gold, "reference_images.npz"))
    files = []
    for c in meta["cases"]:
        pic = c["picture"]
        write_pgm(tmp_path / f"{pic}.pgm", arr[pic + "_noisy"])
        write_pgm(tmp_path / f"{pic}_ref.pgm", arr[pic + "_ref"])
        files.append({"name": f"{pic}.pgm", "ref": f"{pic}_ref.pgm", "ksvdIter": c["ksvd_iter"]})
    cfg = {"inputFolder": str(tmp_path), "outputFolder": str(tmp_path),
           "globalParams": dict(GLOBAL, type="CUDA_K_GESVDJ"), "files": files}
    r = run(denoising_bin, cfg, tmp_path)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    before, after = printed(r.stdout, "PSNR before:"), printed(r.stdout, "PSNR after:")
    for i, c in enumerate(meta["cases"]):
        pic = c["picture"]
        noisy = arr[pic + "_noisy"].astype(np.float32)
        ref = arr[pic + "_ref"].astype(np.float32)
        ks = kl.KSvd(patch=(8, 8), slide=(2, 2), atoms=256, omp_iter=5, ksvd_iter=c["ksvd_iter"])
        ks.set_picture(noisy)
        ks.extract()
        ks.init_dictionary()
        ks.run()
        img = ks.rebuild_picture(np.float32(kl.noise_variance(noisy)))
        ks.close()
        p_abi = kl.psnr(ref, kl.normalize(img))
        print(f"{pic}: CLI PSNR before {before[i]} after {after[i]}; C ABI {p_abi:.6g}; published {c['gpu_psnr']} "
              f"(un-normalised output)")
        assert abs(before[i] - kl.psnr(ref, noisy)) < 1e-3
        assert abs(after[i] - p_abi) < 1e-3
        out = read_pgm(tmp_path / f"8_8_{pic}.pgm")
        assert out.shape == noisy.shape and out.min() == 0 and out.max() == 255


@pytest.mark.gpu
def test_cli_colour_speckle_non_square(denoising_bin, tmp_path, kl):
    """A colour picture (B&W false) with multiplicative speckle (speckle
    true), 96 x 80: the front end keeps Y of YCbCr, estimates sigma^2 on Y,
    Cb and Cr pooled, denoises log2 Y, takes 2^v of the result and stretches
    it (denoiser.cpp:37-99, cuda_svd_denoiser.cu:483-484).  Its output must be
    the bits of the same steps taken through the C ABI."""
    W, H = 96, 80
    rng = np.random.default_rng(5)
    clean = np.stack([kl.synth_image(s, W, H, 0.0)[0] for s in (11, 12, 13)], axis=-1) + 10.0
    speck = rng.gamma(4.0, 0.25, size=clean.shape).astype(np.float32)  # Gamma(L = 4), mean 1
    noisy = (clean * speck).astype(np.float32)
    write_pfm_rgb(tmp_path / "rgb.pfm", noisy)
    g = dict(GLOBAL, **{"B&W": False, "speckle": True, "ksvdIter": 3, "atoms": 64, "patchWidthDim": 6,
                        "patchHeightDim": 6})
    cfg = {"inputFolder": str(tmp_path), "outputFolder": str(tmp_path), "globalParams": g,
           "files": [{"name": "rgb.pfm"}]}
    r = run(denoising_bin, cfg, tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    out = read_pfm(tmp_path / "6_6_rgb.pfm")
    # the same steps through the C ABI
    ycc = kl.rgb_to_ycbcr(noisy)
    s2 = np.float32(kl.noise_variance_channels(ycc))
    frame = kl.speckle_log(kl.image_to_frame(ycc[0]))
    ks = kl.KSvd(patch=(6, 6), slide=(2, 2), atoms=64, omp_iter=5, ksvd_iter=3)
    ks.set_frame(frame, W, H)
    ks.extract()
    ks.init_dictionary()
    ks.run()
    den = kl.speckle_exp(ks.rebuild(s2).ravel())
    ks.close()
    expect = kl.normalize(kl.frame_to_image(den, W, H))
    print(f"\ncolour speckle 96x80: sigma^2 (Y, Cb, Cr pooled) {s2:.2f}, Y alone {kl.noise_variance(ycc[0]):.2f}; "
          f"output range [{out.min():.3f}, {out.max():.3f}]")
    assert out.shape == (H, W)
    np.testing.assert_array_equal(out, expect)


@pytest.mark.gpu
@pytest.mark.parametrize("W,H,th,tw,geom", [(96, 64, 24, 32, "true"), (64, 64, 32, 32, "compat"),
                                            (100, 70, 30, 24, "true")],
                         ids=["96x64_true", "64x64_compat", "100x70_true_ragged"])
def test_cli_subimage_tiling(denoising_bin, tmp_path, kl, W, H, th, tw, geom):
    """subImageHeightDim x subImageWidthDim tiles (cuda_svd_denoiser.cu:130-188,
    503-528): each tile denoised on its own with its own noise variance and
    put back in place; the front end's output must be the bits of those
    steps taken through the C ABI tile by tile.  In the true geometry the
    pixels no tile covers keep their input (the reference leaves 0 there)."""
    clean, noisy = kl.synth_image(31, W, H, 18.0)
    write_pfm(tmp_path / "n.pfm", noisy)
    g = dict(GLOBAL, **{"subImageWidthDim": tw, "subImageHeightDim": th, "ksvdIter": 2, "atoms": 32,
                        "geometry": geom})
    cfg = {"inputFolder": str(tmp_path), "outputFolder": str(tmp_path), "globalParams": g,
           "files": [{"name": "n.pfm"}]}
    r = run(denoising_bin, cfg, tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    out = read_pfm(tmp_path / "8_8_n.pfm")
    compat = geom == "compat"
    frame = kl.image_to_frame(noisy)
    fm, fn = (W, H) if compat else (H, W)
    tiles = tiles_cut(frame, fm, fn, th, tw)
    res = np.zeros((fn, fm), np.float32) if compat else frame.reshape(fn, fm).copy()
    ks = kl.KSvd(patch=(8, 8), slide=(2, 2), atoms=32, omp_iter=5, ksvd_iter=2,
                 geometry=kl.GEOM_COMPAT if compat else kl.GEOM_TRUE)
    k = 0
    for ti in range(fn // th):
        for tj in range(fm // tw):
            t = tiles[k]
            k += 1
            # the tile's noise variance: read as rows of length th (reference) or tw (true)
            s2 = np.float32(kl.noise_variance(t.reshape(tw, th) if compat else t))
            ks.set_frame(t.ravel(), th, tw)
            ks.extract()
            ks.init_dictionary()
            ks.run()
            res[ti * th:(ti + 1) * th, tj * tw:(tj + 1) * tw] = ks.rebuild(s2).reshape(th, tw)
    ks.close()
    expect = kl.normalize(kl.frame_to_image(res.ravel(), W, H))
    gain = kl.psnr(clean, kl.frame_to_image(res.ravel(), W, H)) - kl.psnr(clean, noisy)
    print(f"\ntiles {W}x{H} / {th}x{tw} {geom}: {k} tiles, raw PSNR gain {gain:.2f} dB")
    np.testing.assert_array_equal(out, expect)
    assert gain > 1.0


@pytest.mark.gpu
def test_cli_batch_uneven(denoising_bin, tmp_path, kl, orc):
    """A reduced batch of the uneven workload through `denoising config.json`:
    a colour picture with speckle, a grayscale one at 16 x 12 patches, one
    with a per-file 22 x 22 override and a non-square one, in one config
    (batch_denoiser.cpp:77-112 overrides).  The three 16 x 12 files share
    their sweep launches.  Each output must be the bits of the same steps
    through the C ABI for that file alone, and each C ABI result's raw PSNR
    within 0.05 dB of the FP32 oracle's (3 K-SVD iterations)."""
    rng = np.random.default_rng(17)
    specs = [dict(name="rgb_speckle.pfm", W=192, H=160, rgb=True, seed=61),
             dict(name="gray16x12.pfm", W=256, H=256, sigma=30.0, seed=62),
             dict(name="gray22x22.pfm", W=224, H=160, sigma=40.0, seed=63, patch=(22, 22)),
             dict(name="wide.pfm", W=320, H=192, sigma=15.0, seed=64)]
    files = []
    for sp in specs:
        W, H = sp["W"], sp["H"]
        if sp.get("rgb"):
            clean = np.stack([kl.synth_image(sp["seed"] + 100 * c, W, H, 0.0)[0] for c in range(3)], axis=-1) + 10.0
            noisy = (clean * rng.gamma(4.0, 0.25, size=clean.shape)).astype(np.float32)
            write_pfm_rgb(tmp_path / sp["name"], noisy)
            sp["clean_y"] = kl.rgb_to_ycbcr(clean)[0]
            sp["noisy_planes"] = kl.rgb_to_ycbcr(noisy)
        else:
            clean, noisy = kl.synth_image(sp["seed"], W, H, sp["sigma"])
            write_pfm(tmp_path / sp["name"], noisy)
            sp["clean_y"] = clean
            sp["noisy_planes"] = noisy[None]
        f = {"name": sp["name"]}
        if sp.get("rgb"):
            f.update({"B&W": False, "speckle": True})
        if "patch" in sp:
            f.update({"patchWidthDim": sp["patch"][0], "patchHeightDim": sp["patch"][1]})
        files.append(f)
    g = dict(GLOBAL, **{"patchWidthDim": 16, "patchHeightDim": 12, "slidingWidth": 3, "slidingHeight": 3,
                        "ksvdIter": 3, "atoms": 256})
    cfg = {"inputFolder": str(tmp_path), "outputFolder": str(tmp_path), "globalParams": g, "files": files}
    r = run(denoising_bin, cfg, tmp_path)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    # the three 16 x 12 files share their sweep launches (ksvd_run_batch)
    assert r.stdout.count("sweeps shared by 3 pictures") == 3
    for sp in specs:
        pw, ph = sp.get("patch", (16, 12))
        W, H = sp["W"], sp["H"]
        planes = sp["noisy_planes"]
        s2 = np.float32(kl.noise_variance_channels(planes))
        frame = kl.image_to_frame(planes[0])
        if sp.get("rgb"):
            frame = kl.speckle_log(frame)
        ks = kl.KSvd(patch=(pw, ph), slide=(3, 3), atoms=256, omp_iter=5, ksvd_iter=3)
        ks.set_frame(frame, W, H)
        ks.extract()
        ks.init_dictionary()
        ks.run()
        raw = ks.rebuild(s2).ravel()
        ks.close()
        den = kl.speckle_exp(raw) if sp.get("rgb") else raw
        out = read_pfm(tmp_path / f"{pw}_{ph}_{sp['name']}")
        np.testing.assert_array_equal(out, kl.normalize(kl.frame_to_image(den, W, H)), err_msg=sp["name"])
        # the oracle on the same frame (the log domain for speckle)
        Y = orc.extract(frame, W, H, pw, ph, 3, 3)
        D, idx, coef, cnt, _ = orc.ksvd(Y, orc.init_dict(Y, 256), 5, 3)
        oo, *_ = orc.rebuild(idx, coef, cnt, D, frame, W, H, pw, ph, 3, 3, s2)
        oo = oo.astype(np.float32).ravel()
        if sp.get("rgb"):
            oo = kl.speckle_exp(oo)
        ref = sp["clean_y"]
        pg = kl.psnr(ref, kl.frame_to_image(den, W, H))
        po = kl.psnr(ref, kl.frame_to_image(oo, W, H))
        print(f"{sp['name']} {W}x{H} {pw}x{ph}: raw PSNR gpu {pg:.4f} oracle {po:.4f} ({pg - po:+.4f} dB)")
        assert abs(pg - po) <= 0.05
