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


def test_subimage_tiling_reports_denoise_failure(denoising_bin, tmp_path, kl):
    clean, noisy = kl.synth_image(1, 64, 64, 10.0)
    write_pfm(tmp_path / "a.pfm", noisy)
    cfg = {"inputFolder": str(tmp_path), "outputFolder": str(tmp_path),
           "globalParams": dict(GLOBAL, subImageWidthDim=32, subImageHeightDim=32), "files": [{"name": "a.pfm"}]}
    r = run(denoising_bin, cfg, tmp_path)
    assert r.returncode == 2 and "Denoising Failed" in r.stdout


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
