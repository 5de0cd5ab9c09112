"""CPU tests of the product library: it loads, exports every symbol the public
header declares, its host helpers follow their stated maths, and without a
GPU the C ABI fails cleanly instead of crashing."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ksvd_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(ksvd_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_abi():
    names = declared_functions()
    for must in ("ksvd_create", "ksvd_destroy", "ksvd_set_image", "ksvd_extract_patches", "ksvd_init_dictionary",
                 "ksvd_omp", "ksvd_update_dictionary", "ksvd_iterate", "ksvd_run", "ksvd_rebuild"):
        assert must in names


def test_library_exports_every_declared_symbol(kl):
    lib = kl.lib()
    exported = subprocess.run(["nm", "-D", "--defined-only", kl.LIB_PATH], capture_output=True, text=True).stdout
    for name in declared_functions():
        assert re.search(r"\bT " + name + r"\b", exported), f"{name} not exported"
        assert hasattr(lib, name)
    # the ctypes signature table covers the header exactly
    assert sorted(kl.SIGNATURES) == declared_functions()


def test_default_params_are_the_reference_defaults(kl):
    p = kl.default_params()
    # denoisingLib.h:44-53
    assert (p.patch_w, p.patch_h, p.slide_w, p.slide_h) == (8, 8, 2, 2)
    assert (p.atoms, p.omp_iter, p.ksvd_iter, p.min_omp_iter_batch) == (256, 5, 10, 0)
    assert p.user_cap == 1024 and p.precision == kl.PREC_FP32 and p.geometry == kl.GEOM_TRUE


def test_synth_image_is_seeded(kl):
    c1, n1 = kl.synth_image(42, 96, 64, 20.0)
    c2, n2 = kl.synth_image(42, 96, 64, 20.0)
    c3, _ = kl.synth_image(43, 96, 64, 20.0)
    np.testing.assert_array_equal(c1, c2)
    np.testing.assert_array_equal(n1, n2)
    assert not np.array_equal(c1, c3)
    assert c1.min() == 0.0 and abs(c1.max() - 255.0) < 1e-3
    assert abs(float(np.std(n1 - c1)) - 20.0) < 0.5
    # sigma 0: noisy == clean
    c0, n0 = kl.synth_image(42, 32, 32, 0.0)
    np.testing.assert_array_equal(c0, n0)


def test_noise_variance_white_noise(kl):
    # the scaled Laplacian keeps the variance of white noise and the MAD
    # estimator is consistent for a Gaussian: estimate ~ sigma^2
    rng = np.random.default_rng(0)
    for s in (5.0, 20.0, 50.0):
        img = (100.0 + s * rng.standard_normal((256, 256))).astype(np.float32)
        est = kl.noise_variance(img)
        assert abs(est / s ** 2 - 1.0) < 0.05


def test_noise_variance_constant_and_ramp(kl):
    assert kl.noise_variance(np.full((32, 32), 7.0, np.float32)) == 0.0
    # a linear ramp has zero Laplacian inside; the border rows add a few
    # nonzero values, far below half the pixels: the median is 0
    ramp = np.add.outer(np.arange(64.0), 2 * np.arange(64.0)).astype(np.float32)
    assert kl.noise_variance(ramp) == 0.0


def test_normalize_matches_minmax(kl):
    x = np.array([[-3.0, 1.0], [5.0, 2.0]], np.float32)
    y = kl.normalize(x)
    expect = (x - np.float32(-3)) / np.float32(8) * np.float32(255)
    np.testing.assert_array_equal(y, expect.astype(np.float32))
    np.testing.assert_array_equal(kl.normalize(np.full((2, 2), 3.0, np.float32)), np.zeros((2, 2), np.float32))


def test_psnr_closed_form(kl):
    a = np.zeros(100, np.float32)
    b = np.full(100, 5.0, np.float32)
    assert abs(kl.psnr(a, b) - 10 * np.log10(255.0 ** 2 / 25.0)) < 1e-12
    assert kl.psnr(a, a) == float("inf")


def test_frame_round_trip(kl):
    img = np.arange(35, dtype=np.float32).reshape(5, 7)  # H = 5, W = 7
    buf = kl.image_to_frame(img)
    # buf[x*H + y] = img[y, x]
    assert buf[3 * 5 + 2] == img[2, 3]
    np.testing.assert_array_equal(buf, img.T.ravel())
    np.testing.assert_array_equal(kl.frame_to_image(buf, 7, 5), img)


def test_invalid_params_fail_with_message(kl):
    with pytest.raises(kl.KsvdError) as e:
        kl.KSvd(patch=(0, 8))
    assert e.value.code == kl.KSVD_ERR_ARG and "patch" in str(e.value)
    with pytest.raises(kl.KsvdError) as e:
        kl.KSvd(omp_iter=40)
    assert e.value.code == kl.KSVD_ERR_ARG
    with pytest.raises(kl.KsvdError) as e:
        kl.KSvd(patch=(22, 22))
    assert e.value.code == kl.KSVD_ERR_UNSUPPORTED


def test_create_without_gpu_fails_cleanly(kl):
    import ctypes
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"):
        pytest.skip("a GPU may be visible here")
    with pytest.raises(kl.KsvdError) as e:
        kl.KSvd()
    assert e.value.code == kl.KSVD_ERR_HIP
    assert kl.lib().ksvd_last_error(None) == b"null context"
    assert kl.lib().ksvd_destroy(None) == 0
    del ctypes
