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


def test_noise_variance_stripes_exact(kl):
    """Vertical stripes 0, c, 0, c, ... (W even): the scaled Laplacian
    (s = 1/sqrt(20)) is +2cs on the inner 0-columns, -2cs on the inner
    c-columns, +cs on column 0 and -cs on column W-1 (replicated borders: the
    missing neighbour is the pixel itself).  Half the values are negative, so
    the median (element N/2 of the sorted values) is +cs; the deviations are
    0 (H), cs (H(W-2)/2), 2cs (H), 3cs (H(W-2)/2), whose element N/2 is 2cs.
    Variance = (1.4828 * 2cs)^2.  Zero-padded borders would give 3cs, an
    unscaled Laplacian 20x the value, the 1.4826 constant a 2.7e-4 change."""
    for W, H, c in ((64, 31, 10.0), (10, 7, 3.0)):
        img = np.zeros((H, W), np.float32)
        img[:, 1::2] = c
        expect = (1.4828 * 2 * c / np.sqrt(20.0)) ** 2
        assert abs(kl.noise_variance(img) - expect) <= 1e-12 * expect, (W, H)
        # transposed (horizontal stripes): the Laplacian is symmetric in x and y
        assert abs(kl.noise_variance(np.ascontiguousarray(img.T)) - expect) <= 1e-12 * expect


def test_rgb_to_y_closed_forms(kl):
    # BT.601 luma on the 8-bit scale, Y = (66 R + 129 G + 25 B + 128) / 256 + 16
    rgb = np.array([[[255, 255, 255], [0, 0, 0], [255, 0, 0], [0, 255, 0], [0, 0, 255], [400, 400, 400],
                     [-100, -100, -100]]], np.float32)
    y = kl.rgb_to_y(rgb)
    expect = [235.640625, 16.5, 82.2421875, 144.99609375, 41.40234375, 255.0, 0.0]
    np.testing.assert_array_equal(y[0], np.array(expect, np.float32))


def test_rgb_to_ycbcr_closed_forms(kl):
    # Cb = (-38 R - 74 G + 112 B + 128) / 256 + 128, Cr = (112 R - 94 G - 18 B + 128) / 256 + 128
    rgb = np.array([[[255, 255, 255], [255, 0, 0], [0, 0, 255], [0, 0, 0]]], np.float32)
    ycc = kl.rgb_to_ycbcr(rgb)
    np.testing.assert_array_equal(ycc[0, 0], np.array([235.640625, 82.2421875, 41.40234375, 16.5], np.float32))
    np.testing.assert_array_equal(ycc[1, 0], np.array([128.5, 90.6484375, 240.0625, 128.5], np.float32))
    np.testing.assert_array_equal(ycc[2, 0], np.array([128.5, 240.0625, 110.5703125, 128.5], np.float32))


def test_noise_variance_channels_pooled(kl):
    """Stripes (as above) in plane 0 and a constant plane 1: the pooled
    Laplacian values are WH/2 negatives, WH zeros and WH/2 positives, so
    the median is 0, the deviations are 0 (WH), cs (2H) and 2cs, and element
    WH of them is cs: variance (1.4828 cs)^2, a quarter of plane 0's alone."""
    W, H, c = 64, 31, 10.0
    p0 = np.zeros((H, W), np.float32)
    p0[:, 1::2] = c
    planes = np.stack([p0, np.full((H, W), 5.0, np.float32)])
    expect = (1.4828 * c / np.sqrt(20.0)) ** 2
    assert abs(kl.noise_variance_channels(planes) - expect) <= 1e-12 * expect
    assert abs(kl.noise_variance_channels(planes[:1]) - kl.noise_variance(p0)) == 0


def test_speckle_transforms(kl):
    x = np.array([0.0, 1.0, 2.0, 8.0, 0.5, 1024.0], np.float32)
    lg = kl.speckle_log(x)
    np.testing.assert_array_equal(lg, np.array([0, 0, 1, 3, -1, 10], np.float32))
    np.testing.assert_array_equal(kl.speckle_exp(lg[1:]), x[1:])
    assert kl.speckle_exp(np.zeros(1, np.float32))[0] == 1.0  # a zero pixel comes back as 1


def test_speckle_transforms_threaded_same_bits(kl):
    """Buffers of a million elements or more are split over host threads:
    the same bits as the single-thread loop over pieces below that size."""
    rng = np.random.default_rng(3)
    x = rng.uniform(0.0, 300.0, 5_000_003).astype(np.float32)
    x[::1000] = 0.0
    lg = kl.speckle_log(x)
    ex = kl.speckle_exp(lg)
    piece = 700_001
    lg1 = np.concatenate([kl.speckle_log(x[i:i + piece]) for i in range(0, x.size, piece)])
    ex1 = np.concatenate([kl.speckle_exp(lg[i:i + piece]) for i in range(0, x.size, piece)])
    np.testing.assert_array_equal(lg, lg1)
    np.testing.assert_array_equal(ex, ex1)
    assert (lg[::1000] == 0).all()
    nz = x != 0
    assert np.max(np.abs(ex[nz] - x[nz]) / x[nz]) < 1e-6


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
        kl.KSvd(patch=(33, 32))
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
