"""ctypes binding of libksvd_hip.so (the C ABI declared in include/ksvd_hip.h).

Used by the tests, __graft_entry__ and bench.py.  The GPU path lives entirely
in the shared library; this module only marshals numpy arrays.  If the library
has not been built, loading fails loudly (there is no CPU fallback).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libksvd_hip.so")

KSVD_OK = 0
KSVD_ERR_ARG = -1
KSVD_ERR_STATE = -2
KSVD_ERR_HIP = -3
KSVD_ERR_ALLOC = -4
KSVD_ERR_UNSUPPORTED = -5

PREC_FP32 = 0
PREC_BF16 = 1
GEOM_TRUE = 0
GEOM_COMPAT = 1


class Params(ctypes.Structure):
    _fields_ = [
        ("patch_w", ctypes.c_int),
        ("patch_h", ctypes.c_int),
        ("slide_w", ctypes.c_int),
        ("slide_h", ctypes.c_int),
        ("atoms", ctypes.c_int),
        ("omp_iter", ctypes.c_int),
        ("ksvd_iter", ctypes.c_int),
        ("min_omp_iter_batch", ctypes.c_int),
        ("user_cap", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("geometry", ctypes.c_int),
        ("device", ctypes.c_int),
        ("proposer", ctypes.c_int),
    ]


class Timings(ctypes.Structure):
    _fields_ = [
        ("gather_ms", ctypes.c_double),
        ("init_ms", ctypes.c_double),
        ("omp_ms", ctypes.c_double),
        ("omp_kernel_ms", ctypes.c_double),
        ("omp_calls", ctypes.c_int64),
        ("update_ms", ctypes.c_double),
        ("update_calls", ctypes.c_int64),
        ("rebuild_ms", ctypes.c_double),
        ("omp_dup_stops", ctypes.c_int64),
        ("omp_degenerate", ctypes.c_int64),
        ("eig_iters", ctypes.c_int64),
        ("omp_rescored", ctypes.c_int64),
        ("omp_f64_refactors", ctypes.c_int64),
        ("omp_candidates_cut", ctypes.c_int64),
        ("train_ms", ctypes.c_double),
        ("proposed", ctypes.c_int64),
        ("fallback", ctypes.c_int64),
        ("train_loss", ctypes.c_double),
        ("proposer_gated", ctypes.c_int64),
        ("sweep_fallbacks", ctypes.c_int64),
    ]


# name -> (restype, argtypes); every function include/ksvd_hip.h declares.
_P = ctypes.c_void_p
_FP = ctypes.POINTER(ctypes.c_float)
_IP = ctypes.POINTER(ctypes.c_int32)
SIGNATURES = {
    "ksvd_default_params": (None, [ctypes.POINTER(Params)]),
    "ksvd_create": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.POINTER(_P)]),
    "ksvd_destroy": (ctypes.c_int, [_P]),
    "ksvd_last_error": (ctypes.c_char_p, [_P]),
    "ksvd_set_image": (ctypes.c_int, [_P, _FP, ctypes.c_int, ctypes.c_int]),
    "ksvd_set_image_device": (ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.c_int]),
    "ksvd_extract_patches": (ctypes.c_int, [_P]),
    "ksvd_init_dictionary": (ctypes.c_int, [_P]),
    "ksvd_omp": (ctypes.c_int, [_P]),
    "ksvd_omp_noalloc": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int64, _P, ctypes.c_int, ctypes.c_int, _P, _P, _P,
                                        _P]),
    "ksvd_update_dictionary": (ctypes.c_int, [_P]),
    "ksvd_iterate": (ctypes.c_int, [_P, ctypes.c_int]),
    "ksvd_run": (ctypes.c_int, [_P]),
    "ksvd_iterate_batch": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int, ctypes.c_int]),
    "ksvd_run_batch": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int]),
    "ksvd_run_final_code": (ctypes.c_int, [_P]),
    "ksvd_rebuild": (ctypes.c_int, [_P, ctypes.c_float, _FP]),
    "ksvd_rebuild_speckle": (ctypes.c_int, [_P, ctypes.c_float, _FP]),
    "ksvd_get_rebuild_parts": (ctypes.c_int, [_P, _FP, _FP, _FP, ctypes.POINTER(ctypes.c_double)]),
    "ksvd_rebuild_device_ptr": (ctypes.c_int, [_P, ctypes.POINTER(_P)]),
    "ksvd_get_dims": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int64),
                                     ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "ksvd_get_patches": (ctypes.c_int, [_P, _FP]),
    "ksvd_get_dictionary": (ctypes.c_int, [_P, _FP]),
    "ksvd_set_dictionary": (ctypes.c_int, [_P, _FP]),
    "ksvd_get_codes": (ctypes.c_int, [_P, _IP, _FP, _IP]),
    "ksvd_set_codes": (ctypes.c_int, [_P, _IP, _FP, _IP]),
    "ksvd_get_codes_dense": (ctypes.c_int, [_P, _FP]),
    "ksvd_get_timings": (ctypes.c_int, [_P, ctypes.POINTER(Timings)]),
    "ksvd_reset_timings": (ctypes.c_int, [_P]),
    "ksvd_synchronize": (ctypes.c_int, [_P]),
    "ksvd_proposer_dims": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int64)]),
    "ksvd_proposer_init": (ctypes.c_int, [_P, ctypes.c_uint64]),
    "ksvd_proposer_get_params": (ctypes.c_int, [_P, _FP]),
    "ksvd_proposer_set_params": (ctypes.c_int, [_P, _FP]),
    "ksvd_proposer_train": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_float, ctypes.c_uint64]),
    "ksvd_proposer_omp": (ctypes.c_int, [_P]),
    "ksvd_proposer_loss_grad": (ctypes.c_int, [_P, _FP, _FP, ctypes.c_int, ctypes.POINTER(ctypes.c_double), _FP]),
    "ksvd_synth_image": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_float, _FP, _FP]),
    "ksvd_noise_variance": (ctypes.c_double, [_FP, ctypes.c_int, ctypes.c_int]),
    "ksvd_normalize": (None, [_FP, ctypes.c_int64, ctypes.c_float, ctypes.c_float]),
    "ksvd_psnr": (ctypes.c_double, [_FP, _FP, ctypes.c_int64]),
    "ksvd_rgb_to_y": (None, [_FP, _FP, _FP, ctypes.c_int64, _FP]),
    "ksvd_rgb_to_ycbcr": (None, [_FP, _FP, _FP, ctypes.c_int64, _FP, _FP, _FP]),
    "ksvd_noise_variance_channels": (ctypes.c_double, [_FP, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "ksvd_speckle_log": (None, [_FP, ctypes.c_int64]),
    "ksvd_speckle_exp": (None, [_FP, ctypes.c_int64]),
    "ksvd_image_to_frame": (None, [_FP, ctypes.c_int, ctypes.c_int, _FP]),
    "ksvd_frame_to_image": (None, [_FP, ctypes.c_int, ctypes.c_int, _FP]),
}

_lib = None


def lib():
    """Load libksvd_hip.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C {HERE}` "
                               "(or __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(_FP)


def _ip(a):
    return a.ctypes.data_as(_IP)


class KsvdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ksvd error {code}: {msg}")
        self.code = code


def default_params():
    p = Params()
    lib().ksvd_default_params(ctypes.byref(p))
    return p


class KSvd:
    """One context of the HIP path: image -> patches -> K-SVD -> rebuild."""

    def __init__(self, patch=(8, 8), slide=(2, 2), atoms=256, omp_iter=5, ksvd_iter=10, user_cap=1024,
                 geometry=GEOM_TRUE, precision=PREC_FP32, device=0, proposer=0):
        L = lib()
        p = default_params()
        p.patch_w, p.patch_h = patch
        p.slide_w, p.slide_h = slide
        p.atoms, p.omp_iter, p.ksvd_iter = atoms, omp_iter, ksvd_iter
        p.user_cap, p.geometry, p.precision, p.device = user_cap, geometry, precision, device
        p.proposer = proposer
        self.params = p
        h = _P()
        rc = L.ksvd_create(ctypes.byref(p), ctypes.byref(h))
        self._h = h
        if rc != KSVD_OK:
            msg = L.ksvd_last_error(h).decode() if h else ""
            L.ksvd_destroy(h)
            self._h = None
            raise KsvdError(rc, msg)
        self.kmax = omp_iter
        self.K = atoms
        self.dim = patch[0] * patch[1]

    # -- plumbing ------------------------------------------------------------
    def _check(self, rc):
        if rc != KSVD_OK:
            raise KsvdError(rc, lib().ksvd_last_error(self._h).decode())

    def close(self):
        if self._h:
            lib().ksvd_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- stages --------------------------------------------------------------
    def set_frame(self, buf, m, n):
        """buf: the reference's frame buffer, n rows of length m (m*n floats)."""
        b = np.ascontiguousarray(buf, dtype=np.float32).ravel()
        assert b.size == m * n
        self.m, self.n = m, n
        self._check(lib().ksvd_set_image(self._h, _fp(b), m, n))
        d, P, K, km = self.dims()
        self.P = P

    def set_picture(self, img):
        """img: (H, W) row-major picture; transposed into the frame as Denoiser::loadImage does."""
        img = np.asarray(img, dtype=np.float32)
        H, W = img.shape
        self.set_frame(np.ascontiguousarray(img.T), W, H)

    def set_frame_device(self, dptr, m, n):
        self.m, self.n = m, n
        self._check(lib().ksvd_set_image_device(self._h, _P(dptr), m, n))
        self.P = self.dims()[1]

    def extract(self):
        self._check(lib().ksvd_extract_patches(self._h))

    def init_dictionary(self):
        self._check(lib().ksvd_init_dictionary(self._h))

    def omp(self):
        self._check(lib().ksvd_omp(self._h))

    def update(self):
        self._check(lib().ksvd_update_dictionary(self._h))

    def iterate(self, iters):
        self._check(lib().ksvd_iterate(self._h, iters))

    def run(self):
        self._check(lib().ksvd_run(self._h))

    def rebuild(self, sigma2, copy=True, speckle=False):
        """speckle: the frame is log-domain and the output returns by 2^v on
        the device (ksvd_rebuild_speckle)."""
        out = np.empty(self.m * self.n, dtype=np.float32) if copy else None
        fn = lib().ksvd_rebuild_speckle if speckle else lib().ksvd_rebuild
        self._check(fn(self._h, ctypes.c_float(sigma2), _fp(out) if copy else None))
        return out.reshape(self.n, self.m) if copy else None

    def rebuild_picture(self, sigma2):
        """Denoised (H, W) picture: the frame output transposed back."""
        fr = self.rebuild(sigma2)
        # buf[x*H + y] = picture[y, x] with m = W, n = H
        return np.ascontiguousarray(fr.ravel().reshape(self.m, self.n).T)

    def rebuild_parts(self):
        N = self.m * self.n
        s = np.empty(N, np.float32)
        w = np.empty(N, np.float32)
        lam = ctypes.c_float()
        ds = ctypes.c_double()
        self._check(lib().ksvd_get_rebuild_parts(self._h, _fp(s), _fp(w), ctypes.byref(lam), ctypes.byref(ds)))
        return s.reshape(self.n, self.m), w.reshape(self.n, self.m), lam.value, ds.value

    # -- accessors -----------------------------------------------------------
    def dims(self):
        d, P, K, km = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
        self._check(lib().ksvd_get_dims(self._h, ctypes.byref(d), ctypes.byref(P), ctypes.byref(K),
                                        ctypes.byref(km)))
        return d.value, P.value, K.value, km.value

    def patches(self):
        Y = np.empty((self.P, self.dim), np.float32)
        self._check(lib().ksvd_get_patches(self._h, _fp(Y)))
        return Y

    def dictionary(self):
        """(K, dim): row k is atom k (the column-major dim x K matrix)."""
        D = np.empty((self.K, self.dim), np.float32)
        self._check(lib().ksvd_get_dictionary(self._h, _fp(D)))
        return D

    def set_dictionary(self, D):
        D = np.ascontiguousarray(D, dtype=np.float32)
        assert D.shape == (self.K, self.dim)
        self._check(lib().ksvd_set_dictionary(self._h, _fp(D)))

    def codes(self):
        idx = np.empty((self.P, self.kmax), np.int32)
        coef = np.empty((self.P, self.kmax), np.float32)
        cnt = np.empty(self.P, np.int32)
        self._check(lib().ksvd_get_codes(self._h, _ip(idx), _fp(coef), _ip(cnt)))
        return idx, coef, cnt

    def set_codes(self, idx, coef, cnt):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        coef = np.ascontiguousarray(coef, dtype=np.float32)
        cnt = np.ascontiguousarray(cnt, dtype=np.int32)
        self._check(lib().ksvd_set_codes(self._h, _ip(idx), _fp(coef), _ip(cnt)))

    def codes_dense(self):
        """(P, K): row p is column p of the reference's K x P sparse code."""
        X = np.empty((self.P, self.K), np.float32)
        self._check(lib().ksvd_get_codes_dense(self._h, _fp(X)))
        return X

    def timings(self):
        t = Timings()
        self._check(lib().ksvd_get_timings(self._h, ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in Timings._fields_}

    # -- learned support proposer (BF16 form) ------------------------------
    def proposer_dims(self):
        kin, ko, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
        self._check(lib().ksvd_proposer_dims(self._h, ctypes.byref(kin), ctypes.byref(ko), ctypes.byref(n)))
        return kin.value, ko.value, n.value

    def proposer_init(self, seed):
        self._check(lib().ksvd_proposer_init(self._h, ctypes.c_uint64(seed)))

    def proposer_params(self):
        """(W1 [512 x KIN], W2 [512 x 544], W3 [KO x 544]) FP32 master weights."""
        kin, ko, n = self.proposer_dims()
        th = np.empty(n, np.float32)
        self._check(lib().ksvd_proposer_get_params(self._h, _fp(th)))
        a, b = 512 * kin, 512 * 544
        return th[:a].reshape(512, kin), th[a:a + b].reshape(512, 544), th[a + b:].reshape(ko, 544)

    def set_proposer_params(self, W1, W2, W3):
        th = np.ascontiguousarray(np.concatenate([np.ravel(W1), np.ravel(W2), np.ravel(W3)]), dtype=np.float32)
        assert th.size == self.proposer_dims()[2]
        self._check(lib().ksvd_proposer_set_params(self._h, _fp(th)))

    def proposer_train(self, steps, lr=2e-3, seed=1):
        self._check(lib().ksvd_proposer_train(self._h, steps, ctypes.c_float(lr), ctypes.c_uint64(seed)))

    def proposer_omp(self):
        self._check(lib().ksvd_proposer_omp(self._h))

    def proposer_loss_grad(self, X, T):
        """Mean loss and gradients (W1, W2, W3 shaped) of one batch on the GPU."""
        kin, ko, n = self.proposer_dims()
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        nb = X.shape[0]
        assert X.shape == (nb, kin) and T.shape == (nb, ko)
        g = np.empty(n, np.float32)
        loss = ctypes.c_double()
        self._check(lib().ksvd_proposer_loss_grad(self._h, _fp(X), _fp(T), nb, ctypes.byref(loss), _fp(g)))
        a, b = 512 * kin, 512 * 544
        return loss.value, (g[:a].reshape(512, kin), g[a:a + b].reshape(512, 544), g[a + b:].reshape(ko, 544))

    def reset_timings(self):
        self._check(lib().ksvd_reset_timings(self._h))

    def synchronize(self):
        self._check(lib().ksvd_synchronize(self._h))


def _handles(ctxs):
    return (_P * len(ctxs))(*[k._h for k in ctxs])


def _batch_check(ctxs, rc):
    if rc != KSVD_OK:
        msgs = [lib().ksvd_last_error(k._h).decode() for k in ctxs]
        raise KsvdError(rc, next((m for m in msgs if m), ""))


def iterate_batch(ctxs, iters):
    """ksvd_iterate_batch: several pictures' iterations at once (one
    dictionary sweep launch with picture i on XCD i)."""
    _batch_check(ctxs, lib().ksvd_iterate_batch(_handles(ctxs), len(ctxs), iters))


def run_batch(ctxs):
    """ksvd_run_batch: ksvd_iter iterations of every picture at once, then
    each picture's final code."""
    _batch_check(ctxs, lib().ksvd_run_batch(_handles(ctxs), len(ctxs)))


def omp_noalloc(Y, dim, P, D, K, kmax, idx, coef, cnt, stream=0):
    """ksvd_omp_noalloc on device pointers (ints); raises KsvdError on failure."""
    rc = lib().ksvd_omp_noalloc(_P(Y), dim, P, _P(D), K, kmax, _P(idx), _P(coef), _P(cnt), _P(stream or None))
    if rc != KSVD_OK:
        raise KsvdError(rc, lib().ksvd_last_error(None).decode())


# ---- host helpers ------------------------------------------------------------
def synth_image(seed, W, H, sigma):
    """(clean, noisy), each (H, W) float32."""
    clean = np.empty((H, W), np.float32)
    noisy = np.empty((H, W), np.float32)
    rc = lib().ksvd_synth_image(ctypes.c_uint64(seed), W, H, ctypes.c_float(sigma), _fp(clean), _fp(noisy))
    if rc != KSVD_OK:
        raise KsvdError(rc, "synth_image")
    return clean, noisy


def noise_variance(img):
    img = np.ascontiguousarray(img, dtype=np.float32)
    H, W = img.shape
    return lib().ksvd_noise_variance(_fp(img), W, H)


def normalize(img, lo=0.0, hi=255.0):
    out = np.array(img, dtype=np.float32, copy=True, order="C")
    lib().ksvd_normalize(_fp(out), out.size, ctypes.c_float(lo), ctypes.c_float(hi))
    return out


def psnr(ref, img):
    a = np.ascontiguousarray(ref, dtype=np.float32)
    b = np.ascontiguousarray(img, dtype=np.float32)
    assert a.size == b.size
    return lib().ksvd_psnr(_fp(a), _fp(b), a.size)


def rgb_to_y(rgb):
    """(H, W, 3) colour picture -> (H, W) luma as Denoiser::loadImage keeps it."""
    rgb = np.asarray(rgb, dtype=np.float32)
    planes = [np.ascontiguousarray(rgb[..., c]) for c in range(3)]
    y = np.empty(rgb.shape[:2], np.float32)
    lib().ksvd_rgb_to_y(_fp(planes[0]), _fp(planes[1]), _fp(planes[2]), y.size, _fp(y))
    return y


def rgb_to_ycbcr(rgb):
    """(H, W, 3) colour picture -> (3, H, W) Y, Cb, Cr planes (CImg RGBtoYCbCr)."""
    rgb = np.asarray(rgb, dtype=np.float32)
    planes = [np.ascontiguousarray(rgb[..., c]) for c in range(3)]
    out = np.empty((3,) + rgb.shape[:2], np.float32)
    lib().ksvd_rgb_to_ycbcr(_fp(planes[0]), _fp(planes[1]), _fp(planes[2]), planes[0].size, _fp(out[0]),
                            _fp(out[1]), _fp(out[2]))
    return out


def noise_variance_channels(planes):
    """CImg variance_noise over (C, H, W) planes pooled."""
    p = np.ascontiguousarray(planes, dtype=np.float32)
    C, H, W = p.shape
    return lib().ksvd_noise_variance_channels(_fp(p), W, H, C)


def speckle_log(a):
    out = np.array(a, dtype=np.float32, copy=True, order="C")
    lib().ksvd_speckle_log(_fp(out), out.size)
    return out


def speckle_exp(a):
    out = np.array(a, dtype=np.float32, copy=True, order="C")
    lib().ksvd_speckle_exp(_fp(out), out.size)
    return out


def image_to_frame(img):
    img = np.ascontiguousarray(img, dtype=np.float32)
    H, W = img.shape
    buf = np.empty(W * H, np.float32)
    lib().ksvd_image_to_frame(_fp(img), W, H, _fp(buf))
    return buf


def frame_to_image(buf, W, H):
    buf = np.ascontiguousarray(buf, dtype=np.float32).ravel()
    img = np.empty((H, W), np.float32)
    lib().ksvd_frame_to_image(_fp(buf), W, H, _fp(img))
    return img


def patch_count(m, n, pw, ph, sw, sh):
    if pw > m or ph > n:
        return 0
    return ((n - ph) // sh + 1) * ((m - pw) // sw + 1)
