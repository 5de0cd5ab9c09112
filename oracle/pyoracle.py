"""ctypes binding of the CPU oracle (oracle/build/libksvd_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker (or the timed CPU
baseline), never as the product path.

Every function takes/returns numpy arrays; `real` selects the FP32
('f32', the reference's semantics) or FP64 ('f64') instantiation.
Layouts follow the C ABI: Y is (P, dim) (patch-contiguous), D is (K, dim)
(atom-contiguous, i.e. the column-major dim x K matrix), codes are compact
(P, kmax) idx/coef plus cnt (P,).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libksvd_oracle.so")

_lib = None
_L = ctypes.c_long
_I = ctypes.c_int
_P = ctypes.c_void_p


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_patch_count.restype = _L
        L.orc_patch_count.argtypes = [_I] * 6
        L.orc_kmax.restype = _I
        for suf in ("_f32", "_f64"):
            getattr(L, "orc_extract" + suf).argtypes = [_P, _I, _I, _I, _I, _I, _I, _P]
            getattr(L, "orc_init_dict" + suf).argtypes = [_P, _I, _L, _I, _P]
            getattr(L, "orc_omp" + suf).argtypes = [_P, _I, _L, _P, _I, _I, _P, _P, _P, _P]
            getattr(L, "orc_users" + suf).argtypes = [_P, _P, _P, _L, _I, _I, _P, _P]
            getattr(L, "orc_users" + suf).restype = _L
            getattr(L, "orc_update" + suf).argtypes = [_P, _I, _L, _P, _I, _I, _P, _P, _P, _I, _P, _P]
            getattr(L, "orc_rebuild" + suf).argtypes = [_P, _P, _P, _P, _I, _L, _I, _P, _I, _I, _I, _I, _I, _I,
                                                        _I, ctypes.c_float, _P, _P, _P, _P]
            getattr(L, "orc_ksvd" + suf).argtypes = [_P, _I, _L, _P, _I, _I, _I, _I, _I, _P, _P, _P, _P]
        _lib = L
    return _lib


def _dt(real):
    return np.float32 if real == "f32" else np.float64


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _ptr(a):
    return a.ctypes.data_as(_P)


def patch_count(m, n, pw, ph, sw, sh):
    return lib().orc_patch_count(m, n, pw, ph, sw, sh)


def frame_dims(m, n, compat=False):
    """Row length and row count the patch walk uses for a frame buffer given as
    (m = picture width W, n = height H; buf[x*H + y] = picture[y][x]).  The
    reference's walk (compat) reads it as n rows of length m
    (cuda_svd_denoiser.cu:231-239); the true geometry reads it as what it
    holds, m rows of length n, so that every patch is a block of the picture.
    For a square picture the two agree."""
    return (m, n) if compat else (n, m)


def extract(buf, m, n, pw, ph, sw, sh, real="f32", compat=False):
    """Patches of the frame buffer (m = W, n = H) -> (P, pw*ph), in the true
    geometry or the reference's (compat) walk."""
    buf = _c(buf, np.float32).ravel()
    assert buf.size == m * n
    m, n = frame_dims(m, n, compat)
    P = patch_count(m, n, pw, ph, sw, sh)
    Y = np.empty((P, pw * ph), _dt(real))
    getattr(lib(), "orc_extract_" + real)(_ptr(buf), m, n, pw, ph, sw, sh, _ptr(Y))
    return Y


def init_dict(Y, K, real="f32"):
    Y = _c(Y, _dt(real))
    P, dim = Y.shape
    D = np.empty((K, dim), _dt(real))
    getattr(lib(), "orc_init_dict_" + real)(_ptr(Y), dim, P, K, _ptr(D))
    return D


def omp(Y, D, kmax, real="f32"):
    """-> idx (P, kmax) int32, coef (P, kmax), cnt (P,) int32, stats [dup, degenerate, early]."""
    Y = _c(Y, _dt(real))
    D = _c(D, _dt(real))
    P, dim = Y.shape
    K = D.shape[0]
    idx = np.empty((P, kmax), np.int32)
    coef = np.empty((P, kmax), _dt(real))
    cnt = np.empty(P, np.int32)
    stats = np.zeros(3, np.int64)
    getattr(lib(), "orc_omp_" + real)(_ptr(Y), dim, P, _ptr(D), K, kmax, _ptr(idx), _ptr(coef), _ptr(cnt),
                                      _ptr(stats))
    return idx, coef, cnt, stats


def users(idx, coef, cnt, K, real="f32"):
    idx = _c(idx, np.int32)
    coef = _c(coef, _dt(real))
    cnt = _c(cnt, np.int32)
    P, kmax = idx.shape
    off = np.empty(K + 1, np.int64)
    us = np.empty(int(cnt.sum()) + 1, np.int32)
    getattr(lib(), "orc_users_" + real)(_ptr(idx), _ptr(coef), _ptr(cnt), P, K, kmax, _ptr(off), _ptr(us))
    return off, us[: off[-1]]


def update(Y, D, idx, coef, cnt, cap=1024, real="f32", want_gap=False):
    """One Gauss-Seidel sweep.  Returns (D', coef', gap or None, eig iterations)."""
    Y = _c(Y, _dt(real))
    D = np.array(D, dtype=_dt(real), copy=True, order="C")
    idx = _c(idx, np.int32)
    coef = np.array(coef, dtype=_dt(real), copy=True, order="C")
    cnt = _c(cnt, np.int32)
    P, dim = Y.shape
    K = D.shape[0]
    kmax = idx.shape[1]
    gap = np.empty(K, np.float64) if want_gap else None
    it = np.zeros(1, np.int64)
    getattr(lib(), "orc_update_" + real)(_ptr(Y), dim, P, _ptr(D), K, kmax, _ptr(idx), _ptr(coef), _ptr(cnt),
                                         cap, _ptr(gap) if want_gap else None, _ptr(it))
    return D, coef, gap, int(it[0])


def rebuild(idx, coef, cnt, D, buf, m, n, pw, ph, sw, sh, sigma2, compat=False, real="f32"):
    """-> out, sum, weight (each (n, m), the buffer's own layout), dsum
    (double), lambda.  compat: the reference's walks (its reading of the
    frame and its transposed rebuild grid for rectangular patches)."""
    n0, m0 = n, m
    m, n = frame_dims(m, n, compat)
    idx = _c(idx, np.int32)
    coef = _c(coef, _dt(real))
    cnt = _c(cnt, np.int32)
    D = _c(D, _dt(real))
    buf = _c(buf, np.float32).ravel()
    P, kmax = idx.shape
    dim = D.shape[1]
    out = np.empty(m * n, _dt(real))
    s = np.empty(m * n, _dt(real))
    w = np.empty(m * n, _dt(real))
    dl = np.zeros(2, np.float64)
    getattr(lib(), "orc_rebuild_" + real)(_ptr(idx), _ptr(coef), _ptr(cnt), _ptr(D), dim, P, kmax, _ptr(buf), m, n,
                                          pw, ph, sw, sh, 1 if compat else 0, ctypes.c_float(sigma2), _ptr(out),
                                          _ptr(s), _ptr(w), _ptr(dl))
    return out.reshape(n0, m0), s.reshape(n0, m0), w.reshape(n0, m0), dl[0], dl[1]


def ksvd(Y, D, kmax, iters, cap=1024, final_omp=True, real="f32"):
    """The K-SVD loop.  Returns (D, idx, coef, cnt, [omp_ms, update_ms])."""
    Y = _c(Y, _dt(real))
    D = np.array(D, dtype=_dt(real), copy=True, order="C")
    P, dim = Y.shape
    K = D.shape[0]
    idx = np.empty((P, kmax), np.int32)
    coef = np.empty((P, kmax), _dt(real))
    cnt = np.empty(P, np.int32)
    t = np.zeros(2, np.float64)
    getattr(lib(), "orc_ksvd_" + real)(_ptr(Y), dim, P, _ptr(D), K, kmax, iters, cap, 1 if final_omp else 0,
                                       _ptr(idx), _ptr(coef), _ptr(cnt), _ptr(t))
    return D, idx, coef, cnt, t


def dense(idx, coef, cnt, K):
    """Compact codes -> (P, K) dense (row p = column p of the reference's X)."""
    P, kmax = idx.shape
    X = np.zeros((P, K), coef.dtype)
    for i in range(kmax):
        sel = cnt > i
        X[np.nonzero(sel)[0], idx[sel, i]] = coef[sel, i]
    return X
