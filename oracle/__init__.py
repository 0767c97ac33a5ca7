"""ctypes binding of the CPU FP64 oracle (oracle/ps_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
benchmark's CPU-baseline leg, always as the checker, never as the product.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libps_oracle.so")
_lib = None

MASK_DEFAULT, MASK_DETERMINING, MASK_RANGE, MASK_WATKINSON, MASK_WATKINSON_KN = range(5)

_dp = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_fft1d.argtypes = [ctypes.c_int, _dp, ctypes.c_int]
        L.or_fft3d.argtypes = [ctypes.c_int, _dp, ctypes.c_int]
        L.or_calc_phase.argtypes = [ctypes.c_int, _dp]
        L.or_kfilter.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp,
                                 ctypes.c_double, ctypes.c_double, _u8p]
        L.or_get_corners.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double, _dp, _dp]
        L.or_calc_k3_min_and_max.argtypes = [_dp, _dp, _dp]
        for f in ("or_calc_k3",):
            getattr(L, f).restype = ctypes.c_double
            getattr(L, f).argtypes = [ctypes.c_double] * 3
        for f in ("or_calc_k3min", "or_calc_k3max"):
            getattr(L, f).restype = ctypes.c_double
            getattr(L, f).argtypes = [ctypes.c_double] * 4
        L.or_polyspectrum.restype = ctypes.c_double
        L.or_polyspectrum.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp,
                                      ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_int, _dp]
        L.or_num_polygons.restype = ctypes.c_double
        L.or_num_polygons.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                      ctypes.c_double, ctypes.c_int, ctypes.c_double]
        L.or_direct_polygon_sum.restype = ctypes.c_int64
        L.or_direct_polygon_sum.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _u8p, _u8p, _u8p, _dp]
        L.or_polygon_gradient.restype = ctypes.c_double
        L.or_polygon_gradient.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_double,
                                          ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double, _dp]
        L.or_product.argtypes = [ctypes.c_int64, _dp, _dp, _dp]
        L.or_sum.argtypes = [ctypes.c_int64, _dp, _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def fft1d(x, sign):
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    assert lib().or_fft1d(a.size, _d(a.view(np.float64)), sign) == 0
    return a


def fft3d(x, sign):
    """Unnormalised 3D C2C FFT (sign -1 forward, +1 backward), as fft.c:19-57."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    n = a.shape[0]
    assert a.shape == (n, n, n)
    assert lib().or_fft3d(n, _d(a.view(np.float64)), sign) == 0
    return a


def forward_spectrum(field):
    """get_FT_field (input_field.c:40-81): promote to complex double, forward FFT."""
    return fft3d(np.asarray(field, dtype=np.float64).astype(np.complex128), -1)


def calc_phase(spec):
    a = np.ascontiguousarray(spec, dtype=np.complex128).copy()
    lib().or_calc_phase(a.shape[0], _d(a.view(np.float64)))
    return a


def kfilter(n, kind, k, binwidth, boxsize, kmin=0.0, kmax=0.0):
    """Returns (mask uint8 N^3, kmin, kmax) for one reference mask builder."""
    m = np.zeros((n, n, n), dtype=np.uint8)
    lo = ctypes.c_double(kmin)
    hi = ctypes.c_double(kmax)
    lib().or_kfilter(n, kind, k, ctypes.byref(lo), ctypes.byref(hi), binwidth, boxsize,
                     m.ctypes.data_as(_u8p))
    return m, lo.value, hi.value


def get_corners(mx, my, mz, binwidth, mmin=1e30, mmax=0.0):
    lo = ctypes.c_double(mmin)
    hi = ctypes.c_double(mmax)
    lib().or_get_corners(mx, my, mz, binwidth, ctypes.byref(lo), ctypes.byref(hi))
    return lo.value, hi.value


def calc_k3_min_and_max(k, kmin, kmax):
    k = np.asarray(k, dtype=np.float64).copy()
    lo = np.asarray(kmin, dtype=np.float64).copy()
    hi = np.asarray(kmax, dtype=np.float64).copy()
    lib().or_calc_k3_min_and_max(_d(k), _d(lo), _d(hi))
    return lo[2], hi[2]


def calc_k3(k1, k2, c):
    return lib().or_calc_k3(k1, k2, c)


def calc_k3min(k1, k2, theta, dcos):
    return lib().or_calc_k3min(k1, k2, theta, dcos)


def calc_k3max(k1, k2, theta, dcos):
    return lib().or_calc_k3max(k1, k2, theta, dcos)


def polyspectrum(spec, n, k, kmin, kmax, kbinwidth, kbinning_case, boxsize,
                 diagnostic_sums=False):
    """polyspectrum.c:51-161.  Returns (value, kmin, kmax, sums) where sums =
    (sum prod D, sum prod I) as complex numbers."""
    spec = np.ascontiguousarray(spec, dtype=np.complex128)
    N = spec.shape[0]
    k = np.asarray(k, dtype=np.float64).copy()
    lo = np.asarray(kmin, dtype=np.float64).copy()
    hi = np.asarray(kmax, dtype=np.float64).copy()
    sums = np.zeros(4)
    v = lib().or_polyspectrum(N, _d(spec.view(np.float64)), n, _d(k), _d(lo), _d(hi),
                              kbinwidth, kbinning_case, boxsize, int(diagnostic_sums), _d(sums))
    return v, lo, hi, (complex(sums[0], sums[1]), complex(sums[2], sums[3]))


def num_polygons(N, n, k, kmin, kmax, kbinwidth, kbinning_case, boxsize):
    """polyspectrum.c:163-252."""
    k = np.asarray(k, dtype=np.float64).copy()
    lo = np.asarray(kmin, dtype=np.float64).copy()
    hi = np.asarray(kmax, dtype=np.float64).copy()
    return lib().or_num_polygons(N, n, _d(k), _d(lo), _d(hi), kbinwidth, kbinning_case, boxsize)


def polygon_gradient(field, polys, kbinwidth, kbinning_case, boxsize, weights, phase=False):
    """Values of the polygons (dicts n, k, kmin, kmax) on a real N^3 field and
    sum_p weights[p] * d value_p / d field (N^3 float64), one polygon at a time."""
    f = np.ascontiguousarray(field, dtype=np.float64)
    N = f.shape[0]
    grad = np.zeros((N, N, N))
    vals = []
    for p, w in zip(polys, weights):
        k = np.zeros(3)
        lo = np.zeros(3)
        hi = np.zeros(3)
        n = int(p["n"])
        k[:n] = p["k"][:n]
        lo[:len(p.get("kmin", []))] = p.get("kmin", [])[:3]
        hi[:len(p.get("kmax", []))] = p.get("kmax", [])[:3]
        vals.append(lib().or_polygon_gradient(N, _d(f), n, _d(k), _d(lo), _d(hi), kbinwidth, kbinning_case,
                                              boxsize, int(bool(phase)), float(w), _d(grad)))
    return np.array(vals), grad


def direct_polygon_sum(spec, masks):
    """Direct lattice sum over closed polygons with m_i in masks[i]."""
    spec = np.ascontiguousarray(spec, dtype=np.complex128)
    N = spec.shape[0]
    ms = [np.ascontiguousarray(m, dtype=np.uint8) for m in masks]
    while len(ms) < 3:
        ms.append(ms[-1])
    out = np.zeros(2)
    c = lib().or_direct_polygon_sum(N, len(masks), _d(spec.view(np.float64)),
                                    *[m.ctypes.data_as(_u8p) for m in ms[:3]], _d(out))
    return complex(out[0], out[1]), int(c)
