"""ps_direct_check_static: the transform-free check with every device buffer
in the library's static scratch array (no device allocation), called
through the C ABI and compared with ps_direct_check bit for bit: two calls
in a row, the second on smaller shells, then a set whose shells exceed the
scratch (PS_ERR_SCRATCH, decided on the host before any launch)."""
import ctypes
import math

import numpy as np
import pytest

from polyspectrum_amd.synthetic import lognormal_field

pytestmark = pytest.mark.gpu
PI = math.acos(-1.0)


def _tris(k1, k2, nth, case, dcos=0.05):
    out = []
    for i in range(nth):
        th = PI * i / nth
        k3 = math.sqrt(k1 * k1 + k2 * k2 - 2. * k1 * k2 * math.cos(th))
        lo, hi = [k1, k2, 0.0], [k1, k2, 0.0]
        if case == 2:
            lo[2] = math.sqrt(k1 * k1 + k2 * k2 - 2. * k1 * k2 * min(1., math.cos(th) + 0.5 * dcos))
            hi[2] = math.sqrt(k1 * k1 + k2 * k2 - 2. * k1 * k2 * max(-1., math.cos(th) - 0.5 * dcos))
        out.append({"n": 3, "k": [k1, k2, k3], "kmin": lo, "kmax": hi})
    return out


def _call(P, ctx, polys, bw, case, static):
    arr = P._polygon_array(polys)
    n = len(polys)
    v, npg, s = np.zeros(n), np.zeros(n), np.zeros(2 * n)
    dp = ctypes.POINTER(ctypes.c_double)
    fn = P.load_library().ps_direct_check_static if static else P.load_library().ps_direct_check
    st = fn(ctx._h, arr, n, float(bw), int(case), v.ctypes.data_as(dp), npg.ctypes.data_as(dp),
            s.ctypes.data_as(dp))
    return st, v, npg, s, arr


@pytest.mark.parametrize("prec_name", ["PREC_FP64", "PREC_MIXED"])
def test_static_check_equals_allocating_check(gpu_lib, prec_name):
    P = gpu_lib
    N, L = 128, 100.0
    kf = 2 * PI / L
    f = lognormal_field(N, 71, sigma=0.8, alpha=1.5)
    big = _tris(12.0 * kf, 12.0 * kf, 8, 2) + [{"n": 2, "k": [9 * kf] * 2, "kmin": [9 * kf] * 2,
                                                "kmax": [9 * kf] * 2}]
    small = _tris(3.0 * kf, 4.0 * kf, 5, 0)
    derived = _tris(6.0 * kf, 6.0 * kf, 4, 1)
    with P.Context(N, L, getattr(P, prec_name)) as ctx:
        ctx.load_field(f)
        for polys, case in ((big, 2), (small, 0), (derived, 1)):
            st0, v0, n0, s0, a0 = _call(P, ctx, polys, 1.0, case, False)
            st1, v1, n1, s1, a1 = _call(P, ctx, polys, 1.0, case, True)
            assert st0 == 0 and st1 == 0
            np.testing.assert_array_equal(v1, v0)
            np.testing.assert_array_equal(n1, n0)
            np.testing.assert_array_equal(s1, s0)
            assert np.all(n0 > 0) and np.all(np.isfinite(v0))
            # DERIVED kmin/kmax written back alike
            for p0, p1 in zip(a0, a1):
                assert list(p0.kmin) == list(p1.kmin) and list(p0.kmax) == list(p1.kmax)


def test_static_check_rejects_what_exceeds_scratch(gpu_lib):
    """Two shells of radius 250 cells (about 785k cells each: 2 x 785k x 20 B
    of cells and values plus 4 MiB of partial sums, 35.6 MB) need more than
    the 32 MiB array: PS_ERR_SCRATCH, and the same context still serves a
    fitting call."""
    P = gpu_lib
    N, L = 512, 512.0
    kf = 2 * PI / L
    f = lognormal_field(64, 3)
    with P.Context(N, L, P.PREC_MIXED) as ctx:
        g = np.zeros((N, N, N), dtype=np.float32)
        g[:64, :64, :64] = f
        ctx.load_field(g)
        k = 250.0 * kf
        st, _, _, _, _ = _call(P, ctx, [{"n": 3, "k": [k, k, k], "kmin": [0.0] * 3, "kmax": [0.0] * 3}],
                               1.0, 0, True)
        assert st == -7
        assert b"scratch" in P.load_library().ps_strerror(-7)
        st, v, npg, _, _ = _call(P, ctx, _tris(5.0 * kf, 5.0 * kf, 3, 0), 1.0, 0, True)
        assert st == 0 and np.all(npg > 0)
