"""The hand-written adjoint (ps_batch_adjoint) against
  * an independent numpy FP64 gradient that differentiates every polygon
    separately (per-theta filtered fields, no combined mask):
    dS/d delta = sum_i L_i[prod_{j != i} D_j], L_X[f] = N^3 ifftn(F_X fftn(f)),
    bar 1e-10 relative for the FP64 context;
  * central finite differences of L = sum_t (B_t - B*_t)^2 through the FP64
    path at random voxels (eps = 1e-3), bar 1e-4 relative;
  * the mixed-precision adjoint against the FP64 one, bar 1e-4 relative;
  * the torch.autograd binding.
"""
import math

import numpy as np
import pytest

from polyspectrum_amd.synthetic import lognormal_field

pytestmark = pytest.mark.gpu
PI = math.acos(-1.0)


def _lat(n):
    i = np.arange(n)
    return np.where(i < n // 2, i, i - n)


def _absm(N):
    l = _lat(N)
    return np.sqrt((l[:, None, None] ** 2 + l[None, :, None] ** 2 + l[None, None, :] ** 2).astype(np.float64))


def shell(N, L, k, bw=1.0):
    return (np.abs(k / (2. * PI) * L - _absm(N)) <= 0.5 * bw).astype(np.float64)


def rshell(N, L, kmin, kmax, bw=1.0):
    m = _absm(N)
    return (((kmin / (2. * PI) * L - m) <= 0.5 * bw) & ((m - kmax / (2. * PI) * L) <= 0.5 * bw)).astype(np.float64)


def triangles(k1, k2, nth, case, dcos=0.05):
    out = []
    for i in range(nth):
        th = PI * i / nth
        k3 = math.sqrt(k1 * k1 + k2 * k2 - 2. * k1 * k2 * math.cos(th))
        lo = [k1, k2, 0.0]
        hi = [k1, k2, 0.0]
        if case == 2:
            lo[2] = math.sqrt(k1 * k1 + k2 * k2 - 2. * k1 * k2 * min(1., math.cos(th) + 0.5 * dcos))
            hi[2] = math.sqrt(k1 * k1 + k2 * k2 - 2. * k1 * k2 * max(-1., math.cos(th) - 0.5 * dcos))
        out.append({"n": 3, "k": [k1, k2, k3], "kmin": lo, "kmax": hi})
    return out


def masks_of(p, N, L, case):
    if p["n"] == 2:
        return [shell(N, L, p["k"][0]), shell(N, L, p["k"][1])]
    ms = [shell(N, L, p["k"][0]), shell(N, L, p["k"][1])]
    ms.append(rshell(N, L, p["kmin"][2], p["kmax"][2]) if case == 2 else shell(N, L, p["k"][2]))
    return ms


def numpy_values_and_grad(field, polys, L, case, weights):
    """B_p and sum_p w_p dB_p/d delta, one polygon at a time."""
    N = field.shape[0]
    d = field.astype(np.float64)
    dk = np.fft.fftn(d)
    Lx = lambda F, f: (np.fft.ifftn(F * np.fft.fftn(f)) * N ** 3).real
    vals, grad = [], np.zeros_like(d)
    V = L ** 3
    for p, w in zip(polys, weights):
        ms = masks_of(p, N, L, case)
        n = len(ms)
        D = [(np.fft.ifftn(F * dk) * N ** 3).real for F in ms]
        I = [(np.fft.ifftn(F) * N ** 3).real for F in ms]
        S = np.prod(D, axis=0).sum()
        C = np.prod(I, axis=0).sum()
        cn = V ** (n - 1) * float(N) ** (-3 * n)
        vals.append(cn * S / C)
        g = np.zeros_like(d)
        for i in range(n):
            g += Lx(ms[i], np.prod([D[j] for j in range(n) if j != i], axis=0))
        grad += w * cn / C * g
    return np.array(vals), grad


@pytest.mark.parametrize("N,L,case,pairs", [
    (32, 100.0, 0, [(5.0, 5.0)]),   # prefix squared (config 4's shape)
    (32, 100.0, 2, [(4.0, 6.0)]),   # two distinct prefix legs, GIVEN binning
])
def test_adjoint_fp64_vs_numpy(gpu_lib, N, L, case, pairs):
    P = gpu_lib
    field = lognormal_field(N, 404, sigma=0.8, alpha=1.5)
    kf = 2 * PI / L
    polys = []
    for m1, m2 in pairs:
        polys += triangles(m1 * kf, m2 * kf, 8, case)
    # power-spectrum bins (n = 2, squared legs) in the same batch
    polys += [{"n": 2, "k": [m * kf, m * kf], "kmin": [m * kf] * 2, "kmax": [m * kf] * 2} for m in (2.0, 3.0, 6.0)]
    rng = np.random.default_rng(1)
    w = rng.standard_normal(len(polys))
    ref_vals, ref_grad = numpy_values_and_grad(field, polys, L, case, w)
    with P.Context(N, L, P.PREC_FP64) as ctx:
        ctx.load_field(field)
        b = ctx.batch(polys, 1.0, case)
        vals, _, _ = b.execute()
        g = b.adjoint(w)
        b.close()
    np.testing.assert_allclose(vals, ref_vals, rtol=1e-10)
    err = np.abs(g - ref_grad).max() / np.abs(ref_grad).max()
    print("adjoint vs numpy:", err)
    assert err < 1e-10


def _loss_fn(P, ctx, polys, case, target):
    b = ctx.batch(polys, 1.0, case)

    def loss(field):
        ctx.load_field(field)
        v, _, _ = b.execute()
        return float(((v - target) ** 2).sum()), v
    return loss, b


def test_adjoint_finite_differences(gpu_lib):
    """Config-4 shape: L = sum_t (B_t - B*_t)^2, k1 = k2, 16 theta bins."""
    P = gpu_lib
    N, L = 48, 150.0
    kf = 2 * PI / L
    polys = triangles(6.0 * kf, 6.0 * kf, 16, 0)
    field = lognormal_field(N, 7, sigma=0.8, alpha=1.5).astype(np.float64)
    other = lognormal_field(N, 8, sigma=0.8, alpha=1.5).astype(np.float64)
    with P.Context(N, L, P.PREC_FP64) as ctx:
        ctx.load_field(other)
        bt = ctx.batch(polys, 1.0, 0)
        target, _, _ = bt.execute()
        bt.close()
        loss, b = _loss_fn(P, ctx, polys, 0, target)
        L0, v = loss(field)
        g = b.adjoint(2.0 * (v - target))
        rng = np.random.default_rng(3)
        eps = 1e-3
        errs = []
        for _ in range(16):
            x = tuple(rng.integers(0, N, 3))
            fp = field.copy()
            fp[x] += eps
            fm = field.copy()
            fm[x] -= eps
            fd = (loss(fp)[0] - loss(fm)[0]) / (2 * eps)
            errs.append(abs(fd - g[x]) / max(abs(g).max() * 1e-3, abs(g[x])))
        b.close()
    print("FD rel errors:", max(errs))
    assert max(errs) < 1e-4


def test_adjoint_mixed_vs_fp64(gpu_lib):
    P = gpu_lib
    N, L = 64, 100.0
    kf = 2 * PI / L
    polys = triangles(6.0 * kf, 6.0 * kf, 16, 0)
    field = lognormal_field(N, 9, sigma=0.8, alpha=1.5)
    w = np.random.default_rng(2).standard_normal(len(polys))
    g = {}
    for prec in (P.PREC_FP64, P.PREC_MIXED):
        with P.Context(N, L, prec) as ctx:
            ctx.load_field(field)
            b = ctx.batch(polys, 1.0, 0)
            b.execute()
            g[prec] = b.adjoint(w)
            b.close()
    err = np.abs(g[P.PREC_MIXED] - g[P.PREC_FP64]).max() / np.abs(g[P.PREC_FP64]).max()
    print("mixed adjoint vs fp64:", err)
    assert err < 1e-4


def test_autograd_binding(gpu_lib):
    import torch
    from polyspectrum_amd.autograd import Polyspectra
    P = gpu_lib
    N, L = 32, 100.0
    kf = 2 * PI / L
    polys = triangles(5.0 * kf, 5.0 * kf, 6, 0)
    field = torch.tensor(lognormal_field(N, 11), dtype=torch.float64, device="cuda", requires_grad=True)
    with P.Context(N, L, P.PREC_FP64) as ctx:
        b = ctx.batch_for_field(polys, field.detach(), 1.0, 0)
        vals = Polyspectra.apply(field, b)
        loss = (vals ** 2).sum()
        loss.backward()
        ref_vals, ref_grad = numpy_values_and_grad(field.detach().cpu().numpy(), polys, L, 0,
                                                   2 * vals.detach().cpu().numpy())
        b.close()
    np.testing.assert_allclose(vals.detach().cpu().numpy(), ref_vals, rtol=1e-10)
    gg = field.grad.cpu().numpy()
    assert np.abs(gg - ref_grad).max() / np.abs(ref_grad).max() < 1e-10


# ------------------------------------------------------------------ useOnlyPhase
@pytest.mark.parametrize("case,pairs", [(0, [(5.0, 5.0)]), (2, [(4.0, 6.0)])])
def test_adjoint_phase_only_vs_oracle(gpu_lib, case, pairs):
    """Phase-only fields (p = z/|z|, utils_fftw.c:78-93): the FP64 adjoint
    against the oracle's CPU gradient (which pulls R = sum_i m_i B(Q_i) back
    through the normalisation), bar 1e-10 of max |grad|; power spectra are
    constant (V N^-6) and have no gradient."""
    import oracle as O
    P = gpu_lib
    N, L = 32, 100.0
    field = lognormal_field(N, 505, sigma=0.8, alpha=1.5).astype(np.float64)
    kf = 2 * PI / L
    polys = []
    for m1, m2 in pairs:
        polys += triangles(m1 * kf, m2 * kf, 6, case)
    polys += [{"n": 2, "k": [3 * kf] * 2, "kmin": [3 * kf] * 2, "kmax": [3 * kf] * 2}]
    w = np.random.default_rng(4).standard_normal(len(polys))
    ref_vals, ref_grad = O.polygon_gradient(field, polys, 1.0, case, L, w, phase=True)
    with P.Context(N, L, P.PREC_FP64) as ctx:
        ctx.load_field(field)
        ctx.set_phase_only(1)
        b = ctx.batch(polys, 1.0, case)
        vals, _, _ = b.execute()
        g = b.adjoint(w)
        b.close()
    np.testing.assert_allclose(vals, ref_vals, rtol=1e-10)
    err = np.abs(g - ref_grad).max() / np.abs(ref_grad).max()
    print("phase-only adjoint vs oracle:", err)
    assert err < 1e-10


def test_adjoint_phase_only_finite_differences(gpu_lib):
    """Central differences (eps = 1e-3) of L = sum_t (B_t - B*_t)^2 through the
    FP64 phase-only path at random voxels, bar 1e-4 relative."""
    P = gpu_lib
    N, L = 32, 100.0
    kf = 2 * PI / L
    polys = triangles(5.0 * kf, 5.0 * kf, 8, 0)
    field = lognormal_field(N, 17, sigma=0.8, alpha=1.5).astype(np.float64)
    other = lognormal_field(N, 18, sigma=0.8, alpha=1.5).astype(np.float64)
    with P.Context(N, L, P.PREC_FP64) as ctx:
        ctx.set_phase_only(1)
        ctx.load_field(other)
        bt = ctx.batch(polys, 1.0, 0)
        target, _, _ = bt.execute()
        bt.close()
        loss, b = _loss_fn(P, ctx, polys, 0, target)
        _, v = loss(field)
        g = b.adjoint(2.0 * (v - target))
        rng = np.random.default_rng(6)
        errs = []
        for _ in range(12):
            x = tuple(rng.integers(0, N, 3))
            fp, fm = field.copy(), field.copy()
            fp[x] += 1e-3
            fm[x] -= 1e-3
            fd = (loss(fp)[0] - loss(fm)[0]) / 2e-3
            errs.append(abs(fd - g[x]) / abs(g[x]))
        b.close()
    print("phase-only FD rel errors:", max(errs))
    assert max(errs) < 1e-4


# ------------------------------------------------------------------ 512^3 workload
def _config4(N, L):
    k = 0.4
    return [{"n": 3, "k": [k, k, math.sqrt(2 * k * k - 2 * k * k * math.cos(PI * i / 16))],
             "kmin": [0.0] * 3, "kmax": [0.0] * 3} for i in range(16)]


@pytest.fixture(scope="module")
def config4_fields():
    N = 512
    return (lognormal_field(N, 4000, sigma=0.8, alpha=1.5), lognormal_field(N, 4001, sigma=0.8, alpha=1.5))


def test_config4_finite_differences_512(gpu_lib, config4_fields):
    """L = sum_t (B_t - B*_t)^2 over 16 triangles (k1 = k2 = 0.4 h/Mpc, 16 theta
    bins, DEFAULT binning) of a 512^3 log-normal field, L = 400 Mpc/h, B* from a
    second seeded field: the FP64 adjoint against central differences
    (eps = 1e-3) at 64 random voxels, each within 1e-4 of the gradient there.
    The perturbed fields are uploaded from device memory."""
    import torch
    P = gpu_lib
    N, L = 512, 400.0
    field, other = config4_fields
    polys = _config4(N, L)
    with P.Context(N, L, P.PREC_FP64) as ctx:
        ctx.load_field(other)
        b = ctx.batch(polys, 1.0, 0)
        target, _, _ = b.execute()
        ctx.load_field(field)
        v, _, _ = b.execute()
        g = b.adjoint(2.0 * (v - target))
        t = torch.from_numpy(field.astype(np.float64)).cuda()
        rng = np.random.default_rng(44)
        errs = []
        for _ in range(64):
            x = tuple(int(i) for i in rng.integers(0, N, 3))
            x0 = float(t[x])
            losses = []
            for xv in (x0 + 1e-3, x0 - 1e-3):
                t[x] = xv
                ctx.load_field(t)
                vv, _, _ = b.execute()
                losses.append(float(((vv - target) ** 2).sum()))
            t[x] = x0
            fd = (losses[0] - losses[1]) / 2e-3
            errs.append(abs(fd - g[x]) / abs(g[x]))
        b.close()
    errs = np.array(errs)
    print(f"512^3 FD: max rel err {errs.max():.3e}, median {np.median(errs):.3e}")
    assert errs.max() < 1e-4


def test_config4_crop_vs_cpu_gradient(gpu_lib, config4_fields):
    """The same setup on the 128^3 crop (L = 100 Mpc/h, same cell size): the
    FP64 adjoint against the oracle's CPU FP64 gradient (one polygon at a
    time, no combined mask), bar 1e-10 of max |grad|."""
    import oracle as O
    P = gpu_lib
    N, L = 128, 100.0
    field = np.ascontiguousarray(config4_fields[0][:N, :N, :N]).astype(np.float64)
    other = np.ascontiguousarray(config4_fields[1][:N, :N, :N]).astype(np.float64)
    polys = _config4(N, L)
    with P.Context(N, L, P.PREC_FP64) as ctx:
        ctx.load_field(other)
        b = ctx.batch(polys, 1.0, 0)
        target, _, _ = b.execute()
        ctx.load_field(field)
        v, _, _ = b.execute()
        w = 2.0 * (v - target)
        g = b.adjoint(w)
        b.close()
    ref_vals, ref = O.polygon_gradient(field, polys, 1.0, 0, L, w)
    np.testing.assert_allclose(v, ref_vals, rtol=1e-10)
    err = np.abs(g - ref).max() / np.abs(ref).max()
    print("128^3 crop: adjoint vs CPU gradient", err)
    assert err < 1e-10


def test_config4_mixed_vs_fp64_512(gpu_lib, config4_fields):
    """The mixed-precision adjoint (FP32 transforms of the float32 field) at
    512^3 against the FP64 one, bar 1e-4 of max |grad|."""
    P = gpu_lib
    N, L = 512, 400.0
    field, other = config4_fields
    polys = _config4(N, L)
    g = {}
    for prec in (P.PREC_FP64, P.PREC_MIXED):
        with P.Context(N, L, prec) as ctx:
            ctx.load_field(other)
            b = ctx.batch(polys, 1.0, 0)
            target, _, _ = b.execute()
            ctx.load_field(field)
            v, _, _ = b.execute()
            if prec == P.PREC_FP64:
                w = 2.0 * (v - target)
            g[prec] = b.adjoint(w)
            b.close()
    err = np.abs(g[P.PREC_MIXED] - g[P.PREC_FP64]).max() / np.abs(g[P.PREC_FP64]).max()
    print("512^3 mixed adjoint vs FP64:", err)
    assert err < 1e-4
