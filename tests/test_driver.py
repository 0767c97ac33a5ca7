"""The drop-in driver `driver/polyspectrum iniFile.ini`.

CPU: the reference's error paths (main.c:53-58, parse_ini.h:16-22, kvectors.c:67-71,
run.c:40-44, input_field.c:65-69) fire before any GPU use, with the same messages.
GPU: the `.dat` files and stdout lines of run.c:94-222 for small grids, values
checked against the oracle (FP64): each row is the line the reference's
format prints for the oracle's values, byte for byte.
"""
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "driver", "polyspectrum")

INI = """[Grid]
gridsize = {N}
boxsize = {L}    #in h^-1 cMpc
gasInputsInDoublePrecision = {gdp}
ionInputsInDoublePrecision = {idp}
densityFile = {dens}
ionFile = {ion}

[Cosmology]
hubble_h = 0.7
omega_b = 0.0456
omega_m = 0.27
omega_l = 0.73
sigma8 = 0.8
Y = 0.25

[Polyspectrum]
whichField = {field}
# Possibilities: DENS, XHII, XHI_DENS
{phase}n = {n}
equilateral = {eq}
k1 = {k1}      #in h^-1 cMpc
k2 = {k2}      #in h^-1 cMpc
numValues = {nv}
theta = {theta}
kbinwidth = 1.
kbinningCase = {case}
cosThetaBinwidth = 0.05

[Output]
output_dir = {out}
output_basename = test
write_numpolygons = {wnp}
{extra}"""


def make_ini(tmp, **kw):
    d = dict(N=16, L=100.0, gdp=0, idp=1, dens=str(tmp / "dens.bin"), ion=str(tmp / "ion.bin"), field="DENS",
             phase="", n=3, eq=0, k1=0.25, k2=0.25, nv=4, theta=0.5, case="DEFAULT", out=str(tmp), wnp=1, extra="")
    d.update(kw)
    p = tmp / "run.ini"
    p.write_text(INI.format(**d))
    return p


def run(ini):
    return subprocess.run([DRIVER, str(ini)], capture_output=True, text=True, timeout=120)


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(DRIVER):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "driver")], check=True)


def test_usage_message():
    r = subprocess.run([DRIVER], capture_output=True, text=True)
    assert r.returncode != 0
    assert r.stdout == "cifog: (C)  - Use at own risk...\nUSAGE:\ncifog iniFile\n"


def test_missing_ini(tmp_path):
    r = run(tmp_path / "nope.ini")
    assert r.returncode != 0 and "FATAL:  Could not open" in r.stderr


def test_missing_key(tmp_path):
    ini = make_ini(tmp_path)
    ini.write_text(ini.read_text().replace("sigma8 = 0.8\n", ""))
    r = run(ini)
    assert r.returncode != 0
    assert r.stderr.strip() == "FATAL:  Could not get sigma8 from section Cosmology."


def test_bad_n_and_missing_dir(tmp_path):
    r = run(make_ini(tmp_path, n=1))
    assert r.returncode != 0 and "n>=2" in r.stderr
    r = run(make_ini(tmp_path, n=4))
    assert r.returncode != 0 and "Not implemented yet" in r.stdout
    r = run(make_ini(tmp_path, out=str(tmp_path / "missing")))
    assert r.returncode != 0 and "Directory for output does not exist!" in r.stderr


def test_unknown_field_and_inline_comment(tmp_path):
    # an inline comment stays part of a string value (parse_ini.c:786-799)
    r = run(make_ini(tmp_path, field="DENS   # density"))
    assert r.returncode != 0
    assert "Not supported option" in r.stderr and "Option: DENS   # density" in r.stdout


def test_missing_input_file(tmp_path):
    r = run(make_ini(tmp_path))
    assert r.returncode != 0 and "cannot read input grid" in r.stderr


# ---------------------------------------------------------------- GPU runs
def _ref_rows(field, N, L, n, eq, k1, k2, nv, theta, case, phase):
    """run.c:56-107 with the oracle (FP64)."""
    import oracle as O
    spec = O.forward_spectrum(field)
    if phase:
        spec = O.calc_phase(spec)
    PI = math.acos(-1.0)
    rows = []
    if n == 2:
        ks = [2. * PI / L * (i + 1) for i in range(N // 2)] if nv > 1 else [k1, k1]
        num = N // 2 if nv > 1 else 1
        lo = [ks[0], ks[1], 0.0]
        hi = [ks[0], ks[1], 0.0]
        for i in range(num):
            k = [ks[i], ks[i], 0.0]
            v, lo2, hi2, _ = O.polyspectrum(spec, 2, k, lo, hi, 1.0, case, L)
            npg = O.num_polygons(N, 2, k, lo2, hi2, 1.0, case, L)
            rows.append((ks[i], v, npg))
        return rows
    if eq:
        ks = [2. * PI / L * math.pow(N // 2, (i + 1) / nv) for i in range(nv)]
        ths = [PI / 3.]
    elif nv > 1:
        ths = [PI * i / nv for i in range(nv)]
        ks = [math.sqrt(k1 * k1 + k2 * k2 - 2. * k1 * k2 * math.cos(t)) for t in ths]
    else:
        ths = [theta]
        ks = [math.sqrt(k1 * k1 + k2 * k2 - 2. * k1 * k2 * math.cos(theta))]
    lo = [0.0, 0.0, 0.0]
    hi = [0.0, 0.0, 0.0]
    for i in range(len(ks)):
        kp = [ks[i]] * 3 if eq else [k1, k2, ks[i]]
        th = ths[0] if eq else ths[i]
        if case >= 2:
            c1 = min(1., math.cos(th) + 0.025)
            c2 = max(-1., math.cos(th) - 0.025)
            lo[2] = math.sqrt(kp[0] * kp[0] + kp[1] * kp[1] - 2. * kp[0] * kp[1] * c1)
            hi[2] = math.sqrt(kp[0] * kp[0] + kp[1] * kp[1] - 2. * kp[0] * kp[1] * c2)
        v, lo, hi, _ = O.polyspectrum(spec, 3, kp, lo, hi, 1.0, case, L)
        npg = O.num_polygons(N, 3, kp, lo, hi, 1.0, case, L)
        rows.append((th, ks[i], v, npg))
    return rows


CASES = {"DEFAULT": 0, "DERIVED_BINNING": 1, "GIVEN_BINNING": 2, "WATKINSON": 3}


@pytest.mark.gpu
@pytest.mark.parametrize("kw", [
    dict(n=3, nv=4, case="DEFAULT"),
    dict(n=3, nv=4, case="GIVEN_BINNING", extra="[GPU]\nprecision = FP64\nbatched = 0\n"),
    dict(n=3, nv=1, theta=1.2, case="DERIVED_BINNING"),
    dict(n=3, eq=1, nv=3, case="WATKINSON", wnp=0),
    dict(n=2, nv=1, k1=0.3, case="DEFAULT", phase="useOnlyPhase = 1\n"),
    dict(n=2, nv=2, case="DERIVED_BINNING", field="XHI_DENS"),
])
def test_driver_outputs(tmp_path, gpu_lib, kw):
    from polyspectrum_amd.synthetic import ionisation_field, lognormal_field
    N, L = 16, 100.0
    dens = lognormal_field(N, 21).astype(np.float32)
    ion = ionisation_field(N, 22)
    dens.tofile(tmp_path / "dens.bin")
    ion.tofile(tmp_path / "ion.bin")
    ini = make_ini(tmp_path, **kw)
    r = run(ini)
    assert r.returncode == 0, r.stderr
    n = kw.get("n", 3)
    field = dens.astype(np.float64) if kw.get("field", "DENS") == "DENS" else dens.astype(np.float64) * (1. - ion)
    rows = _ref_rows(field, N, L, n, kw.get("eq", 0), kw.get("k1", 0.25), kw.get("k2", 0.25), kw["nv"],
                     kw.get("theta", 0.5), CASES[kw["case"]], "useOnlyPhase = 1" in kw.get("phase", ""))
    k1, k2 = kw.get("k1", 0.25), kw.get("k2", 0.25)
    if n == 2:
        ending = "_ps.dat"
    elif kw.get("eq", 0):
        ending = "_bs_equilateral.dat"
    elif kw["nv"] == 1:
        ending = "_bs_k_%4.2e_%4.2e_theta_%4.2e.dat" % (k1, k2, kw.get("theta", 0.5))
    else:
        ending = "_bs_k_%4.2e_%4.2e.dat" % (k1, k2)
    path = tmp_path / ("test" + ending)
    assert ("FILENAME : %s/test%s" % (tmp_path, ending)) in r.stdout
    lines = path.read_text().splitlines()
    if n == 2 or kw.get("eq", 0):
        assert lines[0] == "# k [h^-1 Mpc]\t Polyspectrum_%d(k)" % n
    else:
        assert lines[0] == "# theta [rad]\t k [h^-1 Mpc]\t Polyspectrum_%d(k)" % n
    body = [list(map(float, l.split("\t"))) for l in lines[1:]]
    assert len(body) == len(rows)
    wnp = kw.get("wnp", 1)
    for line, got, ref in zip(lines[1:], body, rows):
        ref = list(ref) if wnp else list(ref)[:-1]
        if n == 2 or kw.get("eq", 0):
            ref = ref if n == 2 else [ref[1], ref[2]] + ([ref[3]] if wnp else [])
        assert len(got) == len(ref)
        np.testing.assert_allclose(got, ref, rtol=2e-6, atol=1e-300)
        # the FP64 path prints the reference's rows (run.c:128-222, "%e" joined by tabs) byte for byte
        assert line == "\t".join("%e" % x for x in ref), (line, ref)
    assert "kbinningCase = %d" % CASES[kw["case"]] in r.stdout
    assert "Execution took" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("quirks", [1, 0])
def test_driver_power_spectrum_given_binning(tmp_path, gpu_lib, quirks):
    """n = 2, N/2 bins, GIVEN_BINNING.  The reference keeps the last leg's
    kmin/kmax at k[1] (set once, kvectors.c:92-97; run.c:56-83 updates them for
    n = 3 only), so every bin but k[1] multiplies two disjoint shells: no
    polygon closes and the value is a ratio of roundoff.  referenceQuirks = 1
    (default) reproduces that: bin k[1] equals the oracle and the other bins
    count no polygons.  referenceQuirks = 0 updates the range with k[i]: every
    bin is the DEFAULT power spectrum, as the oracle gives it."""
    import oracle as O
    from polyspectrum_amd.synthetic import lognormal_field
    N, L = 16, 100.0
    dens = lognormal_field(N, 31).astype(np.float32)
    dens.tofile(tmp_path / "dens.bin")
    extra = "[GPU]\nprecision = FP64\nreferenceQuirks = %d\n" % quirks
    r = run(make_ini(tmp_path, n=2, nv=2, case="GIVEN_BINNING", extra=extra))
    assert r.returncode == 0, r.stderr
    body = [list(map(float, l.split("\t"))) for l in (tmp_path / "test_ps.dat").read_text().splitlines()[1:]]
    assert len(body) == N // 2
    spec = O.forward_spectrum(dens.astype(np.float64))
    PI = math.acos(-1.0)
    ks = [2. * PI / L * (i + 1) for i in range(N // 2)]
    for i, row in enumerate(body):
        assert abs(row[0] - ks[i]) <= 1e-6 * ks[i]
        if quirks and i != 1:
            assert abs(row[2]) < 1e-6, row  # no closed polygon: Npolygons = 0
            continue
        lo = [ks[i], ks[1] if quirks else ks[i], 0.0]
        v, lo2, hi2, _ = O.polyspectrum(spec, 2, [ks[i], ks[i], 0.0], lo, list(lo), 1.0, 2, L)
        npg = O.num_polygons(N, 2, [ks[i], ks[i], 0.0], lo2, hi2, 1.0, 2, L)
        np.testing.assert_allclose(row[1:], [v, npg], rtol=2e-6)
