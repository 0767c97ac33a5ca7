"""MI355X (gfx950) polyspectrum estimator -- Python mirror of the C ABI.

The product is the shared library ``libpolyspectrum_gpu.so`` built in-tree
from ``csrc/`` (C ABI declared in ``include/polyspectrum_gpu.h``).  This module
binds it with ctypes and mirrors the reference's estimator interface
(src/polyspectrum.h:7-8, src/run.c): same function names, argument meaning
and normalisation.  There is no CPU fallback: without the built library, or
without a GPU, every entry point raises.
"""
import ctypes
import os
import weakref

import numpy as np

__all__ = [
    "PREC_FP64", "PREC_MIXED", "BINNING", "Polygon", "Context", "Batch", "GPUError",
    "polyspectrum", "num_polygons", "library_path", "load_library", "device_count", "build", "MultiBatch",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_NAME = "libpolyspectrum_gpu.so"

PREC_MIXED = 0
PREC_FP64 = 1
# kbinningCase strings of confObj.c:98-105
BINNING = {"DEFAULT": 0, "DERIVED_BINNING": 1, "GIVEN_BINNING": 2, "WATKINSON": 3}

_lib = None


class GPUError(RuntimeError):
    pass


class Polygon(ctypes.Structure):
    """ps_polygon: n legs with k, kmin, kmax (h/Mpc) and the field slot."""
    _fields_ = [("n", ctypes.c_int), ("slot", ctypes.c_int),
                ("k", ctypes.c_double * 3), ("kmin", ctypes.c_double * 3),
                ("kmax", ctypes.c_double * 3)]


class LaunchParams(ctypes.Structure):
    """ps_launch_params: persistent grid size, work-item size, G2/Fx tile
    divisor, schedule of the waves (0: one after another; 1: side by side on
    up to 4 branch streams; 2 .. 4: on at most that many) and the method
    (METHOD_FOURIER / METHOD_DIRECT)."""
    _fields_ = [("grid_cap", ctypes.c_int), ("chunk_tiles", ctypes.c_int), ("g2_div", ctypes.c_int),
                ("overlap", ctypes.c_int), ("method", ctypes.c_int)]


METHOD_FOURIER = 0  # prefix product forward transformed once, shell sums per last leg
METHOD_DIRECT = 1   # every leg inverse transformed, real-space products (default)
MAX_STAGES = 6


def _params_dict(lp):
    return {"grid_cap": lp.grid_cap, "chunk_tiles": lp.chunk_tiles, "g2_div": lp.g2_div, "overlap": lp.overlap,
            "method": lp.method}


def _polygon_array(polygons):
    arr = (Polygon * len(polygons))()
    for i, p in enumerate(polygons):
        if isinstance(p, Polygon):
            arr[i] = p
            continue
        arr[i].n = int(p["n"])
        arr[i].slot = int(p.get("slot", 0))
        for j in range(3):
            arr[i].k[j] = float(p["k"][j]) if j < len(p["k"]) else 0.0
            arr[i].kmin[j] = float(p.get("kmin", [0, 0, 0])[j]) if j < len(p.get("kmin", [])) else 0.0
            arr[i].kmax[j] = float(p.get("kmax", [0, 0, 0])[j]) if j < len(p.get("kmax", [])) else 0.0
    return arr


def library_path():
    """The in-tree build, or the file PS_LIBRARY names (another build of the
    same sources, e.g. for timing two builds on one machine)."""
    return os.environ.get("PS_LIBRARY") or os.path.join(_HERE, _LIB_NAME)


def build():
    """Compile the HIP library for gfx950 (hipcc cross-compiles without a GPU)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


def load_library():
    """Load the built library; raises if it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise GPUError(f"{path} is not built; run polyspectrum_amd.build() (make -C polyspectrum_amd)")
    # One HIP runtime per process: torch ships its own libamdhip64.so.7.  When
    # torch is importable it is loaded first, and the library's dependency on
    # that soname then resolves to torch's copy instead of a second runtime
    # (which would leave torch without a device).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(path)
    L.ps_strerror.restype = ctypes.c_char_p
    L.ps_strerror.argtypes = [ctypes.c_int]
    L.ps_device_count.restype = ctypes.c_int
    L.ps_grid_supported.argtypes = [ctypes.c_int]
    L.ps_ctx_create.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]
    L.ps_ctx_destroy.argtypes = [_vp]
    L.ps_load_field.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int]
    L.ps_set_phase_only.argtypes = [_vp, ctypes.c_int]
    L.ps_set_method.argtypes = [_vp, ctypes.c_int]
    L.ps_set_refinement.argtypes = [_vp, ctypes.c_double]
    L.ps_batch_refinement.argtypes = [_vp, _ip, _dp]
    L.ps_polyspectrum.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, ctypes.c_double,
                                  ctypes.c_int, _dp, _dp]
    L.ps_num_polygons.argtypes = [_vp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, _dp]
    L.ps_batch_create.argtypes = [_vp, ctypes.POINTER(Polygon), ctypes.c_int, ctypes.c_double,
                                  ctypes.c_int, ctypes.POINTER(_vp)]
    L.ps_batch_execute.argtypes = [_vp, _dp, _dp, _dp]
    L.ps_batch_last_ms.restype = ctypes.c_double
    L.ps_batch_last_ms.argtypes = [_vp]
    L.ps_batch_model_bytes.restype = ctypes.c_double
    L.ps_batch_model_bytes.argtypes = [_vp]
    L.ps_batch_destroy.argtypes = [_vp]
    L.ps_forward_again.argtypes = [_vp, ctypes.c_int]
    L.ps_get_spectrum.argtypes = [_vp, ctypes.c_int, _vp]
    L.ps_batch_profile.argtypes = [_vp, ctypes.c_int, _dp, _dp]
    L.ps_batch_stage_count.argtypes = [_vp]
    L.ps_batch_stage_name.restype = ctypes.c_char_p
    L.ps_batch_stage_name.argtypes = [_vp, ctypes.c_int]
    L.ps_batch_wave_count.restype = ctypes.c_int
    L.ps_batch_wave_count.argtypes = [_vp]
    L.ps_batch_wave_info.restype = ctypes.c_int
    L.ps_batch_wave_info.argtypes = [_vp, ctypes.c_int, _ip, _ip, _ip, _dp]
    L.ps_batch_adjoint.argtypes = [_vp, ctypes.c_int, _dp, _vp, ctypes.c_int, ctypes.c_int]
    L.ps_forward_profile.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp]
    L.ps_batch_results.argtypes = [_vp, _dp, _dp, _dp]
    L.ps_batch_create_ex.argtypes = [_vp, ctypes.POINTER(Polygon), ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                     ctypes.POINTER(LaunchParams), ctypes.POINTER(_vp)]
    L.ps_batch_autotune.argtypes = [_vp, ctypes.POINTER(Polygon), ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                    ctypes.c_int, ctypes.POINTER(LaunchParams), _dp]
    L.ps_batch_params.argtypes = [_vp, ctypes.POINTER(LaunchParams)]
    L.ps_direct_check.argtypes = [_vp, ctypes.POINTER(Polygon), ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                  _dp, _dp, _dp]
    L.ps_direct_check_static.argtypes = [_vp, ctypes.POINTER(Polygon), ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                         _dp, _dp, _dp]
    L.ps_forward_variant_count.argtypes = [_vp]
    L.ps_set_forward_variant.argtypes = [_vp, ctypes.c_int]
    L.ps_forward_autotune.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _ip, _dp]
    L.ps_condition_numbers.argtypes = [_vp, ctypes.POINTER(Polygon), ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                       _dp]
    L.ps_multi_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(_vp)]
    L.ps_multi_execute.argtypes = [_vp]
    L.ps_multi_last_ms.restype = ctypes.c_double
    L.ps_multi_last_ms.argtypes = [_vp]
    L.ps_multi_destroy.argtypes = [_vp]
    L.polyspectrum.restype = ctypes.c_double
    L.polyspectrum.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _dp, _dp, _dp,
                               ctypes.c_double, ctypes.c_int, ctypes.c_double]
    L.num_polygons.restype = ctypes.c_double
    L.num_polygons.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                               ctypes.c_double, ctypes.c_int, ctypes.c_double]
    _lib = L
    return L


def _check(status):
    if status != 0:
        msg = load_library().ps_strerror(status).decode()
        raise GPUError(f"polyspectrum_gpu status {status}: {msg}")


def device_count():
    return load_library().ps_device_count()


def _darr(a, n=3):
    a = np.zeros(n) if a is None else np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


class Context:
    """A field (or several realisation slots) resident on the GPU.

    precision: PREC_FP64 (plain parity path) or PREC_MIXED (FP32 FFTs, FP64
    compensated reductions and normalisation)."""

    def __init__(self, nbins, boxsize, precision=PREC_FP64, nslots=1):
        L = load_library()
        h = _vp()
        _check(L.ps_ctx_create(int(nbins), float(boxsize), int(precision), int(nslots), ctypes.byref(h)))
        self._h = h
        self.nbins = int(nbins)
        self.boxsize = float(boxsize)
        self.precision = precision
        self.nslots = nslots
        self._dependents = weakref.WeakSet()  # batches (and multi-batches over them) of this context

    def close(self):
        """Destroys the context, after every batch of it still open (a
        batch must not outlive its context: ps_batch_destroy reads it)."""
        if getattr(self, "_h", None):
            deps = list(getattr(self, "_dependents", ()))
            for d in deps:
                if isinstance(d, MultiBatch):
                    d.close()
            for d in deps:
                d.close()
            load_library().ps_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def load_field(self, field, slot=0):
        """get_FT_field (input_field.c:40-81): real N^3 float32/float64 field
        (numpy array, or a torch CUDA tensor) -> resident spectrum."""
        try:
            import torch
            if isinstance(field, torch.Tensor):
                if not field.is_cuda:
                    field = field.numpy()
                else:
                    t = field.contiguous()
                    assert t.numel() == self.nbins ** 3 and t.dtype in (torch.float32, torch.float64)
                    _check(load_library().ps_load_field(self._h, slot, ctypes.c_void_p(t.data_ptr()),
                                                        int(t.dtype == torch.float64), 1))
                    return
        except ImportError:
            pass
        a = np.ascontiguousarray(field)
        if a.dtype not in (np.float32, np.float64):
            a = a.astype(np.float64)
        assert a.size == self.nbins ** 3
        _check(load_library().ps_load_field(self._h, slot, a.ctypes.data_as(_vp), int(a.dtype == np.float64), 0))

    def spectrum(self, slot=0):
        """The resident half spectrum as a numpy array [kz][kx][ky] (kz <= N/2)."""
        N = self.nbins
        dt = np.complex128 if self.precision == PREC_FP64 else np.complex64
        out = np.empty((N // 2 + 1, N, N), dtype=dt)
        _check(load_library().ps_get_spectrum(self._h, slot, out.ctypes.data_as(_vp)))
        return out

    def forward_profile(self, reps=5, slot=0):
        """(mean ms, algorithmic bytes) of one forward R2C transform."""
        ms = ctypes.c_double()
        by = ctypes.c_double()
        _check(load_library().ps_forward_profile(self._h, slot, reps, ctypes.byref(ms), ctypes.byref(by)))
        return ms.value, by.value

    def forward_autotune(self, reps=5, slot=0):
        """Times every forward-transform variant (radix plan, tile width) on
        the slot's last host upload; keeps and returns (variant, ms)."""
        v = ctypes.c_int()
        ms = ctypes.c_double()
        _check(load_library().ps_forward_autotune(self._h, slot, reps, ctypes.byref(v), ctypes.byref(ms)))
        return v.value, ms.value

    def set_forward_variant(self, variant):
        _check(load_library().ps_set_forward_variant(self._h, int(variant)))

    @property
    def forward_variants(self):
        return load_library().ps_forward_variant_count(self._h)

    def forward_again(self, slot=0):
        _check(load_library().ps_forward_again(self._h, slot))

    def set_phase_only(self, flag):
        _check(load_library().ps_set_phase_only(self._h, int(flag)))

    def set_method(self, method):
        """METHOD_DIRECT (default) or METHOD_FOURIER for the batches created
        afterwards without explicit launch parameters, and for autotune()."""
        _check(load_library().ps_set_method(self._h, int(method)))
        self._method = int(method)

    @property
    def method(self):
        return getattr(self, "_method", METHOD_DIRECT)

    def set_refinement(self, tol):
        """Mixed precision: evaluate polygons whose estimated relative error
        u32 (16 + 32 kappa / sqrt(#polygons)) exceeds tol again in FP64
        (tol = 0: off; inf: estimates only).  Fields loaded afterwards (and
        the one still in the upload buffer) also get an FP64 spectrum."""
        _check(load_library().ps_set_refinement(self._h, float(tol)))

    def polyspectrum(self, n, k, kmin=None, kmax=None, kbinwidth=1.0, kbinning_case=0, slot=0):
        """polyspectrum.c:51-161 on the GPU.  Returns (value, kmin, kmax,
        (sum prod D, sum prod I)); kmin/kmax as updated for DERIVED binning."""
        k, kp = _darr(k)
        lo, lop = _darr(kmin)
        lo = lo.copy(); lop = lo.ctypes.data_as(_dp)
        hi, hip = _darr(kmax)
        hi = hi.copy(); hip = hi.ctypes.data_as(_dp)
        res = ctypes.c_double()
        sums = np.zeros(2)
        _check(load_library().ps_polyspectrum(self._h, slot, int(n), kp, lop, hip, float(kbinwidth),
                                              int(kbinning_case), ctypes.byref(res), sums.ctypes.data_as(_dp)))
        return res.value, lo, hi, (sums[0], sums[1])

    def num_polygons(self, n, k, kmin=None, kmax=None, kbinwidth=1.0, kbinning_case=0):
        """polyspectrum.c:163-252 on the GPU."""
        k, kp = _darr(k)
        lo = _darr(kmin)[0].copy()
        hi = _darr(kmax)[0].copy()
        res = ctypes.c_double()
        _check(load_library().ps_num_polygons(self._h, int(n), kp, lo.ctypes.data_as(_dp), hi.ctypes.data_as(_dp),
                                              float(kbinwidth), int(kbinning_case), ctypes.byref(res)))
        return res.value

    def batch(self, polygons, kbinwidth=1.0, kbinning_case=0, params=None):
        return Batch(self, polygons, kbinwidth, kbinning_case, params)

    def autotune(self, polygons, kbinwidth=1.0, kbinning_case=0, reps=2):
        """Times candidate launch shapes; later batches of this shape use the best.
        Returns ({grid_cap, chunk_tiles, g2_div, overlap, method}, ms)."""
        arr = _polygon_array(polygons)
        lp = LaunchParams()
        ms = ctypes.c_double()
        _check(load_library().ps_batch_autotune(self._h, arr, len(polygons), float(kbinwidth), int(kbinning_case),
                                                int(reps), ctypes.byref(lp), ctypes.byref(ms)))
        return _params_dict(lp), ms.value

    def direct_check(self, polygons, kbinwidth=1.0, kbinning_case=0, static=False):
        """The polygon values taken without transforms, over the closed
        lattice polygons of the shells (an independent check of a batch).
        Returns (values, npolygons, sums[n,2]) as Batch.execute.  static: the
        entry point that allocates no device memory (ps_direct_check_static)."""
        arr = _polygon_array(polygons)
        n = len(polygons)
        v = np.zeros(n)
        npol = np.zeros(n)
        s = np.zeros(2 * n)
        fn = load_library().ps_direct_check_static if static else load_library().ps_direct_check
        _check(fn(self._h, arr, n, float(kbinwidth), int(kbinning_case), v.ctypes.data_as(_dp),
                  npol.ctypes.data_as(_dp), s.ctypes.data_as(_dp)))
        return v, npol, s.reshape(n, 2)

    def condition_numbers(self, polygons, kbinwidth=1.0, kbinning_case=0):
        """kappa = sum_x |prod D| / |sum_x prod D| of every polygon (the
        amplification of rounding errors in its value), through the
        reference-shaped path of the context's precision."""
        arr = _polygon_array(polygons)
        n = len(polygons)
        k = np.zeros(n)
        _check(load_library().ps_condition_numbers(self._h, arr, n, float(kbinwidth), int(kbinning_case),
                                                   k.ctypes.data_as(_dp)))
        return k

    def batch_for_field(self, polygons, field, kbinwidth=1.0, kbinning_case=0, slot=0):
        """Load a field into the slot, then plan a batch on it."""
        self.load_field(field, slot)
        return Batch(self, polygons, kbinwidth, kbinning_case)


class Batch:
    """Many configurations evaluated by one captured launch sequence."""

    def __init__(self, ctx, polygons, kbinwidth=1.0, kbinning_case=0, params=None):
        arr = _polygon_array(polygons)
        h = _vp()
        lp = None
        if params is not None:
            lp = LaunchParams(int(params["grid_cap"]), int(params["chunk_tiles"]), int(params.get("g2_div", 1)),
                              int(params.get("overlap", 1)), int(params.get("method", ctx.method)))
        _check(load_library().ps_batch_create_ex(ctx._h, arr, len(polygons), float(kbinwidth), int(kbinning_case),
                                                 ctypes.byref(lp) if lp is not None else None, ctypes.byref(h)))
        self._h = h
        self._ctx = ctx  # keeps the context alive
        ctx._dependents.add(self)
        self.polygons = arr
        self.n = len(polygons)

    def execute(self):
        """Returns (values, npolygons, sums[n,2])."""
        v = np.zeros(self.n)
        npol = np.zeros(self.n)
        s = np.zeros(2 * self.n)
        _check(load_library().ps_batch_execute(self._h, v.ctypes.data_as(_dp), npol.ctypes.data_as(_dp),
                                               s.ctypes.data_as(_dp)))
        return v, npol, s.reshape(self.n, 2)

    def refinement(self):
        """(refined flags, error estimates) of the last execute()."""
        f = np.zeros(self.n, dtype=np.int32)
        e = np.zeros(self.n)
        r = load_library().ps_batch_refinement(self._h, f.ctypes.data_as(_ip), e.ctypes.data_as(_dp))
        if r < 0:
            _check(r)
        return f.astype(bool), e

    def adjoint(self, weights, slot=0, out=None):
        """dL/d field for L with dL/dvalues = weights (after execute()).  out:
        None (returns a float64 numpy N^3 array) or a torch CUDA tensor
        (float32/float64, N^3 elements) filled in place."""
        w = np.ascontiguousarray(weights, dtype=np.float64)
        assert w.size == self.n
        N = self._ctx.nbins
        if out is None:
            g = np.empty((N, N, N), dtype=np.float64)
            _check(load_library().ps_batch_adjoint(self._h, slot, w.ctypes.data_as(_dp), g.ctypes.data_as(_vp), 1, 0))
            return g
        import torch
        assert out.is_cuda and out.is_contiguous() and out.numel() == N ** 3
        _check(load_library().ps_batch_adjoint(self._h, slot, w.ctypes.data_as(_dp), ctypes.c_void_p(out.data_ptr()),
                                               int(out.dtype == torch.float64), 1))
        return out

    @property
    def params(self):
        lp = LaunchParams()
        _check(load_library().ps_batch_params(self._h, ctypes.byref(lp)))
        return _params_dict(lp)

    def results(self):
        """(values, npolygons, sums) of the last execution, no launch."""
        v = np.zeros(self.n)
        npol = np.zeros(self.n)
        s = np.zeros(2 * self.n)
        _check(load_library().ps_batch_results(self._h, v.ctypes.data_as(_dp), npol.ctypes.data_as(_dp),
                                               s.ctypes.data_as(_dp)))
        return v, npol, s.reshape(self.n, 2)

    def profile(self, reps=3):
        """Per-pass mean ms and algorithmic bytes, and the pass names
        (direct: G1, G2, Zc, reduce; Fourier: G1, G2, Zf, Fx, dot, reduce)."""
        L = load_library()
        ms = np.zeros(MAX_STAGES)
        by = np.zeros(MAX_STAGES)
        _check(L.ps_batch_profile(self._h, reps, ms.ctypes.data_as(_dp), by.ctypes.data_as(_dp)))
        n = L.ps_batch_stage_count(self._h)
        names = [L.ps_batch_stage_name(self._h, i).decode() for i in range(n)]
        return ms[:n], by[:n], names

    def waves(self):
        """The plan: per wave (transform grid, legs, groups, arena bytes)."""
        L = load_library()
        out = []
        for w in range(L.ps_batch_wave_count(self._h)):
            g, nl, ng = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            by = ctypes.c_double()
            _check(L.ps_batch_wave_info(self._h, w, ctypes.byref(g), ctypes.byref(nl), ctypes.byref(ng),
                                        ctypes.byref(by)))
            out.append((g.value, nl.value, ng.value, by.value))
        return out

    @property
    def last_ms(self):
        return load_library().ps_batch_last_ms(self._h)

    @property
    def model_bytes(self):
        return load_library().ps_batch_model_bytes(self._h)

    def close(self):
        if getattr(self, "_h", None):
            load_library().ps_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def polyspectrum(nbins, local_n0, local_n0_start, fft_array, n, k, kmin, kmax, kbinwidth, kbinningCase, boxsize):
    """Drop-in for polyspectrum() (polyspectrum.h:7): fft_array is the full
    complex128 N^3 spectrum; kmin/kmax (numpy float64) are updated in place."""
    a = np.ascontiguousarray(fft_array, dtype=np.complex128)
    k = np.ascontiguousarray(k, dtype=np.float64)
    assert isinstance(kmin, np.ndarray) and kmin.dtype == np.float64 and kmin.flags.c_contiguous
    assert isinstance(kmax, np.ndarray) and kmax.dtype == np.float64 and kmax.flags.c_contiguous
    v = load_library().polyspectrum(nbins, local_n0, local_n0_start, a.ctypes.data_as(_vp), n,
                                    k.ctypes.data_as(_dp), kmin.ctypes.data_as(_dp), kmax.ctypes.data_as(_dp),
                                    kbinwidth, kbinningCase, boxsize)
    return v


def num_polygons(nbins, local_n0, local_n0_start, n, k, kmin, kmax, kbinwidth, kbinningCase, boxsize):
    """Drop-in for num_polygons() (polyspectrum.h:8)."""
    k = np.ascontiguousarray(k, dtype=np.float64)
    assert isinstance(kmin, np.ndarray) and kmin.dtype == np.float64
    assert isinstance(kmax, np.ndarray) and kmax.dtype == np.float64
    return load_library().num_polygons(nbins, local_n0, local_n0_start, n, k.ctypes.data_as(_dp),
                                       kmin.ctypes.data_as(_dp), kmax.ctypes.data_as(_dp), kbinwidth,
                                       kbinningCase, boxsize)


class MultiBatch:
    """Several batches (any grid sizes and slots) replayed as one hipGraph."""

    def __init__(self, batches):
        arr = (_vp * len(batches))(*[b._h for b in batches])
        h = _vp()
        _check(load_library().ps_multi_create(arr, len(batches), ctypes.byref(h)))
        self._h = h
        self.batches = list(batches)
        for b in self.batches:
            b._ctx._dependents.add(self)

    def execute(self):
        """Replays the graph; returns [(values, npolygons, sums)] per batch."""
        _check(load_library().ps_multi_execute(self._h))
        return [b.results() for b in self.batches]

    @property
    def last_ms(self):
        return load_library().ps_multi_last_ms(self._h)

    def close(self):
        if getattr(self, "_h", None):
            load_library().ps_multi_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
