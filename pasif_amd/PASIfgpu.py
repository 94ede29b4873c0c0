"""``PASIfgpu`` module: the ``__GpuDriver`` class the reference builds with
pybind11 (``src/PASIfgpu.cpp:22-104``), here a thin ctypes layer over the
C ABI of ``libpasif_hip.so``.

Method names, argument names and defaults are the reference's.  Differences:

* failures raise :class:`pasif_amd._lib.PasifError` instead of exiting the
  process (``src/helpers.cuh:43-51``); ``_loadExcitationsSet`` still returns
  1 for ragged excitation sets, as the reference does
  (``src/__GpuDriver.cu:358-362``);
* getters return ``numpy.ndarray`` (float64) instead of Python lists;
* methods after ``_getGradient`` are extensions (per-system layout,
  precision, solver, tuning, fused amplitude window, timing).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib

__all__ = ["__GpuDriver"]


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))


def _u32(a) -> np.ndarray:
    arr = np.asarray(a).reshape(-1)
    if arr.size and (arr.min() < 0 or arr.max() > 0xFFFFFFFF):
        raise ValueError("index out of the uint32 range")
    return np.ascontiguousarray(arr.astype(np.uint32))


def _p(arr, ctype):
    return arr.ctypes.data_as(C.POINTER(ctype))


class __GpuDriver:  # noqa: N801  (the reference's class name)
    """Driver of one GPU: excitations, forward / adjoint systems, solvers.

    ``__GpuDriver(excitationSet_, sampleRate_, numsteps_, dCompute_=False,
    dSystem_=False, dSolver_=False, GPUId_=0)`` -- ``src/__GpuDriver.cu:25-157``.
    """

    def __init__(self, excitationSet_, sampleRate_, numsteps_, dCompute_=False, dSystem_=False,
                 dSolver_=False, GPUId_=0):
        self._pl = _lib.load()
        exc = self._flatten_excitations(excitationSet_)
        if exc is None:
            raise ValueError("Excitations vectors are not of the same size")
        flat, n_exc, length = exc
        h = C.c_void_p()
        rc = self._pl.pasif_create(C.byref(h), _p(flat, C.c_double), n_exc, length,
                                   int(sampleRate_), int(numsteps_), int(bool(dCompute_)),
                                   int(bool(dSystem_)), int(bool(dSolver_)), int(GPUId_))
        self._h = h
        if rc != 0:
            msg = self._pl.pasif_last_error(h)
            self._pl.pasif_destroy(h)
            self._h = None
            raise _lib.PasifError(rc, msg.decode() if msg else "")
        self._layout = {"n": 0, "sizes_f": None, "sizes_a": None, "exc": None, "steps": None}

    # ------------------------------------------------------------ plumbing
    def _chk(self, rc):
        _lib.check(self._pl, self._h, rc)

    @staticmethod
    def _flatten_excitations(excitationSet_):
        rows = [np.asarray(e, dtype=np.float64).reshape(-1) for e in excitationSet_]
        if not rows:
            return None
        L = rows[0].size
        if any(r.size != L for r in rows):
            return None
        return np.ascontiguousarray(np.stack(rows)), len(rows), L

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                self._pl.pasif_destroy(h)
            except Exception:
                pass
            self._h = None

    def _get(self, fn, *args) -> np.ndarray:
        n = C.c_size_t(0)
        self._chk(fn(self._h, *args, None, 0, C.byref(n)))
        out = np.empty(n.value, dtype=np.float64)
        self._chk(fn(self._h, *args, _p(out, C.c_double), out.size, C.byref(n)))
        return out[: n.value]

    # ---------------------------------------------- forward system interface
    def _setFwdK(self, n_, values_, indices_, indptr_):
        self._set_csr(self._pl.pasif_set_fwd_K, n_, values_, indices_, indptr_)

    def _setFwdGamma(self, n_, values_, indices_):
        self._set_coo(self._pl.pasif_set_fwd_gamma, 3, n_, values_, indices_)

    def _setFwdLambda(self, n_, values_, indices_):
        self._set_coo(self._pl.pasif_set_fwd_lambda, 4, n_, values_, indices_)

    def _setFwdPsi(self, n_, values_, indices_):
        self._set_coo(self._pl.pasif_set_fwd_psi, 5, n_, values_, indices_)

    def _setFwdForcePattern(self, forcePattern_):
        f = _f64(forcePattern_)
        self._chk(self._pl.pasif_set_fwd_force_pattern(self._h, _p(f, C.c_double), f.size))

    def _setFwdInitialConditions(self, initialConditions_):
        q = _f64(initialConditions_)
        self._chk(self._pl.pasif_set_fwd_initial_conditions(self._h, _p(q, C.c_double), q.size))

    def _allocateSystemOnDevice(self):
        self._chk(self._pl.pasif_allocate_system(self._h))

    # ---------------------------------------------- backward system interface
    def _setBwdK(self, n_, values_, indices_, indptr_):
        self._set_csr(self._pl.pasif_set_bwd_K, n_, values_, indices_, indptr_)

    def _setBwdGamma(self, n_, values_, indices_):
        self._set_coo(self._pl.pasif_set_bwd_gamma, 3, n_, values_, indices_)

    def _setBwdLambda(self, n_, values_, indices_):
        self._set_coo(self._pl.pasif_set_bwd_lambda, 4, n_, values_, indices_)

    def _setBwdPsi(self, n_, values_, indices_):
        self._set_coo(self._pl.pasif_set_bwd_psi, 5, n_, values_, indices_)

    def _setBwdForcePattern(self, forcePattern_):
        f = _f64(forcePattern_)
        self._chk(self._pl.pasif_set_bwd_force_pattern(self._h, _p(f, C.c_double), f.size))

    def _setBwdInitialConditions(self, initialConditions_):
        q = _f64(initialConditions_)
        self._chk(self._pl.pasif_set_bwd_initial_conditions(self._h, _p(q, C.c_double), q.size))

    def _allocateAdjointOnDevice(self):
        self._chk(self._pl.pasif_allocate_adjoint(self._h))

    # ---------------------------------------------- compute options interface
    def _loadExcitationsSet(self, excitationSet_, sampleRate_) -> int:
        exc = self._flatten_excitations(excitationSet_)
        if exc is None:
            print("[Error] __GpuDriver: Excitations vectors are not of the same size")
            return 1
        flat, n_exc, length = exc
        self._chk(self._pl.pasif_load_excitations(self._h, _p(flat, C.c_double), n_exc, length,
                                                  int(sampleRate_)))
        return 0

    def _setInterpolationMatrix(self, interpolationMatrix_, interpolationWindowSize_):
        m = _f64(interpolationMatrix_)
        self._chk(self._pl.pasif_set_interpolation_matrix(self._h, _p(m, C.c_double), m.size,
                                                          int(interpolationWindowSize_)))

    def _setModulationBuffer(self, modulationBuffer_):
        b = _f64(modulationBuffer_)
        self._chk(self._pl.pasif_set_modulation_buffer(self._h, _p(b, C.c_double), b.size))

    # --------------------------------------------------------- solvers
    def _getAmplitudes(self) -> np.ndarray:
        return self._get(self._pl.pasif_get_amplitudes)

    def _getTrajectory(self, saveSteps_=1) -> np.ndarray:
        return self._get(self._pl.pasif_get_trajectory, C.c_uint32(int(saveSteps_)))

    def _getGradient(self, save=0) -> np.ndarray:
        return self._get(self._pl.pasif_get_gradient, C.c_uint32(int(save)))

    # --------------------------------------------------------- extensions
    def _setSystemsLayout(self, sizes_fwd: Optional[Sequence[int]] = None,
                          sizes_adj: Optional[Sequence[int]] = None,
                          exc_of_system: Optional[Sequence[int]] = None,
                          steps_of_system: Optional[Sequence[int]] = None):
        """Per-system sizes (block boundaries), excitation id and step count;
        every call sets all four (``None`` clears one).  ``exc_of_system``
        switches from the Cartesian product (every system x every
        excitation) to one excitation per system."""
        lay = self._layout
        for k, v in (("sizes_f", sizes_fwd), ("sizes_a", sizes_adj), ("exc", exc_of_system),
                     ("steps", steps_of_system)):
            lay[k] = None if v is None else np.ascontiguousarray(np.asarray(v).reshape(-1))
        ns = [len(v) for v in (lay["sizes_f"], lay["sizes_a"], lay["exc"], lay["steps"])
              if v is not None]
        if ns and any(x != ns[0] for x in ns):
            raise ValueError(f"layout arrays disagree on the number of systems: {ns}")
        n = ns[0] if ns else 0
        sf = _u32(lay["sizes_f"]) if lay["sizes_f"] is not None else None
        sa = _u32(lay["sizes_a"]) if lay["sizes_a"] is not None else None
        ex = (np.ascontiguousarray(lay["exc"].astype(np.int32))
              if lay["exc"] is not None else None)
        st = _u32(lay["steps"]) if lay["steps"] is not None else None
        self._chk(self._pl.pasif_set_systems_layout(
            self._h, n, None if sf is None else _p(sf, C.c_uint32),
            None if sa is None else _p(sa, C.c_uint32),
            None if ex is None else _p(ex, C.c_int32),
            None if st is None else _p(st, C.c_uint32)))

    def _clearSystemsLayout(self):
        self._layout = {"n": 0, "sizes_f": None, "sizes_a": None, "exc": None, "steps": None}
        self._chk(self._pl.pasif_set_systems_layout(self._h, 0, None, None, None, None))

    def _setPrecision(self, bits: int):
        self._chk(self._pl.pasif_set_precision(self._h, int(bits)))
        self._bits = int(bits)

    def _precision(self) -> int:
        """Working precision in bits (64 unless set)."""
        return getattr(self, "_bits", 64)

    def _setSolver(self, solver: str):
        code = {"persistent": _lib.PASIF_SOLVER_PERSISTENT, "plain": _lib.PASIF_SOLVER_PLAIN,
                "graph": _lib.PASIF_SOLVER_GRAPH}[solver]
        self._chk(self._pl.pasif_set_solver(self._h, code))

    def _setTuning(self, tile_lanes: int = 0, slot_config: int = 0, checkpoint_every: int = 0):
        self._chk(self._pl.pasif_set_tuning(self._h, int(tile_lanes), int(slot_config),
                                            int(checkpoint_every)))

    def _setGraphSlicing(self, steps_per_launch: int = 256, launches_per_graph: int = 16):
        self._chk(self._pl.pasif_set_graph_slicing(self._h, int(steps_per_launch),
                                                   int(launches_per_graph)))

    def _setAmplitudeWindow(self, window: int):
        self._chk(self._pl.pasif_set_amplitude_window(self._h, int(window)))

    def _getMaxAmplitudes(self) -> np.ndarray:
        return self._get(self._pl.pasif_get_max_amplitudes)

    _KERNELS = {"auto": 0, "tile": 1, "lane": 2, "big": 4}

    def _setKernel(self, kernel: str = "auto"):
        if kernel not in self._KERNELS:
            raise ValueError(f"kernel must be one of {sorted(self._KERNELS)}")
        self._chk(self._pl.pasif_set_kernel(self._h, self._KERNELS[kernel]))

    def _setLaneGroup(self, lanes: int = 0):
        """Lanes per instance in the lane kernel: 0 = chosen per structure
        class, else a power of two up to 64 for every class it fits."""
        self._chk(self._pl.pasif_set_lane_group(self._h, int(lanes)))

    def _setLaneBias(self, bias: int = 0):
        """Steps along each class's feasible group sizes from the cost
        model's choice (+1: twice the lanes per instance where possible)."""
        self._chk(self._pl.pasif_set_lane_bias(self._h, int(bias)))

    def _laneGroups(self, adjoint: bool = False) -> list:
        n = C.c_size_t(0)
        self._chk(self._pl.pasif_lane_groups(self._h, int(adjoint), None, 0, C.byref(n)))
        buf = (C.c_uint32 * max(1, n.value))()
        self._chk(self._pl.pasif_lane_groups(self._h, int(adjoint), buf, n.value, C.byref(n)))
        return [int(buf[i]) for i in range(n.value)]

    def _kernelInfo(self, adjoint: bool = False) -> dict:
        k, ncls, ck = C.c_int(0), C.c_uint32(0), C.c_uint32(0)
        self._chk(self._pl.pasif_kernel_info(self._h, int(adjoint), C.byref(k), C.byref(ncls),
                                             C.byref(ck)))
        name = {1: "tile", 2: "lane", 3: "mixed", 4: "big"}.get(k.value, "?")
        groups = self._laneGroups(adjoint) if k.value in (2, 3) else []
        return {"kernel": name, "classes": ncls.value, "checkpoint_every": ck.value,
                "lanes_per_instance": groups}

    def _setAdjointMode(self, mode: str = "reference"):
        modes = {"reference": 0, "exact": 1}
        if mode not in modes:
            raise ValueError(f"adjoint mode {mode!r}: 'reference' or 'exact'")
        self._chk(self._pl.pasif_set_adjoint_mode(self._h, modes[mode]))

    def _runFootprint(self, adjoint: bool = False) -> dict:
        """Device bytes of a solve's run buffers and outputs (``run``), the
        checkpoints among them (``ckpt``) and the memory a solve may use now
        (``free``) for the loaded problem and settings."""
        r, c, f = C.c_size_t(0), C.c_size_t(0), C.c_size_t(0)
        self._chk(self._pl.pasif_run_footprint(self._h, int(adjoint), C.byref(r), C.byref(c),
                                               C.byref(f)))
        return {"run": r.value, "ckpt": c.value, "free": f.value}

    def _laneSource(self, adjoint: bool = False, compile: bool = False) -> str:
        """Generated HIP source of the lane kernel (optionally compiled with
        hipRTC for gfx950, which needs no GPU)."""
        n = C.c_size_t(0)
        self._chk(self._pl.pasif_lane_source(self._h, int(adjoint), None, 0, C.byref(n), 0))
        buf = C.create_string_buffer(n.value)
        self._chk(self._pl.pasif_lane_source(self._h, int(adjoint), buf, n.value, C.byref(n),
                                             int(compile)))
        return buf.value.decode()

    def _lastTiming(self) -> dict:
        ms = C.c_double(0)
        launches, tiles, lanes = C.c_uint32(0), C.c_uint32(0), C.c_uint32(0)
        self._chk(self._pl.pasif_last_timing(self._h, C.byref(ms), C.byref(launches),
                                             C.byref(tiles), C.byref(lanes)))
        return {"kernel_ms": ms.value, "launches": launches.value, "tiles": tiles.value,
                "lanes": lanes.value}

    # ------------------------------------------------------------- helpers
    def _set_csr(self, fn, n_, values_, indices_, indptr_):
        n = _u32(list(n_)[:2])
        v = _f64(values_)
        ix = _u32(indices_)
        ip = _u32(indptr_)
        if ix.size != v.size:
            raise ValueError("CSR indices and values differ in length")
        self._chk(fn(self._h, _p(n, C.c_uint32), _p(v, C.c_double), _p(ix, C.c_uint32),
                     _p(ip, C.c_uint32), v.size))

    def _set_coo(self, fn, order, n_, values_, indices_):
        n = _u32(list(n_)[:order])
        if n.size != order:
            raise ValueError(f"tensor of order {order} needs {order} dimensions")
        v = _f64(values_)
        ix = _u32(indices_) if len(indices_) else np.zeros(0, dtype=np.uint32)
        if ix.size != order * v.size:
            raise ValueError(f"tensor of order {order}: {ix.size} indices for {v.size} values")
        self._chk(fn(self._h, _p(n, C.c_uint32), _p(v, C.c_double), _p(ix, C.c_uint32), v.size))
