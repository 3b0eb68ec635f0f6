"""Host side of the MI355X uvh assembly: ctypes bindings of ``libuvh.so`` and a
drop-in ``uvhAssembly`` with the reference's signature and error behaviour.

``uvhAssembly(UserVar, RunInfo, CtrlVar, MUA, F0, F1, l1, BCs1) -> (UserVar, RunInfo, R, K)``
mirrors ``UaMain/uvhAssembly.m:1-70``: the same inputs, ``R`` a dense ``3*Nn`` vector
``[x; y; h]`` and ``K`` a sparse ``3Nn x 3Nn`` matrix (``None`` when
``CtrlVar.uvhMatrixAssembly.Ronly``), with the assembly time added to
``RunInfo.CPU.Assembly.uvh`` (``uvhAssembly.m:6,14-15``). NaN results raise
``UaError('uvhMatrixAssembly:NaNinR')`` as ``uvhMatrixAssembly.m:581-584`` does.
Options the GPU path does not implement raise ``UvhUnsupported``; a MATLAB
gateway would then call Úa's own ``uvhMatrixAssembly`` instead.

There is no CPU fallback here: without ``libuvh.so`` or without a GPU every call
raises.
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np
import scipy.sparse as sp

from .ctrlvar import Struct

_HERE = os.path.dirname(os.path.abspath(__file__))
# UVH_LIBRARY overrides the library path (e.g. to A/B two builds of the kernels).
LIB_PATH = os.environ.get("UVH_LIBRARY") or os.path.join(_HERE, "libuvh.so")

UVH_OK, UVH_E_ARG, UVH_E_HIP, UVH_E_NAN_R, UVH_E_NAN_K, UVH_E_UNSUPPORTED, UVH_E_STATE, UVH_E_NOMEM = range(8)
UVH_RONLY, UVH_ZEROFIELDS, UVH_FP32, UVH_NO_NAN_CHECK = 1, 2, 4, 8
TAU_OPTIONS = {"taus": 0, "taut": 1, "tau1": 2, "tau2": 3, "tau0": 4}

# Every entry point declared in include/uvh.h.
EXPORTED = ("uvh_create", "uvh_destroy", "uvh_last_error", "uvh_set_mesh", "uvh_pattern_info",
            "uvh_get_pattern", "uvh_get_csc_map", "uvh_assemble", "uvh_assemble_device", "uvh_sync_status",
            "uvh_profile", "uvh_kernel_times", "uvh_uv_pattern_info", "uvh_uv_get_pattern", "uvh_uv_assemble",
            "uvh_uv_assemble_device", "uvh_status_async", "uvh_assemble_device_noalloc")


class UaError(RuntimeError):
    """Error with a MATLAB-style identifier (``uvhMatrixAssembly:NaNinR`` ...)."""

    def __init__(self, identifier: str, message: str):
        super().__init__(f"{identifier}: {message}")
        self.identifier = identifier


class UvhUnsupported(UaError):
    """The requested option is outside the GPU path (caller uses Úa's m-file assembly)."""


class UvhLibraryError(RuntimeError):
    pass


class UvhCtrl(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("dt", "theta", "supg_beta0", "supg_tau_multiplier", "speed_zero",
                                          "eps_zero", "eta_zero", "czero", "kH", "Hh0", "g", "rhow", "alpha",
                                          "thick_min")] + \
               [(n, C.c_int32) for n in ("supg_tau", "reset_thickness", "enforce_b_above_B", "sliding_law")] + \
               [("nzero", C.c_double)]


# CtrlVar.SlidingLaw names (UaMain/BasalDrag.m:182-257) -> enum uvh_sliding_law
SLIDING_LAWS = {"Weertman": 0, "W": 0, "Budd": 1, "W-N0": 1, "Coulomb": 2, "C": 2, "Tsai": 3, "minCW-N0": 3,
                "Cornford": 4, "rpCW-N0": 4, "Umbi": 5, "rCW-N0": 5, "Joughin": 6, "rCW-v0": 6}
# the nodal fields each law reads beyond Weertman's (F.q, F.muk, F.V0)
SLIDING_FIELDS = {0: (), 1: ("q",), 2: ("muk",), 3: ("muk",), 4: ("muk",), 5: ("muk",), 6: ("V0",)}


class UvhFields(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("h", "ub", "vb", "S", "B", "rho", "C", "m", "n", "AGlen", "as_", "ab",
                                          "h0", "ub0", "vb0", "as0", "ab0", "dadh", "q", "muk", "V0")]


class UvhUvFields(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("h", "s", "ub", "vb", "S", "B", "rho", "C", "m", "n", "AGlen", "q", "muk",
                                          "V0")]


class UvhUvCtrl(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("cmin", "aglenmin", "nrviscosity")]


UV_KEYS = ("h", "s", "ub", "vb", "S", "B", "rho", "C", "m", "n", "AGlen")


def make_uv_ctrl(CtrlVar) -> UvhUvCtrl:
    """CtrlVar.Cmin, AGlenmin and NRviscosity (Dvisk: 0 with CtrlVar.Picard,
    uvMatrixAssemblySSTREAM.m:72-76)."""
    return UvhUvCtrl(float(CtrlVar.Cmin), float(CtrlVar.AGlenmin),
                     0.0 if CtrlVar.Picard else float(CtrlVar.NRviscosity))


def make_uv_fields(F, keep=None) -> UvhUvFields:
    """Field pointer block of the uv assembly (host arrays or device tensors)."""
    f = UvhUvFields()
    for k in UV_KEYS:
        a = getattr(F, k)
        if isinstance(a, np.ndarray):
            a = np.ascontiguousarray(a, dtype=np.float64)
        if keep is not None:
            keep.append(a)
        setattr(f, k, _ptr(a))
    _sliding_fields(f, F, keep)
    return f


def _sliding_fields(f, F, keep):
    """q, muk and V0 of F, when F has them (the sliding laws that read them check for them)."""
    for k in ("q", "muk", "V0"):
        a = getattr(F, k, None)
        if a is None or (isinstance(a, np.ndarray) and a.size == 0):
            continue
        if isinstance(a, np.ndarray):
            a = np.ascontiguousarray(a, dtype=np.float64)
        if keep is not None:
            keep.append(a)
        setattr(f, k, _ptr(a))


_lib = None


def load_library(path: str = LIB_PATH):
    """Load ``libuvh.so`` (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise UvhLibraryError(f"{path} is missing: build it with `python -m uasource_amd.csrc.build`")
    # One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64 and asks for it by
    # its unversioned name, so if libuvh (which asks for the soname libamdhip64.so.7) were
    # loaded first, torch would load a second runtime and find no device. Loading torch
    # first makes libuvh bind to the runtime torch already holds.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    vp, i32, i64, dp = C.c_void_p, C.c_int32, C.c_int64, C.c_void_p
    lib.uvh_create.argtypes = [C.POINTER(vp), C.c_int]
    lib.uvh_create.restype = C.c_int
    lib.uvh_destroy.argtypes = [vp]
    lib.uvh_destroy.restype = None
    lib.uvh_last_error.argtypes = [vp]
    lib.uvh_last_error.restype = C.c_char_p
    lib.uvh_set_mesh.argtypes = [vp, i32, i32, i32, vp, vp, vp, i32, vp, vp]
    lib.uvh_set_mesh.restype = C.c_int
    lib.uvh_pattern_info.argtypes = [vp, C.POINTER(i64), C.POINTER(i64)]
    lib.uvh_pattern_info.restype = C.c_int
    lib.uvh_get_pattern.argtypes = [vp, vp, vp]
    lib.uvh_get_pattern.restype = C.c_int
    lib.uvh_get_csc_map.argtypes = [vp, vp]
    lib.uvh_get_csc_map.restype = C.c_int
    lib.uvh_assemble.argtypes = [vp, C.POINTER(UvhCtrl), C.POINTER(UvhFields), C.c_uint32, dp, dp, dp, dp]
    lib.uvh_assemble.restype = C.c_int
    lib.uvh_assemble_device.argtypes = [vp, C.POINTER(UvhCtrl), C.POINTER(UvhFields), C.c_uint32, dp, dp, dp, dp, vp]
    lib.uvh_assemble_device.restype = C.c_int
    lib.uvh_uv_pattern_info.argtypes = [vp, C.POINTER(i64), C.POINTER(i64)]
    lib.uvh_uv_pattern_info.restype = C.c_int
    lib.uvh_uv_get_pattern.argtypes = [vp, vp, vp]
    lib.uvh_uv_get_pattern.restype = C.c_int
    lib.uvh_uv_assemble.argtypes = [vp, C.POINTER(UvhCtrl), C.POINTER(UvhUvCtrl), C.POINTER(UvhUvFields), C.c_uint32,
                                    dp, dp, dp, dp]
    lib.uvh_uv_assemble.restype = C.c_int
    lib.uvh_uv_assemble_device.argtypes = [vp, C.POINTER(UvhCtrl), C.POINTER(UvhUvCtrl), C.POINTER(UvhUvFields),
                                           C.c_uint32, dp, dp, dp, dp, vp]
    lib.uvh_uv_assemble_device.restype = C.c_int
    lib.uvh_sync_status.argtypes = [vp]
    lib.uvh_sync_status.restype = C.c_int
    lib.uvh_status_async.argtypes = [vp, vp, vp]
    lib.uvh_status_async.restype = C.c_int
    lib.uvh_profile.argtypes = [vp, C.c_int]
    lib.uvh_profile.restype = C.c_int
    lib.uvh_kernel_times.argtypes = [vp, C.POINTER(C.c_char_p), C.POINTER(C.c_double), C.POINTER(C.c_int64), C.c_int]
    lib.uvh_kernel_times.restype = C.c_int
    _lib = lib
    return lib


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return int(a.data_ptr())  # torch tensor on the device


def check_supported(CtrlVar) -> None:
    """Raise ``UvhUnsupported`` for options outside the GPU path (see include/uvh.h)."""
    cv = CtrlVar
    reasons = []
    if str(cv.FlowApproximation).lower() != "sstream":
        reasons.append(f"FlowApproximation={cv.FlowApproximation}")
    if str(cv.SlidingLaw) not in SLIDING_LAWS:
        reasons.append(f"SlidingLaw={cv.SlidingLaw}")
    if cv.IncludeMelangeModelPhysics:
        reasons.append("IncludeMelangeModelPhysics")
    if cv.LevelSetMethod and cv.LevelSetMethodAutomaticallyApplyMassBalanceFeedback:
        reasons.append("LevelSetMethod with mass-balance feedback")
    if cv.get("ThicknessPenalty", False):
        reasons.append("ThicknessPenalty")
    if cv.get("ThicknessBarrier", False):
        reasons.append("ThicknessBarrier")
    if cv.MassBalanceGeometryFeedback >= 2 and not cv.uvhMatrixAssembly.ZeroFields:
        reasons.append("MassBalanceGeometryFeedback>=2")
    if cv.uvhGroupAssembly or cv.uvGroupAssembly:
        reasons.append("uvhGroupAssembly")
    if cv.IncludeTG3uvhBoundaryTerm and cv.TG3:
        reasons.append("IncludeTG3uvhBoundaryTerm")
    if cv.Development.Pre2025uvhAssembly:
        reasons.append("Development.Pre2025uvhAssembly")
    if cv.OnlyCalcBasalDragAndEffectiveViscosity:
        reasons.append("OnlyCalcBasalDragAndEffectiveViscosity")
    if cv.uvh.SUPG.tau not in TAU_OPTIONS:
        raise UaError("SUPGtau", "in SUPGtau case not found")
    if reasons:
        raise UvhUnsupported("uvhAssembly:unsupported", ", ".join(reasons))


def make_ctrl(CtrlVar, F1) -> UvhCtrl:
    c = UvhCtrl()
    c.dt = CtrlVar.dt
    c.theta = CtrlVar.theta
    c.supg_beta0 = CtrlVar.SUPG.beta0
    c.supg_tau_multiplier = CtrlVar.uvh.SUPG.tauMultiplier
    c.speed_zero = CtrlVar.SpeedZero
    c.eps_zero = CtrlVar.EpsZero
    c.eta_zero = CtrlVar.etaZero
    c.czero = CtrlVar.Czero
    c.kH = CtrlVar.kH
    c.Hh0 = CtrlVar.Hh0
    c.g = F1.g
    c.rhow = F1.rhow
    c.alpha = F1.alpha
    c.thick_min = CtrlVar.ThickMin
    c.supg_tau = TAU_OPTIONS[CtrlVar.uvh.SUPG.tau]
    c.reset_thickness = int(bool(CtrlVar.ResetThicknessToMinThickness) and bool(CtrlVar.ResetThicknessInNonLinLoop))
    c.enforce_b_above_B = int(bool(CtrlVar.Enforce_bAboveB))
    c.sliding_law = SLIDING_LAWS.get(str(CtrlVar.SlidingLaw), -1)
    c.nzero = float(CtrlVar.Nzero)
    return c


F1_KEYS = ("h", "ub", "vb", "S", "B", "rho", "C", "m", "n", "AGlen", "as_", "ab")
F0_KEYS = ("h", "ub", "vb", "as_", "ab")


def make_fields(F0, F1, dadh=None, keep=None) -> UvhFields:
    """Field pointer block; ``keep`` (a list) receives the arrays that must outlive the call."""
    f = UvhFields()
    arrs = []
    for k in F1_KEYS:
        a = getattr(F1, k)
        if isinstance(a, np.ndarray):
            a = np.ascontiguousarray(a, dtype=np.float64)
        arrs.append(a)
        setattr(f, k, _ptr(a))
    for k, name in zip(F0_KEYS, ("h0", "ub0", "vb0", "as0", "ab0")):
        a = getattr(F0, k)
        if isinstance(a, np.ndarray):
            a = np.ascontiguousarray(a, dtype=np.float64)
        arrs.append(a)
        setattr(f, name, _ptr(a))
    f.dadh = _ptr(dadh) if dadh is not None else None
    if keep is not None:
        keep.extend(arrs)
        if dadh is not None:
            keep.append(dadh)
    _sliding_fields(f, F1, keep)
    return f


class UvhContext:
    """A ``libuvh`` context holding one mesh and its device buffers."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        st = self.lib.uvh_create(C.byref(h), int(device))
        if st != UVH_OK:
            raise UvhLibraryError(f"uvh_create failed with status {st} (no usable HIP device {device}?)")
        self.h = h
        self.device = device
        self.mesh_key = None
        self.nn = self.nele = self.nod = self.nip = 0
        self.neq = self.nnz = 0
        self._pattern = None
        self._csc_map = None

    def close(self):
        if self.h:
            self.lib.uvh_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int, what: str):
        if st == UVH_OK:
            return
        msg = self.lib.uvh_last_error(self.h).decode()
        if st == UVH_E_NAN_R:
            raise UaError("uvhMatrixAssembly:NaNinR", " NaN in R ")
        if st == UVH_E_NAN_K:
            raise UaError("", " NaN in K ")  # no identifier (uvhMatrixAssembly.m:577)
        if st == UVH_E_UNSUPPORTED:
            raise UvhUnsupported("uvhAssembly:unsupported", msg)
        raise UvhLibraryError(f"{what}: status {st}: {msg}")

    def set_mesh(self, MUA) -> None:
        conn = np.ascontiguousarray(MUA.connectivity, dtype=np.int32)
        x = np.ascontiguousarray(MUA.coordinates[:, 0], dtype=np.float64)
        y = np.ascontiguousarray(MUA.coordinates[:, 1], dtype=np.float64)
        pts = np.ascontiguousarray(MUA.points, dtype=np.float64)
        wts = np.ascontiguousarray(MUA.weights, dtype=np.float64)
        st = self.lib.uvh_set_mesh(self.h, MUA.Nele, MUA.nod, MUA.Nnodes, _ptr(conn), _ptr(x), _ptr(y),
                                   MUA.nip, _ptr(pts), _ptr(wts))
        self._check(st, "uvh_set_mesh")
        self.nn, self.nele, self.nod, self.nip = MUA.Nnodes, MUA.Nele, MUA.nod, MUA.nip
        neq, nnz = C.c_int64(), C.c_int64()
        self._check(self.lib.uvh_pattern_info(self.h, C.byref(neq), C.byref(nnz)), "uvh_pattern_info")
        self.neq, self.nnz = neq.value, nnz.value
        self._pattern = None
        self._csc_map = None
        self.mesh_key = mesh_key(MUA)

    def pattern(self):
        """(rowptr int64[neq+1], colind int32[nnz]) of K."""
        if self._pattern is None:
            rowptr = np.zeros(self.neq + 1, dtype=np.int64)
            colind = np.zeros(self.nnz, dtype=np.int32)
            self._check(self.lib.uvh_get_pattern(self.h, _ptr(rowptr), _ptr(colind)), "uvh_get_pattern")
            self._pattern = (rowptr, colind)
        return self._pattern

    def assemble(self, CtrlVar, F0, F1, ronly=False, zero_fields=False, want_tf=False, dadh=None,
                 nan_check=True, fp32=False):
        """Host-array assembly; returns dict(R, Kval or None, Tint, Fext). ``fp32`` selects the
        precision form (FP32 element arithmetic, FP64 accumulation)."""
        ctrl = make_ctrl(CtrlVar, F1)
        keep = []
        fields = make_fields(F0, F1, dadh, keep)
        flags = (UVH_RONLY if ronly else 0) | (UVH_ZEROFIELDS if zero_fields else 0) | \
                (0 if nan_check else UVH_NO_NAN_CHECK) | (UVH_FP32 if fp32 else 0)
        R = np.zeros(3 * self.nn)
        Kval = None if ronly else np.zeros(self.nnz)
        Tint = np.zeros(3 * self.nn) if want_tf else None
        Fext = np.zeros(3 * self.nn) if want_tf else None
        st = self.lib.uvh_assemble(self.h, C.byref(ctrl), C.byref(fields), flags, _ptr(R), _ptr(Kval),
                                   _ptr(Tint), _ptr(Fext))
        self._check(st, "uvh_assemble")
        return dict(R=R, Kval=Kval, Tint=Tint, Fext=Fext)

    def assemble_device(self, ctrl: UvhCtrl, fields: UvhFields, flags: int, R, K=None, Tint=None, Fext=None,
                        stream=None):
        """Device-buffer assembly (torch tensors or raw device pointers); enqueue only."""
        st = self.lib.uvh_assemble_device(self.h, C.byref(ctrl), C.byref(fields), int(flags), _ptr(R), _ptr(K),
                                          _ptr(Tint), _ptr(Fext), stream)
        self._check(st, "uvh_assemble_device")

    def sync_status(self):
        self._check(self.lib.uvh_sync_status(self.h), "uvh_sync_status")

    def status_async(self, dst, stream=None):
        """Enqueue a copy of the status word (bit 0 NaN in R, bit 1 NaN in K) into ``dst`` (a
        pinned host or device int32 tensor, or a pointer) on ``stream``."""
        self._check(self.lib.uvh_status_async(self.h, _ptr(dst), stream), "uvh_status_async")

    @staticmethod
    def status_error(word: int):
        """The reference's error for a status word, or None (uvhMatrixAssembly.m:575-584)."""
        if word & 2:
            return UaError("", " NaN in K ")
        if word & 1:
            return UaError("uvhMatrixAssembly:NaNinR", " NaN in R ")
        return None

    # ---------------------------------------------------------------- uv assembly
    def uv_pattern(self):
        """(rowptr int64[2Nn+1], colind int32[4 nnzN]) of the uv K."""
        neq, nnz = C.c_int64(), C.c_int64()
        self._check(self.lib.uvh_uv_pattern_info(self.h, C.byref(neq), C.byref(nnz)), "uvh_uv_pattern_info")
        rowptr = np.zeros(neq.value + 1, dtype=np.int64)
        colind = np.zeros(nnz.value, dtype=np.int32)
        self._check(self.lib.uvh_uv_get_pattern(self.h, _ptr(rowptr), _ptr(colind)), "uvh_uv_get_pattern")
        return rowptr, colind

    def assemble_uv(self, CtrlVar, F, ronly=False, want_tf=False, nan_check=True, fp32=False, BCs=None):
        """uv (diagnostic) assembly from host arrays (``F`` with the upper surface ``s``);
        returns dict(R (2Nn), Kval (4 nnzN) or None, Tint, Fext). With
        ``CtrlVar.uvAssembly.ZeroFields`` the velocities are zero except at the fixed nodes of
        ``BCs``, which take their fixed values (``uvMatrixAssemblySSTREAM.m:24-28``); the
        library assembles those velocities."""
        ctrl = make_ctrl(CtrlVar, F)
        uvc = make_uv_ctrl(CtrlVar)
        if CtrlVar.uvAssembly.ZeroFields:
            F = F.copy()
            F.ub, F.vb = np.zeros(self.nn), np.zeros(self.nn)
            if BCs is not None:
                F.ub[np.asarray(BCs.ubFixedNode, dtype=np.int64)] = BCs.ubFixedValue
                F.vb[np.asarray(BCs.vbFixedNode, dtype=np.int64)] = BCs.vbFixedValue
        keep = []
        fields = make_uv_fields(F, keep)
        flags = (UVH_RONLY if ronly else 0) | (0 if nan_check else UVH_NO_NAN_CHECK) | (UVH_FP32 if fp32 else 0)
        nnzN = self.nnz // 9
        R = np.zeros(2 * self.nn)
        Kval = None if ronly else np.zeros(4 * nnzN)
        Tint = np.zeros(2 * self.nn) if want_tf else None
        Fext = np.zeros(2 * self.nn) if want_tf else None
        st = self.lib.uvh_uv_assemble(self.h, C.byref(ctrl), C.byref(uvc), C.byref(fields), flags, _ptr(R),
                                      _ptr(Kval), _ptr(Tint), _ptr(Fext))
        self._check(st, "uvh_uv_assemble")
        return dict(R=R, Kval=Kval, Tint=Tint, Fext=Fext)

    def assemble_uv_device(self, ctrl: UvhCtrl, uvc: UvhUvCtrl, fields: UvhUvFields, flags: int, R, K=None,
                           Tint=None, Fext=None, stream=None):
        st = self.lib.uvh_uv_assemble_device(self.h, C.byref(ctrl), C.byref(uvc), C.byref(fields), int(flags),
                                             _ptr(R), _ptr(K), _ptr(Tint), _ptr(Fext), stream)
        self._check(st, "uvh_uv_assemble_device")

    def uv_K_matrix(self, Kval):
        rowptr, colind = self.uv_pattern()
        return sp.csr_matrix((Kval, colind, rowptr), shape=(2 * self.nn, 2 * self.nn))

    def profile(self, enable: bool):
        self._check(self.lib.uvh_profile(self.h, int(enable)), "uvh_profile")

    def kernel_times(self):
        cap = 32
        names = (C.c_char_p * cap)()
        tot = (C.c_double * cap)()
        cnt = (C.c_int64 * cap)()
        n = self.lib.uvh_kernel_times(self.h, names, tot, cnt, cap)
        return {names[k].decode(): (tot[k], cnt[k]) for k in range(n)}

    def csc_map(self):
        """int64[nnz]: CSC values of K are ``Kval[map]`` on the CSR index arrays (symmetric pattern)."""
        if self._csc_map is None:
            m = np.zeros(self.nnz, dtype=np.int64)
            self._check(self.lib.uvh_get_csc_map(self.h, _ptr(m)), "uvh_get_csc_map")
            self._csc_map = m
        return self._csc_map

    def K_matrix(self, Kval, fmt="csr"):
        """K as a scipy sparse matrix; ``fmt="csc"`` permutes the values through ``csc_map``
        (the layout of MATLAB's sparse) without a sort."""
        rowptr, colind = self.pattern()
        if fmt == "csc":
            return sp.csc_matrix((Kval[self.csc_map()], colind, rowptr), shape=(self.neq, self.neq))
        return sp.csr_matrix((Kval, colind, rowptr), shape=(self.neq, self.neq))


def mesh_key(MUA):
    c = MUA.connectivity
    return (MUA.Nele, MUA.nod, MUA.Nnodes, MUA.nip, id(MUA), int(c[0, 0]), int(c[-1, -1]))


_contexts: dict = {}


def get_context(MUA, device: int = 0) -> UvhContext:
    """Per-device context with ``MUA`` resident (re-uploaded when the mesh changes)."""
    ctx = _contexts.get(device)
    if ctx is None:
        ctx = UvhContext(device)
        _contexts[device] = ctx
    key = mesh_key(MUA)
    if ctx.mesh_key != key:
        ctx.set_mesh(MUA)
    return ctx


def uvhAssembly(UserVar, RunInfo, CtrlVar, MUA, F0, F1, l1, BCs1, K_format="csr"):
    """Drop-in for ``[UserVar,RunInfo,R,K]=uvhAssembly(UserVar,RunInfo,CtrlVar,MUA,F0,F1,l1,BCs1)``."""
    t0 = time.perf_counter()
    check_supported(CtrlVar)
    ctx = get_context(MUA)
    ronly = bool(CtrlVar.uvhMatrixAssembly.Ronly)
    zf = bool(CtrlVar.uvhMatrixAssembly.ZeroFields)
    out = ctx.assemble(CtrlVar, F0, F1, ronly=ronly, zero_fields=zf)
    R = out["R"]
    K = None if ronly else ctx.K_matrix(out["Kval"], K_format)
    dt = time.perf_counter() - t0
    if RunInfo is not None:
        cpu = getattr(RunInfo, "CPU", None)
        if cpu is None:
            RunInfo.CPU = cpu = Struct(Assembly=Struct(uvh=0.0))
        cpu.Assembly.uvh = cpu.Assembly.uvh + dt
    return UserVar, RunInfo, R, K


def UaRunInfo():
    """Minimal ``RunInfo`` with the assembly timer (``UaMain/UaRunInfo.m:103-106``)."""
    return Struct(CPU=Struct(Assembly=Struct(uvh=0.0), Solution=Struct(uvh=0.0)),
                  Forward=Struct(uvhIterations=0, uvhConverged=False))
