"""ctypes front end of oracle O2 (``oracle/uvh_oracle_c.c``) -- TEST INFRASTRUCTURE ONLY.

Used by ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg.
``build()`` compiles the library with ``oracle/Makefile`` when it is missing or
older than its source.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_uvh.so")
SRC = os.path.join(HERE, "uvh_oracle_c.c")

TAU_OPTIONS = {"taus": 0, "taut": 1, "tau1": 2, "tau2": 3, "tau0": 4}


class O2Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("dt", "theta", "beta0", "tau_multiplier", "speed_zero", "eps_zero",
                                          "eta_zero", "czero", "kH", "Hh0", "g", "rhow", "alpha")] + \
               [("tau_option", C.c_int), ("ronly", C.c_int), ("zero_fields", C.c_int),
                ("reset_thickness", C.c_int), ("thick_min", C.c_double), ("law", C.c_int), ("nzero", C.c_double),
                ("enforce_b_above", C.c_int)]

# CtrlVar.SlidingLaw (UaMain/BasalDrag.m:182-257) -> the restatement's law code
SLIDING_LAWS = {"Weertman": 0, "W": 0, "Budd": 1, "W-N0": 1, "Coulomb": 2, "C": 2, "Tsai": 3, "minCW-N0": 3,
                "Cornford": 4, "rpCW-N0": 4, "Umbi": 5, "rCW-N0": 5, "Joughin": 6, "rCW-v0": 6}


def _optional(F, k):
    """F.q / F.muk as a contiguous array, or None when F does not carry it."""
    a = getattr(F, k, None)
    if a is None or np.size(a) == 0:
        return None
    return np.ascontiguousarray(a, dtype=np.float64)


def build(force: bool = False) -> str:
    stale = (not os.path.exists(LIB)) or os.path.getmtime(LIB) < os.path.getmtime(SRC)
    if force or stale:
        subprocess.run(["make", "-s", "-B", "-C", HERE, "liboracle_uvh.so"], check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        _lib.o2_node_graph.restype = C.c_int64
        _lib.o2_node_graph.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.o2_assemble.restype = C.c_int
        _lib.o2_assemble.argtypes = [C.POINTER(O2Params), C.c_int, C.c_int, C.c_int64, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        _lib.o2_uv_assemble.restype = C.c_int
        _lib.o2_uv_assemble.argtypes = [C.POINTER(O2Params), C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                        C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_int]
        _lib.o2_shape.restype = None
        _lib.o2_shape.argtypes = [C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def shape(nod: int, xi: float, eta: float):
    N, Nxi, Neta = (np.zeros(nod) for _ in range(3))
    lib().o2_shape(nod, xi, eta, _p(N), _p(Nxi), _p(Neta))
    return N, Nxi, Neta


def params_from(CtrlVar, F1, ronly=None, zero_fields=None) -> O2Params:
    P = O2Params()
    P.dt = CtrlVar.dt
    P.theta = CtrlVar.theta
    P.beta0 = CtrlVar.SUPG.beta0
    P.tau_multiplier = CtrlVar.uvh.SUPG.tauMultiplier
    P.speed_zero = CtrlVar.SpeedZero
    P.eps_zero = CtrlVar.EpsZero
    P.eta_zero = CtrlVar.etaZero
    P.czero = CtrlVar.Czero
    P.kH = CtrlVar.kH
    P.Hh0 = CtrlVar.Hh0
    P.g = F1.g
    P.rhow = F1.rhow
    P.alpha = F1.alpha
    P.tau_option = TAU_OPTIONS[CtrlVar.uvh.SUPG.tau]
    P.ronly = int(CtrlVar.uvhMatrixAssembly.Ronly if ronly is None else ronly)
    P.zero_fields = int(CtrlVar.uvhMatrixAssembly.ZeroFields if zero_fields is None else zero_fields)
    P.reset_thickness = int(bool(CtrlVar.ResetThicknessToMinThickness) and bool(CtrlVar.ResetThicknessInNonLinLoop))
    P.thick_min = CtrlVar.ThickMin
    P.law = SLIDING_LAWS[str(CtrlVar.get("SlidingLaw", "Weertman"))]
    P.nzero = float(CtrlVar.get("Nzero", 1e-20))
    P.enforce_b_above = int(bool(CtrlVar.get("Enforce_bAboveB", False)))
    return P


class O2Mesh:
    """Mesh-side arrays for repeated O2 calls (pattern, work buffers)."""

    def __init__(self, MUA):
        self.MUA = MUA
        self.conn = np.ascontiguousarray(MUA.connectivity, dtype=np.int32)
        self.x = np.ascontiguousarray(MUA.coordinates[:, 0])
        self.y = np.ascontiguousarray(MUA.coordinates[:, 1])
        self.pts = np.ascontiguousarray(MUA.points, dtype=np.float64)
        self.wts = np.ascontiguousarray(MUA.weights, dtype=np.float64)
        nele, nod, nn = MUA.Nele, MUA.nod, MUA.Nnodes
        L = lib()
        nnz = L.o2_node_graph(nele, nod, nn, _p(self.conn), None, None)
        self.rowptr_n = np.zeros(nn + 1, dtype=np.int64)
        self.colind_n = np.zeros(nnz, dtype=np.int32)
        L.o2_node_graph(nele, nod, nn, _p(self.conn), _p(self.rowptr_n), _p(self.colind_n))
        self.nnzN = int(nnz)
        self.Rel = np.zeros(nele * nod * 6)
        self.Kel = None
        self.b = np.zeros(nn)

    def csr_pattern(self):
        """3Nn x 3Nn CSR pattern: row (blk, node) = [adj(node); adj+Nn; adj+2Nn]."""
        nn = self.MUA.Nnodes
        deg = np.diff(self.rowptr_n)
        rowptr = np.zeros(3 * nn + 1, dtype=np.int64)
        for bi in range(3):
            rowptr[bi * nn + 1:(bi + 1) * nn + 1] = bi * 3 * self.nnzN + 3 * self.rowptr_n[1:]
        colind = np.empty(9 * self.nnzN, dtype=np.int64)
        node_of_entry = np.repeat(np.arange(nn), deg)
        start3 = 3 * self.rowptr_n[:-1]
        for bi in range(3):
            base = bi * 3 * self.nnzN
            for bj in range(3):
                pos = base + start3[node_of_entry] + bj * deg[node_of_entry] + \
                    (np.arange(self.nnzN) - self.rowptr_n[node_of_entry])
                colind[pos] = self.colind_n + bj * nn
        return rowptr, colind


def assemble(CtrlVar, MUA, F0, F1, mesh: O2Mesh | None = None, nthreads: int = 0,
             ronly=None, zero_fields=None, want_K=True, dadh=None):
    """Run O2. Returns dict(R, Tint, Fext, Kval, K (scipy CSR or None), status, mesh).
    ``dadh`` (Nn, optional): d(as + ab)/dh at the nodes, the mass-balance feedback term of the
    Jacobian (uvhMatrixAssembly.m:151-158); zero when absent."""
    if mesh is None:
        mesh = O2Mesh(MUA)
    P = params_from(CtrlVar, F1, ronly, zero_fields)
    nele, nod, nn = MUA.Nele, MUA.nod, MUA.Nnodes
    f1 = [np.ascontiguousarray(getattr(F1, k), dtype=np.float64)
          for k in ("h", "ub", "vb", "S", "B", "rho", "C", "m", "n", "AGlen", "as_", "ab")]
    f1 += [_optional(F1, "q"), _optional(F1, "muk"), _optional(F1, "V0"),
           None if dadh is None else np.ascontiguousarray(dadh, dtype=np.float64)]
    f0 = [np.ascontiguousarray(getattr(F0, k), dtype=np.float64) for k in ("h", "ub", "vb", "as_", "ab")]
    f1p = (C.c_void_p * 16)(*[None if a is None else a.ctypes.data for a in f1])
    f0p = (C.c_void_p * 5)(*[a.ctypes.data for a in f0])
    R = np.zeros(3 * nn)
    Tint = np.zeros(3 * nn)
    Fext = np.zeros(3 * nn)
    Kval = None
    if not P.ronly and want_K:
        if mesh.Kel is None:
            mesh.Kel = np.zeros(nele * 9 * nod * nod)
        Kval = np.zeros(9 * mesh.nnzN)
    st = lib().o2_assemble(C.byref(P), nele, nod, nn, _p(mesh.conn), _p(mesh.x), _p(mesh.y), MUA.nip,
                           _p(mesh.pts), _p(mesh.wts), C.cast(f1p, C.c_void_p), C.cast(f0p, C.c_void_p),
                           _p(mesh.rowptr_n), _p(mesh.colind_n), _p(R), _p(Tint), _p(Fext), _p(Kval),
                           _p(mesh.Rel), _p(mesh.Kel) if Kval is not None else None, _p(mesh.b), int(nthreads))
    K = None
    if Kval is not None:
        rowptr, colind = mesh.csr_pattern()
        K = sp.csr_matrix((Kval, colind, rowptr), shape=(3 * nn, 3 * nn))
    return dict(R=R, Tint=Tint, Fext=Fext, Kval=Kval, K=K, status=st, mesh=mesh)


UV_FIELDS = ("h", "s", "ub", "vb", "S", "B", "rho", "C", "m", "n", "AGlen")


def uv_csr_pattern(mesh: O2Mesh):
    """CSR pattern of the uv K (2 x 2 node blocks): row (bi, node) = [adj(node); adj(node) + nn]."""
    nn = mesh.MUA.Nnodes
    rowptr_n, colind_n = mesh.rowptr_n, mesh.colind_n
    deg = np.diff(rowptr_n)
    rowptr = np.zeros(2 * nn + 1, dtype=np.int64)
    rowptr[1:] = np.concatenate([np.cumsum(2 * deg), 2 * mesh.nnzN + np.cumsum(2 * deg)])
    colind = np.empty(4 * mesh.nnzN, dtype=np.int64)
    node_of_entry = np.repeat(np.arange(nn), deg)
    start2 = 2 * rowptr_n[:-1]
    for bi in range(2):
        for bj in range(2):
            pos = bi * 2 * mesh.nnzN + start2[node_of_entry] + bj * deg[node_of_entry] + \
                (np.arange(mesh.nnzN) - rowptr_n[node_of_entry])
            colind[pos] = colind_n + bj * nn
    return rowptr, colind


def assemble_uv(CtrlVar, MUA, F, mesh: O2Mesh | None = None, ronly: bool = False, nthreads: int = 0,
                want_K: bool = True, BCs=None):
    """uv (diagnostic) assembly, UaMain/uvMatrixAssemblySSTREAM.m. ``F`` carries the surface
    ``s`` besides the uvh fields. ZeroFields (:24-28): ub = vb = 0, then the fixed values of
    ``BCs`` at their nodes. Returns dict(R, Tint, Fext (2nn), Kval, K (CSR or None), status,
    mesh)."""
    if mesh is None:
        mesh = O2Mesh(MUA)
    zf = CtrlVar.get("uvAssembly")
    if zf is not None and zf.get("ZeroFields", False):
        F = F.copy()
        F.ub = F.ub * 0
        F.vb = F.vb * 0
        if BCs is not None:
            F.ub[BCs.ubFixedNode] = BCs.ubFixedValue
            F.vb[BCs.vbFixedNode] = BCs.vbFixedValue
    P = params_from(CtrlVar, F, ronly=ronly, zero_fields=False)
    P.reset_thickness = 0
    dvisk = 0.0 if CtrlVar.Picard else float(CtrlVar.NRviscosity)
    nn = MUA.Nnodes
    arrs = [np.ascontiguousarray(getattr(F, k), dtype=np.float64) for k in UV_FIELDS]
    arrs += [_optional(F, "q"), _optional(F, "muk"), _optional(F, "V0")]
    fp = (C.c_void_p * len(arrs))(*[None if a is None else a.ctypes.data for a in arrs])
    R, Tint, Fext = np.zeros(2 * nn), np.zeros(2 * nn), np.zeros(2 * nn)
    Kval = np.zeros(4 * mesh.nnzN) if (want_K and not ronly) else None
    st = lib().o2_uv_assemble(C.byref(P), CtrlVar.Cmin, CtrlVar.AGlenmin, dvisk, MUA.Nele, MUA.nod, nn,
                              _p(mesh.conn), _p(mesh.x), _p(mesh.y), MUA.nip, _p(mesh.pts), _p(mesh.wts),
                              C.cast(fp, C.c_void_p), _p(mesh.rowptr_n), _p(mesh.colind_n), _p(R), _p(Tint),
                              _p(Fext), _p(Kval), int(nthreads))
    K = None
    if Kval is not None:
        rowptr, colind = uv_csr_pattern(mesh)
        K = sp.csr_matrix((Kval, colind, rowptr), shape=(2 * nn, 2 * nn))
    return dict(R=R, Tint=Tint, Fext=Fext, Kval=Kval, K=K, status=st, mesh=mesh)
