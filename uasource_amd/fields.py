"""Field containers (Úa's ``UaFields`` / ``BoundaryConditions`` subset) and the
seeded synthetic workloads the assembly is measured on.

Units are Úa's: metres, years, kPa; ``g = 9.81/1000``
(``UaMain/DefineGeometryAndDensities.m:56``). The Python attribute ``as_``
holds Úa's ``F.as`` (``as`` is a Python keyword).

Workloads (all seeded; the seed is recorded in ``Workload.seed``):

* ``cpu_reference``: 100 km x 50 km, 46 x 23 jittered grid, 3-node triangles,
  quadrature degree 2 (3 points): 1,980 elements, 1,058 nodes.
* ``plain_p2``: 400 km x 100 km, 633 x 159 grid, 6-node triangles, degree 5
  (7 points): 199,712 elements, 401,005 nodes.
* ``precision_p3``: 400 km x 100 km, 1001 x 251 grid, 10-node triangles,
  degree 8 (16 points), plus a Gaussian bump in the bed: 500,000 elements.

Field recipe: bed ``B = -500 - 0.002 x'``, thickness ``h = 1000 - 0.006 x'``,
``S = 0``, ``rho = 917``, ``rhow = 1030``, ``AGlen = 1e-8``, ``n = 3``,
``C = 1e-2 * lognormal(sigma)``, ``m = 3``, ``u, v ~ U(0, 500)``; ``F0`` equals
``F1`` except for its own seeded ``u0, v0``; ``as = 0.3``, ``ab = 0``, ``dt = 1``.
On the 400 km domains the recipe is stretched, ``x' = x * 100 km / Lx``, so the
thickness stays positive and the grounding line keeps its relative position
(the literal recipe gives h < 0 beyond x = 166.7 km).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .ctrlvar import Struct
from .mesh import CreateMUA, MUA, jittered_delaunay

G_UA = 9.81 / 1000.0
RHO_ICE = 917.0
RHO_WATER = 1030.0

FIELD_NAMES_F1 = ("h", "ub", "vb", "S", "B", "rho", "C", "m", "n", "AGlen", "as_", "ab")
FIELD_NAMES_F0 = ("h", "ub", "vb", "as_", "ab")


def UaFields(Nnodes: int, **values) -> Struct:
    """Nodal fields (``UaMain/UaFields.m``); unspecified arrays are zeros."""
    F = Struct(g=G_UA, rhow=RHO_WATER, alpha=0.0)
    for name in FIELD_NAMES_F1:
        setattr(F, name, np.zeros(Nnodes))
    for k, v in values.items():
        setattr(F, k, v)
    return F


def BoundaryConditions(**values) -> Struct:
    """Boundary conditions (``UaMain/BoundaryConditions.m``); only the h masks reach the assembly."""
    e = np.zeros(0, dtype=np.int64)
    BCs = Struct(ubFixedNode=e, ubFixedValue=np.zeros(0), vbFixedNode=e, vbFixedValue=np.zeros(0),
                 hFixedNode=e, hFixedValue=np.zeros(0), hPosNode=e, hPosValue=np.zeros(0))
    for k, v in values.items():
        setattr(BCs, k, v)
    return BCs


@dataclass
class Workload:
    name: str
    Lx: float
    Ly: float
    nx: int
    ny: int
    tri_nodes: int
    degree: int
    seed: int
    bump: bool = False
    C_sigma: float = 0.2
    A_sigma: float = 0.0
    jitter: float = 0.1


WORKLOADS = {
    "tiny_p1": Workload("tiny_p1", 20e3, 10e3, 9, 6, 3, 2, seed=11),
    "tiny_p2": Workload("tiny_p2", 20e3, 10e3, 7, 5, 6, 5, seed=12),
    "tiny_p3": Workload("tiny_p3", 20e3, 10e3, 6, 4, 10, 8, seed=13),
    "cpu_reference": Workload("cpu_reference", 100e3, 50e3, 46, 23, 3, 2, seed=1),
    "medium_p2": Workload("medium_p2", 100e3, 50e3, 61, 31, 6, 5, seed=22),
    "medium_p3": Workload("medium_p3", 100e3, 50e3, 41, 21, 10, 8, seed=33, bump=True),
    "plain_p2": Workload("plain_p2", 400e3, 100e3, 633, 159, 6, 5, seed=2),
    "precision_p3": Workload("precision_p3", 400e3, 100e3, 1001, 251, 10, 8, seed=3, bump=True),
    # Newton solve with a host direct solver: the plain_p2 recipe on a 161 x 41 grid
    "newton_p2": Workload("newton_p2", 400e3, 100e3, 161, 41, 6, 5, seed=2),
}


def build_mesh(w: Workload) -> MUA:
    xy, tri = jittered_delaunay(w.Lx, w.Ly, w.nx, w.ny, w.jitter, seed=w.seed)
    return CreateMUA(xy, tri, TriNodes=w.tri_nodes, QuadratureRuleDegree=w.degree)


def make_fields(MUA: MUA, w: Workload, seed_offset: int = 0):
    """Seeded ``(F0, F1, BCs)`` on ``MUA`` following the recipe in the module docstring."""
    ss = np.random.SeedSequence([w.seed, 1000 + seed_offset])
    r_c, r_a, r_uv1, r_uv0, r_bump = (np.random.default_rng(s) for s in ss.spawn(5))
    x = MUA.coordinates[:, 0]
    y = MUA.coordinates[:, 1]
    Nn = MUA.Nnodes
    xs = x * (100e3 / w.Lx) if w.Lx > 100e3 else x
    B = -500.0 - 0.002 * xs
    if w.bump:
        cx = r_bump.uniform(0.3, 0.7) * w.Lx
        cy = r_bump.uniform(0.3, 0.7) * w.Ly
        B = B + 300.0 * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2.0 * 20e3 ** 2))
    h = 1000.0 - 0.006 * xs
    C = 1e-2 * np.exp(w.C_sigma * r_c.standard_normal(Nn))
    A = 1e-8 * np.exp(w.A_sigma * r_a.standard_normal(Nn)) if w.A_sigma > 0 else np.full(Nn, 1e-8)
    u1 = r_uv1.uniform(0.0, 500.0, Nn)
    v1 = r_uv1.uniform(0.0, 500.0, Nn)
    u0 = r_uv0.uniform(0.0, 500.0, Nn)
    v0 = r_uv0.uniform(0.0, 500.0, Nn)
    common = dict(S=np.zeros(Nn), B=B, rho=np.full(Nn, RHO_ICE), C=C, m=np.full(Nn, 3.0),
                  n=np.full(Nn, 3.0), AGlen=A, as_=np.full(Nn, 0.3), ab=np.zeros(Nn))
    F1 = UaFields(Nn, h=h.copy(), ub=u1, vb=v1, **{k: v.copy() for k, v in common.items()})
    F0 = UaFields(Nn, h=h.copy(), ub=u0, vb=v0, **{k: v.copy() for k, v in common.items()})
    return F0, F1, BoundaryConditions()


def newton_problem(MUA: MUA, w: Workload):
    """``(F0, F1, BCs, fixed)`` of the Newton-solve workload on ``MUA``: the geometry and
    physics of ``make_fields``, a smooth start (``ub = 100 + 400 x/Lx`` m/a, ``vb = 0``, the
    same in F0), and Dirichlet conditions: ub, vb and h held at the inflow boundary x = 0 and
    vb held on the sides y = 0 and y = Ly. ``fixed`` indexes those degrees of freedom in
    ``[ub; vb; h]``. (The assembly workloads' velocities are uniform random; Newton from such
    a state, or with no inflow condition on h, needs Úa's Cauchy-step fall-backs.)"""
    F0, F1, _ = make_fields(MUA, w)
    x = MUA.coordinates[:, 0]
    y = MUA.coordinates[:, 1]
    Nn = MUA.Nnodes
    F1.ub = 100.0 + 400.0 * (x - x.min()) / (x.max() - x.min())
    F1.vb = np.zeros(Nn)
    F0.ub, F0.vb = F1.ub.copy(), F1.vb.copy()
    tol = 1e-6 * max(x.max() - x.min(), y.max() - y.min())
    inflow = np.where(x <= x.min() + tol)[0]
    side = np.where((y <= y.min() + tol) | (y >= y.max() - tol))[0]
    fixed = np.unique(np.concatenate([inflow, Nn + inflow, 2 * Nn + inflow, Nn + side]))
    BCs = BoundaryConditions(ubFixedNode=inflow, ubFixedValue=F1.ub[inflow], vbFixedNode=np.union1d(inflow, side),
                             vbFixedValue=np.zeros(len(np.union1d(inflow, side))), hFixedNode=inflow,
                             hFixedValue=F1.h[inflow])
    return F0, F1, BCs, fixed


def batch_workloads(n: int = 64, seed: int = 4, lo: int = 5000, hi: int = 80000, types=((3, 2), (6, 5)),
                    weights=None, name: str = "batch", C_sigma: float = 0.3, A_sigma: float = 0.2) -> list:
    """Ensemble of ``n`` independent meshes: element counts log-uniform in [lo, hi], element
    types drawn from ``types`` ((nodes, quadrature degree) pairs, with ``weights``), each member
    with its own seeded lognormal C (``C_sigma``) and AGlen (``A_sigma``) perturbations, on
    2:1 rectangles of 100 km x 50 km. (On 10-node elements a C perturbation of sigma 0.3 can
    make the interpolated C + Czero negative at a point, which Úa reports as a NaN in R.)"""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        nele = float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        t = types[rng.choice(len(types), p=weights)]
        cells_y = max(2, int(round(np.sqrt(nele / 4.0))))
        out.append(Workload(f"{name}{k}", 100e3, 50e3, 2 * cells_y + 1, cells_y + 1, t[0], t[1],
                            seed=10_000 + seed * 1000 + k, C_sigma=C_sigma, A_sigma=A_sigma))
    return out


def serve_workloads(n_meshes: int = 24, seed: int = 17) -> list:
    """Meshes of the serving workload: element counts lognormal around 20k (sigma 1) clipped to
    [2k, 300k], element types P1/P2/P3 = 50/35/15 % with 3/7/16 points, C lognormal (sigma 0.2)."""
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.exp(rng.normal(np.log(20000), 1.0, n_meshes)), 2000, 300000)
    types = [((3, 2), (6, 5), (10, 8))[rng.choice(3, p=(0.5, 0.35, 0.15))] for _ in range(n_meshes)]
    return [batch_workloads(n=1, seed=100 + k, lo=nele, hi=nele * 1.0001, types=(t,), name=f"srv{k}",
                            C_sigma=0.2, A_sigma=0.0)[0] for k, (nele, t) in enumerate(zip(sizes, types))]


def surface_from_floatation(F, CtrlVar):
    """Upper surface s = b + h with b from the floatation of h over the bed B and sea level S
    (UaMain/Calc_bs_From_hBS.m:71-79): the input the uv assembly takes as ``F.s``."""
    hf = F.rhow * (F.S - F.B) / F.rho
    GF = 0.5 + 0.5 * np.tanh(CtrlVar.kH * ((F.h - hf) - CtrlVar.Hh0))
    b = GF * F.B + (1.0 - GF) * (F.S - F.rho * F.h / F.rhow)
    return b + F.h
