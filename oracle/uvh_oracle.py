"""Oracle O1 -- TEST INFRASTRUCTURE ONLY. FP64 numpy restatement of Úa's uvh assembly.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may
import this module, and only as the checker. The product path
(``uasource_amd``, ``include/uvh.h``, ``libuvh.so``) never calls it.

This file mirrors the MATLAB code's vectorisation and operation order: arrays are
element-vectorised (``Nele``-long columns), there is a ``for Iint`` loop over the
integration points and ``for Inod``/``for Jnod`` loops inside it, ``Tint`` and
``Fext`` are summed separately, and ``K`` is built from COO triplets whose
duplicates are summed (MATLAB ``sparse``). Each function cites the reference lines
it restates:

* ``uvh_matrix_assembly``  -- ``UaMain/uvhMatrixAssembly.m:48-101,128-141,185-266,300-313,356-383,416-509,575-584``
* ``_int_point``           -- ``UaMain/uvhAssemblyIntPointImplicitSUPG.m:27-126,206-308,323-395,409-419``
* ``_nodal_loop``          -- ``UaMain/uvhNodalLoopSSTREAM.m:39-228`` (2025 form, lines 141-157)
* ``effective_viscosity``  -- ``UaMain/EffectiveViscositySSTREAM.m:7-9``
* ``basal_drag_weertman``  -- ``UaMain/BasalDrag.m:54-77,184-186,345-358,410-443``
* ``heaviside_approx``     -- ``UaMain/HeavisideApprox.m:101``
* ``dirac_delta``          -- ``UaMain/DiracDelta.m:29``
* ``supg_tau``             -- ``UaMain/SUPGtau.m:24-25,45-77``
* ``calc_bs_from_hBS``     -- ``UaMain/Calc_bs_From_hBS.m:35-44,71-79,82-109``
* ``shape_fun``            -- ``UaMain/shape_fun.m:52-82`` (2-D, 3/6/10 nodes)
* ``shape_der``            -- ``UaMain/shape_der.m:54-120`` (2-D, 3/6/10 nodes)
* ``deriv_vector``         -- ``UaMain/derivVector.m:25-26,96-109``

How each part is pinned without a runnable reference (the reference is MATLAB and
cannot run here; it ships no fixtures for this path) is recorded in
``tests/test_oracle_invariants.py``: quadrature exactness, shape-function
identities, closed forms of the physics, K = dR/d(u,v,h) by central differences,
block symmetries, rigid-translation, mass and P1 closed-form element matrices,
and agreement with the independent C restatement (oracle O2). Functions of MATLAB
itself (``sparse`` summation order, libm ``tanh``/``sech``/``power`` rounding) are
parity unpinned; they sit far below the 1e-12 tolerance.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


class UaError(RuntimeError):
    """An error with a MATLAB-style identifier, e.g. ``uvhMatrixAssembly:NaNinR``."""

    def __init__(self, identifier: str, message: str):
        super().__init__(f"{identifier}: {message}")
        self.identifier = identifier


# ----------------------------------------------------------------------------- shapes

def shape_fun(Iint: int, nod: int, points: np.ndarray) -> np.ndarray:
    """Values of the ``nod`` form functions at integration point ``Iint`` (0-based)."""
    c1 = points[Iint, 0]
    c2 = points[Iint, 1]
    c3 = 1.0 - c1 - c2
    if nod == 3:
        return np.array([c1, c3, c2])
    if nod == 6:
        return np.array([(2 * c1 - 1) * c1, 4 * c3 * c1, (2 * c3 - 1) * c3,
                         4 * c2 * c3, (2 * c2 - 1) * c2, 4 * c1 * c2])
    if nod == 10:
        f = np.empty(10)
        f[0] = ((3 * c1 - 1) * (3 * c1 - 2) * c1) / 2
        f[1] = -(9 * (3 * c1 - 1) * (c1 + c2 - 1) * c1) / 2
        f[2] = (9 * (3 * c1 + 3 * c2 - 2) * (c1 + c2 - 1) * c1) / 2
        f[3] = -((3 * c1 + 3 * c2 - 1) * (3 * c1 + 3 * c2 - 2) * (c1 + c2 - 1)) / 2
        f[4] = (9 * (3 * c1 + 3 * c2 - 2) * (c1 + c2 - 1) * c2) / 2
        f[5] = -(9 * (c1 + c2 - 1) * (3 * c2 - 1) * c2) / 2
        f[6] = ((3 * c2 - 1) * (3 * c2 - 2) * c2) / 2
        f[7] = (9 * (3 * c2 - 1) * c1 * c2) / 2
        f[8] = (9 * (3 * c1 - 1) * c1 * c2) / 2
        f[9] = -27 * ((c2 - 1) + c1) * c1 * c2
        return f
    raise ValueError("wrong number of nodes in shape_fun")


def shape_der(Iint: int, nod: int, points: np.ndarray) -> np.ndarray:
    """``2 x nod`` derivatives (d/dxi; d/deta) of the form functions at ``Iint``."""
    c1 = points[Iint, 0]
    c2 = points[Iint, 1]
    c3 = 1.0 - c1 - c2
    der = np.zeros((2, nod))
    if nod == 3:
        der[0] = [1.0, -1.0, 0.0]
        der[1] = [0.0, -1.0, 1.0]
    elif nod == 6:
        der[0] = [4 * c1 - 1, 4 * (c3 - c1), -(4 * c3 - 1), -4 * c2, 0.0, 4 * c2]
        der[1] = [0.0, -4 * c1, -(4 * c3 - 1), 4 * (c3 - c2), 4 * c2 - 1, 4 * c1]
    elif nod == 10:
        der[0, 0] = (27 * c1 ** 2 - 18 * c1 + 2) / 2
        der[0, 8] = (9 * (6 * c1 - 1) * c2) / 2
        der[0, 7] = (9 * (3 * c2 - 1) * c2) / 2
        der[0, 6] = 0.0
        der[0, 5] = -(9 * (3 * c2 - 1) * c2) / 2
        der[0, 4] = (9 * (6 * c1 + 6 * c2 - 5) * c2) / 2
        der[0, 3] = -(27 * c1 ** 2 + 54 * c1 * c2 - 36 * c1 + 27 * c2 ** 2 - 36 * c2 + 11) / 2
        der[0, 2] = (9 * (9 * c1 ** 2 + 12 * c1 * c2 - 10 * c1 + 3 * c2 ** 2 - 5 * c2 + 2)) / 2
        der[0, 1] = -(9 * (9 * c1 ** 2 + 6 * c1 * c2 - 8 * c1 - c2 + 1)) / 2
        der[0, 9] = -27 * (((c2 - 1) + c1) + c1) * c2
        der[1, 0] = 0.0
        der[1, 8] = (9 * (3 * c1 - 1) * c1) / 2
        der[1, 7] = (9 * (6 * c2 - 1) * c1) / 2
        der[1, 6] = (27 * c2 ** 2 - 18 * c2 + 2) / 2
        der[1, 5] = -(9 * ((c1 + c2 - 1) * (6 * c2 - 1) + (3 * c2 - 1) * c2)) / 2
        der[1, 4] = (9 * (3 * c1 ** 2 + 12 * c1 * c2 - 5 * c1 + 9 * c2 ** 2 - 10 * c2 + 2)) / 2
        der[1, 3] = -(27 * c1 ** 2 + 54 * c1 * c2 - 36 * c1 + 27 * c2 ** 2 - 36 * c2 + 11) / 2
        der[1, 2] = (9 * (6 * c1 + 6 * c2 - 5) * c1) / 2
        der[1, 1] = -(9 * (3 * c1 - 1) * c1) / 2
        der[1, 9] = -27 * (((c2 - 1) + c1) + c2) * c1
    else:
        raise ValueError("wrong number of nodes in shape_der")
    return der


def deriv_vector(coordinates, connectivity, points, Iint):
    """``Deriv`` (Nele x 2 x nod) and ``detJ`` (Nele) at integration point ``Iint``."""
    Nele, nod = connectivity.shape
    coox = coordinates[connectivity, 0]
    cooy = coordinates[connectivity, 1]
    der = shape_der(Iint, nod, points)
    J11 = coox @ der[0]
    J12 = cooy @ der[0]
    J21 = coox @ der[1]
    J22 = cooy @ der[1]
    detJ = J11 * J22 - J12 * J21
    Deriv = np.zeros((Nele, 2, nod))
    for J in range(nod):
        Deriv[:, 0, J] = (J22 * der[0, J] - J12 * der[1, J]) / detJ
        Deriv[:, 1, J] = (-J21 * der[0, J] + J11 * der[1, J]) / detJ
    return Deriv, detJ


# ----------------------------------------------------------------------------- physics

def heaviside_approx(k, x, x0):
    return 0.5 + 0.5 * np.tanh(k * (x - x0))


def dirac_delta(k, x, x0):
    return 0.5 * k * (1.0 / np.cosh(k * (x - x0))) ** 2


def effective_viscosity(CtrlVar, A, n, exx, eyy, exy):
    e = np.sqrt(CtrlVar.EpsZero ** 2 + exx ** 2 + eyy ** 2 + exx * eyy + exy ** 2)
    eta = 0.5 * A ** (-1.0 / n) * e ** ((1.0 - n) / n) + CtrlVar.etaZero
    E = (1.0 - n) / (4.0 * n) * A ** (-1.0 / n) * e ** ((1.0 - 3.0 * n) / n)
    return eta, E


def basal_drag_weertman(CtrlVar, He, delta, ub, vb, C, m):
    delta = np.where(delta < 0, 0.0, delta)
    He = np.where(He < 0, 0.0, He)
    C0 = CtrlVar.Czero
    u0 = CtrlVar.SpeedZero
    speed = np.sqrt(ub * ub + vb * vb + u0 ** 2)
    Um = speed ** (1.0 / m - 1.0)
    beta2 = (C + C0) ** (-1.0 / m) * Um
    Dbeta2 = (1.0 / m - 1.0) * (C + C0) ** (-1.0 / m) * (ub ** 2 + vb ** 2 + u0 ** 2) ** ((1.0 - 3.0 * m) / (2.0 * m))
    taux = He * beta2 * ub
    tauy = He * beta2 * vb
    dtauxdu = He * (beta2 + Dbeta2 * ub * ub)
    dtauydv = He * (beta2 + Dbeta2 * vb * vb)
    dtauxdv = He * Dbeta2 * ub * vb
    dtauydu = dtauxdv
    dtauxdh = delta * beta2 * ub
    dtauydh = delta * beta2 * vb
    return taux, tauy, dtauxdu, dtauxdv, dtauydu, dtauydv, dtauxdh, dtauydh


def supg_tau(CtrlVar, v, l, dt, option, multiplier):
    taut = dt / 2 + np.finfo(np.float64).eps
    taus = l / (2 * v + CtrlVar.SpeedZero)
    ECN = taut / taus
    if option == "taus":
        tau = taus
    elif option == "taut":
        tau = taut + 0 * v
    elif option == "tau2":
        tau = (dt / 2) * 1.0 / (1.0 + ECN)
    elif option == "tau1":
        with np.errstate(divide="ignore", invalid="ignore"):
            K = 1.0 / np.tanh(ECN) - 1.0 / ECN
        small = ECN < 1e-6
        K = np.where(small, ECN / 3, K)
        tau = K * taus
    elif option == "tau0":
        tau = np.zeros_like(v)
    else:
        raise UaError("SUPGtau", "in SUPGtau case not found")
    return multiplier * tau


def calc_bs_from_hBS(CtrlVar, h, S, B, rho, rhow):
    h = np.array(h, dtype=np.float64, copy=True)
    if CtrlVar.ResetThicknessToMinThickness:
        h = np.where(h < CtrlVar.ThickMin, CtrlVar.ThickMin, h)
    hf = rhow * (S - B) / rho
    GF = heaviside_approx(CtrlVar.kH, h - hf, CtrlVar.Hh0)
    bfloat = S - rho * h / rhow
    b = GF * B + (1 - GF) * bfloat
    if CtrlVar.Enforce_bAboveB:
        b = np.where(b < B, B, b)
    s = b + h
    return b, s


# ----------------------------------------------------------------------------- assembly

def _nodal_loop(detJw, nod, theta, tau0, Ronly, Deriv, fun,
                exx, eyy, exy, exx0, eyy0, dhdx, dhdy, dh0dx, dh0dy, drhodx, drhody, dbdx, dbdy,
                Hposint, Dddhint, ca, sa, g, dt, etaint, Eint,
                taux, tauy, dtauxdu, dtauxdv, dtauydu, dtauydv, dtauxdh, dtauydh,
                rhoint, rhow, uint, vint, u0int, v0int, dint, hint, h0int, a1int, a0int, dadhint):
    Nele = detJw.shape[0]
    h0barr = 0.0
    h1barr = 0.0
    lambda_h = 1.0
    out = {k: np.zeros((Nele, nod)) for k in ("Tx", "Fx", "Ty", "Fy", "Th", "Fh")}
    if not Ronly:
        for k in ("Kxu", "Kxv", "Kyu", "Kyv", "Kxh", "Kyh", "Khu", "Khv", "Khh"):
            out[k] = np.zeros((Nele, nod, nod))

    qx1dx = rhoint * exx * hint + rhoint * uint * dhdx + drhodx * uint * hint
    qy1dy = rhoint * eyy * hint + rhoint * vint * dhdy + drhody * vint * hint
    qx0dx = rhoint * exx0 * h0int + rhoint * u0int * dh0dx + drhodx * u0int * h0int
    qy0dy = rhoint * eyy0 * h0int + rhoint * v0int * dh0dy + drhody * v0int * h0int

    for Inod in range(nod):
        SUPG = fun[Inod] + 0.5 * tau0 * (u0int * Deriv[:, 0, Inod] + v0int * Deriv[:, 1, Inod])
        funI = fun[Inod]
        if not Ronly:
            for Jnod in range(nod):
                Deu = Eint * ((2 * exx + eyy) * Deriv[:, 0, Jnod] + exy * Deriv[:, 1, Jnod])
                Dev = Eint * ((2 * eyy + exx) * Deriv[:, 1, Jnod] + exy * Deriv[:, 0, Jnod])
                E11 = hint * (4 * exx + 2 * eyy) * Deu * Deriv[:, 0, Inod] + 2 * hint * exy * Deu * Deriv[:, 1, Inod]
                E12 = hint * (4 * exx + 2 * eyy) * Dev * Deriv[:, 0, Inod] + 2 * hint * exy * Dev * Deriv[:, 1, Inod]
                E22 = hint * (4 * eyy + 2 * exx) * Dev * Deriv[:, 1, Inod] + 2 * hint * exy * Dev * Deriv[:, 0, Inod]
                E21 = hint * (4 * eyy + 2 * exx) * Deu * Deriv[:, 1, Inod] + 2 * hint * exy * Deu * Deriv[:, 0, Inod]
                out["Kxu"][:, Inod, Jnod] += (4 * hint * etaint * Deriv[:, 0, Inod] * Deriv[:, 0, Jnod]
                                              + hint * etaint * Deriv[:, 1, Inod] * Deriv[:, 1, Jnod]
                                              + E11 + dtauxdu * fun[Jnod] * funI) * detJw
                out["Kyv"][:, Inod, Jnod] += (4 * hint * etaint * Deriv[:, 1, Inod] * Deriv[:, 1, Jnod]
                                              + hint * etaint * Deriv[:, 0, Inod] * Deriv[:, 0, Jnod]
                                              + E22 + dtauydv * fun[Jnod] * funI) * detJw
                out["Kxv"][:, Inod, Jnod] += (etaint * hint * (2 * Deriv[:, 0, Inod] * Deriv[:, 1, Jnod]
                                                               + Deriv[:, 1, Inod] * Deriv[:, 0, Jnod])
                                              + E12 + dtauxdv * fun[Jnod] * funI) * detJw
                out["Kyu"][:, Inod, Jnod] += (etaint * hint * (2 * Deriv[:, 1, Inod] * Deriv[:, 0, Jnod]
                                                               + Deriv[:, 0, Inod] * Deriv[:, 1, Jnod])
                                              + E21 + dtauydu * fun[Jnod] * funI) * detJw
                out["Kxh"][:, Inod, Jnod] += (etaint * (4 * exx + 2 * eyy) * Deriv[:, 0, Inod] * fun[Jnod]
                                              + etaint * 2 * exy * Deriv[:, 1, Inod] * fun[Jnod]
                                              + dtauxdh * funI * fun[Jnod]
                                              + ca * g * (rhoint - rhow * Dddhint) * dbdx * funI * fun[Jnod]
                                              - sa * g * rhoint * funI * fun[Jnod]
                                              - ca * g * (rhoint * hint - rhow * dint * Dddhint) * Deriv[:, 0, Inod] * fun[Jnod]
                                              ) * detJw
                out["Kyh"][:, Inod, Jnod] += (etaint * (4 * eyy + 2 * exx) * Deriv[:, 1, Inod] * fun[Jnod]
                                              + etaint * 2 * exy * Deriv[:, 0, Inod] * fun[Jnod]
                                              + dtauydh * funI * fun[Jnod]
                                              + ca * g * (rhoint - rhow * Dddhint) * dbdy * funI * fun[Jnod]
                                              - ca * g * (rhoint * hint - rhow * dint * Dddhint) * Deriv[:, 1, Inod] * fun[Jnod]
                                              ) * detJw
                out["Khu"][:, Inod, Jnod] += (theta * (rhoint * dhdx * fun[Jnod] + drhodx * hint * fun[Jnod]
                                                       + rhoint * hint * Deriv[:, 0, Jnod])
                                              * SUPG * detJw * dt)
                out["Khv"][:, Inod, Jnod] += (theta * (rhoint * dhdy * fun[Jnod] + drhody * hint * fun[Jnod]
                                                       + rhoint * hint * Deriv[:, 1, Jnod])
                                              * SUPG * detJw * dt)
                out["Khh"][:, Inod, Jnod] += ((rhoint * fun[Jnod]
                                               - dt * theta * rhoint * dadhint * fun[Jnod]
                                               + dt * theta * rhoint * fun[Jnod] * h1barr / lambda_h
                                               + dt * theta * (rhoint * exx * fun[Jnod] + drhodx * uint * fun[Jnod]
                                                               + rhoint * uint * Deriv[:, 0, Jnod]
                                                               + rhoint * eyy * fun[Jnod] + drhody * vint * fun[Jnod]
                                                               + rhoint * vint * Deriv[:, 1, Jnod]))
                                              * SUPG * detJw)
        t1 = -ca * g * (rhoint * hint - rhow * dint) * dbdx * funI + rhoint * g * hint * sa * funI
        t2 = 0.5 * ca * g * (rhoint * hint ** 2 - rhow * dint ** 2) * Deriv[:, 0, Inod]
        t3 = hint * etaint * (4 * exx + 2 * eyy) * Deriv[:, 0, Inod]
        t4 = hint * etaint * 2 * exy * Deriv[:, 1, Inod]
        t5 = taux * funI
        out["Tx"][:, Inod] += (t3 + t4 + t5) * detJw
        out["Fx"][:, Inod] += (t1 + t2) * detJw
        t1 = -ca * g * (rhoint * hint - rhow * dint) * dbdy * funI
        t2 = 0.5 * ca * g * (rhoint * hint ** 2 - rhow * dint ** 2) * Deriv[:, 1, Inod]
        t3 = hint * etaint * (4 * eyy + 2 * exx) * Deriv[:, 1, Inod]
        t4 = hint * etaint * 2 * exy * Deriv[:, 0, Inod]
        t5 = tauy * funI
        out["Ty"][:, Inod] += (t3 + t4 + t5) * detJw
        out["Fy"][:, Inod] += (t1 + t2) * detJw
        qterm = dt * (theta * qx1dx + (1 - theta) * qx0dx + theta * qy1dy + (1 - theta) * qy0dy) * SUPG
        dhdt = rhoint * (h0int - hint + dt * (1 - theta) * h0barr + dt * theta * h1barr) * SUPG
        accterm = dt * rhoint * ((1 - theta) * a0int + theta * a1int) * SUPG
        th = -dhdt
        fh = accterm - qterm
        out["Th"][:, Inod] += th * detJw
        out["Fh"][:, Inod] += fh * detJw
    return out


def _int_point(Iint, MUA, nodal, CtrlVar, rhow, g, Ronly, ca, sa, dt):
    nod = MUA.nod
    theta = CtrlVar.theta
    fun = shape_fun(Iint, nod, MUA.points)
    Deriv, detJ = deriv_vector(MUA.coordinates, MUA.connectivity, MUA.points, Iint)
    N = nodal
    hint = N["h"] @ fun
    uint = N["u"] @ fun
    vint = N["v"] @ fun
    Cint = N["C"] @ fun
    mint = N["m"] @ fun
    AGlenint = N["AGlen"] @ fun
    nint = N["n"] @ fun
    h0int = N["h0"] @ fun
    u0int = N["u0"] @ fun
    v0int = N["v0"] @ fun
    as0int = N["as0"] @ fun
    ab0int = N["ab0"] @ fun
    as1int = N["as1"] @ fun
    ab1int = N["ab1"] @ fun
    a1int = as1int + ab1int
    a0int = as0int + ab0int
    dadhint = N["dadh"] @ fun
    Bint = N["B"] @ fun
    Sint = N["S"] @ fun
    rhoint = N["rho"] @ fun
    Hint = Sint - Bint
    hfint = rhow * Hint / rhoint
    kH, Hh0 = CtrlVar.kH, CtrlVar.Hh0
    Heint = heaviside_approx(kH, hint - hfint, Hh0)
    HEint = heaviside_approx(kH, hfint - hint, Hh0)
    deltaint = dirac_delta(kH, hint - hfint, Hh0)
    Deltaint = dirac_delta(kH, hfint - hint, Hh0)
    Hposint = heaviside_approx(kH, Hint, Hh0) * Hint
    dint = HEint * rhoint * hint / rhow + Heint * Hposint
    Dddhint = HEint * rhoint / rhow - Deltaint * hint * rhoint / rhow + deltaint * Hposint

    Nele = MUA.Nele
    z = lambda: np.zeros(Nele)  # noqa: E731
    dhdx, dhdy, dh0dx, dh0dy = z(), z(), z(), z()
    exx0, eyy0, exx, eyy, exy = z(), z(), z(), z(), z()
    dbdx, dbdy, drhodx, drhody = z(), z(), z(), z()
    for Inod in range(nod):
        d1 = Deriv[:, 0, Inod]
        d2 = Deriv[:, 1, Inod]
        dhdx = dhdx + d1 * N["h"][:, Inod]
        dhdy = dhdy + d2 * N["h"][:, Inod]
        dh0dx = dh0dx + d1 * N["h0"][:, Inod]
        dh0dy = dh0dy + d2 * N["h0"][:, Inod]
        exx0 = exx0 + d1 * N["u0"][:, Inod]
        eyy0 = eyy0 + d2 * N["v0"][:, Inod]
        dbdx = dbdx + d1 * N["b"][:, Inod]
        dbdy = dbdy + d2 * N["b"][:, Inod]
        drhodx = drhodx + d1 * N["rho"][:, Inod]
        drhody = drhody + d2 * N["rho"][:, Inod]
        exx = exx + d1 * N["u"][:, Inod]
        eyy = eyy + d2 * N["v"][:, Inod]
        exy = exy + 0.5 * (d1 * N["v"][:, Inod] + d2 * N["u"][:, Inod])

    etaint, Eint = effective_viscosity(CtrlVar, AGlenint, nint, exx, eyy, exy)
    (taux, tauy, dtauxdu, dtauxdv, dtauydu, dtauydv, dtauxdh, dtauydh) = \
        basal_drag_weertman(CtrlVar, Heint, deltaint, uint, vint, Cint, mint)
    l = np.sqrt(2 * MUA.EleAreas)
    speed0 = np.sqrt(u0int * u0int + v0int * v0int + CtrlVar.SpeedZero ** 2)
    tau = supg_tau(CtrlVar, speed0, l, dt, CtrlVar.uvh.SUPG.tau, CtrlVar.uvh.SUPG.tauMultiplier)
    tau0 = CtrlVar.SUPG.beta0 * tau
    detJw = detJ * MUA.weights[Iint]
    out = _nodal_loop(detJw, nod, theta, tau0, Ronly, Deriv, fun,
                      exx, eyy, exy, exx0, eyy0, dhdx, dhdy, dh0dx, dh0dy, drhodx, drhody, dbdx, dbdy,
                      Hposint, Dddhint, ca, sa, g, dt, etaint, Eint,
                      taux, tauy, dtauxdu, dtauxdv, dtauydu, dtauydv, dtauxdh, dtauydh,
                      rhoint, rhow, uint, vint, u0int, v0int, dint, hint, h0int, a1int, a0int, dadhint)
    out["_ip"] = dict(taux=taux, tauy=tauy, eta=etaint, He=Heint)
    return out


BLOCK_NAMES = ("Kxu", "Kxv", "Kxh", "Kyu", "Kyv", "Kyh", "Khu", "Khv", "Khh")


def uvh_matrix_assembly(CtrlVar, MUA, F0, F1, BCs1=None, return_blocks=False, dadh=None):
    """Residual and Jacobian as Úa's ``uvhMatrixAssembly`` returns them.

    Returns a dict with ``R`` (3*Nn), ``K`` (scipy CSR, 3Nn x 3Nn, duplicates summed,
    explicit zeros kept; ``None`` when ``Ronly``), ``Tint``, ``Fext`` and, with
    ``return_blocks``, the element arrays ``Tx..Fh`` and ``Kxu..Khh``.
    """
    ZeroFields = bool(CtrlVar.uvhMatrixAssembly.ZeroFields)
    Ronly = bool(CtrlVar.uvhMatrixAssembly.Ronly)
    f1 = {k: np.array(getattr(F1, k), dtype=np.float64, copy=True)
          for k in ("h", "ub", "vb", "S", "B", "rho", "C", "m", "n", "AGlen", "as_", "ab")}
    f0 = {k: np.array(getattr(F0, k), dtype=np.float64, copy=True) for k in ("h", "ub", "vb", "as_", "ab")}
    if ZeroFields:
        f1["ub"] = f1["ub"] * 0
        f1["vb"] = f1["vb"] * 0
        f0["ub"] = f0["ub"] * 0
        f0["vb"] = f0["vb"] * 0
        f1["h"] = f0["h"].copy()
        f1["as_"] = np.abs(f1["ab"]) + 1
        f1["ab"] = np.abs(f1["ab"])
        f0["as_"] = np.abs(f0["ab"]) + 1
        f0["ab"] = np.abs(f0["ab"])
    g = F1.g
    alpha = F1.alpha
    rhow = F1.rhow
    dt = CtrlVar.dt
    cv = CtrlVar.copy()
    if not cv.ResetThicknessInNonLinLoop:
        cv.ResetThicknessToMinThickness = False
    b1, _ = calc_bs_from_hBS(cv, f1["h"], f1["S"], f1["B"], f1["rho"], rhow)
    # d(as + ab)/dh: F1.dasdh + F1.dabdh where the mass balance feeds back on the geometry
    # (MassBalanceGeometryFeedback 3, uvhMatrixAssembly.m:151-158), zero otherwise; the
    # caller passes it, the mass-balance callback being outside the assembly
    # (none under ZeroFields, :143, :168)
    dadh = np.zeros(MUA.Nnodes) if dadh is None or ZeroFields else np.asarray(dadh, dtype=np.float64)

    conn = MUA.connectivity
    Nn = MUA.Nnodes
    nod = MUA.nod
    Nele = MUA.Nele
    neq = 3 * Nn
    nodal = {
        "h": f1["h"][conn], "u": f1["ub"][conn], "v": f1["vb"][conn],
        "C": f1["C"][conn], "m": f1["m"][conn], "AGlen": f1["AGlen"][conn], "n": f1["n"][conn],
        "S": f1["S"][conn], "B": f1["B"][conn],
        "h0": f0["h"][conn], "u0": f0["ub"][conn], "v0": f0["vb"][conn],
        "as0": f0["as_"][conn], "ab0": f0["ab"][conn], "as1": f1["as_"][conn], "ab1": f1["ab"][conn],
        "dadh": dadh[conn], "rho": f1["rho"][conn], "b": b1[conn],
    }
    ca, sa = np.cos(alpha), np.sin(alpha)
    acc = {k: np.zeros((Nele, nod)) for k in ("Tx", "Fx", "Ty", "Fy", "Th", "Fh")}
    if not Ronly:
        for k in BLOCK_NAMES:
            acc[k] = np.zeros((Nele, nod, nod))
    ipinfo = []
    for Iint in range(MUA.nip):
        out = _int_point(Iint, MUA, nodal, CtrlVar, rhow, g, Ronly, ca, sa, dt)
        for k in acc:
            acc[k] = acc[k] + out[k]
        ipinfo.append(out["_ip"])

    # right-hand side, fewer-sparse-calls layout (uvhMatrixAssembly.m:416-448)
    iR = np.concatenate([np.concatenate([conn[:, i], conn[:, i] + Nn, conn[:, i] + 2 * Nn]) for i in range(nod)])
    Tval = np.concatenate([np.concatenate([acc["Tx"][:, i], acc["Ty"][:, i], acc["Th"][:, i]]) for i in range(nod)])
    Fval = np.concatenate([np.concatenate([acc["Fx"][:, i], acc["Fy"][:, i], acc["Fh"][:, i]]) for i in range(nod)])
    Tint = np.bincount(iR, weights=Tval, minlength=neq)
    Fext = np.bincount(iR, weights=Fval, minlength=neq)
    R = Tint - Fext

    K = None
    if not Ronly:
        offs = {"x": 0, "y": Nn, "h": 2 * Nn, "u": 0, "v": Nn}
        I, J, V = [], [], []
        for Inod in range(nod):
            for Jnod in range(nod):
                for name in BLOCK_NAMES:
                    ro = offs[name[1]]
                    co = offs["h" if name[2] == "h" else name[2]]
                    I.append(conn[:, Inod] + ro)
                    J.append(conn[:, Jnod] + co)
                    V.append(acc[name][:, Inod, Jnod])
        I = np.concatenate(I)
        J = np.concatenate(J)
        V = np.concatenate(V)
        K = sp.coo_matrix((V, (I, J)), shape=(neq, neq)).tocsr()
        K.sum_duplicates()
        K.sort_indices()
        if np.any(np.isnan(K.diagonal())):
            raise UaError("", " NaN in K ")  # no identifier (uvhMatrixAssembly.m:577)
    if np.any(np.isnan(R)):
        raise UaError("uvhMatrixAssembly:NaNinR", " NaN in R ")
    result = dict(R=R, K=K, Tint=Tint, Fext=Fext, b=b1)
    if return_blocks:
        result["blocks"] = acc
        result["ip"] = ipinfo
    return result
