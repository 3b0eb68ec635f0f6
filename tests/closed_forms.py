"""Closed forms that pin the SUPG parameter tau and Glen's effective viscosity at n = 3
without reusing any oracle code. Shared by the CPU tests (O1, O2) and the device tests.

SUPG tau (UaMain/SUPGtau.m:24-25, 61-77; uvhAssemblyIntPointImplicitSUPG.m:323-325,390;
uvhNodalLoopSSTREAM.m:55). On uniform fields every gradient vanishes, so the thickness row of
node i is R_h,i = -rho [(h0 - h1) + dt ((1-theta) a0 + theta a1)] * sum_e int_e psi_i with the
SUPG weight psi_i = N_i + 1/2 tau0 (u0 dN_i/dx + v0 dN_i/dy). Weighted by the nodal coordinate,
sum_i x_i N_i = x and sum_i x_i dN_i/dx = 1, sum_i x_i dN_i/dy = 0 on straight-sided elements,
so

    sum_i (x_i - xc) R_h,i = -rho [...] * sum_e A_e (xbar_e - xc + 1/2 tau0_e u0)

(and the same in y with v0), where A_e is the element area, xbar_e its centroid and
tau0_e = beta0 * multiplier * tau(l_e = sqrt(2 A_e), speed0 = sqrt(u0^2 + v0^2 + SpeedZero^2)).
Taking xc as the area centroid of the mesh leaves the tau term alone. A shared slip such as
l = sqrt(A), a dropped 1/2 in psi, or speed0 built from u1 changes that term by 20 % or more.

Glen's law (EffectiveViscositySSTREAM.m:7-9): with velocities linear in x and y the strain
rates are constant, e = sqrt(EpsZero^2 + exx^2 + eyy^2 + exx eyy + exy^2) is constant, and so
is eta = 1/2 A^(-1/n) e^((1-n)/n) + etaZero, for any n.
"""
from __future__ import annotations

import numpy as np

from uasource_amd.fields import UaFields

CORNERS = {3: (0, 1, 2), 6: (0, 2, 4), 10: (0, 3, 6)}


def element_areas_centroids(M):
    """Areas and centroids of the straight-sided elements, from their corner coordinates."""
    c = CORNERS[M.nod]
    P = [M.coordinates[M.connectivity[:, k]] for k in c]
    d1, d2 = P[1] - P[0], P[2] - P[0]
    A = 0.5 * np.abs(d1[:, 0] * d2[:, 1] - d1[:, 1] * d2[:, 0])
    xc = (P[0][:, 0] + P[1][:, 0] + P[2][:, 0]) / 3.0
    yc = (P[0][:, 1] + P[1][:, 1] + P[2][:, 1]) / 3.0
    return A, xc, yc


def tau_definition(option, speed, l, dt, speed_zero):
    """tau of SUPGtau.m (before the multiplier), written from its definitions: taut = dt/2 + eps,
    taus = l / (2 v + SpeedZero), ECN = taut / taus; tau1 = (coth ECN - 1/ECN) taus (ECN/3 for
    ECN < 1e-6), tau2 = (dt/2) / (1 + ECN), tau0 = 0."""
    taut = dt / 2.0 + np.finfo(np.float64).eps
    taus = l / (2.0 * speed + speed_zero)
    ecn = taut / taus
    if option == "taus":
        return taus
    if option == "taut":
        return np.full_like(l, taut)
    if option == "tau2":
        return (dt / 2.0) / (1.0 + ecn)
    if option == "tau1":
        k = np.where(ecn < 1e-6, ecn / 3.0, np.cosh(ecn) / np.sinh(np.maximum(ecn, 1e-300)) - 1.0 / ecn)
        return k * taus
    if option == "tau0":
        return np.zeros_like(l)
    raise ValueError(option)


# uniform state of the tau test: h1 != h0, u1 != u0 (so a speed0 built from u1 shows), grounded
UNIFORM = dict(h1=1000.0, h0=990.0, u1=120.0, v1=-30.0, u0=100.0, v0=40.0, rho=917.0, B=-500.0,
               as1=0.3, ab1=-0.1, as0=0.5, ab0=0.05, dt=0.5)


def uniform_fields(M):
    n = M.Nnodes
    s = UNIFORM
    F1 = UaFields(n, h=np.full(n, s["h1"]), ub=np.full(n, s["u1"]), vb=np.full(n, s["v1"]), S=np.zeros(n),
                  B=np.full(n, s["B"]), rho=np.full(n, s["rho"]), C=np.full(n, 0.02), m=np.full(n, 3.0),
                  n=np.full(n, 3.0), AGlen=np.full(n, 1e-8), as_=np.full(n, s["as1"]), ab=np.full(n, s["ab1"]))
    F0 = F1.copy()
    F0.h, F0.ub, F0.vb = np.full(n, s["h0"]), np.full(n, s["u0"]), np.full(n, s["v0"])
    F0.as_, F0.ab = np.full(n, s["as0"]), np.full(n, s["ab0"])
    return F0, F1


def tau_moments(M, cv, option, slip=None):
    """Closed forms of sum_i (x_i - xc) R_h,i and sum_i (y_i - yc) R_h,i on `uniform_fields`,
    for cv's theta, beta0 and tau multiplier and the given tau option. `slip` computes the
    forms with a planted mistake instead: "l=sqrt(A)", "no-half" or "speed-from-u1"."""
    s = UNIFORM
    A, xb, yb = element_areas_centroids(M)
    xc = np.sum(A * xb) / np.sum(A)
    yc = np.sum(A * yb) / np.sum(A)
    l = np.sqrt(A) if slip == "l=sqrt(A)" else np.sqrt(2.0 * A)
    uu, vv = (s["u1"], s["v1"]) if slip == "speed-from-u1" else (s["u0"], s["v0"])
    speed = np.sqrt(uu * uu + vv * vv + cv.SpeedZero ** 2)
    tau0 = cv.SUPG.beta0 * cv.uvh.SUPG.tauMultiplier * tau_definition(option, speed, l, s["dt"], cv.SpeedZero)
    half = 1.0 if slip == "no-half" else 0.5
    th = cv.theta
    a0, a1 = s["as0"] + s["ab0"], s["as1"] + s["ab1"]
    f = -s["rho"] * ((s["h0"] - s["h1"]) + s["dt"] * ((1 - th) * a0 + th * a1))
    mx = f * np.sum(A * (xb - xc + half * tau0 * s["u0"]))
    my = f * np.sum(A * (yb - yc + half * tau0 * s["v0"]))
    # size of the individual terms, for the rounding bar of the tau-free option
    scale = abs(f) * np.sum(A * (np.abs(xb - xc) + np.abs(yb - yc)))
    return mx, my, xc, yc, scale


def measured_tau_moments(M, Rh, xc, yc):
    x, y = M.coordinates[:, 0], M.coordinates[:, 1]
    return float(np.dot(x - xc, Rh)), float(np.dot(y - yc, Rh))


# linear fields of the Glen-law identity (exactly represented by 3-node elements)
LIN = dict(h=(900.0, -0.002, 0.001), u=(100.0, 0.003, -0.001), v=(50.0, -0.002, 0.004))


def _lin(c, X, Y):
    return c[0] + c[1] * X + c[2] * Y


def linear_glen_fields(M, n_glen=3.0, A=2e-8, C=0.02):
    """h, u, v linear; grounded (B = -300); m = 1 so the drag u/C is linear too; n = n_glen."""
    x, y = M.coordinates[:, 0], M.coordinates[:, 1]
    N = M.Nnodes
    F1 = UaFields(N, h=_lin(LIN["h"], x, y), ub=_lin(LIN["u"], x, y), vb=_lin(LIN["v"], x, y), S=np.zeros(N),
                  B=np.full(N, -300.0), rho=np.full(N, 910.0), C=np.full(N, C), m=np.ones(N),
                  n=np.full(N, n_glen), AGlen=np.full(N, A), as_=np.full(N, 0.3), ab=np.zeros(N))
    F0 = F1.copy()
    return F0, F1


def glen_eta(cv, A, n, exx, eyy, exy, slip=None):
    """eta of EffectiveViscositySSTREAM.m:7-9 from its definition; `slip` plants a mistake:
    "exponent" uses e^((1-n)/(2n)), "rate-factor" uses A^(-1/(2n))."""
    e = np.sqrt(cv.EpsZero ** 2 + exx ** 2 + eyy ** 2 + exx * eyy + exy ** 2)
    pe = (1.0 - n) / (2.0 * n) if slip == "exponent" else (1.0 - n) / n
    pa = -1.0 / (2.0 * n) if slip == "rate-factor" else -1.0 / n
    return 0.5 * A ** pa * e ** pe + cv.etaZero


def integrate_poly(M, f):
    """Exact integral of a polynomial f(x, y) of degree <= 5 over the mesh (a degree-5 rule on
    each straight-sided element, mapped affinely)."""
    # Radon's 7-point degree-5 rule, points and weights (summing to 1) from their closed forms
    s15 = np.sqrt(15.0)
    a, b = (9 - 2 * s15) / 21, (6 + s15) / 21
    c, d = (9 + 2 * s15) / 21, (6 - s15) / 21
    pts = [(1 / 3, 1 / 3), (a, b), (b, a), (b, b), (c, d), (d, c), (d, d)]
    w = [9 / 40] + [(155 + s15) / 1200] * 3 + [(155 - s15) / 1200] * 3
    A, _, _ = element_areas_centroids(M)
    cidx = CORNERS[M.nod]
    P = [M.coordinates[M.connectivity[:, k]] for k in cidx]
    total = 0.0
    for (xi, et), wq in zip(pts, w):
        X = P[0] + xi * (P[1] - P[0]) + et * (P[2] - P[0])
        total += wq * np.sum(f(X[:, 0], X[:, 1]) * A)
    return total


def glen_moments(M, cv, n_glen=3.0, A=2e-8, C=0.02, slip=None):
    """Closed forms of sum_i x_i Tx_i, sum_i y_i Tx_i and sum_i y_i Ty_i on `linear_glen_fields`:
    int h eta (4 exx + 2 eyy) + x u / C, int 2 h eta exy + y u / C, int h eta (4 eyy + 2 exx) + y v / C."""
    hc, uc, vc = LIN["h"], LIN["u"], LIN["v"]
    exx, eyy, exy = uc[1], vc[2], 0.5 * (uc[2] + vc[1])
    eta = glen_eta(cv, A, n_glen, exx, eyy, exy, slip)
    fxx = integrate_poly(M, lambda X, Y: _lin(hc, X, Y) * eta * (4 * exx + 2 * eyy) + X * _lin(uc, X, Y) / C)
    fxy = integrate_poly(M, lambda X, Y: 2 * _lin(hc, X, Y) * eta * exy + Y * _lin(uc, X, Y) / C)
    fyy = integrate_poly(M, lambda X, Y: _lin(hc, X, Y) * eta * (4 * eyy + 2 * exx) + Y * _lin(vc, X, Y) / C)
    return fxx, fxy, fyy, eta


def measured_glen_moments(M, Tint):
    n = M.Nnodes
    x, y = M.coordinates[:, 0], M.coordinates[:, 1]
    return float(np.dot(x, Tint[:n])), float(np.dot(y, Tint[:n])), float(np.dot(y, Tint[n:2 * n]))


# Mass-balance feedback term of the Jacobian (uvhAssemblyIntPointImplicitSUPG.m:309-316,
# uvhNodalLoopSSTREAM.m:206: Khh gains -dt theta rho dadh psi_i N_j). With rho and dadh = c
# constant it is the only part of K that changes with dadh, and R does not change at all:
#   K(c) - K(0) on the h-h block = -dt theta rho c  int psi_i N_j.
# Summed over rows (sum_i psi_i = 1 on straight-sided elements) and weighted by the column
# node's coordinate (sum_j x_j N_j = x):
#   sum_ij x_j dKhh_ij = -dt theta rho c sum_e A_e xbar_e,   sum_ij dKhh_ij = -dt theta rho c sum_e A_e.
# A flipped sign, a missing theta or rho, or dadh weighting the row instead of the column
# misses by 100 % or more.
DADH = 0.37


def dadh_moments(M, cv, rho, dt):
    A, xb, _ = element_areas_centroids(M)
    f = -dt * cv.theta * rho * DADH
    return f * np.sum(A * xb), f * np.sum(A)


def measured_dadh_moments(M, K0, K1):
    """K0, K1: scipy sparse Jacobians without and with dadh (3Nn x 3Nn, unknowns u, v, h)."""
    n = M.Nnodes
    d = (K1 - K0).tocsr()[2 * n:, 2 * n:]
    x = M.coordinates[:, 0]
    return float(np.sum(d @ x)), float(np.sum(d @ np.ones(n)))
