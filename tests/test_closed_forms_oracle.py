"""Pins of the SUPG parameter tau and of Glen's viscosity at n = 3 on both oracle restatements
(O1 numpy, O2 C), against closed forms that share no code with them (tests/closed_forms.py).

* tau: the coordinate-weighted thickness rows of R on uniform fields equal
  -rho [(h0 - h1) + dt a] sum_e A_e (xbar_e - xc + tau0_e u0 / 2), for each tau option of
  SUPGtau.m (taus, taut, tau1, tau2, tau0), with beta0 and the tau multiplier away from 1
  (UaMain/SUPGtau.m:24-77, uvhAssemblyIntPointImplicitSUPG.m:323-325,390,
  uvhNodalLoopSSTREAM.m:55).
* eta at n = 3: with linear velocities the coordinate-weighted momentum rows of Tint equal the
  closed-form integrals of h eta (4 exx + 2 eyy) etc. with eta = 1/2 A^(-1/3) e^(-2/3) + etaZero
  (EffectiveViscositySSTREAM.m:7-9), on 3-, 6- and 10-node meshes.
* Planted mistakes (in O1, on the CPU, through monkeypatching): l = sqrt(A), a dropped 1/2 in
  the SUPG weight, and a wrong rate-factor exponent each make the check fail; the closed forms
  computed with a slip (l = sqrt(A), no 1/2, speed0 from u1, wrong exponents) are far from
  what both restatements give.
"""
import numpy as np
import pytest

import oracle.uvh_oracle as O1
from oracle import o2
from uasource_amd.ctrlvar import default_ctrlvar
from uasource_amd.fields import WORKLOADS, build_mesh
from uasource_amd.mesh import CreateMUA, jittered_delaunay

import closed_forms as cf

TAU_OPTIONS = ["taus", "taut", "tau1", "tau2", "tau0"]
_meshes = {}


def mesh(name):
    if name not in _meshes:
        _meshes[name] = build_mesh(WORKLOADS[name])
    return _meshes[name]


def tau_cv(option):
    cv = default_ctrlvar()
    cv.dt = cf.UNIFORM["dt"]
    cv.uvh.SUPG.tau = option
    cv.uvh.SUPG.tauMultiplier = 1.3
    cv.SUPG.beta0 = 0.7
    return cv


def check_tau(M, cv, option, Rh, slip=None):
    mx, my, xc, yc, scale = cf.tau_moments(M, cv, option, slip)
    gx, gy = cf.measured_tau_moments(M, Rh, xc, yc)
    if option == "tau0":
        return abs(gx) <= 1e-12 * scale and abs(gy) <= 1e-12 * scale
    return abs(gx / mx - 1) <= 1e-10 and abs(gy / my - 1) <= 1e-10


def o1_Rh(M, cv, F0, F1):
    return O1.uvh_matrix_assembly(cv, M, F0, F1, None)["R"][2 * M.Nnodes:]


def o2_Rh(M, cv, F0, F1):
    out = o2.assemble(cv, M, F0, F1, ronly=True)
    assert out["status"] == 0
    return out["R"][2 * M.Nnodes:]


@pytest.mark.parametrize("option", TAU_OPTIONS)
@pytest.mark.parametrize("name", ["tiny_p1", "tiny_p2", "tiny_p3"])
def test_supg_tau_closed_form_o1_o2(name, option):
    M = mesh(name)
    cv = tau_cv(option)
    F0, F1 = cf.uniform_fields(M)
    for Rh in (o1_Rh(M, cv, F0, F1), o2_Rh(M, cv, F0, F1)):
        assert check_tau(M, cv, option, Rh)


@pytest.mark.parametrize("name", ["tiny_p1", "tiny_p2", "tiny_p3"])
def test_supg_tau_term_is_resolved(name):
    """The tau term is a visible part of the weighted moment, so the bar of 1e-10 resolves any
    slip in it: the forms with l = sqrt(A), without the 1/2, or with speed0 from u1 differ
    from the right one by far more than the bar, and the oracles match only the right one."""
    M = mesh(name)
    cv = tau_cv("taus")
    F0, F1 = cf.uniform_fields(M)
    Rh = o2_Rh(M, cv, F0, F1)
    for slip in ("l=sqrt(A)", "no-half", "speed-from-u1"):
        mx, my, *_ = cf.tau_moments(M, cv, "taus", slip)
        rx, ry, *_ = cf.tau_moments(M, cv, "taus")
        assert abs(mx / rx - 1) > 1e-3 or abs(my / ry - 1) > 1e-3
        assert not check_tau(M, cv, "taus", Rh, slip)


@pytest.mark.parametrize("planted", ["l=sqrt(A)", "no-half"])
def test_supg_tau_check_catches_a_planted_slip_in_o1(monkeypatch, planted):
    M = mesh("tiny_p2")
    cv = tau_cv("taus")
    F0, F1 = cf.uniform_fields(M)
    real = O1.supg_tau

    def slipped(CtrlVar, v, l, dt, option, multiplier):
        if planted == "l=sqrt(A)":
            return real(CtrlVar, v, l / np.sqrt(2.0), dt, option, multiplier)
        return real(CtrlVar, v, l, dt, option, 2.0 * multiplier)   # the 1/2 of psi dropped

    monkeypatch.setattr(O1, "supg_tau", slipped)
    assert not check_tau(M, cv, "taus", o1_Rh(M, cv, F0, F1))
    monkeypatch.setattr(O1, "supg_tau", real)
    assert check_tau(M, cv, "taus", o1_Rh(M, cv, F0, F1))


def glen_mesh(nod):
    xy, tri = jittered_delaunay(20e3, 10e3, 7, 5, 0.1, seed=5)
    deg = {3: 2, 6: 5, 10: 8}[nod]
    return CreateMUA(xy, tri, TriNodes=nod, QuadratureRuleDegree=deg)


def check_glen(M, cv, Tint, slip=None, n_glen=3.0):
    fxx, fxy, fyy, _ = cf.glen_moments(M, cv, n_glen=n_glen, slip=slip)
    gxx, gxy, gyy = cf.measured_glen_moments(M, Tint)
    return all(abs(g / f - 1) <= 1e-11 for g, f in ((gxx, fxx), (gxy, fxy), (gyy, fyy)))


@pytest.mark.parametrize("n_glen", [3.0, 1.0, 4.0])
@pytest.mark.parametrize("nod", [3, 6, 10])
def test_glen_viscosity_closed_form_o1_o2(nod, n_glen):
    M = glen_mesh(nod)
    cv = default_ctrlvar()
    F0, F1 = cf.linear_glen_fields(M, n_glen=n_glen)
    t1 = O1.uvh_matrix_assembly(cv, M, F0, F1, None)["Tint"]
    t2 = o2.assemble(cv, M, F0, F1, ronly=True)["Tint"]
    assert check_glen(M, cv, t1, n_glen=n_glen)
    assert check_glen(M, cv, t2, n_glen=n_glen)
    if n_glen != 1.0:   # at n = 1 the strain-rate exponent vanishes
        for slip in ("exponent", "rate-factor"):
            assert not check_glen(M, cv, t2, slip, n_glen=n_glen)


def test_glen_check_catches_a_planted_slip_in_o1(monkeypatch):
    M = glen_mesh(6)
    cv = default_ctrlvar()
    F0, F1 = cf.linear_glen_fields(M)
    real = O1.effective_viscosity

    def slipped(CtrlVar, A, n, exx, eyy, exy):   # A^(-1/(2n)) instead of A^(-1/n)
        return real(CtrlVar, np.sqrt(A), n, exx, eyy, exy)

    monkeypatch.setattr(O1, "effective_viscosity", slipped)
    assert not check_glen(M, cv, O1.uvh_matrix_assembly(cv, M, F0, F1, None)["Tint"])
