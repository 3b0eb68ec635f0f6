"""The device R alone against the closed forms of tests/closed_forms.py, with no oracle in
the loop: the coordinate-weighted thickness rows pin the SUPG parameter tau for every option
of UaMain/SUPGtau.m (beta0 and the multiplier away from 1), and the coordinate-weighted
momentum rows of Tint pin Glen's viscosity at n = 3 (EffectiveViscositySSTREAM.m:7-9). FP64
to the bars of the CPU tests (1e-10 on the tau moment, 1e-11 on the viscous moments), on one-
and many-tile meshes of 3-, 6- and 10-node elements; the precision form to 1e-5 on the tau
moment (a slip in tau moves it by 20 % or more)."""
import numpy as np
import pytest

from uasource_amd import uvh
from uasource_amd.ctrlvar import default_ctrlvar
from uasource_amd.fields import WORKLOADS, build_mesh
from uasource_amd.mesh import CreateMUA, jittered_delaunay

import closed_forms as cf

pytestmark = pytest.mark.gpu

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


def device_R(M, cv, F0, F1, fp32=False, want_tf=False):
    ctx = uvh.UvhContext(0)
    try:
        ctx.set_mesh(M)
        return ctx.assemble(cv, F0, F1, ronly=True, want_tf=want_tf, fp32=fp32)
    finally:
        ctx.close()


@pytest.mark.parametrize("option", TAU_OPTIONS)
@pytest.mark.parametrize("name", ["tiny_p1", "tiny_p2", "tiny_p3", "medium_p2", "medium_p3"])
def test_device_supg_tau_closed_form(name, option):
    print(f"[gpu] SUPG tau closed form {name} {option}", flush=True)
    M = mesh(name)
    cv = tau_cv(option)
    F0, F1 = cf.uniform_fields(M)
    Rh = device_R(M, cv, F0, F1)["R"][2 * M.Nnodes:]
    mx, my, xc, yc, scale = cf.tau_moments(M, cv, option)
    gx, gy = cf.measured_tau_moments(M, Rh, xc, yc)
    if option == "tau0":
        print(f"   moments {gx:.2e} {gy:.2e} (scale {scale:.2e})", flush=True)
        assert abs(gx) <= 1e-12 * scale and abs(gy) <= 1e-12 * scale
    else:
        print(f"   x moment / closed form - 1 = {gx / mx - 1:.1e}, y: {gy / my - 1:.1e}", flush=True)
        assert abs(gx / mx - 1) <= 1e-10 and abs(gy / my - 1) <= 1e-10


@pytest.mark.parametrize("name", ["tiny_p2", "medium_p3"])
def test_device_supg_tau_closed_form_precision_form(name):
    print(f"[gpu] SUPG tau closed form, precision form {name}", flush=True)
    M = mesh(name)
    cv = tau_cv("taus")
    F0, F1 = cf.uniform_fields(M)
    Rh = device_R(M, cv, F0, F1, fp32=True)["R"][2 * M.Nnodes:]
    mx, my, xc, yc, _ = cf.tau_moments(M, cv, "taus")
    gx, gy = cf.measured_tau_moments(M, Rh, xc, yc)
    print(f"   x moment / closed form - 1 = {gx / mx - 1:.1e}, y: {gy / my - 1:.1e}", flush=True)
    assert abs(gx / mx - 1) <= 1e-5 and abs(gy / my - 1) <= 1e-5


@pytest.mark.parametrize("nod", [3, 6, 10])
@pytest.mark.parametrize("grid", [(7, 5), (61, 31)])
def test_device_glen_viscosity_n3_closed_form(nod, grid):
    print(f"[gpu] Glen n=3 closed form nod={nod} grid={grid}", flush=True)
    xy, tri = jittered_delaunay(20e3 * grid[0] / 7, 10e3 * grid[1] / 5, grid[0], grid[1], 0.1, seed=5)
    M = CreateMUA(xy, tri, TriNodes=nod, QuadratureRuleDegree={3: 2, 6: 5, 10: 8}[nod])
    cv = default_ctrlvar()
    F0, F1 = cf.linear_glen_fields(M, n_glen=3.0)
    Tint = device_R(M, cv, F0, F1, want_tf=True)["Tint"]
    fxx, fxy, fyy, eta = cf.glen_moments(M, cv, n_glen=3.0)
    gxx, gxy, gyy = cf.measured_glen_moments(M, Tint)
    errs = [abs(g / f - 1) for g, f in ((gxx, fxx), (gxy, fxy), (gyy, fyy))]
    print(f"   eta = {eta:.6e}; moment errors {errs[0]:.1e} {errs[1]:.1e} {errs[2]:.1e}", flush=True)
    assert max(errs) <= 1e-11
