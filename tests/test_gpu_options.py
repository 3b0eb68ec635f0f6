"""GPU parity of assembly options outside the default path, through the C ABI:

* the mass-balance feedback term dadh (uvhNodalLoopSSTREAM.m:206; passed as F1.dasdh +
  F1.dabdh by uvhMatrixAssembly.m:151-158) against the oracle O2 and against the closed form
  of tests/closed_forms.py (sum_ij x_j dKhh_ij = -dt theta rho c sum_e A_e xbar_e), with R
  unchanged by it and no term under ZeroFields (:143, :168);
* Enforce_bAboveB and ResetThicknessToMinThickness with ResetThicknessInNonLinLoop (the lower
  surface b, Calc_bs_From_hBS.m:35-45,82-93) against O2, on a state close to flotation.
Bar as tests/test_gpu_parity.py: 1e-12 normwise.
"""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import o2
from uasource_amd import uvh
from uasource_amd.ctrlvar import default_ctrlvar
from uasource_amd.fields import WORKLOADS, build_mesh, make_fields

import closed_forms as cf
from test_oracle_options import near_flotation_fields, option_cvs

pytestmark = pytest.mark.gpu
TOL = 1e-12
_meshes = {}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mesh(name):
    if name not in _meshes:
        _meshes[name] = build_mesh(WORKLOADS[name])
    return _meshes[name]


@pytest.fixture(scope="module")
def ctx():
    c = uvh.UvhContext(0)
    yield c
    c.close()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def check(M, gpu, ref):
    scale = max(np.linalg.norm(ref["Tint"]), np.linalg.norm(ref["Fext"]))
    assert np.linalg.norm(gpu["R"] - ref["R"]) / scale <= TOL
    assert rel(gpu["Kval"], ref["Kval"]) <= TOL


def k_matrix(M, mesh_o2, Kval):
    rowptr, colind = mesh_o2.csr_pattern()
    n3 = 3 * M.Nnodes
    return sp.csr_matrix((Kval, colind, rowptr), shape=(n3, n3))


@pytest.mark.parametrize("name", ["tiny_p1", "tiny_p2", "tiny_p3", "medium_p2"])
def test_dadh_gpu_vs_oracle(ctx, name):
    M = mesh(name)
    cv = default_ctrlvar()
    F0, F1, _ = make_fields(M, WORKLOADS[name])
    dadh = np.random.default_rng(7).uniform(-0.5, 0.5, M.Nnodes)
    say(f"[gpu] dadh {name}")
    ctx.set_mesh(M)
    gpu = ctx.assemble(cv, F0, F1, want_tf=True, dadh=dadh)
    ref = o2.assemble(cv, M, F0, F1, dadh=dadh)
    check(M, gpu, ref)
    # ZeroFields: no dadh term (the same bits with and without it)
    z1 = ctx.assemble(cv, F0, F1, want_tf=True, zero_fields=True, dadh=dadh)
    z0 = ctx.assemble(cv, F0, F1, want_tf=True, zero_fields=True)
    assert np.array_equal(z1["Kval"], z0["Kval"]) and np.array_equal(z1["R"], z0["R"])


@pytest.mark.parametrize("name", ["tiny_p1", "tiny_p2", "tiny_p3"])
def test_dadh_gpu_closed_form(ctx, name):
    M = mesh(name)
    cv = default_ctrlvar()
    cv.dt = cf.UNIFORM["dt"]
    F0, F1 = cf.uniform_fields(M)
    say(f"[gpu] dadh closed form {name}")
    ctx.set_mesh(M)
    g0 = ctx.assemble(cv, F0, F1, want_tf=True)
    g1 = ctx.assemble(cv, F0, F1, want_tf=True, dadh=np.full(M.Nnodes, cf.DADH))
    mo = o2.O2Mesh(M)
    gx, g1s = cf.measured_dadh_moments(M, k_matrix(M, mo, g0["Kval"]), k_matrix(M, mo, g1["Kval"]))
    mx, m1 = cf.dadh_moments(M, cv, cf.UNIFORM["rho"], cf.UNIFORM["dt"])
    assert abs(gx / mx - 1) <= 1e-10 and abs(g1s / m1 - 1) <= 1e-10
    assert np.array_equal(g0["R"], g1["R"])   # dadh enters the Jacobian only


@pytest.mark.parametrize("opt", ["enforce_b", "reset_h", "both"])
@pytest.mark.parametrize("name", ["tiny_p1", "tiny_p2", "tiny_p3", "medium_p2"])
def test_lower_surface_options_gpu_vs_oracle(ctx, name, opt):
    M = mesh(name)
    F0, F1 = near_flotation_fields(M, name)
    cv = option_cvs()[opt]
    say(f"[gpu] lower-surface option {opt} {name}")
    ctx.set_mesh(M)
    for fp32 in (False,):
        gpu = ctx.assemble(cv, F0, F1, want_tf=True, fp32=fp32)
        check(M, gpu, o2.assemble(cv, M, F0, F1))
