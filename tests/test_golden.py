"""Golden fixtures (inputs + oracle-O1 outputs, written by tests/golden/make_golden.py):
both oracles reproduce them on the CPU, and the HIP path reproduces them on the GPU."""
import numpy as np
import pytest

from golden_io import FIXTURES, load
from oracle import uvh_oracle as O1


def _close(ref, got, tol):
    scale = max(np.linalg.norm(ref["Tint"]), np.linalg.norm(ref["Fext"]))
    assert np.linalg.norm(got["R"] - ref["R"]) <= tol * scale
    for k in ("Tint", "Fext"):
        assert np.linalg.norm(got[k] - ref[k]) <= tol * max(np.linalg.norm(ref[k]), 1e-300)


@pytest.mark.parametrize("name", FIXTURES)
def test_o1_reproduces_fixture(name):
    cv, M, F0, F1, BCs, ref = load(name)
    out = O1.uvh_matrix_assembly(cv, M, F0, F1, BCs)
    _close(ref, out, 1e-15)
    assert np.array_equal(out["K"].indptr, ref["K"].indptr) and np.array_equal(out["K"].indices, ref["K"].indices)
    assert np.linalg.norm(out["K"].data - ref["K"].data) <= 1e-15 * np.linalg.norm(ref["K"].data)


@pytest.mark.parametrize("name", FIXTURES)
def test_o2_reproduces_fixture(name, o2lib):
    cv, M, F0, F1, BCs, ref = load(name)
    out = o2lib.assemble(cv, M, F0, F1)
    _close(ref, out, 1e-13)
    d = out["K"] - ref["K"]
    assert abs(d).max() <= 1e-12 * abs(ref["K"]).max()


@pytest.mark.gpu
@pytest.mark.parametrize("name", FIXTURES)
def test_gpu_reproduces_fixture(name):
    from uasource_amd import uvh
    cv, M, F0, F1, BCs, ref = load(name)
    print(f"[gpu] golden {name}", flush=True)
    ctx = uvh.UvhContext(0)
    try:
        ctx.set_mesh(M)
        out = ctx.assemble(cv, F0, F1, zero_fields=cv.uvhMatrixAssembly.ZeroFields, want_tf=True)
        rowptr, colind = ctx.pattern()
    finally:
        ctx.close()
    _close(ref, out, 1e-12)
    assert np.array_equal(rowptr, ref["K"].indptr) and np.array_equal(colind, ref["K"].indices)
    assert np.max(np.abs(out["Kval"] - ref["K"].data)) <= 1e-12 * np.max(np.abs(ref["K"].data))
