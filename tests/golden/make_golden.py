"""Writes the golden fixtures under tests/golden/ from oracle O1 (the MATLAB-order
numpy restatement). Each fixture holds its inputs (mesh, quadrature, fields) and
outputs (R, Tint, Fext, K as CSR), so the tests do not depend on the mesh
generator reproducing bit-for-bit on another machine.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from oracle import uvh_oracle as O1  # noqa: E402
from uasource_amd.ctrlvar import default_ctrlvar  # noqa: E402
from uasource_amd.fields import WORKLOADS, build_mesh, make_fields, FIELD_NAMES_F1, FIELD_NAMES_F0  # noqa: E402

CASES = [("tiny_p1", False), ("tiny_p2", False), ("tiny_p3", False), ("cpu_reference", False),
         ("cpu_reference", True)]


def main():
    for name, zero_fields in CASES:
        w = WORKLOADS[name]
        M = build_mesh(w)
        F0, F1, BCs = make_fields(M, w)
        cv = default_ctrlvar()
        cv.uvhMatrixAssembly.ZeroFields = zero_fields
        out = O1.uvh_matrix_assembly(cv, M, F0, F1, BCs)
        K = out["K"]
        arrs = dict(coordinates=M.coordinates, connectivity=M.connectivity.astype(np.int32),
                    points=M.points, weights=M.weights, seed=np.array(w.seed), zero_fields=np.array(zero_fields),
                    R=out["R"], Tint=out["Tint"], Fext=out["Fext"],
                    K_indptr=K.indptr.astype(np.int64), K_indices=K.indices.astype(np.int32), K_data=K.data)
        for k in FIELD_NAMES_F1:
            arrs["F1_" + k] = np.asarray(getattr(F1, k))
        for k in FIELD_NAMES_F0:
            arrs["F0_" + k] = np.asarray(getattr(F0, k))
        fn = os.path.join(HERE, f"{name}{'_zerofields' if zero_fields else ''}.npz")
        np.savez_compressed(fn, **arrs)
        print(fn, os.path.getsize(fn))


if __name__ == "__main__":
    main()
