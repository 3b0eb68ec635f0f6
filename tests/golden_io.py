"""Loads a golden fixture (tests/golden/*.npz) into MUA / F0 / F1 objects."""
import os

import numpy as np
import scipy.sparse as sp

from uasource_amd.ctrlvar import default_ctrlvar
from uasource_amd.fields import UaFields, BoundaryConditions, FIELD_NAMES_F1, FIELD_NAMES_F0
from uasource_amd.mesh import MUA, tri_area_fe

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz"))


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    conn = z["connectivity"].astype(np.int64)
    M = MUA(coordinates=z["coordinates"], connectivity=conn, points=z["points"], weights=z["weights"],
            EleAreas=tri_area_fe(z["coordinates"], conn), QuadratureRuleDegree=-1)
    Nn = M.Nnodes
    F1 = UaFields(Nn, **{k: z["F1_" + k] for k in FIELD_NAMES_F1})
    F0 = UaFields(Nn, **{k: z["F0_" + k] for k in FIELD_NAMES_F0})
    for k in ("S", "B", "rho", "C", "m", "n", "AGlen"):
        setattr(F0, k, getattr(F1, k))
    cv = default_ctrlvar()
    cv.uvhMatrixAssembly.ZeroFields = bool(z["zero_fields"])
    K = sp.csr_matrix((z["K_data"], z["K_indices"], z["K_indptr"]), shape=(3 * Nn, 3 * Nn))
    ref = dict(R=z["R"], Tint=z["Tint"], Fext=z["Fext"], K=K)
    return cv, M, F0, F1, BoundaryConditions(), ref
