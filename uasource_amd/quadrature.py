"""Triangle quadrature rules used by the uvh assembly.

Restates the fully symmetric, interior-point ("PI") non-product rules that Úa's
``quadtriangle(Degree,'Type','nonproduct','Points','inside','Domain',[0 0;1 0;0 1])``
returns when ``CtrlVar.QuadRules2021`` is true (``UaMain/CreateMUA.m:36-47``):

* degree 1:  1 point   (``UaMain/quadtriangle.m:447-454``)
* degree 2:  3 points  (``UaMain/quadtriangle.m:456-466``, upper signs of formula 2-1)
* degree 4:  6 points  (``UaMain/quadtriangle.m:505-517``)
* degree 5:  7 points  (``UaMain/quadtriangle.m:519-529``)
* degree 8: 16 points  (``UaMain/quadtriangle.m:700-746``)
* degree 10: 25 points (``UaMain/quadtriangle.m:845-866``)

A rule is a list of symmetry orbits ``(m, b1, b2, w)`` in barycentric form; the
orbit expansion and the mapping onto the reference triangle (0,0),(1,0),(0,1)
follow ``quadTrianglePointsAndWeights`` (``UaMain/quadtriangle.m:3674-3745``):
a point with barycentric weights (b1, b2, b3) on vertices a, b, c is
``a*b1 + b*b2 + c*b3``, and every weight is scaled by the triangle area (1/2).
"""
from __future__ import annotations

import math

import numpy as np

_S10 = math.sqrt(10.0)
_S15 = math.sqrt(15.0)
_D4A = (8.0 - _S10 + math.sqrt(38.0 - 44.0 * math.sqrt(2.0 / 5.0))) / 18.0
_D4B = (8.0 - _S10 - math.sqrt(38.0 - 44.0 * math.sqrt(2.0 / 5.0))) / 18.0
_D4WA = (620.0 + math.sqrt(213125.0 - 53320.0 * _S10)) / 3720.0
_D4WB = (620.0 - math.sqrt(213125.0 - 53320.0 * _S10)) / 3720.0

# Orbits as (m, b1, b2, w); b2 is None where the rule sets b2 = b1.
_RULES = {
    1: [(1, 1.0 / 3.0, None, 1.0)],
    2: [(3, 1.0 / 6.0, 2.0 / 3.0, 1.0 / 3.0)],
    4: [(3, _D4A, None, _D4WA), (3, _D4B, None, _D4WB)],
    5: [
        (1, 1.0 / 3.0, None, 9.0 / 40.0),
        (3, (6.0 - _S15) / 21.0, None, (155.0 - _S15) / 1200.0),
        (3, (6.0 + _S15) / 21.0, None, (155.0 + _S15) / 1200.0),
    ],
    8: [
        (1, 0.3333333333333333333333333333333333, None, 0.1443156076777871682510911104890646),
        (3, 0.1705693077517602066222935014914645, None, 0.1032173705347182502817915502921290),
        (3, 0.0505472283170309754584235505965989, None, 0.0324584976231980803109259283417806),
        (3, 0.4592925882927231560288155144941693, None, 0.0950916342672846247938961043885843),
        (6, 0.2631128296346381134217857862846436, 0.0083947774099576053372138345392944,
         0.0272303141744349942648446900739089),
    ],
    10: [
        (1, 0.3333333333333333333333333333333333, None, 0.0809374287976228802571131238165019),
        (3, 0.4272731788467755380904427175154472, None, 0.0772985880029631216825069823803434),
        (3, 0.1830992224486750205215743848502200, None, 0.0784576386123717313680939208343967),
        (3, 0.4904340197011305874539712223768484, None, 0.0174691679959294869176071632906781),
        (3, 0.0125724455515805327313290850210413, None, 0.0042923741848328280304804020901319),
        (6, 0.6542686679200661406665700955876279, 0.3080460016852477000000000000000000,
         0.0374688582104676429790207654850445),
        (6, 0.1228045770685592734301298174812812, 0.0333718337393047862408164417747804,
         0.0269493525918799596454494795810967),
    ],
}

SUPPORTED_DEGREES = tuple(sorted(_RULES))


def quadtriangle(degree: int):
    """Return ``(points, weights)`` for the reference triangle (0,0),(1,0),(0,1).

    ``points`` is ``nip x 2`` (xi, eta) and ``weights`` sums to 1/2, in the
    point order Úa produces (orbit by orbit, rotations in the order
    (b1,b2,b3), (b2,b3,b1), (b3,b1,b2), then the reflections for 6-orbits).
    """
    if degree not in _RULES:
        raise ValueError(f"quadrature degree {degree} not supported; have {SUPPORTED_DEGREES}")
    a = np.array([0.0, 0.0])
    b = np.array([1.0, 0.0])
    c = np.array([0.0, 1.0])
    area = 0.5
    pts, wts = [], []

    def put(p1, p2, p3, w):
        pts.append(a * p1 + b * p2 + c * p3)
        wts.append(area * w)

    for m, b1, b2, w in _RULES[degree]:
        if b2 is None:
            b2 = b1
        if m == 1:
            b2 = b1
            b3 = 1.0 - (b1 + b2)
            put(b1, b2, b3, w)
        elif m == 3:
            b3 = 1.0 - (b1 + b2)
            put(b1, b2, b3, w)
            put(b2, b3, b1, w)
            put(b3, b1, b2, w)
        elif m == 6:
            b3 = 1.0 - (b1 + b2)
            put(b1, b2, b3, w)
            put(b2, b3, b1, w)
            put(b3, b1, b2, w)
            put(b1, b3, b2, w)
            put(b2, b1, b3, w)
            put(b3, b2, b1, w)
        else:  # pragma: no cover - no such orbit in the tables above
            raise AssertionError(m)
    return np.array(pts, dtype=np.float64), np.array(wts, dtype=np.float64)


def default_degree(tri_nodes: int) -> int:
    """Default rule degree per element type (``UaMain/QuadratureRuleDegree.m:11-35``)."""
    return {3: 4, 6: 8, 10: 10}[tri_nodes]
