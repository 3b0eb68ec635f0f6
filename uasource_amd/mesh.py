"""Mesh inputs for the uvh assembly: a seeded jittered-Delaunay generator and the
parts of Úa's mesh-structure (``MUA``) builder that fix what the assembly reads.

The ``MUA`` built here carries the same fields, node order, element order and
orientation as Úa's ``CreateMUA`` would give for the same input triangulation:

* ``ChangeElementType`` (``UaMain/ChangeElementType.m``) lifts 3-node triangles to
  6 or 10 nodes with ``tri3to6``/``tri3to10`` (``UaMain/tri3to6.m``,
  ``UaMain/tri3to10.m``): corners at local 1,3,5 (P2) or 1,4,7 (P3), edge nodes
  at 1/2 or 1/3, 2/3 along each edge, the P3 centroid as local node 10.
* ``RemoveDuplicateNodes`` (``UaMain/RemoveDuplicateNodes.m``) merges coincident
  nodes, renumbers nodes in lexicographic order of the rounded (x, y), and
  orders elements lexicographically by their sorted node tuples.
* ``TestAndCorrectForInsideOutElements`` / ``FlipElements``
  (``UaMain/TestAndCorrectForInsideOutElements.m``, ``UaMain/FlipElements.m``):
  elements with a negative Jacobian determinant are flipped, so corners end up
  clockwise in (x, y) (Úa's N = [xi, 1-xi-eta, eta]).
* quadrature from ``quadtriangle`` (``UaMain/CreateMUA.m:36-47``) and element
  areas from ``TriAreaFE`` (``UaMain/TriAreaFE.m:16-20``, ``UaMain/TriArea.m:9``).

Connectivity is 0-based here (the MEX gateway converts MATLAB's 1-based arrays).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .quadrature import quadtriangle, default_degree

_EPS = np.finfo(np.float64).eps


def _matlab_round(x: np.ndarray) -> np.ndarray:
    """MATLAB ``round``: halves away from zero (numpy rounds halves to even)."""
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def remove_duplicate_nodes(coordinates: np.ndarray, connectivity: np.ndarray, tolerance: float):
    """Merge coincident nodes and sort nodes and elements (``UaMain/RemoveDuplicateNodes.m``)."""
    if connectivity.shape[0] == 1:
        return coordinates, connectivity
    ix = _matlab_round(coordinates[:, 0] / tolerance)
    iy = _matlab_round(coordinates[:, 1] / tolerance)
    order = np.lexsort((iy, ix))  # sort rows by ix, then iy (MATLAB unique 'rows')
    six, siy = ix[order], iy[order]
    new_group = np.ones(order.size, dtype=bool)
    new_group[1:] = (six[1:] != six[:-1]) | (siy[1:] != siy[:-1])
    group = np.cumsum(new_group) - 1
    first = order[new_group]                      # first occurrence (stable sort)
    m = np.empty(order.size, dtype=np.int64)
    m[order] = group
    coordinates = coordinates[first]
    connectivity = m[connectivity]
    srt = np.sort(connectivity, axis=1)
    eorder = np.lexsort(tuple(srt[:, k] for k in range(srt.shape[1] - 1, -1, -1)))
    s2 = srt[eorder]
    keep = np.ones(eorder.size, dtype=bool)
    keep[1:] = np.any(s2[1:] != s2[:-1], axis=1)
    connectivity = connectivity[eorder[keep]]
    return coordinates, connectivity


def tri3to6(coordinates: np.ndarray, connectivity: np.ndarray):
    """3-node to 6-node triangles (``UaMain/tri3to6.m``)."""
    tol = 1000 * _EPS
    coordinates, connectivity = remove_duplicate_nodes(coordinates, connectivity, tol)
    nn = coordinates.shape[0]
    ne = connectivity.shape[0]
    c6 = np.zeros((ne, 6), dtype=np.int64)
    c6[:, 0], c6[:, 2], c6[:, 4] = connectivity[:, 0], connectivity[:, 1], connectivity[:, 2]
    xy = np.zeros((nn + 3 * ne, 2))
    xy[:nn] = coordinates
    base = nn
    for inod in (1, 3, 5):
        c6[:, inod] = np.arange(base, base + ne)
        base += ne
    p1, p2, p3 = (coordinates[connectivity[:, k]] for k in range(3))
    xy[c6[:, 1]] = (p1 + p2) / 2
    xy[c6[:, 3]] = (p2 + p3) / 2
    xy[c6[:, 5]] = (p3 + p1) / 2
    return remove_duplicate_nodes(xy, c6, tol)


def tri3to10(coordinates: np.ndarray, connectivity: np.ndarray):
    """3-node to 10-node triangles (``UaMain/tri3to10.m``)."""
    tol = 1000 * _EPS
    nn = coordinates.shape[0]
    ne = connectivity.shape[0]
    c10 = np.zeros((ne, 10), dtype=np.int64)
    c10[:, 0], c10[:, 3], c10[:, 6] = connectivity[:, 0], connectivity[:, 1], connectivity[:, 2]
    xy = np.zeros((nn + 7 * ne, 2))
    xy[:nn] = coordinates
    base = nn
    for inod in (1, 2, 4, 5, 7, 8, 9):
        c10[:, inod] = np.arange(base, base + ne)
        base += ne
    p1, p2, p3 = (coordinates[connectivity[:, k]] for k in range(3))
    xy[c10[:, 1]] = (2 * p1 + p2) / 3
    xy[c10[:, 2]] = (p1 + 2 * p2) / 3
    xy[c10[:, 4]] = (2 * p2 + p3) / 3
    xy[c10[:, 5]] = (p2 + 2 * p3) / 3
    xy[c10[:, 7]] = (2 * p3 + p1) / 3
    xy[c10[:, 8]] = (p3 + 2 * p1) / 3
    xy[c10[:, 9]] = (p1 + p2 + p3) / 3
    return remove_duplicate_nodes(xy, c10, tol)


_FLIP = {3: [0, 2, 1], 6: [0, 5, 4, 3, 2, 1], 10: [0, 8, 7, 6, 5, 4, 3, 2, 1, 9]}


def detj_centroid(coordinates: np.ndarray, connectivity: np.ndarray) -> np.ndarray:
    """Jacobian determinant at the centroid from the corner nodes, Úa's N=[xi,1-xi-eta,eta]."""
    nod = connectivity.shape[1]
    corners = {3: (0, 1, 2), 6: (0, 2, 4), 10: (0, 3, 6)}[nod]
    p1, p2, p3 = (coordinates[connectivity[:, k]] for k in corners)
    return (p1[:, 0] - p2[:, 0]) * (p3[:, 1] - p2[:, 1]) - (p1[:, 1] - p2[:, 1]) * (p3[:, 0] - p2[:, 0])


def correct_inside_out(coordinates: np.ndarray, connectivity: np.ndarray) -> np.ndarray:
    """Flip elements with negative detJ (``UaMain/TestAndCorrectForInsideOutElements.m``)."""
    det = detj_centroid(coordinates, connectivity)
    bad = det < 0
    if np.any(bad):
        connectivity = connectivity.copy()
        connectivity[bad] = connectivity[bad][:, _FLIP[connectivity.shape[1]]]
        det = detj_centroid(coordinates, connectivity)
    if np.any(det <= 0):
        raise ValueError("Ua:TestAndCorrectForInsideOutElements:InsideOutelements")
    return connectivity


def tri_area_fe(coordinates: np.ndarray, connectivity: np.ndarray) -> np.ndarray:
    """Element areas from the corner nodes (``UaMain/TriAreaFE.m``, ``UaMain/TriArea.m``)."""
    nod = connectivity.shape[1]
    corners = {3: (0, 1, 2), 6: (0, 2, 4), 10: (0, 3, 6)}[nod]
    x = coordinates[connectivity[:, list(corners)], 0]
    y = coordinates[connectivity[:, list(corners)], 1]
    return 0.5 * np.abs((x[:, 0] - x[:, 2]) * (y[:, 1] - y[:, 0]) - (x[:, 0] - x[:, 1]) * (y[:, 2] - y[:, 0]))


@dataclass
class MUA:
    """The mesh structure the assembly reads (subset of Úa's ``MUA``, ``UaMain/CreateMUA.m:31-107``)."""

    coordinates: np.ndarray          # Nnodes x 2, float64
    connectivity: np.ndarray         # Nele x nod, int64, 0-based
    points: np.ndarray               # nip x 2 quadrature points on (0,0),(1,0),(0,1)
    weights: np.ndarray              # nip quadrature weights (sum 1/2)
    EleAreas: np.ndarray             # Nele
    QuadratureRuleDegree: int
    extra: dict = field(default_factory=dict)

    @property
    def Nnodes(self) -> int:
        return int(self.coordinates.shape[0])

    @property
    def Nele(self) -> int:
        return int(self.connectivity.shape[0])

    @property
    def nod(self) -> int:
        return int(self.connectivity.shape[1])

    @property
    def nip(self) -> int:
        return int(self.points.shape[0])


def CreateMUA(coordinates: np.ndarray, connectivity: np.ndarray, TriNodes: int = 3,
              QuadratureRuleDegree: int | None = None) -> MUA:
    """Build the ``MUA`` the assembly reads from a 3-node triangulation (``UaMain/CreateMUA.m``)."""
    connectivity = np.asarray(connectivity, dtype=np.int64)
    coordinates = np.asarray(coordinates, dtype=np.float64)
    used, inv = np.unique(connectivity.ravel(), return_inverse=True)
    connectivity = inv.reshape(connectivity.shape)
    coordinates = coordinates[used]
    if connectivity.shape[1] != 3:
        raise ValueError("CreateMUA here expects a 3-node triangulation on input")
    if TriNodes == 6:
        coordinates, connectivity = tri3to6(coordinates, connectivity)
    elif TriNodes == 10:
        coordinates, connectivity = tri3to10(coordinates, connectivity)
    elif TriNodes != 3:
        raise ValueError("ChangeElementType: TriNodes must be 3, 6 or 10")
    degree = default_degree(TriNodes) if QuadratureRuleDegree is None else int(QuadratureRuleDegree)
    points, weights = quadtriangle(degree)
    connectivity = correct_inside_out(coordinates, connectivity)
    areas = tri_area_fe(coordinates, connectivity)
    return MUA(coordinates=coordinates, connectivity=connectivity, points=points, weights=weights,
               EleAreas=areas, QuadratureRuleDegree=degree)


def jittered_delaunay(Lx: float, Ly: float, nx: int, ny: int, jitter: float, seed: int,
                      x0: float = 0.0, y0: float = 0.0):
    """Seeded jittered-grid Delaunay triangulation of the rectangle [x0,x0+Lx] x [y0,y0+Ly].

    ``nx x ny`` grid nodes; interior nodes are moved by U(-jitter, jitter) times the
    grid spacing in each direction, boundary nodes stay on the grid. Returns
    ``(coordinates, connectivity)`` with counter-clockwise 3-node triangles as a
    Delaunay mesher gives them.
    """
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    dx, dy = Lx / (nx - 1), Ly / (ny - 1)
    X, Y = np.meshgrid(np.arange(nx) * dx + x0, np.arange(ny) * dy + y0, indexing="ij")
    X, Y = X.ravel(), Y.ravel()
    I, J = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    I, J = I.ravel(), J.ravel()
    interior = (I > 0) & (I < nx - 1) & (J > 0) & (J < ny - 1)
    jx = rng.uniform(-jitter, jitter, size=X.size) * dx
    jy = rng.uniform(-jitter, jitter, size=X.size) * dy
    X = np.where(interior, X + jx, X)
    Y = np.where(interior, Y + jy, Y)
    xy = np.column_stack([X, Y])
    tri = Delaunay(xy).simplices.astype(np.int64)
    # make orientation counter-clockwise explicitly (qhull does not promise it)
    p1, p2, p3 = xy[tri[:, 0]], xy[tri[:, 1]], xy[tri[:, 2]]
    cross = (p2[:, 0] - p1[:, 0]) * (p3[:, 1] - p1[:, 1]) - (p2[:, 1] - p1[:, 1]) * (p3[:, 0] - p1[:, 0])
    cw = cross < 0
    tri[cw] = tri[cw][:, [0, 2, 1]]
    area = 0.5 * np.abs(cross)
    if np.any(area <= 1e-9 * dx * dy):
        raise ValueError("degenerate triangle from Delaunay")
    expected = 2 * (nx - 1) * (ny - 1)
    if tri.shape[0] != expected:
        raise ValueError(f"Delaunay gave {tri.shape[0]} triangles, expected {expected}")
    return xy, tri
