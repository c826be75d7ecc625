"""Seeded Delaunay wedge mesh (host only): it tiles the wedge domain exactly,
its faces are consistent with the reference's conventions (CCW cells, the
side-0 cell left of each face, every boundary face in one zone), it is graded
toward the wall, and the same seed gives the same mesh."""
import numpy as np
import pytest

pytest.importorskip("scipy")

from mallard_amd import meshgen


@pytest.fixture(scope="module")
def mesh():
    return meshgen.wedge_mesh(n_cells=20000, seed=3)


def _tri(m):
    return m["nodes_of_cell"].reshape(-1, 3).astype(np.int64)


def test_tiles_the_domain(mesh):
    xy = mesh["node_coords"]
    t = _tri(mesh)
    p0, p1, p2 = xy[t[:, 0]], xy[t[:, 1]], xy[t[:, 2]]
    cross = (p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p1[:, 1] - p0[:, 1]) * (p2[:, 0] - p0[:, 0])
    assert (cross > 0).all()                                      # counter-clockwise
    tan = np.tan(np.radians(15.0))
    area = 1.5 * 1.0 - 0.5 * 1.0 ** 2 * tan
    assert abs(0.5 * cross.sum() - area) < 1e-12 * area
    assert 15000 < len(t) < 26000


def test_face_conventions_and_zones(mesh):
    xy = mesh["node_coords"]
    t = _tri(mesh)
    nof = mesh["nodes_of_face"].reshape(-1, 2)
    cof = mesh["cells_of_face"].reshape(-1, 2)
    foc = mesh["faces_of_cell"].reshape(-1, 3)
    nf = len(nof)
    # every face: side-0 cell lists it and lies on its left
    c0 = cof[:, 0]
    assert (c0 >= 0).all()
    assert (foc[c0] == np.arange(nf)[:, None]).any(axis=1).all()
    cen = xy[t].mean(axis=1)
    a, b = xy[nof[:, 0]], xy[nof[:, 1]]
    left = (b[:, 0] - a[:, 0]) * (cen[c0, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (cen[c0, 0] - a[:, 0])
    assert (left > 0).all()
    inner = cof[:, 1] >= 0
    assert (foc[cof[inner, 1]] == np.flatnonzero(inner)[:, None]).any(axis=1).all()
    # each face twice over the cells (interior) or once (boundary)
    counts = np.bincount(foc.ravel(), minlength=nf)
    assert np.array_equal(counts, np.where(inner, 2, 1))
    # zones: interior first, then every boundary face exactly once
    off, zf, names = mesh["zone_offsets"], mesh["zone_faces"], mesh["zone_names"]
    assert names[0] == "interior" and np.array_equal(np.sort(zf[off[0]:off[1]]), np.flatnonzero(inner))
    bnd = np.sort(zf[off[1]:])
    assert np.array_equal(bnd, np.flatnonzero(~inner))
    wall = zf[off[names.index("wall")]:off[names.index("wall") + 1]]
    mid = 0.5 * (a[wall] + b[wall])
    assert np.abs(mid[:, 1] - (mid[:, 0] - 0.5) * np.tan(np.radians(15.0))).max() < 1e-9
    assert abs(np.linalg.norm(b[wall] - a[wall], axis=1).sum() - 1.0 / np.cos(np.radians(15.0))) < 1e-9


def test_graded_toward_the_wall(mesh):
    xy = mesh["node_coords"]
    t = _tri(mesh)
    cen = xy[t].mean(axis=1)
    p0, p1, p2 = xy[t[:, 0]], xy[t[:, 1]], xy[t[:, 2]]
    area = 0.5 * ((p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p1[:, 1] - p0[:, 1]) * (p2[:, 0] - p0[:, 0]))
    yb = np.where(cen[:, 0] > 0.5, (cen[:, 0] - 0.5) * np.tan(np.radians(15.0)), 0.0)
    d = (cen[:, 1] - yb) / (1.0 - yb)
    assert area[d < 0.1].mean() * 5 < area[d > 0.8].mean()


def test_seeded(mesh):
    again = meshgen.wedge_mesh(n_cells=20000, seed=3)
    assert all(np.array_equal(np.asarray(mesh[k]), np.asarray(again[k])) for k in mesh)
