"""Host precompute of the product (mesh generators, SFC renumbering, packed
TENO layout) against the oracle, bit for bit, on the CPU."""
import numpy as np
import pytest


def _layout(mlr, kind, nx, ny, jit, seed, **kw):
    num = mlr.numerics(**kw)
    return mlr.Layout(mlr.mesh_gen(kind, nx, ny, 1.0, 1.0, jit, seed), num)


@pytest.mark.parametrize("kind,nx,ny,jit", [("cartesian", 9, 6, 0.0), ("cartesian", 7, 7, 0.2),
                                           ("cartesian_tri", 11, 7, 0.0), ("cartesian_tri", 10, 10, 0.2),
                                           ("wedge", 12, 5, 0.0)])
def test_mesh_bitexact(mlr, oracle, kind, nx, ny, jit):
    L = _layout(mlr, kind, nx, ny, jit, 42, recon="FO")
    m = oracle.Mesh(kind, nx, ny, 1.0, 1.0, jit, 42)
    a = L.mesh_arrays()
    for k in ("node_coords", "cell_coords", "cell_volume", "face_area", "face_normals", "nodes_of_cell",
              "faces_of_cell", "nodes_of_face", "cells_of_face"):
        assert np.array_equal(a[k], getattr(m, k), equal_nan=True), k


@pytest.mark.parametrize("order,nx,ny,jit", [(1, 9, 8, 0.2), (2, 12, 12, 0.0), (2, 13, 9, 0.2), (3, 14, 12, 0.2)])
def test_teno_layout_bitexact(mlr, oracle, order, nx, ny, jit):
    L = _layout(mlr, "cartesian_tri", nx, ny, jit, 7, recon="TENO", order=order)
    m = oracle.Mesh("cartesian_tri", nx, ny, 1.0, 1.0, jit, 7)
    s = oracle.Solver(m, oracle.make_config(order=order))
    T = s.teno()
    K, M = T["n_dof"], T["max_cells"]
    assert (L.K, L.M, L.G) == (K, M, 4)
    ipt, oi = L.teno_tables()
    assert np.array_equal(ipt, T["integral_psi_target"])
    assert np.array_equal(oi.ravel(), T["oscillation_indicator"])
    og, os_, st, orm = T["offsets_stencil_groups"], T["offsets_stencils"], T["stencils"], T["offsets_recon"]
    for c in range(m.n_cells):
        ids = L.cell_stencils(c)
        A, ta, _ = L.cell_matrices(c)
        for sl in range(4):
            si = og[c] + sl
            a, b = os_[si], os_[si + 1]
            got = ids[sl][ids[sl] != 0xFFFFFFFF]
            assert np.array_equal(got, st[a:b]), (c, sl)
            if b > a:
                ref = T["recon"][orm[si]:orm[si] + K * M].reshape(K, M)
                assert np.array_equal(A[sl], ref), (c, sl)
                assert np.array_equal(ta[sl], T["transformed_areas"][a:b]), (c, sl)


def test_sfc_permutation_and_slots(mlr, oracle):
    nx = ny = 16
    L = _layout(mlr, "cartesian_tri", nx, ny, 0.2, 3, recon="TENO", order=2)
    m = oracle.Mesh("cartesian_tri", nx, ny, 1.0, 1.0, 0.2, 3)
    perm = L.permutation()
    assert np.array_equal(np.sort(perm), np.arange(m.n_cells))
    # the curve keeps neighbours close: mean index distance across faces is
    # far below the reference numbering's row-major stride
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    fi = m.zones["interior"]
    d_new = np.abs(inv[m.cells_of_face[fi, 0]].astype(int) - inv[m.cells_of_face[fi, 1]].astype(int))
    d_old = np.abs(m.cells_of_face[fi, 0] - m.cells_of_face[fi, 1])
    assert np.median(d_new) <= np.median(d_old)
    # slot geometry: unit normals of the reference faces, neighbours, sides
    for c in range(0, m.n_cells, 7):
        sl = L.slots(c)
        for l in range(3):
            f = m.faces_of_cell[3 * c + l]
            n = m.face_normals[f]
            inv_n = 1 / np.sqrt(n[0] * n[0] + n[1] * n[1])
            assert sl["nx"][l] == n[0] * inv_n and sl["ny"][l] == n[1] * inv_n
            assert sl["area"][l] == m.face_area[f]
            c0, c1 = m.cells_of_face[f]
            other = c if c1 < 0 else (c1 if c0 == c else c0)
            assert sl["nb"][l] == other
            assert ((sl["info"][l] >> 2) & 1) == (1 if c1 == c else 0)
            assert bool(sl["info"][l] & 8) == (c1 < 0)


def test_rejects_teno_on_quads(mlr):
    with pytest.raises(mlr.MallardError, match="triangular"):
        _layout(mlr, "cartesian", 4, 4, 0.0, 0, recon="TENO", order=2)


def test_rejects_unknown_boundary_zone(mlr):
    with pytest.raises(mlr.MallardError, match="not found"):
        _layout(mlr, "cartesian_tri", 4, 4, 0.0, 0, recon="FO", bcs=[("inlet", "symmetry", {})])


@pytest.mark.parametrize("nx,ny,order", [(1, 1, 1), (1, 1, 2), (2, 1, 2)])
def test_rejects_mesh_smaller_than_stencil(mlr, nx, ny, order):
    """Too few cells for a central stencil of the order's size: a clear error,
    where the oracle's ring search runs out of memory."""
    with pytest.raises(mlr.MallardError, match="mesh too small"):
        _layout(mlr, "cartesian_tri", nx, ny, 0.0, 0, recon="TENO", order=order, teno_variant="corrected")


def test_slot_edge_ids(mlr, oracle):
    """Slot 0's info word holds, for each face slot l in bits 24 + 2 l, the
    edge of the reference triangle the face maps to in the target frame:
    0 for the cell's nodes 0-1, 1 for 1-2, 2 for 2-0 (the compact tables'
    face basis rows), and in bit 30, 31 or 19 whether the face's node 0 is
    the edge's higher cell node (the derived tables' face direction)."""
    L = _layout(mlr, "cartesian_tri", 9, 7, 0.2, 5, recon="TENO", order=2)
    m = oracle.Mesh("cartesian_tri", 9, 7, 1.0, 1.0, 0.2, 5)
    edge = {frozenset((0, 1)): 0, frozenset((1, 2)): 1, frozenset((0, 2)): 2}
    for c in range(m.n_cells):
        info0 = int(L.slots(c)["info"][0])
        nodes = list(m.nodes_of_cell[3 * c:3 * c + 3])
        for l in range(3):
            f = m.faces_of_cell[3 * c + l]
            a, b = m.nodes_of_face[2 * f], m.nodes_of_face[2 * f + 1]
            want = edge[frozenset((nodes.index(a), nodes.index(b)))]
            assert (info0 >> (24 + 2 * l)) & 3 == want, (c, l)
            flip = 1 if nodes.index(a) > nodes.index(b) else 0
            assert (info0 >> (30, 31, 19)[l]) & 1 == flip, (c, l)
