"""Unstructured triangle meshes from seeded Delaunay point sampling.

The reference reads no mesh files (its `file` type throws) and generates only
structured meshes; this module builds the graded wedge mesh of the viscous
wedge problem in the reference's array layout (mesh.h:228-253: CCW nodes per
cell, two nodes per face with the face's side-0 cell on its left, -1 for the
missing side of a boundary face, zone 0 "interior" then the boundary zones),
ready for `Solver(mesh=...)`. The triangulation is scipy's (Qhull).
"""
import numpy as np


def wedge_mesh(n_cells=4_000_000, angle_deg=15.0, x_tip=0.5, x_end=1.5, height=1.0, beta=3.0, seed=0):
    """Triangles filling the region above a wedge: the bottom boundary is flat
    ("symmetry", 0 <= x < x_tip) and then rises at angle_deg ("wall", the
    wedge surface, x_tip <= x <= x_end); "left" is the inflow, "right" and
    "top" the outflow. Points are sampled in rows whose spacing shrinks toward
    the bottom boundary by the factor e^beta (row k at the fraction
    expm1(beta k / n) / expm1(beta) of the local height), with as many points
    per row as keep the spacing isotropic; every interior point is jittered by
    up to 0.3 of its spacing in each direction from a seeded stream, so no two
    points come closer than 0.4 spacings. Boundary points are the rows' ends
    and the end rows themselves. Returns the mesh dictionary Solver(mesh=...)
    takes."""
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    tan = np.tan(np.radians(angle_deg))

    def bottom(x):
        return np.where(x > x_tip, (x - x_tip) * tan, 0.0)

    def rows(n):
        eta = np.expm1(beta * np.arange(n + 1) / n) / np.expm1(beta)
        cols = np.maximum(np.round(x_end / (np.diff(eta) * height)), 4).astype(np.int64)
        return eta, cols

    # rows so that the point count (about half the cell count) matches
    n = 8
    while True:
        eta, cols = rows(n)
        if cols.sum() + cols[-1] >= n_cells // 2:
            break
        n += max(1, n // 16)
    pts = []
    for k in range(n + 1):
        m = cols[min(k, n - 1)]
        if k == 0 or k == n:  # the bottom and top boundaries, evenly spaced
            xs = np.linspace(0.0, x_end, m + 1)
            if k == 0:
                # the wedge tip is a node, so the symmetry and wall zones meet there
                xs = np.unique(np.concatenate([xs, [x_tip]]))
            ys = bottom(xs) if k == 0 else np.full(len(xs), height)
            pts.append(np.stack([xs, ys], 1))
            continue
        d_eta = min(eta[k + 1] - eta[k], eta[k] - eta[k - 1])
        u = (np.arange(m) + 0.5 + 0.6 * (rng.random(m) - 0.5)) / m
        v = eta[k] + 0.6 * (rng.random(m) - 0.5) * d_eta
        xs = u * x_end
        yb = bottom(xs)
        pts.append(np.stack([xs, yb + v * (height - yb)], 1))
        # the row's ends on the left and right boundaries
        yr = bottom(np.array([x_end]))[0]
        pts.append(np.array([[0.0, eta[k] * height], [x_end, yr + eta[k] * (height - yr)]]))
    pts = np.concatenate(pts)
    tri = Delaunay(pts).simplices.astype(np.int64)
    # counter-clockwise nodes
    p0, p1, p2 = pts[tri[:, 0]], pts[tri[:, 1]], pts[tri[:, 2]]
    cross = (p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p1[:, 1] - p0[:, 1]) * (p2[:, 0] - p0[:, 0])
    flip = cross < 0
    tri[flip, 1], tri[flip, 2] = tri[flip, 2].copy(), tri[flip, 1].copy()
    a = 0.5 * np.abs(cross)
    if a.min() <= 1e-12 * a.mean():
        raise RuntimeError("degenerate triangle in the Delaunay mesh")
    return mesh_from_triangles(pts, tri, _wedge_zone(x_tip, x_end, height, tan))


def _wedge_zone(x_tip, x_end, height, tan):
    def zone(mid):
        x, y = mid[:, 0], mid[:, 1]
        eps = 1e-9
        z = np.full(len(mid), -1)
        z[np.abs(x) < eps] = 0                                         # left
        z[np.abs(x - x_end) < eps] = 1                                 # right
        z[np.abs(y - height) < eps] = 2                                # top
        z[(np.abs(y) < eps) & (x < x_tip + eps) & (z < 0)] = 3         # symmetry
        z[(np.abs(y - (x - x_tip) * tan) < 1e-7) & (x >= x_tip - eps) & (z < 0)] = 4  # wall
        return z, ["left", "right", "top", "symmetry", "wall"]
    return zone


def mesh_from_triangles(pts, tri, zone_of):
    """Reference mesh arrays from CCW triangles: faces numbered by first
    appearance (cell-major, edge k from node k to node k+1), each face's
    side-0 cell the first cell that lists it; boundary faces grouped into
    zones by zone_of(midpoints) -> (zone index per face, zone names)."""
    nc = len(tri)
    e0 = tri.reshape(-1)
    e1 = np.roll(tri, -1, axis=1).reshape(-1)
    key = np.minimum(e0, e1) * (len(pts) + 1) + np.maximum(e0, e1)
    _, first, inv = np.unique(key, return_index=True, return_inverse=True)
    # number faces in order of first appearance
    order = np.argsort(first, kind="stable")
    rank = np.empty(len(order), np.int64)
    rank[order] = np.arange(len(order))
    face_of_edge = rank[inv]
    nf = len(order)
    cells_of_face = np.full((nf, 2), -1, np.int32)
    cell_of_edge = np.repeat(np.arange(nc), 3)
    fe = face_of_edge
    is_first = np.zeros(len(fe), bool)
    is_first[first[order]] = True
    cells_of_face[fe[is_first], 0] = cell_of_edge[is_first]
    cells_of_face[fe[~is_first], 1] = cell_of_edge[~is_first]
    nodes_of_face = np.zeros((nf, 2), np.uint32)
    nodes_of_face[fe[is_first], 0] = e0[is_first]
    nodes_of_face[fe[is_first], 1] = e1[is_first]
    bnd = np.flatnonzero(cells_of_face[:, 1] < 0)
    mid = 0.5 * (pts[nodes_of_face[bnd, 0]] + pts[nodes_of_face[bnd, 1]])
    z, names = zone_of(mid)
    if (z < 0).any():
        raise RuntimeError("boundary face outside every zone")
    interior = np.flatnonzero(cells_of_face[:, 1] >= 0)
    zone_faces = [interior] + [bnd[z == k] for k in range(len(names))]
    offsets = np.concatenate([[0], np.cumsum([len(f) for f in zone_faces])]).astype(np.uint32)
    return dict(node_coords=np.ascontiguousarray(pts, np.float64),
                offsets_nodes_of_cell=(3 * np.arange(nc + 1)).astype(np.uint32),
                nodes_of_cell=np.ascontiguousarray(tri.reshape(-1), np.uint32),
                faces_of_cell=np.ascontiguousarray(fe, np.uint32),
                nodes_of_face=np.ascontiguousarray(nodes_of_face.reshape(-1), np.uint32),
                cells_of_face=np.ascontiguousarray(cells_of_face.reshape(-1), np.int32),
                zone_names=["interior"] + names,
                zone_offsets=offsets,
                zone_faces=np.concatenate(zone_faces).astype(np.uint32))
