"""CPU restatement of the viscous (Navier-Stokes) extension of the residual.

TEST INFRASTRUCTURE ONLY: loaded by tests/ to check the GPU's viscous terms.
The reference solves the Euler equations only, so this has no reference
counterpart ("parity unpinned" against the reference); it restates this
build's own design in numpy, and the tests pin it with closed forms (zero
viscous force for a linear shear, viscous heating mu a^2, zero heat-flux
divergence for a linear temperature, zero for uniform flow).

Design restated:
- cell values (u, v, T) from the conserved state (T = e / cv, cv = R / (gamma - 1));
- weighted least-squares cell gradients over the face neighbours, weights
  1/|d|^2, d = neighbour centroid - centroid; a boundary face contributes its
  boundary value at the face midpoint (extrapolation: the cell's own value);
- face gradient = mean of the two cells' gradients with the component along
  t = d/|d| replaced by (phi_R - phi_L)/|d|; face velocity = mean of both
  cells' linear extrapolations to the face midpoint; face viscosity = mean of
  the cells' Sutherland viscosities;
- G = (0, tau n, (u tau + k grad T) n), tau = mu (grad u + grad u^T - 2/3 div u I),
  k = mu cp / Pr; the residual gains +G A / V in the face's side-0 cell and
  -G A / V in its side-1 cell.
Boundary faces (bc: zone name -> (type, params); extrapolation where no zone
is given) hold boundary values of (u, v, T) for the gradients: the no-slip
wall (0, 0, T), symmetry and the slip wall the tangential velocity and T, the
fixed state its (u, T); extrapolation and p_out the cell's own values. The
face flux there takes the cell's gradient, corrected toward the boundary
value along the centroid-to-midpoint direction where one is prescribed,
the boundary velocity and the cell's viscosity; symmetry and the slip wall
carry no viscous flux, the no-slip wall no heat flux.
"""
import numpy as np


def sutherland(T, mu_ref, T_mu, S):
    if S < 0:
        return np.full_like(T, mu_ref)
    t = T / T_mu
    return mu_ref * (t * np.sqrt(t)) * ((T_mu + S) / (T + S))


BC_CODE = {"symmetry": 0, "extrapolation": 1, "wall_adiabatic": 2, "upt": 3, "p_out": 4, "wall_noslip": 5}


def _boundary_types(mesh, bc, nf):
    """Per face: boundary type code (-1 interior or unlisted) and (u0, u1, T)."""
    bt = np.full(nf, 1, np.int64)
    bv = np.zeros((nf, 3))
    if bc and "zone_names" in mesh:
        off, zf = mesh["zone_offsets"], mesh["zone_faces"]
        for k, name in enumerate(mesh["zone_names"]):
            if name not in bc:
                continue
            kind, par = bc[name]
            f = zf[off[k]:off[k + 1]]
            bt[f] = BC_CODE[kind]
            u = par.get("u", (0.0, 0.0))
            bv[f] = (u[0], u[1], par.get("T", 0.0))
    return bt, bv


def _boundary_value(bt, bv, n, ph):
    """(u, v, T) a boundary face holds, from the cell's values ph."""
    out = ph.copy()
    if bt == 5:
        out[0] = 0.0
        out[1] = 0.0
    elif bt in (0, 2):
        un = ph[0] * n[0] + ph[1] * n[1]
        out[0] = ph[0] - un * n[0]
        out[1] = ph[1] - un * n[1]
    elif bt == 3:
        out[:] = bv
    return out


def viscous_residual(mesh, U, gamma, R, mu_ref, T_mu, S, Pr, bc=None):
    """Viscous part of the residual dU/dt, [N][4], reference numbering.
    mesh: the Layout.mesh_arrays() dictionary of a triangle mesh, or a
    meshgen mesh dictionary (with zones); bc: zone name -> (type, params)."""
    U = np.asarray(U, dtype=np.float64)
    cc = mesh["cell_coords"]
    vol = mesh["cell_volume"]
    nof = mesh["nodes_of_face"].reshape(-1, 2)
    cof = mesh["cells_of_face"]
    xy = mesh["node_coords"]
    foc = mesh["faces_of_cell"].reshape(-1, 3)
    N = len(U)
    cv = R / (gamma - 1.0)
    cp = R * gamma / (gamma - 1.0)
    rho = U[:, 0]
    u, v = U[:, 1] / rho, U[:, 2] / rho
    e = U[:, 3] / rho - 0.5 * (u * u + v * v)
    T = e / cv
    mu = sutherland(T, mu_ref, T_mu, S)
    phi = np.stack([u, v, T], axis=1)
    mid = 0.5 * (xy[nof[:, 0]] + xy[nof[:, 1]])
    n = mesh["face_normals"] / np.linalg.norm(mesh["face_normals"], axis=1)[:, None]
    bt, bv = _boundary_types(mesh, bc, len(cof))
    # least-squares gradients
    A = np.zeros((N, 2, 2))
    b = np.zeros((N, 3, 2))
    for l in range(3):
        f = foc[:, l]
        c0, c1 = cof[f, 0], cof[f, 1]
        c = np.arange(N)
        bnd = c1 < 0
        other = np.where(c0 == c, c1, c0)
        d = np.where(bnd[:, None], mid[f] - cc, cc[np.maximum(other, 0)] - cc)
        pb = phi.copy()
        tb, nb_ = bt[f], n[f]
        un = phi[:, 0] * nb_[:, 0] + phi[:, 1] * nb_[:, 1]
        slip = (tb == 0) | (tb == 2)
        pb[slip, 0] = phi[slip, 0] - un[slip] * nb_[slip, 0]
        pb[slip, 1] = phi[slip, 1] - un[slip] * nb_[slip, 1]
        pb[tb == 5, 0:2] = 0.0
        pb[tb == 3] = bv[f][tb == 3]
        pf = np.where(bnd[:, None], pb, phi[np.maximum(other, 0)])
        w = 1.0 / (d[:, 0] ** 2 + d[:, 1] ** 2)
        A[:, 0, 0] += w * d[:, 0] * d[:, 0]
        A[:, 0, 1] += w * d[:, 0] * d[:, 1]
        A[:, 1, 1] += w * d[:, 1] * d[:, 1]
        dp = w[:, None] * (pf - phi)
        b[:, :, 0] += d[:, 0:1] * dp
        b[:, :, 1] += d[:, 1:2] * dp
    det = A[:, 0, 0] * A[:, 1, 1] - A[:, 0, 1] ** 2
    gx = (A[:, 1, 1, None] * b[:, :, 0] - A[:, 0, 1, None] * b[:, :, 1]) / det[:, None]
    gy = (A[:, 0, 0, None] * b[:, :, 1] - A[:, 0, 1, None] * b[:, :, 0]) / det[:, None]
    # face fluxes (vectorised over faces; each cell adds its faces' terms in
    # ascending face order)
    area = mesh["face_area"]
    real = np.zeros(len(cof), bool)
    real[foc.ravel()] = True
    f_all = np.flatnonzero(real)
    L = cof[f_all, 0]
    Rc = cof[f_all, 1]
    inn = Rc >= 0
    nf = len(f_all)
    GX = gx[L].copy()
    GY = gy[L].copy()
    uf = phi[L, 0].copy()
    vf = phi[L, 1].copy()
    m = mu[L].copy()
    # interior faces
    fi, li, ri = f_all[inn], L[inn], Rc[inn]
    d = cc[ri] - cc[li]
    il = 1.0 / np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1])
    t = d * il[:, None]
    gxi = 0.5 * (gx[li] + gx[ri])
    gyi = 0.5 * (gy[li] + gy[ri])
    corr = (phi[ri] - phi[li]) * il[:, None] - (gxi * t[:, 0:1] + gyi * t[:, 1:2])
    GX[inn] = gxi + corr * t[:, 0:1]
    GY[inn] = gyi + corr * t[:, 1:2]
    ml, mr = mid[fi] - cc[li], mid[fi] - cc[ri]
    uf[inn] = 0.5 * ((phi[li, 0] + gx[li, 0] * ml[:, 0] + gy[li, 0] * ml[:, 1]) +
                     (phi[ri, 0] + gx[ri, 0] * mr[:, 0] + gy[ri, 0] * mr[:, 1]))
    vf[inn] = 0.5 * ((phi[li, 1] + gx[li, 1] * ml[:, 0] + gy[li, 1] * ml[:, 1]) +
                     (phi[ri, 1] + gx[ri, 1] * mr[:, 0] + gy[ri, 1] * mr[:, 1]))
    m[inn] = 0.5 * (mu[li] + mu[ri])
    # boundary faces
    carry = np.ones(nf, bool)
    noheat = np.zeros(nf, bool)
    for j in np.flatnonzero(~inn):
        f, l = f_all[j], L[j]
        if bt[f] in (0, 2):  # symmetry, slip wall: no viscous flux
            carry[j] = False
            continue
        pv = _boundary_value(bt[f], bv[f], n[f], phi[l])
        if bt[f] not in (1, 4):  # a prescribed value: correct toward it
            db = mid[f] - cc[l]
            ib = 1.0 / np.sqrt(db @ db)
            tb = db * ib
            cb = (pv - phi[l]) * ib - (GX[j] * tb[0] + GY[j] * tb[1])
            GX[j] = GX[j] + cb * tb[0]
            GY[j] = GY[j] + cb * tb[1]
        uf[j], vf[j] = pv[0], pv[1]
        noheat[j] = bt[f] == 5  # the adiabatic no-slip wall: no heat flux, u = 0
    div = GX[:, 0] + GY[:, 1]
    txx = m * (2.0 * GX[:, 0] - (2.0 / 3.0) * div)
    tyy = m * (2.0 * GY[:, 1] - (2.0 / 3.0) * div)
    txy = m * (GY[:, 0] + GX[:, 1])
    n0, n1 = n[f_all, 0], n[f_all, 1]
    G1 = txx * n0 + txy * n1
    G2 = txy * n0 + tyy * n1
    G3 = np.where(noheat, 0.0, uf * G1 + vf * G2 + m * cp / Pr * (GX[:, 2] * n0 + GY[:, 2] * n1))
    G = np.stack([np.zeros(nf), G1, G2, G3], axis=1) * area[f_all, None]
    G[~carry] = 0.0
    out = np.zeros((N, 4))
    # side 0 gains +G A / V, side 1 loses it, face by face
    cells = np.stack([L, np.where(inn, Rc, L)], axis=1).reshape(-1)
    vals = np.stack([G / vol[L, None], np.where(inn[:, None], -G / vol[np.maximum(Rc, 0), None], 0.0)],
                    axis=1).reshape(-1, 4)
    np.add.at(out, cells, vals)
    return out
