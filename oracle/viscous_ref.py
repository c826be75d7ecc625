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
Only extrapolation boundaries are restated (the tests' meshes use them).
"""
import numpy as np


def sutherland(T, mu_ref, T_mu, S):
    if S < 0:
        return np.full_like(T, mu_ref)
    t = T / T_mu
    return mu_ref * (t * np.sqrt(t)) * ((T_mu + S) / (T + S))


def viscous_residual(mesh, U, gamma, R, mu_ref, T_mu, S, Pr):
    """Viscous part of the residual dU/dt, [N][4], reference numbering.
    mesh: the Layout.mesh_arrays() dictionary of a triangle mesh."""
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
        pf = np.where(bnd[:, None], phi, phi[np.maximum(other, 0)])
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
    # face fluxes
    n = mesh["face_normals"] / np.linalg.norm(mesh["face_normals"], axis=1)[:, None]
    area = mesh["face_area"]
    real = np.zeros(len(cof), bool)
    real[foc.ravel()] = True
    out = np.zeros((N, 4))
    L = cof[:, 0]
    Rc = cof[:, 1]
    f_all = np.flatnonzero(real)
    for f in f_all:
        l = L[f]
        if Rc[f] >= 0:
            r = Rc[f]
            d = cc[r] - cc[l]
            il = 1.0 / np.sqrt(d @ d)
            t = d * il
            GX = 0.5 * (gx[l] + gx[r])
            GY = 0.5 * (gy[l] + gy[r])
            corr = (phi[r] - phi[l]) * il - (GX * t[0] + GY * t[1])
            GX = GX + corr * t[0]
            GY = GY + corr * t[1]
            ml, mr = mid[f] - cc[l], mid[f] - cc[r]
            uf = 0.5 * ((phi[l, 0] + gx[l, 0] * ml[0] + gy[l, 0] * ml[1]) +
                        (phi[r, 0] + gx[r, 0] * mr[0] + gy[r, 0] * mr[1]))
            vf = 0.5 * ((phi[l, 1] + gx[l, 1] * ml[0] + gy[l, 1] * ml[1]) +
                        (phi[r, 1] + gx[r, 1] * mr[0] + gy[r, 1] * mr[1]))
            m = 0.5 * (mu[l] + mu[r])
        else:  # extrapolation
            GX, GY = gx[l].copy(), gy[l].copy()
            uf, vf, m = phi[l, 0], phi[l, 1], mu[l]
        div = GX[0] + GY[1]
        txx = m * (2.0 * GX[0] - (2.0 / 3.0) * div)
        tyy = m * (2.0 * GY[1] - (2.0 / 3.0) * div)
        txy = m * (GY[0] + GX[1])
        n0, n1 = n[f]
        G1 = txx * n0 + txy * n1
        G2 = txy * n0 + tyy * n1
        G3 = uf * G1 + vf * G2 + m * cp / Pr * (GX[2] * n0 + GY[2] * n1)
        G = np.array([0.0, G1, G2, G3]) * area[f]
        out[l] += G / vol[l]
        if Rc[f] >= 0:
            out[Rc[f]] -= G / vol[Rc[f]]
    return out
