"""Python host bindings of the MI355X residual engine (ctypes over the C ABI in
include/mallard_amd.h).

The names mirror the reference's interfaces: ``Solver`` (solver.h:30-279) with
``calc_rhs``, ``calc_dt``, ``take_step``, ``run``; reconstruction, Riemann
solver, integrator and boundary keys are the reference's TOML strings.
Everything computes in lib/libmallard_amd.so (HIP kernels for gfx950); there
is no CPU fallback, and importing the module fails if the library is missing.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmallard_amd.so")

MESH = {"file": 0, "cartesian": 1, "cartesian_tri": 2, "wedge": 3}
RECON = {"FO": 0, "TENO": 1}
BASIS = {"monomial": 0, "legendre": 1}
TENO_VARIANT = {"reference": 0, "corrected": 1}
RIEMANN = {"Rusanov": 0, "HLL": 1, "HLLC": 2, "Roe": 3}
INTEGRATOR = {"FE": 0, "RK4": 1, "SSPRK3": 2}
BC = {"symmetry": 0, "extrapolation": 1, "wall_adiabatic": 2, "upt": 3, "p_out": 4, "wall_noslip": 5}


class MlrBC(C.Structure):
    _fields_ = [("zone", C.c_char * 32), ("type", C.c_int32), ("u", C.c_double * 2),
                ("p", C.c_double), ("T", C.c_double)]


class MlrNumerics(C.Structure):
    _fields_ = [("recon", C.c_int32), ("basis", C.c_int32), ("order", C.c_int32),
                ("stencil_factor", C.c_double), ("quad_order_cell", C.c_int32),
                ("quad_order_face", C.c_int32), ("teno_variant", C.c_int32),
                ("riemann", C.c_int32), ("integrator", C.c_int32),
                ("gamma", C.c_double), ("p_ref", C.c_double), ("T_ref", C.c_double),
                ("rho_ref", C.c_double), ("p_min", C.c_double), ("p_max", C.c_double),
                ("cfl", C.c_double), ("dt", C.c_double), ("n_bc", C.c_int32),
                ("bc", MlrBC * 14), ("n_threads", C.c_int32),
                ("viscous", C.c_int32), ("mu_ref", C.c_double), ("T_mu_ref", C.c_double),
                ("sutherland_S", C.c_double), ("prandtl", C.c_double), ("positivity", C.c_int32)]


class MlrMeshDesc(C.Structure):
    _fields_ = [("n_nodes", C.c_uint32), ("node_coords", C.c_void_p), ("n_cells", C.c_uint32),
                ("offsets_nodes_of_cell", C.c_void_p), ("nodes_of_cell", C.c_void_p),
                ("faces_of_cell", C.c_void_p), ("n_faces", C.c_uint32), ("nodes_of_face", C.c_void_p),
                ("cells_of_face", C.c_void_p), ("n_zones", C.c_uint32), ("zone_names", C.c_void_p),
                ("zone_offsets", C.c_void_p), ("zone_faces", C.c_void_p)]


class MlrMeshGen(C.Structure):
    _fields_ = [("type", C.c_int32), ("nx", C.c_uint32), ("ny", C.c_uint32), ("Lx", C.c_double),
                ("Ly", C.c_double), ("jitter", C.c_double), ("seed", C.c_uint64)]


# every entry point of include/mallard_amd.h with its ctypes signature
_P = C.c_void_p
# int (*mlr_rhs_fn)(const double* U, double* rhs, void* user), device pointers
RHS_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p)
_SIGS = {
    "mlr_ctx_create_toml": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "mlr_ctx_create_generated": (C.c_int, [C.POINTER(MlrMeshGen), C.POINTER(MlrNumerics), C.POINTER(_P)]),
    "mlr_ctx_create": (C.c_int, [_P, C.POINTER(MlrNumerics), C.POINTER(_P)]),
    "mlr_ctx_destroy": (C.c_int, [_P]),
    "mlr_last_error": (C.c_char_p, [_P]),
    "mlr_get_sizes": (C.c_int, [_P, _P]),
    "mlr_init_quadrant": (C.c_int, [_P, C.c_double, C.c_double, _P]),
    "mlr_init_constant": (C.c_int, [_P, _P, C.c_double, C.c_double]),
    "mlr_set_state": (C.c_int, [_P, _P, _P]),
    "mlr_get_state": (C.c_int, [_P, _P, _P]),
    "mlr_calc_rhs": (C.c_int, [_P, _P]),
    "mlr_compute_dt": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "mlr_step": (C.c_int, [_P]),
    "mlr_run_steps": (C.c_int, [_P, C.c_int64, _P]),
    "mlr_run_steps_profiled": (C.c_int, [_P, C.c_int64, _P]),
    "mlr_autotune": (C.c_int, [_P, _P]),
    "mlr_set_tune": (C.c_int, [_P, _P]),
    "mlr_get_tune": (C.c_int, [_P, _P]),
    "mlr_run": (C.c_int, [_P]),
    "mlr_set_cfl": (C.c_int, [_P, C.c_double]),
    "mlr_set_dt_history": (C.c_int, [_P, C.c_int64]),
    "mlr_get_time": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "mlr_get_face_values": (C.c_int, [_P, _P]),
    "mlr_get_decisions": (C.c_int, [_P, _P]),
    "mlr_get_cfl_local": (C.c_int, [_P, _P]),
    "mlr_get_permutation": (C.c_int, [_P, _P]),
    "mlr_get_stencils": (C.c_int, [_P, C.c_uint32, _P, _P]),
    "mlr_get_teno_matrices": (C.c_int, [_P, C.c_uint32, _P, _P, _P, _P]),
    "mlr_get_device_stencils": (C.c_int, [_P, C.c_uint32, _P]),
    "mlr_get_device_matrices": (C.c_int, [_P, C.c_uint32, _P, _P]),
    "mlr_layout_create": (C.c_int, [C.POINTER(MlrMeshGen), C.POINTER(MlrNumerics), C.POINTER(_P)]),
    "mlr_layout_destroy": (C.c_int, [_P]),
    "mlr_layout_sizes": (C.c_int, [_P, _P]),
    "mlr_layout_permutation": (C.c_int, [_P, _P]),
    "mlr_layout_mesh_arrays": (C.c_int, [_P] + [_P] * 9),
    "mlr_layout_cell_stencils": (C.c_int, [_P, C.c_uint32, _P]),
    "mlr_layout_cell_matrices": (C.c_int, [_P, C.c_uint32, _P, _P, _P]),
    "mlr_layout_teno_tables": (C.c_int, [_P, _P, _P]),
    "mlr_layout_slots": (C.c_int, [_P, C.c_uint32] + [_P] * 6),
    "mlr_probe_math": (C.c_int, [C.c_int, _P, _P, _P, C.c_int]),
    "mlr_probe_stream": (C.c_int, [C.c_int, C.c_int64, C.c_int, C.POINTER(C.c_double)]),
    "mlr_config_parse": (C.c_int, [C.c_char_p, C.POINTER(MlrMeshGen), C.POINTER(MlrNumerics), _P, _P]),
    "mlr_ensemble_create": (C.c_int, [_P, C.c_int, C.POINTER(MlrNumerics), C.POINTER(_P)]),
    "mlr_ensemble_destroy": (C.c_int, [_P]),
    "mlr_ensemble_sizes": (C.c_int, [_P, _P]),
    "mlr_ensemble_init_quadrant": (C.c_int, [_P, C.c_int, C.c_double, C.c_double, _P]),
    "mlr_ensemble_get_state": (C.c_int, [_P, C.c_int, _P, _P]),
    "mlr_ensemble_get_time": (C.c_int, [_P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "mlr_ensemble_run_steps": (C.c_int, [_P, C.c_int64]),
    "mlr_ensemble_run_steps_profiled": (C.c_int, [_P, C.c_int64, _P]),
    "mlr_ensemble_autotune": (C.c_int, [_P, _P]),
    "mlr_indicator_cadence": (C.c_int, [_P, C.c_int]),
    "mlr_indicator_split": (C.c_int, [_P, C.c_int]),
    "mlr_indicator_decisions": (C.c_int, [_P, _P]),
    "mlr_ensemble_indicator_cadence": (C.c_int, [_P, C.c_int]),
    "mlr_ensemble_indicator_split": (C.c_int, [_P, C.c_int]),
    "mlr_ensemble_get_volumes": (C.c_int, [_P, C.c_int, _P]),
    "mlr_ensemble_set_tune": (C.c_int, [_P, _P]),
    "mlr_ensemble_set_indicator": (C.c_int, [_P, _P, C.c_double]),
    "mlr_ensemble_gated": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int64)]),
    "mlr_ensemble_set_persistent": (C.c_int, [_P, C.c_int]),
    "mlr_ensemble_persistent_grid": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "mlr_ensemble_persistent_wait_limit": (C.c_int, [_P, C.c_double]),
    "mlr_write_vtu": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_char_p), C.c_int]),
    "mlr_layout_write_vtu": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_char_p), C.c_int, _P, _P, _P]),
    "mlr_expr_eval": (C.c_int, [C.c_char_p, C.c_int64, _P, _P, _P]),
    "mlr_dataset_create": (C.c_int, [C.c_int64, C.POINTER(_P)]),
    "mlr_dataset_destroy": (C.c_int, [_P]),
    "mlr_dataset_size": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "mlr_dataset_get": (C.c_int, [_P, C.c_int64, C.c_int64, _P, _P]),
    "mlr_collect_features": (C.c_int, [_P, _P]),
    "mlr_indicator_features": (C.c_int, [_P, _P, _P]),
    "mlr_mlp_forward": (C.c_int, [_P, _P, C.c_int64, _P]),
    "mlr_mlp_grad": (C.c_int, [_P, _P, _P, C.c_int64, C.c_float, C.c_float, _P, _P]),
    "mlr_mlp_train": (C.c_int, [_P, _P, _P, C.c_int64, C.c_int, C.c_int, C.c_float, C.c_uint64, C.c_float,
                                C.c_float, _P]),
    "mlr_set_indicator": (C.c_int, [_P, _P, C.c_double]),
    "mlr_indicator_gated": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int64)]),
    "mlr_get_volumes": (C.c_int, [_P, _P]),
    "mlr_host_pow": (C.c_int, [_P, _P, _P, C.c_int]),
    "mlr_host_pow6": (C.c_int, [_P, _P, C.c_int]),
    "mlr_calc_face_values": (C.c_int, [_P, _P, _P]),
    "mlr_compact_tables": (C.c_int, [_P]),
    "mlr_calc_rhs_of": (C.c_int, [_P, _P, C.c_int, _P]),
    "mlr_calc_rhs_device": (C.c_int, [_P, _P, C.c_int, _P]),
    "mlr_riemann_flux": (C.c_int, [C.c_int, C.c_int, C.c_int64, _P, _P, _P, _P]),
    "mlr_riemann_flux_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_int64, _P, _P, _P, _P]),
    "mlr_riemann_flux_noalloc_capacity": (C.c_int64, []),
    "mlr_integrator_take_step": (C.c_int, [_P, C.c_int, C.c_double, C.c_int64, _P, _P, RHS_FN, _P]),
}

# residual parts (mlr_calc_rhs_of / mlr_calc_rhs_device)
RHS_INTERIOR, RHS_BOUNDARY, RHS_ALL = 1, 2, 3

_lib = None


def lib():
    """Load lib/libmallard_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"mallard_amd native library missing: {LIB_PATH} "
                               "(build it with `make -C mallard_amd` or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


class MallardError(RuntimeError):
    pass


def numerics(recon="TENO", basis="legendre", order=2, stencil_factor=2.0, quad_order_cell=0,
             quad_order_face=0, teno_variant="reference", riemann="HLLC", integrator="SSPRK3",
             gamma=1.4, p_ref=101325.0, T_ref=298.15, rho_ref=1.225, p_min=-1e20, p_max=1e20,
             cfl=0.4, dt=0.0, bcs=None, n_threads=0, physics="euler", mu_ref=0.0, T_mu_ref=0.0,
             sutherland_S=-1.0, prandtl=0.72, positivity=False):
    """Numerics descriptor with the reference's defaults and key strings.

    bcs: list of (zone, type, params) with the [[boundaries]] keys.
    physics: "euler" (the reference's) or "navier-stokes" (Sutherland
    viscosity mu_ref (T/T_mu_ref)^1.5 (T_mu_ref+S)/(T+S), constant for S < 0;
    T_mu_ref <= 0 takes T_ref).
    positivity: the positivity-preserving flux limiter (Hu, Adams & Shu
    2013) on every stage update (FE and SSPRK3, Euler).
    """
    d = MlrNumerics()
    d.recon, d.basis, d.order = RECON[recon], BASIS[basis], order
    d.stencil_factor, d.quad_order_cell, d.quad_order_face = stencil_factor, quad_order_cell, quad_order_face
    d.teno_variant = TENO_VARIANT[teno_variant]
    d.riemann, d.integrator = RIEMANN[riemann], INTEGRATOR[integrator]
    d.gamma, d.p_ref, d.T_ref, d.rho_ref, d.p_min, d.p_max = gamma, p_ref, T_ref, rho_ref, p_min, p_max
    d.cfl, d.dt = cfl, dt
    bcs = bcs if bcs is not None else [(z, "symmetry", {}) for z in ("left", "right", "top", "bottom")]
    if len(bcs) > 14:
        raise ValueError("at most 14 boundary conditions")
    d.n_bc = len(bcs)
    for i, (zone, kind, par) in enumerate(bcs):
        b = d.bc[i]
        b.zone = zone.encode()
        b.type = BC[kind]
        u = par.get("u", (0.0, 0.0))
        b.u[0], b.u[1] = u
        b.p, b.T = par.get("p", 0.0), par.get("T", 0.0)
    d.n_threads = n_threads
    if physics not in ("euler", "navier-stokes"):
        raise ValueError("Unknown physics type: " + physics)
    d.viscous = 1 if physics == "navier-stokes" else 0
    d.mu_ref, d.T_mu_ref, d.sutherland_S, d.prandtl = mu_ref, T_mu_ref, sutherland_S, prandtl
    d.positivity = 1 if positivity else 0
    return d


def mesh_gen(kind, nx, ny, Lx=1.0, Ly=1.0, jitter=0.0, seed=0):
    g = MlrMeshGen()
    g.type, g.nx, g.ny, g.Lx, g.Ly, g.jitter, g.seed = MESH[kind], nx, ny, Lx, Ly, jitter, seed
    return g


class Layout:
    """Host-only precompute (mesh, SFC order, packed TENO layout); no GPU."""

    def __init__(self, gen, num):
        L = lib()
        h = C.c_void_p()
        if L.mlr_layout_create(C.byref(gen), C.byref(num), C.byref(h)):
            raise MallardError(L.mlr_last_error(None).decode())
        self.h = h
        s = np.zeros(14, np.int64)
        L.mlr_layout_sizes(self.h, _p(s))
        (self.n_cells, self.n_faces, self.n_faces_real, self.S, self.Q, self.K, self.M, self.G,
         self.device_bytes, self.n_nodes, self.n_noc, self.n_zones, self.tile_max, self.tile_pos_bytes) = map(int, s)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.mlr_layout_destroy(self.h)
            self.h = None

    def permutation(self):
        p = np.zeros(self.n_cells, np.uint32)
        lib().mlr_layout_permutation(self.h, _p(p))
        return p

    def mesh_arrays(self):
        nc, nf = self.n_cells, self.n_faces
        out = dict(node_coords=np.zeros((self.n_nodes, 2)), cell_coords=np.zeros((nc, 2)),
                   cell_volume=np.zeros(nc), face_area=np.zeros(nf), face_normals=np.zeros((nf, 2)),
                   nodes_of_cell=np.zeros(self.n_noc, np.uint32), faces_of_cell=np.zeros(self.n_noc, np.uint32),
                   nodes_of_face=np.zeros(2 * nf, np.uint32), cells_of_face=np.zeros((nf, 2), np.int32))
        lib().mlr_layout_mesh_arrays(self.h, *[_p(out[k]) for k in (
            "node_coords", "cell_coords", "cell_volume", "face_area", "face_normals", "nodes_of_cell",
            "faces_of_cell", "nodes_of_face", "cells_of_face")])
        return out

    def cell_stencils(self, cell):
        ids = np.zeros(self.G * self.M, np.uint32)
        if lib().mlr_layout_cell_stencils(self.h, cell, _p(ids)):
            raise MallardError(lib().mlr_last_error(None).decode())
        return ids.reshape(self.G, self.M)

    def cell_matrices(self, cell):
        a = np.zeros((self.G, self.K, self.M))
        t = np.zeros((self.G, self.M))
        fb = np.zeros((self.S, self.Q, self.K - 1))
        if lib().mlr_layout_cell_matrices(self.h, cell, _p(a), _p(t), _p(fb)):
            raise MallardError(lib().mlr_last_error(None).decode())
        return a, t, fb

    def teno_tables(self):
        ipt = np.zeros(self.K)
        oi = np.zeros(self.K * self.K)
        lib().mlr_layout_teno_tables(self.h, _p(ipt), _p(oi))
        return ipt, oi.reshape(self.K, self.K)

    def write_vtu(self, path, variables, U, prim, cfl):
        """The VTU writer on this layout's mesh with caller arrays (host only)."""
        names = (C.c_char_p * len(variables))(*[v.encode() for v in variables])
        if lib().mlr_layout_write_vtu(self.h, os.fspath(path).encode(), names, len(variables), _p(_f64(U)),
                                      _p(_f64(prim)), _p(_f64(cfl))):
            raise MallardError(lib().mlr_last_error(None).decode())

    def slots(self, cell):
        S = self.S
        nb = np.zeros(S, np.uint32)
        info = np.zeros(S, np.uint32)
        nx, ny, area, dxn = (np.zeros(S) for _ in range(4))
        lib().mlr_layout_slots(self.h, cell, _p(nb), _p(info), _p(nx), _p(ny), _p(area), _p(dxn))
        return dict(nb=nb, info=info, nx=nx, ny=ny, area=area, dxn=dxn)


class Solver:
    """GPU solver context (Solver, solver.h:30-279) behind the C ABI."""

    def __init__(self, toml_path=None, gen=None, num=None, mesh=None):
        """From a TOML input, a generated mesh (gen) or mesh arrays in the
        reference layout (mesh: the dictionary meshgen.mesh_from_triangles
        returns), with the numerics descriptor num."""
        L = lib()
        h = C.c_void_p()
        if toml_path is not None:
            rc = L.mlr_ctx_create_toml(os.fspath(toml_path).encode(), C.byref(h))
        elif mesh is not None:
            keep = {k: np.ascontiguousarray(mesh[k], dt) for k, dt in (
                ("node_coords", np.float64), ("offsets_nodes_of_cell", np.uint32), ("nodes_of_cell", np.uint32),
                ("faces_of_cell", np.uint32), ("nodes_of_face", np.uint32), ("cells_of_face", np.int32),
                ("zone_offsets", np.uint32), ("zone_faces", np.uint32))}
            names = (C.c_char_p * len(mesh["zone_names"]))(*[z.encode() for z in mesh["zone_names"]])
            d = MlrMeshDesc()
            d.n_nodes = len(keep["node_coords"]) // 2 if keep["node_coords"].ndim == 1 else len(keep["node_coords"])
            d.node_coords = _p(keep["node_coords"])
            d.n_cells = len(keep["offsets_nodes_of_cell"]) - 1
            d.offsets_nodes_of_cell = _p(keep["offsets_nodes_of_cell"])
            d.nodes_of_cell = _p(keep["nodes_of_cell"])
            d.faces_of_cell = _p(keep["faces_of_cell"])
            d.n_faces = len(keep["nodes_of_face"]) // 2
            d.nodes_of_face = _p(keep["nodes_of_face"])
            d.cells_of_face = _p(keep["cells_of_face"])
            d.n_zones = len(mesh["zone_names"])
            d.zone_names = C.cast(names, C.c_void_p)
            d.zone_offsets = _p(keep["zone_offsets"])
            d.zone_faces = _p(keep["zone_faces"])
            rc = L.mlr_ctx_create(C.byref(d), C.byref(num), C.byref(h))
        else:
            rc = L.mlr_ctx_create_generated(C.byref(gen), C.byref(num), C.byref(h))
        if rc:
            raise MallardError(L.mlr_last_error(None).decode())
        self.h = h
        s = np.zeros(15, np.int64)
        L.mlr_get_sizes(self.h, _p(s))
        self.n_cells, self.n_faces, self.n_faces_real, self.S, self.Q, self.K, self.M, self.G = map(int, s[:8])
        self.n_stencils = int(s[12])
        self.tile_entries, self.tile_pos_bytes = int(s[13]), int(s[14])

    @property
    def device_bytes(self):
        """Device memory the context holds now."""
        s = np.zeros(15, np.int64)
        lib().mlr_get_sizes(self.h, _p(s))
        return int(s[8])

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.mlr_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc:
            raise MallardError(lib().mlr_last_error(self.h).decode())

    def init_quadrant(self, xs, ys, states):
        st = _f64(states).reshape(16)
        self._check(lib().mlr_init_quadrant(self.h, xs, ys, _p(st)))

    def init_constant(self, u, p, T):
        self._check(lib().mlr_init_constant(self.h, _p(_f64(u)), p, T))

    def set_state(self, U, prim=None):
        U = _f64(U)
        prim = _f64(prim) if prim is not None else None
        self._check(lib().mlr_set_state(self.h, _p(U), _p(prim)))

    def state(self):
        U = np.zeros((self.n_cells, 4))
        W = np.zeros((self.n_cells, 5))
        self._check(lib().mlr_get_state(self.h, _p(U), _p(W)))
        return U, W

    def calc_rhs(self):
        r = np.zeros((self.n_cells, 4))
        self._check(lib().mlr_calc_rhs(self.h, _p(r)))
        return r

    def calc_dt(self):
        dt = C.c_double()
        self._check(lib().mlr_compute_dt(self.h, C.byref(dt)))
        return dt.value

    def take_step(self):
        self._check(lib().mlr_step(self.h))

    def run_steps(self, n):
        dts = np.zeros(n)
        self._check(lib().mlr_run_steps(self.h, n, _p(dts)))
        return dts

    def run_steps_profiled(self, n):
        """n steps launched kernel by kernel with HIP events between launches;
        returns event-measured milliseconds and launch counts per kernel kind."""
        st = np.zeros(9)
        self._check(lib().mlr_run_steps_profiled(self.h, n, _p(st)))
        return dict(total_ms=st[0], teno_ms=st[1], teno_n=int(st[2]), flux_ms=st[3], flux_n=int(st[4]),
                    dt_ms=st[5], dt_n=int(st[6]), copy_ms=st[7], copy_n=int(st[8]))

    def autotune(self):
        """Time the launch-shape candidates on the current state and keep the
        fastest: (TENO occupancy target, flux block, spectral block, TENO
        occupancy target of gated passes reusing recorded decisions, plain
        TENO face-basis source: 0 stored, 1 derived from node coordinates)."""
        t = np.zeros(5, np.int32)
        self._check(lib().mlr_autotune(self.h, _p(t)))
        return tuple(int(x) for x in t)

    def set_tune(self, teno_occ=0, flux_block=256, spec_block=256, teno_occ_reuse=0, face_basis=0):
        t = np.array([teno_occ, flux_block, spec_block, teno_occ_reuse, face_basis], np.int32)
        self._check(lib().mlr_set_tune(self.h, _p(t)))

    def tune(self):
        t = np.zeros(5, np.int32)
        self._check(lib().mlr_get_tune(self.h, _p(t)))
        return tuple(int(x) for x in t)

    def run(self):
        self._check(lib().mlr_run(self.h))

    def set_cfl(self, cfl):
        self._check(lib().mlr_set_cfl(self.h, cfl))

    def set_dt_history(self, steps):
        """Length of the device ring of per-step dt values (runs are issued
        in chunks of at most this many steps)."""
        self._check(lib().mlr_set_dt_history(self.h, steps))

    def time(self):
        t, s = C.c_double(), C.c_int64()
        lib().mlr_get_time(self.h, C.byref(t), C.byref(s))
        return t.value, s.value

    def face_values(self):
        out = np.zeros((self.n_faces, self.Q, 2, 4))
        self._check(lib().mlr_get_face_values(self.h, _p(out)))
        return out

    def compact_tables(self):
        """Fold the member areas into A+ and take the face-point basis values
        from the reference triangle's edges (same values in exact arithmetic,
        fewer bytes per TENO pass); one way."""
        self._check(lib().mlr_compact_tables(self.h))

    def calc_face_values(self, U):
        """FaceReconstruction::calc_face_values of a caller's state U[N][4]
        (reference numbering) -> face_solution[n_faces][Q][2][4]; the
        solver's own state is untouched."""
        U = _f64(U)
        out = np.zeros((self.n_faces, self.Q, 2, 4))
        self._check(lib().mlr_calc_face_values(self.h, _p(U), _p(out)))
        return out

    def calc_rhs_of(self, U, parts=RHS_ALL):
        """Residual of a caller's state over the interior faces (1), the
        boundary faces (2) or both (3), divided by the volume."""
        U = _f64(U)
        r = np.zeros((self.n_cells, 4))
        self._check(lib().mlr_calc_rhs_of(self.h, _p(U), parts, _p(r)))
        return r

    def calc_rhs_device(self, U_ptr, rhs_ptr, parts=RHS_ALL):
        """Residual of the device array at U_ptr ([N][4], this solver's
        numbering) into the device array at rhs_ptr, queued on the solver's
        stream."""
        self._check(lib().mlr_calc_rhs_device(self.h, C.c_void_p(U_ptr), parts, C.c_void_p(rhs_ptr)))

    def decisions(self):
        d = np.zeros(self.n_cells, np.uint8)
        self._check(lib().mlr_get_decisions(self.h, _p(d)))
        return d

    def cfl_local(self):
        d = np.zeros(self.n_cells)
        self._check(lib().mlr_get_cfl_local(self.h, _p(d)))
        return d

    def permutation(self):
        p = np.zeros(self.n_cells, np.uint32)
        self._check(lib().mlr_get_permutation(self.h, _p(p)))
        return p

    def write_vtu(self, path, variables=("RHO", "RHOU_X", "RHOU_Y", "RHOE", "U_X", "U_Y", "P", "T", "H", "CFL")):
        names = (C.c_char_p * len(variables))(*[v.encode() for v in variables])
        self._check(lib().mlr_write_vtu(self.h, os.fspath(path).encode(), names, len(variables)))

    # ---- learned troubled-cell indicator
    def set_indicator(self, params, threshold=0.9, every_stage=False, split=True):
        """Gate TENO with the network (params: 7361 FP32 values; None: off).
        every_stage False: the network judges the cells on each step's first
        stage and the later stages reuse the decisions; True: every stage.
        split True: gated and ungated cells are reconstructed by two kernels;
        False: by one (the results are identical)."""
        if params is None:
            self._check(lib().mlr_set_indicator(self.h, None, threshold))
        else:
            p = np.ascontiguousarray(params, dtype=np.float32)
            if p.size != N_PARAMS:
                raise ValueError("indicator parameters must have %d values" % N_PARAMS)
            self._check(lib().mlr_set_indicator(self.h, _p(p), threshold))
            self._check(lib().mlr_indicator_cadence(self.h, int(every_stage)))
            self._check(lib().mlr_indicator_split(self.h, int(split)))

    def gate_decisions(self):
        """The network's last decisions inside the TENO pass (1: gated), per cell."""
        out = np.zeros(self.n_cells, np.uint8)
        self._check(lib().mlr_indicator_decisions(self.h, _p(out)))
        return out

    def gated(self, reset=False):
        n = C.c_int64()
        self._check(lib().mlr_indicator_gated(self.h, int(reset), C.byref(n)))
        return n.value

    def indicator_features(self):
        """bf16 features (uint16 [N][48]) and TENO labels of the current state."""
        X = np.zeros((self.n_cells, N_FEATURES), np.uint16)
        y = np.zeros(self.n_cells, np.float32)
        self._check(lib().mlr_indicator_features(self.h, _p(X), _p(y)))
        return X, y

    def collect_features(self, ds):
        self._check(lib().mlr_collect_features(self.h, ds.h))

    def volumes(self):
        v = np.zeros(self.n_cells)
        self._check(lib().mlr_get_volumes(self.h, _p(v)))
        return v

    def device_matrices(self, cell):
        a = np.zeros((self.G, self.K, self.M))
        t = np.zeros((self.G, self.M))
        self._check(lib().mlr_get_device_matrices(self.h, cell, _p(a), _p(t)))
        return a, t

    def device_stencils(self, cell):
        ids = np.zeros(self.G * self.M, np.uint32)
        self._check(lib().mlr_get_device_stencils(self.h, cell, _p(ids)))
        return ids.reshape(self.G, self.M)


def parse_config(path):
    """Host-only parse of a Mallard TOML input: (mesh_gen, numerics, run dict, init type)."""
    g, d = MlrMeshGen(), MlrNumerics()
    run = np.zeros(7)
    it = np.zeros(1, np.int32)
    if lib().mlr_config_parse(os.fspath(path).encode(), C.byref(g), C.byref(d), _p(run), _p(it)):
        raise MallardError(lib().mlr_last_error(None).decode())
    keys = ("cfl", "dt", "n_steps", "t_stop", "t_wall_stop", "check_interval", "check_nan")
    return g, d, dict(zip(keys, run.tolist())), ("constant", "analytical", "quadrant")[int(it[0])]


def eval_expr(expr, x, y):
    """The analytical-initialisation expression evaluator at points (x, y)."""
    x, y = _f64(x).ravel(), _f64(y).ravel()
    out = np.zeros(len(x))
    if lib().mlr_expr_eval(expr.encode(), len(x), _p(x), _p(y), _p(out)):
        raise MallardError(lib().mlr_last_error(None).decode())
    return out


class Ensemble:
    """Independent cases advanced together, one launch per kernel phase."""

    def __init__(self, gens, num):
        arr = (MlrMeshGen * len(gens))(*gens)
        h = C.c_void_p()
        if lib().mlr_ensemble_create(C.cast(arr, C.c_void_p), len(gens), C.byref(num), C.byref(h)):
            raise MallardError(lib().mlr_last_error(None).decode())
        self.h = h
        self.n = len(gens)
        s = np.zeros(self.n, np.int64)
        lib().mlr_ensemble_sizes(self.h, _p(s))
        self.n_cells = s

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.mlr_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc:
            raise MallardError(lib().mlr_last_error(None).decode())

    def init_quadrant(self, k, xs, ys, states):
        self._check(lib().mlr_ensemble_init_quadrant(self.h, k, xs, ys, _p(_f64(states).reshape(16))))

    def state(self, k):
        U = np.zeros((int(self.n_cells[k]), 4))
        W = np.zeros((int(self.n_cells[k]), 5))
        self._check(lib().mlr_ensemble_get_state(self.h, k, _p(U), _p(W)))
        return U, W

    def volumes(self, k):
        v = np.zeros(int(self.n_cells[k]))
        self._check(lib().mlr_ensemble_get_volumes(self.h, k, _p(v)))
        return v

    def time(self, k):
        t, s = C.c_double(), C.c_int64()
        lib().mlr_ensemble_get_time(self.h, k, C.byref(t), C.byref(s))
        return t.value, s.value

    def run_steps(self, n):
        self._check(lib().mlr_ensemble_run_steps(self.h, n))

    def run_steps_profiled(self, n):
        st = np.zeros(9)
        self._check(lib().mlr_ensemble_run_steps_profiled(self.h, n, _p(st)))
        return dict(total_ms=st[0], teno_ms=st[1], teno_n=int(st[2]), flux_ms=st[3], flux_n=int(st[4]),
                    dt_ms=st[5], dt_n=int(st[6]))

    def set_indicator(self, params, threshold=0.9, every_stage=False, split=True):
        p = None if params is None else _params(params)
        self._check(lib().mlr_ensemble_set_indicator(self.h, _p(p), threshold))
        self._check(lib().mlr_ensemble_indicator_cadence(self.h, int(every_stage)))
        if p is not None:
            self._check(lib().mlr_ensemble_indicator_split(self.h, int(split)))

    def autotune(self):
        """(TENO occupancy target, block, block, reuse occupancy target,
        face-basis source)."""
        t = np.zeros(5, np.int32)
        self._check(lib().mlr_ensemble_autotune(self.h, _p(t)))
        return tuple(int(x) for x in t)

    def set_tune(self, teno_occ=0, block=256, teno_occ_reuse=0, face_basis=0):
        t = np.array([teno_occ, block, block, teno_occ_reuse, face_basis], np.int32)
        self._check(lib().mlr_ensemble_set_tune(self.h, _p(t)))

    def gated(self, reset=False):
        n = C.c_int64()
        self._check(lib().mlr_ensemble_gated(self.h, int(reset), C.byref(n)))
        return n.value

    def set_persistent(self, on=True):
        """Each step as one persistent cooperative launch over every case
        (plain TENO p = 2 or first order; 256-cell flux blocks)."""
        self._check(lib().mlr_ensemble_set_persistent(self.h, int(bool(on))))

    def persistent_wait_limit(self, ms):
        """Longest wait of a block at a grid barrier of the persistent step
        before the step is abandoned and run_steps raises (default 2000 ms)."""
        self._check(lib().mlr_ensemble_persistent_wait_limit(self.h, float(ms)))

    @property
    def persistent_grid(self):
        """Blocks of the last persistent launch (0 before the first)."""
        b = C.c_int32()
        self._check(lib().mlr_ensemble_persistent_grid(self.h, C.byref(b)))
        return b.value


N_FEATURES = 48
N_PARAMS = 7361


class Dataset:
    """Device-resident training rows (bf16 features, TENO label)."""

    def __init__(self, capacity):
        h = C.c_void_p()
        if lib().mlr_dataset_create(int(capacity), C.byref(h)):
            raise MallardError(lib().mlr_last_error(None).decode())
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.mlr_dataset_destroy(self.h)
            self.h = None

    def __len__(self):
        n = C.c_int64()
        lib().mlr_dataset_size(self.h, C.byref(n))
        return n.value

    def get(self, start=0, n=None, features=True):
        """(X, y) of rows start..start+n; X is None when features=False."""
        n = len(self) - start if n is None else n
        X = np.zeros((n, N_FEATURES), np.uint16) if features else None
        y = np.zeros(n, np.float32)
        if lib().mlr_dataset_get(self.h, start, n, _p(X), _p(y)):
            raise MallardError(lib().mlr_last_error(None).decode())
        return X, y


def _params(p):
    p = np.ascontiguousarray(p, dtype=np.float32)
    if p.size != N_PARAMS:
        raise ValueError("indicator parameters must have %d values" % N_PARAMS)
    return p


def mlp_forward(params, X):
    """Logits of bf16 feature rows (uint16 [n][48]) by the GPU inference kernel."""
    X = np.ascontiguousarray(X, dtype=np.uint16)
    z = np.zeros(len(X), np.float32)
    if lib().mlr_mlp_forward(_p(_params(params)), _p(X), len(X), _p(z)):
        raise MallardError(lib().mlr_last_error(None).decode())
    return z


def mlp_grad(params, ds, idx, pos_w=1.0, neg_w=1.0):
    """Gradient and loss of the weighted mean BCE over dataset rows idx (GPU kernels)."""
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    g = np.zeros(N_PARAMS, np.float32)
    loss = np.zeros(1, np.float32)
    if lib().mlr_mlp_grad(_p(_params(params)), ds.h, _p(idx), len(idx), pos_w, neg_w, _p(g), _p(loss)):
        raise MallardError(lib().mlr_last_error(None).decode())
    return g, float(loss[0])


def mlp_train(params, ds, pool, steps, batch, lr, seed=0, pos_w=1.0, neg_w=1.0):
    """Adam on the GPU; returns (trained parameters, loss per step)."""
    p = _params(params).copy()
    pool = np.ascontiguousarray(pool, dtype=np.uint32)
    hist = np.zeros(steps, np.float32)
    if lib().mlr_mlp_train(_p(p), ds.h, _p(pool), len(pool), steps, batch, lr, seed, pos_w, neg_w, _p(hist)):
        raise MallardError(lib().mlr_last_error(None).decode())
    return p, hist


def riemann_flux(rs, n_unit, left, right, guard=0):
    """RiemannSolver::calc_flux over n points on the GPU: n_unit[n][2],
    left/right[n][6] = (rho, u_x, u_y, p, gamma, h) -> flux[n][4]."""
    n_unit, left, right = _f64(n_unit), _f64(left), _f64(right)
    n = n_unit.shape[0]
    if n_unit.shape != (n, 2) or left.shape != (n, 6) or right.shape != (n, 6):
        raise ValueError("riemann_flux: n_unit[n][2], left[n][6], right[n][6]")
    out = np.zeros((n, 4))
    t = RIEMANN[rs] if isinstance(rs, str) else int(rs)
    if lib().mlr_riemann_flux(t, guard, n, _p(n_unit), _p(left), _p(right), _p(out)):
        raise MallardError(lib().mlr_last_error(None).decode())
    return out


def riemann_flux_noalloc(rs, n_unit, left, right, guard=0):
    """riemann_flux staged through the library's fixed device scratch (no
    device allocation; at most riemann_flux_noalloc_capacity() points)."""
    n_unit, left, right = _f64(n_unit), _f64(left), _f64(right)
    n = n_unit.shape[0]
    if n_unit.shape != (n, 2) or left.shape != (n, 6) or right.shape != (n, 6):
        raise ValueError("riemann_flux_noalloc: n_unit[n][2], left[n][6], right[n][6]")
    out = np.zeros((n, 4))
    t = RIEMANN[rs] if isinstance(rs, str) else int(rs)
    if lib().mlr_riemann_flux_noalloc(t, guard, n, _p(n_unit), _p(left), _p(right), _p(out)):
        raise MallardError(lib().mlr_last_error(None).decode())
    return out


def riemann_flux_noalloc_capacity():
    return int(lib().mlr_riemann_flux_noalloc_capacity())


def integrator_take_step(integrator, dt, n, solution_ptrs, rhs_ptrs, rhs_fn, solver=None):
    """TimeIntegrator::take_step on device vectors of n doubles.
    solution_ptrs / rhs_ptrs: device addresses; rhs_fn(U_ptr, rhs_ptr) fills
    the residual (device pointers). With a solver, the updates run on its
    stream."""
    t = INTEGRATOR[integrator] if isinstance(integrator, str) else int(integrator)
    sol = (C.c_void_p * len(solution_ptrs))(*solution_ptrs)
    rhs = (C.c_void_p * len(rhs_ptrs))(*rhs_ptrs)

    def _cb(U, r, _user):
        try:
            rhs_fn(U, r)
            return 0
        except Exception:
            return 1

    cb = RHS_FN(_cb)
    h = solver.h if solver is not None else None
    if lib().mlr_integrator_take_step(h, t, dt, n, sol, rhs, cb, None):
        raise MallardError(lib().mlr_last_error(h).decode())


def probe_stream(kind, nbytes, reps=10):
    """Mean ms of a streaming copy (kind 0) or read (kind 1) of nbytes on the
    GPU; kind 2 hipMemcpyAsync device to device, kind 3 a read of 64
    contiguous bytes per lane, kind 4 a read with non-temporal loads."""
    ms = C.c_double()
    if lib().mlr_probe_stream(kind, int(nbytes), reps, C.byref(ms)):
        raise MallardError(lib().mlr_last_error(None).decode())
    return ms.value


def probe_math(which, x, y=None):
    """Device sqrt (0), division (1), pow (2) or x^6 (3) of arrays."""
    x = _f64(x)
    y = _f64(y) if y is not None else x
    out = np.zeros_like(x)
    rc = lib().mlr_probe_math(which, _p(x), _p(y), _p(out), x.size)
    if rc:
        raise MallardError(lib().mlr_last_error(None).decode())
    return out


def host_pow(x, y):
    x, y = _f64(x), _f64(y)
    out = np.zeros_like(x)
    lib().mlr_host_pow(_p(x), _p(y), _p(out), x.size)
    return out


def host_pow6(x):
    x = _f64(x)
    out = np.zeros_like(x)
    lib().mlr_host_pow6(_p(x), _p(out), x.size)
    return out
