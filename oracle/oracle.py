"""ctypes front end of the CPU oracle (oracle/simplic_oracle.cpp).

Test infrastructure only: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of bench.py, always as the checker.  The product
package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle_simplic.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    src = os.path.join(_HERE, "simplic_oracle.cpp")
    if force or not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(src):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "-s"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    vp = C.c_void_p
    L.or_mesh_new.restype = vp
    L.or_mesh_new.argtypes = [C.c_int, _dp, C.c_int, C.c_int, _ip, _ip, _ip, _ip, C.c_int, C.c_int,
                              _ip, _ip, _ip]
    L.or_mesh_free.argtypes = [vp]
    L.or_mesh_geometry.argtypes = [vp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.or_mesh_stencil_size.argtypes = [vp, C.c_int]
    L.or_mesh_stencil_size.restype = C.c_int
    L.or_calc_sub_face.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp, _dp, C.POINTER(C.c_int),
                                   C.POINTER(C.c_int)]
    L.or_calc_sub_face.restype = C.c_int
    L.or_time_integrated_area.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp, C.c_double, C.c_double,
                                          C.c_double]
    L.or_time_integrated_area.restype = C.c_double
    L.or_time_integrated_face_flux.argtypes = [vp, C.c_int, _dp, C.c_double, C.c_double, C.c_double,
                                               C.c_double, C.c_double]
    L.or_time_integrated_face_flux.restype = C.c_double
    L.or_calc_sub_cell.argtypes = [vp, C.c_int, _dp, C.c_double, C.c_int, _dp]
    L.or_calc_sub_cell.restype = C.c_int
    L.or_find_signed_distance.argtypes = [vp, C.c_int, C.c_double, _dp, C.c_int, _dp]
    L.or_find_signed_distance.restype = C.c_int
    L.or_solver_new.restype = vp
    L.or_solver_new.argtypes = [vp, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int]
    L.or_solver_free.argtypes = [vp]
    L.or_solver_set_orientation.argtypes = [vp, C.c_int]
    L.or_solver_set_grad_scheme.argtypes = [vp, C.c_int]
    L.or_grad_normals.argtypes = [vp, _dp, _dp]
    L.or_set_alpha.argtypes = [vp, _dp]
    L.or_set_flow.argtypes = [vp, _dp, _dp, C.c_void_p]
    L.or_reconstruct.argtypes = [vp]
    L.or_advect.argtypes = [vp, C.c_double, C.c_void_p, C.c_void_p]
    L.or_get_alpha.argtypes = [vp, _dp]
    L.or_get_dvf.argtypes = [vp, _dp]
    L.or_get_alpha_phi.argtypes = [vp, _dp]
    L.or_get_rho_phi.argtypes = [vp, C.c_double, C.c_double, _dp]
    L.or_map_alpha_field.argtypes = [vp, _dp, _dp, C.c_double, C.c_double]
    L.or_solver_set_rdf.argtypes = [vp, C.c_double, C.c_double, C.c_int]
    L.or_rdf_iterations_run.argtypes = [vp]
    L.or_rdf_iterations_run.restype = C.c_int
    L.or_plic_surface.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_void_p, C.c_void_p, C.c_void_p]
    L.or_advect_sub_cycles.argtypes = [vp, C.c_double, C.c_int, C.c_void_p, C.c_void_p, C.c_double,
                                       C.c_double, _dp]
    L.or_n_mixed.argtypes = [vp]
    L.or_n_mixed.restype = C.c_int
    L.or_get_mixed.argtypes = [vp, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp]
    L.or_set_normals.argtypes = [vp, _dp]
    L.or_get_cut_class.argtypes = [vp, _ip, np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")]
    L.or_timers.argtypes = [vp, _dp]
    L.or_bound_log.argtypes = [vp, _dp, C.c_int]
    L.or_bound_log.restype = C.c_int
    L.or_point_u.argtypes = [vp, C.c_int, _dp]
    L.or_interpolate_u.argtypes = [vp, _dp, C.c_int, _dp]
    L.or_ls_grad.argtypes = [C.c_int, _dp, _dp, _dp]
    L.or_set_mutation.argtypes = [C.c_int]
    L.or_limit_flux.argtypes = [vp, C.c_double, _dp, _dp, _dp, _ip, C.c_int, _dp, _dp]
    L.or_deformation_field.argtypes = [vp, _dp, C.c_void_p, _dp]
    L.or_courant.argtypes = [vp, _dp, _dp, _dp]
    _lib = L
    return L


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class OracleMesh:
    """Oracle view of a :class:`geometricvofext_amd.mesh.PolyMesh`."""

    def __init__(self, mesh):
        L = lib()
        self.mesh = mesh
        self._keep = [
            _c(mesh.points.ravel(), np.float64), _c(mesh.face_offsets, np.int32),
            _c(mesh.face_points, np.int32), _c(mesh.owner, np.int32),
            _c(mesh.neighbour if mesh.neighbour.size else np.zeros(1), np.int32),
            _c(mesh.patch_start, np.int32), _c(mesh.patch_size, np.int32), _c(mesh.patch_type, np.int32),
        ]
        k = self._keep
        self.h = L.or_mesh_new(mesh.n_points, k[0], mesh.n_faces, mesh.n_internal_faces, k[1], k[2],
                               k[3], k[4], mesh.n_cells, len(mesh.patch_start), k[5], k[6], k[7])
        nf, nc = mesh.n_faces, mesh.n_cells
        self.Cf = np.zeros((nf, 3))
        self.Sf = np.zeros((nf, 3))
        self.magSf = np.zeros(nf)
        self.C = np.zeros((nc, 3))
        self.V = np.zeros(nc)
        self.flatness = np.zeros(nf)
        L.or_mesh_geometry(self.h, self.Cf.reshape(-1), self.Sf.reshape(-1), self.magSf,
                           self.C.reshape(-1), self.V, self.flatness)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_mesh_free(self.h)
            self.h = None

    # primitives ---------------------------------------------------------
    def calc_sub_cell(self, cell, n, d, split=False):
        out = np.zeros(11)
        st = lib().or_calc_sub_cell(self.h, cell, _c(n, np.float64), d, int(split), out)
        return st, out

    def find_signed_distance(self, cell, alpha, n, split=False):
        out = np.zeros(10)
        st = lib().or_find_signed_distance(self.h, cell, alpha, _c(n, np.float64), int(split), out)
        return st, out

    def time_integrated_face_flux(self, face, n, d, un0, dt, phi, magsf):
        return lib().or_time_integrated_face_flux(self.h, face, _c(n, np.float64), d, un0, dt, phi, magsf)

    def deformation_field(self):
        nc, nf = self.mesh.n_cells, self.mesh.n_faces
        U = np.zeros(3 * nc)
        phi = np.zeros(nf)
        lib().or_deformation_field(self.h, U, None, phi)
        return U.reshape(nc, 3), phi

    def courant(self, phi, alpha):
        out = np.zeros(2)
        lib().or_courant(self.h, _c(phi, np.float64), _c(alpha, np.float64), out)
        return out


def calc_sub_face(pts, n, d):
    pts = _c(np.asarray(pts).reshape(-1), np.float64)
    area = np.zeros(3)
    centre = np.zeros(3)
    ns, nc = C.c_int(0), C.c_int(0)
    st = lib().or_calc_sub_face(pts.size // 3, pts, _c(n, np.float64), d, area, centre, C.byref(ns), C.byref(nc))
    return st, area, centre, ns.value, nc.value


def time_integrated_area(pts, n, d, ptimes, un0, dt, magsf):
    pts = _c(np.asarray(pts).reshape(-1), np.float64)
    return lib().or_time_integrated_area(pts.size // 3, pts, _c(n, np.float64), d, _c(ptimes, np.float64),
                                         un0, dt, magsf)


class mutation:
    """Context manager that switches on one planted mistake of the oracle
    (1: weighted least squares, 2: wrong interpolation tet, 3: boundFlux in
    descending cell order), so that a pinning test can show its check fails."""

    def __init__(self, k):
        self.k = k

    def __enter__(self):
        lib().or_set_mutation(self.k)
        return self

    def __exit__(self, *exc):
        lib().or_set_mutation(0)
        return False


def ls_grad(pos, val):
    pos = _c(np.asarray(pos).reshape(-1), np.float64)
    out = np.zeros(3)
    lib().or_ls_grad(pos.size // 3, pos, _c(val, np.float64), out)
    return out


class OracleSolver:
    """Oracle restatement of ``solveVofEqu`` (reconstruct + advect)."""

    ORIENTATIONS = {"LS": 0, "isoAlphaGrad": 0, "alphaGrad": 1, "isoRDF": 2, "RDF": 2}

    def __init__(self, omesh: OracleMesh, mixedCellTol=1e-8, nAlphaBounds=10, snapTol=0.0, clip=True,
                 splitWarpedFace=False, threads=1, orientationMethod="LS", tol=1e-6, relTol=0.1, iterations=5,
                 gradScheme="Gauss linear"):
        self.om = omesh
        self.h = lib().or_solver_new(omesh.h, mixedCellTol, nAlphaBounds, snapTol, int(clip),
                                     int(splitWarpedFace), threads)
        lib().or_solver_set_orientation(self.h, self.ORIENTATIONS[orientationMethod])
        lib().or_solver_set_rdf(self.h, float(tol), float(relTol), int(iterations))
        lib().or_solver_set_grad_scheme(self.h, {"Gauss linear": 0, "Gauss pointLinear": 1}[" ".join(gradScheme.split())])

    def grad_normals(self, alpha):
        """The alphaGrad normal (normalised -grad(alpha) with the solver's
        gradient scheme) of every cell for the given alpha."""
        out = np.zeros(3 * self.om.mesh.n_cells)
        lib().or_grad_normals(self.h, _c(alpha, np.float64), out)
        return out.reshape(-1, 3)

    def map_alpha_field(self, n, d, lower, upper):
        """mapAlphaField (reconstruction.C:725-784) with per-cell n, d."""
        lib().or_map_alpha_field(self.h, _c(np.asarray(n).reshape(-1), np.float64), _c(d, np.float64),
                                 float(lower), float(upper))

    def rdf_iterations_run(self):
        """Iterations of the last isoRDF reconstruct."""
        return lib().or_rdf_iterations_run(self.h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_solver_free(self.h)
            self.h = None

    def set_alpha(self, a):
        lib().or_set_alpha(self.h, _c(a, np.float64))

    def set_flow(self, phi, U, Ub=None):
        self._ub = None if Ub is None else _c(np.asarray(Ub).reshape(-1), np.float64)
        lib().or_set_flow(self.h, _c(phi, np.float64), _c(np.asarray(U).reshape(-1), np.float64),
                          None if self._ub is None else self._ub.ctypes.data)

    def reconstruct(self):
        lib().or_reconstruct(self.h)

    def advect(self, dt, Sp=None, Su=None):
        sp = None if Sp is None else _c(Sp, np.float64)
        su = None if Su is None else _c(Su, np.float64)
        self._keep = (sp, su)
        lib().or_advect(self.h, dt, None if sp is None else sp.ctypes.data, None if su is None else su.ctypes.data)

    def alpha(self):
        a = np.zeros(self.om.mesh.n_cells)
        lib().or_get_alpha(self.h, a)
        return a

    def dvf(self):
        a = np.zeros(self.om.mesh.n_faces)
        lib().or_get_dvf(self.h, a)
        return a

    def alpha_phi(self):
        a = np.zeros(self.om.mesh.n_faces)
        lib().or_get_alpha_phi(self.h, a)
        return a

    def rho_phi(self, rho1, rho2):
        a = np.zeros(self.om.mesh.n_faces)
        lib().or_get_rho_phi(self.h, rho1, rho2, a)
        return a

    def advect_sub_cycles(self, dt, n_sub, Sp=None, Su=None, rho1=1.0, rho2=0.0):
        """Sub-cycled alpha step (alphaEqnSubCycle.H:30-60); returns rhoPhi."""
        sp = None if Sp is None else _c(Sp, np.float64)
        su = None if Su is None else _c(Su, np.float64)
        a = np.zeros(self.om.mesh.n_faces)
        lib().or_advect_sub_cycles(self.h, dt, int(n_sub), None if sp is None else sp.ctypes.data,
                                   None if su is None else su.ctypes.data, rho1, rho2, a)
        return a

    def plic_surface(self):
        """PLIC polygons of the last reconstruct (reconstruction.C:787-835):
        points (n, 3), face offsets (n_faces + 1) and the cell of each face."""
        nf, npt = C.c_int(), C.c_int()
        lib().or_plic_surface(self.h, C.byref(nf), C.byref(npt), None, None, None)
        cells = np.zeros(max(nf.value, 1), np.int32)
        offs = np.zeros(nf.value + 1, np.int32)
        pts = np.zeros(3 * max(npt.value, 1))
        lib().or_plic_surface(self.h, C.byref(nf), C.byref(npt), cells.ctypes.data, offs.ctypes.data,
                              pts.ctypes.data)
        return dict(points=pts.reshape(-1, 3)[:npt.value], offsets=offs, cells=cells[:nf.value])

    def mixed(self):
        n = lib().or_n_mixed(self.h)
        cells = np.zeros(max(n, 1), np.int32)
        status = np.zeros(max(n, 1), np.int32)
        ncuts = np.zeros(max(n, 1), np.int32)
        nrm = np.zeros(3 * max(n, 1))
        d = np.zeros(max(n, 1))
        Cc = np.zeros(3 * max(n, 1))
        S = np.zeros(3 * max(n, 1))
        un0 = np.zeros(max(n, 1))
        lib().or_get_mixed(self.h, cells, status, ncuts, nrm, d, Cc, S, un0)
        ncp = np.zeros(max(n, 1), np.int32)
        fcl = np.zeros(max(n, 1), np.uint64)
        lib().or_get_cut_class(self.h, ncp, fcl)
        return dict(cells=cells[:n], status=status[:n], ncuts=ncuts[:n], n=nrm.reshape(-1, 3)[:n],
                    d=d[:n], C=Cc.reshape(-1, 3)[:n], S=S.reshape(-1, 3)[:n], un0=un0[:n],
                    ncutpts=ncp[:n], faceclass=fcl[:n])

    def timers(self):
        out = np.zeros(2)
        lib().or_timers(self.h, out)
        return out

    def limit_flux(self, dt, alpha, alpha_old, dvf, mixed):
        """limitFlux alone (advectionTemplates.C:118-215) on the given state;
        phi from set_flow.  Returns the bounded alpha and dVf."""
        a = np.zeros(self.om.mesh.n_cells)
        v = np.zeros(self.om.mesh.n_faces)
        mx = _c(mixed, np.int32)
        lib().or_limit_flux(self.h, dt, _c(alpha, np.float64), _c(alpha_old, np.float64), _c(dvf, np.float64),
                            mx, int(mx.size), a, v)
        return a, v

    def bound_log(self):
        out = np.zeros(64)
        n = lib().or_bound_log(self.h, out, 64)
        return out[:min(n, 64)]
