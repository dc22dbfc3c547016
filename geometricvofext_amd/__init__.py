"""MI355X SimPLIC alpha transport: Python front end of libsimplic_amd.so.

``SolveVofEqu`` mirrors the reference's ``solveVofEqu`` facade
(``src/SimPLIC/solveVofEqu/solveVofEqu.H:67-216``): construct it on a mesh with
the fvSolution ``"alpha.*"`` dictionary, then call ``reconstruct()`` and
``advect(dt, Sp, Su)`` once per time step.  All work runs in the HIP library;
there is no CPU fallback: importing the library fails loudly when it has not
been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .mesh import PolyMesh, agglomerated_box, brick_box, hex_box, refined_hex_box, tet_box  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsimplic_amd.so")
HEADERS = [os.path.join(os.path.dirname(_HERE), "include", h)
           for h in ("simplic_amd.h", "simplic_harness.h", "simplic_learned.h")]

SVP_HOST = 0
SVP_DEVICE = 1

_lib = None


class SvpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class SvpDict(C.Structure):
    _fields_ = [
        ("orientation_method", C.c_char * 32),
        ("mixed_cell_tol", C.c_double),
        ("n_alpha_bounds", C.c_int32),
        ("snap_tol", C.c_double),
        ("clip", C.c_int32),
        ("split_warped_face", C.c_int32),
        ("write_plic_fields", C.c_int32),
        ("learned_residual_tol", C.c_double),
        ("rdf_tol", C.c_double),
        ("rdf_rel_tol", C.c_double),
        ("rdf_iterations", C.c_int32),
        ("grad_scheme", C.c_char * 32),
    ]


class SvpMesh(C.Structure):
    _fields_ = [
        ("n_points", C.c_int32), ("n_faces", C.c_int32), ("n_internal_faces", C.c_int32),
        ("n_cells", C.c_int32), ("n_patches", C.c_int32),
        ("points", C.c_void_p), ("face_offsets", C.c_void_p), ("face_points", C.c_void_p),
        ("owner", C.c_void_p), ("neighbour", C.c_void_p), ("patch_start", C.c_void_p),
        ("patch_size", C.c_void_p), ("patch_type", C.c_void_p),
    ]


class SvpDeviceView(C.Structure):
    _fields_ = [
        ("alpha", C.c_void_p), ("phi", C.c_void_p), ("U", C.c_void_p), ("U_boundary", C.c_void_p),
        ("dVf", C.c_void_p), ("Cf", C.c_void_p), ("Sf", C.c_void_p), ("C", C.c_void_p),
        ("V", C.c_void_p), ("stream", C.c_void_p),
    ]


def lib():
    """Load libsimplic_amd.so (raises if it is missing: no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() or "
                          "python geometricvofext_amd/build.py")
    L = C.CDLL(LIB_PATH)
    vp, i32, dbl = C.c_void_p, C.c_int32, C.c_double
    L.svp_last_error.restype = C.c_char_p
    L.svp_dict_defaults.argtypes = [C.POINTER(SvpDict)]
    L.svp_dict_defaults.restype = None
    L.svp_dict_parse.argtypes = [C.c_char_p, C.POINTER(SvpDict)]
    L.svp_create.argtypes = [C.POINTER(SvpMesh), C.POINTER(SvpDict), C.c_int, C.POINTER(vp)]
    L.svp_destroy.argtypes = [vp]
    L.svp_set_alpha.argtypes = [vp, vp, C.c_int]
    L.svp_set_flow.argtypes = [vp, vp, vp, vp, C.c_int]
    L.svp_reconstruct.argtypes = [vp]
    L.svp_advect.argtypes = [vp, dbl, vp, vp, C.c_int]
    L.svp_get_alpha.argtypes = [vp, vp, C.c_int]
    L.svp_get_alpha_phi.argtypes = [vp, vp, C.c_int]
    L.svp_get_rho_phi.argtypes = [vp, dbl, dbl, vp, C.c_int]
    L.svp_advect_sub_cycles.argtypes = [vp, dbl, C.c_int32, vp, vp, dbl, dbl, vp, C.c_int]
    L.svp_n_mixed.argtypes = [vp, C.POINTER(i32)]
    L.svp_get_interface.argtypes = [vp, vp, vp, vp, vp, vp, vp, C.c_int]
    L.svp_get_plic_surface.argtypes = [vp, vp, vp, vp, vp, vp, C.c_int]
    L.svp_get_plic_surface_device.argtypes = [vp, vp, vp, vp, vp, vp]
    L.svp_get_plic_fields.argtypes = [vp, vp, vp, vp, vp, vp, C.c_int]
    L.svp_batch_create.argtypes = [vp, i32, C.POINTER(vp)]
    L.svp_batch_destroy.argtypes = [vp]
    L.svp_batch_step.argtypes = [vp, vp]
    L.svp_batch_tune.argtypes = [vp, vp, vp]
    L.svp_batch_set_shapes.argtypes = [vp, vp]
    L.svp_batch_timers.argtypes = [vp, C.POINTER(dbl), C.POINTER(C.c_uint64)]
    L.svp_batch_profile_set.argtypes = [vp, C.c_int]
    L.svp_batch_profile_get.argtypes = [vp, vp, vp, C.c_int]
    L.svp_batch_courant.argtypes = [vp, vp, vp]
    L.svp_batch_scale_flow.argtypes = [vp, vp]
    L.svp_alpha_mapping_time.argtypes = [vp, C.POINTER(dbl)]
    L.svp_rdf_iterations.argtypes = [vp, vp]
    L.svp_map_alpha_field.argtypes = [vp, vp, vp, dbl, dbl, C.c_int]
    L.svp_get_cut_class.argtypes = [vp, vp, vp, C.c_int]
    L.svp_get_un0.argtypes = [vp, vp, C.c_int]
    L.svp_get_geometry.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.svp_timers.argtypes = [vp, C.POINTER(dbl), C.POINTER(dbl), C.POINTER(C.c_uint64)]
    L.svp_device_view_get.argtypes = [vp, C.POINTER(SvpDeviceView)]
    L.svp_synchronize.argtypes = [vp]
    L.svp_test_set_error_bits.argtypes = [vp, i32]
    L.svp_harness_deformation.argtypes = [vp, vp, vp]
    L.svp_harness_rotation.argtypes = [vp, vp, vp, vp, vp]
    L.svp_harness_upload_base_flow.argtypes = [vp, vp, vp, vp]
    L.svp_harness_scale_flow.argtypes = [vp, dbl]
    L.svp_harness_courant.argtypes = [vp, C.POINTER(dbl), C.POINTER(dbl)]
    L.svp_harness_init_alpha.argtypes = [vp, vp, C.c_int, C.c_int]
    L.svp_harness_init_alpha_surface.argtypes = [vp, vp, i32, i32]
    L.svp_profile_set.argtypes = [vp, C.c_int]
    L.svp_profile_get.argtypes = [vp, vp, vp, C.c_int]
    i64, u64 = C.c_int64, C.c_uint64
    L.svp_nm_create.argtypes = [C.c_int, u64, C.POINTER(vp)]
    L.svp_nm_destroy.argtypes = [vp]
    L.svp_nm_set_params.argtypes = [vp, vp, i32]
    L.svp_nm_get_params.argtypes = [vp, vp, i32]
    L.svp_nm_generate.argtypes = [vp, i64, u64, i32]
    L.svp_nm_dataset_get.argtypes = [vp, i64, i64, vp, vp, vp, vp]
    L.svp_nm_train.argtypes = [vp, i32, i32, C.c_float, C.c_float, u64, vp, C.POINTER(dbl)]
    L.svp_nm_forward.argtypes = [vp, vp, i64, vp, C.c_int]
    L.svp_nm_gradient.argtypes = [vp, vp, vp, i64, vp, C.POINTER(dbl), C.c_int]
    L.svp_nm_cube_fractions.argtypes = [vp, i64, vp, C.c_int]
    L.svp_bound_stats.argtypes = [vp, vp, C.c_int]
    L.svp_bound_totals.argtypes = [vp, vp]
    L.svp_set_normal_model.argtypes = [vp, vp]
    L.svp_learned_stats.argtypes = [vp, vp]
    L.svp_learned_residuals.argtypes = [vp, vp, C.c_int]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise SvpError(rc, lib().svp_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data


def dict_defaults() -> SvpDict:
    d = SvpDict()
    lib().svp_dict_defaults(C.byref(d))
    return d


def parse_dict(text: str) -> SvpDict:
    """Parse the body of an fvSolution ``"alpha.*"`` entry."""
    d = SvpDict()
    _check(lib().svp_dict_parse(text.encode(), C.byref(d)))
    return d


def make_dict(**kw) -> SvpDict:
    d = dict_defaults()
    names = {"orientationMethod": "orientation_method", "mixedCellTol": "mixed_cell_tol",
             "nAlphaBounds": "n_alpha_bounds", "snapTol": "snap_tol", "clip": "clip",
             "splitWarpedFace": "split_warped_face", "writePlicFields": "write_plic_fields",
             "learnedResidualTol": "learned_residual_tol", "tol": "rdf_tol", "relTol": "rdf_rel_tol",
             "iterations": "rdf_iterations", "gradScheme": "grad_scheme", "grad(alpha1)": "grad_scheme"}
    for k, v in kw.items():
        k = names.get(k, k)
        if k in ("orientation_method", "grad_scheme"):
            v = v.encode()
        elif isinstance(v, bool):
            v = int(v)
        setattr(d, k, v)
    return d


class SolveVofEqu:
    """solveVofEqu(alpha1, phi, U) on a :class:`PolyMesh`, resident on one GPU."""

    def __init__(self, mesh: PolyMesh, dictionary=None, device: int = 0):
        L = lib()
        self.mesh = mesh
        if dictionary is None:
            dictionary = dict_defaults()
        elif isinstance(dictionary, str):
            dictionary = parse_dict(dictionary)
        elif isinstance(dictionary, dict):
            dictionary = make_dict(**dictionary)
        self.dict = dictionary
        self._keep = [
            np.ascontiguousarray(mesh.points, dtype=np.float64),
            np.ascontiguousarray(mesh.face_offsets, dtype=np.int32),
            np.ascontiguousarray(mesh.face_points, dtype=np.int32),
            np.ascontiguousarray(mesh.owner, dtype=np.int32),
            np.ascontiguousarray(mesh.neighbour if mesh.neighbour.size else np.zeros(1), dtype=np.int32),
            np.ascontiguousarray(mesh.patch_start, dtype=np.int32),
            np.ascontiguousarray(mesh.patch_size, dtype=np.int32),
            np.ascontiguousarray(mesh.patch_type, dtype=np.int32),
        ]
        k = self._keep
        sm = SvpMesh(mesh.n_points, mesh.n_faces, mesh.n_internal_faces, mesh.n_cells,
                     len(mesh.patch_start), *[a.ctypes.data for a in k])
        h = C.c_void_p()
        _check(L.svp_create(C.byref(sm), C.byref(self.dict), device, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.svp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # fields ---------------------------------------------------------------
    def set_alpha(self, alpha):
        a = np.ascontiguousarray(alpha, dtype=np.float64)
        assert a.size == self.mesh.n_cells
        _check(lib().svp_set_alpha(self.h, a.ctypes.data, SVP_HOST))

    def set_flow(self, phi, U, U_boundary=None):
        p = np.ascontiguousarray(phi, dtype=np.float64)
        u = np.ascontiguousarray(U, dtype=np.float64).reshape(-1)
        ub = None if U_boundary is None else np.ascontiguousarray(U_boundary, dtype=np.float64).reshape(-1)
        assert p.size == self.mesh.n_faces and u.size == 3 * self.mesh.n_cells
        _check(lib().svp_set_flow(self.h, p.ctypes.data, u.ctypes.data, _ptr(ub), SVP_HOST))

    # the reference's per-step calls ----------------------------------------
    def reconstruct(self):
        _check(lib().svp_reconstruct(self.h))

    def advect(self, dt, Sp=None, Su=None):
        sp = None if Sp is None else np.ascontiguousarray(Sp, dtype=np.float64)
        su = None if Su is None else np.ascontiguousarray(Su, dtype=np.float64)
        _check(lib().svp_advect(self.h, float(dt), _ptr(sp), _ptr(su), SVP_HOST))

    def advect_sub_cycles(self, dt, n_sub, Sp=None, Su=None, rho1=1.0, rho2=0.0):
        """The sub-cycled alpha step of interPlicFoam (alphaEqnSubCycle.H:30-60):
        n_sub reconstruct + advect steps of dt/n_sub; returns rhoPhi summed
        with weights (dt/n_sub)/dt (n_faces)."""
        sp = None if Sp is None else np.ascontiguousarray(Sp, dtype=np.float64)
        su = None if Su is None else np.ascontiguousarray(Su, dtype=np.float64)
        a = np.zeros(self.mesh.n_faces)
        _check(lib().svp_advect_sub_cycles(self.h, float(dt), int(n_sub), _ptr(sp), _ptr(su), float(rho1),
                                           float(rho2), a.ctypes.data, SVP_HOST))
        return a

    # outputs ----------------------------------------------------------------
    def alpha(self):
        a = np.zeros(self.mesh.n_cells)
        _check(lib().svp_get_alpha(self.h, a.ctypes.data, SVP_HOST))
        return a

    def alpha_phi(self):
        a = np.zeros(self.mesh.n_faces)
        _check(lib().svp_get_alpha_phi(self.h, a.ctypes.data, SVP_HOST))
        return a

    def rho_phi(self, rho1, rho2):
        a = np.zeros(self.mesh.n_faces)
        _check(lib().svp_get_rho_phi(self.h, float(rho1), float(rho2), a.ctypes.data, SVP_HOST))
        return a

    def map_alpha_field(self, n, d, lower_refine_level, upper_refine_level):
        """mapAlphaField (reconstruction.C:725-784): cells with alpha in the
        refine levels take the volume fraction of the plane (n, d) given per
        cell (n: (n_cells, 3), d: (n_cells,))."""
        n = np.ascontiguousarray(np.asarray(n, dtype=np.float64).reshape(-1))
        d = np.ascontiguousarray(d, dtype=np.float64)
        if n.size != 3 * self.mesh.n_cells or d.size != self.mesh.n_cells:
            raise ValueError("n must be (n_cells, 3) and d (n_cells,)")
        _check(lib().svp_map_alpha_field(self.h, n.ctypes.data, d.ctypes.data, float(lower_refine_level),
                                         float(upper_refine_level), SVP_HOST))

    def rdf_iterations(self) -> int:
        """Iterations of the last isoRDF / RDF reconstruct."""
        n = C.c_int32()
        _check(lib().svp_rdf_iterations(self.h, C.byref(n)))
        return n.value

    def n_mixed(self) -> int:
        n = C.c_int32()
        _check(lib().svp_n_mixed(self.h, C.byref(n)))
        return n.value

    def plic_surface(self):
        """PLIC polygons of the last reconstruct (reconstruction::interface,
        reconstruction.C:787-835): dict(points (n, 3), offsets (n_faces + 1),
        cells (n_faces)); polygon f is points[offsets[f]:offsets[f + 1]]."""
        nf, npt = C.c_int32(), C.c_int32()
        _check(lib().svp_get_plic_surface(self.h, C.addressof(nf), C.addressof(npt), None, None, None, SVP_HOST))
        cells = np.zeros(max(nf.value, 1), np.int32)
        offs = np.zeros(nf.value + 1, np.int32)
        pts = np.zeros(3 * max(npt.value, 1))
        _check(lib().svp_get_plic_surface(self.h, C.addressof(nf), C.addressof(npt), cells.ctypes.data,
                                          offs.ctypes.data, pts.ctypes.data, SVP_HOST))
        return dict(points=pts.reshape(-1, 3)[:npt.value], offsets=offs, cells=cells[:nf.value])

    def plic_fields(self):
        """The reference's full-mesh interface fields after the last
        reconstruct and the face flatness (reconstruction.C:448-471, 520-584):
        dict(interfaceN (n_cells, 3), interfaceD, interfaceC, interfaceS,
        zeta (n_faces)); zero outside the mixed cells."""
        nc, nf = self.mesh.n_cells, self.mesh.n_faces
        N, D, Cc, S, z = np.zeros(3 * nc), np.zeros(nc), np.zeros(3 * nc), np.zeros(3 * nc), np.zeros(nf)
        _check(lib().svp_get_plic_fields(self.h, N.ctypes.data, D.ctypes.data, Cc.ctypes.data, S.ctypes.data,
                                         z.ctypes.data, SVP_HOST))
        return dict(interfaceN=N.reshape(nc, 3), interfaceD=D, interfaceC=Cc.reshape(nc, 3),
                    interfaceS=S.reshape(nc, 3), zeta=z)

    def alpha_mapping_time(self):
        """GPU seconds spent in map_alpha_field (alphaMappingTime)."""
        t = C.c_double()
        _check(lib().svp_alpha_mapping_time(self.h, C.byref(t)))
        return t.value

    def interface(self):
        n = self.n_mixed()
        k = max(n, 1)
        cells = np.zeros(k, np.int32)
        status = np.zeros(k, np.int32)
        nrm = np.zeros(3 * k)
        d = np.zeros(k)
        Cc = np.zeros(3 * k)
        S = np.zeros(3 * k)
        _check(lib().svp_get_interface(self.h, cells.ctypes.data, status.ctypes.data, nrm.ctypes.data,
                                       d.ctypes.data, Cc.ctypes.data, S.ctypes.data, SVP_HOST))
        fcl = np.zeros(k, np.uint64)
        ncp = np.zeros(k, np.int32)
        _check(lib().svp_get_cut_class(self.h, fcl.ctypes.data, ncp.ctypes.data, SVP_HOST))
        return dict(cells=cells[:n], status=status[:n], n=nrm.reshape(-1, 3)[:n], d=d[:n],
                    C=Cc.reshape(-1, 3)[:n], S=S.reshape(-1, 3)[:n], faceclass=fcl[:n], ncutpts=ncp[:n])

    def un0(self):
        n = self.n_mixed()
        u = np.zeros(max(n, 1))
        _check(lib().svp_get_un0(self.h, u.ctypes.data, SVP_HOST))
        return u[:n]

    def geometry(self):
        nf, nc = self.mesh.n_faces, self.mesh.n_cells
        g = dict(Cf=np.zeros((nf, 3)), Sf=np.zeros((nf, 3)), magSf=np.zeros(nf), flatness=np.zeros(nf),
                 C=np.zeros((nc, 3)), V=np.zeros(nc))
        _check(lib().svp_get_geometry(self.h, g["Cf"].ctypes.data, g["Sf"].ctypes.data, g["magSf"].ctypes.data,
                                      g["flatness"].ctypes.data, g["C"].ctypes.data, g["V"].ctypes.data))
        return g

    def timers(self):
        r, a, n = C.c_double(), C.c_double(), C.c_uint64()
        _check(lib().svp_timers(self.h, C.byref(r), C.byref(a), C.byref(n)))
        return r.value, a.value, n.value

    def synchronize(self):
        _check(lib().svp_synchronize(self.h))

    def set_error_bits(self, bits):
        """Test hook: set device error bits as a failing kernel would."""
        _check(lib().svp_test_set_error_bits(self.h, int(bits)))

    def device_view(self) -> SvpDeviceView:
        v = SvpDeviceView()
        _check(lib().svp_device_view_get(self.h, C.byref(v)))
        return v

    # harness ---------------------------------------------------------------
    def deformation_field(self):
        nc, nf = self.mesh.n_cells, self.mesh.n_faces
        U = np.zeros(3 * nc)
        phi = np.zeros(nf)
        _check(lib().svp_harness_deformation(self.h, U.ctypes.data, phi.ctypes.data))
        return U.reshape(nc, 3), phi

    def rotation_field(self, omega, origin):
        nc, nf = self.mesh.n_cells, self.mesh.n_faces
        U = np.zeros(3 * nc)
        phi = np.zeros(nf)
        w = np.ascontiguousarray(omega, dtype=np.float64)
        o = np.ascontiguousarray(origin, dtype=np.float64)
        _check(lib().svp_harness_rotation(self.h, w.ctypes.data, o.ctypes.data, U.ctypes.data, phi.ctypes.data))
        return U.reshape(nc, 3), phi

    def upload_base_flow(self, U0, phi0, U0_boundary=None):
        u = np.ascontiguousarray(U0, dtype=np.float64).reshape(-1)
        p = np.ascontiguousarray(phi0, dtype=np.float64)
        ub = None if U0_boundary is None else np.ascontiguousarray(U0_boundary, dtype=np.float64).reshape(-1)
        _check(lib().svp_harness_upload_base_flow(self.h, u.ctypes.data, p.ctypes.data, _ptr(ub)))

    def scale_flow(self, factor):
        _check(lib().svp_harness_scale_flow(self.h, float(factor)))

    def courant(self):
        a, b = C.c_double(), C.c_double()
        _check(lib().svp_harness_courant(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def init_alpha(self, shapes, max_depth=4):
        """alpha of a union of shapes; each shape is (centre, semi_axes, R) or
        (centre, radius) for a sphere."""
        rows = []
        for s in shapes:
            if len(s) == 2:
                c, r = s
                rows.append(np.concatenate([np.asarray(c, float), [r, r, r], np.eye(3).ravel()]))
            else:
                c, a, R = s
                rows.append(np.concatenate([np.asarray(c, float), np.asarray(a, float),
                                            np.asarray(R, float).reshape(9)]))
        arr = np.ascontiguousarray(np.stack(rows), dtype=np.float64)
        _check(lib().svp_harness_init_alpha(self.h, arr.ctypes.data, len(rows), int(max_depth)))

    def init_alpha_surface(self, tris, mode="set"):
        """alpha from a closed outward-oriented triangulated surface ((n, 3, 3)
        points, e.g. surface.read_stl): each cell's exact inside fraction
        set, added or subtracted (the setVofField utility's modes), then
        clipped to [0, 1]."""
        t = np.ascontiguousarray(tris, dtype=np.float64).reshape(-1, 9)
        m = {"set": 0, "add": 1, "subtract": 2}[mode]
        _check(lib().svp_harness_init_alpha_surface(self.h, t.ctypes.data, len(t), m))

    STAGES = ("detect", "normal", "findSignedDistance", "upwind", "interfaceFlux", "update",
              "bounding", "snapClip")

    def profile_set(self, mask):
        _check(lib().svp_profile_set(self.h, int(mask)))

    def profile_get(self, reset=False):
        sec = np.zeros(len(self.STAGES))
        cnt = np.zeros(len(self.STAGES), np.int64)
        _check(lib().svp_profile_get(self.h, sec.ctypes.data, cnt.ctypes.data, int(reset)))
        return {n: (s, int(c)) for n, s, c in zip(self.STAGES, sec, cnt)}

    def bound_stats(self):
        """Per bounding pass of the last advect (and since the last call):
        rows of (candidates, active, rounds, corrected faces); resets."""
        out = np.zeros(4 * 64, np.int32)
        _check(lib().svp_bound_stats(self.h, out.ctypes.data, out.size))
        rows = out.reshape(-1, 4)
        return rows[rows[:, 0] > 0]

    def bound_totals(self):
        """Running sums over the bounding passes with candidates: candidates,
        active, rounds, corrected faces, passes."""
        out = np.zeros(5, np.uint64)
        _check(lib().svp_bound_totals(self.h, out.ctypes.data))
        return [int(v) for v in out]


    # learned orientation method (include/simplic_learned.h)
    def set_normal_model(self, model):
        _check(lib().svp_set_normal_model(self.h, model.h))

    def learned_stats(self):
        """(interface cells, learned kept, residual fallbacks, least-squares only)."""
        out = np.zeros(4, np.int64)
        _check(lib().svp_learned_stats(self.h, out.ctypes.data))
        return tuple(int(v) for v in out)

    def learned_residuals(self):
        out = np.zeros(self.n_mixed())
        _check(lib().svp_learned_residuals(self.h, out.ctypes.data, SVP_HOST))
        return out


from .learned import NormalModel, cube_fractions  # noqa: E402,F401


class CaseBatch:
    """Independent :class:`SolveVofEqu` handles advanced together
    (``svp_batch_*``): one launch per stage for all cases, each case with
    its own time step; every case's results equal its handle stepping alone.
    Close the batch before using the handles on their own again."""

    SHAPE_NAMES = ("normal", "findSignedDistance", "fluxList", "interfaceFlux", "boundRecords", "boundFlow",
                   "boundApply")

    def __init__(self, solvers):
        self.solvers = list(solvers)
        arr = (C.c_void_p * len(self.solvers))(*[sv.h for sv in self.solvers])
        self.h = C.c_void_p()
        _check(lib().svp_batch_create(arr, len(self.solvers), C.byref(self.h)))

    def close(self):
        if getattr(self, "h", None):
            _check(lib().svp_batch_destroy(self.h))
            self.h = None

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.svp_batch_destroy(self.h)
            self.h = None

    def _dt(self, dt):
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        assert dt.shape == (len(self.solvers),)
        return dt

    def step(self, dt):
        """One step of every case with dt[k] > 0 (dt[k] <= 0: case left as is)."""
        d = self._dt(dt)
        _check(lib().svp_batch_step(self.h, d.ctypes.data))

    def tune(self, dt):
        d = self._dt(dt)
        out = np.zeros(len(self.SHAPE_NAMES), np.int32)
        _check(lib().svp_batch_tune(self.h, d.ctypes.data, out.ctypes.data))
        return dict(zip(self.SHAPE_NAMES, out.tolist()))

    def set_shapes(self, shapes):
        a = np.ascontiguousarray([shapes[k] for k in self.SHAPE_NAMES], dtype=np.int32)
        _check(lib().svp_batch_set_shapes(self.h, a.ctypes.data))

    def timers(self):
        t, n = C.c_double(), C.c_uint64()
        _check(lib().svp_batch_timers(self.h, C.byref(t), C.byref(n)))
        return t.value, n.value

    def courant(self):
        n = len(self.solvers)
        co, aco = np.zeros(n), np.zeros(n)
        _check(lib().svp_batch_courant(self.h, co.ctypes.data, aco.ctypes.data))
        return co, aco

    def scale_flow(self, factors):
        f = np.ascontiguousarray(factors, dtype=np.float64)
        _check(lib().svp_batch_scale_flow(self.h, f.ctypes.data))

    def profile_set(self, mask):
        _check(lib().svp_batch_profile_set(self.h, int(mask)))

    def profile_get(self, reset=False):
        sec = np.zeros(8)
        cnt = np.zeros(8, np.int64)
        _check(lib().svp_batch_profile_get(self.h, sec.ctypes.data, cnt.ctypes.data, int(reset)))
        return {name: (float(sec[k]), int(cnt[k])) for k, name in enumerate(SolveVofEqu.STAGES)}
