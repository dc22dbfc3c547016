"""svp_get_plic_surface_device, the PLIC export that allocates nothing, called
through the C-ABI with device output buffers, against svp_get_plic_surface
(host outputs) on the same reconstruct.  Bar: sizes, cells and offsets
equal, points bit-identical.  Two calls in a row, the second on a smaller
interface, then one without any interface and one above the scratch limit."""
import ctypes as C
import sys

import numpy as np
import pytest

from tests.helpers import bits_equal, ramp_alpha

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _hip():
    """The HIP runtime the library is linked against (already loaded in the
    process), for the test's own device buffers."""
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class _DevBuf:
    """Device buffer filled from a host array (hipMalloc + hipMemcpy H2D)."""

    def __init__(self, hip, host):
        self.hip, self.host = hip, np.ascontiguousarray(host)
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(self.host.nbytes, 8)) == 0
        assert hip.hipMemcpy(self.p, self.host.ctypes.data, self.host.nbytes, 1) == 0

    def get(self):
        out = np.empty_like(self.host)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _device_export(solver, sentinel=-7):
    import geometricvofext_amd as g
    L = g.lib()
    hip = _hip()
    nf, npt = C.c_int32(-1), C.c_int32(-1)
    rc = L.svp_get_plic_surface_device(solver.h, C.addressof(nf), C.addressof(npt), None, None, None)
    assert rc == 0, L.svp_last_error().decode()
    # outputs prefilled with a sentinel, so every entry read back was written
    cells = _DevBuf(hip, np.full(max(nf.value, 1), sentinel, np.int32))
    offs = _DevBuf(hip, np.full(nf.value + 1, sentinel, np.int32))
    pts = _DevBuf(hip, np.full(3 * max(npt.value, 1), float(sentinel)))
    assert hip.hipDeviceSynchronize() == 0
    nf2, npt2 = C.c_int32(-1), C.c_int32(-1)
    rc = L.svp_get_plic_surface_device(solver.h, C.addressof(nf2), C.addressof(npt2), cells.p, offs.p, pts.p)
    assert rc == 0, L.svp_last_error().decode()
    assert (nf2.value, npt2.value) == (nf.value, npt.value)
    out = dict(points=pts.get().reshape(-1, 3)[:npt.value], offsets=offs.get(), cells=cells.get()[:nf.value])
    for b in (cells, offs, pts):
        b.free()
    return out


def _compare(solver):
    host = solver.plic_surface()
    dev = _device_export(solver)
    assert np.array_equal(dev["cells"], host["cells"])
    assert np.array_equal(dev["offsets"], host["offsets"])
    assert bits_equal(dev["points"], host["points"])
    return len(host["cells"])


def test_plic_export_without_allocation(oracle_lib):
    import geometricvofext_amd as g
    mesh = g.hex_box(20, perturb=0.2, seed=11)
    om = oracle_lib.OracleMesh(mesh)
    s = g.SolveVofEqu(mesh, dict(orientationMethod="LS", nAlphaBounds=3))
    sizes = []
    for r in (0.3, 0.15):   # the second call sees fewer mixed cells
        s.set_alpha(ramp_alpha(om.C, [((0.47, 0.5, 0.52), r)], 1.5 / 20))
        s.reconstruct()
        _say("no-alloc PLIC export, sphere r", r, "mixed cells", s.n_mixed())
        sizes.append(_compare(s))
    assert sizes[0] > sizes[1] > 0
    # no interface at all: no polygon, the one offset 0
    s.set_alpha(np.zeros(mesh.n_cells))
    s.reconstruct()
    _say("no-alloc PLIC export, empty interface")
    dev = _device_export(s)
    assert len(dev["cells"]) == 0 and dev["offsets"].tolist() == [0]


def test_plic_export_without_allocation_rejects_oversize():
    """More mixed cells than the static scratch holds: SVP_E_ARG and a message
    naming the limit; the allocating export still works."""
    import geometricvofext_amd as g
    L = g.lib()
    limit = 1 << 20
    mesh = g.hex_box(102)   # 1,061,208 cells, every one mixed at alpha 0.5
    assert mesh.n_cells > limit
    s = g.SolveVofEqu(mesh, dict(orientationMethod="LS", nAlphaBounds=3))
    s.set_alpha(np.full(mesh.n_cells, 0.5))
    s.reconstruct()
    _say("no-alloc PLIC export above the limit:", s.n_mixed(), "mixed cells")
    assert s.n_mixed() > limit
    nf, npt = C.c_int32(), C.c_int32()
    rc = L.svp_get_plic_surface_device(s.h, C.addressof(nf), C.addressof(npt), None, None, None)
    assert rc == -1   # SVP_E_ARG
    assert str(limit) in L.svp_last_error().decode()
    assert s.plic_surface()["offsets"][0] == 0
