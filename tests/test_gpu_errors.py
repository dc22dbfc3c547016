"""Device error bits reach the caller on the solver path: every call that
hands back a host buffer (alpha, alphaPhi, rhoPhi, the sub-cycled step)
reports SVP_E_DEVICE and clears the bits, and svp_synchronize does for device
outputs.  The bits are set from the host (test hook), so no kernel is made to
fail."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SVP_E_DEVICE = -6
ERR_BOUND_STALL = 16


def test_host_outputs_report_device_errors():
    import geometricvofext_amd as g
    mesh = g.hex_box(8)
    s = g.SolveVofEqu(mesh, dict(orientationMethod="LS", nAlphaBounds=3))
    C = (np.arange(mesh.n_cells) % 8 + 0.5) / 8
    s.set_alpha(np.clip(0.5 - (C - 0.5) * 4, 0.0, 1.0))
    U, phi = s.deformation_field()
    s.set_flow(phi * 0.5, U * 0.5)
    s.reconstruct()
    s.advect(0.002)
    calls = {
        "alpha": lambda: s.alpha(),
        "alpha_phi": lambda: s.alpha_phi(),
        "rho_phi": lambda: s.rho_phi(1000.0, 1.0),
        "advect_sub_cycles": lambda: s.advect_sub_cycles(0.002, 2, rho1=1000.0, rho2=1.0),
        "synchronize": lambda: s.synchronize(),
        "n_mixed": lambda: s.n_mixed(),
    }
    for name, call in calls.items():
        print("error path of", name, flush=True)
        s.set_error_bits(ERR_BOUND_STALL)
        with pytest.raises(g.SvpError) as e:
            call()
        assert e.value.code == SVP_E_DEVICE, name
        assert "0x10" in str(e.value), name
        call()   # the bits were cleared
