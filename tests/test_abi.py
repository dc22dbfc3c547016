"""The C-ABI library loads and exports every function include/*.h declares;
host-side dictionary handling matches the reference's keys, defaults and
error behaviour.  No GPU work (CPU only)."""
import ctypes as C
import glob
import os
import re

import pytest

import geometricvofext_amd as g

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(svp_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_all_declared_symbols():
    names = declared_functions()
    assert len(names) >= 25
    L = g.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_dict_defaults():
    d = g.dict_defaults()
    assert d.orientation_method == b"isoAlphaGrad"      # reconstruction.C:505-508
    assert d.mixed_cell_tol == 1e-8                    # reconstruction.C:502
    assert d.n_alpha_bounds == 10 and d.snap_tol == 0.0 and d.clip == 1   # advection.C:455-457
    assert d.split_warped_face == 0 and d.write_plic_fields == 0


def test_dict_parse_reference_case():
    """The alpha.* entry of tutorials/test/plicVofAdvectionFoam/system/fvSolution."""
    text = """
        nAlphaBounds        3;
        snapTol             0;
        clip                false;

        mixedCellTol        1e-8;
        orientationMethod   LS;     // or isoAlphaGrad
        splitWarpedFace     false;
        writePlicFields     true;

        nAlphaSubCycles     1;
        cAlpha              1;

        period              6.0;
        reverseTime         0.0;
    """
    d = g.parse_dict(text)
    assert d.orientation_method == b"LS" and d.n_alpha_bounds == 3 and d.clip == 0
    assert d.snap_tol == 0.0 and d.write_plic_fields == 1 and d.mixed_cell_tol == 1e-8


def test_dict_reconstruction_scheme_alias_and_errors():
    assert g.parse_dict("reconstructionScheme isoAlphaGrad;").orientation_method == b"isoAlphaGrad"
    # orientationMethod wins over the alias
    assert g.parse_dict("reconstructionScheme LS; orientationMethod isoAlphaGrad;").orientation_method == \
        b"isoAlphaGrad"
    with pytest.raises(g.SvpError) as e:
        g.parse_dict("orientationMethod nonsense;")
    assert e.value.code == -2                             # FatalError in the reference
    assert g.parse_dict("orientationMethod alphaGrad;").orientation_method == b"alphaGrad"
    for m in ("isoRDF", "RDF"):
        d = g.parse_dict(f"orientationMethod {m}; tol 1e-7; relTol 0.05; iterations 8;")
        assert d.orientation_method == m.encode()
        assert d.rdf_tol == 1e-7 and d.rdf_rel_tol == 0.05 and d.rdf_iterations == 8
    d = g.parse_dict("orientationMethod RDF;")
    assert d.rdf_tol == 1e-6 and d.rdf_rel_tol == 0.1 and d.rdf_iterations == 5   # reconstruction.C:513-515
    with pytest.raises(g.SvpError):
        g.parse_dict("clip maybe;")


def test_create_rejects_bad_mesh_before_device_work():
    L = g.lib()
    h = C.c_void_p()
    assert L.svp_create(None, None, 0, C.byref(h)) == -1
