"""Writes tests/golden/exact_solutions_64.npz from the reference's exact
solutions (tutorials/test/exactSolutions/<t>/alpha.water.exact, 64^3, FP64,
renumberMesh cell order), mapped back to natural blockMesh cell order.

Run in a container that has the reference mounted:
    python tests/golden/make_exact_fixtures.py /root/reference
The fixture holds data only (times and alpha arrays); nothing here is needed
at test time.
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", ".."))
from geometricvofext_amd import foamio  # noqa: E402

TIMES = ["0", "0.25", "0.5", "0.75", "1", "1.25", "1.5", "1.75", "2", "2.25", "2.5", "2.75", "3"]


def main(ref_root):
    base = os.path.join(ref_root, "tutorials", "test", "exactSolutions")
    order = foamio.renumber_order(64)
    fields = []
    for t in TIMES:
        a = foamio.read_internal_field(os.path.join(base, t, "alpha.water.exact"))
        assert a.size == 64 ** 3
        fields.append(foamio.to_natural(a, 64, order))
    out = os.path.join(os.path.dirname(__file__), "exact_solutions_64.npz")
    np.savez_compressed(out, times=np.array([float(t) for t in TIMES]), alpha=np.stack(fields))
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
