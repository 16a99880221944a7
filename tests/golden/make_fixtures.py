"""Write the small fixtures under tests/golden/ from the reference's own test data.

* sphere_192_o03.npz -- srcvals(12, 1920) of geometries/sphere_192_o03.go3 (a cube
  refined twice by midpoint subdivision and projected to the unit sphere, order-3 RV
  nodes); pins the RV node table and its ordering.
* triaints_golden.json -- the Mathematica-exact integrals of 1/r against the first
  three Koornwinder polynomials hard-coded in test/tria_routs/test_triaintrouts.f:160-170
  (geometry and targets of :116-151), with the reference's tolerances.
Run once in the build container:  python tests/golden/make_fixtures.py /root/reference
"""
import json
import sys
from pathlib import Path

import numpy as np

ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
out = Path(__file__).resolve().parent

with open(ref / "geometries" / "sphere_192_o03.go3") as fh:
    norder = int(fh.readline())
    npat = int(fh.readline())
    vals = np.array([float(x) for x in fh.read().split()])
srcvals = vals.reshape(12, -1).T.copy()
np.savez_compressed(out / "sphere_192_o03.npz", norder=norder, npatches=npat, srcvals=srcvals)

rr = 1.0e-2
golden = {
    "source": "test/tria_routs/test_triaintrouts.f:116-170",
    "norder": 8, "nporder": 10, "rr": rr, "thet": "pi/4",
    "patch2_shift": [2.1, 1.1, 0.0],
    "targets": [[0.0, 0.1 * rr, 0.4 * rr], [0.019 * rr, -0.019 * rr, 0.0], [2.817, 1.0, 0.0]],
    "target_patch": [1, 1, 2],
    "cintex": [[1.156374590854795 * rr, 0.0942702604583580 * rr, -0.239019699324178 * rr],
               [1.7158925715165438 * rr, 0.588248265036294074 * rr, -0.957021021711130743 * rr],
               [2.341262814070374, -0.05030309850820814, 0.8574902331922906]],
    "tol_adaptive": 1.0e-6, "nqorder_adaptive": 16, "eps_adaptive": 1.0e-6,
    "tol_wnodes": 1.0e-3, "nlev_wnodes": 3, "nqorder_wnodes": 20,
}
(out / "triaints_golden.json").write_text(json.dumps(golden, indent=1))
print("wrote", out / "sphere_192_o03.npz", out / "triaints_golden.json")
