"""Write tests/golden/materials.json: refractive indices of the reference's
tabulated materials at the wavelengths its test cases use.

The reference interpolates its data tables linearly (scipy interp1d,
backend/materials.py:161).  Only the interpolated values at the fixture
wavelengths are committed, not the tables.  Run from the repo root in a
container that has the reference mounted:

    python tests/golden/make_material_fixtures.py /root/reference
"""
import json
import os
import sys

import numpy as np

# material name -> data file (backend/materials.py:225-282)
TABLES = {'Si_c': 'Si_c.txt', 'SiO2': 'SiO2.txt', 'Ag': 'Ag_JC.txt', 'Au': 'Au_JC.txt'}
# wavelengths (nm) used by the reference tests (tests/test_case_*.py)
WLS = {
    'Si_c': [400.0, 500.0, 700.0, 1000.0],
    'SiO2': [500.0, 1000.0],
    'Ag': [400.0, 700.0, 1000.0],
    'Au': [1.11 * 940],
}


def main(ref_root):
    out = {}
    for name, fn in TABLES.items():
        d = np.loadtxt(os.path.join(ref_root, 'backend', 'data', fn))
        wl = d[:, 0]
        n = d[:, 1] + 1j * d[:, 2] if d.shape[1] == 3 else d[:, 1]
        vals = {}
        for w in WLS[name]:
            v = complex(np.interp(w, wl, n.real) + 1j * np.interp(w, wl, np.imag(n)))
            vals[repr(float(w))] = [v.real, v.imag]
        out[name] = vals
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, 'materials.json'), 'w') as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
