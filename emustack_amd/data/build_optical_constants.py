"""Pack the reference's tabulated optical constants into
optical_constants.npz (next to this file), one pair of arrays per material:
<name>__wl (wavelengths in nm) and <name>__n (complex index), sorted by
wavelength with a stable sort (several tables are stored descending or hold
repeated wavelengths; scipy's interp1d, which the reference interpolates
with, sorts them the same way before use).

The tables are measured data (sources in materials.py); the names are the
attribute names of the reference's backend/materials.py:193-330, so that
`materials.Si_c` and the like resolve as they do there.  Run from a checkout
that has the reference next to it:

    python emustack_amd/data/build_optical_constants.py /root/reference/backend/data
"""
import os
import sys

import numpy as np

# attribute name -> table file (reference backend/data/)
TABLES = {
    'H2O': 'H2O', 'TiO2': 'TiO2', 'TiO2_anatase': 'TiO2_anatase', 'ITO': 'ITO', 'ZnO': 'ZnO',
    'Si_c': 'Si_c', 'Si_a': 'Si_a', 'SiO2': 'SiO2', 'CuO': 'CuO', 'CdTe': 'CdTe', 'FeS2': 'FeS2',
    'Zn3P2': 'Zn3P2', 'Sb2S3': 'Sb2S3', 'AlGaAs': 'AlGaAs', 'ZnS': 'ZnS', 'SnO2': 'SnO2',
    'Glass': 'Soda_lime_glass_nk_Pil', 'Al2O3': 'Al2O3', 'Al2O3_PV': 'Al2O3_PV', 'GaAs': 'GaAs',
    'InGaAs': 'InGaAs', 'Si3N4': 'Si3N4', 'MgF2': 'MgF2', 'InP': 'InP', 'InAs': 'InAs', 'GaP': 'GaP',
    'GaN': 'GaN', 'AlN': 'AlN', 'Ge': 'Ge', 'MoO3': 'MoO3', 'Spiro': 'Spiro', 'Spiro_nk': 'Spiro_nk_Filipic',
    'FTO_Wenger': 'FTO_Wenger', 'FTO_Wengerk5': 'FTO_Wengerk5', 'AlN_PV': 'AlN_PV',
    'Au': 'Au_JC', 'Au_Palik': 'Au_Palik', 'Au_CRC': 'Au_CRC', 'Ag': 'Ag_JC', 'Ag_Palik': 'Ag_Palik',
    'Cu': 'Cu_JC', 'Cu_Palik': 'Cu_Palik', 'Al': 'Al', 'Al_Palik': 'Al_Palik', 'Al_CRC': 'Al_CRC',
    'CH3NH3PbI3': 'CH3NH3PbI3', 'Sb2S3_ANU2014': 'Sb2S3_ANU2014', 'Sb2S3_ANU2015': 'Sb2S3_ANU2015',
    'GO_2014': 'GO_2014', 'GO_2015': 'GO_2015', 'rGO_2015': 'rGO_2015', 'SiON_Low': 'SiON_Low',
    'SiON_High': 'SiON_High', 'Low_Fe_Glass': 'Low_Fe_Glass_Pil', 'Perovskite_00': 'Perovskite_E_u_00',
    'Perovskite': 'Perovskite_Loper_E_u_080', 'Perovskite_b2b': 'Perovskite_b2b_nk', 'Ge_Doped': 'Ge_Doped',
    'ITO_annealed': 'ITO_anneal_Gen_Osc',
}


def main(src):
    out = {}
    for name, fname in TABLES.items():
        t = np.loadtxt(os.path.join(src, fname + '.txt'))
        n = t[:, 1] + 1j * t[:, 2] if t.shape[1] == 3 else t[:, 1].astype(complex)
        order = np.argsort(t[:, 0], kind='mergesort')
        out[name + '__wl'] = np.ascontiguousarray(t[order, 0], np.float64)
        out[name + '__n'] = np.ascontiguousarray(n[order], np.complex128)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'optical_constants.npz')
    np.savez_compressed(dst, **out)
    print(dst, len(TABLES), 'tables', os.path.getsize(dst), 'bytes')


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference/backend/data')
