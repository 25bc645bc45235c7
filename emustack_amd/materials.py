"""Refractive-index models (reference: backend/materials.py:31-330).

A Material holds a scalar index, a table (wavelength_nm, n) or
(wavelength_nm, Re n, Im n) evaluated by linear interpolation between the
neighbouring points (the scheme of scipy's interp1d, which the reference
uses, materials.py:161: out-of-range wavelengths raise), or Drude
parameters (omega_p, omega_gamma, eps_inf) in rad/s giving
n = sqrt(eps_inf - omega_p^2 / (omega^2 + i omega_gamma omega)) with
omega = 2 pi c / lambda (materials.py:135-145).

The named materials of the reference (Si_c, SiO2, Au, Ag, ITO, ...) are
module attributes, read on first use from data/optical_constants.npz (the
reference's measured tables, packed by data/build_optical_constants.py).

n_batch() evaluates many (material, wavelength) pairs in one GPU launch, for
wavelength sweeps that feed the batched mode solver.
"""
import os
import threading

import numpy as np

__all__ = ['Material', 'Air', 'Au_drude', 'n_batch', 'named_materials']

_C = 299792458.0
_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data', 'optical_constants.npz')

CONSTANT, TABLE, DRUDE = 0, 1, 2


def _interp_linear(xs, ys, x):
    """Linear interpolation with interp1d's choice of segment: the first
    table point at or above x is the upper end (ties at a repeated
    wavelength take the left pair), clamped to the first segment."""
    x = np.asarray(x, float)
    if np.any(x < xs[0]) or np.any(x > xs[-1]):
        raise ValueError("wavelength outside the tabulated range %g-%g nm" % (xs[0], xs[-1]))
    hi = np.clip(np.searchsorted(xs, x, side='left'), 1, xs.size - 1)
    lo = hi - 1
    slope = (ys[hi] - ys[lo]) / (xs[hi] - xs[lo])
    return slope * (x - xs[lo]) + ys[lo]


class Material(object):
    """A (possibly dispersive, possibly lossy) refractive index."""

    def __init__(self, n):
        if np.shape(n) == ():
            self.kind = CONSTANT
            self.data_wls = None
            self.data_ns = n
        elif np.shape(n) == (3,):
            self.kind = DRUDE
            self.data_wls = 'Drude'
            self.data_ns = [float(n[0]), float(n[1]), float(n[2]), _C]
        else:
            t = np.asarray(n)
            if t.ndim != 2 or t.shape[1] not in (2, 3):
                raise ValueError("tabulated data must be rows of (wavelength, n) or (wavelength, Re n, Im n)")
            order = np.argsort(t[:, 0], kind='mergesort')
            t = t[order]
            self.kind = TABLE
            self.data_wls = t[:, 0].astype(float)
            self.data_ns = t[:, 1] if t.shape[1] == 2 else t[:, 1] + 1j * t[:, 2]

    @classmethod
    def from_file(cls, path):
        """Material from a whitespace table (wavelength_nm, n [, k])."""
        return cls(np.loadtxt(path))

    @classmethod
    def _from_sorted(cls, wl, n):
        m = cls.__new__(cls)
        m.kind = TABLE
        m.data_wls = wl
        m.data_ns = n
        return m

    def n(self, wl_nm):
        """Refractive index at wl_nm (scalar or array, nm)."""
        if self.kind == CONSTANT:
            return self.data_ns
        if self.kind == DRUDE:
            wp, wg, eps_inf, c = self.data_ns
            om = 2 * np.pi * c / (np.asarray(wl_nm, float) * 1e-9)
            return np.sqrt(eps_inf - wp ** 2 / (om ** 2 + 1j * wg * om))
        v = _interp_linear(self.data_wls, self.data_ns, wl_nm)
        return v if np.ndim(v) else v[()]

    def __getstate__(self):
        return dict(self.__dict__)

    def __setstate__(self, d):
        self.__dict__.update(d)


Air = Material(1.00 + 0.0j)
# Drude gold: omega_p = 1.36e16 rad/s, omega_gamma = 1.05e14 rad/s, eps_inf = 9.5
Au_drude = Material([1.36e16, 1.05e14, 9.5])

_named = {}
_named_lock = threading.Lock()
_data = None


def named_materials():
    """Names of the tabulated materials available as module attributes."""
    global _data
    with _named_lock:
        if _data is None:
            _data = np.load(_DATA, allow_pickle=False)
        return sorted(k[:-4] for k in _data.files if k.endswith('__wl'))


def __getattr__(name):
    """Named tabulated materials (materials.Si_c, materials.Au, ...), read
    from the packed tables on first use."""
    if name.startswith('__'):
        raise AttributeError(name)
    global _data
    with _named_lock:
        m = _named.get(name)
        if m is not None:
            return m
        if _data is None:
            _data = np.load(_DATA, allow_pickle=False)
        if name + '__wl' not in _data.files:
            raise AttributeError("module 'materials' has no material %r" % name)
        m = Material._from_sorted(_data[name + '__wl'], _data[name + '__n'])
        _named[name] = m
        return m


def n_batch(materials, wl_nm, ctx=None):
    """Refractive indices of every material in `materials` at every
    wavelength in `wl_nm`, shape (len(materials), len(wl_nm)), evaluated on
    the GPU in one launch (emu_material_index): constant, tabulated (same
    segment choice and formula as Material.n) and Drude materials."""
    from . import _lib, EMUstack
    wl = np.ascontiguousarray(np.asarray(wl_nm, np.float64).reshape(-1))
    kinds, offs, xs, ys = [], [0], [], []
    for m in materials:
        kinds.append(m.kind)
        if m.kind == CONSTANT:
            xs.append(np.zeros(1))
            ys.append(np.array([complex(m.data_ns)]))
        elif m.kind == DRUDE:
            xs.append(np.array(m.data_ns[:3], float))
            ys.append(np.zeros(3, complex))
        else:
            if wl.size and (wl.min() < m.data_wls[0] or wl.max() > m.data_wls[-1]):
                raise ValueError("wavelength outside the tabulated range %g-%g nm"
                                 % (m.data_wls[0], m.data_wls[-1]))
            xs.append(np.asarray(m.data_wls, float))
            ys.append(np.asarray(m.data_ns, complex))
        offs.append(offs[-1] + xs[-1].size)
    nm = len(kinds)
    q_mat = np.repeat(np.arange(nm, dtype=np.int64), wl.size)
    q_wl = np.tile(wl, nm)
    out = (ctx or EMUstack.context()).material_index(np.array(kinds, np.int64), np.array(offs, np.int64),
                                                      np.concatenate(xs) if xs else np.zeros(0),
                                                      np.concatenate(ys) if ys else np.zeros(0, complex),
                                                      q_mat, q_wl)
    return out.reshape(nm, wl.size)
