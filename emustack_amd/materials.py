"""Refractive-index models (mirror of backend/materials.py:31-186).

A Material is built from a scalar index, a table (wavelength_nm, n) or
(wavelength_nm, Re n, Im n) interpolated linearly (the reference uses scipy
interp1d, materials.py:161), or Drude parameters (omega_p, omega_g, eps_inf)
in rad/s (materials.py:135-145).  Tabulated data files are not shipped; load
them with Material.from_file().
"""
import numpy as np

__all__ = ['Material', 'Air']

_C = 299792458.0


class Material(object):
    def __init__(self, n):
        if np.shape(n) == ():
            self.data_wls = None
            self.data_ns = n
        elif np.shape(n) == (3,):
            self.data_wls = 'Drude'
            self.data_ns = [n[0], n[1], n[2], _C]
        else:
            n = np.asarray(n)
            self.data_wls = n[:, 0].astype(float)
            if n.shape[1] == 2:
                self.data_ns = n[:, 1]
            elif n.shape[1] == 3:
                self.data_ns = n[:, 1] + 1j * n[:, 2]
            else:
                raise ValueError("tabulated data must have 2 or 3 columns")

    @classmethod
    def from_file(cls, path):
        """Material from a whitespace table (wavelength_nm, n [, k])."""
        return cls(np.loadtxt(path))

    def n(self, wl_nm):
        """Refractive index at wavelength wl_nm (nm)."""
        if self.data_wls is None:
            return self.data_ns
        if isinstance(self.data_wls, str):
            wp, wg, eps_inf, c = self.data_ns
            om = 2 * np.pi * c / (wl_nm * 1e-9)
            val = np.sqrt(eps_inf - wp ** 2 / (om ** 2 + 1j * wg * om))
            return val if np.imag(val) >= 0 else -val
        wl = np.asarray(wl_nm, float)
        if np.any(wl < self.data_wls[0]) or np.any(wl > self.data_wls[-1]):
            raise ValueError("wavelength %s nm outside the tabulated range" % wl_nm)
        re = np.interp(wl, self.data_wls, np.real(self.data_ns))
        im = np.interp(wl, self.data_wls, np.imag(self.data_ns))
        v = re + 1j * im
        return v if np.ndim(v) else complex(v)


Air = Material(1.00 + 0.0j)
