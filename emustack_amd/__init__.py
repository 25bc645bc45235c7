"""MI355X-native implementation of EMUstack's per-wavelength FEM +
scattering-matrix path (drop-in for NanoStruct / ThinFilm / Light /
Simmo.calc_modes / Stack.calc_scat; the FEM runs in libemu_mi355x.so)."""
from . import materials, paths  # noqa: F401
from .objects import NanoStruct, ThinFilm, Light  # noqa: F401
from .mode_calcs import Simmo, Anallo, r_t_mat  # noqa: F401
from .stack import Stack  # noqa: F401
