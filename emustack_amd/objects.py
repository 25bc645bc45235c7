"""Layer and light objects (mirror of backend/objects.py).

NanoStruct  structured layer solved by the FEM Bloch-mode path (objects.py:42-1612)
ThinFilm    homogeneous layer solved analytically (objects.py:1614-1669)
Light       incident light, wavelength and in-plane wave vector (objects.py:1672-1777)

Constructor arguments, defaults and their meaning follow the reference.  Mesh
generation: the reference builds .geo templates and runs gmsh; here
make_mesh_now=True calls the project mesher (emustack_amd.mesher) for the
'circle' and 'ellipse' inclusion shapes, writing a .mail file with the same
conventions; any other shape needs a premade mesh_file.
"""
import os

import numpy as np

from . import materials
from . import paths
from .mode_calcs import Simmo, Anallo

__all__ = ['NanoStruct', 'ThinFilm', 'Light', 'calculate_ff']


class NanoStruct(object):
    """A periodically structured layer (2D_array or 1D_array periodicity)."""

    def __init__(self, periodicity, period, diameter1, period_y=None, inc_shape='circle', is_hex=False,
                 ellipticity=0.0, ff=0, ff_rand=False, small_space=None, edge_spacing=False,
                 split_touching_incs=False, len_vertical=0, len_horizontal=0,
                 background=materials.Material(1.0 + 0.0j), inclusion_a=materials.Material(1.0 + 0.0j),
                 inclusion_b=materials.Material(1.0 + 0.0j), inclusion_c=materials.Material(1.0 + 0.0j),
                 inclusion_d=materials.Material(1.0 + 0.0j), inclusion_e=materials.Material(1.0 + 0.0j),
                 inclusion_f=materials.Material(1.0 + 0.0j), loss=True, height_nm=100.0, diameter2=0,
                 diameter3=0, diameter4=0, diameter5=0, diameter6=0, diameter7=0, diameter8=0, diameter9=0,
                 diameter10=0, diameter11=0, diameter12=0, diameter13=0, diameter14=0, diameter15=0,
                 diameter16=0, gap=0, smooth=0, t=0, hyperbolic=False, world_1d=None, posx=0, posy=0,
                 xshift=None, make_mesh_now=True, force_mesh=True, mesh_file='NEED_FILE.mail', geo_file=None,
                 geo_params=None, lc_bkg=0.09, lc2=1.0, lc3=1.0, lc4=1.0, lc5=1.0, lc6=1.0, nb_typ_el=None,
                 plotting_fields=False, plot_real=1, plot_imag=0, plot_abs=0, plot_field_conc=False,
                 plt_msh=True, rotation_deg=0.0, mesh_seed=0, inclusion_edge='code'):
        self.periodicity = periodicity
        self.period = float(period)
        self.diameter1 = diameter1
        self.period_y = float(period) if period_y is None else float(period_y)
        self.inc_shape = inc_shape
        self.is_hex = is_hex
        self.height_nm = height_nm
        self.background = background
        self.inclusion_a = inclusion_a
        self.inclusion_b = inclusion_b
        self.inclusion_c = inclusion_c
        self.inclusion_d = inclusion_d
        self.inclusion_e = inclusion_e
        self.inclusion_f = inclusion_f
        self.loss = loss
        self.hyperbolic = hyperbolic
        for i in range(2, 17):
            setattr(self, 'diameter%d' % i, locals()['diameter%d' % i])
        self.geo_file = geo_file
        self.geo_params = geo_params
        self.gap = gap
        self.smooth = smooth
        self.xshift = xshift
        self.t = t
        self.len_vertical = len_vertical
        self.len_horizontal = len_horizontal
        self.ellipticity = ellipticity
        self.rotation_deg = rotation_deg
        self.mesh_seed = mesh_seed
        # 1D gratings: 'code' = the reference's current element rule for the
        # centre inclusion, 'exact' = elements inside [0.5 - r, 0.5 + r] only
        self.inclusion_edge = inclusion_edge
        if ellipticity > 1.0:
            raise ValueError("ellipticity must be less than 1.0")
        if nb_typ_el is None:  # objects.py:296-304
            if diameter3 != 0:
                self.nb_typ_el = 4
            elif diameter2 != 0:
                self.nb_typ_el = 3
            else:
                self.nb_typ_el = 2
        else:
            self.nb_typ_el = nb_typ_el
        if ff == 0:
            if periodicity == '2D_array':
                self.ff = calculate_ff(inc_shape, period, self.period_y, diameter1, diameter2, diameter3,
                                       diameter4, diameter5, diameter6, diameter7, diameter8, diameter9,
                                       diameter10, diameter11, diameter12, diameter13, diameter14, diameter15,
                                       diameter16, ellipticity)
            elif periodicity == '1D_array':
                self.ff = (diameter1 + diameter2) / period
        else:
            self.ff = ff
            if diameter2 != 0:
                self.diameter2 = 2 * ((ff * (period) ** 2) / np.pi - ((diameter1 / 2) ** 2)) ** 0.5
            else:
                self.diameter1 = 2 * np.sqrt((ff * (period) ** 2) / np.pi)
        self.ff_rand = ff_rand
        if world_1d is None:
            self.world_1d = periodicity == '1D_array'
        else:
            self.world_1d = world_1d
        self.posx = posx
        self.posy = posy
        self.lc = lc_bkg
        self.lc2 = lc2
        self.lc3 = lc3
        self.lc4 = lc4
        self.lc5 = lc5
        self.lc6 = lc6
        self.force_mesh = force_mesh
        self.small_space = small_space
        self.edge_spacing = edge_spacing
        self.split_touching_incs = split_touching_incs
        self.plt_msh = plt_msh
        if make_mesh_now is True:
            self.make_mesh()
        else:
            self.mesh_file = mesh_file
        self.plotting_fields = 1 if plotting_fields is True else 0
        self.plot_real = plot_real
        self.plot_imag = plot_imag
        self.plot_abs = plot_abs
        self.plot_field_conc = plot_field_conc

    def make_mesh(self):
        """Generate the mesh: a .mail file from the project mesher for 2D
        arrays (replaces the reference's gmsh templates, objects.py:366-1052),
        the in-memory element arrays for 1D gratings (objects.py:1054-1530)."""
        if self.periodicity == '1D_array':
            from . import mesher
            ds = [self.diameter1] + [getattr(self, 'diameter%d' % i) for i in range(2, 7)]
            x, table_nod, type_el, name = mesher.grating_1d(self.period, ds, self.lc, self.small_space,
                                                            self.split_touching_incs, self.inclusion_edge,
                                                            self.edge_spacing)
            self.n_msh_el = type_el.size
            self.n_msh_pts = x.size
            self.table_nod = table_nod
            self.type_el = type_el
            self.x_arr = x
            self.mesh_file = name
            return
        if self.periodicity != '2D_array':
            raise ValueError("periodicity must be '1D_array' or '2D_array'")
        if self.inc_shape not in ('circle', 'ellipse'):
            raise NotImplementedError("mesher supports 'circle' and 'ellipse' inclusions; "
                                      "pass make_mesh_now=False with a premade mesh_file")
        from . import mesher
        a = self.diameter1 / 2.0
        b = a * np.sqrt(1 - self.ellipticity) if self.ellipticity else a
        name = 'emu_%s_%g_%g_%g_%g_%g_%g_s%d.mail' % (self.inc_shape, self.period, self.period_y, self.diameter1,
                                                      self.ellipticity, self.lc, self.rotation_deg,
                                                      self.mesh_seed)
        path = os.path.join(paths.msh_path or '.', name)
        if self.force_mesh or not os.path.exists(path):
            mesher.write_mail(path, mesher.ellipse_cell(self.period, self.period_y, a, b,
                                                        rotation_deg=self.rotation_deg, lc=self.lc,
                                                        seed=self.mesh_seed))
        self.mesh_file = name

    def calc_modes(self, light, **args):
        """Run the FEM mode solve; returns a Simmo (objects.py:1594-1608)."""
        simmo = Simmo(self, light)
        simmo.calc_modes(**args)
        return simmo


class ThinFilm(object):
    """A homogeneous film (objects.py:1614-1669)."""

    def __init__(self, period, period_y=None, height_nm=1.0, num_pw_per_pol=0, world_1d=False,
                 material=materials.Material(3.0 + 0.001), loss=True):
        self.period = float(period)
        self.period_y = float(period) if period_y is None else float(period_y)
        self.world_1d = world_1d
        self.height_nm = height_nm
        self.num_pw_per_pol = num_pw_per_pol
        self.material = material
        self.loss = loss

    def calc_modes(self, light):
        an = Anallo(self, light)
        an.calc_modes()
        return an


class Light(object):
    """Incident light (objects.py:1672-1777).  Normal incidence is nudged by
    1e-5 degrees in theta and phi as in the reference (objects.py:1736-1737)."""

    def __init__(self, wl_nm, max_order_PWs=2, k_parallel=None, theta=None, phi=None, n_inc=1.):
        if np.imag(wl_nm) != 0:
            self.wl_nm = complex(wl_nm)
        else:
            self.wl_nm = float(np.real(wl_nm))
        self._air_anallos = {}
        self.max_order_PWs = max_order_PWs
        if theta is None and k_parallel is None:
            raise ValueError("Specify incident angle either by k_parallel OR by theta, phi and n_inc.")
        if theta is None:
            self.k_pll = np.array(k_parallel, dtype='float64')
        else:
            if k_parallel is not None or phi is None:
                raise ValueError("Specify incident angle either by k_parallel OR by theta, phi and n_inc.")
            if abs(theta) < 1e-5:
                theta += 1e-5
            if abs(phi) < 1e-5:
                phi += 1e-5
            k = 2 * np.pi * np.real(n_inc) / self.wl_nm
            theta *= np.pi / 180
            phi *= np.pi / 180
            self.k_pll = k * np.sin(theta) * np.array([np.cos(phi), np.sin(phi)], dtype='float64')

    def _air_ref(self, period, period_y, world_1d):
        """Cached air Anallo defining the plane-wave order (objects.py:1745-1777)."""
        if period in self._air_anallos:
            return self._air_anallos[period]
        air = ThinFilm(period=period, period_y=period_y, material=materials.Air, world_1d=world_1d)
        an = Anallo(air, self)
        an.is_air_ref = True
        kz = an.calc_kz()
        an.k_z = np.append(kz, kz)
        self._air_anallos[period] = an
        return an


def calculate_ff(inc_shape, d, dy, a1, a2=0, a3=0, a4=0, a5=0, a6=0, a7=0, a8=0, a9=0, a10=0, a11=0, a12=0,
                 a13=0, a14=0, a15=0, a16=0, el1=0):
    """Fill fraction of the inclusions (objects.py:1788-1838)."""
    if inc_shape in ('circle', 'ellipse'):
        return np.pi * ((a1 / 2) ** 2 * np.sqrt(1 - el1) + sum((a / 2) ** 2 for a in
                        (a2, a3, a4, a5, a6, a7, a8, a9, a10, a11, a12, a13, a14, a15, a16))) / (d * dy)
    if inc_shape == 'square':
        return sum(a ** 2 for a in (a1, a2, a3, a4, a5, a6, a7, a8, a9, a10, a11, a12, a13, a14, a15,
                                    a16)) / (d * dy)
    if inc_shape == 'dimer':
        return np.pi * ((a1 / 2.0) ** 2 + (a2 / 2.0) ** 2) / (d * dy)
    return 0.0
