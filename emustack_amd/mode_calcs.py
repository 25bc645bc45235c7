"""Plane-wave and Bloch-mode bases of the layers of a stack, and the
scattering matrices of the interfaces between them.

Public surface, kept because Stack and user scripts address it by name
(reference: backend/mode_calcs.py):
  Modes    k_pll_norm, wl_norm, air_ref, calc_1d_grating_orders,
           calc_2d_grating_orders, prop_fwd, shear_transform  (:36-96)
  Anallo   a ThinFilm under one Light: k_z of the retained grating orders,
           wave impedances Z, specular incidence vector  (:99-250)
  Simmo    a NanoStruct under one Light: Bloch modes (k_z, J, J_dag) from
           the GPU library through the EMUstack seam  (:253-414)
  r_t_mat, r_t_mat_anallo, r_t_mat_tf_ns   interface R12, T12, R21, T21
           (:416-515)

Organisation of this module: the plane-wave bookkeeping (which orders are
kept, how they are sorted, their wavevectors) is done by free functions
that the films, the FEM layers, the batched solver and the GPU cascade all
use, so the orders can never disagree between them.  A Simmo only prepares
the FEM request (material indices, shift, mode count) and hands it to the
seam; the request is also what the batched path (batch.py) submits.
"""
import collections

import numpy as np

from . import EMUstack
from . import paths

TWO_PI = 2.0 * np.pi

# The FEM call's fixed settings (the reference passes the same values):
# E-field formulation, periodic boundaries, 30 Krylov restarts at most.
E_FIELD_FORMULATION = 1
PERIODIC_BOUNDARIES = 2
MAX_RESTARTS = 30

FemRequest = collections.namedtuple('FemRequest', 'n_effs num_pw_per_pol num_BMs shift')


# ---------------------------------------------------------------- orders


def grating_orders(max_order, period, period_y, world_1d):
    """Integer diffraction orders (px, py) of the retained plane waves.

    1D world: px = -p..p, py = 0.  2D world: px outer, py inner (the loop
    order of the reference's overlap routines), kept when
    (px/period)^2 + (py/period_y)^2 <= (p/max(period, period_y))^2, which
    gives 5, 13, 29, 49, 81, 113 orders for p = 1..6 on a square lattice
    (mode_calcs.py:55-74)."""
    p = np.arange(-max_order, max_order + 1)
    if world_1d:
        return p, np.zeros_like(p)
    px = np.repeat(p, p.size)
    py = np.tile(p, p.size)
    inside = (px / period) ** 2 + (py / period_y) ** 2 <= (max_order / max(period, period_y)) ** 2
    return px[inside], py[inside]


def order_wavevectors(k_par_norm, px, py, period_ratio_y, world_1d):
    """In-plane wavevectors (alpha, beta) of the orders, in units of
    1/period: the incident k_parallel plus the reciprocal lattice vectors.
    In the 1D world every order keeps the incident beta."""
    alpha = k_par_norm[0] + TWO_PI * px
    if world_1d:
        beta = np.full(alpha.shape, float(k_par_norm[1]))
    else:
        beta = k_par_norm[1] + TWO_PI * py / period_ratio_y
    return alpha, beta


def propagation_order(kz):
    """Sort key of the reference's air reference (mode_calcs.py:174-176):
    propagating orders first by decreasing k_z, then evanescent ones by
    increasing decay rate."""
    return np.argsort(-kz.real + kz.imag)


# ------------------------------------------------------------ base class

# Interface matrices already computed, by (id(layer1), id(layer2)).  A
# layer object drops its entries when it is garbage collected, so that a
# recycled id can never return another layer's matrices.
_known_interfaces = {}


class Modes(object):
    """What Anallo and Simmo share: the normalised incidence parameters, the
    air reference of the light and the propagation/shear operators."""

    def k_pll_norm(self):
        """Incident k_parallel times the period (period-normalised units)."""
        return self.light.k_pll * self.structure.period

    def wl_norm(self):
        """Wavelength over period, moved by 1e-10 when the wavelength is an
        exact multiple of the period (a Wood anomaly of the normal-incidence
        orders)."""
        ratio = self.light.wl_nm / self.structure.period
        on_anomaly = self.light.wl_nm % self.structure.period == 0
        return ratio + 1e-10 if on_anomaly else ratio

    def air_ref(self):
        """The Anallo of free space for this light and period, cached on the
        light; it fixes the plane-wave order of every layer."""
        st = self.structure
        return self.light._air_ref(st.period, st.period_y, st.world_1d)

    def calc_1d_grating_orders(self, max_order):
        return grating_orders(max_order, 1.0, 1.0, True)[0]

    def calc_2d_grating_orders(self, max_order):
        return grating_orders(max_order, self.structure.period, self.structure.period_y, False)

    def prop_fwd(self, height_norm):
        """Diagonal propagation matrix exp(i k_z h) over a height in units
        of the period."""
        return np.asmatrix(np.diag(np.exp(1j * self.k_z * height_norm)))

    def shear_transform(self, coords):
        """Diagonal phase matrix of a lateral shift by `coords` (one or two
        components) in the plane-wave basis of the air reference."""
        air = self.air_ref()
        phase = np.append(air.alphas, air.alphas) * coords[0]
        if np.shape(coords) != (1,):
            phase = phase + np.append(air.betas, air.betas) * coords[1]
        return np.asmatrix(np.diag(np.exp(1j * phase)))

    def __del__(self):
        table = _known_interfaces
        if table:
            me = id(self)
            for key in [k for k in list(table) if me in k]:
                table.pop(key, None)


# ------------------------------------------------------------- thin films


class Anallo(Modes):
    """Analytic plane-wave modes of a homogeneous film (ThinFilm) for one
    Light (reference mode_calcs.py:99-250)."""

    def __init__(self, thin_film, light):
        self.structure = thin_film
        self.light = light
        self.max_order_PWs = light.max_order_PWs
        self.is_air_ref = False

    def calc_modes(self):
        """k_z of every retained order, once per polarisation (TE block, then
        TM block)."""
        kz = self.calc_kz()
        self.k_z = np.concatenate((kz, kz))
        self.structure.num_pw_per_pol = kz.size

    def calc_kz(self):
        """k_z of the retained orders in the air reference's order; also sets
        alphas, betas (unsorted), sort_order, specular_order and
        num_prop_pw_per_pol.  An order with k_z exactly 0 (grazing) moves
        k_parallel by a factor 1 - 1e-9."""
        px, py, alpha, beta = self.wavevectors()
        kz = np.sqrt(self.k() ** 2 - alpha ** 2 - beta ** 2)
        if np.count_nonzero(kz) < kz.size:
            px, py, alpha, beta = self.wavevectors(grazing=True)
            kz = np.sqrt(self.k() ** 2 - alpha ** 2 - beta ** 2)
        return self.take_kz(kz, px, py, alpha, beta)

    def wavevectors(self, grazing=False):
        """Orders (px, py) and their in-plane wavevectors (alpha, beta) in
        units of 1/period; grazing=True scales k_parallel by 1 - 1e-9 (the
        reference's step off a Wood anomaly, with its warning)."""
        st = self.structure
        if st.world_1d is not True and st.world_1d is not False:
            raise ValueError("ThinFilm.world_1d must be set to True or False.")
        # the circle test in nm, as the reference evaluates it (the rounding
        # of boundary orders depends on the units)
        px, py = grating_orders(self.max_order_PWs, st.period, st.period_y, st.world_1d)
        k_par = np.asarray(self.k_pll_norm(), float)
        if grazing:
            print("Warning: [k_x, k_y] falls on a Wood anomaly of order (px, py); "
                  "k_parallel is scaled by (1 - 1e-9).")
            k_par = (1 - 1e-9) * k_par
        alpha, beta = order_wavevectors(k_par, px, py, st.period_y / st.period, st.world_1d)
        return px, py, alpha, beta

    def take_kz(self, kz, px, py, alpha, beta, presorted=False):
        """Record the orders' k_z (unsorted, or already in the air reference's
        order when presorted) with their wavevectors; returns k_z sorted."""
        self.alphas, self.betas = alpha, beta
        if self.is_air_ref:
            if hasattr(self, 'sort_order'):
                raise AssertionError("the air reference's plane-wave order is already fixed")
            order = propagation_order(kz)
        else:
            order = self.air_ref().sort_order
            if order.shape != kz.shape:
                raise AssertionError("film and air reference disagree on the number of orders")
        self.sort_order = order
        zeroth = (px[order] == 0) & (py[order] == 0)
        self.specular_order = int(np.flatnonzero(zeroth)[0])
        self.num_prop_pw_per_pol = int(np.count_nonzero(kz.imag == 0))
        return kz if presorted else kz[order]

    def n(self):
        """Refractive index at the light's wavelength (real part only for a
        lossless film)."""
        n = self.structure.material.n(self.light.wl_nm)
        return n if self.structure.loss else n.real

    def k(self):
        """Wavenumber in the film, in units of 1/period."""
        return np.complex128(TWO_PI * self.n() / self.wl_norm())

    def Z(self):
        """Wave impedances relative to free space, TE block (E in the
        interface plane: Z = (1/n) k/k_z) then TM block (Z = (1/n) k_z/k).
        Films evaluated on the GPU (batch.calc_films_many) return the
        impedances that launch computed."""
        if getattr(self, '_device_Z', None) is not None:
            return self._device_Z.copy()
        half = self.k_z.size // 2
        kz = self.k_z[:half]
        if not np.array_equal(kz, self.k_z[half:]):
            raise AssertionError("the two polarisation blocks of k_z differ")
        ratio = self.k() / kz
        inv_n = 1.0 / self.n()
        return np.concatenate((inv_n * ratio, inv_n / ratio))

    _POLARISATIONS = {
        'TE': (1.0, 0.0),
        'TM': (0.0, 1.0),
        'un': (2 ** -0.5, 2 ** -0.5),
        'R Circ': (2 ** -0.5, 1j * 2 ** -0.5),
        'L Circ': (2 ** -0.5, -1j * 2 ** -0.5),
    }

    def specular_incidence(self, pol='TE'):
        """Column vector of plane-wave amplitudes with only the zeroth order
        excited, in the polarisation `pol`."""
        if pol not in self._POLARISATIONS:
            raise NotImplementedError("Unknown polarisation %r; choose TE, TM, un, R Circ or L Circ." % (pol,))
        te, tm = self._POLARISATIONS[pol]
        npw = self.structure.num_pw_per_pol
        amp = np.zeros((2 * npw, 1), np.complex128)
        amp[self.specular_order, 0] = te
        amp[self.specular_order + npw, 0] = tm
        return np.asmatrix(amp)


# ------------------------------------------------------ structured layers


class Simmo(Modes):
    """Bloch modes of a NanoStruct for one Light, solved on the GPU
    (reference mode_calcs.py:253-414)."""

    def __init__(self, structure, light):
        self.structure = structure
        self.light = light
        self.max_order_PWs = light.max_order_PWs
        self.prop_consts = None
        self.mode_pol = None

    def fem_inputs(self, num_BMs=None):
        """The FEM request of this light: material indices by element type,
        plane waves per polarisation, number of Bloch modes (default
        2 x plane waves + 20) and the shift of the shift-and-invert solve,
        1.1 max(Re n)^2 k0^2 - |k_parallel|^2 (the |k_parallel|^2 term is
        dropped for hyperbolic layers); the real parts of the shift's terms
        put complex-conjugate eigenvalue pairs equally far from it."""
        st = self.structure
        if st.periodicity not in ('1D_array', '2D_array'):
            raise ValueError("NanoStruct.periodicity must be '1D_array' or '2D_array', not %r" % (st.periodicity,))
        # background, then inclusions a..f by element type (the reference stops
        # at inclusion_e, mode_calcs.py:273-276, which leaves the sixth
        # inclusion of a 1D grating, type 7, without a material)
        materials = (st.background, st.inclusion_a, st.inclusion_b, st.inclusion_c, st.inclusion_d,
                     st.inclusion_e, st.inclusion_f)
        if st.nb_typ_el > len(materials):
            raise ValueError("nb_typ_el = %d, but a NanoStruct has %d materials" % (st.nb_typ_el, len(materials)))
        n_effs = np.array([m.n(self.light.wl_nm) for m in materials[:st.nb_typ_el]])
        if st.loss is False:
            n_effs = n_effs.real
        npw = grating_orders(self.max_order_PWs, st.period, st.period_y, st.periodicity == '1D_array')[0].size
        nbm = 2 * npw + 20 if num_BMs is None else int(num_BMs)
        if nbm <= 2 * npw:
            raise AssertionError("num_BMs = %d must exceed the %d plane waves (both polarisations)" % (nbm, 2 * npw))
        k0 = TWO_PI * self.air_ref().n() / self.wl_norm()
        shift = 1.1 * np.real(n_effs).max() ** 2 * k0 ** 2
        if not st.hyperbolic == True:  # noqa: E712  (1 and True both mean hyperbolic)
            kx, ky = self.k_pll_norm()
            shift = shift - kx ** 2 - ky ** 2
        return FemRequest(n_effs, npw, nbm, shift)

    def calc_modes(self, num_BMs=None):
        """Solve for the Bloch modes: sets k_z, J, J_dag (and the field
        arrays when the structure asks for plotting)."""
        req = self.fem_inputs(num_BMs)
        self.num_BMs = req.num_BMs
        self.n_effs = req.n_effs
        self.E_H_field = E_FIELD_FORMULATION
        solver = self._solve_1d if self.structure.periodicity == '1D_array' else self._solve_2d
        solver(req)
        st = self.structure
        if not st.plot_field_conc:
            self.mode_pol = None
        if st.plotting_fields != 1:
            self.sol1 = None
            self.n_effs = None
            self.E_H_field = None
            if st.periodicity == '2D_array':
                self.table_nod = self.type_el = self.x_arr = None
                self.n_msh_pts = self.n_msh_el = None

    def _solve_2d(self, req):
        st = self.structure
        with open(paths.msh_path + st.mesh_file) as f:
            self.n_msh_pts, self.n_msh_el = (int(v) for v in f.readline().split()[:2])
        # the reference returns mode_pol from every call; here the field
        # arrays are produced on request, so field-concentration plots ask
        out = EMUstack.calc_modes_2d(
            self.wl_norm(), req.num_BMs, self.max_order_PWs, 0, paths.msh_path, st.mesh_file, self.n_msh_pts,
            self.n_msh_el, st.nb_typ_el, req.n_effs, self.k_pll_norm(), req.shift, E_FIELD_FORMULATION,
            PERIODIC_BOUNDARIES, MAX_RESTARTS, 1 if (st.plotting_fields == 1 or st.plot_field_conc) else 0,
            st.plot_real, st.plot_imag, st.plot_abs, req.num_pw_per_pol, 0, 0, 0)
        (self.k_z, J, J_dag, self.sol1, self.mode_pol, self.table_nod, self.type_el, self.x_arr) = out
        # J is an overlap per unit cell area; the reference divides by the
        # area in units of period^2 (1 for square cells)
        self.J = np.asmatrix(J) / (st.period_y / st.period)
        self.J_dag = np.asmatrix(J_dag)

    def _solve_1d(self, req):
        st = self.structure
        world_1d = st.world_1d is True
        npw_2d = 1 if world_1d else grating_orders(self.max_order_PWs, st.period, st.period_y, False)[0].size
        kx, ky = self.k_pll_norm()
        out = EMUstack.calc_modes_1d(
            self.wl_norm(), req.num_BMs, self.max_order_PWs, st.nb_typ_el, st.n_msh_pts, st.n_msh_el,
            st.table_nod, st.type_el, st.x_arr, MAX_RESTARTS, 0, st.mesh_file, req.n_effs, kx, ky, req.shift,
            st.plotting_fields, st.plot_real, st.plot_imag, st.plot_abs, req.num_pw_per_pol, npw_2d,
            int(world_1d))
        self.k_z, J, J_dag, J_2d, J_dag_2d, self.sol1 = out
        pick = (J, J_dag) if world_1d else (J_2d, J_dag_2d)
        self.J, self.J_dag = np.asmatrix(pick[0]), np.asmatrix(pick[1])


# ---------------------------------------------------- interface matrices


def r_t_mat(lay1, lay2):
    """R12, T12, R21, T21 of the interface from lay1 to lay2 (1 = the side
    the light comes from).  Results are remembered per pair of layer
    objects; the reversed pair reuses them with the roles swapped."""
    if lay1.structure.period != lay2.structure.period:
        raise AssertionError("interface between layers of different periods")
    hit = _known_interfaces.get((id(lay1), id(lay2)))
    if hit is not None:
        return hit
    hit = _known_interfaces.get((id(lay2), id(lay1)))
    if hit is not None:
        return hit[2], hit[3], hit[0], hit[1]
    film1, film2 = isinstance(lay1, Anallo), isinstance(lay2, Anallo)
    if film1 and film2:
        mats = r_t_mat_anallo(lay1, lay2)
    elif film1:
        mats = r_t_mat_tf_ns(lay1, lay2)
    elif film2:
        back = r_t_mat_tf_ns(lay2, lay1)
        mats = back[2], back[3], back[0], back[1]
    else:
        raise NotImplementedError("Two NanoStruct layers cannot touch: put a (possibly very thin) "
                                  "ThinFilm between them.")
    _known_interfaces[id(lay1), id(lay2)] = mats
    return mats


def r_t_mat_anallo(an1, an2):
    """Diagonal Fresnel matrices between two films from their impedances:
    R12 = (Z2 - Z1)/(Z2 + Z1), T12 = 2 sqrt(Z1) sqrt(Z2)/(Z1 + Z2) (fields
    normalised to unit flux, same square-root branch as the impedances),
    R21 = -R12, T21 = T12."""
    if an1.k_z.size != an2.k_z.size:
        raise ValueError("films %r and %r have different plane-wave counts" % (an1, an2))
    z1, z2 = an1.Z(), an2.Z()
    total = z1 + z2
    r = np.asmatrix(np.diag((z2 - z1) / total))
    t = np.asmatrix(np.diag(2.0 * np.sqrt(z2) * np.sqrt(z1) / total))
    return r, t, -r, t


def r_t_mat_tf_ns(an1, sim2):
    """Film -> structured layer (Dossou et al., JOSA A 29, 817 (2012), with
    the film's impedances in place of the free-space X so that the film need
    not be air).  With A = Z1^-1/2 J and B = J_dag Z1^-1/2 and
    D = I + B A:  R12 = 2 A D^-1 B - I,  T12 = 2 D^-1 B,
    R21 = D^-1 (I - B A),  T21 = 2 A D^-1."""
    w = np.sqrt(1.0 / an1.Z())
    J, Jd = np.asarray(sim2.J), np.asarray(sim2.J_dag)
    if w.size != J.shape[0]:
        raise ValueError("plane-wave bases differ between a film (%d) and a structured layer (%d): "
                         "mixing 1D and 2D worlds? check world_1d" % (w.size, J.shape[0]))
    A = w[:, None] * J
    B = Jd * w[None, :]
    D = np.eye(B.shape[0]) + B @ A
    DinvB = np.linalg.solve(D, B)
    R12 = 2.0 * (A @ DinvB) - np.eye(A.shape[0])
    T12 = 2.0 * DinvB
    R21 = np.linalg.solve(D, np.eye(B.shape[0]) - B @ A)
    T21 = 2.0 * A @ np.linalg.inv(D)
    return np.asmatrix(R12), np.asmatrix(T12), np.asmatrix(R21), np.asmatrix(T21)
