"""Layer modes and interface scattering matrices (mirror of backend/mode_calcs.py).

Anallo  analytic plane-wave modes of a ThinFilm (mode_calcs.py:99-250)
Simmo   FEM Bloch modes of a NanoStruct through EMUstack.calc_modes_2d
        (mode_calcs.py:253-410); the FEM call goes to the GPU library
r_t_mat / r_t_mat_anallo / r_t_mat_tf_ns   interface R/T matrices
        (mode_calcs.py:416-515)

Behaviour (plane-wave order, Wood-anomaly nudge, shift choice, area
normalisation of J, memoisation of interfaces) follows the reference.
"""
import numpy as np

from . import EMUstack
from . import paths

_interfaces_i_have_known = {}
pi = np.pi


class Modes(object):
    """Common parts of Simmo and Anallo (mode_calcs.py:36-96)."""

    def k_pll_norm(self):
        return self.light.k_pll * self.structure.period

    def wl_norm(self):
        wl = self.light.wl_nm / self.structure.period
        if self.light.wl_nm % self.structure.period == 0:  # avoid Wood anomalies
            wl += 1e-10
        return wl

    def air_ref(self):
        return self.light._air_ref(self.structure.period, self.structure.period_y, self.structure.world_1d)

    def calc_1d_grating_orders(self, max_order):
        return np.arange(-max_order, max_order + 1)

    def calc_2d_grating_orders(self, max_order):
        """Orders inside the circle (px/dx)^2 + (py/dy)^2 <= (p/max(dx,dy))^2,
        px outer loop, py inner loop (mode_calcs.py:61-74)."""
        pxs = np.arange(-max_order, max_order + 1)
        py_m, px_m = np.meshgrid(pxs, pxs)
        dx, dy = self.structure.period, self.structure.period_y
        keep = (px_m / dx) ** 2 + (py_m / dy) ** 2 <= (max_order / max(dx, dy)) ** 2
        return px_m[keep], py_m[keep]

    def prop_fwd(self, height_norm):
        return np.asmatrix(np.diag(np.exp(1j * self.k_z * height_norm)))

    def shear_transform(self, coords):
        alphas = np.append(self.air_ref().alphas, self.air_ref().alphas)
        if np.shape(coords) == (1,):
            return np.asmatrix(np.diag(np.exp(1j * (alphas * coords[0]))))
        betas = np.append(self.air_ref().betas, self.air_ref().betas)
        return np.asmatrix(np.diag(np.exp(1j * (alphas * coords[0] + betas * coords[1]))))

    def __del__(self):
        if _interfaces_i_have_known is not None:
            for key in list(_interfaces_i_have_known.keys()):
                if id(self) in key:
                    _interfaces_i_have_known.pop(key, None)


class Anallo(Modes):
    """Plane-wave modes of a homogeneous film."""

    def __init__(self, thin_film, light):
        self.structure = thin_film
        self.light = light
        self.max_order_PWs = light.max_order_PWs
        self.is_air_ref = False

    def calc_modes(self):
        kzs = self.calc_kz()
        self.k_z = np.append(kzs, kzs)
        self.structure.num_pw_per_pol = len(kzs)

    def _orders(self):
        d = 1
        dy = self.structure.period_y / self.structure.period
        a0, b0 = self.k_pll_norm()
        if self.structure.world_1d is True:
            pxs = self.calc_1d_grating_orders(self.max_order_PWs)
            pys = np.zeros_like(pxs)
            return pxs, pys, a0 + pxs * 2 * pi / d, b0
        if self.structure.world_1d is False:
            pxs, pys = self.calc_2d_grating_orders(self.max_order_PWs)
            return pxs, pys, a0 + pxs * 2 * pi / d, b0 + pys * 2 * pi / dy
        raise ValueError("must specify world_1d status of ThinFilm.")

    def calc_kz(self):
        """Sorted k_z of the grating orders (mode_calcs.py:116-189)."""
        pxs, pys, alphas, betas = self._orders()
        k_z_unsrt = np.sqrt(self.k() ** 2 - alphas ** 2 - betas ** 2)
        if np.count_nonzero(k_z_unsrt) != k_z_unsrt.size:
            # Wood anomaly: nudge k_parallel by (1 - 1e-9)
            d = 1
            dy = self.structure.period_y / self.structure.period
            a0, b0 = (1 - 1e-9) * self.k_pll_norm()
            alphas = a0 + pxs * 2 * pi / d
            betas = b0 if self.structure.world_1d is True else b0 + pys * 2 * pi / dy
            k_z_unsrt = np.sqrt(self.k() ** 2 - alphas ** 2 - betas ** 2)
        self.alphas, self.betas = alphas, betas
        if self.is_air_ref:
            assert not hasattr(self, 'sort_order'), "Are you sure you want to reset the sort_order?"
            s = np.argsort(-1 * k_z_unsrt.real + k_z_unsrt.imag)
        else:
            s = self.air_ref().sort_order
            assert s.shape == k_z_unsrt.shape, (s.shape, k_z_unsrt.shape)
        self.sort_order = s
        if self.structure.world_1d is True:
            self.specular_order = np.nonzero(pxs[s] == 0)[0][0]
        else:
            self.specular_order = np.nonzero((pxs[s] == 0) * (pys[s] == 0))[0][0]
        self.num_prop_pw_per_pol = (k_z_unsrt.imag == 0).sum()
        return k_z_unsrt[s]

    def n(self):
        if self.structure.loss:
            return self.structure.material.n(self.light.wl_nm)
        return self.structure.material.n(self.light.wl_nm).real

    def k(self):
        return np.complex128(2 * pi * self.n() / self.wl_norm())

    def Z(self):
        """Relative wave impedances, TE block then TM block (mode_calcs.py:198-218)."""
        Zcr = 1. / self.n()
        half = int(len(self.k_z) / 2)
        k_z = self.k_z[:half]
        assert (k_z == self.k_z[half:]).all()
        k_on_kz = self.k() / k_z
        return np.concatenate((Zcr * k_on_kz, Zcr / k_on_kz))

    def specular_incidence(self, pol='TE'):
        """Unit-amplitude zeroth-order incidence vector (mode_calcs.py:220-250)."""
        te = self.specular_order
        tm = self.specular_order + self.structure.num_pw_per_pol
        inc = np.asmatrix(np.zeros(self.structure.num_pw_per_pol * 2, dtype='complex128')).T
        r2 = 1 / np.sqrt(2.)
        amps = {'TE': ((te, 1),), 'TM': ((tm, 1),), 'un': ((te, r2), (tm, r2)),
                'R Circ': ((te, r2), (tm, 1j * r2)), 'L Circ': ((te, r2), (tm, -1j * r2))}
        if pol not in amps:
            raise NotImplementedError("Must select from the currently implemented polarisations; "
                                      "TE, TM, R Circ, L Circ.")
        for idx, val in amps[pol]:
            inc[idx] = val
        return inc


class Simmo(Modes):
    """FEM Bloch modes of a NanoStruct for one Light."""

    def __init__(self, structure, light):
        self.structure = structure
        self.light = light
        self.max_order_PWs = light.max_order_PWs
        self.prop_consts = None
        self.mode_pol = None

    def fem_inputs(self, num_BMs=None):
        """Everything calc_modes passes to the FEM (mode_calcs.py:268-312)."""
        st = self.structure
        wl = self.light.wl_nm
        mats = [st.background, st.inclusion_a, st.inclusion_b, st.inclusion_c, st.inclusion_d,
                st.inclusion_e][:st.nb_typ_el]
        n_effs = np.array([m.n(wl) for m in mats])
        if st.loss is False:
            n_effs = n_effs.real
        if st.periodicity == '1D_array':
            pxs = self.calc_1d_grating_orders(self.max_order_PWs)
        elif st.periodicity == '2D_array':
            pxs, pys = self.calc_2d_grating_orders(self.max_order_PWs)
        else:
            raise ValueError("NanoStruct layer must have periodicity of either '1D_array' or '2D_array'.")
        num_pw_per_pol = pxs.size
        nbm = num_pw_per_pol * 2 + 20 if num_BMs is None else num_BMs
        assert nbm > num_pw_per_pol * 2, (
            "You must include at least as many BMs as PWs. \n"
            "Currently you have %(bm)i BMs < %(np)i PWs." % {'bm': nbm, 'np': num_pw_per_pol * 2})
        max_n = np.real(n_effs).max()
        k_0 = 2 * pi * self.air_ref().n() / self.wl_norm()
        if st.hyperbolic is True:
            shift = 1.1 * max_n ** 2 * k_0 ** 2
        else:
            shift = 1.1 * max_n ** 2 * k_0 ** 2 - self.k_pll_norm()[0] ** 2 - self.k_pll_norm()[1] ** 2
        return n_effs, num_pw_per_pol, nbm, shift

    def calc_modes(self, num_BMs=None):
        """Solve for the Bloch modes (mode_calcs.py:268-410)."""
        n_effs, num_pw_per_pol, self.num_BMs, shift = self.fem_inputs(num_BMs)
        self.n_effs = n_effs
        self.E_H_field = 1
        i_cond = 2
        itermax = 30
        FEM_debug = 0
        st = self.structure
        if st.periodicity == '1D_array':
            if st.world_1d is True:
                world_1d, num_pw_per_pol_2d = 1, 1
            else:
                world_1d = 0
                num_pw_per_pol_2d = self.calc_2d_grating_orders(self.max_order_PWs)[0].size
            resm = EMUstack.calc_modes_1d(
                self.wl_norm(), self.num_BMs, self.max_order_PWs, st.nb_typ_el, st.n_msh_pts, st.n_msh_el,
                st.table_nod, st.type_el, st.x_arr, itermax, FEM_debug, st.mesh_file, self.n_effs,
                self.k_pll_norm()[0], self.k_pll_norm()[1], shift, st.plotting_fields, st.plot_real,
                st.plot_imag, st.plot_abs, num_pw_per_pol, num_pw_per_pol_2d, world_1d)
            self.k_z, J, J_dag, J_2d, J_dag_2d, self.sol1 = resm
            if st.world_1d is True:
                self.J, self.J_dag = np.asmatrix(J), np.asmatrix(J_dag)
            else:
                self.J, self.J_dag = np.asmatrix(J_2d), np.asmatrix(J_dag_2d)
            if st.plotting_fields != 1:
                self.sol1 = None
                self.n_effs = None
                self.E_H_field = None
            return
        mesh_path = paths.msh_path + st.mesh_file
        with open(mesh_path) as f:
            self.n_msh_pts, self.n_msh_el = [int(i) for i in f.readline().split()]
        cmplx_max, real_max, int_max = 2 ** 27, 2 ** 23, 2 ** 22
        resm = EMUstack.calc_modes_2d(
            self.wl_norm(), self.num_BMs, self.max_order_PWs, FEM_debug, paths.msh_path, st.mesh_file,
            self.n_msh_pts, self.n_msh_el, st.nb_typ_el, self.n_effs, self.k_pll_norm(), shift,
            self.E_H_field, i_cond, itermax, st.plotting_fields, st.plot_real, st.plot_imag, st.plot_abs,
            num_pw_per_pol, cmplx_max, real_max, int_max)
        self.k_z, J, J_dag, self.sol1, self.mode_pol, self.table_nod, self.type_el, self.x_arr = resm
        area_norm = st.period * st.period_y / st.period ** 2
        self.J, self.J_dag = np.asmatrix(J) / area_norm, np.asmatrix(J_dag)
        if not st.plot_field_conc:
            self.mode_pol = None
        if st.plotting_fields != 1:
            self.sol1 = None
            self.n_effs = None
            self.E_H_field = None
            self.table_nod = None
            self.type_el = None
            self.x_arr = None
            self.n_msh_pts = None
            self.n_msh_el = None


def r_t_mat(lay1, lay2):
    """R12, T12, R21, T21 at the lay1 -> lay2 interface, memoised by object id
    in both orientations (mode_calcs.py:416-448)."""
    assert lay1.structure.period == lay2.structure.period
    key = (id(lay1), id(lay2))
    if key in _interfaces_i_have_known:
        return _interfaces_i_have_known[key]
    rkey = (id(lay2), id(lay1))
    if rkey in _interfaces_i_have_known:
        R21, T21, R12, T12 = _interfaces_i_have_known[rkey]
        return R12, T12, R21, T21
    if isinstance(lay1, Anallo) and isinstance(lay2, Anallo):
        out = r_t_mat_anallo(lay1, lay2)
    elif isinstance(lay1, Anallo) and isinstance(lay2, Simmo):
        out = r_t_mat_tf_ns(lay1, lay2)
    elif isinstance(lay1, Simmo) and isinstance(lay2, Anallo):
        R21, T21, R12, T12 = r_t_mat_tf_ns(lay2, lay1)
        out = R12, T12, R21, T21
    else:
        raise NotImplementedError("Sorry! For, now you can put an extremely thin film between your NanoStructs")
    _interfaces_i_have_known[key] = out
    return out


def r_t_mat_anallo(an1, an2):
    """Diagonal interface matrices between two thin films (mode_calcs.py:451-477)."""
    if len(an1.k_z) != len(an2.k_z):
        raise ValueError("Need the same number of plane waves in Anallos %s and %s" % (an1, an2))
    Z1, Z2 = an1.Z(), an2.Z()
    R12 = np.asmatrix(np.diag((Z2 - Z1) / (Z2 + Z1)))
    T12 = np.asmatrix(np.diag(2. * np.sqrt(Z2) * np.sqrt(Z1) / (Z2 + Z1)))
    return R12, T12, -R12, T12


def r_t_mat_tf_ns(an1, sim2):
    """Thin film -> NanoStruct interface from J, J^dagger (mode_calcs.py:480-515;
    Dossou et al., JOSA A 29, 817 (2012))."""
    zs = np.sqrt(1 / an1.Z()).reshape((1, -1))
    if np.shape(zs)[1] != np.shape(sim2.J.A)[0]:
        raise ValueError("Scattering matrices of layers are not consistent,\n"
                         "some layers are 1D and others 2D. Check that world_1d status.")
    A = np.asmatrix(zs.T * sim2.J.A)
    B = np.asmatrix(sim2.J_dag.A * zs)
    D = np.eye(len(B)) + B.dot(A)
    DB = np.linalg.solve(D, B)
    R12 = -np.eye(len(A)) + 2 * A * DB
    T12 = 2 * DB
    R21 = np.linalg.solve(D, (np.eye(len(B)) - B * A))
    T21 = 2 * A * D.I
    return np.asmatrix(R12), np.asmatrix(T12), np.asmatrix(R21), np.asmatrix(T21)
