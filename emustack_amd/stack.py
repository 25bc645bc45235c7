"""Stack of layers: star-product cascade and energy fluxes (mirror of
backend/stack.py:27-435).

Stack.calc_scat builds the net reflection/transmission matrices bottom-up
through infinitesimal air gaps (stack.py:204-278) and then the downward /
upward fluxes layer by layer (stack.py:285-427), giving the per-layer
absorptance, reflectance and transmittance lists a_list, r_list, t_list.
Scattering matrices are organised as
    | TE -> TE | TM -> TE |
    | TE -> TM | TM -> TM |
"""
import numpy as np

from .mode_calcs import r_t_mat, Anallo

__all__ = ['Stack']


class Stack(object):
    """Layers ordered from bottom (substrate) to top (superstrate)."""

    def __init__(self, layers, heights_nm=None, shears=None):
        self.layers = tuple(layers)
        self._heights_nm = heights_nm
        self.shears = shears
        self.period = float(layers[0].structure.period)
        self.period_y = float(layers[0].structure.period_y)
        self._check_periods_are_consistent()
        if self._heights_nm is not None:
            for i_lay, lay in enumerate(self.layers[1:-1]):
                lay.structure.height_nm = heights_nm[i_lay]

    def heights_nm(self):
        if self._heights_nm is not None:
            return self._heights_nm
        return [float(lay.structure.height_nm) for lay in self.layers[1:-1]]

    def heights_norm(self):
        return [h / self.period for h in self.heights_nm()]

    def total_height(self):
        return sum(self.heights_nm())

    def structures(self):
        return (lay.structure for lay in self.layers)

    def _shear_pair(self, lay, coord_diff):
        return lay.shear_transform(coord_diff), lay.shear_transform(-1 * coord_diff)

    def calc_scat(self, pol='TE', incoming_amplitudes=None, calc_fluxes=True, save_scat_list=False):
        """Net R/T of the stack and, if calc_fluxes, per-layer A/R/T."""
        self._check_periods_are_consistent()
        r12, r21, t12, t21, P_list = [], [], [], [], []
        for lay in self.layers:
            R12, T12, R21, T21 = r_t_mat(lay.air_ref(), lay)
            r12.append(R12)
            r21.append(R21)
            t12.append(T12)
            t21.append(T21)
            lay.R12, lay.T12, lay.R21, lay.T21 = R12, T12, R21, T21
        PW_pols = int(np.shape(R12)[0])
        neq_PW = int(PW_pols / 2.0)
        I_air = np.matrix(np.eye(PW_pols), dtype='D')
        last = self.layers[-1]

        tnet, rnet = t12[0], r12[0]
        tnet_list, rnet_list = [tnet], [rnet]
        inv_t21_list, inv_t12_list = [], []

        def air_gap(i, tnet, rnet, coord_diff):
            if self.shears is None:
                it21 = np.linalg.solve(I_air - r12[i] * rnet, t21[i])
                return it21, tnet * it21, r21[i] + t12[i] * rnet * it21
            Q, Qi = self._shear_pair(last, coord_diff)
            it21 = np.linalg.solve(I_air - r12[i] * Qi * rnet * Q, t21[i])
            return it21, tnet * Q * it21, r21[i] + t12[i] * Qi * rnet * Q * it21

        for i in range(1, len(self.layers) - 1):
            lay = self.layers[i]
            cd = None if self.shears is None else np.asarray(self.shears[i - 1]) - np.asarray(self.shears[i])
            it21, tnet, rnet = air_gap(i, tnet, rnet, cd)
            inv_t21_list.append(it21)
            tnet_list.append(tnet)
            rnet_list.append(rnet)
            # propagation through the layer
            P = lay.prop_fwd(self.heights_nm()[i - 1] / self.period)
            I_TF = np.matrix(np.eye(len(P)), dtype='D')
            it12 = np.linalg.solve(I_TF - r21[i] * P * rnet * P, t12[i])
            P_it12 = P * it12
            tnet = tnet * P_it12
            rnet = r12[i] + t21[i] * P * rnet * P_it12
            P_list.append(P)
            inv_t12_list.append(it12)
            tnet_list.append(tnet)
            rnet_list.append(rnet)
        cd = None if self.shears is None else np.asarray(self.shears[-1])
        it21, tnet, rnet = air_gap(len(self.layers) - 1, tnet, rnet, cd)
        inv_t21_list.append(it21)
        tnet_list.append(tnet)
        rnet_list.append(rnet)
        self.R_net, self.T_net = rnet, tnet
        if save_scat_list:
            self.tnet_list, self.rnet_list = tnet_list, rnet_list
        if calc_fluxes:
            self._fluxes(pol, incoming_amplitudes, PW_pols, neq_PW, rnet_list, tnet_list, inv_t21_list,
                         inv_t12_list, P_list)

    def _fluxes(self, pol, incoming_amplitudes, PW_pols, neq_PW, rnet_list, tnet_list, inv_t21_list,
                inv_t12_list, P_list):
        """Net downward flux in every air gap and the semi-infinite media
        (stack.py:285-427; appendix C of Dossou et al., JOSA A 2012)."""
        self.t_list, self.r_list, self.a_list = [], [], []
        top, bottom = self.layers[-1], self.layers[0]
        num_prop_air = top.air_ref().num_prop_pw_per_pol
        num_prop_in = top.num_prop_pw_per_pol
        down, up = [], []
        self.vec_coef_down, self.vec_coef_up = [], []
        # U: 1 for propagating, +-i couplings for evanescent air-gap orders
        U = np.matrix(np.zeros((2 * PW_pols, 2 * PW_pols), complex))
        for i in range(num_prop_air):
            U[i, i] = 1.0
            U[neq_PW + i, neq_PW + i] = 1.0
            U[PW_pols + i, PW_pols + i] = -1.0
            U[PW_pols + neq_PW + i, PW_pols + neq_PW + i] = -1.0
        for i in range(num_prop_air, neq_PW):
            U[i, PW_pols + i] = -1.0j
            U[neq_PW + i, PW_pols + neq_PW + i] = 1.0j
            U[PW_pols + i, i] = 1.0j
            U[PW_pols + neq_PW + i, neq_PW + i] = -1.0j
        d_minus = top.specular_incidence(pol) if incoming_amplitudes is None else incoming_amplitudes

        def prop_flux(v, nprop):
            return np.linalg.norm(v[0:nprop]) ** 2 + np.linalg.norm(v[neq_PW:neq_PW + nprop]) ** 2

        down.append(prop_flux(d_minus, num_prop_in))
        d_plus = rnet_list[-1] * d_minus
        self.vec_coef_down.append(d_minus)
        self.vec_coef_up.append(d_plus)
        up.append(prop_flux(d_plus, num_prop_in))
        f1_minus = inv_t21_list[-1] * d_minus
        for i in range(len(self.layers) - 2):
            f1_plus = rnet_list[-2 * i - 2] * f1_minus
            f_mat = np.matrix(np.concatenate((f1_minus, f1_plus)))
            down.append(f_mat.H * U * f_mat)
            f2_minus = inv_t12_list[-i - 1] * f1_minus
            f2_plus = rnet_list[-2 * i - 3] * P_list[-i - 1] * f2_minus
            self.vec_coef_down.append(f2_minus)
            self.vec_coef_up.append(f2_plus)
            f1_minus = inv_t21_list[-i - 2] * P_list[-i - 1] * f2_minus
        f2_minus = tnet_list[0] * f1_minus
        self.vec_coef_down.append(f2_minus)
        if isinstance(bottom, Anallo):
            num_prop_out = bottom.num_prop_pw_per_pol
            if num_prop_out != 0:
                down.append(prop_flux(f2_minus, num_prop_out))
                self._tally(down, up, absolute_t=True)
            else:
                self.a_list = np.zeros(len(self.layers) - 2)
                self.r_list = np.zeros(len(self.layers) - 2)
                self.t_list = np.zeros(len(self.layers) - 2)
        else:
            self._tally(down, up, absolute_t=False)

    def _tally(self, down, up, absolute_t):
        for i in range(1, len(down) - 1):
            self.a_list.append(abs(abs(down[i]) - abs(down[i + 1])))
        self.a_list.append(abs(down[0] - down[-1] - up[0]))
        for i in range(1, len(up) - 1):
            self.r_list.append(abs(up[i]) / abs(down[i]))
        self.r_list.append(abs(up[0] / down[0]))
        for i in range(0, len(down) - 2):
            self.t_list.append(abs(abs(down[i + 2]) / abs(down[i])) if absolute_t
                               else abs(down[i + 2]) / abs(down[i]))
        self.t_list.append(abs(down[-1] / down[0]))

    def _check_periods_are_consistent(self):
        for lay in self.layers:
            assert lay.structure.period == self.period, \
                "All layers in a multilayer stack must have the same period."
            assert lay.structure.period_y == self.period_y, \
                "All layers in a multilayer stack must have the same period_y."
