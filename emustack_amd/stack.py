"""A stack of layers: net scattering matrices and energy fluxes.

Public surface kept from the reference's backend/stack.py:27-435: the
Stack class with layers, heights_nm(), total_height(), structures(),
calc_scat(pol, incoming_amplitudes, calc_fluxes, save_scat_list), and after
calc_scat the attributes R_net, T_net, a_list, r_list, t_list,
vec_coef_down, vec_coef_up (plus tnet_list, rnet_list when asked).

How it works.  Every layer sits between two infinitesimal air gaps, so each
layer contributes the matrices of its interface with air (layer.air_ref()).
The net reflection and transmission seen from the top of the substrate are
grown upwards one step at a time: an air gap (multiple reflections between
the net structure below and the next interface) and then the layer body
(propagation over its height and back), Redheffer star products written as
linear solves (stack.py:204-278).  The flux pass then walks down again,
rebuilding the amplitudes in every air gap from the stored solves and
taking the net downward power there with the gap's power matrix U
(Dossou et al., JOSA A 29, 817 (2012), appendix C; stack.py:285-427).

Matrix layout of every scattering matrix (TE block first):
    | TE -> TE | TM -> TE |
    | TE -> TM | TM -> TM |
"""
import numpy as np

from .mode_calcs import r_t_mat, Anallo

__all__ = ['Stack']


def _air_gap_power_matrix(n_pw, n_prop):
    """U with f^H U f = net downward power of the air-gap amplitudes
    f = [down; up] (2 polarisations x n_pw orders each): +1 / -1 on the
    propagating down / up orders, and the evanescent orders coupled
    between down and up with +-i."""
    n = 2 * n_pw
    U = np.zeros((2 * n, 2 * n), complex)
    for pol in (0, n_pw):
        prop = np.arange(pol, pol + n_prop)
        evan = np.arange(pol + n_prop, pol + n_pw)
        U[prop, prop] = 1.0
        U[n + prop, n + prop] = -1.0
        sign = -1.0 if pol == 0 else 1.0
        U[evan, n + evan] = sign * 1j
        U[n + evan, evan] = -sign * 1j
    return np.asmatrix(U)


def _propagating_power(v, n_pw, n_prop):
    """Power carried by the first n_prop orders of both polarisations."""
    v = np.asarray(v).reshape(-1)
    return np.linalg.norm(v[:n_prop]) ** 2 + np.linalg.norm(v[n_pw:n_pw + n_prop]) ** 2


class Stack(object):
    """Layers (Anallo / Simmo objects) ordered from the bottom (substrate)
    to the top (superstrate)."""

    def __init__(self, layers, heights_nm=None, shears=None):
        self.layers = tuple(layers)
        self._heights_nm = heights_nm
        self.shears = shears
        self.period = float(layers[0].structure.period)
        self.period_y = float(layers[0].structure.period_y)
        self._check_periods_are_consistent()
        if heights_nm is not None:
            # heights given here override (and are written to) the layers
            for lay, h in zip(self.layers[1:-1], heights_nm):
                lay.structure.height_nm = h

    def heights_nm(self):
        """Heights of the interior layers, bottom to top."""
        if self._heights_nm is not None:
            return self._heights_nm
        return [float(lay.structure.height_nm) for lay in self.layers[1:-1]]

    def heights_norm(self):
        return [h / self.period for h in self.heights_nm()]

    def total_height(self):
        return sum(self.heights_nm())

    def structures(self):
        return (lay.structure for lay in self.layers)

    # ------------------------------------------------------------ cascade

    def _gap_shear(self, index):
        """Lateral offset across the air gap below layer `index` (None when
        the stack has no shears)."""
        if self.shears is None:
            return None
        if index == len(self.layers) - 1:
            return np.asarray(self.shears[-1])
        return np.asarray(self.shears[index - 1]) - np.asarray(self.shears[index])

    def calc_scat(self, pol='TE', incoming_amplitudes=None, calc_fluxes=True, save_scat_list=False):
        """Net R/T of the stack (R_net, T_net) and, with calc_fluxes, the
        per-layer absorptance, reflectance and transmittance lists."""
        self._check_periods_are_consistent()
        faces = []
        for lay in self.layers:
            R12, T12, R21, T21 = r_t_mat(lay.air_ref(), lay)
            lay.R12, lay.T12, lay.R21, lay.T21 = R12, T12, R21, T21
            faces.append((R12, T12, R21, T21))
        n_pols = int(np.shape(faces[-1][0])[0])
        eye_gap = np.asmatrix(np.eye(n_pols, dtype=complex))
        heights = self.heights_nm()
        shear_ref = self.layers[-1]

        # net matrices seen from above, after the substrate's own face
        t_net, r_net = faces[0][1], faces[0][0]
        t_nets, r_nets = [t_net], [r_net]
        gap_solves, body_solves, props = [], [], []

        def cross_gap(i, t_net, r_net):
            R12, T12, R21, T21 = faces[i]
            shift = self._gap_shear(i)
            if shift is None:
                g = np.linalg.solve(eye_gap - R12 * r_net, T21)
                return g, t_net * g, R21 + T12 * r_net * g
            Q, Qi = shear_ref.shear_transform(shift), shear_ref.shear_transform(-1 * shift)
            g = np.linalg.solve(eye_gap - R12 * Qi * r_net * Q, T21)
            return g, t_net * Q * g, R21 + T12 * Qi * r_net * Q * g

        for i in range(1, len(self.layers) - 1):
            g, t_net, r_net = cross_gap(i, t_net, r_net)
            gap_solves.append(g)
            t_nets.append(t_net)
            r_nets.append(r_net)
            R12, T12, R21, T21 = faces[i]
            P = self.layers[i].prop_fwd(heights[i - 1] / self.period)
            eye_body = np.asmatrix(np.eye(len(P), dtype=complex))
            b = np.linalg.solve(eye_body - R21 * P * r_net * P, T12)
            Pb = P * b
            t_net = t_net * Pb
            r_net = R12 + T21 * P * r_net * Pb
            props.append(P)
            body_solves.append(b)
            t_nets.append(t_net)
            r_nets.append(r_net)
        g, t_net, r_net = cross_gap(len(self.layers) - 1, t_net, r_net)
        gap_solves.append(g)
        t_nets.append(t_net)
        r_nets.append(r_net)

        self.R_net, self.T_net = r_net, t_net
        if save_scat_list:
            self.tnet_list, self.rnet_list = t_nets, r_nets
        if calc_fluxes:
            self._fluxes(pol, incoming_amplitudes, n_pols, r_nets, t_nets, gap_solves, body_solves, props)

    # ------------------------------------------------------------- fluxes

    def _fluxes(self, pol, incoming, n_pols, r_nets, t_nets, gap_solves, body_solves, props):
        """Walk down from the superstrate, rebuilding the down/up amplitudes
        in every air gap, and turn the net downward powers into a_list,
        r_list, t_list (stack.py:285-427)."""
        n_pw = n_pols // 2
        top, bottom = self.layers[-1], self.layers[0]
        U = _air_gap_power_matrix(n_pw, top.air_ref().num_prop_pw_per_pol)
        n_prop_top = top.num_prop_pw_per_pol
        d_minus = top.specular_incidence(pol) if incoming is None else incoming
        d_plus = r_nets[-1] * d_minus
        self.vec_coef_down, self.vec_coef_up = [d_minus], [d_plus]
        down = [_propagating_power(d_minus, n_pw, n_prop_top)]
        up = [_propagating_power(d_plus, n_pw, n_prop_top)]

        f_down = gap_solves[-1] * d_minus  # amplitudes in the gap under the top
        n_inner = len(self.layers) - 2
        for j in range(n_inner):
            f_up = r_nets[-2 * j - 2] * f_down
            f = np.asmatrix(np.concatenate((f_down, f_up)))
            down.append(f.H * U * f)
            body_down = body_solves[-j - 1] * f_down
            body_up = r_nets[-2 * j - 3] * props[-j - 1] * body_down
            self.vec_coef_down.append(body_down)
            self.vec_coef_up.append(body_up)
            f_down = gap_solves[-j - 2] * props[-j - 1] * body_down
        out = t_nets[0] * f_down
        self.vec_coef_down.append(out)

        self.a_list, self.r_list, self.t_list = [], [], []
        if isinstance(bottom, Anallo):
            n_prop_out = bottom.num_prop_pw_per_pol
            if n_prop_out == 0:
                # nothing propagates into the substrate: no flux bookkeeping
                self.a_list = np.zeros(n_inner)
                self.r_list = np.zeros(n_inner)
                self.t_list = np.zeros(n_inner)
                return
            down.append(_propagating_power(out, n_pw, n_prop_out))
            self._tally(down, up, take_abs=True)
        else:
            self._tally(down, up, take_abs=False)

    def _tally(self, down, up, take_abs):
        """Per-layer A, R, T from the net downward powers `down` (superstrate,
        every air gap, substrate) and the reflected powers `up`; the last
        entry of each list is the whole stack's."""
        for i in range(1, len(down) - 1):
            self.a_list.append(abs(abs(down[i]) - abs(down[i + 1])))
        self.a_list.append(abs(down[0] - down[-1] - up[0]))
        for i in range(1, len(up) - 1):
            self.r_list.append(abs(up[i]) / abs(down[i]))
        self.r_list.append(abs(up[0] / down[0]))
        for i in range(len(down) - 2):
            ratio = abs(down[i + 2]) / abs(down[i])
            self.t_list.append(abs(ratio) if take_abs else ratio)
        self.t_list.append(abs(down[-1] / down[0]))

    def _check_periods_are_consistent(self):
        for lay in self.layers:
            if lay.structure.period != self.period:
                raise AssertionError("every layer of a stack needs the same period (x)")
            if lay.structure.period_y != self.period_y:
                raise AssertionError("every layer of a stack needs the same period (y)")
