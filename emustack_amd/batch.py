"""Batched mode solves: many Light objects on one NanoStruct in one library
call (the reference solves one wavelength per process, see
tests/test_case_1_simple_struct.py:97-100).  Returns ordinary Simmo objects,
so Stack.calc_scat is unchanged."""
import numpy as np

from . import EMUstack, paths
from .mode_calcs import Simmo

__all__ = ['calc_modes_many']


def calc_modes_many(structure, lights, num_BMs=None, tol=0.0, itermax=30):
    """Simmo objects for every light in `lights` (same max_order_PWs),
    computed together on the GPU."""
    if structure.periodicity != '2D_array':
        raise NotImplementedError("batched path supports 2D_array NanoStructs")
    sims = [Simmo(structure, l) for l in lights]
    orders = {l.max_order_PWs for l in lights}
    if len(orders) != 1:
        raise ValueError("all lights of a batch need the same max_order_PWs")
    inputs = [s.fem_inputs(num_BMs) for s in sims]
    nbms = {x[2] for x in inputs}
    if len(nbms) != 1:
        raise ValueError("all lights of a batch need the same num_BMs")
    nbm = nbms.pop()
    P = inputs[0][1]
    beta, J, Jd, stats = EMUstack.calc_modes_2d_batch(
        [s.wl_norm() for s in sims], nbm, orders.pop(), paths.msh_path + structure.mesh_file,
        [x[0] for x in inputs], [s.k_pll_norm() for s in sims], [x[3] for x in inputs], P, tol=tol,
        itermax=itermax)
    area_norm = structure.period * structure.period_y / structure.period ** 2
    for i, s in enumerate(sims):
        s.num_BMs = nbm
        s.k_z = beta[i]
        s.J = np.mat(J[i]) / area_norm
        s.J_dag = np.mat(Jd[i])
        s.sol1 = s.mode_pol = s.table_nod = s.type_el = s.x_arr = None
    calc_modes_many.last_stats = stats
    return sims


calc_modes_many.last_stats = None
