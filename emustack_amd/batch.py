"""Batched mode solves: many Light objects on one NanoStruct in one library
call (the reference solves one wavelength per process, see
tests/test_case_1_simple_struct.py:97-100).  Returns ordinary Simmo objects,
so Stack.calc_scat is unchanged.

With groups > 1 the lights are split into that many groups, each solved by
its own host thread on its own context (HIP stream and work buffers): while
one group's projected Schur problems run on the host cores, another group's
Krylov expansion runs on the GPU."""
from concurrent.futures import ThreadPoolExecutor
import threading

import numpy as np

from . import EMUstack, paths
from .mode_calcs import Simmo

__all__ = ['calc_modes_many', 'calc_modes_stack', 'calc_films_many', 'calc_scat_many']

_pools = {}
_pool_lock = threading.Lock()


def _pool(n):
    with _pool_lock:
        p = _pools.get(n)
        if p is None:
            p = ThreadPoolExecutor(max_workers=n, thread_name_prefix='emu-group')
            _pools[n] = p
        return p


PRECISIONS = {'fp64': 0, 'mixed': 1}


def _solve(structure, sims, inputs, nbm, order, P, tol, itermax, ncv, precision='fp64'):
    st = structure
    lam = [s.wl_norm() for s in sims]
    n_effs = [x[0] for x in inputs]
    kps = [s.k_pll_norm() for s in sims]
    shifts = [x[3] for x in inputs]
    if st.periodicity == '2D_array':
        beta, J, Jd, stats = EMUstack.calc_modes_2d_batch(lam, nbm, order, paths.msh_path + st.mesh_file, n_effs,
                                                          kps, shifts, P, tol=tol, itermax=itermax, ncv=ncv,
                                                          precision=PRECISIONS[precision])
        area_norm = st.period * st.period_y / st.period ** 2
        return beta, [np.asmatrix(j) / area_norm for j in J], [np.asmatrix(j) for j in Jd], stats
    h = EMUstack.mesh_1d(st.x_arr, st.table_nod, st.type_el)
    w1 = st.world_1d is True
    P2 = 1 if w1 else sims[0].calc_2d_grating_orders(order)[0].size
    beta, J, Jd, J2, Jd2, stats = EMUstack.context().calc_modes_1d_batch(
        h, lam, nbm, order, n_effs, kps, shifts, P, neq_pw_2d=P2, world_1d=w1, itermax=itermax, tol=tol, ncv=ncv,
        precision=PRECISIONS[precision])
    if not w1:
        J, Jd = J2, Jd2
    return beta, [np.asmatrix(j) for j in J], [np.asmatrix(j) for j in Jd], stats


def _prepare(structure, lights, num_BMs):
    if structure.periodicity not in ('1D_array', '2D_array'):
        raise ValueError("NanoStruct layer must have periodicity of either '1D_array' or '2D_array'.")
    sims = [Simmo(structure, l) for l in lights]
    orders = {l.max_order_PWs for l in lights}
    if len(orders) != 1:
        raise ValueError("all lights of a batch need the same max_order_PWs")
    inputs = [s.fem_inputs(num_BMs) for s in sims]
    nbms = {x[2] for x in inputs}
    if len(nbms) != 1:
        raise ValueError("all lights of a batch need the same num_BMs")
    return sims, inputs, nbms.pop(), orders.pop(), inputs[0][1]


def _parts(n, groups):
    groups = max(1, min(int(groups), n))
    cuts = np.linspace(0, n, groups + 1).round().astype(int)
    return [(cuts[g], cuts[g + 1]) for g in range(groups)]


def _fill(sims, nbm, parts, results):
    stats = np.zeros(8, np.int64)
    for (a, b), (beta, J, Jd, st) in zip(parts, results):
        stats[0] += st[0]
        stats[1] = max(stats[1], st[1])
        stats[2:] += st[2:]
        for i, s in enumerate(sims[a:b]):
            s.num_BMs = nbm
            s.k_z = beta[i]
            s.J = J[i]
            s.J_dag = Jd[i]
            s.sol1 = s.mode_pol = s.table_nod = s.type_el = s.x_arr = None
    return stats


def calc_modes_many(structure, lights, num_BMs=None, tol=0.0, itermax=30, groups=1, ncv=0, precision='fp64'):
    """Simmo objects for every light in `lights` (same max_order_PWs),
    computed together on the GPU (`groups` concurrent sub-batches; ncv the
    Krylov dimension, 0 for the reference's rule; precision 'fp64' or
    'mixed')."""
    if precision not in PRECISIONS:
        raise ValueError("precision must be 'fp64' or 'mixed'")
    sims, inputs, nbm, order, P = _prepare(structure, lights, num_BMs)
    parts = _parts(len(sims), groups)
    if len(parts) == 1:
        results = [_solve(structure, sims, inputs, nbm, order, P, tol, itermax, ncv, precision)]
    else:
        futs = [_pool(len(parts)).submit(_solve, structure, sims[a:b], inputs[a:b], nbm, order, P, tol, itermax, ncv,
                                         precision) for a, b in parts]
        results = [f.result() for f in futs]
    calc_modes_many.last_stats = _fill(sims, nbm, parts, results)
    return sims


def calc_modes_stack(structures, lights, num_BMs=None, tol=0.0, itermax=30, groups=2, ncv=0, workers=None,
                     precision='fp64'):
    """calc_modes_many for several NanoStruct layers at once: every (layer,
    sub-batch) job goes to one pool of `workers` host threads (default
    groups + 1), each with its own context, so that one job's host Schur
    steps overlap another's GPU work and the GPU does not idle between
    layers.  Returns one list of Simmo objects per structure."""
    if precision not in PRECISIONS:
        raise ValueError("precision must be 'fp64' or 'mixed'")
    preps = [_prepare(st, lights, num_BMs) for st in structures]
    pool = _pool(int(workers) if workers else groups + 1)
    jobs = []
    for st, (sims, inputs, nbm, order, P) in zip(structures, preps):
        parts = _parts(len(sims), groups)
        futs = [pool.submit(_solve, st, sims[a:b], inputs[a:b], nbm, order, P, tol, itermax, ncv, precision)
                for a, b in parts]
        jobs.append((parts, futs))
    stats = np.zeros(8, np.int64)
    out = []
    for (sims, inputs, nbm, order, P), (parts, futs) in zip(preps, jobs):
        st = _fill(sims, nbm, parts, [f.result() for f in futs])
        stats[0] += st[0]
        stats[1] = max(stats[1], st[1])
        stats[2:] += st[2:]
        out.append(sims)
    calc_modes_stack.last_stats = stats
    return out


def preload_meshes(structures, groups=2, workers=None):
    """Load the meshes of the 2D layers in `structures` into the context of
    every thread of the pool that calc_modes_stack(..., groups, workers=)
    uses, so that later batches find them loaded (a load is a one-time
    setup: mesh file, symbolic analysis, upload).  One task per pool thread
    (the tasks wait for each other at a barrier)."""
    n = int(workers) if workers else groups + 1
    files = sorted({paths.msh_path + st.mesh_file for st in structures if st.periodicity == '2D_array'})
    pool = _pool(n)
    bar = threading.Barrier(n)

    def load():
        bar.wait(timeout=600)
        for f in files:
            EMUstack.mesh(f)

    for f in [pool.submit(load) for _ in range(n)]:
        f.result()


calc_modes_stack.last_stats = None
calc_modes_many.last_stats = None


def calc_films_many(film, lights):
    """ThinFilm.calc_modes for many lights (mode_calcs.py:110-212 of the
    reference, once per light there): the k_z and impedances of every
    (light, order) in one GPU launch (emu_film_modes_batch).  The orders,
    their in-plane wavevectors and the air reference's ordering are host
    bookkeeping; a light whose orders include a grazing one (k_z exactly 0)
    is evaluated again with k_parallel scaled by 1 - 1e-9, as Anallo does.
    Returns one Anallo per light, as film.calc_modes(light) would."""
    from .mode_calcs import Anallo
    if not lights:
        return []
    films = [Anallo(film, light) for light in lights]
    waves = [f.wavevectors() for f in films]
    grazing = np.zeros(len(films), bool)
    for _ in range(2):
        n = np.array([f.n() for f in films], np.complex128)
        wl = np.array([f.wl_norm() for f in films])
        alpha = np.array([w[2] for w in waves])
        beta = np.array([np.broadcast_to(w[3], w[2].shape) for w in waves])
        order = np.array([f.air_ref().sort_order for f in films])
        if alpha.ndim != 2 or order.shape != alpha.shape:
            raise AssertionError("film and air reference disagree on the number of orders")
        kz, Z = EMUstack.context().film_modes(n, wl, alpha, beta, order)
        hit = np.any(kz == 0, axis=1) & ~grazing
        if not hit.any():
            break
        for i in np.flatnonzero(hit):
            waves[i] = films[i].wavevectors(grazing=True)
        grazing |= hit
    npw = alpha.shape[1]
    for i, f in enumerate(films):
        px, py, a, b = waves[i]
        f.k_z = f.take_kz(kz[i, :npw], px, py, a, b, presorted=True)
        f.k_z = np.concatenate((f.k_z, f.k_z))
        f._device_Z = Z[i]
        f.structure.num_pw_per_pol = npw
    return films


def calc_scat_many(stacks, pol='TE', want_nets=False):
    """Stack.calc_scat (stack.py:162-427, shears = None, fluxes on) for many
    stacks of identical layer structure at once, as a batched star-product
    cascade on the GPU.  Fills a_list, r_list, t_list (and R_net, T_net when
    want_nets) on every stack, as Stack.calc_scat does.  The bottom and top
    layers must be thin films."""
    from .mode_calcs import Anallo
    if not stacks:
        return stacks
    first = stacks[0]
    L = len(first.layers)
    kinds = [0 if isinstance(l, Anallo) else 1 for l in first.layers]
    if kinds[0] != 0 or kinds[-1] != 0:
        raise ValueError("calc_scat_many needs thin films as top and bottom layers")
    for st in stacks:
        if [0 if isinstance(l, Anallo) else 1 for l in st.layers] != kinds:
            raise ValueError("all stacks of a batch need the same layer structure")
        if st.shears is not None:
            raise NotImplementedError("sheared stacks are not batched")
    kz, Z, J, Jd = [], [], [], []
    for li in range(L):
        lays = [st.layers[li] for st in stacks]
        if kinds[li] == 0:
            kz.append(np.array([l.k_z for l in lays]))  # both polarisations (Anallo.calc_modes)
            Z.append(np.array([l.Z() for l in lays]))
            J.append(None)
            Jd.append(None)
        else:
            kz.append(np.array([np.asarray(l.k_z).reshape(-1) for l in lays]))
            Z.append(None)
            J.append(np.array([np.asarray(l.J) for l in lays]))
            Jd.append(np.array([np.asarray(l.J_dag) for l in lays]))
    for st in stacks:
        if st.period != first.period or st.period_y != first.period_y:
            raise ValueError("all stacks of a batch need the same period")
    # heights may differ from stack to stack (a sweep over layer thickness)
    heights = np.array([[0.0] + [h / st.period for h in st.heights_nm()] + [0.0] for st in stacks])
    top = [st.layers[-1] for st in stacks]
    Zair = np.array([t.air_ref().Z() for t in top])
    d_minus = np.array([np.asarray(t.specular_incidence(pol)).reshape(-1) for t in top])
    nprop = np.array([[t.air_ref().num_prop_pw_per_pol, t.num_prop_pw_per_pol,
                       st.layers[0].num_prop_pw_per_pol] for t, st in zip(top, stacks)], np.int64)
    a_l, r_l, t_l, R, T = EMUstack.context().star_cascade(kinds, kz, Z, J, Jd, heights, Zair, d_minus, nprop,
                                                         want_nets=want_nets)
    for i, st in enumerate(stacks):
        st.a_list = list(a_l[i])
        st.r_list = [r_l[i]]
        st.t_list = list(t_l[i])
        if want_nets:
            st.R_net, st.T_net = np.asmatrix(R[i]), np.asmatrix(T[i])
    return stacks
