#!/usr/bin/env python
"""Throughput of the per-wavelength FEM + scattering-matrix path on one MI355X.

Workload (stacked gratings, 8k-triangle meshes): square period 500 nm; three
structured layers with circular inclusions r = 100, 150, 80 nm (heights
100 / 150 / 120 nm, inclusion index drawn from U[2,4] + i U[0,0.05] with
seed 3, air host) separated by two 50 nm films of n = 1.46, on a substrate of
n = 1.5 under air.  Each layer mesh has about 8,000 P2 triangles (project
mesher, seeded); num_BMs = 200, max_order_PWs = 4 (49 plane waves per
polarisation, S-matrix dimension 98); 512 wavelengths in 400-1000 nm at
normal incidence.  Data are synthetic.

A step solves the modes of all three layers for a batch of wavelength
points (prime and adjoint problems, batched library calls per layer) and
runs the star-product cascade with fluxes for every point (one batched
cascade on the GPU).  `value` is wavelength-points per second of the
precision form given by --precision (default: FP64 throughout).  One step of
the other form on the same batch gives the mixed/FP64 ratio and the R/T/A
difference between the two (mixed: the Krylov solves read a complex64 copy
of the FP64 inverted factor, and every beta is the FP64 two-sided Rayleigh
quotient), and both are compared with the oracle's FP64 result at the
fixture wavelength (tests/golden/configs/config3.npz) when the batch holds
it.  The default is FP64 because the mixed form misses the 1e-6 R/T/A bar
on this workload (its eigenvectors carry the complex64 operator's
first-order error).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
                    [--precision fp64|mixed] [--no-cpu] [--no-compare]
                    [--copy-calibration] [--trace-summary DIR]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PERIOD = 500.0
RADII = (100.0, 150.0, 80.0)
HEIGHTS = (100.0, 150.0, 120.0)
NTRI = 8000
NUM_BMS = 200
ORDER = 4
NWL = 512
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FIXTURE = os.path.join(ROOT, 'tests', 'golden', 'configs', 'config3.npz')
# HBM bytes per dispatch of the kernels that can be the roofline kernel, from
# two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of this workload at 16
# points in 2 sub-batches (8 points = 16 instances per dispatch, as in the
# default step), each scaled by what the counter reads for a plain 2 GiB copy;
# summary and commands in profiles/kernels_stacked_gratings_8k.md.
TRAFFIC_PER_DISPATCH = {
    'k_atb_mfma16<true>': 7475500000,
    'k_combine<16>': 7323900000,
    'k_fwd': 433500000,
    'k_bwd': 412900000,
    'k_spmm_il': 2810600000,
    'k_zgemm_nn_tall': 3957700000,
}
TRAFFIC_SOURCE = 'profiles/kernels_stacked_gratings_8k.md'


def inclusion_indices():
    rng = np.random.default_rng(3)
    return [complex(rng.uniform(2.0, 4.0), rng.uniform(0.0, 0.05)) for _ in RADII]


def build_meshes(workdir):
    from emustack_amd import mesher
    paths = []
    for i, r in enumerate(RADII):
        p = os.path.join(workdir, 'stacked_r%d_s%d.mail' % (int(r), i + 1))
        if not os.path.exists(p):
            tmp = p + '.%d.tmp' % os.getpid()
            mesher.write_mail(tmp, mesher.ellipse_cell_ntri(PERIOD, PERIOD, r, r, NTRI, seed=i + 1))
            os.replace(tmp, p)
        paths.append(p)
    return paths


def layers_for(mesh_paths):
    from emustack_amd import objects, materials
    nincl = inclusion_indices()
    ns = [objects.NanoStruct('2D_array', PERIOD, 2 * r, height_nm=h, background=materials.Air,
                             inclusion_a=materials.Material(n), loss=True, make_mesh_now=False, mesh_file=p)
          for r, h, n, p in zip(RADII, HEIGHTS, nincl, mesh_paths)]
    film = objects.ThinFilm(PERIOD, height_nm=50.0, material=materials.Material(1.46 + 0j))
    sub = objects.ThinFilm(PERIOD, height_nm='semi_inf', material=materials.Material(1.5 + 0j), loss=False)
    sup = objects.ThinFilm(PERIOD, height_nm='semi_inf', material=materials.Air, loss=False)
    return ns, film, sub, sup


def run_step(ns, film, sub, sup, wls, groups=1, ncv=0, workers=None, precision='fp64'):
    """All three layers for the wavelength batch `wls` (`groups` concurrent
    sub-batches per layer, Krylov dimension ncv), then the cascade.  Returns
    R/T/A per point and the summed solver stats."""
    from emustack_amd import objects, Stack
    from emustack_amd.batch import calc_films_many, calc_modes_stack, calc_scat_many
    lights = [objects.Light(wl, max_order_PWs=ORDER, theta=0.0, phi=0.0) for wl in wls]
    sims = calc_modes_stack(ns, lights, num_BMs=NUM_BMS, groups=groups, ncv=ncv, workers=workers,
                            precision=precision)
    stats = calc_modes_stack.last_stats.copy()
    # film modes of all points, one launch per film layer
    fl = [calc_films_many(f, lights) for f in (sub, film, film, sup)]
    stacks = [Stack((fl[0][i], sims[0][i], fl[1][i], sims[1][i], fl[2][i], sims[2][i], fl[3][i]))
              for i in range(len(lights))]
    calc_scat_many(stacks, pol='TE')  # batched star-product cascade on the GPU
    out = np.array([(float(st.a_list[-1]), float(st.r_list[-1]), float(st.t_list[-1])) for st in stacks])
    return out, stats


# ---------------------------------------------------------------- roofline


def symbolic_flops(mesh_path):
    """Real FP64 flops of one numeric factorisation and of its inverted form
    (per front with s pivots and u update rows: LU s f^2 - s^2 f + s^3/3,
    inversion s^3/3 + u s^2 complex multiply-adds, 8 flops each), and the
    complex entries of the inverted factor."""
    from emustack_amd import _lib
    mh = _lib.MeshHandle(mesh_path)
    fr = mh.symbolic()['fronts'].reshape(-1, 8).astype(np.float64)
    s, u, f = fr[:, 0], fr[:, 1], fr[:, 2]
    lu = 8.0 * np.sum(s * f * f - s * s * f + s ** 3 / 3.0)
    inv = 8.0 * np.sum(s ** 3 / 3.0 + u * s * s)
    return lu, inv, mh.neq, mh.nnz


def whole_step_roofline(prof, mesh_info, npts, ops_per_inst, m, keep, step_s):
    """T_roofline / T_measured of the timed region: the profiled kernel
    classes at their algorithmic bytes over HBM peak, plus the
    factorisations, triangular inversions, restart and Ritz-vector products
    at their flop counts over the FP64 peak, plus assembly writes."""
    bw = HBM_PEAK_GBS * 1e9
    fp = FP64_PEAK_TFLOPS * 1e12
    t_classes = sum(v[2] for v in prof.values()) / bw
    t_flops = 0.0
    t_asm = 0.0
    for lu, inv, neq, nnz in mesh_info:
        inst = 2 * npts  # prime + adjoint per point
        restarts = max(0.0, (ops_per_inst - m) / max(m - keep, 1))
        t_flops += inst * (lu + inv + 8.0 * neq * m * keep * restarts + 8.0 * neq * m * NUM_BMS) / fp
        t_asm += inst * 2 * nnz * 16.0 / bw
    t_roof = t_classes + t_flops + t_asm
    return dict(T_roofline_s=t_roof, T_measured_s=step_s, achieved=t_roof / step_s,
                parts_s=dict(profiled_classes=t_classes, factor_inverse_restart_flops=t_flops, assembly=t_asm))


# ------------------------------------------------------------- CPU baseline

CPU_SAMPLE = r'''
import sys, time, json
import numpy as np
sys.path.insert(0, ROOT)
from oracle import fem2d
from emustack_amd import _lib
import scipy.sparse.linalg as spla
mesh_path, wl, n_incl, nops = sys.argv[1], float(sys.argv[2]), complex(sys.argv[3]), int(sys.argv[4])
t0 = time.perf_counter()
m = fem2d.build_mesh(mesh_path)
t_mesh = time.perf_counter() - t0
perm = _lib.MeshHandle(mesh_path).arrays()['perm']
wln = wl / PERIOD; k0 = 2 * np.pi / wln
n_eff = np.array([1.0, n_incl]); th = np.deg2rad(1e-5)
kpn = 2 * np.pi / wl * np.sin(th) * np.array([np.cos(th), np.sin(th)]) * PERIOD
shift = 1.1 * np.real(n_eff).max() ** 2 * k0 ** 2 - kpn[0] ** 2 - kpn[1] ** 2
t0 = time.perf_counter()
A, B = fem2d.assemble(m, np.ones(2, complex), n_eff ** 2 * k0 ** 2, shift, kpn)
t_asm = time.perf_counter() - t0
Ap = A.tocsr()[perm][:, perm].tocsc(); Bp = B.tocsr()[perm][:, perm].tocsr()
t0 = time.perf_counter()
# threshold partial pivoting with diagonal preference on the nested-dissection
# order (the symmetric strategy UMFPACK applies to this structurally
# symmetric pencil, pivot tolerance 0.001)
lu = spla.splu(Ap, permc_spec='NATURAL', diag_pivot_thresh=0.001, options=dict(SymmetricMode=True))
t_lu = time.perf_counter() - t0
v = np.random.default_rng(0).standard_normal(m.neq) + 0j
t0 = time.perf_counter()
for i in range(nops):
    v = lu.solve(Bp @ v); v /= np.linalg.norm(v)
t_op = (time.perf_counter() - t0) / nops
mcv = 2 * NBM + NBM // 2 + 3; j = (NBM + mcv) // 2
V = np.random.default_rng(1).standard_normal((j, m.neq)) + 0j
w = v.copy()
t0 = time.perf_counter()
for i in range(4):
    for p in range(2):
        h = V.conj() @ w; w = w - V.T @ h
t_orth = (time.perf_counter() - t0) / 4
print(json.dumps(dict(t_mesh=t_mesh, t_asm=t_asm, t_lu=t_lu, t_op=t_op, t_orth=t_orth, neq=m.neq)))
'''


def _cpu_worker(args):
    code, mesh_path, wl, n_incl, nops = args
    env = dict(os.environ, OMP_NUM_THREADS='1', OPENBLAS_NUM_THREADS='1', MKL_NUM_THREADS='1')
    r = subprocess.run([sys.executable, '-c', code, mesh_path, str(wl), str(n_incl), str(nops)],
                       capture_output=True, text=True, env=env, timeout=900)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-2000:])
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baseline(mesh_paths, wls, n_incls, ops_per_instance, cores, nops=40):
    """The oracle's FP64 path on the host cores: `cores` single-threaded
    processes at once, each timing on one layer mesh and wavelength the mesh
    preprocessing, assembly, SuperLU factorisation (threshold partial
    pivoting, tolerance 0.001 with diagonal preference, on the
    nested-dissection order) and `nops` shift-invert OP
    applications and CGS2 steps of an ARPACK-sized basis; an eigensolve is
    assembly + factorisation + ops_per_instance x (OP + CGS2), a point 6
    eigensolves, and the cores work on independent points."""
    from concurrent.futures import ThreadPoolExecutor
    code = 'ROOT=%r; PERIOD=%r; NBM=%r\n' % (ROOT, PERIOD, NUM_BMS) + CPU_SAMPLE
    jobs = [(code, mesh_paths[i % 3], wls[i % len(wls)], n_incls[i % 3], nops) for i in range(cores)]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cores) as ex:
        res = list(ex.map(_cpu_worker, jobs))
    wall = time.perf_counter() - t0
    t_eig = np.mean([d['t_asm'] + d['t_lu'] + ops_per_instance * (d['t_op'] + d['t_orth']) for d in res])
    value = cores / (6.0 * t_eig)
    d = res[0]
    sample = ('%d processes at once (one core each, %.0f s wall): per process one layer mesh at one wavelength, '
              'assembly %.1f s + SuperLU (threshold pivoting 0.001, ND order) %.1f s + %d OP applications '
              '(%.1f ms each) and CGS2 steps (%.1f ms each); an eigensolve = assembly + LU + %d x (OP + CGS2) '
              '(the GPU run\'s OP count), 6 eigensolves per point, points independent across cores'
              % (cores, wall, d['t_asm'], d['t_lu'], nops, 1e3 * d['t_op'], 1e3 * d['t_orth'], ops_per_instance))
    return dict(value=value, unit='wavelength-points/s', cores=cores, kind='port', sample=sample)


def trace_summary(trace_dir, top=30):
    """GPU-busy fraction and per-kernel time of the timed step in a rocprofv3
    --kernel-trace CSV of `bench.py --steps 1 --warmup 1`: the step runs from
    the end of the warmup step's cascade (k_prop dispatches, one batched
    cascade per step) to the end of the timed step's; busy = union of the
    kernel intervals inside it."""
    import csv
    import glob
    import collections
    files = glob.glob(os.path.join(trace_dir, '**', '*kernel_trace.csv'), recursive=True)
    if not files:
        raise SystemExit('no kernel_trace.csv under %s' % trace_dir)
    rows = []
    with open(files[0]) as f:
        for r in csv.DictReader(f):
            rows.append((int(r['Start_Timestamp']), int(r['End_Timestamp']), r['Kernel_Name']))
    rows.sort()
    casc = [i for i, r in enumerate(rows) if 'k_prop' in r[2]]
    groups = []
    for i in casc:
        if groups and i - groups[-1][-1] <= 1000:
            groups[-1].append(i)
        else:
            groups.append([i])
    if len(groups) < 2:
        raise SystemExit('expected a warmup and a timed step (cascade groups: %d)' % len(groups))
    lo, hi = rows[groups[-2][-1]][1], rows[groups[-1][-1]][1]
    win = sorted((max(a, lo), min(b, hi), n) for a, b, n in rows if b > lo and a < hi)
    busy, cs, ce = 0, None, None
    for a, b, _ in win:
        if ce is None or a > ce:
            if ce is not None:
                busy += ce - cs
            cs, ce = a, b
        else:
            ce = max(ce, b)
    busy += (ce - cs) if ce is not None else 0
    agg = collections.defaultdict(lambda: [0, 0])
    for a, b, n in win:
        k = agg[n.split('(')[0]]
        k[0] += 1
        k[1] += b - a
    tot = sum(v[1] for v in agg.values())
    print('timed step %.1f ms, GPU busy %.1f ms = %.1f%%, kernel time %.1f ms'
          % ((hi - lo) / 1e6, busy / 1e6, 100.0 * busy / (hi - lo), tot / 1e6))
    for name, (c, t) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
        print('| `%s` | %d | %.1f | %.1f | %.1f%% |' % (name.replace('void ', '').replace('emu::', ''), c, t / 1e6,
                                                       t / 1e3 / c, 100.0 * t / tot))


def trace_levels(trace_dir, workdir):
    """Factorisation time and TFLOP/s by elimination-tree level from a
    rocprofv3 --kernel-trace CSV of the bench: each stream's kernels from
    k_scatter_fronts to k_triinv are one factorisation, split into levels
    (leaves first) after each level's k_extend_add launches; flops from the
    symbolic count of the bench meshes (averaged over the three)."""
    import csv
    import glob
    import collections
    from emustack_amd import _lib
    files = glob.glob(os.path.join(trace_dir, '**', '*kernel_trace.csv'), recursive=True)
    if not files:
        raise SystemExit('no kernel_trace.csv under %s' % trace_dir)
    lu = ('k_scatter_fronts', 'k_panel', 'k_trsm', 'k_panel_rows', 'k_gemm_update', 'k_extend_add')
    streams = collections.defaultdict(list)
    with open(files[0]) as f:
        for r in csv.DictReader(f):
            n = r['Kernel_Name'].split('(')[0].replace('void ', '').replace('emu::', '')
            streams[r['Stream_Id']].append((int(r['Start_Timestamp']), int(r['End_Timestamp']), n,
                                            int(r.get('Grid_Size_Y', 1) or 1)))
    facts = []
    for ks in streams.values():
        cur = None
        for a, b, n, gy in sorted(ks):
            if n.startswith('k_scatter_fronts'):
                cur = dict(nb=gy, levels=[[]])
                facts.append(cur)
            elif cur is not None and (n.startswith('k_triinv') or n.startswith('k_gather_index')):
                cur = None
            elif cur is not None and any(n.startswith(x) for x in lu):
                lv = cur['levels']
                if n.startswith('k_extend_add') and not lv[-1] and len(lv) > 1:
                    lv[-2].append((a, b))  # the level's second extend-add launch
                    continue
                lv[-1].append((a, b))
                if n.startswith('k_extend_add'):
                    lv.append([])
    fl = []
    for p in build_meshes(workdir):
        fr = _lib.MeshHandle(p).symbolic()['fronts']
        sz, u, lev = fr[:, 0].astype(float), fr[:, 1].astype(float), fr[:, 5]
        fl.append([8.0 * np.sum((sz ** 3 / 3 + sz ** 2 * u + sz * u ** 2)[lev == l]) for l in range(lev.max() + 1)])
    fl = np.mean(np.array(fl), axis=0)
    nlev = len(fl)
    acc = collections.defaultdict(lambda: [0.0, 0, 0])
    for fct in facts:
        levs = [l for l in fct['levels'] if l]
        if len(levs) != nlev:
            continue
        for i, l in enumerate(levs):
            k = acc[nlev - 1 - i]
            k[0] += l[-1][1] - l[0][0]
            k[1] += 1
            k[2] = fct['nb']
    print('| level | flops per instance | ms per factorisation | TFLOP/s |')
    tf = tw = 0.0
    for level in sorted(acc):
        w, c, nb = acc[level]
        w = w / c / 1e9
        tf += fl[level] * nb
        tw += w
        print('| %d | %.2e | %.2f | %.2f |' % (level, fl[level], w * 1e3, fl[level] * nb / w / 1e12))
    if tw:
        print('all levels: %.2e flops per %d-instance factorisation in %.1f ms = %.2f TFLOP/s (%d factorisations)'
              % (tf, nb, tw * 1e3, tf / tw / 1e12, acc[0][1]))


def pmc_summary(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir):
    """HBM bytes per dispatch of each kernel from rocprofv3 --pmc runs: the
    average FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of every kernel,
    each scaled by what the same counter reads for the --copy-calibration
    copy of known size (bytes moved / counter bytes)."""
    import csv
    import glob
    import collections

    def read(d):
        files = glob.glob(os.path.join(d, '**', '*counter_collection.csv'), recursive=True)
        if not files:
            raise SystemExit('no counter_collection.csv under %s' % d)
        agg = collections.defaultdict(lambda: [0, 0.0])
        with open(files[0]) as f:
            for r in csv.DictReader(f):
                k = r['Kernel_Name'].split('(')[0].replace('void ', '').replace('emu::', '')
                agg[k][0] += 1
                agg[k][1] += float(r['Counter_Value'])
        return agg

    copy_bytes = 2 * 2 ** 30  # bench.py --copy-calibration: a 2 GiB device copy (read 2 GiB, write 2 GiB)
    cf, cw = read(copy_fetch_dir), read(copy_write_dir)
    kf = max(cf.values(), key=lambda v: v[1])
    kw = max(cw.values(), key=lambda v: v[1])
    sf = copy_bytes / (kf[1] / kf[0] * 1024.0)
    sw = copy_bytes / (kw[1] / kw[0] * 1024.0)
    print('copy calibration: FETCH_SIZE %.0f KiB, WRITE_SIZE %.0f KiB per 2 GiB copy -> scale %.3f / %.3f'
          % (kf[1] / kf[0], kw[1] / kw[0], sf, sw))
    fe, wr = read(fetch_dir), read(write_dir)
    print('| kernel | dispatches | fetch MB / dispatch | write MB / dispatch | total GB / dispatch |')
    print('|---|---:|---:|---:|---:|')
    for name in sorted(fe, key=lambda k: -fe[k][1]):
        c, v = fe[name]
        fb = v / c * 1024.0 * sf
        wb = (wr[name][1] / wr[name][0] * 1024.0 * sw) if name in wr else 0.0
        print('| `%s` | %d | %.1f | %.1f | %.3f |' % (name, c, fb / 1e6, wb / 1e6, (fb + wb) / 1e9))


def host_cores():
    """Cores this process may use (the box grants 16 per GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, int(os.environ.get('OMP_NUM_THREADS', n)), 16))


# --------------------------------------------------------------------- main


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=2)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--batch', type=int, default=64, help='wavelength points per step')
    ap.add_argument('--groups', type=int, default=8,
                    help='concurrent sub-batches (host threads, each with its own context and stream)')
    ap.add_argument('--ncv', type=int, default=384,
                    help='Krylov dimension (0: the reference rule 2 num_BMs + num_BMs/2 + 3 = 503)')
    ap.add_argument('--workers', type=int, default=4,
                    help='host threads (contexts) for the layer sub-batches (0: groups + 1); each context '
                         'holds ~40 GB at 8 points per sub-batch, and 16 host cores serve the Schur steps')
    ap.add_argument('--precision', choices=('fp64', 'mixed'), default='fp64')
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline sample')
    ap.add_argument('--no-compare', action='store_true', help='skip the step of the other precision form')
    ap.add_argument('--copy-calibration', action='store_true',
                    help='only copy 2 GiB device to device (the reference kernel for scaling HBM counters)')
    ap.add_argument('--pmc-summary', nargs=4, metavar=('FETCH', 'WRITE', 'COPY_FETCH', 'COPY_WRITE'),
                    help='HBM bytes per dispatch per kernel from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of the '
                         'bench and of --copy-calibration')
    ap.add_argument('--trace-levels', metavar='DIR',
                    help='factorisation time and TFLOP/s per elimination-tree level from a rocprofv3 kernel trace')
    ap.add_argument('--trace-summary', metavar='DIR',
                    help='only summarise a rocprofv3 kernel trace of a --steps 1 --warmup 1 run (no GPU use)')
    ap.add_argument('--workdir', default=os.path.join(tempfile.gettempdir(), 'emustack_amd_bench_meshes'),
                    help='where the seeded bench meshes are written (regenerated when missing)')
    args = ap.parse_args()
    if args.trace_summary:
        trace_summary(args.trace_summary)
        return
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary)
        return
    if args.trace_levels:
        os.makedirs(args.workdir, exist_ok=True)
        trace_levels(args.trace_levels, args.workdir)
        return
    if args.copy_calibration:
        import torch
        x = torch.ones(1 << 27, dtype=torch.complex128, device='cuda')
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({'copy_bytes': int(x.numel() * x.element_size())}))
        return
    if args.gpus != 1:
        print('bench.py: this build runs on one GPU; the workload is independent wavelength points and '
              'would shard by wavelength across GPUs, but only --gpus 1 is measured.',
              file=sys.stderr)
        sys.exit(2)
    os.makedirs(args.workdir, exist_ok=True)
    from emustack_amd import EMUstack
    mesh_paths = build_meshes(args.workdir)
    ns, film, sub, sup = layers_for(mesh_paths)
    wavelengths = np.linspace(400.0, 1000.0, NWL)
    from emustack_amd.batch import preload_meshes
    EMUstack.context()
    for p in mesh_paths:
        EMUstack.mesh(p)
    # one-time setup outside the timed steps: every worker context loads the
    # three meshes (a first load inside a step would stall the other workers)
    preload_meshes(ns, groups=args.groups, workers=args.workers or None)
    step_i = 0

    def next_batch():
        nonlocal step_i
        idx = [(step_i * args.batch + i) % NWL for i in range(args.batch)]
        step_i += 1
        return idx
    kw = dict(groups=args.groups, ncv=args.ncv, workers=args.workers)
    for _ in range(args.warmup):
        idx = next_batch()
        print('warmup batch %s' % np.round(wavelengths[idx], 1), file=sys.stderr, flush=True)
        run_step(ns, film, sub, sup, wavelengths[idx], precision=args.precision, **kw)
    ctxs = EMUstack.all_contexts()
    for c in ctxs:
        c.profile(True, every=1)
    ops = 0
    results = {}
    t0 = time.perf_counter()
    for k in range(args.steps):
        idx = next_batch()
        res, st = run_step(ns, film, sub, sup, wavelengths[idx], precision=args.precision, **kw)
        ops += int(st[0])
        for i, r in zip(idx, res):
            results[i] = r
        print('step %d done (%.1f s)' % (k, time.perf_counter() - t0), file=sys.stderr, flush=True)
    T = time.perf_counter() - t0
    prof = {}
    for c in EMUstack.all_contexts():
        for k, v in c.profile(False).items():
            prof[k] = tuple(a + b for a, b in zip(prof.get(k, (0, 0, 0)), v))
    npts = args.steps * args.batch
    value = npts / T
    ops_per_inst = ops / (npts * 6.0)
    dom = max(prof, key=lambda k: prof[k][1])
    cnt, ms, byt = prof[dom]
    achieved = byt / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    kernel_name = {'solve_fwd': 'k_fwd', 'solve_bwd': 'k_bwd', 'gs_project': 'k_atb_mfma16<true>',
                   'gs_update': 'k_combine<16>', 'spmv_M': 'k_spmm_il'}[dom]
    m = args.ncv or 2 * NUM_BMS + NUM_BMS // 2 + 3
    m = ((m + 7) // 8) * 8
    keep = min(m - 8, NUM_BMS + (m - NUM_BMS) // 2)
    keep = m - ((m - keep + 7) // 8) * 8
    mesh_info = [symbolic_flops(p) for p in mesh_paths]
    whole = whole_step_roofline(prof, mesh_info, npts, ops_per_inst, m, keep, T)
    line = {
        'metric': 'wavelength-points/s at 8k-tri mesh; mixed vs FP64 GPU at <=1e-6 R/T/A error',
        'value': value, 'unit': 'wavelength-points/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': 1e3 * T / args.steps, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'c128 (complex64 factor reads)' if args.precision == 'mixed' else 'c128', 'data': 'synthetic',
        'config': {'workload': 'stacked gratings: 3 structured layers (r=100/150/80 nm, ~%d-triangle P2 meshes) '
                               '+ 2 films, period 500 nm, num_BMs=%d, max_order_PWs=%d (98-dim S-matrix), '
                               '%d wavelengths 400-1000 nm, normal incidence' % (NTRI, NUM_BMS, ORDER, NWL),
                   'batch_wavelengths': args.batch, 'groups': args.groups,
                   'workers': args.workers or args.groups + 1, 'krylov_dim': args.ncv or None,
                   'precision': args.precision, 'op_applications_per_instance': ops_per_inst},
        'roofline': {'bound': 'hbm', 'kernel': kernel_name, 'achieved': achieved, 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS, 'bytes': 'algorithmic',
                     'traffic': TRAFFIC_PER_DISPATCH.get(kernel_name), 'traffic_source': TRAFFIC_SOURCE,
                     # the measured per-dispatch traffic over the same launches' average time: what the
                     # kernel actually moves (algorithmic bytes plus re-reads), against the same peak
                     'traffic_rate': (TRAFFIC_PER_DISPATCH[kernel_name] / (1e-3 * ms / max(cnt, 1)) / 1e9
                                      if kernel_name in TRAFFIC_PER_DISPATCH and ms > 0 else None),
                     'launches_timed': int(cnt), 'avg_launch_us': 1e3 * ms / max(cnt, 1)},
        'whole_step_roofline': whole,
        # every timed class with its own roofline (the four largest are within a few per cent of each
        # other, so which one is 'dominant' can change from run to run)
        'kernel_profile': {k: {'launches': int(v[0]), 'ms': v[1], 'GB': v[2] / 1e9,
                               'GBps': v[2] / (v[1] * 1e-3) / 1e9 if v[1] > 0 else None,
                               'frac': v[2] / (v[1] * 1e-3) / 1e9 / HBM_PEAK_GBS if v[1] > 0 else None}
                           for k, v in prof.items()},
    }
    # the other precision form on one batch holding the fixture wavelength
    fx = np.load(FIXTURE, allow_pickle=False) if os.path.exists(FIXTURE) else None
    fix_i = int(np.argmin(np.abs(wavelengths - float(fx['wl'][0])))) if fx is not None else 64
    cmp_idx = [(fix_i + i) % NWL for i in range(args.batch)]
    if not args.no_compare:
        other = 'fp64' if args.precision == 'mixed' else 'mixed'
        mine = np.array([results[i] for i in cmp_idx]) if all(i in results for i in cmp_idx) else \
            run_step(ns, film, sub, sup, wavelengths[cmp_idx], precision=args.precision, **kw)[0]
        print('%s step on the comparison batch...' % other, file=sys.stderr, flush=True)
        t1 = time.perf_counter()
        theirs, st2 = run_step(ns, film, sub, sup, wavelengths[cmp_idx], precision=other, **kw)
        t_other = time.perf_counter() - t1
        by = {args.precision: mine, other: theirs}
        v_other = args.batch / t_other
        cmpd = {'other_form': other, 'other_value': v_other, 'other_ms_per_step': 1e3 * t_other,
                'op_applications_per_instance_other': float(st2[0]) / (args.batch * 6.0),
                'mixed_over_fp64': (value / v_other) if other == 'fp64' else (v_other / value),
                'rta_max_abs_diff_mixed_vs_fp64': float(np.abs(by['mixed'] - by['fp64']).max()),
                'rta_median_abs_diff_mixed_vs_fp64': float(np.median(np.abs(by['mixed'] - by['fp64']).max(axis=1))),
                'points_over_1e-6': int((np.abs(by['mixed'] - by['fp64']).max(axis=1) > 1e-6).sum()),
                'points_compared': args.batch}
        if fx is not None:
            ref = np.asarray(fx['rta_cluster'], float)
            for form in ('mixed', 'fp64'):
                cmpd['rta_err_vs_oracle_%s' % form] = float(np.abs(by[form][0] - ref).max())
            cmpd['oracle_point_nm'] = float(fx['wl'][0])
        line['precision_compare'] = cmpd
    if not args.no_cpu:
        cores = host_cores()
        print('cpu baseline sample on %d cores...' % cores, file=sys.stderr, flush=True)
        line['cpu_baseline'] = cpu_baseline(mesh_paths, [float(w) for w in wavelengths[cmp_idx[:4]]],
                                            inclusion_indices(), max(1, int(round(ops_per_inst))), cores)
    else:
        line['cpu_baseline'] = None
    print(json.dumps(line))


if __name__ == '__main__':
    main()
