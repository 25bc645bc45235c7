// Host orchestration of the per-wavelength mode solve on one MI355X:
// assembly -> multifrontal LU -> Krylov-Schur shift-invert eigensolve ->
// normalisation / orthogonality / J overlaps, batched over instances that
// share a mesh.  Mirrors calc_modes_2d (backend/fortran/py_calc_modes.f:1-958).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <condition_variable>
#include <mutex>
#include <numeric>
#include <random>
#include <thread>

#include <sched.h>
#include <sys/resource.h>
#include <sys/syscall.h>
#include <unistd.h>

#include "emu_host.h"
#include "engine.h"
#include "fem_basis.h"
#include "engine_api.h"

namespace emu {

// One mutex per device: the "turn" that the Krylov cycles of concurrent
// contexts on that device take for their GPU phases.  (Measured on the
// stacked-gratings bench: letting two sub-batches expand at once, or three of
// four, is no faster than strict alternation.)
std::mutex &gpu_turn(int device) {
  static std::mutex turns[64];
  return turns[device & 63];
}

typedef std::complex<double> cplx;

// ------------------------------------------------------------------ LAPACK
typedef int (*zselect_t)(cplx *);
typedef void (*zgees_t)(char *, char *, zselect_t, int *, cplx *, int *, int *, cplx *, cplx *, int *,
                        cplx *, int *, double *, int *, int *);
typedef void (*ztrsen_t)(char *, char *, int *, int *, cplx *, int *, cplx *, int *, cplx *, int *, double *,
                         double *, cplx *, int *, int *);
typedef void (*ztrevc_t)(char *, char *, int *, int *, cplx *, int *, cplx *, int *, cplx *, int *, int *, int *,
                         cplx *, double *, int *);
static zgees_t g_zgees = nullptr;
static ztrsen_t g_ztrsen = nullptr;
static ztrevc_t g_ztrevc = nullptr;

void set_lapack(void *a, void *b, void *c) {
  g_zgees = (zgees_t)a;
  g_ztrsen = (ztrsen_t)b;
  g_ztrevc = (ztrevc_t)c;
}
static int dummy_select(cplx *) { return 0; }
bool lapack_ready() { return g_zgees && g_ztrsen && g_ztrevc; }

// ------------------------------------------------------------ mesh upload
void devmesh_build(DevMesh &dm, hipStream_t s) {
  Mesh &m = dm.host;
  const Symbolic &sy = m.sym;
  const int64_t nel = m.nel, n = m.neq;
  const bool d2 = m.dim == 2;
  const int nloc = d2 ? 22 : 11;  // local DOFs per element
  if (d2) {
    dm.geo.upload(m.geo, s);
    dm.det.upload(m.det, s);
    std::vector<signed char> o(m.orient.begin(), m.orient.end());
    dm.orient.upload(o, s);
    dm.dphase.upload(m.dphase, s);
  } else {
    dm.h1d.upload(m.h1d, s);
    dm.xd1d.upload(m.xd1d, s);
    dm.tab1d.upload(m.tab1d, kTab1d, s);
  }
  {
    std::vector<int32_t> t(nel);
    int64_t mx = 0;
    for (int64_t e = 0; e < nel; ++e) {
      t[e] = (int32_t)(m.type_el[e] - 1);
      mx = std::max<int64_t>(mx, m.type_el[e]);
    }
    dm.ntyp = mx;
    dm.type_el.upload(t, s);
  }
  dm.ldof.upload(m.ldof, s);
  dm.cptr.upload(m.cptr, s);
  dm.citem.upload(m.citem, s);
  dm.row_ptr.upload(m.row_ptr, s);
  dm.col_ind.upload(m.col_ind, s);
  if (d2) {
    dm.xnod.upload(m.x, s);
    dm.table_nod.upload(m.table_nod, s);
  }
  // permuted CSR (new numbering) for M
  {
    std::vector<int64_t> prp(n + 1, 0);
    std::vector<int32_t> pci(m.nnz), psrc(m.nnz);
    std::vector<std::pair<int32_t, int32_t>> row;
    for (int64_t i = 0; i < n; ++i) {
      int32_t old = sy.perm[i];
      row.clear();
      for (int64_t k = m.row_ptr[old]; k < m.row_ptr[old + 1]; ++k)
        row.push_back({sy.iperm[m.col_ind[k]], (int32_t)k});
      std::sort(row.begin(), row.end());
      prp[i + 1] = prp[i] + (int64_t)row.size();
      for (size_t q = 0; q < row.size(); ++q) {
        pci[prp[i] + q] = row[q].first;
        psrc[prp[i] + q] = row[q].second;
      }
    }
    dm.prow_ptr.upload(prp, s);
    dm.pcol_ind.upload(pci, s);
    dm.pnnz_src.upload(psrc, s);
  }
  // transverse DOF contributors (2D: 12 edge/face DOFs per element, item
  // e*12 + l; 1D: the x- and y-component DOFs 0..6, item e*11 + l)
  {
    const int ntr = d2 ? 12 : 7, istr = d2 ? 12 : 11;
    std::vector<int32_t> cnt(n + 1, 0);
    for (int64_t e = 0; e < nel; ++e)
      for (int l = 0; l < ntr; ++l) cnt[m.ldof[nloc * e + l] + 1]++;
    std::vector<int32_t> ptr(n + 1, 0);
    for (int64_t i = 0; i < n; ++i) ptr[i + 1] = ptr[i] + cnt[i + 1];
    std::vector<int32_t> item(ptr[n]);
    std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
    for (int64_t e = 0; e < nel; ++e)
      for (int l = 0; l < ntr; ++l) item[fill[m.ldof[nloc * e + l]]++] = (int32_t)(istr * e + l);
    dm.dcptr.upload(ptr, s);
    dm.dcitem.upload(item, s);
  }
  // transverse field at the 6 P2 nodes from the 12 transverse coefficients
  if (d2) {
    std::vector<double> gt(144 * nel);
    const double xn[6][2] = {{0, 0}, {1, 0}, {0, 1}, {0.5, 0}, {0.5, 0.5}, {0, 0.5}};
    for (int64_t e = 0; e < nel; ++e) {
      const double *T = &m.geo[8 * e + 4];
      const signed char *ori = reinterpret_cast<const signed char *>(&m.orient[6 * e]);
      for (int inod = 0; inod < 6; ++inod)
        for (int l = 0; l < 12; ++l) {
          double vx, vy, c;
          transverse_basis(l, xn[inod][0], xn[inod][1], T, ori, vx, vy, c);
          gt[144 * e + (2 * inod) * 12 + l] = vx;
          gt[144 * e + (2 * inod + 1) * 12 + l] = vy;
        }
    }
    dm.gt.upload(gt, s);
  }
  // symbolic LU arrays
  dm.perm.upload(sy.perm, s);
  dm.iperm.upload(sy.iperm, s);
  const int32_t nf = (int32_t)sy.fronts.size();
  std::vector<int32_t> fs(nf), fu(nf), ff(nf), fb(nf), fp(nf), fsl(nf);
  std::vector<int64_t> foff(nf), fuoff(nf), fpiv(nf);
  for (int32_t i = 0; i < nf; ++i) {
    const FrontInfo &F = sy.fronts[i];
    fs[i] = F.s; fu[i] = F.u; ff[i] = F.f; fb[i] = F.b; fp[i] = F.parent; fsl[i] = F.slot;
    foff[i] = F.off; fuoff[i] = F.uoff; fpiv[i] = F.pivoff;
  }
  dm.fs.upload(fs, s); dm.fu.upload(fu, s); dm.ff.upload(ff, s); dm.fb.upload(fb, s);
  dm.fparent.upload(fp, s); dm.fslot.upload(fsl, s);
  dm.foff.upload(foff, s); dm.fuoff.upload(fuoff, s); dm.fpivoff.upload(fpiv, s);
  dm.uidx.upload(sy.uidx, s);
  dm.pmap.upload(sy.pmap, s);
  dm.nnz2front.upload(sy.nnz2front, s);
  dm.max_f = sy.max_f;
  // inverted-factor form: per front the offset of its forward / backward
  // blocks (2 f s entries) and of its rows in the child update-vector map
  std::vector<int64_t> io(nf), fv(nf);
  {
    int64_t acc = 0, accf = 0;
    for (int32_t i = 0; i < nf; ++i) {
      const FrontInfo &F = sy.fronts[i];
      io[i] = acc;
      fv[i] = accf;
      acc += 2 * (int64_t)F.f * F.s;
      accf += F.f;
    }
    dm.isz = std::max<int64_t>(acc, 1);
  }
  // launch schedule: deepest level first
  std::vector<int32_t> tasks;
  dm.sched.clear();
  const int NBP = 32, TL = 64, EAC = 8;
  for (int lv = (int)sy.levels.size() - 1; lv >= 0; --lv) {
    const LevelPlan &LP = sy.levels[lv];
    LevelSched L;
    L.max_f = LP.max_f;
    L.fronts_off = (int64_t)tasks.size();
    for (int32_t t : LP.fronts) tasks.push_back(t);
    L.fronts_cnt = (int64_t)LP.fronts.size();
    for (int32_t k0 = 0; k0 < LP.max_s; k0 += NBP) {
      StepPlan st;
      st.k0 = k0;
      st.panel_off = (int64_t)tasks.size();
      for (int32_t t : LP.fronts)
        if (sy.fronts[t].s > k0) tasks.push_back(t);
      st.panel_cnt = (int64_t)tasks.size() - st.panel_off;
      st.rows_off = (int64_t)tasks.size();
      for (int32_t t : LP.fronts) {
        const FrontInfo &F = sy.fronts[t];
        if (F.s <= k0) continue;
        for (int32_t r = F.s; r < F.f; r += 64) { tasks.push_back(t); tasks.push_back(r); }
      }
      st.rows_cnt = ((int64_t)tasks.size() - st.rows_off) / 2;
      st.trsm_off = (int64_t)tasks.size();
      for (int32_t t : LP.fronts) {
        const FrontInfo &F = sy.fronts[t];
        if (F.s <= k0) continue;
        int32_t ke = k0 + std::min(NBP, F.s - k0);
        for (int32_t c = ke; c < F.f; c += 64) { tasks.push_back(t); tasks.push_back(c); }
      }
      st.trsm_cnt = ((int64_t)tasks.size() - st.trsm_off) / 2;
      st.gemm_off = (int64_t)tasks.size();
      for (int32_t t : LP.fronts) {
        const FrontInfo &F = sy.fronts[t];
        if (F.s <= k0) continue;
        int32_t ke = k0 + std::min(NBP, F.s - k0);
        for (int32_t c = ke; c < F.f; c += TL)
          for (int32_t r = ke; r < F.f; r += TL) { tasks.push_back(t); tasks.push_back(r); tasks.push_back(c); }
      }
      st.gemm_cnt = ((int64_t)tasks.size() - st.gemm_off) / 3;
      L.steps.push_back(st);
    }
    for (int sl = 0; sl < 2; ++sl) {
      L.ea_off[sl] = (int64_t)tasks.size();
      for (int32_t t : LP.fronts) {
        const FrontInfo &F = sy.fronts[t];
        if (F.parent < 0 || F.slot != sl || F.u == 0) continue;
        for (int32_t c = 0; c < F.u; c += EAC) { tasks.push_back(t); tasks.push_back(c); }
      }
      L.ea_cnt[sl] = ((int64_t)tasks.size() - L.ea_off[sl]) / 2;
    }
    // solve tasks (front, r0, tiles per workgroup, k0, slot): 64-row tiles,
    // 4 per workgroup when the front has at least 4, else 2 or 1 with the
    // waves splitting each tile's columns (k0 = slot = -1).  In the split
    // variant a tile whose column range exceeds the level's span is cut into
    // tasks of `span` columns (from k0) that write partial sums to a slot,
    // and a reduction task (front, r0, first slot, count) adds them.  The
    // span is 64 columns on narrow levels (fewer than 256 unsplit tasks: the
    // few long fronts at the top of the tree, whose tasks would otherwise be
    // long serial column loops on a handful of CUs) and 256 elsewhere.  The
    // split variant serves the Krylov solves (at most 8 right-hand sides),
    // the plain one solves with many right-hand sides.
    auto push_task = [&](int32_t t, int32_t r0, int32_t tpw, int32_t kb, int32_t slot) {
      const FrontInfo &F = sy.fronts[t];
      const SolveTask T{t, r0, tpw, kb, slot, F.s, F.f, F.b, io[t], F.pivoff, fv[t], F.uoff};
      int32_t w[16];
      std::memcpy(w, &T, sizeof(T));
      tasks.insert(tasks.end(), w, w + 16);
    };
    auto solve_tasks = [&](bool fwd, bool split, SolveTasks &st) {
      int64_t unsplit = 0;
      for (int32_t t : LP.fronts) {
        const int32_t R = fwd ? sy.fronts[t].f : sy.fronts[t].s;
        const int32_t T = (R + 63) / 64;
        const int32_t tpw = T >= 4 ? 4 : (T >= 2 ? 2 : 1);
        unsplit += (T + tpw - 1) / tpw;
      }
      st.span = !split ? INT32_MAX : (unsplit < 256 ? 64 : 256);
      st.nslot = 0;
      std::vector<int32_t> red;
      while (tasks.size() % 16) tasks.push_back(-1);  // 64-byte aligned records
      st.off = (int64_t)tasks.size();
      for (int32_t t : LP.fronts) {
        const FrontInfo &F = sy.fronts[t];
        const int32_t R = fwd ? F.f : F.s;
        const int32_t kmax = fwd ? F.s : F.f;
        if (kmax > st.span) {
          for (int32_t r0 = 0; r0 < R; r0 += 64) {
            const int32_t kb = fwd ? 0 : r0, ke = fwd ? std::min(F.s, r0 + 64) : F.f;
            const int32_t nc = (ke - kb + st.span - 1) / st.span;
            if (nc <= 1) {
              push_task(t, r0, 1, kb, -1);
              continue;
            }
            red.insert(red.end(), {t, r0, st.nslot, nc});
            for (int32_t c = 0; c < nc; ++c) push_task(t, r0, 1, kb + c * st.span, st.nslot + c);
            st.nslot += nc;
          }
        } else {
          const int32_t T = (R + 63) / 64;
          const int32_t tpw = T >= 4 ? 4 : (T >= 2 ? 2 : 1);
          for (int32_t r0 = 0; r0 < R; r0 += 64 * tpw) push_task(t, r0, tpw, -1, -1);
        }
      }
      st.cnt = ((int64_t)tasks.size() - st.off) / 16;
      st.roff = (int64_t)tasks.size();
      tasks.insert(tasks.end(), red.begin(), red.end());
      st.rcnt = (int64_t)red.size() / 4;
      dm.part_rhs = std::max<int64_t>(dm.part_rhs, 64 * (int64_t)st.nslot);
    };
    for (int v = 0; v < 2; ++v) {
      solve_tasks(true, v == 1, L.fwd[v]);
      solve_tasks(false, v == 1, L.bwd[v]);
    }
    // algorithmic bytes of one solve sweep of this level (per instance):
    // triangle incl. diagonal + off-diagonal block of the inverted form once
    // per solve; b/y/x/update vectors per RHS
    for (int32_t t : LP.fronts) {
      const FrontInfo &F = sy.fronts[t];
      double ss = F.s, uu = F.u, wc = 0;
      for (int c = 0; c < F.nchild; ++c) wc += sy.fronts[F.child[c]].u;
      L.fwd_mat += 16.0 * (ss * (ss + 1) / 2 + uu * ss) + 4.0 * ss + 8.0 * (ss + uu);
      L.fwd_vec += 16.0 * (2 * ss + uu + wc);
      L.bwd_mat += 16.0 * (ss * (ss + 1) / 2 + ss * uu) + 4.0 * uu;
      L.bwd_vec += 16.0 * (2 * ss + uu);
    }
    dm.sched.push_back(L);
  }
  // inverted-factor form: offsets, child update-vector map, task lists
  {
    const int64_t accf = nf ? fv[nf - 1] + sy.fronts[nf - 1].f : 0;
    std::vector<int32_t> cm(2 * std::max<int64_t>(accf, 1), -1);
    for (int32_t i = 0; i < nf; ++i) {
      const FrontInfo &F = sy.fronts[i];
      if (F.parent < 0) continue;
      for (int32_t k = 0; k < F.u; ++k) {
        int32_t prow = sy.pmap[F.uoff + k];
        cm[2 * (fv[F.parent] + prow) + F.slot] = (int32_t)(F.uoff + k);
      }
    }
    dm.fioff.upload(io, s);
    dm.ffvoff.upload(fv, s);
    dm.cmap.upload(cm, s);
    dm.tri_off = (int64_t)tasks.size();
    for (int32_t i = 0; i < nf; ++i)
      for (int kind = 0; kind < 2; ++kind)
        for (int32_t J = 0; J < (sy.fronts[i].s + 31) / 32; ++J) {
          tasks.push_back(i); tasks.push_back(J); tasks.push_back(kind);
        }
    dm.tri_cnt = ((int64_t)tasks.size() - dm.tri_off) / 3;
    dm.wz_off = (int64_t)tasks.size();
    for (int32_t i = 0; i < nf; ++i) {
      const FrontInfo &F = sy.fronts[i];
      for (int kind = 0; kind < 2; ++kind) {
        const int32_t M = kind ? F.s : F.u, N = kind ? F.u : F.s;
        for (int32_t r0 = 0; r0 < M; r0 += TL)
          for (int32_t c0 = 0; c0 < N; c0 += TL) {
            tasks.push_back(i); tasks.push_back(kind); tasks.push_back(r0); tasks.push_back(c0);
          }
      }
    }
    dm.wz_cnt = ((int64_t)tasks.size() - dm.wz_off) / 4;
  }
  dm.tasks.upload(tasks, s);
  if (d2) launch_elem_integrals(dm, s);
  EMU_HIP(hipStreamSynchronize(s));
}

// ------------------------------------------------------------- workspace
Ctx *ctx_create(int device) {
  int ndev = 0;
  EMU_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) throw HipError("no HIP device " + std::to_string(device));
  EMU_HIP(hipSetDevice(device));
  Ctx *c = new Ctx();
  c->device = device;
  EMU_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  noalloc_scratch(device);  // per-device address of the fixed scratch, with the device current
  return c;
}
void ctx_destroy(Ctx *c) { delete c; }

void profile(Ctx *c, int enable, int every, double *out) {
  if (enable == 1) {
    c->prof.reset();
    c->prof.on = true;
    c->prof.every = std::max(1, every);
  } else if (enable == 0) {
    c->prof.on = false;
  }
  EMU_HIP(hipStreamSynchronize(c->stream));
  c->prof.collect();
  if (out)
    for (int k = 0; k < 8; ++k)
      for (int q = 0; q < 3; ++q) out[3 * k + q] = c->prof.stat[k][q];
}

DevMesh *mesh_load(Ctx *c, const std::string &path, int leaf) {
  EMU_HIP(hipSetDevice(c->device));
  DevMesh *dm = new DevMesh();
  try {
    build_mesh(dm->host, path);
    build_symbolic(dm->host, leaf > 0 ? leaf : 96);
    devmesh_build(*dm, c->stream);
  } catch (...) {
    delete dm;
    throw;
  }
  return dm;
}

// ---------------------------------------------------------------- assembly
__global__ void k_store_norm(int nb, const double *__restrict__ nrm, zd *__restrict__ dst, int64_t sd) {
  int b = threadIdx.x + blockIdx.x * blockDim.x;
  if (b < nb) dst[b * sd] = zmk(nrm[b], 0.0);
}

static void upload_params(Ctx *c, DevMesh &dm, int nb, int ntyp, const double *pp, const double *qq,
                          const double *shift, const double *bloch, zd **dpp, zd **dqq, zd **dsh, double **dbl) {
  *dpp = c->buf<zd>("pp", (size_t)nb * ntyp);
  *dqq = c->buf<zd>("qq", (size_t)nb * ntyp);
  *dsh = c->buf<zd>("shift", nb);
  *dbl = c->buf<double>("bloch", 2 * (size_t)nb);
  EMU_HIP(hipMemcpyAsync(*dpp, pp, sizeof(zd) * nb * ntyp, hipMemcpyHostToDevice, c->stream));
  EMU_HIP(hipMemcpyAsync(*dqq, qq, sizeof(zd) * nb * ntyp, hipMemcpyHostToDevice, c->stream));
  EMU_HIP(hipMemcpyAsync(*dsh, shift, sizeof(zd) * nb, hipMemcpyHostToDevice, c->stream));
  EMU_HIP(hipMemcpyAsync(*dbl, bloch, sizeof(double) * 2 * nb, hipMemcpyHostToDevice, c->stream));
  EMU_HIP(hipStreamSynchronize(c->stream));
}

void assemble(Ctx *c, DevMesh &dm, int nb, int ntyp, const double *pp, const double *qq, const double *shift,
              const double *bloch, double *mat1, double *mat2) {
  EMU_HIP(hipSetDevice(c->device));
  if (ntyp < dm.ntyp) throw std::invalid_argument("nb_typ_el smaller than the mesh's material count");
  zd *dpp, *dqq, *dsh;
  double *dbl;
  upload_params(c, dm, nb, ntyp, pp, qq, shift, bloch, &dpp, &dqq, &dsh, &dbl);
  const int64_t nnz = dm.host.nnz;
  zd *ve = c->buf<zd>("ve", (size_t)nb * dm.host.nel * 14);
  zd *m1 = c->buf<zd>("mat1", (size_t)nb * nnz);
  zd *m2 = c->buf<zd>("mat2", (size_t)nb * nnz);
  if (dm.host.dim == 1) {
    launch_gather_1d(dm, nb, ntyp, dpp, dqq, dsh, dbl, 0, m1, m2, c->stream);
  } else {
    launch_phase(dm, nb, dbl, ve, c->stream);
    launch_gather(dm, nb, ntyp, dpp, dqq, dsh, ve, m1, m2, c->stream);
  }
  if (mat1) EMU_HIP(hipMemcpyAsync(mat1, m1, sizeof(zd) * nb * nnz, hipMemcpyDeviceToHost, c->stream));
  if (mat2) EMU_HIP(hipMemcpyAsync(mat2, m2, sizeof(zd) * nb * nnz, hipMemcpyDeviceToHost, c->stream));
  EMU_HIP(hipStreamSynchronize(c->stream));
}

void lu_solve_host(Ctx *c, DevMesh &dm, int nb, const double *mat1, int nrhs, const double *b, double *x,
                   int64_t *stats) {
  EMU_HIP(hipSetDevice(c->device));
  const Symbolic &sy = dm.host.sym;
  const int64_t nnz = dm.host.nnz, n = dm.host.neq;
  zd *m1 = c->buf<zd>("lu_mat1", (size_t)nb * nnz);
  EMU_HIP(hipMemcpyAsync(m1, mat1, sizeof(zd) * nb * nnz, hipMemcpyHostToDevice, c->stream));
  EMU_HIP(hipStreamSynchronize(c->stream));
  zd *pool = c->buf<zd>("pool", (size_t)nb * sy.pool);
  int32_t *piv = c->buf<int32_t>("piv", (size_t)nb * sy.npiv);
  int32_t *pfl = c->buf<int32_t>("pflag", (size_t)nb * sy.npiv);
  int32_t *sig = c->buf<int32_t>("sig", (size_t)nb * sy.npiv);
  int32_t *gx = c->buf<int32_t>("gx", (size_t)nb * sy.npiv * 4);
  zd *inv = c->buf<zd>("inv", (size_t)nb * dm.isz);
  lu_factor(dm, nb, m1, pool, piv, pfl, sig, gx, inv, c->stream);
  zd *bo = c->buf<zd>("lu_b", (size_t)nb * nrhs * n);
  zd *bn = c->buf<zd>("lu_bn", (size_t)nb * nrhs * n);
  zd *yn = c->buf<zd>("lu_y", (size_t)nb * nrhs * n);
  zd *xn = c->buf<zd>("lu_x", (size_t)nb * nrhs * n);
  zd *w = c->buf<zd>("lu_w", (size_t)nb * nrhs * solve_ws(dm));
  EMU_HIP(hipMemcpyAsync(bo, b, sizeof(zd) * nb * nrhs * n, hipMemcpyHostToDevice, c->stream));
  EMU_HIP(hipStreamSynchronize(c->stream));
  launch_permute_vec(dm, nb, nrhs, bo, bn, true, c->stream);
  lu_solve(dm, nb, nrhs, inv, gx, bn, yn, xn, w, c->stream);
  launch_permute_vec(dm, nb, nrhs, xn, bo, false, c->stream);
  EMU_HIP(hipMemcpyAsync(x, bo, sizeof(zd) * nb * nrhs * n, hipMemcpyDeviceToHost, c->stream));
  std::vector<int32_t> pf((size_t)nb * sy.npiv);
  EMU_HIP(hipMemcpyAsync(pf.data(), pfl, sizeof(int32_t) * pf.size(), hipMemcpyDeviceToHost, c->stream));
  EMU_HIP(hipStreamSynchronize(c->stream));
  if (stats) stats[0] = std::accumulate(pf.begin(), pf.end(), (int64_t)0);
}

// ------------------------------------------------------ Krylov-Schur solve
// Block Krylov-Schur (block size bq) for the nev largest-|mu| eigenpairs of
// OP = (K - shift M)^-1 M, the operator ARPACK iterates on in the reference
// (valpr_umf_64.f:291-459, which = 'LM').  One block step applies OP to bq
// vectors (one multi-RHS solve), orthogonalises them against the basis with
// two block classical Gram-Schmidt passes and Cholesky-QR2 inside the block.
// The projected (m x m) Schur problem is solved on the host with LAPACK.
//
// Relation after the expansion:  OP V[:, :m] = V[:, :m+bq] H[:m+bq, :m].
struct KSState {
  std::vector<cplx> H;     // (m+bq) x m, column-major, ld = m+bq
  std::vector<cplx> T, Q;  // m x m Schur form and vectors
  std::vector<cplx> theta;
  int nlock = 0;           // leading Schur vectors locked (their residual coupling is zero in H)
  int nlock_next = 0;      // locked after this projection (applied at the restart)
  std::vector<zd> Qk, Hk;  // restart blocks: Q[:, :keep] and the restarted H (filled by ks_restart_blocks)
  std::vector<zd> Wz;      // Ritz vectors in the basis (m x nev) once converged
};

// The restart blocks of one instance, computed on its host thread right after
// its Schur step: Q[:, :keep] (m x keep) and the restarted projected matrix
// (ld x m: T[:keep, :keep] on top, the residual coupling B Q[:, :keep] in rows
// keep..keep+bq, zero for the locked columns).
static void ks_restart_blocks(KSState &S, int m, int bq, int keep, int ld, int64_t sH) {
  S.Qk.assign((size_t)m * keep, zmk(0, 0));
  S.Hk.assign((size_t)sH, zmk(0, 0));
  for (int j = 0; j < keep; ++j)
    for (int i = 0; i < m; ++i) {
      const cplx q = S.Q[i + (size_t)j * m];
      S.Qk[i + (size_t)j * m] = zmk(q.real(), q.imag());
    }
  for (int j = 0; j < keep; ++j) {
    for (int i = 0; i <= j; ++i) {
      const cplx t = S.T[i + (size_t)j * m];
      S.Hk[i + (size_t)j * ld] = zmk(t.real(), t.imag());
    }
    cplx acc[16] = {};
    if (bq > 16) throw std::logic_error("restart blocks: block size above 16");
    if (j >= S.nlock_next)  // locked: coupling deflated to zero
      for (int i = 0; i < m; ++i) {
        const cplx q = S.Q[i + (size_t)j * m];
        const cplx *hr = &S.H[m + (size_t)i * ld];
        for (int r = 0; r < bq; ++r) acc[r] += hr[r] * q;
      }
    for (int r = 0; r < bq; ++r) S.Hk[(keep + r) + (size_t)j * ld] = zmk(acc[r].real(), acc[r].imag());
  }
}

// Schur-decompose H[:m,:m], order the `keep` largest |theta| first and the
// `nev` largest among them first of all, test the residuals
// ||B Q y_i|| <= tol max(eps^(2/3), |theta_i|) of the leading nev Ritz pairs (B = last bq rows
// of H).  Optionally return W = Q[:, :nev] Y (Ritz vectors in the basis),
// always or only when converged.
//
// Locking: the first st.nlock columns of H are a converged upper-triangular
// Schur block whose coupling to the next block row was zeroed at the last
// restart, so H = [T_LL X; 0 H_AA] and only the active block H_AA needs a
// Schur decomposition (the expensive O(m^3) host step); T_LA = X Q_AA.  Schur
// vectors of the wanted set whose own residual ||B q_i|| is below tol |theta_i|
// (a leading run) are locked for the next cycle.
// ARPACK's convergence scale (znaup2): a Ritz value's residual estimate is
// measured against max(eps^(2/3), |theta|), so values near zero are not held
// to an unreachable relative bound.
static const double kEps23 = 3.6668528625e-11;  // (2^-53)^(2/3)

static bool ks_project(KSState &st, int m, int bq, int nev, int keep, double tol, std::vector<cplx> *ritzvec,
                       double *maxres, bool ritz_always = true) {
  const int ld = m + bq;
  const int a0 = std::min(st.nlock, nev), na = m - a0;
  std::vector<cplx> TA((size_t)na * na), QA((size_t)na * na);
  for (int j = 0; j < na; ++j)
    for (int i = 0; i < na; ++i) TA[i + (size_t)j * na] = st.H[(a0 + i) + (size_t)(a0 + j) * ld];
  std::vector<cplx> thA(na);
  int n = na, lda = na, sdim = 0, ldvs = na, info = 0, lwork = -1;
  char jobvs = 'V', sort = 'N';
  std::vector<double> rwork(na);
  std::vector<int> bwork(na);
  cplx wq;
  g_zgees(&jobvs, &sort, dummy_select, &n, TA.data(), &lda, &sdim, thA.data(), QA.data(), &ldvs, &wq, &lwork,
          rwork.data(), bwork.data(), &info);
  lwork = std::max(1, (int)wq.real());
  std::vector<cplx> work(lwork);
  g_zgees(&jobvs, &sort, dummy_select, &n, TA.data(), &lda, &sdim, thA.data(), QA.data(), &ldvs, work.data(),
          &lwork, rwork.data(), bwork.data(), &info);
  if (info != 0) throw std::runtime_error("zgees failed, info=" + std::to_string(info));
  st.T.assign((size_t)m * m, 0);
  st.Q.assign((size_t)m * m, 0);
  for (int j = 0; j < a0; ++j) {
    for (int i = 0; i <= j; ++i) st.T[i + (size_t)j * m] = st.H[i + (size_t)j * ld];
    st.Q[j + (size_t)j * m] = 1.0;
  }
  for (int j = 0; j < na; ++j) {
    // T[:a0, a0 + j] = H[:a0, a0:m] QA[:, j], by columns of H (contiguous)
    cplx *tc = &st.T[(size_t)(a0 + j) * m];
    for (int k = 0; k < na; ++k) {
      const cplx q = QA[k + (size_t)j * na];
      const cplx *hc = &st.H[(size_t)(a0 + k) * ld];
      for (int i = 0; i < a0; ++i) tc[i] += hc[i] * q;
    }
    for (int i = 0; i < na; ++i) {
      st.T[(a0 + i) + (size_t)(a0 + j) * m] = TA[i + (size_t)j * na];
      st.Q[(a0 + i) + (size_t)(a0 + j) * m] = QA[i + (size_t)j * na];
    }
  }
  st.theta.assign(m, 0);
  for (int i = 0; i < m; ++i) st.theta[i] = st.T[i + (size_t)i * m];
  n = m;
  lda = m;
  ldvs = m;
  auto reorder = [&](int count) {
    // locked columns stay selected (and first); the rest by largest |theta|
    std::vector<int> idx(na);
    std::iota(idx.begin(), idx.end(), a0);
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int a, int b) { return std::abs(st.theta[a]) > std::abs(st.theta[b]); });
    std::vector<int> sel(m, 0);
    for (int i = 0; i < a0; ++i) sel[i] = 1;
    for (int i = 0; i < count - a0; ++i) sel[idx[i]] = 1;
    char job = 'N', compq = 'V';
    int mm = 0, lw = std::max(1, m), inf = 0;
    double sv = 0, sep = 0;
    std::vector<cplx> wk(lw);
    std::vector<cplx> w(m);
    g_ztrsen(&job, &compq, sel.data(), &n, st.T.data(), &lda, st.Q.data(), &ldvs, w.data(), &mm, &sv, &sep,
             wk.data(), &lw, &inf);
    if (inf != 0) throw std::runtime_error("ztrsen failed, info=" + std::to_string(inf));
    st.theta = w;
  };
  reorder(keep);
  reorder(nev);
  std::vector<cplx> Tn((size_t)nev * nev);
  for (int j = 0; j < nev; ++j)
    for (int i = 0; i < nev; ++i) Tn[i + (size_t)j * nev] = st.T[i + (size_t)j * m];
  std::vector<cplx> VR((size_t)nev * nev), VL(1), wk(2 * nev);
  std::vector<double> rw(nev);
  char side = 'R', howmny = 'A';
  int nn = nev, ldt = nev, ldvl = 1, ldvr = nev, mmv = nev, mout = 0, inf = 0;
  g_ztrevc(&side, &howmny, nullptr, &nn, Tn.data(), &ldt, VL.data(), &ldvl, VR.data(), &ldvr, &mmv, &mout,
           wk.data(), rw.data(), &inf);
  if (inf != 0) throw std::runtime_error("ztrevc failed");
  // BQ = B Q[:, :nev]  (bq x nev)
  std::vector<cplx> BQ((size_t)bq * nev, 0);
  for (int j = 0; j < nev; ++j)
    for (int i = 0; i < m; ++i) {
      cplx qv = st.Q[i + (size_t)j * m];
      if (qv == 0.0) continue;
      for (int r = 0; r < bq; ++r) BQ[r + (size_t)j * bq] += st.H[(m + r) + (size_t)i * ld] * qv;
    }
  // residual bound of pair k: tol relative to |theta_k|
  auto bound = [&](int k) { return tol * std::max(kEps23, std::abs(st.theta[k])); };
  // locking candidates: leading run of Schur vectors with small own residual
  st.nlock_next = 0;
  for (int k = 0; k < nev; ++k) {
    double r2 = 0;
    for (int r = 0; r < bq; ++r) r2 += std::norm(BQ[r + (size_t)k * bq]);
    if (!(std::sqrt(r2) <= bound(k))) break;
    st.nlock_next = k + 1;
  }
  bool conv = true;
  double worst = 0;
  for (int k = 0; k < nev; ++k) {
    double ny = 0, res2 = 0;
    for (int j = 0; j <= k; ++j) ny += std::norm(VR[j + (size_t)k * nev]);
    for (int r = 0; r < bq; ++r) {
      cplx acc = 0;
      for (int j = 0; j <= k; ++j) acc += BQ[r + (size_t)j * bq] * VR[j + (size_t)k * nev];
      res2 += std::norm(acc);
    }
    const double res = std::sqrt(res2 / ny);
    worst = std::max(worst, res / std::max(std::abs(st.theta[k]), kEps23));
    if (!(res <= bound(k))) conv = false;
  }
  if (maxres) *maxres = worst;
  if (ritzvec && (conv || ritz_always)) {
    ritzvec->assign((size_t)m * nev, 0);
    for (int k = 0; k < nev; ++k)
      for (int j = 0; j <= k; ++j) {
        cplx v = VR[j + (size_t)k * nev];
        for (int i = 0; i < m; ++i) (*ritzvec)[i + (size_t)k * m] += st.Q[i + (size_t)j * m] * v;
      }
  }
  return conv;
}

struct EigResult {
  std::vector<cplx> mu;  // nb x nev
  int64_t ops = 0, restarts = 0, steps = 0, rank_fixes = 0;
};

// Opt-in stage timer (environment EMU_TIMING=1): synchronises the stream at
// stage boundaries and prints wall times to stderr.  Off by default because
// the synchronisations serialise host and device.
struct StageTimer {
  bool on = false;
  hipStream_t s = nullptr;
  std::chrono::steady_clock::time_point t0;
  std::vector<std::pair<std::string, double>> acc;
  explicit StageTimer(hipStream_t st) : s(st) {
    const char *e = getenv("EMU_TIMING");
    on = e && e[0] == '1';
    t0 = std::chrono::steady_clock::now();
  }
  void mark(const char *name) {
    if (!on) return;
    (void)hipStreamSynchronize(s);
    auto t = std::chrono::steady_clock::now();
    double dt = std::chrono::duration<double>(t - t0).count();
    t0 = t;
    for (auto &p : acc)
      if (p.first == name) { p.second += dt; return; }
    acc.push_back({name, dt});
  }
  ~StageTimer() {
    if (!on) return;
    double tot = 0;
    for (auto &p : acc) tot += p.second;
    fprintf(stderr, "[emu timing]");
    for (auto &p : acc) fprintf(stderr, " %s=%.3fs", p.first.c_str(), p.second);
    fprintf(stderr, " total=%.3fs\n", tot);
  }
};

// orthonormalise V[:, j:j+bq] in place (Cholesky-QR2); R (bq x bq) of the
// factorisation W = Q R is written to Rout[b] (device, bq x bq) when given;
// xi (8-column blocks): also write the row-interleaved copy of the result
static void block_cholqr2(Ctx *c, int nb, int64_t n, int bq, zd *W, int64_t sV, zd *G, zd *Ri, zd *R1, zd *Rout,
                          int32_t *bad, zd *atbs, size_t atb_elems, hipStream_t s, zd *xi = nullptr) {
  const int64_t sG = (int64_t)bq * bq;
  launch_atb(nb, n, bq, bq, W, n, sV, W, n, sV, G, bq, sG, true, atbs, atb_elems, s);
  launch_cholqr(nb, bq, G, Ri, nullptr, R1, bad, s);
  launch_trmm_right(nb, n, bq, W, n, sV, Ri, W, n, sV, s);
  launch_atb(nb, n, bq, bq, W, n, sV, W, n, sV, G, bq, sG, true, atbs, atb_elems, s);
  launch_cholqr(nb, bq, G, Ri, R1, Rout, bad, s);
  launch_trmm_right(nb, n, bq, W, n, sV, Ri, W, n, sV, s, xi);
}

// First-come admission of host Schur steps to the cores: at most one running
// step per available CPU, admitted in ticket order.  Sub-batches that reach
// their Schur phase together would otherwise share the cores and all finish
// late; in arrival order the first sub-batch's steps finish first and its
// next GPU phase starts while the others compute (the total CPU work is the
// same).
class CoreGate {
 public:
  CoreGate() {
    cpu_set_t cs;
    cap_ = sched_getaffinity(0, sizeof(cs), &cs) == 0 ? std::max(1, CPU_COUNT(&cs)) : 8;
  }
  uint64_t enter() {
    std::unique_lock<std::mutex> lk(mu_);
    const uint64_t t = next_++;
    cv_.wait(lk, [&] { return t < done_ + cap_; });
    return t;
  }
  void leave() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      ++done_;
    }
    cv_.notify_all();
  }

 private:
  std::mutex mu_;
  std::condition_variable cv_;
  uint64_t next_ = 0, done_ = 0;
  uint64_t cap_ = 8;
};
static CoreGate &core_gate() {
  static CoreGate g;
  return g;
}

static EigResult krylov_schur(Ctx *c, DevMesh &dm, int nb, int nev, int m, int bq, int itermax, double tol,
                              const zd *inv, const float2 *inv32, const int32_t *gx, const zd *mvals, zd *Xout,
                              StageTimer &tm) {
  hipStream_t s = c->stream;
  const Symbolic &sy = dm.host.sym;
  const int64_t n = dm.host.neq;
  const int ld = m + bq;                     // rows of H
  const int64_t sV = (int64_t)(m + bq) * n;  // basis stride per instance
  const int64_t sH = (int64_t)ld * m;
  zd *V = c->buf<zd>("ks_V", (size_t)nb * sV);
  zd *V2 = c->buf<zd>("ks_V2", (size_t)nb * sV);
  zd *t1 = c->buf<zd>("ks_t1", (size_t)nb * bq * n);
  zd *ty = c->buf<zd>("ks_ty", (size_t)nb * bq * n);
  zd *xil = bq == 8 ? c->buf<zd>("ks_xil", (size_t)nb * 8 * n) : nullptr;  // interleaved SpMM input
  zd *tw = c->buf<zd>("ks_w", (size_t)nb * bq * solve_ws(dm));
  zd *Hd = c->buf<zd>("ks_H", (size_t)nb * sH);
  zd *H2 = c->buf<zd>("ks_H2", (size_t)nb * ld * bq);
  zd *G = c->buf<zd>("ks_G", (size_t)nb * bq * bq);
  zd *Ri = c->buf<zd>("ks_Ri", (size_t)nb * bq * bq);
  zd *R1 = c->buf<zd>("ks_R1", (size_t)nb * bq * bq);
  zd *Rf = c->buf<zd>("ks_Rf", (size_t)nb * bq * bq);
  int32_t *bad = c->buf<int32_t>("ks_bad", nb);
  size_t atb_elems = (size_t)nb * (m + bq) * std::max(bq, 32) * 128;
  zd *atbs = c->buf<zd>("ks_atb", atb_elems);
  zd *Qd = c->buf<zd>("ks_Q", (size_t)nb * m * m);
  // delayed block Gram-Schmidt (block size 8): projections and Gram matrix
  // P = [V Z W]^H [Z W], A0 = H S, combination E, L
  const bool delayed = bq == 8;
  const int64_t ldp = m + bq, sP = ldp * 2 * bq, sA0 = ldp * bq;
  zd *Pd = delayed ? c->buf<zd>("ks_P", (size_t)nb * sP) : nullptr;
  zd *A0d = delayed ? c->buf<zd>("ks_A0", (size_t)nb * sA0) : nullptr;
  zd *Ed = delayed ? c->buf<zd>("ks_E", (size_t)nb * sP) : nullptr;
  zd *Ld = delayed ? c->buf<zd>("ks_L", (size_t)nb * 4 * bq * bq) : nullptr;
  EMU_HIP(hipMemsetAsync(bad, 0, sizeof(int32_t) * nb, s));
  // start block: deterministic pseudo-random, orthonormalised
  {
    std::vector<zd> v0((size_t)nb * bq * n);
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    for (auto &z : v0) z = zmk(U(rng), U(rng));
    EMU_HIP(hipMemcpyAsync(t1, v0.data(), sizeof(zd) * v0.size(), hipMemcpyHostToDevice, s));
    EMU_HIP(hipStreamSynchronize(s));
    launch_copy(nb, bq * n, t1, bq * n, V, sV, s);
    block_cholqr2(c, nb, n, bq, V, sV, G, Ri, R1, Rf, bad, atbs, atb_elems, s);
  }
  launch_fill(Hd, (int64_t)nb * sH, zmk(0, 0), s);
  std::vector<KSState> st(nb);
  for (auto &x : st) x.H.assign((size_t)sH, 0);
  EigResult res;
  res.mu.assign((size_t)nb * nev, 0);
  std::vector<int> schur_na;  // active Schur sizes, recorded when the stage timer is on
  int k = 0;
  int keep = std::min(m - bq, nev + (m - nev) / 2);
  keep = m - ((m - keep + bq - 1) / bq) * bq;  // (m - keep) multiple of bq
  if (keep < nev) throw std::invalid_argument("Krylov subspace too small for num_BMs");
  std::vector<zd> Hh((size_t)nb * sH);
  std::vector<std::vector<cplx>> ritz(nb);
  Profiler *pf = c->prof.on ? &c->prof : nullptr;
  const double nd = (double)n, nzd = (double)dm.host.nnz;
  // Active instances occupy slots 0..nact-1; slot_inst maps a slot to its
  // instance; converged instances are retired by swapping their slot with
  // the last active one (device data and host state), logged in swaps
  int nact = nb;
  std::vector<int> slot_inst(nb);
  std::iota(slot_inst.begin(), slot_inst.end(), 0);
  std::vector<std::pair<int, int>> swaps;
  zd *inv_m = const_cast<zd *>(inv);          // the caller's buffers: slots are swapped during the
  int32_t *gx_m = const_cast<int32_t *>(gx);  // solve and restored in order before returning
  zd *mv_m = const_cast<zd *>(mvals);
  float2 *inv32_m = const_cast<float2 *>(inv32);
  const int64_t nnzm = dm.host.nnz;
  auto swap_caller = [&](int a1, int a2) {
    launch_swap(dm.isz, inv_m + a1 * dm.isz, inv_m + a2 * dm.isz, s);
    if (inv32_m) launch_swap_f2(dm.isz, inv32_m + a1 * dm.isz, inv32_m + a2 * dm.isz, s);
    launch_swap_i32(4 * sy.npiv, gx_m + 4 * a1 * sy.npiv, gx_m + 4 * a2 * sy.npiv, s);
    launch_swap(nnzm, mv_m + a1 * nnzm, mv_m + a2 * nnzm, s);
  };
  auto swap_slots = [&](int a1, int a2) {
    swap_caller(a1, a2);
    launch_swap(sV, V + a1 * sV, V + a2 * sV, s);
    launch_swap(sH, Hd + a1 * sH, Hd + a2 * sH, s);
    launch_swap_i32(1, bad + a1, bad + a2, s);
  };
  zd *Wd = c->buf<zd>("ks_W", (size_t)nb * m * nev);
  // The expansion of a cycle holds the device's turn; the host Schur phase
  // and the restart products release it, so that concurrent sub-batches on
  // other contexts alternate their expansions on the GPU instead of expanding
  // together and then leaving the GPU idle during their Schur phases.
  std::unique_lock<std::mutex> turn(gpu_turn(c->device), std::defer_lock);
  for (int it = 0; it <= itermax; ++it) {
    if (!turn.owns_lock()) turn.lock();
    const double nbd = nact;
    // xil holds the interleaved copy of the block to expand once the
    // previous block's Cholesky-QR has written it (not for the first block
    // of an expansion: the restart or the start vectors wrote that one)
    bool xil_ready = false;
    for (int j = k; j < m; j += bq) {
      zd *Xj = V + (int64_t)j * n;
      zd *Wn = V + (int64_t)(j + bq) * n;  // next block slot
      if (pf) pf->start_sweep();
      const KTime kspmv = pf ? pf->time(4, nbd * (20.0 * nzd + 8.0 * (nd + 1) + 32.0 * nd * bq)) : KTime{};
      if (bq == 8)
        launch_spmm8_il(nact, n, dm.prow_ptr.p, dm.pcol_ind.p, mvals, dm.host.nnz, Xj, sV, xil, t1, (int64_t)bq * n,
                        s, kspmv, xil_ready);
      else
        launch_spmm(nact, n, bq, dm.prow_ptr.p, dm.pcol_ind.p, mvals, dm.host.nnz, Xj, sV, t1, (int64_t)bq * n, s,
                    kspmv);
      // the solve writes the next block slot of the basis directly
      if (inv32)
        lu_solve(dm, nact, bq, inv32, gx, t1, ty, Wn, tw, s, pf, sV);
      else
        lu_solve(dm, nact, bq, inv, gx, t1, ty, Wn, tw, s, pf, sV);
      if (delayed) {
        // Xj is the provisional block Z; one pass over V[:, :j] for Z's
        // reorthogonalisation and W's projection, then [Q_j, Zraw] in place
        const int c0 = j == k ? 0 : j - bq;
        // algorithmic bytes: one read of [V Z W] (j + 2 bq columns; [Z W] is
        // a slice of it); the combine reads V[:, :j] and [Z W] and writes [Z W]
        launch_atb(nact, n, j + 2 * bq, 2 * bq, V, n, sV, Xj, n, sV, Pd, ldp, sP, true, atbs, atb_elems, s,
                   pf ? pf->time(2, nbd * 16.0 * nd * (j + 2 * bq)) : KTime{});
        if (j > 0) launch_gemm_nn(nact, j, bq, j, zmk(1, 0), Hd, ld, sH, Pd, ldp, sP, zmk(0, 0), A0d, ldp, sA0, s);
        launch_dcgs_coef(nact, false, j, c0, Pd, ldp, sP, A0d, ldp, sA0, Hd, ld, sH, Ed, sP, Ld, bad, s);
        launch_combine(nact, n, j, 2 * bq, V, n, sV, Ed, sP, Ld, Xj, n, sV, s,
                       pf ? pf->time(3, nbd * 16.0 * nd * (j + 4 * bq)) : KTime{});
        block_cholqr2(c, nact, n, bq, Wn, sV, G, Ri, R1, Rf, bad, atbs, atb_elems, s, bq == 8 ? xil : nullptr);
        xil_ready = bq == 8;
        launch_block_copy(nact, bq, bq, Rf, bq, (int64_t)bq * bq, Hd + (int64_t)j * ld + j + bq, ld, sH, false, s);
        res.ops += (int64_t)nact * bq;
        res.steps += 1;
        continue;
      }
      // block CGS, two passes
      for (int pass = 0; pass < 2; ++pass) {
        zd *hc = pass == 0 ? Hd + (int64_t)j * ld : H2;
        int64_t hld = pass == 0 ? ld : ld, hs = pass == 0 ? sH : (int64_t)ld * bq;
        launch_atb(nact, n, j + bq, bq, V, n, sV, Wn, n, sV, hc, hld, hs, true, atbs, atb_elems, s,
                   pf ? pf->time(2, nbd * 16.0 * nd * (j + 2 * bq)) : KTime{});
        launch_sub_vh_blk(nact, n, j + bq, bq, V, n, sV, hc, hld, hs, Wn, n, sV, s,
                          pf ? pf->time(3, nbd * 16.0 * nd * (j + 3 * bq)) : KTime{});
      }
      launch_block_copy(nact, j + bq, bq, H2, ld, (int64_t)ld * bq, Hd + (int64_t)j * ld, ld, sH, true, s);
      block_cholqr2(c, nact, n, bq, Wn, sV, G, Ri, R1, Rf, bad, atbs, atb_elems, s);
      launch_block_copy(nact, bq, bq, Rf, bq, (int64_t)bq * bq, Hd + (int64_t)j * ld + j + bq, ld, sH, false, s);
      res.ops += (int64_t)nact * bq;
      res.steps += 1;
    }
    if (delayed) {
      // reorthogonalise the last provisional block (the residual block of the
      // Krylov-Schur relation) and correct row block m of H
      launch_atb(nact, n, m + bq, bq, V, n, sV, V + (int64_t)m * n, n, sV, Pd, ldp, sP, true, atbs, atb_elems, s);
      launch_dcgs_coef(nact, true, m, m - bq, Pd, ldp, sP, nullptr, ldp, sA0, Hd, ld, sH, Ed, sP, Ld, bad, s);
      launch_combine(nact, n, m, bq, V, n, sV, Ed, sP, Ld, V + (int64_t)m * n, n, sV, s);
    }
    tm.mark("krylov_expand");
    EMU_HIP(hipMemcpyAsync(Hh.data(), Hd, sizeof(zd) * (size_t)nact * sH, hipMemcpyDeviceToHost, s));
    EMU_HIP(hipStreamSynchronize(s));
    turn.unlock();
    if (c->prof.on) c->prof.collect();
    const bool last = (it == itermax);
    std::vector<std::thread> th;
    std::mutex schur_mu;
    std::vector<std::string> errs(nact);
    std::vector<char> convflag(nact, 0);
    for (int b = 0; b < nact; ++b) {
      th.emplace_back([&, b]() {
        // The host Schur steps of concurrent sub-batches can occupy every
        // core; at lower priority they leave the threads that launch GPU
        // work runnable at once (a launch thread waiting for a core leaves
        // the GPU idle).  Raising one's own nice value needs no privilege.
        (void)setpriority(PRIO_PROCESS, (id_t)syscall(SYS_gettid), 10);
        core_gate().enter();
        struct Leave {
          ~Leave() { core_gate().leave(); }
        } leave_gate;
        try {
          KSState &S = st[b];
          if (tm.on) {  // active Schur sizes (printed with the stage times)
            std::lock_guard<std::mutex> g(schur_mu);
            schur_na.push_back(m - std::min(S.nlock, nev));
          }
          for (int64_t i = 0; i < sH; ++i) S.H[i] = cplx(Hh[(size_t)b * sH + i].x, Hh[(size_t)b * sH + i].y);
          // Ritz vectors of the basis are formed once the pairs pass the test
          convflag[b] = ks_project(S, m, bq, nev, keep, tol, &ritz[b], nullptr, false) ? 1 : 0;
          // the restart's host part here, in parallel, not on the launching
          // thread while it holds the GPU turn
          if (!convflag[b] && !last) ks_restart_blocks(S, m, bq, keep, ld, sH);
          if (convflag[b]) {
            const std::vector<cplx> &W = ritz[b];
            S.Wz.resize(W.size());
            for (size_t i = 0; i < W.size(); ++i) S.Wz[i] = zmk(W[i].real(), W[i].imag());
          }
        } catch (std::exception &e) {
          errs[b] = e.what();
        }
      });
    }
    for (auto &t : th) t.join();
    for (int b = 0; b < nact; ++b)
      if (!errs[b].empty()) throw std::runtime_error(errs[b]);
    res.restarts = it;
    tm.mark("host_schur");
    // The Ritz-vector and restart products below run without the turn (it is
    // taken again at the top of the next cycle): they are matrix-core bound
    // and overlap another sub-batch's bandwidth-bound expansion.
    // Retire the converged instances: their Ritz vectors X = V[:, :m] W are
    // formed now, and the slot is filled with the last active instance, so
    // that the next cycles run on the active instances only.
    // All converged instances' Ritz vectors first, from the slots as they are
    // (one copy of the coefficients, the products back to back), then the
    // slot exchanges.
    {
      std::vector<int> conv;
      for (int b = 0; b < nact; ++b)
        if (convflag[b]) conv.push_back(b);
      if (!conv.empty()) {
        std::vector<zd> Wall(conv.size() * (size_t)m * nev);
        for (size_t i = 0; i < conv.size(); ++i) {
          const KSState &S = st[conv[i]];
          if (S.Wz.size() != (size_t)m * nev) throw std::logic_error("Ritz vectors missing after convergence");
          std::copy(S.Wz.begin(), S.Wz.end(), Wall.begin() + i * (size_t)m * nev);
          const int o = slot_inst[conv[i]];
          for (int q = 0; q < nev; ++q) res.mu[(size_t)o * nev + q] = S.theta[q];
        }
        EMU_HIP(hipMemcpyAsync(Wd, Wall.data(), sizeof(zd) * Wall.size(), hipMemcpyHostToDevice, s));
        for (size_t i = 0; i < conv.size(); ++i) {
          const int b = conv[i], o = slot_inst[b];
          launch_gemm_nn(1, n, nev, m, zmk(1, 0), V + (int64_t)b * sV, n, sV, Wd + i * (size_t)m * nev, m,
                         (int64_t)m * nev, zmk(0, 0), Xout + (int64_t)o * n * nev, n, (int64_t)n * nev, s);
        }
        EMU_HIP(hipStreamSynchronize(s));  // Wall is released below
      }
    }
    for (int b = nact - 1; b >= 0; --b) {
      if (!convflag[b]) continue;
      const int e = nact - 1;
      if (b != e) {
        swap_slots(b, e);
        std::swap(st[b], st[e]);
        std::swap(ritz[b], ritz[e]);
        std::swap(slot_inst[b], slot_inst[e]);
        swaps.push_back({b, e});
      }
      --nact;
    }
    if (nact == 0) break;
    if (last) break;
    // restart: V2[:, :keep] = V[:, :m] Q[:, :keep]; V2[:, keep:keep+bq] = V[:, m:m+bq].
    // Every instance's Q is the identity on its first min(nlock, nev)
    // columns (the Schur step and its reordering act on the rest only), so
    // with a0 the smallest such count over the batch, V2[:, :a0] = V[:, :a0]
    // is a copy and the product runs over the trailing (m - a0) x (keep - a0)
    // block: the same values with fewer flops.
    int a0 = keep;
    for (int b = 0; b < nact; ++b) a0 = std::min(a0, std::min(st[b].nlock, nev));
    std::vector<zd> Qh((size_t)nact * m * keep);
    std::vector<zd> Hn((size_t)nact * sH);
    for (int b = 0; b < nact; ++b) {
      KSState &S = st[b];
      if (S.Qk.size() != (size_t)m * keep || S.Hk.size() != (size_t)sH)
        throw std::logic_error("restart blocks missing for an active instance");
      std::copy(S.Qk.begin(), S.Qk.end(), Qh.begin() + (size_t)b * m * keep);
      std::copy(S.Hk.begin(), S.Hk.end(), Hn.begin() + (size_t)b * sH);
      S.nlock = S.nlock_next;
    }
    EMU_HIP(hipMemcpyAsync(Qd, Qh.data(), sizeof(zd) * Qh.size(), hipMemcpyHostToDevice, s));
    EMU_HIP(hipMemcpyAsync(Hd, Hn.data(), sizeof(zd) * Hn.size(), hipMemcpyHostToDevice, s));
    EMU_HIP(hipStreamSynchronize(s));
    if (a0 > 0) launch_copy(nact, (int64_t)a0 * n, V, sV, V2, sV, s);
    launch_gemm_nn(nact, n, keep - a0, m - a0, zmk(1, 0), V + (int64_t)a0 * n, n, sV, Qd + a0 + (int64_t)a0 * m, m,
                   (int64_t)m * keep, zmk(0, 0), V2 + (int64_t)a0 * n, n, sV, s);
    launch_copy(nact, (int64_t)bq * n, V + (int64_t)m * n, sV, V2 + (int64_t)keep * n, sV, s);
    std::swap(V, V2);
    k = keep;
    tm.mark("restart");
  }
  if (turn.owns_lock()) turn.unlock();
  // the caller's per-instance arrays back in their original order
  for (auto it = swaps.rbegin(); it != swaps.rend(); ++it) swap_caller(it->first, it->second);
  EMU_HIP(hipStreamSynchronize(s));
  if (nact > 0) throw NoConvergence("Krylov-Schur: eigenpairs did not converge within itermax restarts");
  std::vector<int32_t> badh(nb);
  EMU_HIP(hipMemcpy(badh.data(), bad, sizeof(int32_t) * nb, hipMemcpyDeviceToHost));
  for (int b = 0; b < nb; ++b) res.rank_fixes += badh[b];
  tm.mark("ritz_vectors");
  if (tm.on && !schur_na.empty()) {
    double sum = 0, cube = 0;
    for (int v : schur_na) {
      sum += v;
      cube += (double)v * v * v;
    }
    fprintf(stderr, "[emu timing] Schur steps %zu, active size mean %.1f min %d max %d, sum of cubes / m^3 %.2f\n",
            schur_na.size(), sum / schur_na.size(), *std::min_element(schur_na.begin(), schur_na.end()),
            *std::max_element(schur_na.begin(), schur_na.end()), cube / ((double)m * m * m));
  }
  return res;
}

// -------------------------------------------------------- host utilities

// Numerical-Recipes quicksort index (z_indexx.f / sort_n.f): ascending key,
// not stable; reproduced so that ties break as in the reference.
static std::vector<int> nr_indexx(const std::vector<double> &key) {
  const int n = (int)key.size();
  std::vector<int> indx(n);
  std::iota(indx.begin(), indx.end(), 0);
  if (n < 2) return indx;
  const int M = 7;
  std::vector<int> istack;
  int l = 0, ir = n - 1;
  for (;;) {
    if (ir - l < M) {
      for (int j = l + 1; j <= ir; ++j) {
        int indxt = indx[j];
        double a = key[indxt];
        int i = j - 1;
        for (; i >= l; --i) {
          if (key[indx[i]] <= a) break;
          indx[i + 1] = indx[i];
        }
        indx[i + 1] = indxt;
      }
      if (istack.empty()) return indx;
      ir = istack.back(); istack.pop_back();
      l = istack.back(); istack.pop_back();
    } else {
      int k = (l + ir) / 2;
      std::swap(indx[k], indx[l + 1]);
      if (key[indx[l]] > key[indx[ir]]) std::swap(indx[l], indx[ir]);
      if (key[indx[l + 1]] > key[indx[ir]]) std::swap(indx[l + 1], indx[ir]);
      if (key[indx[l]] > key[indx[l + 1]]) std::swap(indx[l], indx[l + 1]);
      int i = l + 1, j = ir;
      int indxt = indx[l + 1];
      double a = key[indxt];
      for (;;) {
        do ++i; while (key[indx[i]] < a);
        do --j; while (key[indx[j]] > a);
        if (j < i) break;
        std::swap(indx[i], indx[j]);
      }
      indx[l + 1] = indx[j];
      indx[j] = indxt;
      if (ir - i + 1 >= j - l) {
        istack.push_back(i);
        istack.push_back(ir);
        ir = j - 1;
      } else {
        istack.push_back(l);
        istack.push_back(j - 1);
        l = i;
      }
    }
  }
}

// beta = sqrt(1/mu + shift), forward branch (py_calc_modes.f:664-679)
static cplx beta_from_mu(cplx mu, cplx shift) {
  cplx b = std::sqrt(1.0 / mu + shift);
  if (std::abs(b.imag()) < 1e-8) {
    if (b.real() < 0) b = -b;
  } else if (b.imag() < 0) {
    b = -b;
  }
  return b;
}

// Conjugate-pair reordering of the adjoint modes (orthogonal.f:251-301):
// returns the swaps (j, j2) in the order the reference applies them.
static std::vector<std::pair<int, int>> pair_swaps(const std::vector<cplx> &L, int nval) {
  std::vector<std::pair<int, int>> sw;
  int j = 0, skip = 0;
  while (j < nval) {
    if (std::abs(L[j + (size_t)j * nval]) > 1e-7) {
      ++j;
    } else if (skip == 0) {
      for (int j2 = j + 1; j2 < nval; ++j2)
        if (std::abs(L[j2 + (size_t)j * nval]) > 1e-7) sw.push_back({j, j2});
      skip = 1;
      ++j;
    } else {
      skip = 0;
      ++j;
    }
  }
  return sw;
}

// Near-degenerate modes (a pair split only by the 1e-5 degree tilt of normal
// incidence, relative gap ~1e-7) are individually ill-determined, and their
// prime and adjoint vectors are biorthogonal only to (residual / gap), far
// from working precision; the per-mode normalisation of normalisation.f then
// leaves a coupling that shows in J J^dagger and R/T/A, and it depends on how
// the solve happened to converge (batch, Krylov dimension).  Within each
// group of modes coupled through the overlap matrix (|L_rc| > 1e-10
// sqrt|L_rr L_cc|) the adjoint overlaps are therefore multiplied by the
// inverse of the group's normalised overlap block: J^dagger rows of a group
// become S_g^-1 J^dagger_g, which is what exact eigenvectors would give
// (S_g = I for them); isolated modes are left exactly as the reference's
// normalisation leaves them.  Lp(r, c): overlap of adjoint mode r (after the
// pair swaps) with prime mode c, column-major.  `sym`: both families were
// scaled by 1/sqrt(L_kk) (1D path), else the adjoint by 1/L_kk (2D path).
// Ra: the instance's P2 x nval adjoint overlaps (column k = mode k), updated.
static int biorth_groups(const std::vector<cplx> &Lp, int nval, bool sym, zd *Ra, int P2) {
  std::vector<cplx> z(nval);
  std::vector<char> ok(nval);
  for (int k = 0; k < nval; ++k) {
    z[k] = std::sqrt(Lp[k + (size_t)k * nval]);
    ok[k] = std::abs(z[k]) > 1e-8;
  }
  std::vector<int> parent(nval);
  std::iota(parent.begin(), parent.end(), 0);
  auto find = [&](int i) {
    while (parent[i] != i) i = parent[i] = parent[parent[i]];
    return i;
  };
  for (int c = 0; c < nval; ++c)
    for (int r = 0; r < nval; ++r) {
      if (r == c) continue;
      const double lim = 1e-10 * std::sqrt(std::abs(Lp[r + (size_t)r * nval]) * std::abs(Lp[c + (size_t)c * nval]));
      if (std::abs(Lp[r + (size_t)c * nval]) > lim) {
        const int a = find(r), b = find(c);
        if (a != b) parent[std::max(a, b)] = std::min(a, b);
      }
    }
  std::vector<std::vector<int>> groups(nval);
  for (int k = 0; k < nval; ++k) groups[find(k)].push_back(k);
  int changed = 0;
  for (const auto &g : groups) {
    const int gs = (int)g.size();
    if (gs < 2) continue;
    bool all_ok = true;
    for (int k : g) all_ok = all_ok && ok[k];
    if (!all_ok) continue;
    // S_g (row-major gs x gs) and its inverse by Gauss-Jordan with pivoting
    std::vector<cplx> S((size_t)gs * gs), Inv((size_t)gs * gs, 0.0);
    for (int i = 0; i < gs; ++i) {
      Inv[(size_t)i * gs + i] = 1.0;
      for (int j = 0; j < gs; ++j) {
        const cplx v = Lp[g[i] + (size_t)g[j] * nval];
        S[(size_t)i * gs + j] = sym ? v / (z[g[i]] * z[g[j]]) : v / Lp[g[i] + (size_t)g[i] * nval];
      }
    }
    bool singular = false;
    for (int col = 0; col < gs && !singular; ++col) {
      int piv = col;
      for (int i = col + 1; i < gs; ++i)
        if (std::abs(S[(size_t)i * gs + col]) > std::abs(S[(size_t)piv * gs + col])) piv = i;
      if (std::abs(S[(size_t)piv * gs + col]) < 1e-12) {
        singular = true;
        break;
      }
      if (piv != col)
        for (int j = 0; j < gs; ++j) {
          std::swap(S[(size_t)piv * gs + j], S[(size_t)col * gs + j]);
          std::swap(Inv[(size_t)piv * gs + j], Inv[(size_t)col * gs + j]);
        }
      const cplx d = 1.0 / S[(size_t)col * gs + col];
      for (int j = 0; j < gs; ++j) {
        S[(size_t)col * gs + j] *= d;
        Inv[(size_t)col * gs + j] *= d;
      }
      for (int i = 0; i < gs; ++i) {
        if (i == col) continue;
        const cplx f = S[(size_t)i * gs + col];
        if (f == 0.0) continue;
        for (int j = 0; j < gs; ++j) {
          S[(size_t)i * gs + j] -= f * S[(size_t)col * gs + j];
          Inv[(size_t)i * gs + j] -= f * Inv[(size_t)col * gs + j];
        }
      }
    }
    if (singular) continue;
    std::vector<cplx> col_new((size_t)gs);
    for (int q = 0; q < P2; ++q) {
      for (int i = 0; i < gs; ++i) {
        cplx acc = 0;
        for (int j = 0; j < gs; ++j) {
          const zd v = Ra[q + (size_t)g[j] * P2];
          acc += Inv[(size_t)i * gs + j] * cplx(v.x, v.y);
        }
        col_new[i] = acc;
      }
      for (int i = 0; i < gs; ++i) Ra[q + (size_t)g[i] * P2] = zmk(col_new[i].real(), col_new[i].imag());
    }
    changed += gs;
  }
  return changed;
}

// ------------------------------------------------------------ calc_modes
// Per-instance coefficients (py_calc_modes.f:583-599, py_calc_modes_1d.f:
// 318-372): instance 2i is the prime problem (+bloch), 2i+1 the adjoint
// (-bloch); E-field formulation (eh = 1): q = eps k0^2, p = 1; H-field:
// q = k0^2, p = 1/eps.  Rejects homogeneous layers as the reference does.
static void instance_params(const ModesArgs &a, int eh, std::vector<cplx> &pp, std::vector<cplx> &qq,
                            std::vector<cplx> &sh, std::vector<double> &bl, std::vector<double> &k0) {
  const int nl = (int)a.nl, nb = 2 * nl, ntyp = (int)a.ntyp;
  pp.assign((size_t)nb * ntyp, 0);
  qq.assign((size_t)nb * ntyp, 0);
  sh.assign(nb, 0);
  bl.assign(2 * (size_t)nb, 0);
  k0.assign(nl, 0);
  for (int i = 0; i < nl; ++i) {
    k0[i] = 2.0 * M_PI / a.lambda[i];
    std::vector<cplx> eps(ntyp);
    for (int t = 0; t < ntyp; ++t) {
      cplx ne(a.n_eff[2 * ((size_t)i * ntyp + t)], a.n_eff[2 * ((size_t)i * ntyp + t) + 1]);
      eps[t] = ne * ne;
    }
    bool homog = true;
    for (int t = 0; t + 1 < ntyp; ++t)
      if (eps[t].real() != eps[t + 1].real() || eps[t].imag() != eps[t + 1].imag()) homog = false;
    if (homog) throw Homogeneous("FEM routine cannot handle homogeneous layers; define the layer as a ThinFilm");
    for (int h = 0; h < 2; ++h) {
      int b = 2 * i + h;
      for (int t = 0; t < ntyp; ++t) {
        if (eh == 1) {
          qq[(size_t)b * ntyp + t] = eps[t] * k0[i] * k0[i];
          pp[(size_t)b * ntyp + t] = 1.0;
        } else {
          qq[(size_t)b * ntyp + t] = k0[i] * k0[i];
          pp[(size_t)b * ntyp + t] = 1.0 / eps[t];
        }
      }
      sh[b] = cplx(a.shift[2 * i], a.shift[2 * i + 1]);
      double sg = h == 0 ? 1.0 : -1.0;
      bl[2 * b] = sg * a.bloch[2 * i];
      bl[2 * b + 1] = sg * a.bloch[2 * i + 1];
    }
  }
}

// Steps shared by the 2D and 1D drivers: numeric LU of K - shift M
// (mat1), Krylov-Schur for the nval largest |mu| of (K - shift M)^-1 M,
// beta = sqrt(1/mu + shift) with the reference's branch rule, and the
// z_indexx sort on -Re(beta^2).  The sorted modes (original numbering) end
// in Xn; Xo is scratch.
// default convergence tolerance of the mixed-precision form
static const double kMixedTol = 1e-8;

// beta from the two-sided Rayleigh quotient mu_rq of a mode whose Ritz value
// gave beta_ritz; the quotient is used only when it agrees with the Ritz
// value to 1e-4 (a mode whose adjoint partner is not its left eigenvector,
// e.g. a vanishing overlap, keeps the Ritz value)
static cplx rq_beta(cplx mu_rq, cplx beta_ritz, cplx shift, int64_t &used, int64_t &kept) {
  const cplx mu_ritz = 1.0 / (beta_ritz * beta_ritz - shift);
  if (!(std::abs(mu_rq - mu_ritz) <= 1e-4 * std::abs(mu_ritz))) {
    ++kept;
    return beta_ritz;
  }
  ++used;
  return beta_from_mu(mu_rq, shift);
}

struct SolveOut {
  EigResult er;
  std::vector<cplx> beta;  // nb x nval, sorted
  int32_t *dord = nullptr; // device scratch nb x nval
  int32_t *pflag = nullptr;
};

static SolveOut factor_solve_sort(Ctx *c, DevMesh &dm, const ModesArgs &a, int nb, int m, int bq,
                                  const std::vector<cplx> &sh, const zd *m1, const zd *m2p, zd *Xn, zd *Xo,
                                  StageTimer &tm) {
  hipStream_t s = c->stream;
  const Mesh &mh = dm.host;
  const Symbolic &sy = mh.sym;
  const int64_t n = mh.neq;
  const int nval = (int)a.nval;
  SolveOut o;
  zd *pool = c->buf<zd>("pool", (size_t)nb * sy.pool);
  int32_t *piv = c->buf<int32_t>("piv", (size_t)nb * sy.npiv);
  o.pflag = c->buf<int32_t>("pflag", (size_t)nb * sy.npiv);
  int32_t *sig = c->buf<int32_t>("sig", (size_t)nb * sy.npiv);
  int32_t *gx = c->buf<int32_t>("gx", (size_t)nb * sy.npiv * 4);
  zd *inv = c->buf<zd>("inv", (size_t)nb * dm.isz);
  lu_factor(dm, nb, m1, pool, piv, o.pflag, sig, gx, inv, s);
  float2 *inv32 = nullptr;
  if (a.precision == 1) {
    inv32 = c->buf<float2>("inv32", (size_t)nb * dm.isz);
    lu_narrow(dm, nb, inv, inv32, s);
  }
  tm.mark("lu_factor");
  // mixed form: the pairs converge against the complex64 operator, whose
  // distance from the FP64 one (~1e-7) bounds what a tighter test could buy
  const double tol = a.tol_set || a.precision == 0 ? a.tol : kMixedTol;
  o.er = krylov_schur(c, dm, nb, nval, m, bq, (int)a.itermax, tol, inv, inv32, gx, m2p, Xn, tm);
  launch_permute_vec(dm, nb, nval, Xn, Xo, false, s);
  o.beta.assign((size_t)nb * nval, 0);
  std::vector<int32_t> order((size_t)nb * nval);
  for (int b = 0; b < nb; ++b) {
    std::vector<cplx> bb(nval);
    std::vector<double> key(nval);
    for (int k = 0; k < nval; ++k) {
      bb[k] = beta_from_mu(o.er.mu[(size_t)b * nval + k], sh[b]);
      key[k] = -(bb[k] * bb[k]).real();
    }
    std::vector<int> idx = nr_indexx(key);
    for (int k = 0; k < nval; ++k) {
      o.beta[(size_t)b * nval + k] = bb[idx[k]];
      order[(size_t)b * nval + k] = idx[k];
    }
  }
  o.dord = c->buf<int32_t>("order", (size_t)nb * nval);
  EMU_HIP(hipMemcpyAsync(o.dord, order.data(), sizeof(int32_t) * order.size(), hipMemcpyHostToDevice, s));
  EMU_HIP(hipStreamSynchronize(s));
  launch_permute_cols(nb, n, nval, Xo, n * nval, o.dord, Xn, n * nval, s);
  return o;
}

void calc_modes(Ctx *c, DevMesh &dm, const ModesArgs &a) {
  EMU_HIP(hipSetDevice(c->device));
  if (!lapack_ready()) throw LapackMissing("host LAPACK routines not registered (emu_set_lapack)");
  hipStream_t s = c->stream;
  const Mesh &mh = dm.host;
  const int64_t n = mh.neq, nnz = mh.nnz, nel = mh.nel;
  const int nl = (int)a.nl, nb = 2 * nl, ntyp = (int)a.ntyp, nval = (int)a.nval;
  if (mh.dim != 2) throw std::invalid_argument("calc_modes_2d: the mesh handle holds a 1D grating");
  if (nl < 1 || nval < 1 || ntyp < dm.ntyp) throw std::invalid_argument("calc_modes: bad sizes");
  // Krylov dimension: ARPACK's nvect = 2 nval + nval/2 + 3 (py_calc_modes.f:256),
  // rounded up to a whole number of blocks
  const int bq = (int)a.block;
  int m = a.ncv > 0 ? (int)a.ncv : 2 * nval + nval / 2 + 3;
  m = ((m + bq - 1) / bq) * bq;
  if (m + bq >= n || nval >= m) throw std::invalid_argument("calc_modes: num_BMs too large for the mesh");
  std::vector<cplx> pp, qq, sh;
  std::vector<double> bl, k0;
  instance_params(a, (int)a.eh, pp, qq, sh, bl, k0);
  zd *dpp, *dqq, *dsh;
  double *dbl;
  upload_params(c, dm, nb, ntyp, reinterpret_cast<const double *>(pp.data()),
                reinterpret_cast<const double *>(qq.data()), reinterpret_cast<const double *>(sh.data()), bl.data(),
                &dpp, &dqq, &dsh, &dbl);
  StageTimer tm(s);
  tm.mark("setup");
  // 1. assembly
  zd *ve = c->buf<zd>("ve", (size_t)nb * nel * 14);
  zd *m1 = c->buf<zd>("mat1", (size_t)nb * nnz);
  zd *m2 = c->buf<zd>("mat2", (size_t)nb * nnz);
  zd *m2p = c->buf<zd>("mat2p", (size_t)nb * nnz);
  launch_phase(dm, nb, dbl, ve, s);
  launch_gather(dm, nb, ntyp, dpp, dqq, dsh, ve, m1, m2, s);
  launch_permute_vals(dm, nb, m2, m2p, s);
  tm.mark("assembly");
  // 2.-4. LU of K - shift M, eigensolve, beta and sort
  zd *Xn = c->buf<zd>("Xn", (size_t)nb * n * nval);
  zd *Xo = c->buf<zd>("Xo", (size_t)nb * n * nval);
  SolveOut so = factor_solve_sort(c, dm, a, nb, m, bq, sh, m1, m2p, Xn, Xo, tm);
  const EigResult &er = so.er;
  const std::vector<cplx> &beta = so.beta;
  int32_t *dord = so.dord;
  int32_t *pfl = so.pflag;
  const Symbolic &sy = mh.sym;
  // 5. field normalisation (array_sol.f:297-339)
  zd *zmax = c->buf<zd>("zmax", (size_t)nb * nval);
  size_t nparts = (size_t)((nel + 63) / 64);  // k_field_max_partial's 64-element chunks
  double *fms = c->buf<double>("fmax_scratch", (size_t)nb * nval * nparts * 4);
  launch_field_max(dm, nb, nval, Xn, n, n * nval, ve, false, zmax, fms, s);
  launch_scale_cols(nb, n, nval, Xn, n, n * nval, zmax, nval, true, s);
  if (a.sol || a.mode_pol) {
    // plotting fields: sol1 of the prime modes as array_sol.f leaves them,
    // and mode_pol of the adjoint modes in their own sorted order (the
    // reference's last writer of mode_pol is mode_energy on the adjoint
    // family, py_calc_modes.f:709-714); the core material is type 1 when
    // Re eps_1 > Re eps_2, else type 2 (py_calc_modes.f:546-551)
    const size_t nsol = (size_t)39 * nval * nel;
    zd *fsol = a.sol ? c->buf<zd>("plot_sol", nsol) : nullptr;
    zd *fpart = a.mode_pol ? c->buf<zd>("plot_part", (size_t)4 * nval * nel) : nullptr;
    zd *fpol = c->buf<zd>("plot_pol", (size_t)4 * nval);
    zd *fbe = c->buf<zd>("plot_beta", (size_t)nval + ntyp);
    std::vector<zd> hbe((size_t)nval + ntyp);
    for (int i = 0; i < nl; ++i) {
      for (int k = 0; k < nval; ++k) {
        const cplx b = beta[(size_t)(2 * i + 1) * nval + k];
        hbe[k] = zmk(b.real(), b.imag());
      }
      for (int t = 0; t < ntyp; ++t) {
        const cplx ne(a.n_eff[2 * (i * ntyp + t)], a.n_eff[2 * (i * ntyp + t) + 1]);
        const cplx ep = ne * ne;
        hbe[nval + t] = zmk(ep.real(), ep.imag());
      }
      const int core_t = ntyp > 1 && hbe[nval].x <= hbe[nval + 1].x ? 1 : 0;
      if (a.sol) {
        launch_fields_2d(dm, nval, Xn + (int64_t)(2 * i) * n * nval, n, ve + (int64_t)(2 * i) * nel * 14, fsol, s);
        EMU_HIP(hipMemcpyAsync(a.sol + (size_t)i * 2 * nsol, fsol, sizeof(zd) * nsol, hipMemcpyDeviceToHost, s));
      }
      if (a.mode_pol) {
        EMU_HIP(hipMemcpyAsync(fbe, hbe.data(), sizeof(zd) * hbe.size(), hipMemcpyHostToDevice, s));
        launch_mode_pol_2d(dm, nval, Xn + (int64_t)(2 * i + 1) * n * nval, n, ve + (int64_t)(2 * i + 1) * nel * 14,
                           fbe, fbe + nval, core_t, fpart, fpol, s);
        EMU_HIP(hipMemcpyAsync(a.mode_pol + (size_t)i * 8 * nval, fpol, sizeof(zd) * 4 * nval,
                               hipMemcpyDeviceToHost, s));
      }
      EMU_HIP(hipStreamSynchronize(s));
    }
  }
  // 6. orthogonality overlap L = -(1/k0) X2^T M_prime X1 diag(beta1) (orthogonal.f:95-248)
  zd *Y = c->buf<zd>("ovl_Y", (size_t)nl * n * nval);
  zd *Xrm = c->buf<zd>("spmm_xi", (size_t)nl * n * nval);  // row-major copy for the many-column SpMM
  launch_spmm(nl, n, nval, dm.row_ptr.p, dm.col_ind.p, m2, 2 * nnz, Xn, 2 * n * nval, Y, n * nval, s, {}, Xrm);
  zd *Ld = c->buf<zd>("ovl_L", (size_t)nl * nval * nval);
  size_t atb_elems = (size_t)nl * nval * nval * 64 + (size_t)nb * 2 * 113 * nval * 64;
  zd *atbs = c->buf<zd>("post_atb", atb_elems);
  launch_atb(nl, n, nval, nval, Xn + n * nval, n, 2 * n * nval, Y, n, n * nval, Ld, nval, (int64_t)nval * nval,
             false, atbs, atb_elems, s);
  std::vector<cplx> Lh((size_t)nl * nval * nval);
  EMU_HIP(hipMemcpyAsync(Lh.data(), Ld, sizeof(zd) * Lh.size(), hipMemcpyDeviceToHost, s));
  // the FP64 pencil K - shift M (mat1) applied to the prime modes as well,
  // for the two-sided Rayleigh quotients below
  std::vector<cplx> LAh;
  {
    launch_spmm(nl, n, nval, dm.row_ptr.p, dm.col_ind.p, m1, 2 * nnz, Xn, 2 * n * nval, Y, n * nval, s, {}, Xrm);
    zd *LAd = c->buf<zd>("ovl_LA", (size_t)nl * nval * nval);
    launch_atb(nl, n, nval, nval, Xn + n * nval, n, 2 * n * nval, Y, n, n * nval, LAd, nval, (int64_t)nval * nval,
               false, atbs, atb_elems, s);
    LAh.resize((size_t)nl * nval * nval);
    EMU_HIP(hipMemcpyAsync(LAh.data(), LAd, sizeof(zd) * LAh.size(), hipMemcpyDeviceToHost, s));
  }
  EMU_HIP(hipStreamSynchronize(s));
  std::vector<int32_t> perm2((size_t)nb * nval);
  std::vector<zd> scl((size_t)nb * nval, zmk(1, 0));
  std::vector<std::vector<cplx>> Lswapped(nl);
  std::vector<cplx> beta_c(beta);
  int64_t rq_used = 0, rq_kept = 0;  // corrected in the mixed form
  for (int i = 0; i < nl; ++i) {
    cplx *L = &Lh[(size_t)i * nval * nval];
    {
      // beta from mu = (y^T M x) / (y^T (K - shift M) x) of each prime mode x
      // and its adjoint partner y (the pencil's left eigenvector): FP64, with
      // an error of the order of the product of the two vectors' errors.
      // This matters for modes near cut-off (beta ~ 0), where a residual of
      // tol |mu| in mu becomes tol shift / (2 |beta|) in beta, and in the
      // mixed form, whose vectors carry the complex64 factor's error.
      std::vector<int> p(nval);
      std::iota(p.begin(), p.end(), 0);
      std::vector<cplx> Lv(L, L + (size_t)nval * nval);
      for (int j = 0; j < nval; ++j)
        for (int r = 0; r < nval; ++r) Lv[r + (size_t)j * nval] *= beta[(size_t)(2 * i) * nval + j] * (-1.0 / k0[i]);
      for (auto &sw : pair_swaps(Lv, nval)) std::swap(p[sw.first], p[sw.second]);
      const cplx *LA = &LAh[(size_t)i * nval * nval];
      for (int k = 0; k < nval; ++k) {
        const cplx den = LA[p[k] + (size_t)k * nval];
        if (std::abs(den) == 0.0) continue;
        const cplx mu = L[p[k] + (size_t)k * nval] / den;
        beta_c[(size_t)(2 * i) * nval + k] = rq_beta(mu, beta[(size_t)(2 * i) * nval + k], sh[2 * i], rq_used, rq_kept);
      }
    }
    for (int j = 0; j < nval; ++j) {
      cplx f = beta_c[(size_t)(2 * i) * nval + j] * (-1.0 / k0[i]);
      for (int r = 0; r < nval; ++r) L[r + (size_t)j * nval] *= f;
    }
    std::vector<int> p(nval);
    std::iota(p.begin(), p.end(), 0);
    std::vector<cplx> Lv(L, L + (size_t)nval * nval);
    for (auto &sw : pair_swaps(Lv, nval)) std::swap(p[sw.first], p[sw.second]);
    Lswapped[i].resize((size_t)nval * nval);
    for (int cc = 0; cc < nval; ++cc)
      for (int r = 0; r < nval; ++r) Lswapped[i][r + (size_t)cc * nval] = L[p[r] + (size_t)cc * nval];
    for (int k = 0; k < nval; ++k) {
      perm2[(size_t)(2 * i) * nval + k] = k;
      perm2[(size_t)(2 * i + 1) * nval + k] = p[k];
      cplx d = std::sqrt(L[p[k] + (size_t)k * nval]);  // diagonal after the swaps
      if (std::abs(d) > 1e-8) {
        cplx f = 1.0 / (d * d);
        scl[(size_t)(2 * i + 1) * nval + k] = zmk(f.real(), f.imag());
      }
    }
  }
  EMU_HIP(hipMemcpyAsync(dord, perm2.data(), sizeof(int32_t) * perm2.size(), hipMemcpyHostToDevice, s));
  EMU_HIP(hipMemcpyAsync(zmax, scl.data(), sizeof(zd) * scl.size(), hipMemcpyHostToDevice, s));
  EMU_HIP(hipStreamSynchronize(s));
  launch_permute_cols(nb, n, nval, Xn, n * nval, dord, Xo, n * nval, s);
  launch_scale_cols(nb, n, nval, Xo, n, n * nval, zmax, nval, false, s);
  tm.mark("normalise_orthogonality");
  // 7. plane-wave ordering (pw_ordering.f:48-78) and J / J^dagger
  const double d = mh.lat[0], dy = mh.lat[3];
  std::vector<std::pair<int, int>> pws;
  for (int px = -(int)a.ordre; px <= (int)a.ordre; ++px)
    for (int py = -(int)a.ordre; py <= (int)a.ordre; ++py)
      if ((px / d) * (px / d) + (py / dy) * (py / dy) <= (a.ordre / std::max(d, dy)) * (a.ordre / std::max(d, dy)))
        pws.push_back({px, py});
  const int P = (int)pws.size();
  if (P != a.neq_pw) throw std::invalid_argument("neq_PW does not match the circular plane-wave truncation");
  std::vector<double> al((size_t)nb * P), be((size_t)nb * P);
  std::vector<int32_t> pos((size_t)nb * P), sgn(nb);
  for (int i = 0; i < nl; ++i) {
    std::vector<double> key(P);
    for (int q = 0; q < P; ++q) {
      double aa = a.bloch[2 * i] + 2 * M_PI / d * pws[q].first;
      double bb = a.bloch[2 * i + 1] + 2 * M_PI / dy * pws[q].second;
      cplx kz = std::sqrt(cplx(k0[i] * k0[i] - aa * aa - bb * bb, 0.0));
      key[q] = -(kz * kz).real();
      for (int h = 0; h < 2; ++h) {
        al[(size_t)(2 * i + h) * P + q] = aa;
        be[(size_t)(2 * i + h) * P + q] = bb;
      }
    }
    std::vector<int> idx = nr_indexx(key);
    for (int r = 0; r < P; ++r) {
      pos[(size_t)(2 * i) * P + idx[r]] = r;
      pos[(size_t)(2 * i + 1) * P + idx[r]] = r;
    }
    sgn[2 * i] = -1;
    sgn[2 * i + 1] = 1;
  }
  double *dal = c->buf<double>("pw_al", (size_t)nb * P);
  double *dbe = c->buf<double>("pw_be", (size_t)nb * P);
  int32_t *dpos = c->buf<int32_t>("pw_pos", (size_t)nb * P);
  int32_t *dsg = c->buf<int32_t>("pw_sg", nb);
  EMU_HIP(hipMemcpyAsync(dal, al.data(), sizeof(double) * al.size(), hipMemcpyHostToDevice, s));
  EMU_HIP(hipMemcpyAsync(dbe, be.data(), sizeof(double) * be.size(), hipMemcpyHostToDevice, s));
  EMU_HIP(hipMemcpyAsync(dpos, pos.data(), sizeof(int32_t) * pos.size(), hipMemcpyHostToDevice, s));
  EMU_HIP(hipMemcpyAsync(dsg, sgn.data(), sizeof(int32_t) * sgn.size(), hipMemcpyHostToDevice, s));
  EMU_HIP(hipStreamSynchronize(s));
  zd *JB = c->buf<zd>("pw_JB", (size_t)nb * nel * 12 * 2 * P);
  zd *G = c->buf<zd>("pw_G", (size_t)nb * 2 * P * n);
  launch_pw_load(dm, nb, P, dal, dbe, dpos, dsg, ve, JB, G, s);
  zd *R = c->buf<zd>("pw_R", (size_t)nb * 2 * P * nval);
  launch_atb(nb, n, 2 * P, nval, G, n, 2 * P * n, Xo, n, n * nval, R, 2 * P, (int64_t)2 * P * nval, false, atbs,
             atb_elems, s);
  std::vector<zd> Rh((size_t)nb * 2 * P * nval);
  EMU_HIP(hipMemcpyAsync(Rh.data(), R, sizeof(zd) * Rh.size(), hipMemcpyDeviceToHost, s));
  std::vector<int32_t> pf((size_t)nb * sy.npiv);
  EMU_HIP(hipMemcpyAsync(pf.data(), pfl, sizeof(int32_t) * pf.size(), hipMemcpyDeviceToHost, s));
  EMU_HIP(hipStreamSynchronize(s));
  // outputs
  const int P2 = 2 * P;
  for (int i = 0; i < nl; ++i) biorth_groups(Lswapped[i], nval, false, &Rh[(size_t)(2 * i + 1) * P2 * nval], P2);
  for (int i = 0; i < nl; ++i) {
    for (int k = 0; k < nval; ++k) {
      a.beta[2 * ((size_t)i * nval + k)] = beta_c[(size_t)(2 * i) * nval + k].real();
      a.beta[2 * ((size_t)i * nval + k) + 1] = beta_c[(size_t)(2 * i) * nval + k].imag();
    }
    const zd *Rp = &Rh[(size_t)(2 * i) * P2 * nval];
    const zd *Ra = &Rh[(size_t)(2 * i + 1) * P2 * nval];
    double *J = a.J + (size_t)i * 2 * P2 * nval;
    double *Jd = a.Jd + (size_t)i * 2 * P2 * nval;
    for (int k = 0; k < nval; ++k)
      for (int q = 0; q < P2; ++q) {
        J[2 * (q + (size_t)k * P2)] = Rp[q + (size_t)k * P2].x;
        J[2 * (q + (size_t)k * P2) + 1] = Rp[q + (size_t)k * P2].y;
        Jd[2 * (k + (size_t)q * nval)] = Ra[q + (size_t)k * P2].x;
        Jd[2 * (k + (size_t)q * nval) + 1] = Ra[q + (size_t)k * P2].y;
      }
  }
  tm.mark("J_overlaps");
  if (a.stats) {
    a.stats[0] = er.ops;
    a.stats[1] = er.restarts;
    a.stats[2] = std::accumulate(pf.begin(), pf.end(), (int64_t)0);
    a.stats[3] = er.steps;
    a.stats[4] = er.rank_fixes;
    a.stats[5] = rq_used;
    a.stats[6] = rq_kept;
  }
}

// ----------------------------------------------------------- 1D gratings
// calc_modes_1d (backend/fortran/py_calc_modes_1d.f:1-707) on the GPU: the
// same sparse LU / Krylov-Schur machinery on the 1D pattern, followed by
// the normalisation of array_sol_1d.f, the overlap of orthogonal_1d.f with
// its pair swaps, normalisation_1d.f (both families by 1/sqrt(L_ii)),
// pw_ordering_1d.f and the J / J^dagger overlaps; pw_matrix_1d_to_2d.f when
// the rest of the stack is 2D.
DevMesh *mesh_create_1d(Ctx *c, int64_t npt, int64_t nel, const int64_t *table_nod, const int64_t *type_el,
                        const double *x, int leaf) {
  EMU_HIP(hipSetDevice(c->device));
  DevMesh *dm = new DevMesh();
  try {
    build_mesh_1d(dm->host, npt, nel, table_nod, type_el, x);
    build_symbolic(dm->host, leaf > 0 ? leaf : 96);
    devmesh_build(*dm, c->stream);
  } catch (...) {
    delete dm;
    throw;
  }
  return dm;
}

void calc_modes_1d(Ctx *c, DevMesh &dm, const ModesArgs &a) {
  EMU_HIP(hipSetDevice(c->device));
  if (!lapack_ready()) throw LapackMissing("host LAPACK routines not registered (emu_set_lapack)");
  hipStream_t s = c->stream;
  const Mesh &mh = dm.host;
  if (mh.dim != 1) throw std::invalid_argument("calc_modes_1d: the mesh handle holds a 2D mesh");
  const int64_t n = mh.neq, nnz = mh.nnz, nel = mh.nel;
  const int nl = (int)a.nl, nb = 2 * nl, ntyp = (int)a.ntyp, nval = (int)a.nval;
  if (nl < 1 || nval < 1 || ntyp < dm.ntyp) throw std::invalid_argument("calc_modes_1d: bad sizes");
  const int bq = (int)a.block;
  int m = a.ncv > 0 ? (int)a.ncv : 2 * nval + nval / 2 + 3;  // nvect of py_calc_modes_1d.f:145
  m = ((m + bq - 1) / bq) * bq;
  if (m + bq >= n || nval >= m) throw std::invalid_argument("calc_modes_1d: num_BMs too large for the mesh");
  const int P = 2 * (int)a.ordre + 1;
  if (P != a.neq_pw) throw std::invalid_argument("neq_PW must be 2 ordre_ls + 1 for a 1D grating");
  std::vector<cplx> pp, qq, sh;
  std::vector<double> bl, k0;
  instance_params(a, 1, pp, qq, sh, bl, k0);  // the 1D path is E-field only
  zd *dpp, *dqq, *dsh;
  double *dbl;
  upload_params(c, dm, nb, ntyp, reinterpret_cast<const double *>(pp.data()),
                reinterpret_cast<const double *>(qq.data()), reinterpret_cast<const double *>(sh.data()), bl.data(),
                &dpp, &dqq, &dsh, &dbl);
  StageTimer tm(s);
  tm.mark("setup");
  // 1. assembly of K - shift M, M and the overlap form
  zd *m1 = c->buf<zd>("mat1", (size_t)nb * nnz);
  zd *m2 = c->buf<zd>("mat2", (size_t)nb * nnz);
  zd *m2p = c->buf<zd>("mat2p", (size_t)nb * nnz);
  zd *mov = c->buf<zd>("mat_ovl", (size_t)nb * nnz);
  launch_gather_1d(dm, nb, ntyp, dpp, dqq, dsh, dbl, 0, m1, m2, s);
  launch_gather_1d(dm, nb, ntyp, dpp, dqq, dsh, dbl, 1, mov, nullptr, s);
  launch_permute_vals(dm, nb, m2, m2p, s);
  tm.mark("assembly");
  // 2.-4. LU, eigensolve, beta and sort
  zd *Xn = c->buf<zd>("Xn", (size_t)nb * n * nval);
  zd *Xo = c->buf<zd>("Xo", (size_t)nb * n * nval);
  SolveOut so = factor_solve_sort(c, dm, a, nb, m, bq, sh, m1, m2p, Xn, Xo, tm);
  const std::vector<cplx> &beta = so.beta;
  // 5. field normalisation (array_sol_1d.f:141-292)
  double xlo = mh.xd1d[0], xhi = mh.xd1d[0];
  for (double v : mh.xd1d) {
    xlo = std::min(xlo, v);
    xhi = std::max(xhi, v);
  }
  const double xmid = (xlo + xhi) / 2.0, dq = (xhi - xlo) / (double)nel / 4.0;
  zd *den = c->buf<zd>("zmax", (size_t)nb * nval);
  launch_field_max_1d(dm, nb, nval, Xn, n * nval, dbl, xmid, dq, den, s);
  launch_scale_cols(nb, n, nval, Xn, n, n * nval, den, nval, true, s);
  // 6. overlap L = X2^T Mov X1 diag(beta1) / k0 of the prime instances
  zd *Y = c->buf<zd>("ovl_Y", (size_t)nl * n * nval);
  zd *Xrm = c->buf<zd>("spmm_xi", (size_t)nl * n * nval);  // row-major copy for the many-column SpMM
  launch_spmm(nl, n, nval, dm.row_ptr.p, dm.col_ind.p, mov, 2 * nnz, Xn, 2 * n * nval, Y, n * nval, s, {}, Xrm);
  zd *Ld = c->buf<zd>("ovl_L", (size_t)nl * nval * nval);
  size_t atb_elems = (size_t)nl * nval * nval * 64 + (size_t)nb * 2 * P * nval * 64;
  zd *atbs = c->buf<zd>("post_atb", atb_elems);
  launch_atb(nl, n, nval, nval, Xn + n * nval, n, 2 * n * nval, Y, n, n * nval, Ld, nval, (int64_t)nval * nval,
             false, atbs, atb_elems, s);
  std::vector<cplx> Lh((size_t)nl * nval * nval);
  EMU_HIP(hipMemcpyAsync(Lh.data(), Ld, sizeof(zd) * Lh.size(), hipMemcpyDeviceToHost, s));
  // y^T M x and y^T (K - shift M) x of the prime modes and their adjoint
  // partners, for the two-sided Rayleigh quotients below
  std::vector<cplx> LBh, LAh;
  {
    zd *LAd = c->buf<zd>("ovl_LA", (size_t)nl * nval * nval);
    zd *LBd = c->buf<zd>("ovl_LB", (size_t)nl * nval * nval);
    for (int w = 0; w < 2; ++w) {
      launch_spmm(nl, n, nval, dm.row_ptr.p, dm.col_ind.p, w ? m1 : m2, 2 * nnz, Xn, 2 * n * nval, Y, n * nval, s,
                  {}, Xrm);
      launch_atb(nl, n, nval, nval, Xn + n * nval, n, 2 * n * nval, Y, n, n * nval, w ? LAd : LBd, nval,
                 (int64_t)nval * nval, false, atbs, atb_elems, s);
    }
    LAh.resize((size_t)nl * nval * nval);
    LBh.resize((size_t)nl * nval * nval);
    EMU_HIP(hipMemcpyAsync(LAh.data(), LAd, sizeof(zd) * LAh.size(), hipMemcpyDeviceToHost, s));
    EMU_HIP(hipMemcpyAsync(LBh.data(), LBd, sizeof(zd) * LBh.size(), hipMemcpyDeviceToHost, s));
  }
  EMU_HIP(hipStreamSynchronize(s));
  std::vector<int32_t> perm2((size_t)nb * nval);
  std::vector<zd> scl((size_t)nb * nval, zmk(1, 0));
  std::vector<std::vector<cplx>> Lswapped(nl);
  std::vector<cplx> beta_c(beta);
  int64_t rq_used = 0, rq_kept = 0;  // corrected in the mixed form
  for (int i = 0; i < nl; ++i) {
    cplx *L = &Lh[(size_t)i * nval * nval];
    for (int j = 0; j < nval; ++j) {
      cplx f = beta[(size_t)(2 * i) * nval + j] / k0[i];
      for (int r = 0; r < nval; ++r) L[r + (size_t)j * nval] *= f;
    }
    std::vector<int> p(nval);
    std::iota(p.begin(), p.end(), 0);
    std::vector<cplx> Lv(L, L + (size_t)nval * nval);
    for (auto &sw : pair_swaps(Lv, nval)) std::swap(p[sw.first], p[sw.second]);
    {
      // beta from mu = (y^T M x) / (y^T (K - shift M) x) (FP64 pencil; see
      // calc_modes)
      const cplx *LA = &LAh[(size_t)i * nval * nval], *LB = &LBh[(size_t)i * nval * nval];
      for (int k = 0; k < nval; ++k) {
        const cplx den = LA[p[k] + (size_t)k * nval];
        if (std::abs(den) == 0.0) continue;
        beta_c[(size_t)(2 * i) * nval + k] = rq_beta(LB[p[k] + (size_t)k * nval] / den, beta[(size_t)(2 * i) * nval + k],
                                                     sh[2 * i], rq_used, rq_kept);
      }
      // the overlap carries beta1: rescale it by the corrected values
      for (int j = 0; j < nval; ++j) {
        const cplx f = beta_c[(size_t)(2 * i) * nval + j] / beta[(size_t)(2 * i) * nval + j];
        for (int r = 0; r < nval; ++r) L[r + (size_t)j * nval] *= f;
      }
    }
    Lswapped[i].resize((size_t)nval * nval);
    for (int cc = 0; cc < nval; ++cc)
      for (int r = 0; r < nval; ++r) Lswapped[i][r + (size_t)cc * nval] = L[p[r] + (size_t)cc * nval];
    for (int k = 0; k < nval; ++k) {
      perm2[(size_t)(2 * i) * nval + k] = k;
      perm2[(size_t)(2 * i + 1) * nval + k] = p[k];
      cplx z = std::sqrt(L[p[k] + (size_t)k * nval]);  // diagonal after the swaps
      if (std::abs(z) > 1e-8) {
        cplx f = 1.0 / z;
        scl[(size_t)(2 * i) * nval + k] = zmk(f.real(), f.imag());
        scl[(size_t)(2 * i + 1) * nval + k] = zmk(f.real(), f.imag());
      }
    }
  }
  EMU_HIP(hipMemcpyAsync(so.dord, perm2.data(), sizeof(int32_t) * perm2.size(), hipMemcpyHostToDevice, s));
  EMU_HIP(hipMemcpyAsync(den, scl.data(), sizeof(zd) * scl.size(), hipMemcpyHostToDevice, s));
  EMU_HIP(hipStreamSynchronize(s));
  launch_permute_cols(nb, n, nval, Xn, n * nval, so.dord, Xo, n * nval, s);
  launch_scale_cols(nb, n, nval, Xo, n, n * nval, den, nval, false, s);
  if (a.sol) {
    // sol_P2 of the prime modes after normalisation_1d (array_sol_P2_1d.f)
    const size_t nsol = (size_t)9 * nval * nel;
    zd *fsol = c->buf<zd>("plot_sol", nsol);
    for (int i = 0; i < nl; ++i) {
      launch_sol_p2_1d(dm, nval, Xo + (int64_t)(2 * i) * n * nval, n, a.bloch[2 * i], fsol, s);
      EMU_HIP(hipMemcpyAsync(a.sol + (size_t)i * 2 * nsol, fsol, sizeof(zd) * nsol, hipMemcpyDeviceToHost, s));
      EMU_HIP(hipStreamSynchronize(s));
    }
  }
  tm.mark("normalise_orthogonality");
  // 7. plane-wave order (pw_ordering_1d.f) and J / J^dagger
  std::vector<double> al((size_t)nb * P), kyv(nb);
  std::vector<int32_t> sgn(nb);
  std::vector<std::vector<int>> inv1(nl, std::vector<int>(P));
  for (int i = 0; i < nl; ++i) {
    const double kx = a.bloch[2 * i], ky = a.bloch[2 * i + 1];
    std::vector<double> key(P), alu(P);
    for (int q = 0; q < P; ++q) {
      alu[q] = kx + 2.0 * M_PI / mh.period_x * (q - (int)a.ordre);
      cplx kz = std::sqrt(cplx(k0[i] * k0[i] - alu[q] * alu[q] - ky * ky, 0.0));
      key[q] = -(kz * kz).real();
    }
    std::vector<int> idx = nr_indexx(key);
    for (int r = 0; r < P; ++r) {
      inv1[i][idx[r]] = r;
      al[(size_t)(2 * i) * P + r] = alu[idx[r]];
      al[(size_t)(2 * i + 1) * P + r] = alu[idx[r]];
    }
    kyv[2 * i] = kyv[2 * i + 1] = ky;
    sgn[2 * i] = -1;
    sgn[2 * i + 1] = 1;
  }
  double *dal = c->buf<double>("pw_al", (size_t)nb * P);
  double *dky = c->buf<double>("pw_ky", nb);
  int32_t *dsg = c->buf<int32_t>("pw_sg", nb);
  EMU_HIP(hipMemcpyAsync(dal, al.data(), sizeof(double) * al.size(), hipMemcpyHostToDevice, s));
  EMU_HIP(hipMemcpyAsync(dky, kyv.data(), sizeof(double) * kyv.size(), hipMemcpyHostToDevice, s));
  EMU_HIP(hipMemcpyAsync(dsg, sgn.data(), sizeof(int32_t) * sgn.size(), hipMemcpyHostToDevice, s));
  EMU_HIP(hipStreamSynchronize(s));
  zd *G = c->buf<zd>("pw_G", (size_t)nb * 2 * P * n);
  launch_pw_load_1d(dm, nb, P, dal, dky, dsg, dbl, G, s);
  zd *R = c->buf<zd>("pw_R", (size_t)nb * 2 * P * nval);
  launch_atb(nb, n, 2 * P, nval, G, n, 2 * P * n, Xo, n, n * nval, R, 2 * P, (int64_t)2 * P * nval, false, atbs,
             atb_elems, s);
  std::vector<zd> Rh((size_t)nb * 2 * P * nval);
  EMU_HIP(hipMemcpyAsync(Rh.data(), R, sizeof(zd) * Rh.size(), hipMemcpyDeviceToHost, s));
  std::vector<int32_t> pf((size_t)nb * mh.sym.npiv);
  EMU_HIP(hipMemcpyAsync(pf.data(), so.pflag, sizeof(int32_t) * pf.size(), hipMemcpyDeviceToHost, s));
  EMU_HIP(hipStreamSynchronize(s));
  // outputs: beta, J (2P x nval), J^dagger (nval x 2P) and, for a 2D world,
  // the same embedded in the 2D plane-wave basis (pw_matrix_1d_to_2d.f)
  const int P2 = 2 * P;
  for (int i = 0; i < nl; ++i) biorth_groups(Lswapped[i], nval, true, &Rh[(size_t)(2 * i + 1) * P2 * nval], P2);
  int PW2 = 0;
  std::vector<int> pos2d;  // per 1D order (unsorted): sorted 2D position of (px, 0)
  for (int i = 0; i < nl; ++i) {
    for (int k = 0; k < nval; ++k) {
      a.beta[2 * ((size_t)i * nval + k)] = beta_c[(size_t)(2 * i) * nval + k].real();
      a.beta[2 * ((size_t)i * nval + k) + 1] = beta_c[(size_t)(2 * i) * nval + k].imag();
    }
    const zd *Rp = &Rh[(size_t)(2 * i) * P2 * nval];
    const zd *Ra = &Rh[(size_t)(2 * i + 1) * P2 * nval];
    double *J = a.J + (size_t)i * 2 * P2 * nval;
    double *Jd = a.Jd + (size_t)i * 2 * P2 * nval;
    for (int k = 0; k < nval; ++k)
      for (int q = 0; q < P2; ++q) {
        J[2 * (q + (size_t)k * P2)] = Rp[q + (size_t)k * P2].x;
        J[2 * (q + (size_t)k * P2) + 1] = Rp[q + (size_t)k * P2].y;
        Jd[2 * (k + (size_t)q * nval)] = Ra[q + (size_t)k * P2].x;
        Jd[2 * (k + (size_t)q * nval) + 1] = Ra[q + (size_t)k * P2].y;
      }
    if (a.world_1d == 0 && a.J2d && a.Jd2d) {
      const double kx = a.bloch[2 * i], ky = a.bloch[2 * i + 1], px_ = mh.period_x;
      const int o = (int)a.ordre;
      std::vector<double> key;
      std::vector<int> py0(P, -1);
      for (int px = -o; px <= o; ++px)
        for (int py = -o; py <= o; ++py)
          if ((px / px_) * (px / px_) + (py / px_) * (py / px_) <= (o / px_) * (o / px_)) {
            double aa = kx + 2.0 * M_PI / px_ * px, bb = ky + 2.0 * M_PI / px_ * py;
            cplx kz = std::sqrt(cplx(k0[i] * k0[i] - aa * aa - bb * bb, 0.0));
            if (py == 0) py0[px + o] = (int)key.size();
            key.push_back(-(kz * kz).real());
          }
      PW2 = (int)key.size();
      if (PW2 != a.neq_pw_2d) throw std::invalid_argument("neq_PW_2d does not match the 2D plane-wave truncation");
      std::vector<int> idx = nr_indexx(key), inv2(PW2);
      for (int r = 0; r < PW2; ++r) inv2[idx[r]] = r;
      double *J2 = a.J2d + (size_t)i * 2 * 2 * PW2 * nval;
      double *Jd2 = a.Jd2d + (size_t)i * 2 * 2 * PW2 * nval;
      std::fill(J2, J2 + (size_t)2 * 2 * PW2 * nval, 0.0);
      std::fill(Jd2, Jd2 + (size_t)2 * 2 * PW2 * nval, 0.0);
      for (int ipw = 0; ipw < P; ++ipw) {
        const int s1 = inv1[i][ipw], s2 = inv2[py0[ipw]];
        for (int k = 0; k < nval; ++k)
          for (int hpol = 0; hpol < 2; ++hpol) {
            const zd vj = Rp[(s1 + hpol * P) + (size_t)k * P2];
            const zd vd = Ra[(s1 + hpol * P) + (size_t)k * P2];
            const size_t q2 = (size_t)(s2 + hpol * PW2);
            J2[2 * (q2 + (size_t)k * 2 * PW2)] = vj.x;
            J2[2 * (q2 + (size_t)k * 2 * PW2) + 1] = vj.y;
            Jd2[2 * (k + q2 * nval)] = vd.x;
            Jd2[2 * (k + q2 * nval) + 1] = vd.y;
          }
      }
    }
  }
  tm.mark("J_overlaps");
  if (a.stats) {
    a.stats[0] = so.er.ops;
    a.stats[1] = so.er.restarts;
    a.stats[2] = std::accumulate(pf.begin(), pf.end(), (int64_t)0);
    a.stats[3] = so.er.steps;
    a.stats[4] = so.er.rank_fixes;
    a.stats[5] = rq_used;
    a.stats[6] = rq_kept;
  }
}

void mesh_free(DevMesh *m) { delete m; }

DevMesh *mesh_analyze(const std::string &path, int leaf) {
  DevMesh *dm = new DevMesh();
  try {
    build_mesh(dm->host, path);
    build_symbolic(dm->host, leaf > 0 ? leaf : 96);
  } catch (...) {
    delete dm;
    throw;
  }
  return dm;
}

void mesh_symbolic(const DevMesh *dm, int32_t *fronts, int64_t *foff, int64_t *uoff, int32_t *uidx, int32_t *pmap,
                   int64_t *nnz2front) {
  const Symbolic &sy = dm->host.sym;
  for (size_t i = 0; i < sy.fronts.size(); ++i) {
    const FrontInfo &F = sy.fronts[i];
    if (fronts) {
      int32_t *r = fronts + 8 * i;
      r[0] = F.s; r[1] = F.u; r[2] = F.f; r[3] = F.b; r[4] = F.parent; r[5] = F.level;
      r[6] = F.child[0]; r[7] = F.child[1];
    }
    if (foff) foff[i] = F.off;
    if (uoff) uoff[i] = F.uoff;
  }
  if (uidx) std::copy(sy.uidx.begin(), sy.uidx.end(), uidx);
  if (pmap) std::copy(sy.pmap.begin(), sy.pmap.end(), pmap);
  if (nnz2front) std::copy(sy.nnz2front.begin(), sy.nnz2front.end(), nnz2front);
}

void mesh_info(const DevMesh *dm, int64_t *info) {
  const Mesh &m = dm->host;
  const Symbolic &sy = m.sym;
  int64_t v[16] = {m.npt, m.nel, m.n_ddl, m.neq, m.nnz, (int64_t)sy.fronts.size(), (int64_t)sy.levels.size(),
                   sy.pool, sy.max_f, sy.factor_entries, (int64_t)sy.flops, m.n_edge, m.npt_p3, sy.nw, 0, 0};
  std::memcpy(info, v, sizeof(v));
}

void mesh_arrays(const DevMesh *dm, int64_t *row_ptr, int32_t *col_ind, int32_t *ldof, int32_t *perm, double *lat) {
  const Mesh &m = dm->host;
  if (row_ptr) std::copy(m.row_ptr.begin(), m.row_ptr.end(), row_ptr);
  if (col_ind) std::copy(m.col_ind.begin(), m.col_ind.end(), col_ind);
  if (ldof) std::copy(m.ldof.begin(), m.ldof.end(), ldof);
  if (perm) std::copy(m.sym.perm.begin(), m.sym.perm.end(), perm);
  if (lat) std::copy(m.lat, m.lat + 4, lat);
}

void mesh_geometry(const DevMesh *dm, int64_t *table_nod, int64_t *type_el, double *x_arr) {
  const Mesh &m = dm->host;
  if (table_nod) std::copy(m.table_nod.begin(), m.table_nod.end(), table_nod);
  if (type_el) std::copy(m.type_el.begin(), m.type_el.end(), type_el);
  if (x_arr) std::copy(m.x.begin(), m.x.end(), x_arr);
}

}  // namespace emu
