// Assembly of K - shift M and M into caller-provided device buffers without
// any device allocation, for callers that cannot allocate (code under HIP
// graph capture, tight serving loops).  The only scratch the assembly
// kernels need, the per-element Bloch phases of every instance (nbatch x
// nel x 14 complex values), lives in a file-scope device array of fixed
// size, one copy per GPU, held until the process ends.
//
// Every call on a GPU shares that GPU's array and leaves its contents
// behind; calls that overlap on one GPU (two streams, two host threads)
// race on it.  k_phase writes every phase an instance reads before k_gather
// reads it, so nothing needs clearing between calls.
#include <mutex>

#include "common.h"
#include "engine.h"
#include "engine_api.h"

namespace emu {

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

static std::mutex g_scratch_mu;
static void *g_scratch_dev[kMaxDevices];

// Device address of `scratch` on `device`, looked up once per device with
// that device current.  Host code passes this address to kernels; the
// array's name itself is a host-side stand-in and must not be.
void *noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("noalloc scratch: device out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    int cur = -1;
    EMU_HIP(hipGetDevice(&cur));
    if (cur != device) EMU_HIP(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      throw HipError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[device] = p;
  }
  return g_scratch_dev[device];
}

size_t assemble_noalloc_scratch(const DevMesh &dm, int64_t nb) {
  return dm.host.dim == 1 ? 0 : (size_t)nb * (size_t)dm.host.nel * 14 * sizeof(zd);
}

void assemble_noalloc(Ctx *c, DevMesh &dm, int nb, int ntyp, const double *pp_, const double *qq_,
                      const double *shift_, const double *bloch, double *mat1_, double *mat2_, void *stream) {
  const zd *pp = reinterpret_cast<const zd *>(pp_), *qq = reinterpret_cast<const zd *>(qq_);
  const zd *shift = reinterpret_cast<const zd *>(shift_);
  zd *mat1 = reinterpret_cast<zd *>(mat1_), *mat2 = reinterpret_cast<zd *>(mat2_);
  if (ntyp < dm.ntyp) throw std::invalid_argument("nb_typ_el smaller than the mesh's material count");
  const size_t need = assemble_noalloc_scratch(dm, nb);
  if (need > kNoallocScratchBytes)
    throw std::invalid_argument("emu_assemble_noalloc: nbatch x nel x 224 bytes = " + std::to_string(need) +
                                " exceeds the fixed scratch of " + std::to_string(kNoallocScratchBytes) + " bytes");
  EMU_HIP(hipSetDevice(c->device));
  hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : c->stream;
  if (dm.host.dim == 1) {
    launch_gather_1d(dm, nb, ntyp, pp, qq, shift, bloch, 0, mat1, mat2, s);
    return;
  }
  zd *ve = reinterpret_cast<zd *>(noalloc_scratch(c->device));
  launch_phase(dm, nb, bloch, ve, s);
  launch_gather(dm, nb, ntyp, pp, qq, shift, ve, mat1, mat2, s);
}

}  // namespace emu
