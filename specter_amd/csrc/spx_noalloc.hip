// spx_gradre_noalloc: gradre with its work arrays in a static device array
// instead of memory the library allocates, for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops).
//
// In host code `scratch` names a host-side stand-in, not device memory: its
// device address is looked up once per device (hipGetSymbolAddress, with that
// device current, on the entry point's first call there) and only that
// address is handed to kernels.  Every call on one GPU shares that GPU's
// array: calls that overlap on one GPU (different streams or threads) race.
#include <mutex>
#include <string>

#include "spx_common.h"

using namespace spx;

#define SPX_REQ(cond, msg) \
  do {                     \
    if (!(cond)) return fail(SPX_ERR_ARG, msg); \
  } while (0)

// the scratch array: 128 MiB on every GPU the library runs on, held until the
// process ends (HD float64 gradre of a 64x64x(64+25) grid needs 45 MB)
constexpr size_t kScratchBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices] = {nullptr};
static std::mutex g_scratch_mu;

static int scratch_address(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(SPX_ERR_UNSUPPORTED, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(SPX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                   hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = static_cast<unsigned char*>(g_scratch_ptr[dev]);
  return SPX_OK;
}

// the gradre work set carved out of the scratch, each array on a 256-byte
// boundary: XV, XZ, XY (3 each), RV, RO (3 each), and in mixed plans with
// walls the six float64 boundary-plane buffers
// (in the order of the pointer slots of spx_gradre_noalloc); returns the count
static size_t round256(size_t b) { return (b + 255) / 256 * 256; }
static int noalloc_sizes(const spx_plan* p, size_t sz[21]) {
  const size_t cb = p->nc * p->nb * p->csize, rb = p->nr * p->nb * p->rsize;
  for (int i = 0; i < 9; ++i) sz[i] = cb;
  for (int i = 9; i < 15; ++i) sz[i] = rb;
  if (!(p->fp32 && p->g.Cz > 0)) return 15;
  const size_t d2 = 2 * size_t(p->g.oz), nbm = size_t(p->nb);
  const size_t bcb = size_t(p->nkx) * p->g.ny * d2 * nbm * 16;
  const size_t brb = size_t(p->g.nx) * p->g.ny * d2 * nbm * 8;
  const size_t bsz[6] = {4 * bcb, 4 * bcb, 4 * bcb, 3 * brb, 4 * brb, 4 * bcb};
  for (int i = 0; i < 6; ++i) sz[15 + i] = bsz[i];
  return 21;
}

}  // namespace

extern "C" {

int spx_gradre_scratch_bytes(const spx_plan* p, size_t* need, size_t* limit) {
  SPX_REQ(p && need && limit, "spx_gradre_scratch_bytes: null argument");
  size_t sz[21];
  const int nset = noalloc_sizes(p, sz);
  size_t n = 0;
  for (int i = 0; i < nset; ++i) n += round256(sz[i]);
  *need = n;
  *limit = kScratchBytes;
  return SPX_OK;
}

int spx_gradre_noalloc(spx_plan* p, const void* a, const void* b, const void* c, void* d,
                       void* e, void* f) {
  SPX_REQ(p && a && b && c && d && e && f, "spx_gradre_noalloc: null argument");
  size_t need, limit;
  int rc = spx_gradre_scratch_bytes(p, &need, &limit);
  if (rc) return rc;
  if (need > limit)
    return fail(SPX_ERR_UNSUPPORTED,
                "spx_gradre_noalloc: the grid needs " + std::to_string(need) +
                    " bytes of work arrays; the static scratch holds " +
                    std::to_string(limit) + " (use spx_gradre)");
  int dev = -1;
  SPX_HIP(hipGetDevice(&dev));
  if (dev != p->device)
    return fail(SPX_ERR_UNSUPPORTED, "spx_gradre_noalloc: the plan's device is not current");
  unsigned char* base = nullptr;
  if ((rc = scratch_address(dev, &base))) return rc;
  // carve the work set out of the scratch (every array is written before it
  // is read, as with the allocated set, so nothing is cleared); the plan's
  // own work pointers are swapped for the call and restored after it
  void** slots[21] = {&p->XV[0], &p->XV[1], &p->XV[2], &p->XZ[0], &p->XZ[1], &p->XZ[2],
                      &p->XY[0], &p->XY[1], &p->XY[2], &p->RV[0], &p->RV[1], &p->RV[2],
                      &p->RO[0], &p->RO[1], &p->RO[2], &p->XVb, &p->XZb, &p->XYb,
                      &p->RVb,   &p->ROb,   &p->XFb};
  void* sv[21];
  for (int i = 0; i < 21; ++i) sv[i] = *slots[i];
  size_t sz[21];
  const int nset = noalloc_sizes(p, sz);
  size_t off = 0;
  for (int i = 0; i < nset; ++i) {
    *slots[i] = base + off;
    off += round256(sz[i]);
  }
  {
    MemberView mv(p, p->nact);
    const void* v[3] = {a, b, c};
    void* o[3] = {d, e, f};
    rc = gradre_impl(p, v, o, false);
  }
  for (int i = 0; i < 21; ++i) *slots[i] = sv[i];
  return rc;
}

}  // extern "C"
