// Allocation-free Coulomb entry point (hf_atomic_coulomb_dev_noalloc).
//
// The atomic Coulomb build needs per-call work space beside its inputs and
// outputs: the Gaunt-weighted in-element bands Jin [n_lm][Nel][256], the two
// multipole traces per ((L,M), element) and their prefix sums.  The ordinary
// entry point keeps these in buffers the handle allocates at upload; this one
// takes them from `scratch`, a file-scope device array every GPU holds from
// module load to process end, so the call itself allocates nothing (for
// callers under stream capture or in loops that must not allocate).
//
// In host code `scratch` names a host-side stand-in, not device memory: the
// device address is looked up once per device with hipGetSymbolAddress (from
// hf_ctx_create, with that device current) and only that address is used.
// Every call on one GPU shares its array and leaves the last call's values
// in it; calls that overlap on one GPU (two streams, two host threads) race.
// Nothing in the array needs to start at zero: each call writes every entry
// it reads before reading it.
#include "capi_common.hpp"
#include <mutex>

using namespace hfg;

namespace {

// 16 MiB: an atomic basis fits when n_lm * Nel * (256 + 4) doubles do (Xe at
// lmax = mmax = 6 on 25 elements needs 8.8 MB).
constexpr size_t kScratchDoubles = (size_t)2 << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// Work-space doubles of one atomic Coulomb build: n_lm counts the (L,M)
// pairs L = 0 .. N_L-1, |M| <= min(L, Mmax) with Mmax = max m - min m (the
// (L,M) list the upload builds, TwoDBasis.cpp:790-822).
size_t coulomb_work_doubles(const AtomicSetup &S) {
  const int Mmax = *std::max_element(S.mval.begin(), S.mval.end()) -
                   *std::min_element(S.mval.begin(), S.mval.end());
  size_t nlm = 0;
  for (int L = 0; L < S.N_L; ++L) nlm += 2 * (size_t)std::min(L, Mmax) + 1;
  return std::max<size_t>(nlm, 1) * S.rad.Nel() * (TILE2 + 4);
}

}  // namespace

namespace hfg {

// Look up this device's copy of `scratch` (the current device must be
// `device`); called once per device from hf_ctx_create.
void noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error(HF_EINVAL, "device index above the scratch table");
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch_ptr[device]) return;
  void *p = nullptr;
  hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  if (!p) throw Error(HF_EDEVICE, "hipGetSymbolAddress returned no address for the scratch array");
  g_scratch_ptr[device] = static_cast<double *>(p);
}

}  // namespace hfg

extern "C" {

int hf_atomic_coulomb_dev_noalloc(hf_atomic *h, const double *dP, int64_t ldp, double *dJ, int64_t ldj) {
  return guarded([&] {
    need(h != nullptr, "NULL handle");
    need(h->ctx != nullptr, "handle was created without a device context");
    const size_t need_doubles = coulomb_work_doubles(h->host);
    if (need_doubles > kScratchDoubles)
      throw Error(HF_EINVAL, "basis needs " + std::to_string(need_doubles * sizeof(double)) +
                                 " bytes of Coulomb work space; the allocation-free entry point has " +
                                 std::to_string(kScratchDoubles * sizeof(double)));
    if (!h->host.have_tei) throw Error(HF_ESTATE, "Primitive teis have not been computed!");
    if (!h->uploaded)
      throw Error(HF_ESTATE, "the handle's tables are not on the device yet: make one ordinary J or K call first");
    if (atomic_coulomb_work_doubles(h->dev) > kScratchDoubles)
      throw Error(HF_EINTERNAL, "Coulomb work space of the uploaded tables exceeds the scratch array");
    check_matrix(dP, ldp, h->host.Nbf(), "P");
    check_matrix(dJ, ldj, h->host.Nbf(), "J");
    const int dev = h->ctx->device;
    double *work = nullptr;
    {
      std::lock_guard<std::mutex> lk(g_mu);
      work = (dev >= 0 && dev < kMaxDevices) ? g_scratch_ptr[dev] : nullptr;
    }
    if (!work) throw Error(HF_ESTATE, "the scratch array of this device was not looked up");
    DeviceGuard g(dev);
    atomic_coulomb_launch(h->dev, dP, ldp, dJ, ldj, h->ctx->cur, nullptr, work);
  });
}

}  // extern "C"
