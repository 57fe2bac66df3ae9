// Shared pieces of the C-ABI layer: the context handle, error translation
// and argument checks.  Every extern "C" entry point runs inside guarded(),
// so no C++ exception crosses the ABI.
#pragma once
#include "hfgpu_internal.hpp"
#include "../../include/hfgpu.h"
#include <string>

struct hf_ctx {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t cur = nullptr;
  hfg::KernelTimer timer;
  void *blas = nullptr;  // rocblas_handle of the batched eigensolver, created on first use
};
// destroys ctx->blas (capi_eig.hip)
void hf_ctx_release_blas(hf_ctx *ctx);

struct hf_atomic {
  hf_ctx *ctx = nullptr;
  hfg::AtomicSetup host;
  hfg::AtomicDevice dev;
  bool uploaded = false;
  bool rs_uploaded = false;  // range-separated tables on the device
  // radial one-electron matrices (Nrad x Nrad) on the device, per kind, and
  // the kinetic l(l+1) factor's X; the shells' l values
  hfg::DevArray<double> one_M[3], one_X;
  hfg::DevArray<int> lval_dev;
};

struct hf_diatomic {
  hf_ctx *ctx = nullptr;
  hfg::DiatomicSetup host;
  hfg::DiatomicDevice dev;
  bool absm = false;
  bool uploaded = false;
};

namespace hfg {

extern thread_local std::string g_err;

template <typename F>
int guarded(F &&f) {
  try {
    f();
    return HF_OK;
  } catch (const Error &e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_err = "host allocation failed";
    return HF_ENOMEM;
  } catch (const std::logic_error &e) {
    g_err = e.what();
    return HF_EINVAL;
  } catch (const std::exception &e) {
    g_err = e.what();
    return HF_EINTERNAL;
  }
}

inline void need(bool cond, const char *msg) {
  if (!cond) throw Error(HF_EINVAL, msg);
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    hip_check(hipGetDevice(&prev), "hipGetDevice");
    if (prev != dev) hip_check(hipSetDevice(dev), "hipSetDevice");
  }
  ~DeviceGuard() {
    int now = -1;
    if (hipGetDevice(&now) == hipSuccess && now != prev && prev >= 0) (void)hipSetDevice(prev);
  }
};

inline void check_matrix(const void *p, int64_t ld, int64_t n, const char *name) {
  if (!p) throw Error(HF_EINVAL, std::string(name) + " is NULL");
  if (ld < n) throw Error(HF_EINVAL, std::string("leading dimension of ") + name + " is smaller than Nbf");
}

// Host-pointer entry points stage P and the result through device buffers.
template <typename Launch>
void host_roundtrip(hipStream_t s, int64_t n, DevArray<double> &dP, DevArray<double> &dO, const double *P,
                    int64_t ldp, double *O, int64_t ldo, Launch launch) {
  check_matrix(P, ldp, n, "P");
  check_matrix(O, ldo, n, "output");
  dP.alloc((size_t)n * n);
  dO.alloc((size_t)n * n);
  hip_check(hipMemcpy2DAsync(dP.p, n * sizeof(double), P, ldp * sizeof(double), n * sizeof(double), n,
                             hipMemcpyHostToDevice, s),
            "copy P to device");
  launch(dP.p, n, dO.p, n);
  hip_check(hipMemcpy2DAsync(O, ldo * sizeof(double), dO.p, n * sizeof(double), n * sizeof(double), n,
                             hipMemcpyDeviceToHost, s),
            "copy result to host");
  hip_check(hipStreamSynchronize(s), "stream sync");
}

constexpr double kFourPi = 12.566370614359172953850573533118;

}  // namespace hfg
