// vjx_lu_solve_noalloc: the triangular solves (and FP32 refinement) of
// vjx_lu_solve with every intermediate vector in a fixed device array, so a
// call allocates nothing.  For callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops).
//
// The array `scratch` below is file-scope device memory of a fixed size; each
// GPU holds its own copy until the process ends.  Host code never passes its
// name to a kernel: on the host it is only a stand-in, and the device address
// of each GPU's copy comes from hipGetSymbolAddress, looked up once per device
// (with that device current) and kept in g_scratch[device].
//
// Every call on one GPU shares that GPU's array and leaves its contents
// behind.  Nothing in it has to start at zero: the forward solve writes every
// intermediate entry before it is read, and every hub partial is written
// before its row is finished.  Calls that overlap on one GPU (two host threads,
// or two streams) race on the array.
#include <mutex>

#include "circuit.h"

namespace vjx {
namespace {

// Largest system (unknowns) and largest number of 4096-entry hub chunks per
// hub plan the array serves.
constexpr size_t kScratchMaxN = 1u << 19;
constexpr size_t kScratchMaxChunks = 1u << 14;
// layout in doubles: y, z (FP32 solves store floats in the same space), the
// refinement residual r and correction dx, then the forward solve's, the
// backward solve's and the product's hub chunk partials
constexpr size_t kOffY = 0, kOffZ = kScratchMaxN, kOffR = 2 * kScratchMaxN,
                 kOffDX = 3 * kScratchMaxN, kOffFwd = 4 * kScratchMaxN,
                 kOffBwd = kOffFwd + kScratchMaxChunks, kOffSpmv = kOffBwd + kScratchMaxChunks,
                 kScratchDoubles = kOffSpmv + kScratchMaxChunks;

__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// Device address of the current device's copy of scratch (looked up once).
int scratch_address(double** out) {
  int dev = 0;
  VJX_TRY_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("vjx_lu_solve_noalloc: device index beyond the scratch table");
    return VJX_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    VJX_TRY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      set_error("vjx_lu_solve_noalloc: no device address for the scratch array");
      return VJX_ERR_HIP;
    }
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return VJX_OK;
}

}  // namespace
}  // namespace vjx

using namespace vjx;

extern "C" int vjx_lu_solve_noalloc(vjx_pattern* pat, const double* data_dev, const double* b_dev,
                                    double* x_dev, int32_t refine_steps) {
  if (!pat || !b_dev || !x_dev || refine_steps < 0) {
    set_error("vjx_lu_solve_noalloc: null argument or negative refine_steps");
    return VJX_ERR_ARG;
  }
  if ((size_t)pat->n > kScratchMaxN) {
    set_error("vjx_lu_solve_noalloc: " + std::to_string(pat->n) +
              " unknowns exceed the scratch limit of " + std::to_string(kScratchMaxN));
    return VJX_ERR_ARG;
  }
  DevLU& L = pat->lu;
  if (L.factored_prec < 0 || !pat->stream) {
    set_error("vjx_lu_solve_noalloc: call vjx_lu_refactor first");
    return VJX_ERR_STATE;
  }
  const bool refine = L.factored_prec == VJX_PREC_F32 && refine_steps > 0;
  if (refine && !data_dev) {
    set_error("vjx_lu_solve_noalloc: refinement needs the FP64 matrix");
    return VJX_ERR_ARG;
  }
  if (refine && !pat->d_indptr.p) {
    set_error("vjx_lu_solve_noalloc: the pattern's CSR arrays are not on the device");
    return VJX_ERR_STATE;
  }
  if (L.fwd_band.part.n > kScratchMaxChunks || L.bwd_band.part.n > kScratchMaxChunks ||
      (size_t)pat->hub_rows.n_chunks > kScratchMaxChunks) {
    set_error("vjx_lu_solve_noalloc: hub chunks exceed the scratch limit of " +
              std::to_string(kScratchMaxChunks));
    return VJX_ERR_ARG;
  }
  double* base = nullptr;
  int rc = scratch_address(&base);
  if (rc < 0) return rc;
  hipStream_t s = pat->stream;
  void* parts[2] = {base + kOffFwd, base + kOffBwd};
  if (L.factored_prec == VJX_PREC_F64) {
    launch_lu_solve<double>(L, L.lu64.p, b_dev, 1.0, x_dev, base + kOffY, base + kOffZ, s, parts);
  } else {
    float* y = reinterpret_cast<float*>(base + kOffY);
    float* z = reinterpret_cast<float*>(base + kOffZ);
    launch_lu_solve<float>(L, L.lu32.p, b_dev, 1.0, x_dev, y, z, s, parts);
    if (refine) {
      HubChunks hub = pat->hub_rows.view();
      hub.part = base + kOffSpmv;
      for (int r = 0; r < refine_steps; ++r) {
        // r = b - A x, then A dx = r, x += dx (as vjx_lu_solve)
        launch_spmv(pat->n, pat->d_indptr.p, pat->d_indices.p, data_dev, x_dev, b_dev, -1.0,
                    base + kOffR, hub, s);
        launch_lu_solve<float>(L, L.lu32.p, base + kOffR, 1.0, base + kOffDX, y, z, s, parts);
        launch_axpy(pat->n, x_dev, base + kOffDX, s);
      }
    }
  }
  VJX_TRY_HIP(hipGetLastError());
  VJX_TRY_HIP(hipStreamSynchronize(s));
  return VJX_OK;
}
