// Allocation-free entry point: the squared 2-norm of an FP64 vector (the Newton
// convergence test's reduction, gm_norm2) with its partial sums in a file-scope
// device array instead of a caller workspace, for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops).
//
// The array's host-side name is only a stand-in for its device address: each
// device holds its own copy, whose address is looked up once per device with
// hipGetSymbolAddress (that device current) and kept; kernels receive that
// address.  Every call on one device shares the array, so calls that overlap on
// one device (different streams, no ordering) race; calls ordered on a stream
// do not.  The partials are fully rewritten by every call (nothing is read
// before it is written), so no clearing is needed.
//
// The reduction is the one gm_norm2 runs: at most 1024 workgroups of 256 lanes
// with a grid-stride loop, per-workgroup tree sums, then one workgroup summing
// the partials in index order -- the same arithmetic in the same order, so the
// result is bitwise equal to gm_norm2's.
#include <hip/hip_runtime.h>

#include <mutex>

#include "gm_common.h"

namespace {
constexpr int kT = 256;          // lanes per workgroup (gm_norm2's RED_T)
constexpr int kParts = 1024;     // most workgroups a call launches
constexpr long kMaxN = 1L << 40; // largest vector length accepted

__device__ __align__(256) double scratch[2 * kParts];

std::mutex g_mu;
double* g_addr[256];

__global__ void k_sq_parts(long n, const double* v, double* part) {
  __shared__ double a[kT], b[kT];
  double s0 = 0.0;
  for (long i = blockIdx.x * (long)kT + threadIdx.x; i < n; i += (long)gridDim.x * kT) s0 += v[i] * v[i];
  a[threadIdx.x] = s0;
  b[threadIdx.x] = 0.0;
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      a[threadIdx.x] += a[threadIdx.x + o];
      b[threadIdx.x] += b[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a[0];
    part[2 * blockIdx.x + 1] = b[0];
  }
}

__global__ void k_sum_pairs(const double* part, int n, double* out) {
  __shared__ double a[kT];
  double s0 = 0.0;
  for (int i = threadIdx.x; i < n; i += kT) s0 += part[2 * i];
  a[threadIdx.x] = s0;
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) a[threadIdx.x] += a[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = a[0];
}

// Device address of `scratch` on the current device (looked up on first use).
int scratch_address(double** out) {
  int dev = -1;
  GM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 256) {
    gm::set_error("gm_norm2_noalloc: device index %d out of range", dev);
    return -1;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_addr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      gm::set_error("gm_norm2_noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                    hipGetErrorString(e));
      return -2;
    }
    g_addr[dev] = (double*)p;
  }
  *out = g_addr[dev];
  return 0;
}
}  // namespace

extern "C" int gm_norm2_noalloc(long n, const double* v, double* out, hipStream_t s) {
  if (n < 0 || n > kMaxN) {
    gm::set_error("gm_norm2_noalloc: n = %ld outside [0, 2^40]", n);
    return -1;
  }
  GM_REQUIRE(out && (v || n == 0), "gm_norm2_noalloc: null vector or output");
  double* part = nullptr;
  int rc = scratch_address(&part);
  if (rc) return rc;
  long b = (n + kT - 1) / kT;
  int blocks = (int)(b > kParts ? kParts : (b < 1 ? 1 : b));
  k_sq_parts<<<blocks, kT, 0, s>>>(n, v, part);
  k_sum_pairs<<<1, kT, 0, s>>>(part, blocks, out);
  GM_LAUNCH_CHECK();
  return 0;
}
