/*
 * Allocation-free prediction of the learned cost model (no counterpart in
 * the reference): adatm_mlp_predict_noalloc gives adatm_mlp_predict's
 * results without any device allocation, for callers that cannot allocate
 * (code that sets up HIP graph captures, tight planning loops).
 *
 * The activations of the forward pass live in `scratch`, a file-scope device
 * array of ADATM_NOALLOC_MAX_ROWS rows (6.7 MB); every GPU the library runs
 * on holds its own copy until the process ends.  Host code never passes the
 * array's name to a kernel or a copy: on the first call on a device, with
 * that device current, hipGetSymbolAddress gives that device's address of
 * its copy, which is kept per device and used from then on.  All calls on
 * one device share its copy: calls that overlap on one GPU (from several
 * host threads) race; one call finishes (stream synchronised) before it
 * returns, so calls in sequence are safe.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <mutex>

#include "adatm_kern.h"
#include "../../../include/adatm.h"

namespace {

constexpr uint32_t NF = ADATM_MLP_IN;
constexpr uint32_t NH = ADATM_MLP_HID;
/* floats per row: the standardised input, then Z and H of the four layers */
constexpr uint32_t ROWF = NF + 2u * (NH + NH + NH + 1u);
constexpr uint32_t MAXDEV = 64;

static __device__ __align__(256) float scratch[(size_t)ADATM_NOALLOC_MAX_ROWS * ROWF];

float *g_ptr[MAXDEV];
std::mutex g_mu;

struct norm_t {
  float mean[NF], sd[NF];
};

/* x = (x - mean) / sd in place, the host's arithmetic of adatm_mlp_predict
 * (a subtraction and a correctly rounded division per element) */
__global__ __launch_bounds__(256) void k_standardize(float *__restrict__ x, uint32_t n,
                                                     norm_t nm)
{
  uint32_t const e = blockIdx.x * 256u + threadIdx.x;
  if (e >= n * NF) return;
  uint32_t const i = e % NF;
  x[e] = (x[e] - nm.mean[i]) / nm.sd[i];
}

/* this device's address of `scratch`, looked up once per device with the
 * device current */
int scratch_ptr(int device, float **out)
{
  if (device < 0 || (uint32_t)device >= MAXDEV) return (int)hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) {
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return (int)e;
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return (int)(e != hipSuccess ? e : hipErrorInvalidSymbol);
    g_ptr[device] = (float *)p;
  }
  *out = g_ptr[device];
  return 0;
}

} // namespace

extern "C" {

char const *adatm_last_error(void);
void adatm_set_error(char const *fmt, ...);

int adatm_mlp_predict_noalloc(adatm_mlp *M, float const *X, uint32_t n, float *out)
{
  if (!M || !X || !out || n == 0) return SPLATT_ERROR_BADINPUT;
  if (n > ADATM_NOALLOC_MAX_ROWS) {
    adatm_set_error("adatm_mlp_predict_noalloc: %u rows, at most %u", n,
                    (unsigned)ADATM_NOALLOC_MAX_ROWS);
    return SPLATT_ERROR_BADINPUT;
  }
  int device;
  hipStream_t s;
  norm_t nm;
  float ymean, ystd;
  adatm_mlp_dev_info(M, &device, &s, nm.mean, nm.sd, &ymean, &ystd);
  float *base = nullptr;
  int const k = scratch_ptr(device, &base);
  if (k != 0) {
    adatm_set_error("adatm_mlp_predict_noalloc: scratch address on device %d: %s", device,
                    hipGetErrorString((hipError_t)k));
    return SPLATT_ERROR_DEVICE;
  }
  if (hipSetDevice(device) != hipSuccess) return SPLATT_ERROR_DEVICE;
  /* rows of n: x, then Z[l] and H[l] of each layer */
  float *x = base;
  float *Z[4], *H[4];
  float *q = x + (size_t)n * NF;
  uint32_t const out_w[4] = {NH, NH, NH, 1u};
  for (int l = 0; l < 4; ++l) {
    Z[l] = q;
    q += (size_t)n * out_w[l];
    H[l] = q;
    q += (size_t)n * out_w[l];
  }
  if (hipMemcpyAsync(x, X, (size_t)n * NF * sizeof(float), hipMemcpyHostToDevice, s) !=
      hipSuccess) {
    adatm_set_error("adatm_mlp_predict_noalloc: input upload failed");
    return SPLATT_ERROR_DEVICE;
  }
  hipLaunchKernelGGL(k_standardize, dim3((n * NF + 255u) / 256u), dim3(256), 0, s, x, n,
                     nm);
  if (hipGetLastError() != hipSuccess) {
    adatm_set_error("adatm_mlp_predict_noalloc: standardisation launch failed");
    return SPLATT_ERROR_DEVICE;
  }
  int const rc = adatm_mlp_forward_dev(M, x, Z, H, n);
  if (rc != SPLATT_SUCCESS) return rc;
  if (hipMemcpyAsync(out, H[3], (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    adatm_set_error("adatm_mlp_predict_noalloc: download failed");
    return SPLATT_ERROR_DEVICE;
  }
  for (uint32_t r = 0; r < n; ++r) out[r] = out[r] * ystd + ymean;
  return SPLATT_SUCCESS;
}

} /* extern "C" */
