// Test harness for the device's correctly rounded math (cylag_amd/csrc/cyl_crmath.h):
// evaluates, for each argument, the fast path (its value and whether its error
// bound decided the rounding), the double-double evaluation alone, and the
// function as the step kernels call it.  Test infrastructure only.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../cylag_amd/csrc/cyl_crmath.h"

namespace {

__global__ void k_eval(int fn, int n, const double* x, const double* y, double* fast, double* full, double* fin,
                       int* decided) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = x[i], b = y ? y[i] : 0.;
  double f = 0., d = 0., c = 0.;
  bool ok = false;
  if (fn == 0) {
    ok = cyl::cr_exp_try(a, f);
    d = cyl::cr_exp_dd(a);
    c = cyl::cr_exp(a);
  } else if (fn == 1) {
    ok = cyl::cr_log_try(a, f);
    d = cyl::cr_log_dd(a);
    c = cyl::cr_log(a);
  } else if (fn == 2) {
    ok = cyl::cr_pow_try(a, b, f);
    d = cyl::cr_pow_dd(a, b);
    c = cyl::cr_pow(a, b);
  } else {
    ok = cyl::cr_expm1_try(a, f);
    d = cyl::cr_expm1_dd(a);
    c = cyl::cr_expm1(a);
  }
  fast[i] = f;
  full[i] = d;
  fin[i] = c;
  decided[i] = ok ? 1 : 0;
}

}  // namespace

// fn: 0 exp, 1 log, 2 pow (x^y), 3 expm1.  Arguments must lie where the fast
// and double-double paths are defined: exp [-708, 709.78]; log positive normal;
// pow x positive normal, x != 1, y != 0, y log x in [-708, 709.78]; expm1 [-38, 709.78].
extern "C" int crm_eval(int fn, int n, const double* x, const double* y, double* fast, double* full, double* fin,
                        int* decided) {
  if (n <= 0 || fn < 0 || fn > 3 || (fn == 2 && !y)) return 1;
  double *dx = nullptr, *dy = nullptr, *df = nullptr, *dd = nullptr, *dc = nullptr;
  int* dk = nullptr;
  const size_t b = (size_t)n * sizeof(double);
  int rc = 0;
  if (hipMalloc(&dx, b) || hipMalloc(&df, b) || hipMalloc(&dd, b) || hipMalloc(&dc, b) ||
      hipMalloc(&dk, (size_t)n * sizeof(int)) || (y && hipMalloc(&dy, b)))
    rc = 2;
  if (!rc && hipMemcpy(dx, x, b, hipMemcpyHostToDevice)) rc = 3;
  if (!rc && y && hipMemcpy(dy, y, b, hipMemcpyHostToDevice)) rc = 3;
  if (!rc) {
    k_eval<<<(n + 255) / 256, 256>>>(fn, n, dx, dy, df, dd, dc, dk);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = 4;
  }
  if (!rc && (hipMemcpy(fast, df, b, hipMemcpyDeviceToHost) || hipMemcpy(full, dd, b, hipMemcpyDeviceToHost) ||
              hipMemcpy(fin, dc, b, hipMemcpyDeviceToHost) ||
              hipMemcpy(decided, dk, (size_t)n * sizeof(int), hipMemcpyDeviceToHost)))
    rc = 5;
  for (void* p : {(void*)dx, (void*)dy, (void*)df, (void*)dd, (void*)dc, (void*)dk})
    if (p) (void)hipFree(p);
  return rc;
}
