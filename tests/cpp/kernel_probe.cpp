// Short run for hardware-counter passes (rocprofv3 --pmc): one 64^3 Q2 SUPG
// system, 6 FP32 stencil SpMV launches and 3 FP32 SSOR sweeps (54 colour
// launches each) on the finest level, then 5 device copies of 1 GiB
// (hipMemcpy device to device, the calibration kernel of known size).
// Prints the time per SpMV launch and per sweep.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "pmx.h"

int main() {
  pmx_ctx *ctx = nullptr;
  pmx_system *sys = nullptr;
  if (pmx_ctx_create(0, &ctx) || pmx_system_create(ctx, 2, 5, 2, &sys)) {
    std::printf("setup failed: %s\n", pmx_last_error());
    return 1;
  }
  std::vector<double> prm(PMX_PRM_LEN, 0.0);
  prm[PMX_PRM_EXAMPLE] = PMX_EX_SINE_CD;
  prm[PMX_PRM_EPS] = 1e-4;
  prm[PMX_PRM_B0] = 0.2672612419124244;
  prm[PMX_PRM_B1] = 0.5345224838248488;
  prm[PMX_PRM_B2] = 0.8017837257372732;
  prm[PMX_PRM_FORM] = PMX_FORM_SDFEM;
  prm[PMX_PRM_SDFEM_TYPE] = 2;
  prm[PMX_PRM_DELTA0] = 1.0;
  prm[PMX_PRM_DELTA1] = 1.0;
  prm[PMX_PRM_LINEAR] = 1;
  prm[PMX_PRM_POLYDEG] = 2;
  if (pmx_system_assemble(sys, prm.data())) {
    std::printf("assemble failed: %s\n", pmx_last_error());
    return 1;
  }
  double ms = 0;
  if (pmx_time_kernel(sys, 0, -1, 1, 5, &ms)) {
    std::printf("time_kernel failed: %s\n", pmx_last_error());
    return 1;
  }
  std::printf("fp32 spmv %.4f ms per launch\n", ms);
  if (pmx_time_kernel(sys, 3, -1, 1, 2, &ms)) {
    std::printf("time_kernel failed: %s\n", pmx_last_error());
    return 1;
  }
  std::printf("fp32 ssor sweep %.4f ms\n", ms);
  const size_t bytes = size_t(1) << 30;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
  (void)hipMemset(a, 0, bytes);
  for (int k = 0; k < 5; k++) (void)hipMemcpy(b, a, bytes, hipMemcpyDeviceToDevice);
  (void)hipDeviceSynchronize();
  std::printf("copied 5 x %zu bytes\n", bytes);
  (void)hipFree(a);
  (void)hipFree(b);
  pmx_system_destroy(sys);
  pmx_ctx_destroy(ctx);
  return 0;
}
