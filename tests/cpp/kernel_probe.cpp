// Short run for hardware-counter passes (rocprofv3 --pmc) on the 64^3 Q2 SUPG
// system of the benchmark: the parameter block comes from a file of
// PMX_PRM_LEN doubles (written by `python -m parmoon_amd.prm config2 FILE`,
// the curl-Fourier convection of the benchmark).  Five levels 4^3 .. 64^3,
// then on the finest level: 6 FP32 stencil SpMV launches, 3 FP32 SSOR sweeps
// (53 colour launches each; the first SpMV and sweep are untimed), 3 SSOR
// sweeps on the row-scaled FP16 values (k_sor_scaled, the first untimed), the
// calibration dispatches of the colour launches (empty dispatches with their
// workgroup counts, streaming reads of each colour's FP32 and FP16 values) and
// 5 device copies of 1 GiB (hipMemcpy device to device, the calibration kernel
// of known size).  Prints the average dispatch durations.
//
//   kernel_probe PRM_FILE
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "pmx.h"

int main(int argc, char **argv) {
  if (argc < 2) {
    std::printf("usage: kernel_probe PRM_FILE\n");
    return 2;
  }
  std::vector<double> prm(PMX_PRM_LEN, 0.0);
  FILE *fp = std::fopen(argv[1], "rb");
  if (!fp || std::fread(prm.data(), sizeof(double), PMX_PRM_LEN, fp) != (size_t)PMX_PRM_LEN) {
    std::printf("cannot read %d doubles from %s\n", PMX_PRM_LEN, argv[1]);
    return 2;
  }
  std::fclose(fp);
  pmx_ctx *ctx = nullptr;
  pmx_system *sys = nullptr;
  if (pmx_ctx_create(0, &ctx) || pmx_system_create(ctx, 2, 5, 2, &sys)) {
    std::printf("setup failed: %s\n", pmx_last_error());
    return 1;
  }
  if (pmx_system_assemble(sys, prm.data())) {
    std::printf("assemble failed: %s\n", pmx_last_error());
    return 1;
  }
  double ms = 0;
  if (pmx_time_kernel(sys, 11, -1, 1, 5, &ms)) {
    std::printf("time_kernel failed: %s\n", pmx_last_error());
    return 1;
  }
  std::printf("fp32 spmv %.4f ms per dispatch\n", ms);
  if (pmx_time_kernel(sys, 10, -1, 1, 2, &ms)) {
    std::printf("time_kernel failed: %s\n", pmx_last_error());
    return 1;
  }
  std::printf("fp32 ssor colour launch %.4f ms per dispatch\n", ms);
  if (pmx_time_kernel(sys, 10, -1, 3, 2, &ms)) {
    std::printf("time_kernel failed: %s\n", pmx_last_error());
    return 1;
  }
  std::printf("fp32 ssor colour launch, FP16 values %.4f ms per dispatch\n", ms);
  // calibration of the colour launches: an empty dispatch with each colour
  // launch's workgroup count, and a plain streaming read of each colour's
  // stored values (FP32, then the FP16 smoother values)
  const struct {
    int kernel, precision;
    const char *what;
  } cal[] = {{12, 1, "empty dispatch, colour-launch grids"},
             {13, 1, "stream read of a colour's FP32 values"},
             {13, 3, "stream read of a colour's FP16 values"}};
  for (const auto &c : cal) {
    if (pmx_time_kernel(sys, c.kernel, -1, c.precision, 2, &ms)) {
      std::printf("time_kernel failed: %s\n", pmx_last_error());
      return 1;
    }
    std::printf("%s %.4f ms per dispatch\n", c.what, ms);
  }
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
