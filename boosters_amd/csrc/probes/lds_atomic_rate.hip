// lds_atomic_rate.hip — LDS throughput probe (not part of the library).
//
// Measures lane-operations per clock per CU of ds_add_u64, ds_add_u32,
// ds_write_b64 and two ds_add_u32 on the histogram kernel's [bin][32] LDS
// layout (1024-thread blocks, random bins).  Build and run:
//   hipcc -O3 --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops \
//       boosters_amd/csrc/probes/lds_atomic_rate.hip -o /tmp/lds_atomic_rate && /tmp/lds_atomic_rate
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
#define CHECK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("%s: %s\n", #x, hipGetErrorString(e)); return 1; } } while (0)
constexpr int kT = 1024, kIters = 4096;

template <int MODE>
__global__ __launch_bounds__(kT) void k(unsigned long long* out, uint32_t seed) {
  __shared__ unsigned long long s64[256 * 32];
  __shared__ uint32_t s32[256 * 32 * 2];
  for (int i = threadIdx.x; i < 256 * 32; i += kT) { s64[i] = 0; s32[2 * i] = 0; s32[2 * i + 1] = 0; }
  __syncthreads();
  const int lf = threadIdx.x & 31;
  uint32_t x = seed ^ (threadIdx.x * 2654435761u) ^ (blockIdx.x * 40503u);
  for (int it = 0; it < kIters; it++) {
    x = x * 1664525u + 1013904223u;
    const uint32_t bin = x >> 24;
    if (MODE == 0) atomicAdd(&s64[bin * 32 + lf], (unsigned long long)x);
    if (MODE == 1) atomicAdd(&s32[bin * 64 + lf], x);
    if (MODE == 3) s64[bin * 32 + lf] = x;  // plain store
    if (MODE == 4) { atomicAdd(&s32[bin * 64 + 2 * lf], x); atomicAdd(&s32[bin * 64 + 2 * lf + 1], x >> 7); }
  }
  __syncthreads();
  unsigned long long acc = 0;
  for (int i = threadIdx.x; i < 256 * 32; i += kT) acc += s64[i] + s32[2 * i];
  if (acc == 0x1234567ull) out[blockIdx.x] = acc;  // keep the work alive
}

int main() {
  unsigned long long* out;
  CHECK(hipMalloc(&out, 65536 * 8));
  hipDeviceProp_t p;
  CHECK(hipGetDeviceProperties(&p, 0));
  const int cus = p.multiProcessorCount;
  const double clk = p.clockRate * 1e3;  // Hz
  printf("CUs %d clock %.0f MHz\n", cus, clk / 1e6);
  const char* names[] = {"ds_add_u64", "ds_add_u32", "", "ds_write_b64", "2x ds_add_u32"};
  for (int mode = 0; mode < 5; mode++) {
    if (mode == 2) continue;
    for (int bpc : {1, 2, 4}) {
      const int grid = cus * bpc * 2;
      hipEvent_t a, b;
      CHECK(hipEventCreate(&a)); CHECK(hipEventCreate(&b));
      for (int rep = 0; rep < 2; rep++) {
        CHECK(hipEventRecord(a));
        switch (mode) {
          case 0: hipLaunchKernelGGL(k<0>, dim3(grid), dim3(kT), 0, 0, out, 7u); break;
          case 1: hipLaunchKernelGGL(k<1>, dim3(grid), dim3(kT), 0, 0, out, 7u); break;
          case 3: hipLaunchKernelGGL(k<3>, dim3(grid), dim3(kT), 0, 0, out, 7u); break;
          case 4: hipLaunchKernelGGL(k<4>, dim3(grid), dim3(kT), 0, 0, out, 7u); break;
        }
        CHECK(hipEventRecord(b));
        CHECK(hipEventSynchronize(b));
        float ms; CHECK(hipEventElapsedTime(&ms, a, b));
        if (rep == 1) {
          const double ops = (double)grid * kT * kIters * (mode == 4 ? 2 : 1);
          printf("%-14s grid %6d  %.3f ms  %.2f lane-ops/clk/CU\n", names[mode], grid, ms, ops / (ms * 1e-3) / clk / cus);
        }
      }
    }
  }
  return 0;
}
