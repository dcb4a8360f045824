// HBM bandwidth probe: copy, read-only and write-only kernels over 1 GiB and
// 4 GiB buffers (beyond the 256 MB Infinity Cache), with 4 or 8 16-byte loads
// in flight per lane, plain or non-temporal, at 4..32 workgroups per CU or one
// pass per workgroup; hipMemcpy device-to-device for comparison.  A stand-alone
// measurement program, not part of the library:
//
//   hipcc -O3 --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         -o /tmp/hbm_probe sparse_matrix_algorithms_amd/hbm_probe.hip
//   /tmp/hbm_probe
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <vector>
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP %s line %d\n", hipGetErrorString(e), __LINE__); exit(1); } } while (0)

template <int U, bool NT>
__global__ __launch_bounds__(256) void copy_k(const float4* __restrict__ a, float4* __restrict__ b, size_t n4) {
  const size_t per = (size_t)U * 256;
  for (size_t base = (size_t)blockIdx.x * per; base + per <= n4; base += (size_t)gridDim.x * per) {
    float4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      typedef float f4v __attribute__((ext_vector_type(4)));
      const f4v* p = reinterpret_cast<const f4v*>(a + base + (size_t)u * 256 + threadIdx.x);
      f4v t = NT ? __builtin_nontemporal_load(p) : *p;
      v[u] = make_float4(t.x, t.y, t.z, t.w);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      typedef float f4v __attribute__((ext_vector_type(4)));
      f4v* p = reinterpret_cast<f4v*>(b + base + (size_t)u * 256 + threadIdx.x);
      f4v t = {v[u].x, v[u].y, v[u].z, v[u].w};
      if (NT) __builtin_nontemporal_store(t, p); else *p = t;
    }
  }
}
template <int U>
__global__ __launch_bounds__(256) void read_k(const float4* __restrict__ a, float* __restrict__ out, size_t n4) {
  const size_t per = (size_t)U * 256;
  float s = 0.f;
  for (size_t base = (size_t)blockIdx.x * per; base + per <= n4; base += (size_t)gridDim.x * per) {
    float4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = a[base + (size_t)u * 256 + threadIdx.x];
#pragma unroll
    for (int u = 0; u < U; ++u) s += v[u].x + v[u].y + v[u].z + v[u].w;
  }
  if (s == 12345.f) out[threadIdx.x] = s;  // keeps the loads; never true for the fill value
}
template <int U>
__global__ __launch_bounds__(256) void write_k(float4* __restrict__ b, size_t n4) {
  const size_t per = (size_t)U * 256;
  for (size_t base = (size_t)blockIdx.x * per; base + per <= n4; base += (size_t)gridDim.x * per)
#pragma unroll
    for (int u = 0; u < U; ++u) b[base + (size_t)u * 256 + threadIdx.x] = make_float4(1, 2, 3, 4);
}

int main() {
  int cu = 0; CK(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, 0));
  const size_t maxbytes = size_t(4) << 30;
  float4 *a, *b; float* o;
  CK(hipMalloc(&a, maxbytes)); CK(hipMalloc(&b, maxbytes)); CK(hipMalloc(&o, 4096));
  CK(hipMemset(a, 0, maxbytes)); CK(hipMemset(b, 0, maxbytes));
  hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
  for (size_t mib : {1024, 4096}) {
    const size_t bytes = mib << 20, n4 = bytes / 16;
    auto timeit = [&](const char* name, auto launch, double moved) {
      launch(); CK(hipDeviceSynchronize());
      CK(hipEventRecord(e0)); for (int r = 0; r < 10; ++r) launch(); CK(hipEventRecord(e1));
      CK(hipEventSynchronize(e1)); float ms; CK(hipEventElapsedTime(&ms, e0, e1));
      printf("%5zu MiB %-28s %7.1f us  %6.0f GB/s\n", mib, name, ms * 100.0, moved * 10 / (ms * 1e-3) / 1e9);
      fflush(stdout);
    };
    for (int wpc : {4, 8, 16, 32, 0}) {  // workgroups per CU (0: one trip each)
      auto grid = [&](int U) { size_t ch = n4 / (U * 256); return (int)(wpc ? std::min<size_t>(ch, (size_t)wpc * cu) : ch); };
      char nm[64];
      snprintf(nm, 64, "copy U4 wg/cu=%d", wpc); timeit(nm, [&] { hipLaunchKernelGGL((copy_k<4, false>), dim3(grid(4)), dim3(256), 0, 0, a, b, n4); }, 2.0 * bytes);
      snprintf(nm, 64, "copy U8 wg/cu=%d", wpc); timeit(nm, [&] { hipLaunchKernelGGL((copy_k<8, false>), dim3(grid(8)), dim3(256), 0, 0, a, b, n4); }, 2.0 * bytes);
      snprintf(nm, 64, "copy U4 NT wg/cu=%d", wpc); timeit(nm, [&] { hipLaunchKernelGGL((copy_k<4, true>), dim3(grid(4)), dim3(256), 0, 0, a, b, n4); }, 2.0 * bytes);
      snprintf(nm, 64, "read U8 wg/cu=%d", wpc); timeit(nm, [&] { hipLaunchKernelGGL((read_k<8>), dim3(grid(8)), dim3(256), 0, 0, a, o, n4); }, 1.0 * bytes);
      snprintf(nm, 64, "write U8 wg/cu=%d", wpc); timeit(nm, [&] { hipLaunchKernelGGL((write_k<8>), dim3(grid(8)), dim3(256), 0, 0, b, n4); }, 1.0 * bytes);
    }
    timeit("hipMemcpyDtoD", [&] { CK(hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, 0)); }, 2.0 * bytes);
  }
  return 0;
}
