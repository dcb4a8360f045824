// Grid-barrier probe: the cost of one cooperative-groups grid.sync() at 1, 2,
// 4 and 8 resident workgroups per CU (hipLaunchCooperativeKernel, which fails
// instead of launching a grid that cannot be co-resident), against 1000
// back-to-back launches of an empty kernel on one stream.  It prices a
// persistent batched filter (one launch, one grid barrier per Clenshaw step)
// against the graph of per-step launches.  A stand-alone measurement program,
// not part of the library:
//
//   hipcc -O3 --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         -o /tmp/gridsync_probe sparse_matrix_algorithms_amd/gridsync_probe.hip
//   /tmp/gridsync_probe
#include <hip/hip_runtime.h>
#include <hip/hip_cooperative_groups.h>
#include <cstdio>
#include <vector>
namespace cg = cooperative_groups;
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { std::printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)

__global__ __launch_bounds__(256) void sync_kernel(int iters, float* out) {
  cg::grid_group g = cg::this_grid();
  float acc = 0.f;
  for (int i = 0; i < iters; ++i) { acc += 1.f; g.sync(); }
  if (threadIdx.x == 0) out[blockIdx.x] = acc;
}
__global__ __launch_bounds__(256) void empty_kernel(float* out) {
  if (threadIdx.x == 0) out[blockIdx.x] += 1.f;
}

int main() {
  hipDeviceProp_t p; CK(hipGetDeviceProperties(&p, 0));
  int occ = 0; CK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, sync_kernel, 256, 0));
  std::printf("CUs %d, cooperative %d, max resident blocks/CU %d\n", p.multiProcessorCount, p.cooperativeLaunch, occ);
  const int maxg = p.multiProcessorCount * occ;
  float* out; CK(hipMalloc(&out, sizeof(float) * maxg));
  CK(hipMemset(out, 0, sizeof(float) * maxg));
  hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
  for (int per : {1, 2, 4, 8}) {
    if (per > occ) break;
    const int grid = p.multiProcessorCount * per;
    for (int iters : {10, 1000}) {
      void* args[] = {(void*)&iters, (void*)&out};
      CK(hipLaunchCooperativeKernel((void*)sync_kernel, dim3(grid), dim3(256), args, 0, 0));
      CK(hipDeviceSynchronize());
      CK(hipEventRecord(e0, 0));
      CK(hipLaunchCooperativeKernel((void*)sync_kernel, dim3(grid), dim3(256), args, 0, 0));
      CK(hipEventRecord(e1, 0));
      CK(hipEventSynchronize(e1));
      float ms = 0.f; CK(hipEventElapsedTime(&ms, e0, e1));
      std::printf("grid %d blocks (%d/CU), %d grid syncs: %.3f ms -> %.2f us per sync\n", grid, per, iters, ms, ms * 1e3 / iters);
    }
  }
  for (int per : {1, 8}) {
    const int grid = p.multiProcessorCount * per;
    if (grid > maxg) break;
    for (int r = 0; r < 2; ++r) {
      CK(hipEventRecord(e0, 0));
      for (int i = 0; i < 1000; ++i) hipLaunchKernelGGL(empty_kernel, dim3(grid), dim3(256), 0, 0, out);
      CK(hipEventRecord(e1, 0));
      CK(hipEventSynchronize(e1));
      float ms = 0.f; CK(hipEventElapsedTime(&ms, e0, e1));
      std::printf("1000 back-to-back launches of %d blocks: %.3f ms -> %.2f us per launch\n", grid, ms, ms);
    }
  }
  CK(hipFree(out));
  return 0;
}
