// hbm_probe.hip -- achievable HBM bandwidth on this GPU, the practical ceiling
// the gather-reduce kernel is compared with: grid-stride copy, read-only and
// write-only kernels over 2 GiB, best over workgroup counts and sizes; and the
// non-temporal copy with its source, or its destination, 4 bytes past a 16-byte
// boundary (the cost of unaligned 16-byte loads and stores).
// Built by bine_trees_fugaku_amd/build.py into bine_trees_fugaku_amd/hbm_probe.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
typedef unsigned int u4 __attribute__((ext_vector_type(4)));
template <int V, bool NT> __global__ void copy_k(const u4* __restrict__ a, u4* __restrict__ b, size_t n) {
  size_t stride = (size_t)gridDim.x * blockDim.x * V;
  for (size_t i = (size_t)blockIdx.x * blockDim.x * V + threadIdx.x; i < n; i += stride) {
    u4 x[V];
#pragma unroll
    for (int v = 0; v < V; ++v) { size_t j = i + (size_t)v * blockDim.x; if (j < n) x[v] = NT ? __builtin_nontemporal_load(a + j) : a[j]; }
#pragma unroll
    for (int v = 0; v < V; ++v) { size_t j = i + (size_t)v * blockDim.x; if (j < n) { if (NT) __builtin_nontemporal_store(x[v], b + j); else b[j] = x[v]; } }
  }
}
template <int V> __global__ void read_k(const u4* __restrict__ a, unsigned* out, size_t n) {
  size_t stride = (size_t)gridDim.x * blockDim.x * V;
  unsigned acc = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x * V + threadIdx.x; i < n; i += stride) {
#pragma unroll
    for (int v = 0; v < V; ++v) { size_t j = i + (size_t)v * blockDim.x; if (j < n) { u4 x = a[j]; acc ^= x.x ^ x.y ^ x.z ^ x.w; } }
  }
  if (acc == 0x12345678u) out[0] = acc;
}
template <int V> __global__ void write_k(u4* __restrict__ b, size_t n) {
  size_t stride = (size_t)gridDim.x * blockDim.x * V;
  for (size_t i = (size_t)blockIdx.x * blockDim.x * V + threadIdx.x; i < n; i += stride) {
#pragma unroll
    for (int v = 0; v < V; ++v) { size_t j = i + (size_t)v * blockDim.x; if (j < n) b[j] = u4{(unsigned)j, 1u, 2u, 3u}; }
  }
}
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("err %s line %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)
int main() {
  const size_t bytes = 2ull << 30, n = bytes / 16;
  u4 *a, *b; unsigned* o;
  // 64 bytes of slack: the shifted copies move n - 1 vectors starting 4 bytes in
  CK(hipMalloc(&a, bytes + 64)); CK(hipMalloc(&b, bytes + 64)); CK(hipMalloc(&o, 64));
  CK(hipMemset(a, 1, bytes + 64)); CK(hipMemset(b, 0, bytes + 64));
  const u4* a4 = reinterpret_cast<const u4*>(reinterpret_cast<const char*>(a) + 4);
  u4* b4 = reinterpret_cast<u4*>(reinterpret_cast<char*>(b) + 4);
  hipDeviceProp_t p; CK(hipGetDeviceProperties(&p, 0)); int cus = p.multiProcessorCount;
  printf("CUs %d\n", cus);
  hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
  int mults[] = {4, 8, 16, 32, 64};
  int thr[] = {256, 512, 1024};
  for (int kind = 0; kind < 7; ++kind) {
    float best = 1e9; int bm = 0, bt = 0;
    for (int m : mults) for (int t : thr) {
      int grid = cus * m;
      for (int rep = 0; rep < 6; ++rep) {
        if (rep == 1) CK(hipEventRecord(e0));
        switch (kind) {
          case 0: hipLaunchKernelGGL((copy_k<4, false>), dim3(grid), dim3(t), 0, 0, a, b, n); break;
          case 1: hipLaunchKernelGGL((copy_k<4, true>), dim3(grid), dim3(t), 0, 0, a, b, n); break;
          case 2: hipLaunchKernelGGL((read_k<4>), dim3(grid), dim3(t), 0, 0, a, o, n); break;
          case 3: hipLaunchKernelGGL((write_k<4>), dim3(grid), dim3(t), 0, 0, b, n); break;
          case 4: hipLaunchKernelGGL((copy_k<1, false>), dim3(grid), dim3(t), 0, 0, a, b, n); break;
          case 5: hipLaunchKernelGGL((copy_k<4, true>), dim3(grid), dim3(t), 0, 0, a4, b, n - 1); break;
          case 6: hipLaunchKernelGGL((copy_k<4, true>), dim3(grid), dim3(t), 0, 0, a, b4, n - 1); break;
        }
      }
      CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1)); CK(hipGetLastError());
      float ms; CK(hipEventElapsedTime(&ms, e0, e1)); ms /= 5;
      if (ms < best) { best = ms; bm = m; bt = t; }
    }
    double moved = (kind == 2 || kind == 3) ? (double)bytes : 2.0 * bytes;
    const char* names[] = {"copy v4", "copy v4 nt", "read v4", "write v4", "copy v1", "copy nt src+4", "copy nt dst+4"};
    printf("%-12s best %.3f ms  %.1f GB/s  (grid %d x CUs, %d threads)\n", names[kind], best, moved / best / 1e6, bm, bt);
  }
  return 0;
}
