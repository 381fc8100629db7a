// Host launchers of the fused RR-LBM step (the kernels are in fused_impl.hpp,
// instantiated per storage type by fused_storage.hip): dispatch on the storage
// type of a launch, the deterministic final sum of the diagnostics and the
// step counter of captured FP16 steps.
#include "kernels.hpp"

namespace rrlbm {

hipError_t launch_fused(int storage, int q, bool walls, bool les, int mode, const FusedArgs& a, uint32_t grid,
                        hipStream_t s) {
  switch (storage) {
    case kF64: return launch_fused_s(StorageTag<kF64>{}, q, walls, les, mode, a, grid, s);
    case kF32: return launch_fused_s(StorageTag<kF32>{}, q, walls, les, mode, a, grid, s);
    case kF16: return launch_fused_s(StorageTag<kF16>{}, q, walls, les, mode, a, grid, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_fused_op(int op, int storage, bool walls, int mode, const FusedArgs& a, uint32_t grid,
                           hipStream_t s) {
  switch (storage) {
    case kF64: return launch_fused_op_s(StorageTag<kF64>{}, op, walls, mode, a, grid, s);
    case kF32: return launch_fused_op_s(StorageTag<kF32>{}, op, walls, mode, a, grid, s);
    case kF16: return launch_fused_op_s(StorageTag<kF16>{}, op, walls, mode, a, grid, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_fused_pair(int storage, bool walls, const FusedArgs& a, uint32_t grid, hipStream_t s) {
  switch (storage) {
    case kF64: return launch_fused_pair_s(StorageTag<kF64>{}, walls, a, grid, s);
    case kF32: return launch_fused_pair_s(StorageTag<kF32>{}, walls, a, grid, s);
    case kF16: return launch_fused_pair_s(StorageTag<kF16>{}, walls, a, grid, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_fused2(int storage, bool walls, int op, const FusedArgs& a, const Fused2Args& t, hipStream_t s) {
  switch (storage) {
    case kF64: return launch_fused2_s(StorageTag<kF64>{}, walls, op, a, t, s);
    case kF32: return launch_fused2_s(StorageTag<kF32>{}, walls, op, a, t, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_batched(int storage, const FusedArgs* table, const uint2* work, uint32_t ntiles, uint32_t grid,
                          uint32_t parity, uint32_t step, const uint32_t* step_base, hipStream_t s) {
  switch (storage) {
    case kF64: return launch_batched_s(StorageTag<kF64>{}, table, work, ntiles, grid, parity, step, step_base, s);
    case kF32: return launch_batched_s(StorageTag<kF32>{}, table, work, ntiles, grid, parity, step, step_base, s);
    case kF16: return launch_batched_s(StorageTag<kF16>{}, table, work, ntiles, grid, parity, step, step_base, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_fdbc_rows(int storage, const FusedArgs& a, hipStream_t s) {
  switch (storage) {
    case kF64: return launch_fdbc_rows_s(StorageTag<kF64>{}, a, s);
    case kF32: return launch_fdbc_rows_s(StorageTag<kF32>{}, a, s);
    case kF16: return launch_fdbc_rows_s(StorageTag<kF16>{}, a, s);
    default: return hipErrorInvalidValue;
  }
}

// The step counter of captured FP16 steps: set before a series of graph
// replays (outside the graph) and advanced at the end of each replay (a
// kernel node: the graph holds no memset node).  One lane writes.
__global__ void counter_set_kernel(uint32_t* p, uint32_t v) {
  if (threadIdx.x == 0) *p = v;
}

__global__ void counter_add_kernel(uint32_t* p, uint32_t inc) {
  if (threadIdx.x == 0) *p = *p + inc;
}

hipError_t launch_counter_set(uint32_t* p, uint32_t v, hipStream_t s) {
  hipLaunchKernelGGL(counter_set_kernel, dim3(1), dim3(64), 0, s, p, v);
  return hipGetLastError();
}

hipError_t launch_counter_add(uint32_t* p, uint32_t inc, hipStream_t s) {
  hipLaunchKernelGGL(counter_add_kernel, dim3(1), dim3(64), 0, s, p, inc);
  return hipGetLastError();
}

// Deterministic final sum of the per-workgroup diagnostics: workgroup k sums
// quantity k.  Thread t adds partials t, t + 1024, ... in four interleaved
// accumulators (independent loads in flight), combined in a fixed order, then
// a fixed-shape tree over the workgroup: the result does not depend on timing.
constexpr int kFinBlock = 1024;
__global__ __launch_bounds__(kFinBlock) void diag_finalize_kernel(const double* partials, uint32_t nparts, double* out) {
  __shared__ double red[kFinBlock];
  const uint32_t k = blockIdx.x;
  double t4[4] = {0.0, 0.0, 0.0, 0.0};
  uint32_t p = threadIdx.x;
  for (; p + 3 * kFinBlock < nparts; p += 4 * kFinBlock) {
#pragma unroll
    for (int j = 0; j < 4; ++j) t4[j] += partials[(size_t)(p + j * kFinBlock) * kDiagWidth + k];
  }
  for (int j = 0; p < nparts; p += kFinBlock, ++j) t4[j] += partials[(size_t)p * kDiagWidth + k];
  red[threadIdx.x] = (t4[0] + t4[1]) + (t4[2] + t4[3]);
  __syncthreads();
  for (int s = kFinBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = red[0];
}

hipError_t launch_diag_finalize(const double* partials, uint32_t nparts, double* out, hipStream_t s) {
  hipLaunchKernelGGL(diag_finalize_kernel, dim3(kDiagWidth), dim3(kFinBlock), 0, s, partials, nparts, out);
  return hipGetLastError();
}

}  // namespace rrlbm
