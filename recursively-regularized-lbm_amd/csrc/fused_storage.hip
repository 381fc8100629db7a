// The fused kernels of one storage type: fused_impl.hpp instantiated for
// RRLBM_FUSED_STORAGE (0 FP64, 1 FP32, 2 FP16).  The build compiles this file
// once per storage type, so the three sets of kernels build in parallel; the
// storage-dispatching launchers are in kernels_fused.hip.
#include "fused_impl.hpp"

#ifndef RRLBM_FUSED_STORAGE
#error "compile with -DRRLBM_FUSED_STORAGE=0 (FP64), 1 (FP32) or 2 (FP16)"
#endif

namespace rrlbm {

namespace {
constexpr int kS = RRLBM_FUSED_STORAGE;
using Tag = StorageTag<kS>;
}  // namespace

hipError_t launch_fused_s(Tag, int q, bool walls, bool les, int mode, const FusedArgs& a, uint32_t grid,
                          hipStream_t s) {
  return launch_storage<kS>(q, walls, les, force_mode_of(a), mode, a, grid, s);
}

hipError_t launch_fused_op_s(Tag, int op, bool walls, int mode, const FusedArgs& a, uint32_t grid, hipStream_t s) {
  switch (op) {
    case kOpBGK: return launch_op_storage<kS, kOpBGK>(walls, mode, a, grid, s);
    case kOpRR6: return launch_op_storage<kS, kOpRR6>(walls, mode, a, grid, s);
    case kOpRRFD: return launch_op_storage<kS, kOpRRFD>(walls, mode, a, grid, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_fused_pair_s(Tag, bool walls, const FusedArgs& a, uint32_t grid, hipStream_t s) {
  const int fm = force_mode_of(a);
  return walls ? launch_pair_force<kS, true>(fm, a, grid, s) : launch_pair_force<kS, false>(fm, a, grid, s);
}

hipError_t launch_batched_s(Tag, const FusedArgs* table, const uint2* work, uint32_t ntiles, uint32_t grid,
                            uint32_t parity, uint32_t step, const uint32_t* step_base, hipStream_t s) {
  return launch_batched_s<kS>(table, work, ntiles, grid, parity, step, step_base, s);
}

hipError_t launch_fused2_s(Tag, bool walls, int op, const FusedArgs& a, const Fused2Args& t, hipStream_t s) {
  return launch_fused2_storage<kS>(walls, op, force_mode_of(a), a, t, s);
}

hipError_t launch_fdbc_rows_s(Tag, const FusedArgs& a, hipStream_t s) {
  const dim3 grid((4 * a.lx * a.lz + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(fdbc_rows_kernel<kS>, grid, dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

}  // namespace rrlbm
