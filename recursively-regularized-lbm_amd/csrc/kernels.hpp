// Shared declarations of the gfx950 kernels: argument blocks, launch helpers,
// storage traits and fast integer division.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lattice.hpp"

namespace rrlbm {

constexpr int kBlock = 256;  // threads per workgroup (4 wavefronts of 64)

// Division by a runtime-constant divisor: q = (umulhi(n, magic) + n) >> shift,
// valid for n < 2^31 (round-up multiplier method).
struct FastDiv {
  uint32_t d, magic, shift;
};
inline FastDiv make_fastdiv(uint32_t d) {
  uint32_t l = 0;
  while ((1ull << l) < d) ++l;
  FastDiv f;
  f.d = d;
  f.shift = l;
  f.magic = (uint32_t)(((1ull << 32) * ((1ull << l) - d)) / d + 1);
  return f;
}
RRLBM_HD RRLBM_INLINE uint32_t fastdiv(uint32_t n, const FastDiv& f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (__umulhi(n, f.magic) + n) >> f.shift;
#else
  return (uint32_t)((((uint64_t)n * f.magic) >> 32) + n) >> f.shift;
#endif
}

enum Storage : int { kF64 = 0, kF32 = 1, kF16 = 2 };
enum Mode : int { kStep = 0, kMacro = 1, kFinish = 2, kDiag = 3 };

template <int S>
struct StoreT;
template <>
struct StoreT<kF64> {
  using T = double;
  using C = double;
};
template <>
struct StoreT<kF32> {
  using T = float;
  using C = float;
};
template <>
struct StoreT<kF16> {
  using T = _Float16;
  using C = float;
};

constexpr int kDiagWidth = 9;  // sum(rho-1), sum u(3), sum rho u(3), sum KE, nonfinite count

// Arguments of the fused kernel (passed by value in the kernarg segment).
struct FusedArgs {
  const void* src;      // post-collision populations of the previous step, SoA [Q][N]
  void* dst;            // post-collision populations of this step
  const void* u_in;     // LES: velocity field read by the WALE stencil, AoS [N][3] (compute type)
  void* u_out;          // LES lagged / macro pass: velocity written by this launch
  double* ref_f;        // finish pass: reference-layout FP64 outputs
  double* ref_rho;
  double* ref_ux;
  double* ref_uy;
  double* ref_uz;
  double* partials;     // diagnostics pass: per-workgroup FP64 partial sums
  uint32_t lx, ly, lz, n;
  FastDiv div_lx, div_lxly;
  uint32_t walls;       // bit 0: x faces, bit 1: y faces, bit 2: z faces
  int32_t lid_face;     // 2*axis + side, or -1
  int32_t wall_rr;      // RR regularisation of wall-incoming populations
  int32_t write_u;      // STEP with LES: also write the velocity to u_out
  double fx, fy, fz;    // uniform body force
  double tau, cw;
  double omega;         // 1 - 1/tau (without LES)
  double lid_x, lid_y, lid_z;
  double f16_scale;     // FP16: stored value = (f_i - w_i) * f16_scale / w_i
  uint32_t f16_round;   // FP16 store: 0 stochastic, 1 nearest
  uint32_t sr_step;     // step counter of the stochastic-rounding dither
  uint32_t sr_salt;     // per-instance offset of the dither (batched ensembles)
  uint32_t nblocks;     // workgroups covering the domain
  uint32_t blocks_per_xcd;
  uint32_t xcd_remap;   // nonzero: launch is padded to 8*blocks_per_xcd and remapped per XCD
};

// Host launchers (kernels_fused.hip)
hipError_t launch_fused(int storage, int q, bool walls, bool les, int mode, const FusedArgs& a, uint32_t grid,
                        hipStream_t s);
// Two cells per thread, STEP mode, D3Q27 without LES; a.nblocks = ceil(n / 512).
hipError_t launch_fused_pair(int storage, bool walls, const FusedArgs& a, uint32_t grid, hipStream_t s);
hipError_t launch_batched(int storage, const FusedArgs* table, const uint32_t* blk2inst, const uint32_t* first_block,
                          uint32_t total_blocks, uint32_t parity, uint32_t step, hipStream_t s);
// Dither offset of batch instance k (the FP16 stochastic rounding of instance
// k then differs from that of the other instances at the same cell index).
inline uint32_t instance_salt(uint32_t k) { return k * 0x68E31DA5u; }
hipError_t launch_diag_finalize(const double* partials, uint32_t nparts, double* out, hipStream_t s);

}  // namespace rrlbm
