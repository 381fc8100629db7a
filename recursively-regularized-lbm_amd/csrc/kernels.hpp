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
// collision operator of a fused kernel: third-order RR, BGK baseline, RR with
// the fourth- to sixth-order terms, third-order RR with the FDBC wall rows
enum FusedOp : int { kOpRR = 0, kOpBGK = 1, kOpRR6 = 2, kOpRRFD = 3 };

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

// Layout of the fused path's population buffers.  The cells (x fastest) are
// split into groups of `grows` whole rows (gcells = grows*lx cells); a group
// holds its Q populations one after another, each a segment of `seg` bytes
// (gcells*sizeof + pad), so
//   byte offset of population i of cell c = i*seg + c*sizeof + (c / gcells)*gpad,
//   gpad = Q*seg - gcells*sizeof (the extra bytes per group).
// One group (gcells = N) is structure-of-arrays with `pad` bytes between the
// population arrays.  Groups of a few rows keep the populations one workgroup
// reads and writes close together, which the HBM serves faster than Q arrays
// a whole field apart.
struct PopLayout {
  uint64_t seg;        // bytes from population i to i+1 within a group
  uint64_t gpad;       // extra bytes per group
  uint64_t bytes;      // size of one population buffer
  uint32_t grows;      // rows per group
  uint32_t gcells;     // cells per group
  FastDiv div_gcells;  // cell -> group
  FastDiv div_grows;   // row (z*ly + y) -> group
};
// Byte offset of cell c within a population (without the term i*seg).  The
// buffer of a layout with several groups is below 4 GiB (checked on the host);
// with one group (structure of arrays) c / gcells = 0.
RRLBM_HD RRLBM_INLINE uint32_t pop_cell_off(const PopLayout& l, uint32_t c, uint32_t sz) {
  return c * sz + fastdiv(c, l.div_gcells) * (uint32_t)l.gpad;
}
// rows = ly*lz; grows in [1, rows]; pad in bytes (multiple of 16)
inline PopLayout make_layout(uint32_t lx, uint32_t rows, int q, uint32_t sz, uint32_t grows, uint32_t pad) {
  PopLayout l;
  l.grows = grows;
  l.gcells = grows * lx;
  l.seg = (uint64_t)l.gcells * sz + pad;
  l.gpad = (uint64_t)q * l.seg - (uint64_t)l.gcells * sz;
  const uint64_t ngroups = (rows + grows - 1) / grows;
  l.bytes = ngroups * q * l.seg;
  l.div_gcells = make_fastdiv(l.gcells);
  l.div_grows = make_fastdiv(grows);
  return l;
}

// Arguments of the fused kernel (passed by value in the kernarg segment).
struct FusedArgs {
  const void* src;      // post-collision populations of the previous step (layout `lay`)
  void* dst;            // post-collision populations of this step
  const void* u_in;     // LES: velocity field read by the WALE stencil, AoS [N][3] (compute type)
  void* u_out;          // LES lagged / macro pass: velocity written by this launch
  double* ref_f;        // finish pass: reference-layout FP64 outputs
  double* ref_rho;
  double* ref_ux;
  double* ref_uy;
  double* ref_uz;
  double* partials;     // diagnostics pass: per-workgroup FP64 partial sums
  void* u_wall;         // FDBC: velocity of rows 1, 2, LY-2, LY-3 after this step's streaming,
                        // [4][LZ][LX] x (ux, uy, uz) in the compute type
  int32_t gp;           // forcing_GP body force computed per cell (force_in_x = fx, fy = fz = 0):
  int32_t gp_on[2];     //   [0] the step the launch completes, [1] the step it collides;
  int32_t gp_istep[2];  //   off: the uniform force; on: forcing_GP(istep, phase)
  double gp_phase[2];
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
  float f16_unscale;    // FP16: 1 / f16_scale in FP32 (set on the host; no per-wavefront division)
  uint32_t f16_round;   // FP16 store: 0 stochastic, 1 nearest
  uint32_t sr_step;     // step counter of the stochastic-rounding dither
  const uint32_t* sr_base;  // graph replays: device counter added to sr_step (advanced by each replay), or null
  uint32_t sr_salt;     // per-instance offset of the dither (batched ensembles)
  uint32_t nblocks;     // workgroups covering the domain
  uint32_t blocks_per_xcd;
  uint32_t xcd_remap;   // 0 linear block order; 1 one contiguous slab per XCD; C > 1: runs of C blocks per XCD
  FastDiv div_remap;    // division by xcd_remap (C > 1)
  uint32_t tiles;       // length of the block sequence (nblocks rounded up for the remapped orders);
                        // the grid may be smaller (persistent workgroups stride over it)
  PopLayout lay;        // population buffer layout
};

// Two steps per launch (fused2_kernel): the domain is cut into tiles of
// kT2X x kT2Y cells in (x, y) and chunks of `zc` planes in z; a workgroup
// marches through its tile's chunk plane by plane, computing the first step
// on the tile and a one-cell halo around it into a ring of planes in LDS and
// the second step on the tile from that ring.
constexpr uint32_t kT2X = 64, kT2Y = 4;
constexpr uint32_t kT2Threads = 448;  // 7 wavefronts: 6 halo rows + the two halo columns
struct Fused2Args {
  uint32_t ntx, nty;  // tiles along x (lx / kT2X) and y (ly / kT2Y)
  uint32_t zc, nzc;   // planes per chunk, chunks (the last may be shorter)
  uint32_t items;     // ntx * nty * nzc
  uint32_t per_xcd;   // items per XCD: XCD j takes items [j * per_xcd, (j + 1) * per_xcd)
};

// Launchers of one storage type's fused kernels (fused_storage.hip, built once
// per storage type); the storage is a tag type, one overload per translation unit.
template <int S>
struct StorageTag {};
#define RRLBM_STORAGE_LAUNCHERS(S)                                                                                  \
  hipError_t launch_fused_s(StorageTag<S>, int q, bool walls, bool les, int mode, const FusedArgs& a,               \
                            uint32_t grid, hipStream_t s);                                                          \
  hipError_t launch_fused_op_s(StorageTag<S>, int op, bool walls, int mode, const FusedArgs& a, uint32_t grid,      \
                               hipStream_t s);                                                                      \
  hipError_t launch_fused_pair_s(StorageTag<S>, bool walls, const FusedArgs& a, uint32_t grid, hipStream_t s);      \
  hipError_t launch_batched_s(StorageTag<S>, const FusedArgs* table, const uint2* work, uint32_t ntiles,            \
                              uint32_t grid, uint32_t parity, uint32_t step, const uint32_t* step_base,             \
                              hipStream_t s);                                                                       \
  hipError_t launch_fdbc_rows_s(StorageTag<S>, const FusedArgs& a, hipStream_t s);                                 \
  hipError_t launch_fused2_s(StorageTag<S>, bool walls, int op, const FusedArgs& a, const Fused2Args& t,            \
                             hipStream_t s);
RRLBM_STORAGE_LAUNCHERS(kF64)
RRLBM_STORAGE_LAUNCHERS(kF32)
RRLBM_STORAGE_LAUNCHERS(kF16)
#undef RRLBM_STORAGE_LAUNCHERS

// Host launchers (kernels_fused.hip)
hipError_t launch_fused(int storage, int q, bool walls, bool les, int mode, const FusedArgs& a, uint32_t grid,
                        hipStream_t s);
// The alternative operators op = kOpBGK / kOpRR6 / kOpRRFD (D3Q27, no LES):
// step, finish and diagnostics modes.  kOpRRFD needs launch_fdbc_rows on the
// same source buffer first.
hipError_t launch_fused_op(int op, int storage, bool walls, int mode, const FusedArgs& a, uint32_t grid,
                           hipStream_t s);
// FDBC pre-pass: macroscopic velocity of rows 1, 2, LY-2, LY-3 into a.u_wall.
hipError_t launch_fdbc_rows(int storage, const FusedArgs& a, hipStream_t s);
// Two steps in one launch (D3Q27, op = kOpRR / kOpBGK / kOpRR6, without LES or
// per-cell force, FP32 / FP64 storage, lx a multiple of kT2X, ly of kT2Y):
// a.src -> a.dst.
hipError_t launch_fused2(int storage, bool walls, int op, const FusedArgs& a, const Fused2Args& t, hipStream_t s);
// Two cells per thread, STEP mode, D3Q27 without LES; a.nblocks = ceil(n / 512).
hipError_t launch_fused_pair(int storage, bool walls, const FusedArgs& a, uint32_t grid, hipStream_t s);
// work[t] = (instance, block); grid workgroups (<= ntiles) stride over the list;
// the dither step is step (+ *step_base when step_base is not null)
hipError_t launch_batched(int storage, const FusedArgs* table, const uint2* work, uint32_t ntiles, uint32_t grid,
                          uint32_t parity, uint32_t step, const uint32_t* step_base, hipStream_t s);
// Device step counter of captured FP16 steps: *p = v, and *p += inc (one lane each).
hipError_t launch_counter_set(uint32_t* p, uint32_t v, hipStream_t s);
hipError_t launch_counter_add(uint32_t* p, uint32_t inc, hipStream_t s);
// Dither offset of batch instance k (the FP16 stochastic rounding of instance
// k then differs from that of the other instances at the same cell index).
inline uint32_t instance_salt(uint32_t k) { return k * 0x68E31DA5u; }
hipError_t launch_diag_finalize(const double* partials, uint32_t nparts, double* out, hipStream_t s);

}  // namespace rrlbm
