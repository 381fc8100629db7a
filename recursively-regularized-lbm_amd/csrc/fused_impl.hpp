// The fused RR-LBM step for gfx950: device code and per-storage launch templates
// (instantiated by fused_storage.hip, once per storage type).
//
// One launch = one time step of the channel driver's sequence
// (RR27V33_code_DNS_LES/main.cu:50-84), reordered so that everything after the
// streaming is cell-local:
//
//   pull-stream (streaming.cu:3-24 as a gather) -> half-way bounce-back
//   (streaming.cu:26-55) -> macrovar (macrovar.cu:4-31) -> RRBC at wall cells
//   (BC.cu:4-126) -> macrovar_BN (macrovar.cu:33-62) -> [WALE, LES.cu:4-98]
//   -> RR collision with forcing (collision.cu:4-119) -> store.
//
// The stored state between launches is the post-collision populations, kept
// shifted (f_i - w_i) in FP64, FP32 or per-direction-scaled FP16, SoA with x
// fastest.  Each of the 27 (19) populations is read once (gathered from the
// upwind neighbour, periodic wrap in every axis; wall cells then overwrite
// their incoming populations with the bounce-back value) and written once,
// coalesced.  Bytes per lattice update: 2*Q*sizeof(storage) (+ velocity field
// traffic with LES).  One thread per cell, 256-thread workgroups, workgroup
// order remapped so that each XCD works on a contiguous slab (neighbouring
// rows share its L2).
#pragma once
#include "kernels.hpp"
#include "rr_cell.hpp"

namespace rrlbm {

// Population i at byte offset `off` from `base` (the population buffer, or a
// row of it); populations are `seg` bytes apart.  The 64-bit scalar part and
// the 32-bit per-lane offset give one saddr global load.
template <int S>
RRLBM_INLINE __device__ typename StoreT<S>::C load_pop_b(const void* base, uint64_t seg, int i, uint32_t off,
                                                         typename StoreT<S>::C unscale) {
  using T = typename StoreT<S>::T;
  using C = typename StoreT<S>::C;
  const char* p = reinterpret_cast<const char*>(base) + (uint64_t)i * seg;
  const T v = *reinterpret_cast<const T*>(p + off);
  if constexpr (S == kF16)
    return C(v) * unscale;
  else
    return C(v);
}

// Factor from a stored FP16 value to h_i / w_i (1 / f16_scale); 1 otherwise.
template <int S>
RRLBM_INLINE __device__ typename StoreT<S>::C inv_scale_of(const FusedArgs& a) {
  using C = typename StoreT<S>::C;
  if constexpr (S == kF16)
    return C(a.f16_unscale);
  else
    return C(1);
}

// Raw moments of a D3Q27 cell from its FP16-stored values g_i = h_i * S / w_i
// (moments_q27_wscaled; unscale216 = 1 / (216 S)).  The FP16 values enter the
// first stage as operands of gfx950's mixed-precision FMA (v_fma_mix_f32
// converts them in the same instruction), so a population costs one
// instruction here instead of a conversion, a scaling multiply and its share
// of the line sums.  The 1.0 is made opaque (a scalar register the compiler
// cannot see into): fma(g+, 1, g-) would otherwise fold into an add of two
// separately converted values.
RRLBM_INLINE __device__ Moments<float> moments_q27_f16(const _Float16 (&g)[27], float unscale216) {
  float one = 1.0f;
  asm volatile("" : "+s"(one));
  return moments_q27_wscaled<float, _Float16>(g, one, unscale216);
}

template <int S>
RRLBM_INLINE __device__ uint32_t cell_off(const PopLayout& l, uint32_t c) {
  return pop_cell_off(l, c, (uint32_t)sizeof(typename StoreT<S>::T));
}

// T in the global address space (device memory): an address formed from an
// integer keeps its global load / store instructions.
template <class T>
using GlobalT = __attribute__((address_space(1))) T;

// Post-collision populations are written once and read by the next launch:
// non-temporal stores keep them from displacing the gather's cache lines.
template <class T>
RRLBM_INLINE __device__ void store_raw(uint64_t base, uint32_t off, T v) {
  __builtin_nontemporal_store(v, reinterpret_cast<GlobalT<T>*>(base + off));
}

// Stochastic rounding of an FP32 value to FP16 with gfx950's v_cvt_sr_f16_f32:
// the 13 dither bits (top bits of r) are added to the 13 low mantissa bits that
// FP16 drops and the sum is truncated, i.e. (bits(v) + (r >> 19)) & ~0x1FFF
// (one instruction; it matched that integer form on 2^20 random (v, r) pairs
// on MI355X).  The result is the upper FP16 neighbour with probability
// (dropped part) / (FP16 step), so the stored value equals v in expectation.
// A carry into the exponent gives the next binade's first value, as it should.
// The instruction writes one half of its destination and keeps the other:
// `z` carries the register from one conversion to the next, so that no
// conversion needs a cleared register of its own (only the low half is used).
typedef _Float16 rr_half2 __attribute__((ext_vector_type(2)));
RRLBM_INLINE __device__ _Float16 f16_stochastic(rr_half2& z, float v, uint32_t r) {
  z = __builtin_amdgcn_cvt_sr_f16_f32(z, v, r, false);
  return z[0];
}

// 32-bit integer mixer ("lowbias32", from C. Wellons' hash-prospector search).
RRLBM_INLINE __device__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// Dither of one cell at one step, hashed from both, so that a cell's dither
// at the next step is independent of this one.  (A dither that advances by a
// fixed increment per step -- a Weyl sequence -- correlates with the value it
// rounds: the next value depends on this step's rounding, and the mean flow
// drifts.)  Population i adds i times an odd constant.
RRLBM_INLINE __device__ uint32_t dither_base(uint32_t cell, uint32_t step, uint32_t salt) {
  return mix32(mix32(cell + salt) + step * 0x9E3779B9u);
}

// Store the Q post-collision values of one cell (FP16: scaled values) at
// base + i*seg + off; base is uniform across the wavefront.
template <int S, int Q>
RRLBM_INLINE __device__ void store_cell(const FusedArgs& a, void* base, uint32_t off, uint32_t cell,
                                        const typename StoreT<S>::C (&out)[Q]) {
  using T = typename StoreT<S>::T;
  // Each branch takes a fresh 32-bit copy of the lane offset: the offset is
  // also used by the wall path, and a 64-bit extension of it shared across
  // blocks turns every store into a per-lane 64-bit address add instead of a
  // scalar-base + 32-bit-offset store.  The population base advances by seg
  // in a scalar register pair held opaque, as in the gather.
  uint64_t pb = reinterpret_cast<uint64_t>(base);
  auto put = [&](T v) {
    store_raw<T>(pb, off, v);
    pb += a.lay.seg;
    asm volatile("" : "+s"(pb));
  };
  if constexpr (S == kF16) {
    if (a.f16_round == 0) {
      asm volatile("" : "+v"(off));
      // captured steps take the step count from the device counter that each
      // graph replay advances, so replays do not repeat the dither
      const uint32_t step = a.sr_base ? *a.sr_base + a.sr_step : a.sr_step;
      const uint32_t r0 = dither_base(cell, step, a.sr_salt);
      rr_half2 z = {_Float16(0), _Float16(0)};
      static_for<Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        put(f16_stochastic(z, out[i], r0 + uint32_t(i) * 0xC2B2AE3Du));
      });
      return;
    }
  }
  asm volatile("" : "+v"(off));
  static_for<Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    put(T(out[i]));
  });
}

// Does population i (velocity x,y,z) enter the domain through face f?
RRLBM_HD constexpr bool enters(int face, int x, int y, int z) {
  switch (face) {
    case 0: return x > 0;
    case 1: return x < 0;
    case 2: return y > 0;
    case 3: return y < 0;
    case 4: return z > 0;
    default: return z < 0;
  }
}

template <class L, typename C, int face, int FM, bool O6>
RRLBM_INLINE __device__ void regularise_face(C (&h)[L::Q], C Fx, C Fy, C Fz) {
  Moments<C> m;
  if constexpr (L::Q == 27) m = moments_q27<C>(h);
  else m = moments_generic<L, C>(h);
  const Macro<C> q = macro_of<C, FM>(m, Fx, Fy, Fz);
  C out[L::Q];
  if constexpr (L::Q == 27) {
    const NeqTensor<C> N = neq_q27<C, FM>(m, q, Fx, Fy, Fz);
    if constexpr (O6)
      rr_post_q27_o6<C, true, FM>(q, N, Fx, Fy, Fz, C(1), C(-1), C(1), out);
    else
      rr_post_q27<C, true, FM>(q, N, Fx, Fy, Fz, C(1), C(-1), C(1), out);
  } else if constexpr (FM == kForceNone) {
    const NeqTensor<C> N = neq_q27<C, kForceNone>(m, q, C(0), C(0), C(0));
    rr_post_q19<C, true>(q, N, C(1), C(1), out);
  } else {
    rr_post_generic<L, C>(h, q, Fx, Fy, Fz, C(1), C(-1), out);
  }
  static_for<L::Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    if constexpr (enters(face, L::cx(i), L::cy(i), L::cz(i))) h[i] = out[i];
  });
}

// Wall treatment of one cell (h: its pulled populations, replaced in place).
// own(o): the cell's own post-collision population o of the previous step
// (unscaled, shifted), the source of the half-way bounce-back of population
// opp(o) (+ moving-lid term 6 w_i c_i.u_lid, Cavity/streaming.f90:142-152).
// O6: regularise with the fourth- to sixth-order terms.
template <class L, typename C, int FM, bool O6, class Own>
RRLBM_INLINE __device__ void wall_treatment_own(const FusedArgs& a, C (&h)[L::Q], uint32_t x, uint32_t y,
                                                uint32_t z, Own&& own, C Fx, C Fy, C Fz) {
  const bool on[6] = {(a.walls & 1) && x == 0,        (a.walls & 1) && x == a.lx - 1,
                      (a.walls & 2) && y == 0,        (a.walls & 2) && y == a.ly - 1,
                      (a.walls & 4) && z == 0,        (a.walls & 4) && z == a.lz - 1};
  // Face by face in the order x-,x+,y-,y+,z-,z+ (Cavity/streaming.f90:89-292):
  // the face's incoming populations take their half-way bounce-back value,
  // then (wall_rr) the RR regularisation replaces them, with moments of the
  // cell's current populations.  Populations entering through a later face
  // still hold the periodically pulled value while an earlier face is
  // treated, as in the reference, whose bounce-back of that face runs later.
  static_for<6>([&](auto fc) {
    constexpr int face = decltype(fc)::value;
    if (!on[face]) return;
    static_for<L::Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      if constexpr (enters(face, L::cx(i), L::cy(i), L::cz(i))) {
        C v = own(L::opp(i));
        if (face == a.lid_face)
          v += C(6.0 * L::w(i)) * (C(L::cx(i)) * C(a.lid_x) + C(L::cy(i)) * C(a.lid_y) + C(L::cz(i)) * C(a.lid_z));
        h[i] = v;
      }
    });
    if (a.wall_rr) regularise_face<L, C, face, FM, O6>(h, Fx, Fy, Fz);
  });
}

// The same with the previous step's populations in the source buffer:
// `base` + population offset + off (this cell's byte offset, without i*seg).
template <int S, class L, int FM, bool O6 = false>
RRLBM_INLINE __device__ void wall_treatment(const FusedArgs& a, typename StoreT<S>::C (&h)[L::Q],
                                            uint32_t x, uint32_t y, uint32_t z, const void* base, uint32_t off,
                                            typename StoreT<S>::C Fx, typename StoreT<S>::C Fy,
                                            typename StoreT<S>::C Fz) {
  using C = typename StoreT<S>::C;
  const C inv_scale = inv_scale_of<S>(a);
  wall_treatment_own<L, C, FM, O6>(
      a, h, x, y, z, [&](int o) { return load_pop_b<S>(base, a.lay.seg, o, off, C(L::w(o)) * inv_scale); }, Fx, Fy,
      Fz);
}

// LES velocity of one cell: the three components are stored together (AoS,
// [N][3]) so that each stencil point is one 12-byte (FP32) load.
template <typename C>
struct alignas(sizeof(C)) Vel3 {
  C x, y, z;
};

template <typename C>
RRLBM_INLINE __device__ Vel3<C> ld_v(const void* u, uint32_t cell) {
  return reinterpret_cast<const Vel3<C>*>(u)[cell];
}

template <typename C>
RRLBM_INLINE __device__ void st_v(void* u, uint32_t cell, C x, C y, C z) {
  reinterpret_cast<Vel3<C>*>(u)[cell] = Vel3<C>{x, y, z};
}

// Twice the velocity derivative along one axis (LES.cu:19-33): the undivided
// central difference v(+1) - v(-1), or at a wall cell twice the one-sided form
// with the half-way wall velocity taken as zero, (3 v0 + v(+1)) / 3 at the low
// wall and -(3 v0 + v(-1)) / 3 at the high one.
template <typename C>
RRLBM_INLINE __device__ Vel3<C> d2_central(const Vel3<C>& vm, const Vel3<C>& vp) {
  return Vel3<C>{vp.x - vm.x, vp.y - vm.y, vp.z - vm.z};
}
template <typename C>
RRLBM_INLINE __device__ Vel3<C> d2_low_wall(const Vel3<C>& v0, const Vel3<C>& vp) {
  const C k = C(2.0 / 3.0);
  return Vel3<C>{(C(3) * v0.x + vp.x) * k, (C(3) * v0.y + vp.y) * k, (C(3) * v0.z + vp.z) * k};
}
template <typename C>
RRLBM_INLINE __device__ Vel3<C> d2_high_wall(const Vel3<C>& v0, const Vel3<C>& vm) {
  const C k = C(2.0 / 3.0);
  return Vel3<C>{(C(-3) * v0.x - vm.x) * k, (C(-3) * v0.y - vm.y) * k, (C(-3) * v0.z - vm.z) * k};
}

// Byte offset (without the population term) of the first cell of row
// (yy, zz): the whole rows before it plus the pads of the groups before it.
RRLBM_INLINE __device__ uint64_t row_offset(const FusedArgs& a, uint32_t row_b, uint32_t yy, uint32_t zz) {
  const uint32_t r = zz * a.ly + yy;
  return (uint64_t)fastdiv(r, a.lay.div_grows) * a.lay.gpad + (uint64_t)r * row_b;
}

// The same in 32 bits: within one population segment the offset is below
// 4 GiB (a grouped buffer stays below 4 GiB; with one group r / grows = 0 and
// one population array is below 4 GiB).
RRLBM_INLINE __device__ uint32_t row_offset32(const FusedArgs& a, uint32_t row_b, uint32_t yy, uint32_t zz) {
  const uint32_t r = zz * a.ly + yy;
  return fastdiv(r, a.lay.div_grows) * (uint32_t)a.lay.gpad + r * row_b;
}

// Pull gather of the Q upwind populations of the cell (x, y, z), per-lane
// addressing (any row length): nine source-row offsets, below 4 GiB.
template <int S, class L>
RRLBM_INLINE __device__ void gather_lane(const FusedArgs& a, uint32_t x, uint32_t y, uint32_t z,
                                         typename StoreT<S>::C (&h)[L::Q]) {
  using C = typename StoreT<S>::C;
  constexpr uint32_t sz = sizeof(typename StoreT<S>::T);
  const uint32_t row_b = a.lx * sz;
  const C inv_scale = inv_scale_of<S>(a);
  const uint32_t xb[3] = {(x + 1 == a.lx ? 0u : x + 1) * sz, x * sz, (x == 0 ? a.lx - 1 : x - 1) * sz};
  const uint32_t ys[3] = {y + 1 == a.ly ? 0u : y + 1, y, y == 0 ? a.ly - 1 : y - 1};
  const uint32_t zs[3] = {z + 1 == a.lz ? 0u : z + 1, z, z == 0 ? a.lz - 1 : z - 1};
  uint32_t ro[3][3];
  static_for<3>([&](auto jy) {
    static_for<3>([&](auto jz) {
      constexpr int iy = decltype(jy)::value, iz = decltype(jz)::value;
      ro[iy][iz] = (uint32_t)row_offset(a, row_b, ys[iy], zs[iz]);
    });
  });
  static_for<L::Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    constexpr int cx = L::cx(i), cy = L::cy(i), cz = L::cz(i);
    h[i] = load_pop_b<S>(a.src, a.lay.seg, i, ro[cy + 1][cz + 1] + xb[cx + 1], C(L::w(i)) * inv_scale);
  });
}

// forcing_GP (macrovar.cu:76-114) at cell (x, y, z) for part k of the launch
// (0: the step being completed, 1: the step being collided), in the compute
// type, with the drop-in k_forcing_gp's expressions: a time-periodic force in
// two bands near the walls; the upper band's x phase is the integer ix / LX
// (zero) plus the random phase, as the reference writes it.  Off: the
// uniform force (force_in_x, 0, 0).
template <typename C>
RRLBM_INLINE __device__ void gp_force(const FusedArgs& a, int k, uint32_t x, uint32_t y, uint32_t z, C& fx, C& fy,
                                      C& fz) {
  const C f0 = C(a.fx);
  fx = f0; fy = C(0); fz = C(0);
  if (!a.gp_on[k]) return;
  const C pi2 = C(6.283185307179586476925);
  const C LYq = C(a.ly / 4);
  const int h0 = 2;
  const C beta9 = C(3), gamma9 = C(2);
  const C LX = C(a.lx), LY = C(a.ly), LZ = C(a.lz);
  const C Amp0 = (C(100) * beta9 / LY) * sin(pi2 * C(a.gp_istep[k]) / C(2000));
  const int iy = (int)y;
  const C ix = C(x), izc = C(z);
  if (iy > h0 && C(iy) < LYq + C(h0)) {
    const C sy = sin(pi2 * C(iy - h0) / LYq), cy = C(1) - cos(pi2 * C(iy - h0) / LYq);
    const C cz = cos(gamma9 * pi2 * izc / LZ);
    fx = f0 * (C(1) - Amp0 * LX / beta9 * sy * sin(beta9 * pi2 * ix / LX) * cz);
    fy = f0 * Amp0 * C(0.5) * LYq * cy * cos(beta9 * pi2 * ix / LX) * cz;
    fz = f0 * Amp0 * C(0.5) * LZ / gamma9 * sy * cos(beta9 * pi2 * ix / LX) * sin(gamma9 * pi2 * izc / LZ);
  } else if (C(iy) > C(0.75) * LY - C(h0) && iy < (int)a.ly - h0) {
    const C yy = -LY + C(iy) + C(h0) + LYq;
    const C xph = C(0) + C(a.gp_phase[k]);
    const C sy = sin(pi2 * yy / LYq), cy = C(1) - cos(pi2 * yy / LYq);
    const C cz = cos(gamma9 * pi2 * izc / LZ);
    fx = f0 * (C(1) + Amp0 * LX / beta9 * sy * sin(beta9 * pi2 * xph) * cz);
    fy = -f0 * Amp0 * C(0.5) * LYq * cy * cos(beta9 * pi2 * xph) * cz;
    fz = -f0 * Amp0 * C(0.5) * LZ / gamma9 * sy * cos(beta9 * pi2 * xph) * sin(gamma9 * pi2 * izc / LZ);
  }
}

// The WALE stencil of one cell: the velocity field a.u_in (previous step) at
// the six face neighbours (periodic wrap).
template <typename C>
struct LesStencil {
  Vel3<C> xm, xp, ym, yp, zm, zp;
};

// Load the stencil of cell (x, y, z); xs, ys, zs: the periodic neighbour
// coordinates {+1, 0, -1}.  With ROWS the five stencil rows are
// wavefront-uniform (scalar bases) and the lane adds its x offset.
template <typename C>
RRLBM_INLINE __device__ LesStencil<C> les_load(const FusedArgs& a, uint32_t x, uint32_t y, uint32_t z,
                                               const uint32_t (&xs)[3], const uint32_t (&ys)[3],
                                               const uint32_t (&zs)[3]) {
  constexpr uint32_t vb = sizeof(Vel3<C>);
  auto vrow = [&](uint32_t yy, uint32_t zz) {
    return reinterpret_cast<const char*>(a.u_in) + (uint64_t)((zz * a.ly + yy) * a.lx) * vb;
  };
  auto vat = [](const char* row, uint32_t off) { return *reinterpret_cast<const Vel3<C>*>(row + off); };
  const char* r00 = vrow(y, z);
  const uint32_t o0 = x * vb;
  LesStencil<C> st;
  st.xm = vat(r00, xs[2] * vb);
  st.xp = vat(r00, xs[0] * vb);
  st.ym = vat(vrow(ys[2], z), o0);
  st.yp = vat(vrow(ys[0], z), o0);
  st.zm = vat(vrow(y, zs[2]), o0);
  st.zp = vat(vrow(y, zs[0]), o0);
  return st;
}

// LES relaxation factor 1 - 1/(tau + tau_t) of cell (x, y, z) from its WALE
// stencil (LES.cu:4-98); wall cells load their own velocity here for the
// one-sided forms.
template <typename C>
RRLBM_INLINE __device__ C les_omega_of(const FusedArgs& a, const LesStencil<C>& st, uint32_t x, uint32_t y,
                                       uint32_t z) {
  const bool wy = (a.walls & 2) != 0, wx = (a.walls & 1) != 0, wz = (a.walls & 4) != 0;
  const bool xlo = wx && x == 0, xhi = wx && x == a.lx - 1;
  const bool ylo = wy && y == 0, yhi = wy && y == a.ly - 1;
  const bool zlo = wz && z == 0, zhi = wz && z == a.lz - 1;
  // D = 2 grad u; rows: Dx = d/dx of (u, v, w), ...
  Vel3<C> Dx = d2_central(st.xm, st.xp), Dy = d2_central(st.ym, st.yp), Dz = d2_central(st.zm, st.zp);
  // wall cells (with ROWS the y and z tests are uniform per wavefront):
  // the one-sided forms, which need the cell's own velocity
  if (xlo || xhi || ylo || yhi || zlo || zhi) {
    const Vel3<C> v0 = reinterpret_cast<const Vel3<C>*>(a.u_in)[(z * a.ly + y) * a.lx + x];
    if (xlo) Dx = d2_low_wall(v0, st.xp);
    else if (xhi) Dx = d2_high_wall(v0, st.xm);
    if (ylo) Dy = d2_low_wall(v0, st.yp);
    else if (yhi) Dy = d2_high_wall(v0, st.ym);
    if (zlo) Dz = d2_low_wall(v0, st.zp);
    else if (zhi) Dz = d2_high_wall(v0, st.zm);
  }
  const C tau_eff = C(a.tau) + wale_taut_d2<C>(Dx.x, Dy.x, Dz.x, Dx.y, Dy.y, Dz.y, Dx.z, Dy.z, Dz.z, C(a.cw));
  return C(1) - recip(tau_eff);
}

// FDBC at a y-wall cell (BC.cu:129-309, the drop-in k_fdbc_* in one place):
// wall velocities extrapolated from rows 1, 2 (or LY-2, LY-3) of a.u_wall for
// this cell and its x and z neighbours, strain rate by finite differences,
// then every population replaced by the reconstruction at the density that
// keeps the known populations' sum (c_y pointing into the wall).  h: the
// cell's gathered shifted populations, replaced in place.  With the
// reconstruction per unit density g = w + r (r from rr_post_q27 with rho = 1),
// rho_l - 1 = (sum_k h - sum_k r) / (W + sum_k r), W = sum_k w = 1/6, and the
// new h = r + (rho_l - 1) g: only small quantities enter.
template <int S, int FM>
RRLBM_INLINE __device__ void fdbc_treatment(const FusedArgs& a, typename StoreT<S>::C (&h)[27], uint32_t x,
                                            uint32_t y, uint32_t z) {
  using C = typename StoreT<S>::C;
  using L = D3Q27;
  const bool top = y != 0;
  const uint32_t k1 = top ? 2u : 0u, k2 = top ? 3u : 1u;
  const Vel3<C>* U = reinterpret_cast<const Vel3<C>*>(a.u_wall);
  auto row = [&](uint32_t k, uint32_t xx, uint32_t zz) { return U[(k * a.lz + zz) * a.lx + xx]; };
  auto wall = [&](uint32_t xx, uint32_t zz) {
    const Vel3<C> u1 = row(k1, xx, zz), u2 = row(k2, xx, zz);
    const C c1 = C(2.0 / 3.0), c2 = C(1.0 / 5.0);
    return Vel3<C>{c1 * u1.x - c2 * u2.x, c1 * u1.y - c2 * u2.y, c1 * u1.z - c2 * u2.z};
  };
  const uint32_t xp = x + 1 == a.lx ? 0u : x + 1, xm = x == 0 ? a.lx - 1 : x - 1;
  const uint32_t zp = z + 1 == a.lz ? 0u : z + 1, zm = z == 0 ? a.lz - 1 : z - 1;
  const Vel3<C> u0 = wall(x, z), uxp = wall(xp, z), uxm = wall(xm, z), uzp = wall(x, zp), uzm = wall(x, zm);
  const Vel3<C> u1 = row(k1, x, z), u2 = row(k2, x, z);
  const C hs = top ? C(-0.5) : C(0.5);
  const C dudy = hs * (C(-3) * u0.x + C(4) * u1.x - u2.x);
  const C dvdy = hs * (C(-3) * u0.y + C(4) * u1.y - u2.y);
  const C dwdy = hs * (C(-3) * u0.z + C(4) * u1.z - u2.z);
  const C dudx = C(0.5) * (uxp.x - uxm.x), dvdx = C(0.5) * (uxp.y - uxm.y), dwdx = C(0.5) * (uxp.z - uxm.z);
  const C dudz = C(0.5) * (uzp.x - uzm.x), dvdz = C(0.5) * (uzp.y - uzm.y), dwdz = C(0.5) * (uzp.z - uzm.z);
  // non-equilibrium part: Hermite coefficients -tau S without the collision's
  // 1/2, i.e. N = -(2/3) tau S in the moment-space form
  const C k = C(-2.0 / 3.0 * a.tau);
  NeqTensor<C> N;
  N.xx = k * dudx; N.yy = k * dvdy; N.zz = k * dwdz;
  N.xy = k * C(0.5) * (dudy + dvdx);
  N.xz = k * C(0.5) * (dudz + dwdx);
  N.yz = k * C(0.5) * (dvdz + dwdy);
  Macro<C> q;
  q.rho = C(1); q.dr = C(0);
  q.ux = u0.x; q.uy = u0.y; q.uz = u0.z;
  C r[27];
  rr_post_q27<C, true, FM>(q, N, C(a.fx), C(a.fy), C(a.fz), C(1), C(-1), C(1), r);
  C dps = C(0), din = C(0);
  static_for<27>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    if constexpr (L::cy(i) != 0) {
      if ((L::cy(i) > 0) == top) {
        dps += h[i];
        din += r[i];
      }
    }
  });
  const C d = (dps - din) / (C(1.0 / 6.0) + din);
  static_for<27>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    h[i] = r[i] + d * (r[i] + C(L::w(i)));
  });
}

// FDBC pre-pass: macrovar (macrovar.cu:4-31) of rows 1, 2, LY-2, LY-3 from
// this step's pulled populations, stored for the wall rows' differences.
// Thread t -> x = t % LX, slot k = (t / LX) % 4 (rows 1, 2, LY-2, LY-3), z.
template <int S>
__global__ __launch_bounds__(kBlock) void fdbc_rows_kernel(FusedArgs a) {
  using C = typename StoreT<S>::C;
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= 4 * a.lx * a.lz) return;
  const uint32_t x = t % a.lx, r = t / a.lx, k = r & 3u, z = r >> 2;
  const uint32_t y = k == 0 ? 1u : (k == 1 ? 2u : (k == 2 ? a.ly - 2 : a.ly - 3));
  C h[27];
  gather_lane<S, D3Q27>(a, x, y, z, h);
  const Moments<C> m = moments_q27<C>(h);
  const Macro<C> q = macro_of<C, kForceAll>(m, C(a.fx), C(a.fy), C(a.fz));
  reinterpret_cast<Vel3<C>*>(a.u_wall)[(k * a.lz + z) * a.lx + x] = Vel3<C>{q.ux, q.uy, q.uz};
}

// One workgroup's cells of one step: workgroup b of the domain described by a.
//
// ROWS: the row length lx is a multiple of 64, so every wavefront covers 64
// cells of one row.  Its y, z and the byte offsets of its nine source rows are
// then wavefront-uniform (scalar registers): each population's source address
// is a scalar row base plus one of three per-lane x offsets, instead of a
// per-lane three-term sum.  Otherwise (any lx) every lane computes its own.
// OP: kOpRR the RR collision; kOpBGK the baseline BGK operator (bgk_post_q27) with
// the BGK solver's velocity u = j + F/2 (D3Q27, no LES, no wall
// regularisation); kOpRR6 the RR collision and wall regularisation with the
// fourth- to sixth-order terms (rr_post_q27_o6, D3Q27); kOpRRFD the RR
// collision with FDBC at the y-wall rows (fdbc_treatment, D3Q27).
template <int S, class L, bool WALLS, bool LES, int MODE, int FM, bool ROWS, int OP = kOpRR>
RRLBM_INLINE __device__ void fused_body(const FusedArgs& a, uint32_t b) {
  using C = typename StoreT<S>::C;
  using T = typename StoreT<S>::T;
  constexpr int Q = L::Q;
  const uint32_t lxly = a.lx * a.ly;
  uint32_t cell, x, y, z;
  bool active;
  if constexpr (ROWS) {
    // n is a multiple of 64: a wavefront is wholly inside or wholly outside
    uint32_t w0 = __builtin_amdgcn_readfirstlane(b * kBlock + (threadIdx.x & ~63u));
    active = w0 < a.n;
    if (!active) w0 = a.n - 64;
    z = fastdiv(w0, a.div_lxly);
    const uint32_t rem = w0 - z * lxly;
    y = fastdiv(rem, a.div_lx);
    const uint32_t lane = threadIdx.x & 63u;
    x = rem - y * a.lx + lane;
    cell = w0 + lane;
  } else {
    const uint32_t raw = b * kBlock + threadIdx.x;
    active = raw < a.n;
    cell = active ? raw : a.n - 1;
    z = fastdiv(cell, a.div_lxly);
    const uint32_t rem = cell - z * lxly;
    y = fastdiv(rem, a.div_lx);
    x = rem - y * a.lx;
  }
  // upwind source coordinates for c = -1, 0, +1 (periodic wrap)
  const uint32_t xs[3] = {x + 1 == a.lx ? 0u : x + 1, x, x == 0 ? a.lx - 1 : x - 1};
  const uint32_t ys[3] = {y + 1 == a.ly ? 0u : y + 1, y, y == 0 ? a.ly - 1 : y - 1};
  const uint32_t zs[3] = {z + 1 == a.lz ? 0u : z + 1, z, z == 0 ? a.lz - 1 : z - 1};
  const C inv_scale = inv_scale_of<S>(a);
  constexpr uint32_t sz = sizeof(T);
  const uint32_t row_b = a.lx * sz;
  const uint32_t xb[3] = {xs[0] * sz, xs[1] * sz, xs[2] * sz};
  const char* src = reinterpret_cast<const char*>(a.src);
  char* dst = reinterpret_cast<char*>(a.dst);

  // LES step: the WALE stencil's velocity loads go out before the population
  // gather, and the eddy viscosity is computed right after it, so that the
  // arithmetic runs while the population loads are still in flight (loads
  // complete in order)
  [[maybe_unused]] LesStencil<C> les_st;
  if constexpr (LES && MODE == kStep) les_st = les_load<C>(a, x, y, z, xs, ys, zs);
  // FP16 D3Q27 collisions that need only the moments (not the populations
  // after the wall treatment) keep the gathered FP16 values and take their
  // moments with moments_q27_f16; wall cells convert them first.
  constexpr bool kRawF16 = ROWS && S == kF16 && Q == 27 && OP != kOpBGK && MODE != kFinish;
  C h[Q];
  [[maybe_unused]] _Float16 g[Q];
  if constexpr (ROWS) {
    // nine scalar source-row offsets, then population i at
    // src + i*seg + row offset + the lane's x offset.  The population base
    // advances by seg in a scalar register pair held opaque: left to the
    // compiler, the 27 products i*seg of the gather and the 27 of the store are
    // formed up front, overflow the scalar register file and spill into
    // vector-register lanes (about 100 vector instructions per wavefront in
    // the FP16 LES step; 1.5-3 % of the step time).
    uint32_t ro[3][3];
    static_for<3>([&](auto jy) {
      static_for<3>([&](auto jz) {
        constexpr int iy = decltype(jy)::value, iz = decltype(jz)::value;
        ro[iy][iz] = row_offset32(a, row_b, ys[iy], zs[iz]);
      });
    });
    uint64_t pb = reinterpret_cast<uint64_t>(src);
    static_for<Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      constexpr int cx = L::cx(i), cy = L::cy(i), cz = L::cz(i);
      const T v = *reinterpret_cast<const GlobalT<T>*>(pb + ro[cy + 1][cz + 1] + xb[cx + 1]);
      if constexpr (kRawF16)
        g[i] = v;
      else if constexpr (S == kF16)
        h[i] = C(v) * (C(L::w(i)) * inv_scale);
      else
        h[i] = C(v);
      pb += a.lay.seg;
      asm volatile("" : "+s"(pb));
    });
  } else {
    gather_lane<S, L>(a, x, y, z, h);
  }
  [[maybe_unused]] C les_om;
  if constexpr (LES && MODE == kStep) les_om = les_omega_of<C>(a, les_st, x, y, z);
  // this cell in the source and destination buffers: base (scalar with ROWS) + offset
  const char* own_src = src;
  char* own_dst = dst;
  uint32_t own_off;
  if constexpr (ROWS) {
    const uint64_t r0 = row_offset(a, row_b, y, z);
    own_src += r0;
    own_dst += r0;
    own_off = x * sz;
  } else {
    own_off = cell_off<S>(a.lay, cell);
  }

  // body force of the step this launch completes (F) and of the step it
  // collides (G): the uniform force, or forcing_GP per cell
  C Fx = C(a.fx), Fy = C(a.fy), Fz = C(a.fz);
  C Gx = Fx, Gy = Fy, Gz = Fz;
  if constexpr (FM == kForceAll && OP == kOpRR && Q == 27) {
    if (a.gp) {
      gp_force<C>(a, 0, x, y, z, Fx, Fy, Fz);
      gp_force<C>(a, 1, x, y, z, Gx, Gy, Gz);
    }
  }

  bool treated = false;  // wall cell: h holds the populations after the wall treatment
  if constexpr (WALLS) {
    const bool on_wall = ((a.walls & 1) && (x == 0 || x == a.lx - 1)) || ((a.walls & 2) && (y == 0 || y == a.ly - 1)) ||
                         ((a.walls & 4) && (z == 0 || z == a.lz - 1));
    treated = OP == kOpRRFD ? (y == 0 || y == a.ly - 1) : on_wall;
    if (treated) {
      if constexpr (kRawF16) {
        static_for<Q>([&](auto ic) {
          constexpr int i = decltype(ic)::value;
          h[i] = C(g[i]) * (C(L::w(i)) * inv_scale);
        });
      }
      if constexpr (OP == kOpRRFD)
        fdbc_treatment<S, FM>(a, h, x, y, z);
      else
        wall_treatment<S, L, FM, OP == kOpRR6>(a, h, x, y, z, own_src, own_off, Fx, Fy, Fz);
    }
  }

  Moments<C> m;
  if constexpr (kRawF16) {
    if (treated)
      m = moments_q27<C>(h);
    else
      m = moments_q27_f16(g, float(a.f16_unscale * (1.0f / 216.0f)));
  } else if constexpr (Q == 27) {
    m = moments_q27<C>(h);
  } else {
    m = moments_generic<L, C>(h);
  }
  Macro<C> q;
  if constexpr (OP == kOpBGK) {
    q.dr = m.dr;
    q.rho = C(1) + m.dr;
    q.ux = m.jx + C(0.5) * Fx;
    q.uy = m.jy + C(0.5) * Fy;
    q.uz = m.jz + C(0.5) * Fz;
  } else {
    q = macro_of<C, FM>(m, Fx, Fy, Fz);
  }

  if constexpr (MODE == kMacro) {
    if (active) st_v<C>(a.u_out, cell, q.ux, q.uy, q.uz);
    return;
  } else if constexpr (MODE == kFinish) {
    if (active) {
      static_for<Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        a.ref_f[(size_t)i * a.n + cell] = double(h[i]) + L::w(i);
      });
      a.ref_rho[cell] = double(q.rho);
      a.ref_ux[cell] = double(q.ux);
      a.ref_uy[cell] = double(q.uy);
      a.ref_uz[cell] = double(q.uz);
    }
    return;
  } else if constexpr (MODE == kDiag) {
    // FP64 wavefront + workgroup reduction of the end-of-step fields
    double v[kDiagWidth];
    const double r = double(q.rho), ux = double(q.ux), uy = double(q.uy), uz = double(q.uz);
    v[0] = double(q.dr);
    v[1] = ux; v[2] = uy; v[3] = uz;
    v[4] = r * ux; v[5] = r * uy; v[6] = r * uz;
    v[7] = 0.5 * r * (ux * ux + uy * uy + uz * uz);
    v[8] = (__builtin_isfinite(v[0]) && __builtin_isfinite(v[7])) ? 0.0 : 1.0;
    __shared__ double red[kBlock / 64][kDiagWidth];
#pragma unroll
    for (int k = 0; k < kDiagWidth; ++k) {
      double t = active ? v[k] : 0.0;
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) t += __shfl_xor(t, off, 64);
      v[k] = t;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < kDiagWidth; ++k) red[wave][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x < kDiagWidth) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < kBlock / 64; ++w) t += red[w][threadIdx.x];
      a.partials[(size_t)b * kDiagWidth + threadIdx.x] = t;
    }
    __syncthreads();  // red is reused by the workgroup's next block (persistent grids)
    return;
  } else {
    C omega;
    if constexpr (LES) {
      omega = les_om;
      constexpr uint32_t vb = sizeof(Vel3<C>);
      const uint32_t o0 = x * vb;
      auto vrow = [&](const void* base, uint32_t yy, uint32_t zz) {
        return reinterpret_cast<const char*>(base) + (uint64_t)((zz * a.ly + yy) * a.lx) * vb;
      };
      if (a.write_u && active) {
        uint32_t ow = o0;
        asm volatile("" : "+v"(ow));  // 32-bit lane offset in this block (scalar-base store)
        *reinterpret_cast<Vel3<C>*>(const_cast<char*>(vrow(a.u_out, y, z)) + ow) = Vel3<C>{q.ux, q.uy, q.uz};
      }
    }
    if constexpr (!LES) omega = C(a.omega);
    C out[Q];
    if constexpr (OP == kOpBGK) {
      const C inv_tau = C(1) - omega;
      const C fcoef = C(1) - C(0.5) * inv_tau;
      if constexpr (S == kF16)
        bgk_post_q27<C, false>(h, q.dr, q.ux, q.uy, q.uz, Fx, Fy, Fz, inv_tau, fcoef, C(a.f16_scale), out);
      else
        bgk_post_q27<C, true>(h, q.dr, q.ux, q.uy, q.uz, Fx, Fy, Fz, inv_tau, fcoef, C(1), out);
    } else if constexpr (OP == kOpRR6) {
      static_assert(Q == 27, "the sixth-order terms need D3Q27");
      const NeqTensor<C> N = neq_q27<C, FM>(m, q, Fx, Fy, Fz);
      if constexpr (S == kF16)
        rr_post_q27_o6<C, false, FM>(q, N, Fx, Fy, Fz, omega, C(1), C(a.f16_scale), out);
      else
        rr_post_q27_o6<C, true, FM>(q, N, Fx, Fy, Fz, omega, C(1), C(1), out);
    } else if constexpr (Q == 27) {
      const NeqTensor<C> N = neq_q27<C, FM>(m, q, Gx, Gy, Gz);
      if constexpr (S == kF16)
        rr_post_q27<C, false, FM>(q, N, Gx, Gy, Gz, omega, C(1), C(a.f16_scale), out);
      else
        rr_post_q27<C, true, FM>(q, N, Gx, Gy, Gz, omega, C(1), C(1), out);
    } else if constexpr (FM == kForceNone) {
      // D3Q19 without force: moment-space form
      const NeqTensor<C> N = neq_q27<C, kForceNone>(m, q, C(0), C(0), C(0));
      if constexpr (S == kF16)
        rr_post_q19<C, false>(q, N, omega, C(a.f16_scale), out);
      else
        rr_post_q19<C, true>(q, N, omega, C(1), out);
    } else {
      rr_post_generic<L, C>(h, q, Fx, Fy, Fz, omega, C(1), out);
      if constexpr (S == kF16) {
        const C sc = C(a.f16_scale);
        static_for<Q>([&](auto ic) {
          constexpr int i = decltype(ic)::value;
          out[i] = out[i] * (sc * C(1.0 / L::w(i)));
        });
      }
    }
    if (active) store_cell<S, Q>(a, own_dst, own_off, cell, out);
  }
}

// Block t of the launch's tile sequence -> block of cells.  The dispatcher
// deals consecutive workgroup IDs round-robin to the 8 XCDs (each with its own
// L2).  xcd_remap 0: linear; 1: each XCD sweeps one contiguous 1/8 of the
// domain; C > 1: each XCD takes runs of C consecutive blocks, so the
// cache lines two neighbouring blocks share (the +-1 x shift of the gather)
// mostly stay in one L2 while all XCDs still work on nearby memory.
RRLBM_INLINE __device__ uint32_t block_of(const FusedArgs& a, uint32_t t) {
  if (a.xcd_remap == 0) return t;
  if (a.xcd_remap == 1) return (t & 7u) * a.blocks_per_xcd + (t >> 3);
  const uint32_t j = t >> 3, c = a.xcd_remap, q = fastdiv(j, a.div_remap);
  return (q * 8u + (t & 7u)) * c + (j - q * c);
}

// One workgroup per block of 256 cells (grid = a.tiles).
template <int S, class L, bool WALLS, bool LES, int MODE, int FM, bool ROWS, int OP = kOpRR>
__global__ __launch_bounds__(kBlock) void fused_kernel(FusedArgs a) {
  const uint32_t b = block_of(a, blockIdx.x);
  if (b < a.nblocks) fused_body<S, L, WALLS, LES, MODE, FM, ROWS, OP>(a, b);
}

// Persistent grid (step mode): fewer workgroups than blocks, each striding
// over the block sequence; every block is computed once per launch.  A
// separate kernel, because the loop raises the register count of the body
// (loop-invariant addresses stay live), which one-block workgroups avoid.
template <int S, class L, bool WALLS, bool LES, int FM, bool ROWS, int OP = kOpRR>
__global__ __launch_bounds__(kBlock) void fused_kernel_persistent(FusedArgs a) {
  for (uint32_t t = blockIdx.x; t < a.tiles; t += gridDim.x) {
    const uint32_t b = block_of(a, t);
    if (b < a.nblocks) fused_body<S, L, WALLS, LES, kStep, FM, ROWS, OP>(a, b);
  }
}

// Two cells per thread (STEP mode, D3Q27, no LES): lane l of wavefront w of
// workgroup b takes cells base + l and base + l + 64, base = b*512 + w*128, so
// every load and store instruction still covers 64 consecutive cells while
// each wavefront keeps twice as many population loads in flight -- the better
// layout when the per-cell arithmetic is a large share of the time.
template <int S, bool WALLS, int FM>
RRLBM_INLINE __device__ void fused_pair_body(const FusedArgs& a, uint32_t b) {
  using C = typename StoreT<S>::C;
  using L = D3Q27;
  constexpr int Q = 27;
  const uint32_t base = b * (2 * kBlock) + (threadIdx.x >> 6) * 128 + (threadIdx.x & 63);
  const uint32_t craw[2] = {base, base + 64};
  const uint32_t lxly = a.lx * a.ly;
  C h[2][Q];
  uint32_t cell[2], cx_[2], cy_[2], cz_[2];
  bool active[2];
  static_for<2>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    active[k] = craw[k] < a.n;
    cell[k] = active[k] ? craw[k] : a.n - 1;
    const uint32_t z = fastdiv(cell[k], a.div_lxly);
    const uint32_t rem = cell[k] - z * lxly;
    const uint32_t y = fastdiv(rem, a.div_lx);
    const uint32_t x = rem - y * a.lx;
    cx_[k] = x; cy_[k] = y; cz_[k] = z;
    gather_lane<S, L>(a, x, y, z, h[k]);
  });
  const C Fx = C(a.fx), Fy = C(a.fy), Fz = C(a.fz);
  const C omega = C(a.omega);
  static_for<2>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    const uint32_t off = cell_off<S>(a.lay, cell[k]);
    if constexpr (WALLS) {
      const uint32_t x = cx_[k], y = cy_[k], z = cz_[k];
      const bool on_wall = ((a.walls & 1) && (x == 0 || x == a.lx - 1)) ||
                           ((a.walls & 2) && (y == 0 || y == a.ly - 1)) || ((a.walls & 4) && (z == 0 || z == a.lz - 1));
      if (on_wall) wall_treatment<S, L, FM>(a, h[k], x, y, z, a.src, off, C(a.fx), C(a.fy), C(a.fz));
    }
    const Moments<C> m = moments_q27<C>(h[k]);
    const Macro<C> q = macro_of<C, FM>(m, Fx, Fy, Fz);
    const NeqTensor<C> N = neq_q27<C, FM>(m, q, Fx, Fy, Fz);
    C out[Q];
    if constexpr (S == kF16)
      rr_post_q27<C, false, FM>(q, N, Fx, Fy, Fz, omega, C(1), C(a.f16_scale), out);
    else
      rr_post_q27<C, true, FM>(q, N, Fx, Fy, Fz, omega, C(1), C(1), out);
    if (active[k]) store_cell<S, Q>(a, a.dst, off, cell[k], out);
  });
}

template <int S, bool WALLS, int FM, bool PERSIST>
__global__ __launch_bounds__(kBlock) void fused_pair_kernel(FusedArgs a) {
  if constexpr (PERSIST) {
    for (uint32_t t = blockIdx.x; t < a.tiles; t += gridDim.x) {
      const uint32_t b = block_of(a, t);
      if (b < a.nblocks) fused_pair_body<S, WALLS, FM>(a, b);
    }
  } else {
    const uint32_t b = block_of(a, blockIdx.x);
    if (b < a.nblocks) fused_pair_body<S, WALLS, FM>(a, b);
  }
}

// ---- Two steps per launch ----------------------------------------------
//
// The stored state of step t (post-collision populations in a.src) goes to
// that of step t + 2 (a.dst) in one pass, so each population crosses HBM once
// per two steps.  A workgroup owns a tile of kT2X x kT2Y cells in (x, y) and a
// chunk of planes [z0, z1) in z, and marches k = z0 - 1 .. z1.  In FP64
// (fused2_kernel), per plane:
//   A. step t+1 at plane k on the tile and a one-cell halo around it
//      ((kT2X + 2) x (kT2Y + 2) cells), gathered from a.src as the one-step
//      kernel does, into registers; the populations with c_z = -1 go to LDS;
//   B. step t+2 at plane k - 1 on the tile, gathered from the LDS ring
//      (c_z = +1 from plane k - 2, 0 from k - 1, -1 from k), stored to a.dst;
//   C. the populations of plane k with c_z = 0 and +1 go to LDS.
// The ring holds plane k - 1 in full while B runs (the wall cells' bounce-back
// reads their own populations of that plane): c_z = -1 and +1 in two slots
// each, c_z = 0 in one, 45 x 396 values.  FP32 (fused2p_kernel) overlaps the
// two steps in different wavefronts.  Both steps call the one-step kernel's
// per-cell functions, so the result equals two one-step launches: bit for bit
// where the compiler contracts the same multiply-adds in both kernels, to an
// ulp where it does not.  Planes z0 - 1 and z1 of step t+1 are computed by
// both chunks that border them.
constexpr uint32_t kH2X = kT2X + 2, kH2Y = kT2Y + 2, kH2N = kH2X * kH2Y;

// Index of population i among the nine with the same c_z.
template <class L>
RRLBM_HD constexpr int zgroup_index(int i) {
  return (L::cx(i) + 1) * 3 + (L::cy(i) + 1);
}

// The collision of a cell in the two-step kernels, whose pulled (and wall-
// treated) populations are h: OP = kOpRR the RR collision, kOpBGK the BGK
// operator with the BGK solver's velocity j + F/2, kOpRR6 the RR collision
// with the fourth- to sixth-order terms; the per-cell functions and force
// forms of the one-step kernel's collision.
template <typename C, int FM, int OP>
RRLBM_INLINE __device__ void collide_cell(const C (&h)[27], C Fx, C Fy, C Fz, C omega, C (&out)[27]) {
  const Moments<C> m = moments_q27<C>(h);
  if constexpr (OP == kOpBGK) {
    const C inv_tau = C(1) - omega;
    const C fcoef = C(1) - C(0.5) * inv_tau;
    bgk_post_q27<C, true>(h, m.dr, m.jx + C(0.5) * Fx, m.jy + C(0.5) * Fy, m.jz + C(0.5) * Fz, Fx, Fy, Fz, inv_tau,
                          fcoef, C(1), out);
  } else {
    const Macro<C> q = macro_of<C, FM>(m, Fx, Fy, Fz);
    const NeqTensor<C> N = neq_q27<C, FM>(m, q, Fx, Fy, Fz);
    if constexpr (OP == kOpRR6)
      rr_post_q27_o6<C, true, FM>(q, N, Fx, Fy, Fz, omega, C(1), C(1), out);
    else
      rr_post_q27<C, true, FM>(q, N, Fx, Fy, Fz, omega, C(1), C(1), out);
  }
}

// FP64 (the FP32 form is fused2p_kernel below): 142.6 KB of LDS, one
// workgroup of 7 wavefronts per compute unit.
template <int S, bool WALLS, int FM, int OP>
__global__ __launch_bounds__(kT2Threads) __attribute__((amdgpu_waves_per_eu(2))) void fused2_kernel(FusedArgs a,
                                                                                                   Fused2Args t) {
  using C = typename StoreT<S>::C;
  using T = typename StoreT<S>::T;
  using L = D3Q27;
  constexpr int Q = 27;
  static_assert(S == kF64, "two steps per launch in FP64 storage (FP32: fused2p_kernel)");
  // work item: XCD (blockIdx.x mod 8) takes a contiguous run of items, x
  // fastest, so the workgroups an XCD runs at one time cover whole rows of
  // tiles and share their halo cache lines in its L2
  const uint32_t f = (blockIdx.x & 7u) * t.per_xcd + (blockIdx.x >> 3);
  if (f >= t.items) return;  // the whole workgroup: no barrier is left half-reached
  const uint32_t tx = f % t.ntx, ty = (f / t.ntx) % t.nty, ch = f / (t.nty * t.ntx);
  const uint32_t x0 = tx * kT2X, y0 = ty * kT2Y, z0 = ch * t.zc;
  const uint32_t zlen = min(t.zc, a.lz - z0);

  __shared__ C sM[2][9][kH2N], sZ[9][kH2N], sP[2][9][kH2N];

  // w is uniform across a wavefront; readfirstlane lets the compiler keep the
  // store row's base in scalar registers
  const uint32_t tid = threadIdx.x, w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63u;
  // step t+1: wavefronts 0-5 take halo row w, columns 1..64; wavefront 6
  // lanes 0-11 the two halo columns 0 and 65 of the six rows
  uint32_t hx, hy;
  bool h_on = true;
  if (w < kH2Y) {
    hy = w;
    hx = lane + 1;
  } else {
    h_on = lane < 2 * kH2Y;
    hy = h_on ? lane >> 1 : 0u;
    hx = (lane & 1u) ? kH2X - 1 : 0u;
  }
  const uint32_t hidx = hy * kH2X + hx;
  const uint32_t gx = (x0 + hx + a.lx - 1) % a.lx, gy = (y0 + hy + a.ly - 1) % a.ly;
  // step t+2: wavefronts 0-3, row w of the tile, column lane
  const bool i_on = w < kT2Y;
  const uint32_t iidx = (w + 1) * kH2X + lane + 1;
  const uint32_t ix = x0 + lane, iy = y0 + w;

  const C Fx = C(a.fx), Fy = C(a.fy), Fz = C(a.fz);
  const C omega = C(a.omega);
  constexpr uint32_t sz = sizeof(T);
  const uint32_t row_b = a.lx * sz;
  auto on_wall = [&](uint32_t x, uint32_t y, uint32_t z) {
    return ((a.walls & 1) && (x == 0 || x == a.lx - 1)) || ((a.walls & 2) && (y == 0 || y == a.ly - 1)) ||
           ((a.walls & 4) && (z == 0 || z == a.lz - 1));
  };
  auto collide = [&](C (&h)[Q], C (&out)[Q]) { collide_cell<C, FM, OP>(h, Fx, Fy, Fz, omega, out); };

  // Halo rows: (y, z) is uniform across the wavefront, so the nine source rows
  // of the step-1 gather are scalar bases and each lane adds its x offset.
  const uint32_t gyr = (y0 + w + a.ly - 1) % a.ly;  // = gy in the halo-row wavefronts, uniform
  const uint32_t ysr[3] = {gyr + 1 == a.ly ? 0u : gyr + 1, gyr, gyr == 0 ? a.ly - 1 : gyr - 1};
  const uint32_t xbr[3] = {(gx + 1 == a.lx ? 0u : gx + 1) * sz, gx * sz, (gx == 0 ? a.lx - 1 : gx - 1) * sz};
  auto row_gather = [&](uint32_t kk, C (&dst)[Q]) {
    const uint32_t zs[3] = {kk + 1 == a.lz ? 0u : kk + 1, kk, kk == 0 ? a.lz - 1 : kk - 1};
    uint32_t ro[3][3];
    static_for<3>([&](auto jy) {
      static_for<3>([&](auto jz) {
        constexpr int iy2 = decltype(jy)::value, iz2 = decltype(jz)::value;
        ro[iy2][iz2] = row_offset32(a, row_b, ysr[iy2], zs[iz2]);
      });
    });
    // the lane's x offsets are made opaque per call: a zero extension of them
    // hoisted out of the z loop would make every load a 64-bit per-lane
    // address add instead of a scalar base plus a 32-bit lane offset
    uint32_t xo[3] = {xbr[0], xbr[1], xbr[2]};
    asm volatile("" : "+v"(xo[0]), "+v"(xo[1]), "+v"(xo[2]));
    uint64_t pb = reinterpret_cast<uint64_t>(a.src);
    static_for<Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      constexpr int cx = L::cx(i), cy = L::cy(i), cz = L::cz(i);
      dst[i] = C(*reinterpret_cast<const GlobalT<T>*>(pb + ro[cy + 1][cz + 1] + xo[cx + 1]));
      pb += a.lay.seg;
      asm volatile("" : "+s"(pb));
    });
  };

  for (uint32_t it = 0; it < zlen + 2; ++it) {
    const uint32_t k = (z0 + a.lz - 1 + it) % a.lz;  // plane of step t+1
    const uint32_t p = it & 1u;                      // ring slot of plane k (k - 2 before C)
    // A. step t+1 at plane k
    C o1[Q];
    if (h_on) {
      C h[Q];
      if (w < kH2Y) {
        row_gather(k, h);
      } else {
        gather_lane<S, L>(a, gx, gy, k, h);
      }
      if constexpr (WALLS) {
        if (on_wall(gx, gy, k)) {
          const uint32_t off = cell_off<S>(a.lay, (k * a.ly + gy) * a.lx + gx);
          wall_treatment<S, L, FM, OP == kOpRR6>(a, h, gx, gy, k, a.src, off, Fx, Fy, Fz);
        }
      }
      collide(h, o1);
      static_for<Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        if constexpr (L::cz(i) < 0) sM[p][zgroup_index<L>(i)][hidx] = o1[i];
      });
    }
    __syncthreads();
    // B. step t+2 at plane k - 1
    if (it >= 2 && i_on) {
      const uint32_t q = 1u - p;  // ring slot of plane k - 1
      const uint32_t kz = (k + a.lz - 1) % a.lz;
      C h[Q];
      static_for<Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        constexpr int cx = L::cx(i), cy = L::cy(i), cz = L::cz(i), j = zgroup_index<L>(i);
        const uint32_t src = iidx - (uint32_t)(cy * (int)kH2X + cx);
        if constexpr (cz > 0)
          h[i] = sP[p][j][src];
        else if constexpr (cz == 0)
          h[i] = sZ[j][src];
        else
          h[i] = sM[p][j][src];
      });
      if constexpr (WALLS) {
        if (on_wall(ix, iy, kz)) {
          wall_treatment_own<L, C, FM, OP == kOpRR6>(
              a, h, ix, iy, kz,
              [&](int o) -> C {
                const int j = (L::cx(o) + 1) * 3 + (L::cy(o) + 1);
                return L::cz(o) > 0 ? sP[q][j][iidx] : (L::cz(o) == 0 ? sZ[j][iidx] : sM[q][j][iidx]);
              },
              Fx, Fy, Fz);
        }
      }
      C o2[Q];
      collide(h, o2);
      char* own_dst = reinterpret_cast<char*>(a.dst) + row_offset(a, row_b, iy, kz);
      store_cell<S, Q>(a, own_dst, ix * sz, (kz * a.ly + iy) * a.lx + ix, o2);
    }
    __syncthreads();
    // C. the rest of plane k into the ring
    if (h_on) {
      static_for<Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        if constexpr (L::cz(i) == 0)
          sZ[zgroup_index<L>(i)][hidx] = o1[i];
        else if constexpr (L::cz(i) > 0)
          sP[p][zgroup_index<L>(i)][hidx] = o1[i];
      });
    }
  }
}

// FP32: the two steps overlap in different wavefronts of the workgroup.
// Wavefronts 0-6 compute step t+1 at plane k (the six halo rows and the two
// halo columns, as above) and write all its populations into the ring;
// wavefronts 7-10 compute step t+2 at plane k - 2 from planes k - 3 .. k - 1
// of the ring, which the barrier that ends each plane has made visible.  The
// ring keeps the c_z = +1 populations of four planes and the c_z = 0 and -1
// populations of three (90 x 396 values, 142.6 KB: one workgroup of 11
// wavefronts per compute unit).  The step-t+1 wavefronts issue the gather of
// the next plane before the barrier.
constexpr uint32_t kT2PThreads = 11 * 64;

template <bool WALLS, int FM, int OP>
__global__ __launch_bounds__(kT2PThreads) __attribute__((amdgpu_waves_per_eu(3))) void fused2p_kernel(
    FusedArgs a, Fused2Args t) {
  constexpr int S = kF32;
  using C = float;
  using T = float;
  using L = D3Q27;
  constexpr int Q = 27;
  constexpr uint32_t kH2Y = kT2Y + 2, kH2N = kH2X * kH2Y;
  const uint32_t f = (blockIdx.x & 7u) * t.per_xcd + (blockIdx.x >> 3);
  if (f >= t.items) return;  // the whole workgroup: no barrier is left half-reached
  const uint32_t tx = f % t.ntx, ty = (f / t.ntx) % t.nty, ch = f / (t.nty * t.ntx);
  const uint32_t x0 = tx * kT2X, y0 = ty * kT2Y, z0 = ch * t.zc;
  const uint32_t zlen = min(t.zc, a.lz - z0);

  __shared__ C sP[4][9][kH2N], sZ[3][9][kH2N], sM[3][9][kH2N];

  const uint32_t tid = threadIdx.x, w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63u;
  const bool step1 = w <= kH2Y;  // wavefronts 0-6
  // The halo-column wavefront gathers with per-lane addresses when its plane
  // starts (prefetched like the rows' gathers it measured 10 % slower, with no
  // register spills), so it finishes its plane last; issue priority over the
  // other wavefronts of its SIMD shortens each plane by 6 %, and the step-t+2
  // wavefronts ahead of the halo rows (whose gathers were issued a plane
  // earlier) by another 2 %.
  if (w == kH2Y)
    __builtin_amdgcn_s_setprio(3);
  else if (w > kH2Y)
    __builtin_amdgcn_s_setprio(1);
  // step t+1 cell (wavefronts 0-5: halo row w, columns 1..64; wavefront 6
  // lanes 0-11: the halo columns 0 and 65)
  uint32_t hx, hy;
  bool h_on = step1;
  if (w < kH2Y) {
    hy = w;
    hx = lane + 1;
  } else {
    h_on = step1 && lane < 2 * kH2Y;
    hy = h_on ? lane >> 1 : 0u;
    hx = (lane & 1u) ? kH2X - 1 : 0u;
  }
  const uint32_t hidx = hy * kH2X + hx;
  const uint32_t gx = (x0 + hx + a.lx - 1) % a.lx, gy = (y0 + hy + a.ly - 1) % a.ly;
  // step t+2 cell (wavefronts 7-10: tile row r, column lane)
  const uint32_t r = step1 ? 0u : w - (kH2Y + 1);
  const uint32_t iidx = (r + 1) * kH2X + lane + 1;
  const uint32_t ix = x0 + lane, iy = y0 + r;

  const C Fx = C(a.fx), Fy = C(a.fy), Fz = C(a.fz);
  const C omega = C(a.omega);
  constexpr uint32_t sz = sizeof(T);
  const uint32_t row_b = a.lx * sz;
  auto on_wall = [&](uint32_t x, uint32_t y, uint32_t z) {
    return ((a.walls & 1) && (x == 0 || x == a.lx - 1)) || ((a.walls & 2) && (y == 0 || y == a.ly - 1)) ||
           ((a.walls & 4) && (z == 0 || z == a.lz - 1));
  };
  auto collide = [&](C (&h)[Q], C (&out)[Q]) { collide_cell<C, FM, OP>(h, Fx, Fy, Fz, omega, out); };
  const uint32_t gyr = (y0 + w + a.ly - 1) % a.ly;  // = gy in the halo-row wavefronts, uniform
  const uint32_t ysr[3] = {gyr + 1 == a.ly ? 0u : gyr + 1, gyr, gyr == 0 ? a.ly - 1 : gyr - 1};
  const uint32_t xbr[3] = {(gx + 1 == a.lx ? 0u : gx + 1) * sz, gx * sz, (gx == 0 ? a.lx - 1 : gx - 1) * sz};
  auto row_gather = [&](uint32_t kk, C (&dst)[Q]) {
    const uint32_t zs[3] = {kk + 1 == a.lz ? 0u : kk + 1, kk, kk == 0 ? a.lz - 1 : kk - 1};
    uint32_t ro[3][3];
    static_for<3>([&](auto jy) {
      static_for<3>([&](auto jz) {
        constexpr int iy2 = decltype(jy)::value, iz2 = decltype(jz)::value;
        ro[iy2][iz2] = row_offset32(a, row_b, ysr[iy2], zs[iz2]);
      });
    });
    // the lane's x offsets are made opaque per call: a zero extension of them
    // hoisted out of the z loop would make every load a 64-bit per-lane
    // address add instead of a scalar base plus a 32-bit lane offset
    uint32_t xo[3] = {xbr[0], xbr[1], xbr[2]};
    asm volatile("" : "+v"(xo[0]), "+v"(xo[1]), "+v"(xo[2]));
    uint64_t pb = reinterpret_cast<uint64_t>(a.src);
    static_for<Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      constexpr int cx = L::cx(i), cy = L::cy(i), cz = L::cz(i);
      dst[i] = C(*reinterpret_cast<const GlobalT<T>*>(pb + ro[cy + 1][cz + 1] + xo[cx + 1]));
      pb += a.lay.seg;
      asm volatile("" : "+s"(pb));
    });
  };
  C hn[Q];
  if (w < kH2Y) row_gather((z0 + a.lz - 1) % a.lz, hn);

  // iteration it: step t+1 at plane z0 - 1 + it (it <= zlen + 1), step t+2 at
  // plane z0 - 3 + it (it >= 3); ring slots of the plane of iteration i:
  // sP[i % 4], sZ[i % 3], sM[i % 3]
  for (uint32_t it = 0; it < zlen + 3; ++it) {
    const uint32_t p4 = it & 3u, p3 = it % 3u;
    if (step1) {
      if (it < zlen + 2 && h_on) {
        const uint32_t k = (z0 + a.lz - 1 + it) % a.lz;
        C h[Q];
        if (w < kH2Y)
          static_for<Q>([&](auto ic) { h[decltype(ic)::value] = hn[decltype(ic)::value]; });
        else
          gather_lane<S, L>(a, gx, gy, k, h);
        if constexpr (WALLS) {
          if (on_wall(gx, gy, k)) {
            const uint32_t off = cell_off<S>(a.lay, (k * a.ly + gy) * a.lx + gx);
            wall_treatment<S, L, FM, OP == kOpRR6>(a, h, gx, gy, k, a.src, off, Fx, Fy, Fz);
          }
        }
        C o1[Q];
        collide(h, o1);
        static_for<Q>([&](auto ic) {
          constexpr int i = decltype(ic)::value;
          constexpr int j = zgroup_index<L>(i);
          if constexpr (L::cz(i) > 0)
            sP[p4][j][hidx] = o1[i];
          else if constexpr (L::cz(i) == 0)
            sZ[p3][j][hidx] = o1[i];
          else
            sM[p3][j][hidx] = o1[i];
        });
        if (w < kH2Y && it + 1 < zlen + 2) {
          const uint32_t kn = k + 1 == a.lz ? 0u : k + 1;
          row_gather(kn, hn);
        }
      }
    } else if (it >= 3) {
      // step t+2 at plane kz = z0 - 3 + it; planes kz - 1, kz, kz + 1 were
      // written in iterations it - 3, it - 2, it - 1
      const uint32_t kz = (z0 + it - 3) % a.lz;
      const uint32_t pm3 = (it + 1) & 3u, pm2 = (it + 2) & 3u;  // P slots of planes kz - 1, kz
      const uint32_t qm2 = (it + 1) % 3u, qm1 = (it + 2) % 3u;  // Z / M slots of planes kz, kz + 1
      C h[Q];
      static_for<Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        constexpr int cx = L::cx(i), cy = L::cy(i), cz = L::cz(i), j = zgroup_index<L>(i);
        const uint32_t src = iidx - (uint32_t)(cy * (int)kH2X + cx);
        if constexpr (cz > 0)
          h[i] = sP[pm3][j][src];
        else if constexpr (cz == 0)
          h[i] = sZ[qm2][j][src];
        else
          h[i] = sM[qm1][j][src];
      });
      if constexpr (WALLS) {
        if (on_wall(ix, iy, kz)) {
          wall_treatment_own<L, C, FM, OP == kOpRR6>(
              a, h, ix, iy, kz,
              [&](int o) -> C {
                const int j = (L::cx(o) + 1) * 3 + (L::cy(o) + 1);
                return L::cz(o) > 0 ? sP[pm2][j][iidx] : (L::cz(o) == 0 ? sZ[qm2][j][iidx] : sM[qm2][j][iidx]);
              },
              Fx, Fy, Fz);
        }
      }
      C o2[Q];
      collide(h, o2);
      char* own_dst = reinterpret_cast<char*>(a.dst) + row_offset(a, row_b, iy, kz);
      store_cell<S, Q>(a, own_dst, ix * sz, (kz * a.ly + iy) * a.lx + ix, o2);
    }
    __syncthreads();
  }
}

// Two-step launch of one storage type, walls flag and operator; the force
// forms are those of the one-step kernels (RR: none, x only or general; the
// sixth-order RR: x only or general; BGK: general), so that both give the same
// result.
template <int S, bool W, int OP>
hipError_t launch_fused2_op(int fm, const FusedArgs& a, const Fused2Args& t, hipStream_t s) {
  const dim3 grid(8u * t.per_xcd), block(S == kF32 ? kT2PThreads : kT2Threads);
  auto go = [&](auto fmc) {
    constexpr int F = decltype(fmc)::value;
    if constexpr (S == kF32)
      hipLaunchKernelGGL((fused2p_kernel<W, F, OP>), grid, block, 0, s, a, t);
    else
      hipLaunchKernelGGL((fused2_kernel<S, W, F, OP>), grid, block, 0, s, a, t);
  };
  if constexpr (OP == kOpBGK) {
    go(std::integral_constant<int, kForceAll>{});
  } else if constexpr (OP == kOpRR6) {
    if (fm == kForceX)
      go(std::integral_constant<int, kForceX>{});
    else
      go(std::integral_constant<int, kForceAll>{});
  } else {
    switch (fm) {
      case kForceNone: go(std::integral_constant<int, kForceNone>{}); break;
      case kForceX: go(std::integral_constant<int, kForceX>{}); break;
      default: go(std::integral_constant<int, kForceAll>{}); break;
    }
  }
  return hipGetLastError();
}

template <int S, int OP>
hipError_t launch_fused2_w(bool walls, int fm, const FusedArgs& a, const Fused2Args& t, hipStream_t s) {
  return walls ? launch_fused2_op<S, true, OP>(fm, a, t, s) : launch_fused2_op<S, false, OP>(fm, a, t, s);
}

template <int S>
hipError_t launch_fused2_storage(bool walls, int op, int fm, const FusedArgs& a, const Fused2Args& t,
                                 hipStream_t s) {
  if constexpr (S == kF16) {
    return hipErrorInvalidValue;
  } else {
    switch (op) {
      case kOpRR: return launch_fused2_w<S, kOpRR>(walls, fm, a, t, s);
      case kOpBGK: return launch_fused2_w<S, kOpBGK>(walls, fm, a, t, s);
      case kOpRR6: return launch_fused2_w<S, kOpRR6>(walls, fm, a, t, s);
      default: return hipErrorInvalidValue;
    }
  }
}

// Batched step over many periodic D3Q27 instances (ensembles).  The launch
// walks a work list of tiles, work[t] = (instance, block of 256 cells of that
// instance): with one workgroup per tile, or a persistent grid whose
// workgroup g takes tiles g, g + grid, ...  Each instance's domain, relaxation
// time, layout and A/B buffers come from its FusedArgs in `table` (parity 1
// swaps src and dst, so one captured graph alternates buffers; `step` drives
// the FP16 stochastic-rounding dither).
template <int S>
RRLBM_INLINE __device__ void batched_tile(const FusedArgs* __restrict__ table, uint2 w, uint32_t parity,
                                         uint32_t step, const uint32_t* step_base) {
  FusedArgs a = table[w.x];
  a.sr_step = step;
  a.sr_base = step_base;
  if (parity) {
    const void* s = a.src;
    a.src = a.dst;
    a.dst = const_cast<void*>(s);
  }
  if (a.lx % 64u == 0)
    fused_body<S, D3Q27, false, false, kStep, kForceNone, true>(a, w.y);
  else
    fused_body<S, D3Q27, false, false, kStep, kForceNone, false>(a, w.y);
}

template <int S, bool PERSIST>
__global__ __launch_bounds__(kBlock) void batched_kernel(const FusedArgs* __restrict__ table,
                                                         const uint2* __restrict__ work, uint32_t ntiles,
                                                         uint32_t parity, uint32_t step, const uint32_t* step_base) {
  if constexpr (PERSIST) {
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) batched_tile<S>(table, work[t], parity, step, step_base);
  } else {
    if (blockIdx.x < ntiles) batched_tile<S>(table, work[blockIdx.x], parity, step, step_base);
  }
}

template <int S>
static hipError_t launch_batched_s(const FusedArgs* table, const uint2* work, uint32_t ntiles, uint32_t grid,
                                   uint32_t parity, uint32_t step, const uint32_t* step_base, hipStream_t s) {
  if (grid < ntiles)
    hipLaunchKernelGGL((batched_kernel<S, true>), dim3(grid), dim3(kBlock), 0, s, table, work, ntiles, parity, step,
                       step_base);
  else
    hipLaunchKernelGGL((batched_kernel<S, false>), dim3(grid), dim3(kBlock), 0, s, table, work, ntiles, parity, step,
                       step_base);
  return hipGetLastError();
}

template <int S, class L, bool W, bool LES, int FM, bool R>
static hipError_t launch_mode(int mode, const FusedArgs& a, uint32_t grid, hipStream_t s) {
  // only the step kernel strides over the blocks: a smaller grid for another
  // pass would leave blocks out
  if (mode != kStep && grid < a.tiles) return hipErrorInvalidConfiguration;
  if (mode == kStep && grid < a.tiles) {
    hipLaunchKernelGGL((fused_kernel_persistent<S, L, W, LES, FM, R>), dim3(grid), dim3(kBlock), 0, s, a);
    return hipGetLastError();
  }
  switch (mode) {
    case kStep: hipLaunchKernelGGL((fused_kernel<S, L, W, LES, kStep, FM, R>), dim3(grid), dim3(kBlock), 0, s, a); break;
    case kMacro: hipLaunchKernelGGL((fused_kernel<S, L, W, false, kMacro, FM, R>), dim3(grid), dim3(kBlock), 0, s, a); break;
    case kFinish: hipLaunchKernelGGL((fused_kernel<S, L, W, false, kFinish, FM, R>), dim3(grid), dim3(kBlock), 0, s, a); break;
    case kDiag: hipLaunchKernelGGL((fused_kernel<S, L, W, false, kDiag, FM, R>), dim3(grid), dim3(kBlock), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int S, class L, bool W, bool LES, int FM>
static hipError_t launch_rows(int mode, const FusedArgs& a, uint32_t grid, hipStream_t s) {
  return a.lx % 64u == 0 ? launch_mode<S, L, W, LES, FM, true>(mode, a, grid, s)
                         : launch_mode<S, L, W, LES, FM, false>(mode, a, grid, s);
}

template <int S, class L, bool W, bool LES>
static hipError_t launch_force(int fm, int mode, const FusedArgs& a, uint32_t grid, hipStream_t s) {
  switch (fm) {
    case kForceNone: return launch_rows<S, L, W, LES, kForceNone>(mode, a, grid, s);
    case kForceX: return launch_rows<S, L, W, LES, kForceX>(mode, a, grid, s);
    default: return launch_rows<S, L, W, LES, kForceAll>(mode, a, grid, s);
  }
}

template <int S>
static hipError_t launch_storage(int q, bool walls, bool les, int fm, int mode, const FusedArgs& a, uint32_t grid,
                                 hipStream_t s) {
  if (q == 27) {
    if (walls)
      return les ? launch_force<S, D3Q27, true, true>(fm, mode, a, grid, s)
                 : launch_force<S, D3Q27, true, false>(fm, mode, a, grid, s);
    return les ? launch_force<S, D3Q27, false, true>(fm, mode, a, grid, s)
               : launch_force<S, D3Q27, false, false>(fm, mode, a, grid, s);
  }
  if (q == 19) {
    if (les) return hipErrorInvalidValue;  // WALE is provided on D3Q27 only
    // D3Q19: the moment-space form without force, the per-direction
    // formulation with one
    if (fm == kForceNone)
      return walls ? launch_rows<S, D3Q19, true, false, kForceNone>(mode, a, grid, s)
                   : launch_rows<S, D3Q19, false, false, kForceNone>(mode, a, grid, s);
    return walls ? launch_rows<S, D3Q19, true, false, kForceAll>(mode, a, grid, s)
                 : launch_rows<S, D3Q19, false, false, kForceAll>(mode, a, grid, s);
  }
  return hipErrorInvalidValue;
}

template <int S, bool W, int FM>
static hipError_t launch_pair_grid(const FusedArgs& a, uint32_t grid, hipStream_t s) {
  if (grid < a.tiles)
    hipLaunchKernelGGL((fused_pair_kernel<S, W, FM, true>), dim3(grid), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL((fused_pair_kernel<S, W, FM, false>), dim3(grid), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

template <int S, bool W>
static hipError_t launch_pair_force(int fm, const FusedArgs& a, uint32_t grid, hipStream_t s) {
  switch (fm) {
    case kForceNone: return launch_pair_grid<S, W, kForceNone>(a, grid, s);
    case kForceX: return launch_pair_grid<S, W, kForceX>(a, grid, s);
    default: return launch_pair_grid<S, W, kForceAll>(a, grid, s);
  }
}

inline int force_mode_of(const FusedArgs& a) {
  if (a.gp) return kForceAll;  // forcing_GP: all three components, per cell
  return (a.fx == 0.0 && a.fy == 0.0 && a.fz == 0.0) ? kForceNone : (a.fy == 0.0 && a.fz == 0.0) ? kForceX : kForceAll;
}

// The alternative operators (BGK baseline, sixth-order RR): D3Q27, no LES,
// general force (step, finish and diagnostics passes).
template <int S, bool W, bool R, int OP>
static hipError_t launch_op_mode(int mode, const FusedArgs& a, uint32_t grid, hipStream_t s) {
  if constexpr (OP == kOpRR6 || OP == kOpRRFD) {
    // a force along x only (the channel's): the step kernel without the y and
    // z force terms of the collision and the wall treatment
    if (mode == kStep && force_mode_of(a) == kForceX) {
      if (grid < a.tiles)
        hipLaunchKernelGGL((fused_kernel_persistent<S, D3Q27, W, false, kForceX, R, OP>), dim3(grid), dim3(kBlock),
                           0, s, a);
      else
        hipLaunchKernelGGL((fused_kernel<S, D3Q27, W, false, kStep, kForceX, R, OP>), dim3(grid), dim3(kBlock), 0, s,
                           a);
      return hipGetLastError();
    }
  }
  if (mode != kStep && grid < a.tiles) return hipErrorInvalidConfiguration;
  if (mode == kStep && grid < a.tiles) {
    hipLaunchKernelGGL((fused_kernel_persistent<S, D3Q27, W, false, kForceAll, R, OP>), dim3(grid), dim3(kBlock), 0,
                       s, a);
    return hipGetLastError();
  }
  switch (mode) {
    case kStep: hipLaunchKernelGGL((fused_kernel<S, D3Q27, W, false, kStep, kForceAll, R, OP>), dim3(grid), dim3(kBlock), 0, s, a); break;
    case kFinish: hipLaunchKernelGGL((fused_kernel<S, D3Q27, W, false, kFinish, kForceAll, R, OP>), dim3(grid), dim3(kBlock), 0, s, a); break;
    case kDiag: hipLaunchKernelGGL((fused_kernel<S, D3Q27, W, false, kDiag, kForceAll, R, OP>), dim3(grid), dim3(kBlock), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int S, int OP>
static hipError_t launch_op_storage(bool walls, int mode, const FusedArgs& a, uint32_t grid, hipStream_t s) {
  const bool rows = a.lx % 64u == 0;
  if constexpr (OP == kOpRRFD) {
    // y walls always
    (void)walls;
    return rows ? launch_op_mode<S, true, true, OP>(mode, a, grid, s)
                : launch_op_mode<S, true, false, OP>(mode, a, grid, s);
  } else {
    if (walls)
      return rows ? launch_op_mode<S, true, true, OP>(mode, a, grid, s)
                  : launch_op_mode<S, true, false, OP>(mode, a, grid, s);
    return rows ? launch_op_mode<S, false, true, OP>(mode, a, grid, s)
                : launch_op_mode<S, false, false, OP>(mode, a, grid, s);
  }
}

}  // namespace rrlbm
