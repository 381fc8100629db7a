// Per-cell arithmetic of the recursively regularised (RR) collision.
//
// Two formulations of the same operator, both compiled for gfx950 device code:
//
//  * rr_post_q27 -- the D3Q27 fast path.  It works in moment space: the 10 raw
//    moments of the shifted populations come from a 3-stage tensor-product
//    reduction (D3Q27 = {-1,0,1}^3), the post-collision state is assembled as
//    the 23 non-zero coefficients of a polynomial in (c_x, c_y, c_z) (degree
//    <= 2 per axis, which spans every function on the 27 velocities), and the
//    27 outputs come from a 3-stage tensor-product evaluation of that
//    polynomial.  About 400 VALU operations per cell instead of the ~3000 of
//    the per-direction formulation.
//  * rr_post_generic -- the per-direction formulation of
//    RR27V33_code_DNS_LES/collision.cu:52-118 with compile-time Hermite
//    constants.  Used for D3Q19 (Malaspinas third-order pairs,
//    Taylor_Green/D3Q19RR.f90:131-136,230-235) and as an independent check of
//    the fast path.
//
// The operator (collision.cu:17-118, BC.cu:20-125), with u~ = u/cs, cs^2 = 1/3:
//   Fbar_i = (c_i.F - u.F)/(2 cs^2)
//   feq_i  = w_i rho [1 + H1.u~ + (H2:u~u~ + H3:u~u~u~)/2]     (3rd order)
//   A2     = sum_i H2_i (f_i + Fbar_i feq_i) - rho u~u~
//   A3     = recursive: A_aab = A_aa u~_b + 2 u~_a A_ab,  A_xyz = A_xy u~_z + A_xz u~_y + A_yz u~_x
//   fneq_i = w_i (H2_i:A2 + H3_i:A3)/2
//   collision: f*_i = feq_i + (1 - 1/tau) fneq_i + Fbar_i feq_i
//   RRBC:      f_i  = feq_i + fneq_i - Fbar_i feq_i
// Populations are handled shifted, h_i = f_i - w_i (rho0 = 1).
#pragma once
#include "lattice.hpp"

namespace rrlbm {

template <typename C>
struct Moments {
  C dr;              // sum h_i = rho - 1
  C jx, jy, jz;      // sum c_i h_i
  C pxx, pyy, pzz;   // sum c_a c_a h_i
  C pxy, pxz, pyz;   // sum c_a c_b h_i
};

// Raw moments of the 27 shifted populations (channel ordering) by three
// tensor-product stages over x, y, z.
template <typename C>
RRLBM_HD RRLBM_INLINE Moments<C> moments_q27(const C (&h)[27]) {
  // stage x: for each (cy, cz), sums over cx of 1, cx, cx^2
  C X0[3][3], X1[3][3], X2[3][3];
  static_for<3>([&](auto jy) {
    static_for<3>([&](auto jz) {
      constexpr int cy = decltype(jy)::value - 1, cz = decltype(jz)::value - 1;
      const C hm = h[D3Q27::index(-1, cy, cz)];
      const C h0 = h[D3Q27::index(0, cy, cz)];
      const C hp = h[D3Q27::index(1, cy, cz)];
      const C s = hp + hm;
      X0[cy + 1][cz + 1] = s + h0;
      X1[cy + 1][cz + 1] = hp - hm;
      X2[cy + 1][cz + 1] = s;
    });
  });
  // stage y: (a,b) in {(0,0),(0,1),(0,2),(1,0),(1,1),(2,0)}
  C Y00[3], Y01[3], Y02[3], Y10[3], Y11[3], Y20[3];
  static_for<3>([&](auto kz) {
    constexpr int z = decltype(kz)::value;
    const C s0 = X0[2][z] + X0[0][z];
    Y02[z] = s0;
    Y00[z] = s0 + X0[1][z];
    Y01[z] = X0[2][z] - X0[0][z];
    Y10[z] = (X1[2][z] + X1[0][z]) + X1[1][z];
    Y11[z] = X1[2][z] - X1[0][z];
    Y20[z] = (X2[2][z] + X2[0][z]) + X2[1][z];
  });
  // stage z
  Moments<C> m;
  const C s00 = Y00[2] + Y00[0];
  m.dr = s00 + Y00[1];
  m.jz = Y00[2] - Y00[0];
  m.pzz = s00;
  m.jx = (Y10[2] + Y10[0]) + Y10[1];
  m.pxz = Y10[2] - Y10[0];
  m.jy = (Y01[2] + Y01[0]) + Y01[1];
  m.pyz = Y01[2] - Y01[0];
  m.pxx = (Y20[2] + Y20[0]) + Y20[1];
  m.pyy = (Y02[2] + Y02[0]) + Y02[1];
  m.pxy = (Y11[2] + Y11[0]) + Y11[1];
  return m;
}

// Raw moments from weight-scaled values g_i = h_i * S / w_i (the FP16
// storage form; T is the stored type).  The D3Q27 weights are products of the
// 1-D weights (1/6, 2/3, 1/6) = (1, 4, 1) / 6, so h_i = g_i k(cx) k(cy) k(cz)
// / (216 S) with k(0) = 4, k(+-1) = 1: the three stages of moments_q27 run on
// g with the zero-velocity term of each line sum taken four times, and the ten
// moments take unscale216 = 1 / (216 S) at the end.  `one` is 1; the device
// code passes it opaque so that fma(g+, one, g-) stays one mixed-precision FMA
// reading both stored values.
template <typename C, typename T>
RRLBM_HD RRLBM_INLINE Moments<C> moments_q27_wscaled(const T (&g)[27], C one, C unscale216) {
  C X0[3][3], X1[3][3], X2[3][3];
  static_for<3>([&](auto jy) {
    static_for<3>([&](auto jz) {
      constexpr int cy = decltype(jy)::value - 1, cz = decltype(jz)::value - 1;
      const T gm = g[D3Q27::index(-1, cy, cz)];
      const T g0 = g[D3Q27::index(0, cy, cz)];
      const T gp = g[D3Q27::index(1, cy, cz)];
      const C s = fma(C(gp), one, C(gm));
      X2[cy + 1][cz + 1] = s;
      X1[cy + 1][cz + 1] = fma(C(gm), C(-2), s);  // g+ - g-
      X0[cy + 1][cz + 1] = fma(C(g0), C(4), s);
    });
  });
  C Y00[3], Y01[3], Y02[3], Y10[3], Y11[3], Y20[3];
  static_for<3>([&](auto kz) {
    constexpr int z = decltype(kz)::value;
    const C s0 = X0[2][z] + X0[0][z];
    Y02[z] = s0;
    Y00[z] = fma(X0[1][z], C(4), s0);
    Y01[z] = X0[2][z] - X0[0][z];
    Y10[z] = fma(X1[1][z], C(4), X1[2][z] + X1[0][z]);
    Y11[z] = X1[2][z] - X1[0][z];
    Y20[z] = fma(X2[1][z], C(4), X2[2][z] + X2[0][z]);
  });
  auto k3 = [](const C (&v)[3]) { return fma(v[1], C(4), v[2] + v[0]); };
  Moments<C> m;
  const C s00 = Y00[2] + Y00[0];
  m.dr = fma(Y00[1], C(4), s00) * unscale216;
  m.jz = (Y00[2] - Y00[0]) * unscale216;
  m.pzz = s00 * unscale216;
  m.jx = k3(Y10) * unscale216;
  m.pxz = (Y10[2] - Y10[0]) * unscale216;
  m.jy = k3(Y01) * unscale216;
  m.pyz = (Y01[2] - Y01[0]) * unscale216;
  m.pxx = k3(Y20) * unscale216;
  m.pyy = k3(Y02) * unscale216;
  m.pxy = k3(Y11) * unscale216;
  return m;
}

// Raw moments for any lattice, direction by direction.
template <class L, typename C>
RRLBM_HD RRLBM_INLINE Moments<C> moments_generic(const C (&h)[L::Q]) {
  Moments<C> m;
  m.dr = h[0];
  m.jx = m.jy = m.jz = m.pxx = m.pyy = m.pzz = m.pxy = m.pxz = m.pyz = C(-0.0);
  static_for<L::Q - 1>([&](auto ic) {
    constexpr int i = decltype(ic)::value + 1;
    constexpr int x = L::cx(i), y = L::cy(i), z = L::cz(i);
    const C v = h[i];
    m.dr += v;
    if constexpr (x > 0) m.jx += v; else if constexpr (x < 0) m.jx -= v;
    if constexpr (y > 0) m.jy += v; else if constexpr (y < 0) m.jy -= v;
    if constexpr (z > 0) m.jz += v; else if constexpr (z < 0) m.jz -= v;
    if constexpr (x != 0) m.pxx += v;
    if constexpr (y != 0) m.pyy += v;
    if constexpr (z != 0) m.pzz += v;
    if constexpr (x * y > 0) m.pxy += v; else if constexpr (x * y < 0) m.pxy -= v;
    if constexpr (x * z > 0) m.pxz += v; else if constexpr (x * z < 0) m.pxz -= v;
    if constexpr (y * z > 0) m.pyz += v; else if constexpr (y * z < 0) m.pyz -= v;
  });
  return m;
}

template <typename C>
struct Macro {
  C rho, dr, ux, uy, uz;
};

// Which body-force components a kernel variant carries: the channel's force is
// along x only (macrovar.cu:70-72), Taylor-Green and cavity have none.  Terms
// of absent components are removed at compile time.
enum ForceMode : int { kForceNone = 0, kForceX = 1, kForceAll = 2 };

// 1/x: correctly rounded in FP64; the hardware reciprocal (1 ulp) in FP32.
RRLBM_HD RRLBM_INLINE double recip(double x) { return 1.0 / x; }
RRLBM_HD RRLBM_INLINE float recip(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}

// macrovar (RR27V33_code_DNS_LES/macrovar.cu:17-29): rho = sum f, u = (sum c f + F/2)/rho
template <typename C, int FM = kForceAll>
RRLBM_HD RRLBM_INLINE Macro<C> macro_of(const Moments<C>& m, C Fx, C Fy, C Fz) {
  Macro<C> r;
  r.dr = m.dr;
  r.rho = C(1) + m.dr;
  const C inv = recip(r.rho);
  if constexpr (FM == kForceNone) {
    r.ux = m.jx * inv;
    r.uy = m.jy * inv;
    r.uz = m.jz * inv;
  } else if constexpr (FM == kForceX) {
    r.ux = (m.jx + C(0.5) * Fx) * inv;
    r.uy = m.jy * inv;
    r.uz = m.jz * inv;
  } else {
    r.ux = (m.jx + C(0.5) * Fx) * inv;
    r.uy = (m.jy + C(0.5) * Fy) * inv;
    r.uz = (m.jz + C(0.5) * Fz) * inv;
  }
  return r;
}

// Non-equilibrium second moment N = A2/3 in physical units (A2 of
// collision.cu:72-86), using the D3Q27 closed form of the force projection
//   sum_i c_a c_b Fbar_i feq_i = S_ab,  S_aa = rho F_a u_a (1 - 3/2 u_a^2),
//   S_ab = rho (F_a u_b + F_b u_a)/2,
// which follows from the third lattice moments of the RR equilibrium on D3Q27
// (sum feq c_a^2 c_b = rho(u_a^2 u_b + u_b/3), sum feq c_x c_y c_z = rho u_x u_y u_z,
// sum feq c_a^3 = rho u_a).
template <typename C>
struct NeqTensor {
  C xx, yy, zz, xy, xz, yz;
};

template <typename C, int FM = kForceAll>
RRLBM_HD RRLBM_INLINE NeqTensor<C> neq_q27(const Moments<C>& m, const Macro<C>& q, C Fx, C Fy, C Fz) {
  const C rho = q.rho, ux = q.ux, uy = q.uy, uz = q.uz;
  // sum_i H_aa f_i = 3 P_aa - (sum_i f_i - 1): the populations' own density
  // excess m.dr (equal to q.dr in the fused step; the drop-in collision takes
  // rho from the macrovar field, as collision.cu:20 does).
  const C third = C(1.0 / 3.0) * m.dr;
  const C rx = rho * ux, ry = rho * uy, rz = rho * uz;
  NeqTensor<C> n;
  n.xx = m.pxx - rx * ux - third;
  n.yy = m.pyy - ry * uy - third;
  n.zz = m.pzz - rz * uz - third;
  n.xy = m.pxy - rx * uy;
  n.xz = m.pxz - rx * uz;
  n.yz = m.pyz - ry * uz;
  if constexpr (FM != kForceNone) {
    n.xx += rx * Fx * (C(1) - C(1.5) * ux * ux);
    n.xy += C(0.5) * ry * Fx;
    n.xz += C(0.5) * rz * Fx;
  }
  if constexpr (FM == kForceAll) {
    n.yy += ry * Fy * (C(1) - C(1.5) * uy * uy);
    n.zz += rz * Fz * (C(1) - C(1.5) * uz * uz);
    n.xy += C(0.5) * rx * Fy;
    n.xz += C(0.5) * rx * Fz;
    n.yz += C(0.5) * (ry * Fz + rz * Fy);
  }
  return n;
}

// Values of the polynomial k0 + k1 c + k2 c^2 at c = -1, 0, +1.
template <typename C>
struct Line3 {
  C m, z, p;
};
template <typename C>
RRLBM_HD RRLBM_INLINE Line3<C> line3(C k0, C k1, C k2) {
  const C t = k0 + k2;
  return {t - k1, k0, t + k1};
}
template <typename C>
RRLBM_HD RRLBM_INLINE Line3<C> line3_nok2(C k0, C k1) {
  return {k0 - k1, k0, k0 + k1};
}

// Post-collision (or RRBC) state of one D3Q27 cell.  Writes
//   out_i = scale_i * (g_i - w_i)/w_i,   g = feq (1 + gs Fbar) + omega fneq,
// with scale_i = w_i when per_direction_w is true (giving g_i - w_i, the shifted
// population) and scale_i = scale otherwise (FP16 storage: (g_i - w_i) S / w_i).
// Collision: omega = 1 - 1/tau_eff, gs = +1.  RRBC: omega = 1, gs = -1.
//
// With per_direction_w the weights are folded into the polynomial instead of
// 27 products at the end.  D3Q27's weights are products of the 1-D weights
// (1, 4, 1)/6, so w_i = k(c_x) k(c_y) k(c_z) / 216 with k(0) = 4, k(+-1) = 1.
// The coefficient of c_x^a c_y^b c_z^c is scaled by 4^(number of zero
// exponents) / 216 -- the weight of a velocity with as many non-zero
// components as (a, b, c) has non-zero exponents -- so that in each stage's
// lines (k0, k1, k2) the constant term k0 carries a factor 4 against k1 and
// k2, and the line's values at c = -1, 0, +1 are (k0/4 + k2 - k1, k0,
// k0/4 + k2 + k1): each stage multiplies by k(c) at no cost, and the last one
// leaves k(c_x) k(c_y) k(c_z) / 216 times the polynomial.  The powers of 4 go
// into the constants (kF below; factors between terms are powers of 4, exact)
// and the common 1/216 into rho, rho - 1 and omega, which scale every term:
// all terms of a cell then carry the same rounding of that factor, and the
// terms that cancel in the mass and momentum sums still cancel.  Without the
// weights the output scale goes in the same way.
template <typename C, bool per_direction_w, int FM = kForceAll>
RRLBM_HD RRLBM_INLINE void rr_post_q27(const Macro<C>& q, const NeqTensor<C>& N, C Fx, C Fy, C Fz, C omega,
                                       C gs, C scale, C (&out)[27]) {
  // scale of a coefficient with n non-zero exponents, without the common
  // 1/216 (all 1 without the weights)
  constexpr double kF[4] = {per_direction_w ? 64.0 : 1.0, per_direction_w ? 16.0 : 1.0,
                            per_direction_w ? 4.0 : 1.0, 1.0};
  constexpr double kQ = per_direction_w ? 0.25 : 1.0;  // kF[n + 1] / kF[n]
  // the common factor: 1/216 with the weights, the output scale without
  C sc;
  if constexpr (per_direction_w)
    sc = C(1.0 / 216.0);
  else
    sc = scale;
  const C ux = q.ux, uy = q.uy, uz = q.uz;
  const C rho = q.rho * sc, dr = q.dr * sc, wn = omega * sc;
  // feq / w as a polynomial in c: e_abc multiplies c_x^a c_y^b c_z^c
  const C ux2 = ux * ux, uy2 = uy * uy, uz2 = uz * uz;
  const C rx = rho * ux, ry = rho * uy, rz = rho * uz;
  const C e000m1 = C(kF[0]) * dr - C(1.5 * kF[0]) * rho * (ux2 + uy2 + uz2);
  const C e100 = rx * (C(3 * kF[1]) - C(4.5 * kF[1]) * (uy2 + uz2));
  const C e010 = ry * (C(3 * kF[1]) - C(4.5 * kF[1]) * (ux2 + uz2));
  const C e001 = rz * (C(3 * kF[1]) - C(4.5 * kF[1]) * (ux2 + uy2));
  const C e200 = C(4.5 * kF[1]) * rx * ux, e020 = C(4.5 * kF[1]) * ry * uy, e002 = C(4.5 * kF[1]) * rz * uz;
  const C e110 = C(9 * kF[2]) * rx * uy, e101 = C(9 * kF[2]) * rx * uz, e011 = C(9 * kF[2]) * ry * uz;
  const C e210 = C(3 * kQ) * e200 * uy, e201 = C(3 * kQ) * e200 * uz;
  const C e120 = C(3 * kQ) * e020 * ux, e021 = C(3 * kQ) * e020 * uz;
  const C e102 = C(3 * kQ) * e002 * ux, e012 = C(3 * kQ) * e002 * uy;
  const C e111 = C(3 * kQ) * e110 * uz;
  // fneq / w: same structure with N and the recursive third-order tensor R
  const C Rxy = N.xx * uy + C(2) * ux * N.xy, Rxz = N.xx * uz + C(2) * ux * N.xz;
  const C Ryx = N.yy * ux + C(2) * uy * N.xy, Ryz = N.yy * uz + C(2) * uy * N.yz;
  const C Rzx = N.zz * ux + C(2) * uz * N.xz, Rzy = N.zz * uy + C(2) * uz * N.yz;
  const C Rxyz = N.xy * uz + N.xz * uy + N.yz * ux;
  // Coefficients K of feq (1 + gs Fbar) + omega fneq, Fbar_i = 3/2 (c_i - u).F:
  // K = b0 e + [phi.c e reduced with c^3 = c] + omega n, b0 = 1 - 3/2 gs u.F,
  // phi = 3/2 gs F.  Products with absent force components are not formed.
  // omega k x as (omega k) x: omega is uniform in the kernels with a constant
  // relaxation time, so omega k is formed once per launch where a workgroup
  // loops over cells
  auto wt = [&](double k, C x) -> C { return (wn * C(k)) * x; };
  C K000 = e000m1 + wt(-1.5 * kF[0], N.xx + N.yy + N.zz);
  C K100 = e100 + wt(-4.5 * kF[1], Ryx + Rzx);
  C K010 = e010 + wt(-4.5 * kF[1], Rxy + Rzy);
  C K001 = e001 + wt(-4.5 * kF[1], Rxz + Ryz);
  C K200 = e200 + wt(4.5 * kF[1], N.xx);
  C K020 = e020 + wt(4.5 * kF[1], N.yy);
  C K002 = e002 + wt(4.5 * kF[1], N.zz);
  C K110 = e110 + wt(9 * kF[2], N.xy);
  C K101 = e101 + wt(9 * kF[2], N.xz);
  C K011 = e011 + wt(9 * kF[2], N.yz);
  C K210 = e210 + wt(13.5 * kF[2], Rxy);
  C K201 = e201 + wt(13.5 * kF[2], Rxz);
  C K120 = e120 + wt(13.5 * kF[2], Ryx);
  C K021 = e021 + wt(13.5 * kF[2], Ryz);
  C K102 = e102 + wt(13.5 * kF[2], Rzx);
  C K012 = e012 + wt(13.5 * kF[2], Rzy);
  C K111 = e111 + wt(27 * kF[3], Rxyz);
  C K220, K202, K022, K211, K121, K112;
  if constexpr (FM == kForceNone) {
    K220 = K202 = K022 = K211 = K121 = K112 = C(0);
  } else {
    const C k15 = C(1.5) * gs;
    const C phx = k15 * Fx;
    const C e000 = C(kF[0]) * sc + e000m1;
    C bm1 = -(phx * ux);
    if constexpr (FM == kForceAll) bm1 -= k15 * (Fy * uy + Fz * uz);
    // b0 e - e = bm1 e
    K000 += bm1 * (C(kF[0]) * sc) + bm1 * e000m1;
    K100 += bm1 * e100 + phx * (C(kQ) * e000 + e200);
    K010 += bm1 * e010;
    K001 += bm1 * e001;
    K200 += bm1 * e200 + phx * e100;
    K020 += bm1 * e020;
    K002 += bm1 * e002;
    K110 += bm1 * e110 + phx * (C(kQ) * e010 + e210);
    K101 += bm1 * e101 + phx * (C(kQ) * e001 + e201);
    K011 += bm1 * e011;
    K210 += bm1 * e210 + phx * e110;
    K201 += bm1 * e201 + phx * e101;
    K120 += bm1 * e120 + phx * C(kQ) * e020;
    K021 += bm1 * e021;
    K102 += bm1 * e102 + phx * C(kQ) * e002;
    K012 += bm1 * e012;
    K111 += bm1 * e111 + phx * C(kQ) * e011;
    K220 = phx * e120;
    K202 = phx * e102;
    K022 = C(0);
    K211 = phx * e111;
    K121 = phx * C(kQ) * e021;
    K112 = phx * C(kQ) * e012;
    if constexpr (FM == kForceAll) {
      const C phy = k15 * Fy, phz = k15 * Fz;
      K010 += phy * (C(kQ) * e000 + e020);
      K001 += phz * (C(kQ) * e000 + e002);
      K020 += phy * e010;
      K002 += phz * e001;
      K110 += phy * (C(kQ) * e100 + e120);
      K101 += phz * (C(kQ) * e100 + e102);
      K011 += phy * (C(kQ) * e001 + e021) + phz * (C(kQ) * e010 + e012);
      K210 += phy * C(kQ) * e200;
      K201 += phz * C(kQ) * e200;
      K120 += phy * e110;
      K021 += phy * e011 + phz * C(kQ) * e020;
      K102 += phz * e101;
      K012 += phy * C(kQ) * e002 + phz * e011;
      K111 += phy * C(kQ) * e101 + phz * C(kQ) * e110;
      K220 += phy * e210;
      K202 += phz * e201;
      K022 = phy * e012 + phz * e021;
      K211 += phy * C(kQ) * e201 + phz * C(kQ) * e210;
      K121 += phy * e111 + phz * C(kQ) * e120;
      K112 += phy * C(kQ) * e102 + phz * e111;
    }
  }
  // tensor-product evaluation: x, then y, then z
  constexpr bool kq = FM != kForceNone;   // K220, K202, K211, K121, K112 present
  constexpr bool k22 = FM == kForceAll;   // K022 present
  auto ln = [](C k0, C k1, C k2) -> Line3<C> {
    if constexpr (per_direction_w) {
      const C t = fma(k0, C(0.25), k2);
      return {t - k1, k0, t + k1};
    } else {
      return line3(k0, k1, k2);
    }
  };
  auto ln_nok2 = [](C k0, C k1) -> Line3<C> {
    if constexpr (per_direction_w)
      return {C(fma(k0, C(0.25), -k1)), k0, C(fma(k0, C(0.25), k1))};
    else
      return line3_nok2(k0, k1);
  };
  const Line3<C> L00 = ln(K000, K100, K200);
  const Line3<C> L10 = ln(K010, K110, K210);
  const Line3<C> L01 = ln(K001, K101, K201);
  const Line3<C> L20 = kq ? ln(K020, K120, K220) : ln_nok2(K020, K120);
  const Line3<C> L02 = kq ? ln(K002, K102, K202) : ln_nok2(K002, K102);
  const Line3<C> L11 = kq ? ln(K011, K111, K211) : ln_nok2(K011, K111);
  const Line3<C> L21 = ln_nok2(K021, K121);
  const Line3<C> L12 = ln_nok2(K012, K112);
  // the line of K022 alone: K022 at c_x = 0, K022 / 4 at c_x = +-1 (weighted)
  const C L22 = K022, L22q = C(kQ) * K022;
  auto pick = [](const Line3<C>& l, int c) -> C { return c < 0 ? l.m : (c == 0 ? l.z : l.p); };
  static_for<3>([&](auto jx) {
    constexpr int cx = decltype(jx)::value - 1;
    const Line3<C> M0 = ln(pick(L00, cx), pick(L10, cx), pick(L20, cx));
    const Line3<C> M1 = ln(pick(L01, cx), pick(L11, cx), pick(L21, cx));
    const Line3<C> M2 = k22 ? ln(pick(L02, cx), pick(L12, cx), cx == 0 ? L22 : L22q)
                            : ln_nok2(pick(L02, cx), pick(L12, cx));
    static_for<3>([&](auto jy) {
      constexpr int cy = decltype(jy)::value - 1;
      const Line3<C> V = ln(pick(M0, cy), pick(M1, cy), pick(M2, cy));
      static_for<3>([&](auto jz) {
        constexpr int cz = decltype(jz)::value - 1;
        constexpr int i = D3Q27::index(cx, cy, cz);
        out[i] = pick(V, cz);
      });
    });
  });
}

// The same state with every Hermite order D3Q27 carries: the fourth- to
// sixth-order terms of the RR expansion, which the reference keeps commented
// out (collision.cu:48-49,62-67,95-104,112-116; BC.cu:37-38,50-54,79-88,97-101).
// Outputs as rr_post_q27.  With these orders the equilibrium is a product over
// the axes, G = prod_a P_a(c_a), P(c) = 1 + 3uc + 9/2 u^2 c^2 - 3/2 u^2, and the
// recursive non-equilibrium part is (1/2) sum_ab A_ab d2G/du~_a du~_b:
//   feq / w  = rho G
//   fneq / w = 3/2 sum_a N_aa h_a prod_{b!=a} P_b + 9 sum_{a<b} N_ab D_a D_b P_c,
//              h = 3c^2 - 1, D = c + h u
//   force:     Fbar times the equilibrium truncated after fourth order
//              (forcefeq of collision.cu:67), Q = G - T2 T2 T1 (3 ways) - T2 T2 T2,
//              T1 = 3uc, T2 = 3/2 u^2 h (the per-axis first / second-order parts of P).
// Everything is evaluated at c in {-1, 0, 1} per axis, then combined over
// (x, y) pairs and finally z; the shift by w is taken analytically
// (out / w - 1 = dr + rho (G - 1) + ...) with G - 1 formed from the small
// per-axis parts p = P - 1, so FP32 keeps the relative precision of h.
template <typename C, bool per_direction_w, int FM = kForceAll>
RRLBM_HD RRLBM_INLINE void rr_post_q27_o6(const Macro<C>& q, const NeqTensor<C>& N, C Fx, C Fy, C Fz, C omega,
                                          C gs, C scale, C (&out)[27]) {
  struct Axis {
    C p[3], P[3], T1[3], T2[3], D[3], g[3];
  };
  auto axis = [](C u, C F) {
    Axis a;
    const C u2 = u * u, t = C(3) * u, s = C(3) * u2;
    a.p[0] = s - t; a.p[1] = C(-1.5) * u2; a.p[2] = s + t;
    a.T1[0] = -t; a.T1[1] = C(0); a.T1[2] = t;
    a.T2[0] = s; a.T2[1] = C(-1.5) * u2; a.T2[2] = s;
    a.D[0] = C(2) * u - C(1); a.D[1] = -u; a.D[2] = C(2) * u + C(1);
    const C f15 = C(1.5) * F;
    a.g[0] = -f15 * (C(1) + u); a.g[1] = -f15 * u; a.g[2] = f15 * (C(1) - u);
    for (int j = 0; j < 3; ++j) a.P[j] = C(1) + a.p[j];
    return a;
  };
  constexpr bool fx = FM != kForceNone, fyz = FM == kForceAll;
  const Axis X = axis(q.ux, fx ? Fx : C(0)), Y = axis(q.uy, fyz ? Fy : C(0)), Z = axis(q.uz, fyz ? Fz : C(0));
  // h(0) = -1, h(+-1) = 2
  auto hval = [](int j) { return j == 1 ? C(-1) : C(2); };
  // The value is linear in (dr, rho, N, gs rho): a uniform output scale (FP16
  // storage) multiplies those once instead of each of the 27 results.
  const C sc = per_direction_w ? C(1) : scale;
  const C on = omega * sc;
  const C nxx = C(1.5) * on * N.xx, nyy = C(1.5) * on * N.yy, nzz = C(1.5) * on * N.zz;
  const C nxy = C(9) * on * N.xy, nxz = C(9) * on * N.xz, nyz = C(9) * on * N.yz;
  const C dr_s = q.dr * sc, rho_s = q.rho * sc;
  const C grho = gs * rho_s;
  static_for<3>([&](auto jxc) {
    constexpr int jx = decltype(jxc)::value;
    const C ax = nxx * hval(jx), dx = nxy * X.D[jx], ex = nxz * X.D[jx];
    // force along x only: the y and z factors g vanish (kept out rather than
    // added as zeros, which the compiler may not drop)
    const C gxr = grho * X.g[jx];
    static_for<3>([&](auto jyc) {
      constexpr int jy = decltype(jyc)::value;
      const C Exy = X.p[jx] + Y.p[jy] + X.p[jx] * Y.p[jy];
      const C PP = C(1) + Exy;
      const C Ap = ax * Y.P[jy] + (nyy * hval(jy)) * X.P[jx] + dx * Y.D[jy];
      const C Bp = nzz * PP;
      const C Cp = ex * Y.P[jy] + (nyz * Y.D[jy]) * X.P[jx];
      C TT = C(0), TU = C(0), Sxy = C(0);
      if constexpr (fx) {
        TT = X.T2[jx] * Y.T2[jy];
        TU = X.T2[jx] * Y.T1[jy] + X.T1[jx] * Y.T2[jy];
        if constexpr (fyz) Sxy = X.g[jx] + Y.g[jy];
      }
      if constexpr (!fyz) {
        // Force along x or none: with P_z = 1 + p_z the column over c_z is
        //   val = al + be p_z + Cp D_z + ga T2_z + Bp h_z,
        // whose z factors take at c_z = +-1 the values p = s +- t, D = 2u +- 1,
        // T2 = s, h = 2 and at c_z = 0 p = T2 = -3/2 u^2, D = -u, h = -1
        // (s = 3u^2, t = 3u): the column is an even part plus or minus an odd one.
        C al = dr_s + rho_s * Exy + Ap;
        C be = rho_s * PP + Ap;
        C ga = C(0);
        if constexpr (fx) {
          al += gxr * PP;
          be += gxr * (PP - TT);
          ga = -gxr * TU;
        }
        const C bg = be + ga;
        const C uz = q.uz;
        const C v0 = al + bg * Z.T2[1] - uz * Cp - Bp;
        const C ev = al + bg * Z.T2[0] + (C(2) * uz) * Cp + C(2) * Bp;
        const C od = be * Z.T1[2] + Cp;
        const C col[3] = {ev - od, v0, ev + od};
        static_for<3>([&](auto jzc) {
          constexpr int jz = decltype(jzc)::value;
          constexpr int i = D3Q27::index(jx - 1, jy - 1, jz - 1);
          if constexpr (per_direction_w)
            out[i] = col[jz] * C(D3Q27::w(i));
          else
            out[i] = col[jz];
        });
      } else {
        static_for<3>([&](auto jzc) {
          constexpr int jz = decltype(jzc)::value;
          constexpr int i = D3Q27::index(jx - 1, jy - 1, jz - 1);
          const C E = Exy + Z.p[jz] * PP;
          C val = dr_s + rho_s * E + Ap * Z.P[jz] + Bp * hval(jz) + Cp * Z.D[jz];
          if constexpr (fx) {
            const C Qv = PP * Z.P[jz] - TT * Z.p[jz] - TU * Z.T2[jz];
            if constexpr (fyz)
              val += grho * (Sxy + Z.g[jz]) * Qv;
            else
              val += gxr * Qv;
          }
          if constexpr (per_direction_w)
            out[i] = val * C(D3Q27::w(i));
          else
            out[i] = val;
        });
      }
    });
  });
}

// D3Q19 without body force, in the same moment-space form.  On the 19
// velocities every monomial containing c_x c_y c_z vanishes, so the
// post-collision state is a polynomial in the remaining monomials.  The third
// order is Malaspinas' pairing (Taylor_Green/D3Q19RR.f90:131-136,230-235): for
// a linear axis b with quadratic axes a, a' the Hermite pair enters as
//   (4/3)(H_aab X_ab + H_a'a'b X_a'b) + (2/3)(H_aab X_a'b + H_a'a'b X_ab),
// so with E_ab = rho u_a^2 u_b + omega R_ab (R_ab = N_aa u_b + 2 u_a N_ab, the
// recursive third-order non-equilibrium tensor):
//   K[c_a^2 c_b] = 18 E_ab + 9 E_a'b,   K[c_b] = 3 rho u_b - 9 (E_ab + E_a'b)
// (on D3Q27 the same terms are 13.5 E_ab and -4.5 (E_ab + E_a'b)).  Outputs as
// rr_post_q27, in D3Q19 order; gs plays no role without a force.
template <typename C, bool per_direction_w>
RRLBM_HD RRLBM_INLINE void rr_post_q19(const Macro<C>& q, const NeqTensor<C>& N, C omega, C scale,
                                       C (&out)[19]) {
  const C rho = q.rho, ux = q.ux, uy = q.uy, uz = q.uz;
  const C rx = rho * ux, ry = rho * uy, rz = rho * uz;
  const C wn = omega;
  // E_ab: a quadratic axis, b linear axis
  const C Exy = rx * ux * uy + wn * (N.xx * uy + C(2) * ux * N.xy);
  const C Exz = rx * ux * uz + wn * (N.xx * uz + C(2) * ux * N.xz);
  const C Eyx = ry * uy * ux + wn * (N.yy * ux + C(2) * uy * N.xy);
  const C Eyz = ry * uy * uz + wn * (N.yy * uz + C(2) * uy * N.yz);
  const C Ezx = rz * uz * ux + wn * (N.zz * ux + C(2) * uz * N.xz);
  const C Ezy = rz * uz * uy + wn * (N.zz * uy + C(2) * uz * N.yz);
  const C K000 = q.dr - C(1.5) * (rx * ux + ry * uy + rz * uz) - C(1.5) * wn * (N.xx + N.yy + N.zz);
  const C K100 = C(3) * rx - C(9) * (Eyx + Ezx);
  const C K010 = C(3) * ry - C(9) * (Exy + Ezy);
  const C K001 = C(3) * rz - C(9) * (Exz + Eyz);
  const C K200 = C(4.5) * (rx * ux + wn * N.xx);
  const C K020 = C(4.5) * (ry * uy + wn * N.yy);
  const C K002 = C(4.5) * (rz * uz + wn * N.zz);
  const C K110 = C(9) * (rx * uy + wn * N.xy);
  const C K101 = C(9) * (rx * uz + wn * N.xz);
  const C K011 = C(9) * (ry * uz + wn * N.yz);
  const C K210 = C(18) * Exy + C(9) * Ezy, K012 = C(18) * Ezy + C(9) * Exy;
  const C K120 = C(18) * Eyx + C(9) * Ezx, K102 = C(18) * Ezx + C(9) * Eyx;
  const C K201 = C(18) * Exz + C(9) * Eyz, K021 = C(18) * Eyz + C(9) * Exz;
  // tensor-product evaluation (x, then y, then z) at the 19 velocities
  const Line3<C> L00 = line3(K000, K100, K200);
  const Line3<C> L10 = line3(K010, K110, K210);
  const Line3<C> L01 = line3(K001, K101, K201);
  const Line3<C> L20 = line3_nok2(K020, K120);
  const Line3<C> L02 = line3_nok2(K002, K102);
  auto pick = [](const Line3<C>& l, int c) -> C { return c < 0 ? l.m : (c == 0 ? l.z : l.p); };
  static_for<19>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    constexpr int cx = D3Q19::cx(i), cy = D3Q19::cy(i), cz = D3Q19::cz(i);
    C val;
    if constexpr (cz == 0) {
      // (cx, cy, 0): L00 + L10 cy + L20 cy^2
      val = pick(L00, cx);
      if constexpr (cy != 0) val += C(cy) * pick(L10, cx) + pick(L20, cx);
    } else if constexpr (cy == 0) {
      // (cx, 0, cz): L00 + L01 cz + L02 cz^2
      val = pick(L00, cx) + C(cz) * pick(L01, cx) + pick(L02, cx);
    } else {
      // (0, cy, cz): all terms with c_x = 0
      val = K000 + C(cy) * K010 + K020 + C(cz) * (K001 + K021) + K002 + C(cy * cz) * K011 + C(cy) * K012;
    }
    if constexpr (per_direction_w)
      out[i] = val * C(D3Q19::w(i));
    else
      out[i] = val * scale;
  });
}

// Per-direction formulation, following collision.cu:41-118 (D3Q27) and the
// D3Q19 third-order pairs of D3Q19RR.f90:131-136,221-235,263-281.  h are the
// shifted populations of the cell; writes out_i = g_i - w_i with
// g = feq + omega fneq + gs Fbar feq.
template <class L, typename C>
RRLBM_HD RRLBM_INLINE C third_order(int /*unused*/, C a210, C a201, C a021, C a120, C a102, C a012, C a111,
                                    double H210, double H201, double H021, double H120, double H102, double H012,
                                    double H111) {
  if constexpr (L::Q == 27) {
    return C(H210) * a210 + C(H201) * a201 + C(H021) * a021 + C(H120) * a120 + C(H012) * a012 + C(H102) * a102 +
           C(2.0 * H111) * a111;
  } else {
    (void)a111; (void)H111;
    return C(H210 + H012) * (a210 + a012) + C(H102 + H120) * (a102 + a120) + C(H021 + H201) * (a021 + a201) +
           C((H210 - H012) / 3.0) * (a210 - a012) + C((H102 - H120) / 3.0) * (a102 - a120) +
           C((H021 - H201) / 3.0) * (a021 - a201);
  }
}

template <class L, typename C>
RRLBM_HD RRLBM_INLINE void rr_post_generic(const C (&h)[L::Q], const Macro<C>& q, C Fx, C Fy, C Fz, C omega, C gs,
                                           C (&out)[L::Q]) {
  constexpr double isq3 = 1.7320508075688772935;  // 1/cs
  const C rho = q.rho;
  const C uu = q.ux * C(isq3), vv = q.uy * C(isq3), ww = q.uz * C(isq3);
  const C a110 = uu * vv, a101 = uu * ww, a011 = vv * ww, a200 = uu * uu, a020 = vv * vv, a002 = ww * ww;
  const C a210 = a200 * vv, a201 = a200 * ww, a120 = a020 * uu, a021 = a020 * ww, a012 = a002 * vv,
          a102 = a002 * uu, a111 = a110 * ww;
  const C G3 = q.ux * Fx + q.uy * Fy + q.uz * Fz;
  C feqm1[L::Q], G[L::Q];
  C n110 = C(-0.0), n101 = C(-0.0), n011 = C(-0.0), n200 = C(-0.0), n020 = C(-0.0), n002 = C(-0.0);
  static_for<L::Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    constexpr int x = L::cx(i), y = L::cy(i), z = L::cz(i);
    constexpr double w = L::w(i);
    constexpr double H100 = Hermite::h1(x), H010 = Hermite::h1(y), H001 = Hermite::h1(z);
    constexpr double H110 = Hermite::h2(x, y, false), H101 = Hermite::h2(x, z, false), H011 = Hermite::h2(y, z, false);
    constexpr double H200 = Hermite::h2(x, x, true), H020 = Hermite::h2(y, y, true), H002 = Hermite::h2(z, z, true);
    constexpr double H210 = Hermite::h3(x, y), H201 = Hermite::h3(x, z), H021 = Hermite::h3(y, z);
    constexpr double H120 = Hermite::h3(y, x), H102 = Hermite::h3(z, x), H012 = Hermite::h3(z, y);
    constexpr double H111 = Hermite::hxyz(x, y, z);
    const C Fbar = C(1.5) * (C(x) * Fx + C(y) * Fy + C(z) * Fz - G3);
    const C fir = C(H100) * uu + C(H010) * vv + C(H001) * ww;
    const C sec = C(2.0 * H110) * a110 + C(2.0 * H101) * a101 + C(2.0 * H011) * a011 + C(H200) * a200 +
                  C(H020) * a020 + C(H002) * a002;
    const C thir = third_order<L, C>(0, a210, a201, a021, a120, a102, a012, a111, H210, H201, H021, H120, H102,
                                     H012, H111);
    const C fm1 = C(w) * (q.dr + rho * (fir + C(0.5) * (sec + thir)));  // feq - w
    const C g = Fbar * (C(w) + fm1);                                  // Fbar feq
    feqm1[i] = fm1;
    G[i] = g;
    const C proj = h[i] + g;  // (f + Fbar feq) - w; sum_i H2_i w_i = 0
    if constexpr (H110 != 0.0) n110 += C(H110) * proj;
    if constexpr (H101 != 0.0) n101 += C(H101) * proj;
    if constexpr (H011 != 0.0) n011 += C(H011) * proj;
    n200 += C(H200) * proj;
    n020 += C(H020) * proj;
    n002 += C(H002) * proj;
  });
  // collision.cu:81-94
  n110 -= rho * a110; n101 -= rho * a101; n011 -= rho * a011;
  n200 -= rho * a200; n020 -= rho * a020; n002 -= rho * a002;
  const C n210 = n200 * vv + C(2) * uu * n110;
  const C n201 = n200 * ww + C(2) * uu * n101;
  const C n120 = n020 * uu + C(2) * vv * n110;
  const C n021 = n020 * ww + C(2) * vv * n011;
  const C n102 = n002 * uu + C(2) * ww * n101;
  const C n012 = n002 * vv + C(2) * ww * n011;
  const C n111 = n110 * ww + n101 * vv + n011 * uu;
  static_for<L::Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    constexpr int x = L::cx(i), y = L::cy(i), z = L::cz(i);
    constexpr double w = L::w(i);
    constexpr double H110 = Hermite::h2(x, y, false), H101 = Hermite::h2(x, z, false), H011 = Hermite::h2(y, z, false);
    constexpr double H200 = Hermite::h2(x, x, true), H020 = Hermite::h2(y, y, true), H002 = Hermite::h2(z, z, true);
    constexpr double H210 = Hermite::h3(x, y), H201 = Hermite::h3(x, z), H021 = Hermite::h3(y, z);
    constexpr double H120 = Hermite::h3(y, x), H102 = Hermite::h3(z, x), H012 = Hermite::h3(z, y);
    constexpr double H111 = Hermite::hxyz(x, y, z);
    const C sec = C(2.0 * H110) * n110 + C(2.0 * H101) * n101 + C(2.0 * H011) * n011 + C(H200) * n200 +
                  C(H020) * n020 + C(H002) * n002;
    const C thir = third_order<L, C>(0, n210, n201, n021, n120, n102, n012, n111, H210, H201, H021, H120, H102,
                                     H012, H111);
    const C fneq = C(0.5 * w) * (sec + thir);
    out[i] = feqm1[i] + omega * fneq + gs * G[i];
  });
}

// BGK baseline operator (BGK/collision.cu:3-33), the incompressible form of
// the reference's BGK solver on shifted populations h = f - w, with the density
// deviation dr = sum h and u = j + F/2 (no division by rho, BGK/macrovar.cu:27):
//   feq_i - w_i = w_i (dr + 3 c.u + 9/2 (c.u)^2 - 3/2 u^2)
//   out_i = h_i - (h_i - (feq_i - w_i))/tau + (1 - 1/(2 tau)) 3 (c_i - u).F feq0_i,
//   feq0_i = w_i (1 + 3 c.u + 9/2 (c.u)^2 - 3/2 u^2),
// each written times scale_i (w_i / w_i = 1 for shifted values, S / w_i for
// the per-direction-scaled FP16 store).
template <class L, typename C, bool per_direction_w>
RRLBM_HD RRLBM_INLINE void bgk_post(const C (&h)[L::Q], C dr, C ux, C uy, C uz, C Fx, C Fy, C Fz, C inv_tau,
                                    C fcoef, C scale, C (&out)[L::Q]) {
  const C uu = C(1.5) * (ux * ux + uy * uy + uz * uz);
  const C uF = ux * Fx + uy * Fy + uz * Fz;
  static_for<L::Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    constexpr int x = L::cx(i), y = L::cy(i), z = L::cz(i);
    const C cu = C(x) * ux + C(y) * uy + C(z) * uz;
    const C poly = C(3) * cu + C(4.5) * cu * cu - uu;          // 3 c.u + 9/2 (c.u)^2 - 3/2 u^2
    const C cF = C(x) * Fx + C(y) * Fy + C(z) * Fz;
    const C g = C(L::w(i));
    const C feq = g * (dr + poly);                                // feq_i - w_i
    const C frc = fcoef * C(3) * (cF - uF) * g * (C(1) + poly);
    const C v = h[i] - (h[i] - feq) * inv_tau + frc;
    if constexpr (per_direction_w)
      out[i] = v;
    else
      out[i] = v * (scale * C(1.0 / L::w(i)));
  });
}

// The same BGK update on D3Q27 in moment space: out_i = (1 - 1/tau) h_i +
// w_i P(c_i), where P = (1/tau)(dr + Q - 1) + 3 fcoef (c - u).F Q with
// Q = 1 + 3 c.u + 9/2 (c.u)^2 - 3/2 u^2, is a polynomial in c of degree <= 2 per
// axis (c_a^3 = c_a on the lattice): its 17 coefficients are formed once and
// evaluated at the 27 velocities by the tensor-product stages of rr_post_q27.
// Outputs as bgk_post (h scaled by S / w_i when per_direction_w is false).
template <typename C, bool per_direction_w>
RRLBM_HD RRLBM_INLINE void bgk_post_q27(const C (&h)[27], C dr, C ux, C uy, C uz, C Fx, C Fy, C Fz, C inv_tau,
                                        C fcoef, C scale, C (&out)[27]) {
  const C q0 = C(1) - C(1.5) * (ux * ux + uy * uy + uz * uz);
  const C f3 = C(3) * fcoef;
  const C px = f3 * Fx, py = f3 * Fy, pz = f3 * Fz;   // phi = 3 fcoef F
  const C beta = -(px * ux + py * uy + pz * uz);       // -3 fcoef u.F
  const C t = inv_tau + beta;
  const C K000 = inv_tau * (dr + q0 - C(1)) + beta * q0;
  const C K100 = C(3) * ux * t + px * (q0 + C(4.5) * ux * ux);
  const C K010 = C(3) * uy * t + py * (q0 + C(4.5) * uy * uy);
  const C K001 = C(3) * uz * t + pz * (q0 + C(4.5) * uz * uz);
  const C K200 = C(4.5) * ux * ux * t + C(3) * px * ux;
  const C K020 = C(4.5) * uy * uy * t + C(3) * py * uy;
  const C K002 = C(4.5) * uz * uz * t + C(3) * pz * uz;
  const C K110 = C(9) * ux * uy * t + C(3) * (py * ux + px * uy);
  const C K101 = C(9) * ux * uz * t + C(3) * (pz * ux + px * uz);
  const C K011 = C(9) * uy * uz * t + C(3) * (pz * uy + py * uz);
  // c_a^2 c_b: 9/2 phi_b u_a^2 + 9 phi_a u_a u_b
  const C K210 = C(4.5) * py * ux * ux + C(9) * px * ux * uy;
  const C K201 = C(4.5) * pz * ux * ux + C(9) * px * ux * uz;
  const C K120 = C(4.5) * px * uy * uy + C(9) * py * uy * ux;
  const C K021 = C(4.5) * pz * uy * uy + C(9) * py * uy * uz;
  const C K102 = C(4.5) * px * uz * uz + C(9) * pz * uz * ux;
  const C K012 = C(4.5) * py * uz * uz + C(9) * pz * uz * uy;
  const C K111 = C(9) * (px * uy * uz + py * ux * uz + pz * ux * uy);
  const Line3<C> L00 = line3(K000, K100, K200);
  const Line3<C> L10 = line3(K010, K110, K210);
  const Line3<C> L01 = line3(K001, K101, K201);
  const Line3<C> L20 = line3_nok2(K020, K120);
  const Line3<C> L02 = line3_nok2(K002, K102);
  const Line3<C> L11 = line3_nok2(K011, K111);
  auto pick = [](const Line3<C>& l, int c) -> C { return c < 0 ? l.m : (c == 0 ? l.z : l.p); };
  const C keep = C(1) - inv_tau;
  static_for<3>([&](auto jx) {
    constexpr int cx = decltype(jx)::value - 1;
    const Line3<C> M0 = line3(pick(L00, cx), pick(L10, cx), pick(L20, cx));
    const Line3<C> M1 = line3(pick(L01, cx), pick(L11, cx), K021);
    const Line3<C> M2 = line3_nok2(pick(L02, cx), K012);
    static_for<3>([&](auto jy) {
      constexpr int cy = decltype(jy)::value - 1;
      const Line3<C> V = line3(pick(M0, cy), pick(M1, cy), pick(M2, cy));
      static_for<3>([&](auto jz) {
        constexpr int cz = decltype(jz)::value - 1;
        constexpr int i = D3Q27::index(cx, cy, cz);
        const C p = pick(V, cz);
        if constexpr (per_direction_w)
          out[i] = keep * h[i] + C(D3Q27::w(i)) * p;
        else
          out[i] = (keep * C(1.0 / D3Q27::w(i))) * (h[i] * scale) + scale * p;
      });
    });
  });
}

// Shifted RR equilibrium of (rho, u): feq_i - w_i (collision.cu:57-69).
template <class L, typename C>
RRLBM_HD RRLBM_INLINE void feq_shifted(C rho, C ux, C uy, C uz, C (&out)[L::Q]) {
  constexpr double isq3 = 1.7320508075688772935;
  const C uu = ux * C(isq3), vv = uy * C(isq3), ww = uz * C(isq3);
  const C a110 = uu * vv, a101 = uu * ww, a011 = vv * ww, a200 = uu * uu, a020 = vv * vv, a002 = ww * ww;
  const C a210 = a200 * vv, a201 = a200 * ww, a120 = a020 * uu, a021 = a020 * ww, a012 = a002 * vv,
          a102 = a002 * uu, a111 = a110 * ww;
  const C dr = rho - C(1);
  static_for<L::Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    constexpr int x = L::cx(i), y = L::cy(i), z = L::cz(i);
    constexpr double w = L::w(i);
    const C fir = C(Hermite::h1(x)) * uu + C(Hermite::h1(y)) * vv + C(Hermite::h1(z)) * ww;
    const C sec = C(2.0 * Hermite::h2(x, y, false)) * a110 + C(2.0 * Hermite::h2(x, z, false)) * a101 +
                  C(2.0 * Hermite::h2(y, z, false)) * a011 + C(Hermite::h2(x, x, true)) * a200 +
                  C(Hermite::h2(y, y, true)) * a020 + C(Hermite::h2(z, z, true)) * a002;
    const C thir = third_order<L, C>(0, a210, a201, a021, a120, a102, a012, a111, Hermite::h3(x, y),
                                     Hermite::h3(x, z), Hermite::h3(y, z), Hermite::h3(y, x), Hermite::h3(z, x),
                                     Hermite::h3(z, y), Hermite::hxyz(x, y, z));
    out[i] = C(w) * (dr + rho * (fir + C(0.5) * (sec + thir)));
  });
}

// Square root for the WALE powers: correctly rounded in FP64; the hardware
// square root (1 ulp) in FP32, which removes the correction sequence of a
// correctly rounded FP32 sqrt (measured 2-3 % of the FP16 LES step on
// MI355X); tau_t is an eddy-viscosity model value, not a conserved quantity.
RRLBM_HD RRLBM_INLINE double wale_sqrt(double x) { return sqrt(x); }
RRLBM_HD RRLBM_INLINE float wale_sqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sqrtf(x);
#else
  return sqrtf(x);
#endif
}

// WALE eddy relaxation time (Nicoud & Ducros 1999, as used by
// RR27V33_code_DNS_LES/LES.cu:79-97 with filter width 1):
//   tau_t = nu_t / cs^2 = 3 Cw^2 (Sd:Sd)^(3/2) / ((S:S)^(5/2) + (Sd:Sd)^(5/4)),
// S the symmetric part of the velocity gradient g (g_ij = du_i/dx_j) and Sd
// the traceless symmetric part of g.g.  Written with the rotation vector w of
// g (its antisymmetric part is the cross-product matrix [w]x, and
// [w]x^2 = w w^T - |w|^2 I): sym(g.g) = S.S + w w^T - |w|^2 I, so
// Sd = M - (tr M / 3) I with M = S.S + w w^T, and S:S = tr(S.S).  The powers
// use x^1.5 = x sqrt(x), x^2.5 = x^2 sqrt(x), x^1.25 = x sqrt(sqrt(x)); a
// vanishing gradient (0/0 in the reference) gives tau_t = 0.
//
// The _d2 forms take D = 2 g, the undivided central differences
// v(x+1) - v(x-1): tau_t is homogeneous of degree 1 in g (numerator degree 6,
// denominator degree 5), so tau_t(g) = tau_t(D) / 2, and the factor 1/2 goes
// into the constant.  Scaling by 2 is exact in binary floating point (and so
// are the square roots of 4x and 16x), so both forms give the same bits.
template <typename C>
RRLBM_HD RRLBM_INLINE C wale_taut_d2(C dudx, C dudy, C dudz, C dvdx, C dvdy, C dvdz, C dwdx, C dwdy, C dwdz, C cw) {
  const C h = C(0.5);
  // S (diagonal a, off-diagonal b) and w
  const C a1 = dudx, a2 = dvdy, a3 = dwdz;
  const C b12 = h * (dudy + dvdx), b13 = h * (dudz + dwdx), b23 = h * (dvdz + dwdy);
  const C w1 = h * (dwdy - dvdz), w2 = h * (dudz - dwdx), w3 = h * (dvdx - dudy);
  // S.S
  const C q12 = b12 * b12, q13 = b13 * b13, q23 = b23 * b23;
  const C p11 = a1 * a1 + q12 + q13, p22 = a2 * a2 + q12 + q23, p33 = a3 * a3 + q13 + q23;
  const C ss = p11 + p22 + p33;  // S:S
  const C m12 = b12 * (a1 + a2) + b13 * b23 + w1 * w2;
  const C m13 = b13 * (a1 + a3) + b12 * b23 + w1 * w3;
  const C m23 = b23 * (a2 + a3) + b12 * b13 + w2 * w3;
  const C m11 = p11 + w1 * w1, m22 = p22 + w2 * w2, m33 = p33 + w3 * w3;
  const C t = (m11 + m22 + m33) * C(1.0 / 3.0);
  const C d1 = m11 - t, d2 = m22 - t, d3 = m33 - t;
  const C sd = d1 * d1 + d2 * d2 + d3 * d3 + C(2) * (m12 * m12 + m13 * m13 + m23 * m23);  // Sd:Sd
  const C rs = wale_sqrt(ss), rd = wale_sqrt(sd);
  const C den = ss * ss * rs + sd * wale_sqrt(rd);
  const C num = C(1.5) * cw * cw * sd * rd;
  return den > C(0) ? num * recip(den) : C(0);
}

// The same from the velocity gradient itself (g_ij = du_i/dx_j).
template <typename C>
RRLBM_HD RRLBM_INLINE C wale_taut(C dudx, C dudy, C dudz, C dvdx, C dvdy, C dvdz, C dwdx, C dwdy, C dwdz, C cw) {
  const C t = C(2);
  return wale_taut_d2<C>(t * dudx, t * dudy, t * dudz, t * dvdx, t * dvdy, t * dvdz, t * dwdx, t * dwdy, t * dwdz, cw);
}

}  // namespace rrlbm
