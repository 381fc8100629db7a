// Host build of the kernels' per-cell arithmetic (csrc/rr_cell.hpp) for CPU
// tests: the same templates the gfx950 kernels instantiate, compiled with g++,
// so the moment-space D3Q27 formulation and the per-direction D3Q19 one can be
// compared with the oracle without a GPU.  Test infrastructure only.
#include <cmath>
#include "rr_cell.hpp"

using namespace rrlbm;

template <typename C, int FM>
static void fast_q27(const C (&h)[27], const Macro<C>& q, const double* F, C omega, C gs, C (&o)[27]) {
  const Moments<C> m = moments_q27<C>(h);
  const NeqTensor<C> N = neq_q27<C, FM>(m, q, C(F[0]), C(F[1]), C(F[2]));
  rr_post_q27<C, true, FM>(q, N, C(F[0]), C(F[1]), C(F[2]), omega, gs, C(1), o);
}

// The FP16-storage form: the polynomial yields (f_i - w_i) * scale / w_i.
template <typename C, int FM>
static void scaled_q27(const C (&h)[27], const Macro<C>& q, const double* F, C omega, C gs, C (&o)[27]) {
  constexpr double kScale = 4096.0;
  const Moments<C> m = moments_q27<C>(h);
  const NeqTensor<C> N = neq_q27<C, FM>(m, q, C(F[0]), C(F[1]), C(F[2]));
  rr_post_q27<C, false, FM>(q, N, C(F[0]), C(F[1]), C(F[2]), omega, gs, C(kScale), o);
  for (int i = 0; i < 27; ++i) o[i] = C(double(o[i]) * D3Q27::w(i) / kScale);
}

// The same with the fourth- to sixth-order terms (rr_post_q27_o6); scaled:
// the FP16-storage form.
template <typename C, int FM, bool scaled>
static void o6_q27(const C (&h)[27], const Macro<C>& q, const double* F, C omega, C gs, C (&o)[27]) {
  constexpr double kScale = 4096.0;
  const Moments<C> m = moments_q27<C>(h);
  const NeqTensor<C> N = neq_q27<C, FM>(m, q, C(F[0]), C(F[1]), C(F[2]));
  if constexpr (scaled) {
    rr_post_q27_o6<C, false, FM>(q, N, C(F[0]), C(F[1]), C(F[2]), omega, gs, C(kScale), o);
    for (int i = 0; i < 27; ++i) o[i] = C(double(o[i]) * D3Q27::w(i) / kScale);
  } else {
    rr_post_q27_o6<C, true, FM>(q, N, C(F[0]), C(F[1]), C(F[2]), omega, gs, C(1), o);
  }
}

template <class L, typename C>
static void cell(const double* f, double rho, double ux, double uy, double uz, const double* F, double tau_eff,
                 int mode, int fast, double* out) {
  C h[L::Q];
  for (int i = 0; i < L::Q; ++i) h[i] = C(f[i] - L::w(i));
  Macro<C> q{C(rho), C(rho - 1.0), C(ux), C(uy), C(uz)};
  const C omega = mode == 0 ? C(1.0 - 1.0 / tau_eff) : C(1);
  const C gs = mode == 0 ? C(1) : C(-1);
  C o[L::Q];
  if constexpr (L::Q == 27) {
    if (fast == 1) {
      fast_q27<C, kForceAll>(h, q, F, omega, gs, o);
    } else if (fast == 2) {
      fast_q27<C, kForceX>(h, q, F, omega, gs, o);
    } else if (fast == 3) {
      fast_q27<C, kForceNone>(h, q, F, omega, gs, o);
    } else if (fast == 4) {
      scaled_q27<C, kForceAll>(h, q, F, omega, gs, o);
    } else if (fast == 5) {
      scaled_q27<C, kForceX>(h, q, F, omega, gs, o);
    } else if (fast == 6) {
      scaled_q27<C, kForceNone>(h, q, F, omega, gs, o);
    } else if (fast == 7) {
      o6_q27<C, kForceAll, false>(h, q, F, omega, gs, o);
    } else if (fast == 8) {
      o6_q27<C, kForceX, false>(h, q, F, omega, gs, o);
    } else if (fast == 9) {
      o6_q27<C, kForceNone, false>(h, q, F, omega, gs, o);
    } else if (fast == 10) {
      o6_q27<C, kForceAll, true>(h, q, F, omega, gs, o);
    } else {
      rr_post_generic<L, C>(h, q, C(F[0]), C(F[1]), C(F[2]), omega, gs, o);
    }
  } else if (fast == 11 || fast == 12) {
    // D3Q19 moment-space path (no body force); 12: the FP16-storage form
    const Moments<C> m = moments_generic<L, C>(h);
    const NeqTensor<C> N = neq_q27<C, kForceNone>(m, q, C(0), C(0), C(0));
    if (fast == 11) {
      rr_post_q19<C, true>(q, N, omega, C(1), o);
    } else {
      rr_post_q19<C, false>(q, N, omega, C(4096.0), o);
      for (int i = 0; i < L::Q; ++i) o[i] = C(double(o[i]) * L::w(i) / 4096.0);
    }
  } else {
    rr_post_generic<L, C>(h, q, C(F[0]), C(F[1]), C(F[2]), omega, gs, o);
  }
  for (int i = 0; i < L::Q; ++i) out[i] = double(o[i]) + L::w(i);
}

template <class L>
static void moments(const double* f, double* m) {
  double h[L::Q];
  for (int i = 0; i < L::Q; ++i) h[i] = f[i] - L::w(i);
  Moments<double> r;
  if constexpr (L::Q == 27) r = moments_q27<double>(h);
  else r = moments_generic<L, double>(h);
  const double v[10] = {r.dr, r.jx, r.jy, r.jz, r.pxx, r.pyy, r.pzz, r.pxy, r.pxz, r.pyz};
  for (int k = 0; k < 10; ++k) m[k] = v[k];
}

extern "C" {
// q: 27 or 19; single: 1 = FP32 arithmetic; fast: 0 per-direction path, 1/2/3 moment-space D3Q27
// path with all force components / force along x only / no force, 4/5/6 the same in the
// FP16-storage form (values scaled by 4096 / w_i); 7/8/9 the sixth-order path with all force
// components / x only / none, 10 its FP16-storage form; q = 19: 11 the moment-space D3Q19 path
// without force, 12 its FP16-storage form
void hc_rr_cell(int q, int single, int fast, const double* f, double rho, double ux, double uy, double uz,
                const double* F, double tau_eff, int mode, double* out) {
  if (q == 27) {
    if (single) cell<D3Q27, float>(f, rho, ux, uy, uz, F, tau_eff, mode, fast, out);
    else cell<D3Q27, double>(f, rho, ux, uy, uz, F, tau_eff, mode, fast, out);
  } else {
    if (single) cell<D3Q19, float>(f, rho, ux, uy, uz, F, tau_eff, mode, fast, out);
    else cell<D3Q19, double>(f, rho, ux, uy, uz, F, tau_eff, mode, fast, out);
  }
}
// BGK update of one D3Q27 cell, shifted populations in and out.  poly: 0 the
// per-direction bgk_post, 1 the moment-space bgk_post_q27, 2 its FP16-storage
// form (values scaled by 4096 / w_i, unscaled here).
void hc_bgk_cell(int single, int poly, const double* h, double dr, double ux, double uy, double uz, const double* F,
                 double tau, double* out) {
  auto run = [&](auto zero) {
    using C = decltype(zero);
    C hh[27], o[27];
    const double kS = 4096.0;
    for (int i = 0; i < 27; ++i) hh[i] = C(poly == 2 ? h[i] * kS / D3Q27::w(i) : h[i]);
    const C it = C(1.0 / tau), fc = C(1.0 - 0.5 / tau);
    if (poly == 0)
      bgk_post<D3Q27, C, true>(hh, C(dr), C(ux), C(uy), C(uz), C(F[0]), C(F[1]), C(F[2]), it, fc, C(1), o);
    else if (poly == 1)
      bgk_post_q27<C, true>(hh, C(dr), C(ux), C(uy), C(uz), C(F[0]), C(F[1]), C(F[2]), it, fc, C(1), o);
    else {
      // the kernels unscale on load: h = stored * w / S
      for (int i = 0; i < 27; ++i) hh[i] = C(h[i]);
      bgk_post_q27<C, false>(hh, C(dr), C(ux), C(uy), C(uz), C(F[0]), C(F[1]), C(F[2]), it, fc, C(kS), o);
      for (int i = 0; i < 27; ++i) o[i] = C(double(o[i]) * D3Q27::w(i) / kS);
    }
    for (int i = 0; i < 27; ++i) out[i] = double(o[i]);
  };
  if (single) run(0.0f);
  else run(0.0);
}
// Raw moments of a D3Q27 cell from weight-scaled values g_i = h_i * scale / w_i
// (moments_q27_wscaled, the FP16 kernels' moment sum); single: FP32 arithmetic.
void hc_moments_wscaled(int single, const double* g, double scale, double* m) {
  auto run = [&](auto zero) {
    using C = decltype(zero);
    C gg[27];
    for (int i = 0; i < 27; ++i) gg[i] = C(g[i]);
    const Moments<C> r = moments_q27_wscaled<C, C>(gg, C(1), C(1.0 / (216.0 * scale)));
    const C v[10] = {r.dr, r.jx, r.jy, r.jz, r.pxx, r.pyy, r.pzz, r.pxy, r.pxz, r.pyz};
    for (int k = 0; k < 10; ++k) m[k] = double(v[k]);
  };
  if (single) run(0.0f);
  else run(0.0);
}
void hc_moments(int q, const double* f, double* m) {
  if (q == 27) moments<D3Q27>(f, m);
  else moments<D3Q19>(f, m);
}
double hc_wale(const double* g, double cw) {
  return wale_taut<double>(g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7], g[8], cw);
}
void hc_feq(int q, double rho, double ux, double uy, double uz, double* out) {
  if (q == 27) {
    double o[27];
    feq_shifted<D3Q27, double>(rho, ux, uy, uz, o);
    for (int i = 0; i < 27; ++i) out[i] = o[i] + D3Q27::w(i);
  } else {
    double o[19];
    feq_shifted<D3Q19, double>(rho, ux, uy, uz, o);
    for (int i = 0; i < 19; ++i) out[i] = o[i] + D3Q19::w(i);
  }
}
}
