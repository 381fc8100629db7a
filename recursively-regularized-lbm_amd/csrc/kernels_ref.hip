// Drop-in kernels with the channel solver's data layout: FP64 populations
// SoA f[ip*N + idx], idx = (iz*LY + iy)*LX + ix, separate FP64 rho, ux, uy, uz,
// force and tau_t fields (RR27V33_code_DNS_LES/LBMLES.h:58-81).  Each kernel
// performs one reference routine over the whole grid, one thread per cell, so
// a driver calling them in the main.cu order gets the reference's step.
// They also provide initialisation, the pack step into the fused storage and
// the FP64 diagnostics.
#include "kernels.hpp"
#include "kernels_ref.hpp"
#include "rr_cell.hpp"

namespace rrlbm {

struct Cell3 {
  uint32_t x, y, z;
};
RRLBM_INLINE __device__ Cell3 coords(uint32_t cell, const Grid& g) {
  const uint32_t z = fastdiv(cell, g.div_lxly);
  const uint32_t r = cell - z * g.lx * g.ly;
  const uint32_t y = fastdiv(r, g.div_lx);
  return {r - y * g.lx, y, z};
}
RRLBM_INLINE __device__ uint32_t lin(uint32_t x, uint32_t y, uint32_t z, const Grid& g) {
  return (z * g.ly + y) * g.lx + x;
}
RRLBM_INLINE __device__ uint32_t tid() { return blockIdx.x * kBlock + threadIdx.x; }
RRLBM_INLINE __device__ bool on_face(int face, const Cell3& c, const Grid& g) {
  switch (face) {
    case 0: return (g.walls & 1) && c.x == 0;
    case 1: return (g.walls & 1) && c.x == g.lx - 1;
    case 2: return (g.walls & 2) && c.y == 0;
    case 3: return (g.walls & 2) && c.y == g.ly - 1;
    case 4: return (g.walls & 4) && c.z == 0;
    default: return (g.walls & 4) && c.z == g.lz - 1;
  }
}
RRLBM_HD constexpr bool enters_face(int face, int x, int y, int z) {
  return face == 0 ? x > 0 : face == 1 ? x < 0 : face == 2 ? y > 0 : face == 3 ? y < 0 : face == 4 ? z > 0 : z < 0;
}

// collision_RR (collision.cu:4-119), in place; tau_eff = tau + tau_t with LES.
template <class L>
__global__ __launch_bounds__(kBlock) void k_collision(RefFields r, Grid g, double tau, int use_les) {
  const uint32_t cell = tid();
  if (cell >= g.n) return;
  constexpr int Q = L::Q;
  double h[Q];
  static_for<Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    h[i] = r.f[(size_t)i * g.n + cell] - L::w(i);
  });
  Macro<double> q;
  q.rho = r.rho[cell];
  q.dr = q.rho - 1.0;
  q.ux = r.ux[cell]; q.uy = r.uy[cell]; q.uz = r.uz[cell];
  const double Fx = r.fx[cell], Fy = r.fy[cell], Fz = r.fz[cell];
  const double tau_eff = tau + (use_les ? r.taut[cell] : 0.0);
  const double omega = 1.0 - 1.0 / tau_eff;
  double out[Q];
  if constexpr (Q == 27) {
    const Moments<double> m = moments_q27<double>(h);
    const NeqTensor<double> N = neq_q27<double>(m, q, Fx, Fy, Fz);
    if (g.rr6)
      rr_post_q27_o6<double, true>(q, N, Fx, Fy, Fz, omega, 1.0, 1.0, out);
    else
      rr_post_q27<double, true>(q, N, Fx, Fy, Fz, omega, 1.0, 1.0, out);
  } else {
    rr_post_generic<L, double>(h, q, Fx, Fy, Fz, omega, 1.0, out);
  }
  static_for<Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    r.f[(size_t)i * g.n + cell] = out[i] + L::w(i);
  });
}

// collision_BGK (BGK/collision.cu:3-33), in place, in this library's field
// convention (f unshifted, rho = 1 + the BGK solver's density deviation):
// the BGK solver's shifted populations are f - w.
__global__ __launch_bounds__(kBlock) void k_collision_bgk(RefFields r, Grid g, double tau) {
  const uint32_t cell = tid();
  if (cell >= g.n) return;
  using L = D3Q27;
  double h[27], out[27];
  static_for<27>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    h[i] = r.f[(size_t)i * g.n + cell] - L::w(i);
  });
  bgk_post<L, double, true>(h, r.rho[cell] - 1.0, r.ux[cell], r.uy[cell], r.uz[cell], r.fx[cell], r.fy[cell],
                            r.fz[cell], 1.0 / tau, 1.0 - 0.5 / tau, 1.0, out);
  static_for<27>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    r.f[(size_t)i * g.n + cell] = out[i] + L::w(i);
  });
}

// streaming (streaming.cu:3-24): push into ftmp, periodic in x, y, z.
template <class L>
__global__ __launch_bounds__(kBlock) void k_streaming(const double* f, double* ftmp, Grid g) {
  const uint32_t cell = tid();
  if (cell >= g.n) return;
  const Cell3 c = coords(cell, g);
  static_for<L::Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    const uint32_t x = (c.x + g.lx + L::cx(i)) % g.lx, y = (c.y + g.ly + L::cy(i)) % g.ly,
                   z = (c.z + g.lz + L::cz(i)) % g.lz;
    ftmp[(size_t)i * g.n + lin(x, y, z, g)] = f[(size_t)i * g.n + cell];
  });
}

// BB (streaming.cu:26-55): half-way bounce-back of the wall-incoming
// populations from the post-collision field f; faces x-,x+,y-,y+,z-,z+, lid
// term 6 w_i c_i.u_lid on the lid face (Cavity/streaming.f90:142-152).
template <class L>
__global__ __launch_bounds__(kBlock) void k_bb(const double* f, double* ftmp, Grid g) {
  const uint32_t cell = tid();
  if (cell >= g.n) return;
  const Cell3 c = coords(cell, g);
  static_for<6>([&](auto fc) {
    constexpr int face = decltype(fc)::value;
    if (!on_face(face, c, g)) return;
    static_for<L::Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      if constexpr (enters_face(face, L::cx(i), L::cy(i), L::cz(i))) {
        double v = f[(size_t)L::opp(i) * g.n + cell];
        if (face == g.lid_face)
          v += 6.0 * L::w(i) * (L::cx(i) * g.lid_x + L::cy(i) * g.lid_y + L::cz(i) * g.lid_z);
        ftmp[(size_t)i * g.n + cell] = v;
      }
    });
  });
}

RRLBM_INLINE __device__ bool on_any_wall(const Cell3& c, const Grid& g) {
  return ((g.walls & 1) && (c.x == 0 || c.x == g.lx - 1)) || ((g.walls & 2) && (c.y == 0 || c.y == g.ly - 1)) ||
         ((g.walls & 4) && (c.z == 0 || c.z == g.lz - 1));
}

// macrovar (macrovar.cu:4-31); wall_only: macrovar_BN (macrovar.cu:33-62).
// bgk: the BGK solver's velocity u = j + F/2 (BGK/macrovar.cu:4-31).
template <class L>
__global__ __launch_bounds__(kBlock) void k_macrovar(RefFields r, Grid g, int wall_only, int bgk) {
  const uint32_t cell = tid();
  if (cell >= g.n) return;
  if (wall_only && !on_any_wall(coords(cell, g), g)) return;
  double rho = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
  static_for<L::Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    const double v = r.f[(size_t)i * g.n + cell];
    rho += v;
    if constexpr (L::cx(i) != 0) jx += L::cx(i) * v;
    if constexpr (L::cy(i) != 0) jy += L::cy(i) * v;
    if constexpr (L::cz(i) != 0) jz += L::cz(i) * v;
  });
  r.rho[cell] = rho;
  const double inv = bgk ? 1.0 : 1.0 / rho;
  r.ux[cell] = (jx + 0.5 * r.fx[cell]) * inv;
  r.uy[cell] = (jy + 0.5 * r.fy[cell]) * inv;
  r.uz[cell] = (jz + 0.5 * r.fz[cell]) * inv;
}

// RRBC (BC.cu:4-126): at wall cells replace the wall-incoming populations by
// feq + fneq - Fbar feq.  Several faces (cavity edges) are treated as the
// reference's Cavity solver does (Cavity/streaming.f90:89-292): face by face
// in the order x-,x+,y-,y+,z-,z+, the face's incoming populations take their
// bounce-back values (from the post-collision field fpost, plus the lid term)
// and are regularised with rho, u of the cell's current populations, while
// populations entering through a later face still hold the values the
// periodic streaming brought.  BB has already written every face's
// bounce-back values, so such a cell first takes the streamed values back
// from fpost at the upwind cell.  A cell on one face reads rho, u from the
// macrovar output, as the reference RRBC does.
template <class L>
__global__ __launch_bounds__(kBlock) void k_rrbc(RefFields r, const double* fpost, Grid g) {
  const uint32_t cell = tid();
  if (cell >= g.n) return;
  const Cell3 c = coords(cell, g);
  if (!on_any_wall(c, g)) return;
  constexpr int Q = L::Q;
  int nactive = 0;
  static_for<6>([&](auto fc) { nactive += on_face(decltype(fc)::value, c, g) ? 1 : 0; });
  double h[Q];
  static_for<Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    h[i] = r.f[(size_t)i * g.n + cell] - L::w(i);
  });
  if (nactive > 1) {
    static_for<Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      bool entering = false;
      static_for<6>([&](auto fc) {
        constexpr int face = decltype(fc)::value;
        if constexpr (enters_face(face, L::cx(i), L::cy(i), L::cz(i))) entering = entering || on_face(face, c, g);
      });
      if (entering) {
        const uint32_t x = (c.x + g.lx - L::cx(i)) % g.lx, y = (c.y + g.ly - L::cy(i)) % g.ly,
                       z = (c.z + g.lz - L::cz(i)) % g.lz;
        h[i] = fpost[(size_t)i * g.n + lin(x, y, z, g)] - L::w(i);
      }
    });
  }
  const double Fx = r.fx[cell], Fy = r.fy[cell], Fz = r.fz[cell];
  static_for<6>([&](auto fc) {
    constexpr int face = decltype(fc)::value;
    if (!on_face(face, c, g)) return;
    Macro<double> q;
    if (nactive == 1) {
      q.rho = r.rho[cell]; q.dr = q.rho - 1.0;
      q.ux = r.ux[cell]; q.uy = r.uy[cell]; q.uz = r.uz[cell];
    } else {
      static_for<Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        if constexpr (enters_face(face, L::cx(i), L::cy(i), L::cz(i))) {
          double v = fpost[(size_t)L::opp(i) * g.n + cell] - L::w(i);
          if (face == g.lid_face)
            v += 6.0 * L::w(i) * (L::cx(i) * g.lid_x + L::cy(i) * g.lid_y + L::cz(i) * g.lid_z);
          h[i] = v;
        }
      });
      Moments<double> m0;
      if constexpr (Q == 27) m0 = moments_q27<double>(h);
      else m0 = moments_generic<L, double>(h);
      q = macro_of<double>(m0, Fx, Fy, Fz);
    }
    double out[Q];
    if constexpr (Q == 27) {
      const Moments<double> m = moments_q27<double>(h);
      const NeqTensor<double> N = neq_q27<double>(m, q, Fx, Fy, Fz);
      if (g.rr6)
        rr_post_q27_o6<double, true>(q, N, Fx, Fy, Fz, 1.0, -1.0, 1.0, out);
      else
        rr_post_q27<double, true>(q, N, Fx, Fy, Fz, 1.0, -1.0, 1.0, out);
    } else {
      rr_post_generic<L, double>(h, q, Fx, Fy, Fz, 1.0, -1.0, out);
    }
    static_for<Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      if constexpr (enters_face(face, L::cx(i), L::cy(i), L::cz(i))) h[i] = out[i];
    });
  });
  static_for<Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    r.f[(size_t)i * g.n + cell] = h[i] + L::w(i);
  });
}

// FDBC (BC.cu:129-309), the finite-difference wall reconstruction at the y
// walls, in two passes over the 2*LX*LZ wall cells (t -> x = t % LX, then the
// side, then z).  Pass 1 extrapolates every wall velocity from the two next
// rows (BC.cu:149-161, quadratic through u = 0 on the half-way wall); pass 2
// differentiates (one-sided second order in y, BC.cu:167-180; central in x
// and z over the extrapolated wall row, BC.cu:182-196) and replaces all
// populations by rho_l (feq + fneq(-tau_eff S) - Fbar feq)(rho = 1), rho_l
// keeping the sum of the known populations (c_y pointing into the wall,
// BC.cu:203-265).  The reference does both in one kernel, where the x/z
// differences race with the neighbours' extrapolation; the two passes give
// every difference the extrapolated values.
struct WallCell {
  uint32_t x, y, z, y1, y2;
  bool top;
};
RRLBM_INLINE __device__ WallCell wall_cell(uint32_t t, const Grid& g) {
  WallCell w;
  w.x = t % g.lx;
  const uint32_t r = t / g.lx;
  w.top = (r & 1u) != 0;
  w.z = r >> 1;
  w.y = w.top ? g.ly - 1 : 0;
  w.y1 = w.top ? g.ly - 2 : 1;
  w.y2 = w.top ? g.ly - 3 : 2;
  return w;
}

__global__ __launch_bounds__(kBlock) void k_fdbc_velocity(RefFields r, Grid g) {
  const uint32_t t = tid();
  if (t >= 2 * g.lx * g.lz) return;
  const WallCell w = wall_cell(t, g);
  const uint32_t c = lin(w.x, w.y, w.z, g), c1 = lin(w.x, w.y1, w.z, g), c2 = lin(w.x, w.y2, w.z, g);
  r.ux[c] = 2.0 / 3.0 * r.ux[c1] - 1.0 / 5.0 * r.ux[c2];
  r.uy[c] = 2.0 / 3.0 * r.uy[c1] - 1.0 / 5.0 * r.uy[c2];
  r.uz[c] = 2.0 / 3.0 * r.uz[c1] - 1.0 / 5.0 * r.uz[c2];
}

__global__ __launch_bounds__(kBlock) void k_fdbc_reconstruct(RefFields r, Grid g, double tau, int use_les) {
  const uint32_t t = tid();
  if (t >= 2 * g.lx * g.lz) return;
  const WallCell w = wall_cell(t, g);
  const uint32_t c = lin(w.x, w.y, w.z, g), c1 = lin(w.x, w.y1, w.z, g), c2 = lin(w.x, w.y2, w.z, g);
  const uint32_t xp = lin(w.x + 1 == g.lx ? 0 : w.x + 1, w.y, w.z, g);
  const uint32_t xm = lin(w.x == 0 ? g.lx - 1 : w.x - 1, w.y, w.z, g);
  const uint32_t zp = lin(w.x, w.y, w.z + 1 == g.lz ? 0 : w.z + 1, g);
  const uint32_t zm = lin(w.x, w.y, w.z == 0 ? g.lz - 1 : w.z - 1, g);
  const double sg = w.top ? -1.0 : 1.0;
  auto ddy = [&](const double* u) { return sg * (-3.0 * u[c] + 4.0 * u[c1] - u[c2]) * 0.5; };
  auto ddx = [&](const double* u) { return 0.5 * (u[xp] - u[xm]); };
  auto ddz = [&](const double* u) { return 0.5 * (u[zp] - u[zm]); };
  const double tau_eff = tau + (use_les ? r.taut[c] : 0.0);
  // the reconstruction per unit density: feq + fneq - Fbar feq with the
  // non-equilibrium Hermite coefficients -tau_eff S used without the 1/2 of
  // the collision's fneq, i.e. the moment-space form with N = -(2/3) tau_eff S
  const double k = -2.0 / 3.0 * tau_eff;
  NeqTensor<double> N;
  N.xx = k * ddx(r.ux);
  N.yy = k * ddy(r.uy);
  N.zz = k * ddz(r.uz);
  N.xy = k * 0.5 * (ddy(r.ux) + ddx(r.uy));
  N.xz = k * 0.5 * (ddz(r.ux) + ddx(r.uz));
  N.yz = k * 0.5 * (ddz(r.uy) + ddy(r.uz));
  Macro<double> q;
  q.rho = 1.0; q.dr = 0.0;
  q.ux = r.ux[c]; q.uy = r.uy[c]; q.uz = r.uz[c];
  using L = D3Q27;
  double g1[27];
  rr_post_q27<double, true>(q, N, r.fx[c], r.fy[c], r.fz[c], 1.0, -1.0, 1.0, g1);
  double sum_ps = 0.0, sum_in = 0.0;
  static_for<27>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    g1[i] += L::w(i);
    if constexpr (L::cy(i) != 0) {
      if ((L::cy(i) > 0) == w.top) {
        sum_ps += r.f[(size_t)i * g.n + c];
        sum_in += g1[i];
      }
    }
  });
  const double rho_l = sum_ps / sum_in;
  static_for<27>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    r.f[(size_t)i * g.n + c] = rho_l * g1[i];
  });
}

// LES_WALE (LES.cu:4-98): central differences, one-sided at walled faces.
__global__ __launch_bounds__(kBlock) void k_les_wale(RefFields r, Grid g, double cw) {
  const uint32_t cell = tid();
  if (cell >= g.n) return;
  const Cell3 c = coords(cell, g);
  const uint32_t xp = c.x + 1 == g.lx ? 0 : c.x + 1, xm = c.x == 0 ? g.lx - 1 : c.x - 1;
  const uint32_t yp = c.y + 1 == g.ly ? 0 : c.y + 1, ym = c.y == 0 ? g.ly - 1 : c.y - 1;
  const uint32_t zp = c.z + 1 == g.lz ? 0 : c.z + 1, zm = c.z == 0 ? g.lz - 1 : c.z - 1;
  const uint32_t ixp = lin(xp, c.y, c.z, g), ixm = lin(xm, c.y, c.z, g);
  const uint32_t iyp = lin(c.x, yp, c.z, g), iym = lin(c.x, ym, c.z, g);
  const uint32_t izp = lin(c.x, c.y, zp, g), izm = lin(c.x, c.y, zm, g);
  auto d = [&](const double* u, uint32_t m, uint32_t p, bool lo, bool hi) -> double {
    if (lo) return (3.0 * u[cell] + u[p]) / 3.0;
    if (hi) return (-3.0 * u[cell] - u[m]) / 3.0;
    return 0.5 * (u[p] - u[m]);
  };
  const bool xlo = (g.walls & 1) && c.x == 0, xhi = (g.walls & 1) && c.x == g.lx - 1;
  const bool ylo = (g.walls & 2) && c.y == 0, yhi = (g.walls & 2) && c.y == g.ly - 1;
  const bool zlo = (g.walls & 4) && c.z == 0, zhi = (g.walls & 4) && c.z == g.lz - 1;
  const double dudx = d(r.ux, ixm, ixp, xlo, xhi), dvdx = d(r.uy, ixm, ixp, xlo, xhi), dwdx = d(r.uz, ixm, ixp, xlo, xhi);
  const double dudy = d(r.ux, iym, iyp, ylo, yhi), dvdy = d(r.uy, iym, iyp, ylo, yhi), dwdy = d(r.uz, iym, iyp, ylo, yhi);
  const double dudz = d(r.ux, izm, izp, zlo, zhi), dvdz = d(r.uy, izm, izp, zlo, zhi), dwdz = d(r.uz, izm, izp, zlo, zhi);
  r.taut[cell] = wale_taut<double>(dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy, dwdz, cw);
}

// forcing (macrovar.cu:63-74), with the force vector of the parameters.
__global__ __launch_bounds__(kBlock) void k_forcing(RefFields r, Grid g, double fx, double fy, double fz) {
  const uint32_t cell = tid();
  if (cell >= g.n) return;
  r.fx[cell] = fx; r.fy[cell] = fy; r.fz[cell] = fz;
}

// forcing_GP (macrovar.cu:76-114): time-periodic streak-triggering force in two
// wall bands; the upper band's x phase uses the integer ix/LX (always 0), as
// the reference writes it.
__global__ __launch_bounds__(kBlock) void k_forcing_gp(RefFields r, Grid g, double force_in_x, int istep,
                                                       double rand_phase) {
  const uint32_t cell = tid();
  if (cell >= g.n) return;
  const Cell3 c = coords(cell, g);
  const double pi2 = 6.283185307179586476925;
  const double Tpd = 2000.0;
  const double LYq = (double)(g.ly / 4);
  const int h0 = 2;
  const double beta9 = 3.0, gamma9 = 2.0;
  const double Amp0 = (100.0 * beta9 / g.ly) * sin(pi2 * istep / Tpd);
  const int ix = c.x, iy = c.y, iz = c.z;
  const double LX = g.lx, LY = g.ly, LZ = g.lz;
  double fx = force_in_x, fy = 0.0, fz = 0.0;
  if (iy > h0 && iy < LYq + h0) {
    fx = force_in_x * (1.0 - Amp0 * LX / beta9 * sin(pi2 * (iy - h0) / LYq) * sin(beta9 * pi2 * ix / LX) *
                                 cos(gamma9 * pi2 * iz / LZ));
    fy = force_in_x * Amp0 * 0.5 * LYq * (1 - cos(pi2 * (iy - h0) / LYq)) * cos(beta9 * pi2 * ix / LX) *
         cos(gamma9 * pi2 * iz / LZ);
    fz = force_in_x * Amp0 * 0.5 * LZ / gamma9 * sin(pi2 * (iy - h0) / LYq) * cos(beta9 * pi2 * ix / LX) *
         sin(gamma9 * pi2 * iz / LZ);
  } else if (iy > 0.75 * LY - h0 && iy < LY - h0) {
    const double xph = (double)(ix / (int)g.lx) + rand_phase;
    fx = force_in_x * (1.0 + Amp0 * LX / beta9 * sin(pi2 * (-LY + iy + h0 + LYq) / LYq) * sin(beta9 * pi2 * xph) *
                                 cos(gamma9 * pi2 * iz / LZ));
    fy = -force_in_x * Amp0 * 0.5 * LYq * (1 - cos(pi2 * (-LY + iy + h0 + LYq) / LYq)) * cos(beta9 * pi2 * xph) *
         cos(gamma9 * pi2 * iz / LZ);
    fz = -force_in_x * Amp0 * 0.5 * LZ / gamma9 * sin(pi2 * (-LY + iy + h0 + LYq) / LYq) * cos(beta9 * pi2 * xph) *
         sin(gamma9 * pi2 * iz / LZ);
  }
  r.fx[cell] = fx; r.fy[cell] = fy; r.fz[cell] = fz;
}

// init_arrays (initial.cu:268-325): rho = 1, u_x from the viscous / log10 law
// plus a sinusoidal perturbation, F = (force_in_x, 0, 0), populations = rest
// equilibrium shifted by half the force.
template <class L>
__global__ __launch_bounds__(kBlock) void k_init_arrays(RefFields r, Grid g, double ustar, double ystar,
                                                        double force_in_x) {
  const uint32_t cell = tid();
  if (cell >= g.n) return;
  const Cell3 c = coords(cell, g);
  const double pi2 = 6.283185307179586476925;
  const double A9 = 0.1 * sqrt(2.0), alpha = 1.0, beta9 = 1.0, cc = 60.0;
  const double ccc1 = -(double)g.lx / pi2 / alpha / ystar * A9 * ustar / cc / cc;
  const double z9 = pi2 * (c.z + 0.5) / g.lz, x9 = pi2 * (c.x + 0.5) / g.lx;
  const int iy = c.y, LY = g.ly;
  const double yplus = iy <= LY / 2 ? (iy + 0.5) / ystar : (LY - iy - 0.5) / ystar;
  double u = yplus <= 10.8 ? ustar * yplus : (log10(yplus) / 0.41 + 5.0) * ustar;
  u += ccc1 * yplus * sin(alpha * x9 + beta9 * z9);
  r.rho[cell] = 1.0;
  r.ux[cell] = u; r.uy[cell] = 0.0; r.uz[cell] = 0.0;
  r.fx[cell] = force_in_x; r.fy[cell] = 0.0; r.fz[cell] = 0.0;
  static_for<L::Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    const double feq = L::w(i) * 1.0;
    r.f[(size_t)i * g.n + cell] = feq - 0.5 * force_in_x * L::cx(i) * 3.0 * feq;
  });
}

// Populations = RR equilibrium of the rho, u fields.
// bgk: the BGK solver's second-order incompressible equilibrium
// w (rho + 3 c.u + 9/2 (c.u)^2 - 3/2 u^2).
template <class L>
__global__ __launch_bounds__(kBlock) void k_init_equilibrium(RefFields r, Grid g, int bgk) {
  const uint32_t cell = tid();
  if (cell >= g.n) return;
  double out[L::Q];
  if (bgk) {
    const double ux = r.ux[cell], uy = r.uy[cell], uz = r.uz[cell];
    const double uu = 1.5 * (ux * ux + uy * uy + uz * uz);
    static_for<L::Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const double cu = L::cx(i) * ux + L::cy(i) * uy + L::cz(i) * uz;
      out[i] = L::w(i) * (r.rho[cell] - 1.0 + 3.0 * cu + 4.5 * cu * cu - uu);
    });
  } else if constexpr (L::Q == 27) {
    if (g.rr6) {
      // the full-order equilibrium: the sixth-order post-collision state of a
      // cell without non-equilibrium part or force
      Macro<double> q;
      q.rho = r.rho[cell]; q.dr = q.rho - 1.0;
      q.ux = r.ux[cell]; q.uy = r.uy[cell]; q.uz = r.uz[cell];
      const NeqTensor<double> N{0, 0, 0, 0, 0, 0};
      rr_post_q27_o6<double, true, kForceNone>(q, N, 0.0, 0.0, 0.0, 0.0, 1.0, 1.0, out);
    } else {
      feq_shifted<L, double>(r.rho[cell], r.ux[cell], r.uy[cell], r.uz[cell], out);
    }
  } else {
    feq_shifted<L, double>(r.rho[cell], r.ux[cell], r.uy[cell], r.uz[cell], out);
  }
  static_for<L::Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    r.f[(size_t)i * g.n + cell] = out[i] + L::w(i);
  });
}

RRLBM_HD RRLBM_INLINE uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
RRLBM_HD RRLBM_INLINE double uniform01(uint64_t seed, uint64_t stream, uint64_t counter) {
  const uint64_t h = splitmix64(seed ^ splitmix64(stream * 0xD1B54A32D192ED03ull + counter));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}
RRLBM_HD RRLBM_INLINE double uplus_profile(double yplus) {
  return yplus <= 10.8 ? yplus : log(yplus) / 0.41 + 5.0;
}

// Seeded initial fields of the configurations (rho, u only; populations follow
// from k_init_equilibrium).
__global__ __launch_bounds__(kBlock) void k_init_fields(RefFields r, Grid g, InitSpec s) {
  const uint32_t cell = tid();
  if (cell >= g.n) return;
  const Cell3 c = coords(cell, g);
  const double pi2 = 6.283185307179586476925;
  double rho = 1.0, u = 0.0, v = 0.0, w = 0.0;
  if (s.kind == kInitTaylorGreen2D) {
    // u = u0 sin(kx x) cos(ky y), v = -u0 cos(kx x) sin(ky y), rho = 1 + u0^2/(4 cs^2)(cos 2kx x + cos 2ky y)
    const double kx = pi2 / g.lx, ky = pi2 / g.ly;
    const double x = c.x + 0.5, y = c.y + 0.5;
    u = s.u0 * sin(kx * x) * cos(ky * y);
    v = -s.u0 * cos(kx * x) * sin(ky * y);
    rho = 1.0 + s.u0 * s.u0 * 0.75 * (cos(2.0 * kx * x) + cos(2.0 * ky * y));
  } else if (s.kind == kInitTaylorGreen3D) {
    // u = u0 sin x cos y cos z, v = -u0 cos x sin y cos z, w = 0 on [0, 2 pi)^3,
    // p = u0^2/16 (cos 2x + cos 2y)(cos 2z + 2), rho = 1 + 3 p
    const double x = pi2 * (c.x + 0.5) / g.lx, y = pi2 * (c.y + 0.5) / g.ly, z = pi2 * (c.z + 0.5) / g.lz;
    u = s.u0 * sin(x) * cos(y) * cos(z);
    v = -s.u0 * cos(x) * sin(y) * cos(z);
    rho = 1.0 + 3.0 * s.u0 * s.u0 / 16.0 * (cos(2.0 * x) + cos(2.0 * y)) * (cos(2.0 * z) + 2.0);
  } else if (s.kind == kInitChannel) {
    const uint64_t seed = s.seed;
    const int kx = 2 + (int)(splitmix64(seed + 1) % 3), kz = 2 + (int)(splitmix64(seed + 2) % 3);
    const double phi = pi2 * uniform01(seed, 3, 0);
    const double ystar = s.visc / s.u0;
    const int LY = g.ly;
    const double dist = (int)c.y < LY / 2 ? c.y + 0.5 : LY - c.y - 0.5;
    const double U = s.u0 * uplus_profile(dist / ystar);
    const double streak = s.amp_streak * U * sin(pi2 * kz * (c.z + 0.5) / g.lz) * cos(pi2 * kx * (c.x + 0.5) / g.lx + phi);
    const double an = s.amp_noise * s.ubulk;
    u = U + streak + an * (2.0 * uniform01(seed, 10, cell) - 1.0);
    v = an * (2.0 * uniform01(seed, 11, cell) - 1.0);
    w = an * (2.0 * uniform01(seed, 12, cell) - 1.0);
  }
  r.rho[cell] = rho;
  r.ux[cell] = u; r.uy[cell] = v; r.uz[cell] = w;
}

// Pack: the reference-layout state at the end of a step (populations after
// streaming / bounce-back / RRBC, with its rho and u) is collided (the next
// step's collision_RR with the uniform force of the parameters) and stored in
// the fused kernel's storage.
// cell_force: the collision takes the force fields (as after forcing_GP)
// instead of the uniform force fx, fy, fz.
template <int S, class L>
__global__ __launch_bounds__(kBlock) void k_pack(RefFields r, Grid g, void* dst, PopLayout lay, double tau, int use_les,
                                                 double fx, double fy, double fz, double f16_scale, int bgk,
                                                 int cell_force) {
  const uint32_t cell = tid();
  if (cell >= g.n) return;
  if (cell_force) {
    fx = r.fx[cell]; fy = r.fy[cell]; fz = r.fz[cell];
  }
  constexpr int Q = L::Q;
  double h[Q];
  static_for<Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    h[i] = r.f[(size_t)i * g.n + cell] - L::w(i);
  });
  Macro<double> q;
  q.rho = r.rho[cell]; q.dr = q.rho - 1.0;
  q.ux = r.ux[cell]; q.uy = r.uy[cell]; q.uz = r.uz[cell];
  const double omega = 1.0 - 1.0 / (tau + (use_les ? r.taut[cell] : 0.0));
  double out[Q];
  if constexpr (Q == 27) {
    if (bgk) {
      bgk_post<L, double, true>(h, q.dr, q.ux, q.uy, q.uz, fx, fy, fz, 1.0 - omega, 1.0 - 0.5 * (1.0 - omega), 1.0,
                                out);
    } else {
      const Moments<double> m = moments_q27<double>(h);
      const NeqTensor<double> N = neq_q27<double>(m, q, fx, fy, fz);
      if (g.rr6)
        rr_post_q27_o6<double, true>(q, N, fx, fy, fz, omega, 1.0, 1.0, out);
      else
        rr_post_q27<double, true>(q, N, fx, fy, fz, omega, 1.0, 1.0, out);
    }
  } else {
    rr_post_generic<L, double>(h, q, fx, fy, fz, omega, 1.0, out);
  }
  using T = typename StoreT<S>::T;
  // the fused path's population layout
  char* d = reinterpret_cast<char*>(dst) + pop_cell_off(lay, cell, (uint32_t)sizeof(T));
  static_for<Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    double v = out[i];
    if constexpr (S == kF16) v = v * (f16_scale / L::w(i));
    *reinterpret_cast<T*>(d + (uint64_t)i * lay.seg) = T(v);
  });
}

// Velocity field in the fused path's compute type (for the WALE stencil).
template <typename C>
__global__ __launch_bounds__(kBlock) void k_copy_u(RefFields r, Grid g, C* u) {
  const uint32_t cell = tid();
  if (cell >= g.n) return;
  // AoS [N][3], the fused path's LES velocity layout
  u[3 * (size_t)cell] = C(r.ux[cell]);
  u[3 * (size_t)cell + 1] = C(r.uy[cell]);
  u[3 * (size_t)cell + 2] = C(r.uz[cell]);
}

// Per-plane statistics of output_profile (output.cu:182-215), one workgroup
// per y plane, FP64, two passes (means, then covariances of the fluctuations).
// out[iy*10 + k]: <u>,<v>,<w>,<uu>,<vv>,<ww>,<uv>,<uw>,<vw>,<tau_t>.
__global__ __launch_bounds__(kBlock) void k_plane_profiles(RefFields r, Grid g, int with_taut, double* out) {
  const uint32_t iy = blockIdx.x;
  __shared__ double red[kBlock];
  __shared__ double mean[4];
  const uint32_t np = g.lx * g.lz;
  auto block_sum = [&](double v) -> double {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    const double t = red[0];
    __syncthreads();
    return t;
  };
  double s[4] = {0, 0, 0, 0};
  for (uint32_t k = threadIdx.x; k < np; k += kBlock) {
    const uint32_t x = k % g.lx, z = k / g.lx;
    const uint32_t c = lin(x, iy, z, g);
    s[0] += r.ux[c]; s[1] += r.uy[c]; s[2] += r.uz[c];
    if (with_taut) s[3] += r.taut[c];
  }
  for (int k = 0; k < 4; ++k) {
    const double t = block_sum(s[k]);
    if (threadIdx.x == 0) mean[k] = t / np;
  }
  __syncthreads();
  double cv[6] = {0, 0, 0, 0, 0, 0};
  for (uint32_t k = threadIdx.x; k < np; k += kBlock) {
    const uint32_t x = k % g.lx, z = k / g.lx;
    const uint32_t c = lin(x, iy, z, g);
    const double ur = r.ux[c] - mean[0], vr = r.uy[c] - mean[1], wr = r.uz[c] - mean[2];
    cv[0] += ur * ur; cv[1] += vr * vr; cv[2] += wr * wr;
    cv[3] += ur * vr; cv[4] += ur * wr; cv[5] += vr * wr;
  }
  double res[6];
  for (int k = 0; k < 6; ++k) res[k] = block_sum(cv[k]) / np;
  if (threadIdx.x == 0) {
    double* o = out + (size_t)iy * 10;
    o[0] = mean[0]; o[1] = mean[1]; o[2] = mean[2];
    for (int k = 0; k < 6; ++k) o[3 + k] = res[k];
    o[9] = mean[3];
  }
}

// diag_flow sums (output.cu:24-35) over the reference fields, per-workgroup
// FP64 partials in the layout of the fused diagnostics pass.
__global__ __launch_bounds__(kBlock) void k_diag_ref(RefFields r, Grid g, double* partials) {
  const uint32_t cell = tid();
  const bool active = cell < g.n;
  double v[kDiagWidth];
  const double rho = active ? r.rho[cell] : 1.0;
  const double ux = active ? r.ux[cell] : 0.0, uy = active ? r.uy[cell] : 0.0, uz = active ? r.uz[cell] : 0.0;
  v[0] = rho - 1.0;
  v[1] = ux; v[2] = uy; v[3] = uz;
  v[4] = rho * ux; v[5] = rho * uy; v[6] = rho * uz;
  v[7] = 0.5 * rho * (ux * ux + uy * uy + uz * uz);
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
  if (lane == 0)
    for (int k = 0; k < kDiagWidth; ++k) red[wave][k] = v[k];
  __syncthreads();
  if (threadIdx.x < kDiagWidth) {
    double t = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) t += red[w][threadIdx.x];
    partials[(size_t)blockIdx.x * kDiagWidth + threadIdx.x] = t;
  }
}

// ---------------------------------------------------------------- launchers
static inline uint32_t grid_of(const Grid& g) { return (g.n + kBlock - 1) / kBlock; }
#define RRLBM_LAUNCH(kern, g, s, ...)                                                  \
  do {                                                                                \
    hipLaunchKernelGGL(kern, dim3(grid_of(g)), dim3(kBlock), 0, s, __VA_ARGS__);      \
    return hipGetLastError();                                                         \
  } while (0)

hipError_t ref_collision(int q, const RefFields& r, const Grid& g, double tau, int use_les, hipStream_t s) {
  if (q == 27) RRLBM_LAUNCH(k_collision<D3Q27>, g, s, r, g, tau, use_les);
  RRLBM_LAUNCH(k_collision<D3Q19>, g, s, r, g, tau, use_les);
}
hipError_t ref_streaming(int q, const double* f, double* ftmp, const Grid& g, hipStream_t s) {
  if (q == 27) RRLBM_LAUNCH(k_streaming<D3Q27>, g, s, f, ftmp, g);
  RRLBM_LAUNCH(k_streaming<D3Q19>, g, s, f, ftmp, g);
}
hipError_t ref_bb(int q, const double* f, double* ftmp, const Grid& g, hipStream_t s) {
  if (q == 27) RRLBM_LAUNCH(k_bb<D3Q27>, g, s, f, ftmp, g);
  RRLBM_LAUNCH(k_bb<D3Q19>, g, s, f, ftmp, g);
}
hipError_t ref_macrovar(int q, const RefFields& r, const Grid& g, int wall_only, hipStream_t s, int bgk) {
  if (q == 27) RRLBM_LAUNCH(k_macrovar<D3Q27>, g, s, r, g, wall_only, bgk);
  RRLBM_LAUNCH(k_macrovar<D3Q19>, g, s, r, g, wall_only, bgk);
}
hipError_t ref_collision_bgk(const RefFields& r, const Grid& g, double tau, hipStream_t s) {
  RRLBM_LAUNCH(k_collision_bgk, g, s, r, g, tau);
}
hipError_t ref_rrbc(int q, const RefFields& r, const double* fpost, const Grid& g, hipStream_t s) {
  if (q == 27) RRLBM_LAUNCH(k_rrbc<D3Q27>, g, s, r, fpost, g);
  RRLBM_LAUNCH(k_rrbc<D3Q19>, g, s, r, fpost, g);
}
hipError_t ref_fdbc(const RefFields& r, const Grid& g, double tau, int use_les, hipStream_t s) {
  const uint32_t nw = 2 * g.lx * g.lz;
  const dim3 grid((nw + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_fdbc_velocity, grid, dim3(kBlock), 0, s, r, g);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_fdbc_reconstruct, grid, dim3(kBlock), 0, s, r, g, tau, use_les);
  return hipGetLastError();
}
hipError_t ref_les_wale(const RefFields& r, const Grid& g, double cw, hipStream_t s) {
  RRLBM_LAUNCH(k_les_wale, g, s, r, g, cw);
}
hipError_t ref_forcing(const RefFields& r, const Grid& g, double fx, double fy, double fz, hipStream_t s) {
  RRLBM_LAUNCH(k_forcing, g, s, r, g, fx, fy, fz);
}
hipError_t ref_forcing_gp(const RefFields& r, const Grid& g, double force_in_x, int istep, double phase,
                          hipStream_t s) {
  RRLBM_LAUNCH(k_forcing_gp, g, s, r, g, force_in_x, istep, phase);
}
hipError_t ref_init_arrays(int q, const RefFields& r, const Grid& g, double ustar, double ystar, double fx,
                           hipStream_t s) {
  if (q == 27) RRLBM_LAUNCH(k_init_arrays<D3Q27>, g, s, r, g, ustar, ystar, fx);
  RRLBM_LAUNCH(k_init_arrays<D3Q19>, g, s, r, g, ustar, ystar, fx);
}
hipError_t ref_init_equilibrium(int q, const RefFields& r, const Grid& g, hipStream_t s, int bgk) {
  if (q == 27) RRLBM_LAUNCH(k_init_equilibrium<D3Q27>, g, s, r, g, bgk);
  RRLBM_LAUNCH(k_init_equilibrium<D3Q19>, g, s, r, g, bgk);
}
hipError_t ref_init_fields(const RefFields& r, const Grid& g, const InitSpec& spec, hipStream_t s) {
  RRLBM_LAUNCH(k_init_fields, g, s, r, g, spec);
}
hipError_t ref_pack(int storage, int q, const RefFields& r, const Grid& g, void* dst, const PopLayout& lay, double tau,
                    int use_les, double fx, double fy, double fz, double f16_scale, hipStream_t s, int bgk,
                    int cell_force) {
#define RRLBM_PACK(S, LT) \
  RRLBM_LAUNCH((k_pack<S, LT>), g, s, r, g, dst, lay, tau, use_les, fx, fy, fz, f16_scale, bgk, cell_force)
  if (q == 27) {
    if (storage == kF64) RRLBM_PACK(kF64, D3Q27);
    if (storage == kF32) RRLBM_PACK(kF32, D3Q27);
    RRLBM_PACK(kF16, D3Q27);
  }
  if (storage == kF64) RRLBM_PACK(kF64, D3Q19);
  if (storage == kF32) RRLBM_PACK(kF32, D3Q19);
  RRLBM_PACK(kF16, D3Q19);
#undef RRLBM_PACK
}
hipError_t ref_copy_u(int storage, const RefFields& r, const Grid& g, void* u, hipStream_t s) {
  if (storage == kF64) RRLBM_LAUNCH(k_copy_u<double>, g, s, r, g, reinterpret_cast<double*>(u));
  RRLBM_LAUNCH(k_copy_u<float>, g, s, r, g, reinterpret_cast<float*>(u));
}
hipError_t ref_plane_profiles(const RefFields& r, const Grid& g, int with_taut, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_plane_profiles, dim3(g.ly), dim3(kBlock), 0, s, r, g, with_taut, out);
  return hipGetLastError();
}
hipError_t ref_diag(const RefFields& r, const Grid& g, double* partials, hipStream_t s) {
  RRLBM_LAUNCH(k_diag_ref, g, s, r, g, partials);
}

}  // namespace rrlbm
