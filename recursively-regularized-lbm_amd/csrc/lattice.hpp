// Lattice velocity sets used by the RR-LBM hot path, as compile-time tables.
//
// D3Q27 follows the channel solver's ordering (RR27V33_code_DNS_LES/initial.cu:67-77):
// rest, 6 faces, 12 edges, 8 corners, with weights 8/27, 2/27, 1/54, 1/216.
// D3Q19 follows the Taylor-Green D3Q19 ordering (Taylor_Green/D3Q19RR.f90:75-81):
// rest, +x,-x,+y,-y,+z,-z, then 12 edges, with weights 1/3, 1/18, 1/36.
// Every table is reached through a constexpr function called with a template
// argument, so kernels see only folded constants (no constant-memory lookups,
// no local arrays).
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RRLBM_HD __host__ __device__
#define RRLBM_INLINE __attribute__((always_inline)) inline
#else
#define RRLBM_HD
#define RRLBM_INLINE inline
#endif

#include <utility>

namespace rrlbm {

// Compile-time loop: calls fn(std::integral_constant<int, I>) for I = 0..N-1.
template <typename Fn, int... Is>
RRLBM_HD RRLBM_INLINE void static_for_impl(Fn&& fn, std::integer_sequence<int, Is...>) {
  (fn(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename Fn>
RRLBM_HD RRLBM_INLINE void static_for(Fn&& fn) {
  static_for_impl(fn, std::make_integer_sequence<int, N>{});
}

struct D3Q27 {
  static constexpr int Q = 27;
  RRLBM_HD static constexpr int cx(int i) {
    constexpr int t[27] = {0, 1, 0, -1, 0, 0, 0, 1, -1, -1, 1, 1, 0, -1, 0, 1, 0, -1, 0, 1, -1, -1, 1, 1, -1, -1, 1};
    return t[i];
  }
  RRLBM_HD static constexpr int cy(int i) {
    constexpr int t[27] = {0, 0, 1, 0, -1, 0, 0, 1, 1, -1, -1, 0, 1, 0, -1, 0, 1, 0, -1, 1, 1, -1, -1, 1, 1, -1, -1};
    return t[i];
  }
  RRLBM_HD static constexpr int cz(int i) {
    constexpr int t[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1};
    return t[i];
  }
  RRLBM_HD static constexpr double w(int i) {
    return i == 0 ? 8.0 / 27.0 : (i <= 6 ? 2.0 / 27.0 : (i <= 18 ? 1.0 / 54.0 : 1.0 / 216.0));
  }
  RRLBM_HD static constexpr int opp(int i) {
    for (int j = 0; j < Q; ++j)
      if (cx(j) == -cx(i) && cy(j) == -cy(i) && cz(j) == -cz(i)) return j;
    return -1;
  }
  // Index of the velocity (x, y, z) with components in {-1, 0, 1}.
  RRLBM_HD static constexpr int index(int x, int y, int z) {
    for (int j = 0; j < Q; ++j)
      if (cx(j) == x && cy(j) == y && cz(j) == z) return j;
    return -1;
  }
};

struct D3Q19 {
  static constexpr int Q = 19;
  RRLBM_HD static constexpr int cx(int i) {
    constexpr int t[19] = {0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1, 0, 0, 0, 0};
    return t[i];
  }
  RRLBM_HD static constexpr int cy(int i) {
    constexpr int t[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1};
    return t[i];
  }
  RRLBM_HD static constexpr int cz(int i) {
    constexpr int t[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
    return t[i];
  }
  RRLBM_HD static constexpr double w(int i) {
    return i == 0 ? 1.0 / 3.0 : (i <= 6 ? 1.0 / 18.0 : 1.0 / 36.0);
  }
  RRLBM_HD static constexpr int opp(int i) {
    for (int j = 0; j < Q; ++j)
      if (cx(j) == -cx(i) && cy(j) == -cy(i) && cz(j) == -cz(i)) return j;
    return -1;
  }
  RRLBM_HD static constexpr int index(int x, int y, int z) {
    for (int j = 0; j < Q; ++j)
      if (cx(j) == x && cy(j) == y && cz(j) == z) return j;
    return -1;
  }
};

// Scaled Hermite polynomials of one velocity, cs = 1/sqrt(3)
// (RR27V33_code_DNS_LES/initial.cu:79-110): first order c/cs, second order
// c_a c_b/cs^2 - delta_ab, third order c_a^2 c_b/cs^3 - c_b/cs and c_x c_y c_z/cs^3.
struct Hermite {
  static constexpr double cs = 0.57735026918962576451;   // 1/sqrt(3)
  static constexpr double cs3 = cs / 3.0;                // cs^3
  RRLBM_HD static constexpr double h1(int c) { return c / cs; }
  RRLBM_HD static constexpr double h2(int ca, int cb, bool diag) { return ca * cb * 3.0 - (diag ? 1.0 : 0.0); }
  RRLBM_HD static constexpr double h3(int ca, int cb) { return (ca * ca * cb) / cs3 - cb / cs; }
  RRLBM_HD static constexpr double hxyz(int x, int y, int z) { return (x * y * z) / cs3; }
};

}  // namespace rrlbm
