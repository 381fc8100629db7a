// Declarations of the reference-layout (FP64) kernels and their launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"

namespace rrlbm {

struct Grid {
  uint32_t lx, ly, lz, n;
  FastDiv div_lx, div_lxly;
  uint32_t walls;  // bit 0 x faces, bit 1 y faces, bit 2 z faces
  int32_t lid_face;
  double lid_x, lid_y, lid_z;
  int32_t rr6;     // RR operators with the fourth- to sixth-order terms (D3Q27)
};

// The channel solver's device fields (RR27V33_code_DNS_LES/globals.cu:67-71).
struct RefFields {
  double* f;
  double *rho, *ux, *uy, *uz;
  double *fx, *fy, *fz;
  double* taut;
};

enum InitKind : int { kInitRest = 0, kInitTaylorGreen2D = 1, kInitTaylorGreen3D = 2, kInitChannel = 3 };

struct InitSpec {
  int32_t kind;
  double u0;          // TG amplitude, or u_tau for the channel
  double visc;        // channel: kinematic viscosity (y* = visc/u_tau)
  double amp_streak;  // channel streak amplitude relative to U(y)
  double amp_noise;   // channel noise amplitude relative to the bulk velocity
  double ubulk;       // bulk velocity of the mean profile
  uint64_t seed;
};

hipError_t ref_collision(int q, const RefFields& r, const Grid& g, double tau, int use_les, hipStream_t s);
hipError_t ref_streaming(int q, const double* f, double* ftmp, const Grid& g, hipStream_t s);
hipError_t ref_bb(int q, const double* f, double* ftmp, const Grid& g, hipStream_t s);
// bgk: the BGK solver's velocity u = j + F/2 (BGK/macrovar.cu:4-31)
hipError_t ref_macrovar(int q, const RefFields& r, const Grid& g, int wall_only, hipStream_t s, int bgk = 0);
// collision_BGK (BGK/collision.cu:3-33), D3Q27
hipError_t ref_collision_bgk(const RefFields& r, const Grid& g, double tau, hipStream_t s);
hipError_t ref_rrbc(int q, const RefFields& r, const double* fpost, const Grid& g, hipStream_t s);
// FDBC (BC.cu:129-309) at the y walls, D3Q27, ly >= 4
hipError_t ref_fdbc(const RefFields& r, const Grid& g, double tau, int use_les, hipStream_t s);
hipError_t ref_les_wale(const RefFields& r, const Grid& g, double cw, hipStream_t s);
hipError_t ref_forcing(const RefFields& r, const Grid& g, double fx, double fy, double fz, hipStream_t s);
hipError_t ref_forcing_gp(const RefFields& r, const Grid& g, double force_in_x, int istep, double phase,
                          hipStream_t s);
hipError_t ref_init_arrays(int q, const RefFields& r, const Grid& g, double ustar, double ystar, double fx,
                           hipStream_t s);
hipError_t ref_init_equilibrium(int q, const RefFields& r, const Grid& g, hipStream_t s, int bgk = 0);
hipError_t ref_init_fields(const RefFields& r, const Grid& g, const InitSpec& spec, hipStream_t s);
// cell_force: collide with the force fields instead of the uniform fx, fy, fz
hipError_t ref_pack(int storage, int q, const RefFields& r, const Grid& g, void* dst, const PopLayout& lay, double tau,
                    int use_les, double fx, double fy, double fz, double f16_scale, hipStream_t s, int bgk = 0,
                    int cell_force = 0);
hipError_t ref_copy_u(int storage, const RefFields& r, const Grid& g, void* u, hipStream_t s);
hipError_t ref_plane_profiles(const RefFields& r, const Grid& g, int with_taut, double* out, hipStream_t s);
hipError_t ref_diag(const RefFields& r, const Grid& g, double* partials, hipStream_t s);

}  // namespace rrlbm
