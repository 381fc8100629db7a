// C ABI of the gfx950 RR-LBM library (declarations and reference citations in
// include/rrlbm.h).  A context owns its device buffers and a HIP stream; every
// entry point validates its arguments, enqueues work on that stream and
// returns a status code with a per-context message -- the library never exits
// the process (the reference aborts through CHECK_CUDA_ERROR, LBMLES.h:89-97).
#include <errno.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "kernels.hpp"
#include "kernels_ref.hpp"
#include "rrlbm.h"

using namespace rrlbm;

struct rrlbm_ctx {
  rrlbm_params p;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint32_t n = 0;
  int q = 27;
  Grid grid{};
  RefFields ref{};
  double* f_temp = nullptr;
  void* pop_a = nullptr;  // post-collision populations read by the next step
  void* pop_b = nullptr;
  size_t pop_bytes = 0;   // bytes of the current layout
  size_t pop_cap = 0;     // bytes allocated per buffer (>= pop_bytes)
  PopLayout lay{};        // layout of pop_a / pop_b
  void* u_a = nullptr;  // LES velocity buffers (compute type)
  void* u_b = nullptr;
  void* u_wall = nullptr;  // FDBC: velocities of rows 1, 2, LY-2, LY-3 (compute type)
  // body force of the fused state's last collision, and what the force fields
  // hold: the uniform force or forcing_GP(istep, phase)
  struct Forcing {
    int gp = 0, istep = 0;
    double phase = 0.0;
  } force_cur, force_fields;
  double* partials = nullptr;
  double* diag_out = nullptr;
  double* profiles = nullptr;
  uint32_t nblocks = 0;
  int cus = 256;  // compute units of the device
  bool packed = false;
  uint32_t sr_step = 0;  // fused steps taken since creation (stochastic-rounding dither)
  uint32_t* d_step = nullptr;    // device step counter read by captured FP16 steps
  bool capture_dev_step = false;  // set while capturing FP16 stochastic-rounding steps
  hipGraph_t graph = nullptr;  // captured fused steps (rrlbm_graph_capture)
  hipGraphExec_t exec = nullptr;
  int graph_steps = 0;
  bool graph_dev_step = false;  // the captured steps read d_step (advanced by every replay)
  bool graph_swap_pop = false;  // a replay leaves the state in the other population buffer
  bool graph_swap_u = false;    // ... and the lagged LES velocity in the other velocity buffer
  void* graph_pop_a = nullptr;  // buffer the captured sequence starts from
  Fused2Args pass{};   // steps_per_pass 2: tiles and z chunks of the two-step launch
  bool ref_ok = true;  // reference-layout fields allocated
  size_t bytes = 0;
  std::string err;
};

namespace {

thread_local std::string g_create_error;

int fail(rrlbm_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c)
    c->err = buf;
  else
    g_create_error = buf;
  return code;
}

#define CK(ctx, call)                                                                                 \
  do {                                                                                                \
    hipError_t e_ = (call);                                                                           \
    if (e_ != hipSuccess) return fail((ctx), RRLBM_E_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define NEED(ctx)                                          \
  do {                                                     \
    if (!(ctx)) return fail(nullptr, RRLBM_E_INVALID, "null context"); \
  } while (0)

#define NEED_REF(ctx)                                                                          \
  do {                                                                                         \
    NEED(ctx);                                                                                 \
    if (!(ctx)->ref_ok) return fail((ctx), RRLBM_E_STATE, "reference fields were released"); \
  } while (0)

size_t storage_size(int s) { return s == RRLBM_STORE_F64 ? 8 : (s == RRLBM_STORE_F32 ? 4 : 2); }
size_t compute_size(int s) { return s == RRLBM_STORE_F64 ? 8 : 4; }

int dalloc(rrlbm_ctx* c, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(c, RRLBM_E_HIP, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
  c->bytes += bytes;
  e = hipMemsetAsync(*p, 0, bytes, c->stream);
  if (e != hipSuccess) return fail(c, RRLBM_E_HIP, "hipMemsetAsync failed: %s", hipGetErrorString(e));
  return RRLBM_OK;
}

void dfree(void* p) {
  if (p) (void)hipFree(p);
}

double* field_ptr(rrlbm_ctx* c, rrlbm_field f, size_t* count) {
  const size_t n = c->n, qn = (size_t)c->q * c->n;
  switch (f) {
    case RRLBM_FIELD_F: *count = qn; return c->ref.f;
    case RRLBM_FIELD_F_TEMP: *count = qn; return c->f_temp;
    case RRLBM_FIELD_RHO: *count = n; return c->ref.rho;
    case RRLBM_FIELD_UX: *count = n; return c->ref.ux;
    case RRLBM_FIELD_UY: *count = n; return c->ref.uy;
    case RRLBM_FIELD_UZ: *count = n; return c->ref.uz;
    case RRLBM_FIELD_FORCEX: *count = n; return c->ref.fx;
    case RRLBM_FIELD_FORCEY: *count = n; return c->ref.fy;
    case RRLBM_FIELD_FORCEZ: *count = n; return c->ref.fz;
    case RRLBM_FIELD_TAUT: *count = n; return c->ref.taut;
  }
  *count = 0;
  return nullptr;
}

bool has_walls(const rrlbm_params& p) { return p.wall_x || p.wall_y || p.wall_z; }
bool is_bgk(const rrlbm_params& p) { return p.collision_model == RRLBM_COLLISION_BGK; }
bool is_rr6(const rrlbm_params& p) { return p.collision_model == RRLBM_COLLISION_RR && p.rr_order == 6; }

void set_blocks(FusedArgs& a, uint32_t nblocks, int remap) {
  a.nblocks = nblocks;
  a.blocks_per_xcd = (nblocks + 7) / 8;
  // a run at least as long as one XCD's share is one contiguous slab per XCD
  if (remap > 1 && (uint32_t)remap >= a.blocks_per_xcd) remap = 1;
  a.xcd_remap = (uint32_t)remap;
  a.div_remap = make_fastdiv(remap > 1 ? (uint32_t)remap : 1u);
  if (remap == 0)
    a.tiles = nblocks;
  else if (remap == 1)
    a.tiles = 8u * a.blocks_per_xcd;
  else  // whole rounds of 8 runs of `remap` blocks
    a.tiles = (nblocks + 8u * remap - 1) / (8u * remap) * (8u * remap);
}

// Population layout from pop_group_rows / pop_pad_bytes; writes the values used
// back into p.  Default: groups of about 6144 cells (whole rows, no padding)
// when rows are whole wavefronts (LX a multiple of 64); otherwise one group
// (structure of arrays) with 2 KiB between the arrays.  A layout of several
// groups must stay below 4 GiB per buffer (32-bit lane offsets).
int resolve_layout(rrlbm_params& p, PopLayout* out) {
  const uint32_t rows = (uint32_t)p.ly * (uint32_t)p.lz;
  const uint32_t sz = (uint32_t)storage_size(p.storage);
  const int q = p.npop;
  if (p.pop_group_rows < -1) return fail(nullptr, RRLBM_E_INVALID, "pop_group_rows must be -1, 0 or positive");
  if (p.pop_pad_bytes < -1 || p.pop_pad_bytes > (1 << 20) || (p.pop_pad_bytes > 0 && p.pop_pad_bytes % 16))
    return fail(nullptr, RRLBM_E_INVALID, "pop_pad_bytes must be -1 or a multiple of 16 in [0, 1 MiB]");
  uint32_t grows;
  if (p.pop_group_rows == -1)
    grows = rows;
  else if (p.pop_group_rows == 0)
    grows = p.lx % 64 == 0 ? std::max(1u, std::min(rows, (6144u + (uint32_t)p.lx / 2) / (uint32_t)p.lx)) : rows;
  else
    grows = std::min<uint32_t>((uint32_t)p.pop_group_rows, rows);
  uint32_t pad;
  if (p.pop_pad_bytes >= 0)
    pad = (uint32_t)p.pop_pad_bytes;
  else
    pad = grows == rows ? 2048u : 0u;
  PopLayout l = make_layout((uint32_t)p.lx, rows, q, sz, grows, pad);
  if (grows < rows && l.bytes >= (1ull << 32)) {
    if (p.pop_group_rows > 0)
      return fail(nullptr, RRLBM_E_INVALID, "a population buffer of %d-row groups would exceed 4 GiB",
                  p.pop_group_rows);
    grows = rows;  // default: fall back to one group
    l = make_layout((uint32_t)p.lx, rows, q, sz, grows, pad);
  }
  p.pop_group_rows = grows == rows ? -1 : (int32_t)grows;
  p.pop_pad_bytes = (int32_t)pad;
  *out = l;
  return RRLBM_OK;
}

FusedArgs fused_args(rrlbm_ctx* c) {
  FusedArgs a{};
  a.src = c->pop_a;
  a.dst = c->pop_b;
  a.lx = c->grid.lx; a.ly = c->grid.ly; a.lz = c->grid.lz; a.n = c->n;
  a.div_lx = c->grid.div_lx; a.div_lxly = c->grid.div_lxly;
  a.walls = c->grid.walls;
  a.lid_face = c->p.lid_face;
  a.wall_rr = c->p.wall_rr;
  a.fx = c->p.force[0]; a.fy = c->p.force[1]; a.fz = c->p.force[2];
  a.tau = c->p.tau; a.cw = c->p.cw;
  a.omega = 1.0 - 1.0 / c->p.tau;
  a.lid_x = c->p.lid_u[0]; a.lid_y = c->p.lid_u[1]; a.lid_z = c->p.lid_u[2];
  a.f16_scale = c->p.f16_scale;
  a.f16_unscale = (float)(1.0 / c->p.f16_scale);
  a.f16_round = (uint32_t)c->p.f16_rounding;
  a.sr_step = c->sr_step;
  a.sr_base = c->capture_dev_step ? c->d_step : nullptr;
  set_blocks(a, c->nblocks, c->p.xcd_remap);
  a.ref_f = c->ref.f; a.ref_rho = c->ref.rho; a.ref_ux = c->ref.ux; a.ref_uy = c->ref.uy; a.ref_uz = c->ref.uz;
  a.partials = c->partials;
  a.u_wall = c->u_wall;
  // the step this launch completes collided under force_cur; the collision of
  // a plain step takes the uniform force (fused_steps overrides part 1)
  a.gp_on[0] = c->force_cur.gp;
  a.gp_istep[0] = c->force_cur.istep;
  a.gp_phase[0] = c->force_cur.phase;
  a.gp_on[1] = 0;
  a.gp = a.gp_on[0];
  a.lay = c->lay;
  return a;
}

// Configurations the two-step kernel covers.
bool two_step_ok(const rrlbm_params& p) {
  return p.npop == 27 && !p.use_les && p.wall_rr != 2 && p.storage != RRLBM_STORE_F16 && p.lx % (int)kT2X == 0 &&
         p.ly % (int)kT2Y == 0;
}

// Geometry of the two-step launch: tiles of kT2X x kT2Y cells, z chunks of
// pass_planes planes (default 16: on the 384 x 128 x 192 channel the fastest of
// 8-48, nine work items per compute unit), the items dealt to the 8 XCDs in
// contiguous runs.
Fused2Args pass_geometry(const rrlbm_params& p) {
  Fused2Args t{};
  t.ntx = (uint32_t)p.lx / kT2X;
  t.nty = (uint32_t)p.ly / kT2Y;
  const uint32_t tiles = t.ntx * t.nty, lz = (uint32_t)p.lz;
  const uint32_t zc = p.pass_planes > 0 ? (uint32_t)p.pass_planes : 16u;
  t.zc = std::min(zc, lz);
  t.nzc = (lz + t.zc - 1) / t.zc;
  t.items = tiles * t.nzc;
  t.per_xcd = (t.items + 7) / 8;
  return t;
}

// Workgroups of a step launch over a.tiles blocks: one per block, or a
// persistent grid of persistent_wg_per_cu workgroups per compute unit striding
// over them.  Only the step kernels stride; the macro, finish and diagnostics
// passes always launch one workgroup per block (launch_grid).
uint32_t fused_grid(const rrlbm_ctx* c, const FusedArgs& a) {
  const int k = c->p.persistent_wg_per_cu;
  if (k <= 0) return a.tiles;
  return std::max(1u, std::min<uint32_t>(a.tiles, (uint32_t)k * (uint32_t)c->cus));
}

uint32_t launch_grid(const rrlbm_ctx* c, int mode, const FusedArgs& a) {
  return mode == kStep ? fused_grid(c, a) : a.tiles;
}

int launch(rrlbm_ctx* c, int mode, const FusedArgs& a) {
  const bool les = c->p.use_les && mode == kStep;
  int op = is_bgk(c->p) ? kOpBGK : (is_rr6(c->p) ? kOpRR6 : kOpRR);
  if (c->p.wall_rr == 2) {
    // FDBC: the velocities of rows 1, 2, LY-2, LY-3 from the same source first
    hipError_t e = launch_fdbc_rows(c->p.storage, a, c->stream);
    if (e != hipSuccess) return fail(c, RRLBM_E_HIP, "FDBC row pass launch failed: %s", hipGetErrorString(e));
    op = kOpRRFD;
  }
  const uint32_t grid = launch_grid(c, mode, a);
  hipError_t e = op != kOpRR ? launch_fused_op(op, c->p.storage, has_walls(c->p), mode, a, grid, c->stream)
                             : launch_fused(c->p.storage, c->q, has_walls(c->p), les, mode, a, grid, c->stream);
  if (e != hipSuccess) return fail(c, RRLBM_E_HIP, "fused kernel launch failed: %s", hipGetErrorString(e));
  return RRLBM_OK;
}

// Two cells per thread for the plain D3Q27 step only when requested: on
// MI355X one cell per thread is faster for every storage (higher occupancy).
bool use_pair(const rrlbm_ctx* c) {
  if (c->q != 27 || c->p.use_les || is_bgk(c->p) || is_rr6(c->p) || c->p.wall_rr == 2) return false;
  return c->p.cells_per_thread == 2;
}

int launch_step(rrlbm_ctx* c, const FusedArgs& a) {
  if (use_pair(c) && !a.gp) {
    FusedArgs p = a;
    set_blocks(p, (c->n + 2 * kBlock - 1) / (2 * kBlock), c->p.xcd_remap);
    hipError_t e = launch_fused_pair(c->p.storage, has_walls(c->p), p, fused_grid(c, p), c->stream);
    if (e != hipSuccess) return fail(c, RRLBM_E_HIP, "fused kernel launch failed: %s", hipGetErrorString(e));
    return RRLBM_OK;
  }
  return launch(c, kStep, a);
}

// next: the force of the collided steps (nullptr: the uniform force)
int fused_steps(rrlbm_ctx* c, int nsteps, const rrlbm_ctx::Forcing* next = nullptr) {
  for (int s = 0; s < nsteps; ++s) {
    if (c->p.steps_per_pass == 2 && !next && !c->force_cur.gp && s + 2 <= nsteps) {
      // two steps in one launch (the step it completes and both it collides
      // take the uniform force)
      const FusedArgs a = fused_args(c);
      const int op = is_bgk(c->p) ? kOpBGK : (is_rr6(c->p) ? kOpRR6 : kOpRR);
      hipError_t e = launch_fused2(c->p.storage, has_walls(c->p), op, a, c->pass, c->stream);
      if (e != hipSuccess) return fail(c, RRLBM_E_HIP, "two-step kernel launch failed: %s", hipGetErrorString(e));
      c->sr_step += 2;
      std::swap(c->pop_a, c->pop_b);
      ++s;
      continue;
    }
    FusedArgs a = fused_args(c);
    if (next && next->gp) {
      a.gp_on[1] = 1;
      a.gp_istep[1] = next->istep;
      a.gp_phase[1] = next->phase;
      a.gp = 1;
    }
    c->force_cur = next ? *next : rrlbm_ctx::Forcing{};
    if (c->p.use_les) {
      if (c->p.les_lagged) {
        a.u_in = c->u_a;
        a.u_out = c->u_b;
        a.write_u = 1;
      } else {
        FusedArgs m = a;
        m.u_out = c->u_a;
        int rc = launch(c, kMacro, m);
        if (rc) return rc;
        a.u_in = c->u_a;
        a.write_u = 0;
      }
    }
    int rc = c->p.use_les ? launch(c, kStep, a) : launch_step(c, a);
    if (rc) return rc;
    ++c->sr_step;
    std::swap(c->pop_a, c->pop_b);
    if (c->p.use_les && c->p.les_lagged) std::swap(c->u_a, c->u_b);
  }
  return RRLBM_OK;
}

// Forget a captured graph: its kernel arguments hold the population buffers
// and the layout of the time of capture.
void drop_graph(rrlbm_ctx* c) {
  if (c->exec) (void)hipGraphExecDestroy(c->exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->exec = nullptr;
  c->graph = nullptr;
  c->graph_steps = 0;
  c->graph_pop_a = nullptr;
  c->graph_dev_step = false;
  c->graph_swap_pop = false;
  c->graph_swap_u = false;
}

bool f16_dither(const rrlbm_params& p) {
  return p.storage == RRLBM_STORE_F16 && p.f16_rounding == RRLBM_F16_STOCHASTIC;
}

void diag_from_sums(const double (&h)[kDiagWidth], double ncells, rrlbm_diag* out) {
  out->mass_excess = h[0];
  out->u_sum[0] = h[1]; out->u_sum[1] = h[2]; out->u_sum[2] = h[3];
  out->momentum[0] = h[4]; out->momentum[1] = h[5]; out->momentum[2] = h[6];
  out->kinetic_energy = h[7];
  out->nonfinite_cells = h[8];
  out->ncells = ncells;
}

int finalize_diag(rrlbm_ctx* c, uint32_t nparts, rrlbm_diag* out) {
  CK(c, launch_diag_finalize(c->partials, nparts, c->diag_out, c->stream));
  double h[kDiagWidth];
  CK(c, hipMemcpyAsync(h, c->diag_out, sizeof h, hipMemcpyDeviceToHost, c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  diag_from_sums(h, (double)c->n, out);
  if (h[8] != 0.0) return fail(c, RRLBM_E_NUMERIC, "%.0f cells hold non-finite values", h[8]);
  return RRLBM_OK;
}

int make_dir(rrlbm_ctx* c, const char* dir) {
  struct stat st;
  if (stat(dir, &st) == 0) return RRLBM_OK;
  if (mkdir(dir, 0755) != 0 && errno != EEXIST) return fail(c, RRLBM_E_IO, "cannot create directory %s: %s", dir, strerror(errno));
  return RRLBM_OK;
}

int write_fields(rrlbm_ctx* c, const char* path, double* const* dev, int nf, size_t count) {
  FILE* fp = fopen(path, "wb");
  if (!fp) return fail(c, RRLBM_E_IO, "cannot open %s: %s", path, strerror(errno));
  std::string buf(count * sizeof(double), '\0');
  for (int k = 0; k < nf; ++k) {
    hipError_t e = hipMemcpyAsync(&buf[0], dev[k], count * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      fclose(fp);
      return fail(c, RRLBM_E_HIP, "device to host copy failed: %s", hipGetErrorString(e));
    }
    if (fwrite(buf.data(), sizeof(double), count, fp) != count) {
      fclose(fp);
      return fail(c, RRLBM_E_IO, "short write to %s", path);
    }
  }
  if (fclose(fp) != 0) return fail(c, RRLBM_E_IO, "cannot close %s: %s", path, strerror(errno));
  return RRLBM_OK;
}

int read_fields(rrlbm_ctx* c, const char* path, double* const* dev, int nf, size_t count) {
  FILE* fp = fopen(path, "rb");
  if (!fp) return fail(c, RRLBM_E_IO, "cannot open %s: %s", path, strerror(errno));
  std::string buf(count * sizeof(double), '\0');
  for (int k = 0; k < nf; ++k) {
    size_t got = fread(&buf[0], sizeof(double), count, fp);
    if (got != count) {
      fclose(fp);
      return fail(c, RRLBM_E_IO, "short read from %s: %zu of %zu values", path, got, count);
    }
    hipError_t e = hipMemcpyAsync(dev[k], buf.data(), count * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      fclose(fp);
      return fail(c, RRLBM_E_HIP, "host to device copy failed: %s", hipGetErrorString(e));
    }
  }
  fclose(fp);
  return RRLBM_OK;
}

}  // namespace

extern "C" {

const char* rrlbm_version(void) { return "rrlbm-gfx950 1.1 (abi 1)"; }

int rrlbm_params_default(rrlbm_params* p) {
  if (!p) return fail(nullptr, RRLBM_E_INVALID, "null params");
  memset(p, 0, sizeof *p);
  p->lx = p->ly = p->lz = 64;  // parameters.h:9-11
  p->npop = 27;                // parameters.h:13
  p->use_les = 0;              // parameters.h:6
  p->les_lagged = 1;
  p->storage = RRLBM_STORE_F32;
  p->wall_y = 1;
  p->wall_rr = 1;
  p->lid_face = -1;
  const double Re = 180.0, visc = 0.000622;  // globals.cu:57-58
  const double u0 = 2.0 * Re * visc / p->ly;   // globals.cu:59
  p->tau = 3.0 * visc + 0.5;                   // globals.cu:60
  p->force[0] = 2.0 * u0 * u0 / p->ly;         // globals.cu:63
  p->cw = 0.5;                                 // LES.cu:12
  p->f16_scale = 4096.0;
  p->xcd_remap = -1;  // library default: runs of 32 blocks per XCD on grids of >= 2048 blocks
  p->pop_group_rows = 0;
  p->pop_pad_bytes = -1;
  return RRLBM_OK;
}

int rrlbm_lattice(int npop, int32_t* c, double* w, int32_t* opp) {
  if (npop != 27 && npop != 19) return fail(nullptr, RRLBM_E_INVALID, "npop must be 27 or 19");
  auto fill = [&](auto lat) {
    using L = decltype(lat);
    for (int i = 0; i < L::Q; ++i) {
      if (c) { c[3 * i] = L::cx(i); c[3 * i + 1] = L::cy(i); c[3 * i + 2] = L::cz(i); }
      if (w) w[i] = L::w(i);
      if (opp) opp[i] = L::opp(i);
    }
  };
  if (npop == 27) fill(D3Q27{});
  else fill(D3Q19{});
  return RRLBM_OK;
}

int rrlbm_device_count(int* n) {
  if (!n) return fail(nullptr, RRLBM_E_INVALID, "null output");
  hipError_t e = hipGetDeviceCount(n);
  if (e != hipSuccess) {
    *n = 0;
    return fail(nullptr, RRLBM_E_HIP, "hipGetDeviceCount failed: %s", hipGetErrorString(e));
  }
  return RRLBM_OK;
}

int rrlbm_create(const rrlbm_params* pin, rrlbm_ctx** out) {
  if (!pin || !out) return fail(nullptr, RRLBM_E_INVALID, "null argument");
  *out = nullptr;
  rrlbm_params p = *pin;
  if (p.lx < 1 || p.ly < 1 || p.lz < 1) return fail(nullptr, RRLBM_E_INVALID, "grid dimensions must be >= 1");
  const uint64_t n64 = (uint64_t)p.lx * p.ly * p.lz;
  if (n64 >= (1ull << 31)) return fail(nullptr, RRLBM_E_INVALID, "grid of %llu cells exceeds 2^31", (unsigned long long)n64);
  if (p.npop != 27 && p.npop != 19) return fail(nullptr, RRLBM_E_INVALID, "npop must be 27 or 19, got %d", p.npop);
  if (p.use_les && p.npop != 27) return fail(nullptr, RRLBM_E_UNSUPPORTED, "WALE LES is provided for D3Q27 only");
  if (p.storage < 0 || p.storage > 2) return fail(nullptr, RRLBM_E_INVALID, "storage must be 0 (F64), 1 (F32) or 2 (F16)");
  if (!(p.tau > 0.5)) return fail(nullptr, RRLBM_E_INVALID, "tau must exceed 0.5, got %g", p.tau);
  if ((p.wall_x && p.lx < 2) || (p.wall_y && p.ly < 2) || (p.wall_z && p.lz < 2))
    return fail(nullptr, RRLBM_E_INVALID, "a walled axis needs at least 2 cells");
  if (p.lid_face < -1 || p.lid_face > 5) return fail(nullptr, RRLBM_E_INVALID, "lid_face must be -1 or 0..5");
  if (p.lid_face >= 0) {
    const int ax = p.lid_face / 2;
    const int walled = ax == 0 ? p.wall_x : (ax == 1 ? p.wall_y : p.wall_z);
    if (!walled) return fail(nullptr, RRLBM_E_INVALID, "lid_face %d lies on a periodic axis", p.lid_face);
  }
  if (p.f16_scale == 0.0) p.f16_scale = 4096.0;
  if (p.xcd_remap < -1 || p.xcd_remap > 65536)
    return fail(nullptr, RRLBM_E_INVALID, "xcd_remap must be -1, 0, 1 or a run length up to 65536");
  // default block order: runs of 32 blocks per XCD keep the blocks that share
  // cache lines in one L2 (measured: FP16 channel reads 1.06x instead of 1.23x
  // the algorithmic bytes); small grids keep the linear order
  if (p.xcd_remap == -1) p.xcd_remap = (n64 + kBlock - 1) / kBlock >= 2048 ? 32 : 0;
  if (p.collision_model != RRLBM_COLLISION_RR && p.collision_model != RRLBM_COLLISION_BGK)
    return fail(nullptr, RRLBM_E_INVALID, "collision_model must be 0 (RR) or 1 (BGK)");
  if (p.collision_model == RRLBM_COLLISION_BGK && (p.npop != 27 || p.use_les || p.wall_rr))
    return fail(nullptr, RRLBM_E_UNSUPPORTED, "the BGK baseline is D3Q27 without LES and without RR wall regularisation");
  if (p.wall_rr < 0 || p.wall_rr > 2) return fail(nullptr, RRLBM_E_INVALID, "wall_rr must be 0, 1 (RRBC) or 2 (FDBC)");
  if (p.wall_rr == 2 && (p.npop != 27 || !p.wall_y || p.wall_x || p.wall_z || p.lid_face >= 0 || p.ly < 4 ||
                         p.collision_model != RRLBM_COLLISION_RR || p.rr_order == 6))
    return fail(nullptr, RRLBM_E_UNSUPPORTED,
                "FDBC (wall_rr 2) is the D3Q27 third-order RR channel's y-wall treatment (y walls only, ly >= 4)");
  if (p.rr_order == 0) p.rr_order = 3;
  if (p.rr_order != 3 && p.rr_order != 6) return fail(nullptr, RRLBM_E_INVALID, "rr_order must be 0, 3 or 6");
  if (p.rr_order == 6 && (p.collision_model != RRLBM_COLLISION_RR || p.npop != 27 || p.use_les))
    return fail(nullptr, RRLBM_E_UNSUPPORTED, "rr_order 6 is the D3Q27 RR operator without LES");
  if (p.cells_per_thread < 0 || p.cells_per_thread > 2)
    return fail(nullptr, RRLBM_E_INVALID, "cells_per_thread must be 0, 1 or 2");
  if (p.persistent_wg_per_cu < 0 || p.persistent_wg_per_cu > 64)
    return fail(nullptr, RRLBM_E_INVALID, "persistent_wg_per_cu must be in [0, 64]");
  if (p.steps_per_pass < 0 || p.steps_per_pass > 2) return fail(nullptr, RRLBM_E_INVALID, "steps_per_pass must be 0, 1 or 2");
  if (p.steps_per_pass == 0) p.steps_per_pass = 1;
  if (p.steps_per_pass == 2 && !two_step_ok(p))
    return fail(nullptr, RRLBM_E_UNSUPPORTED,
                "two steps per pass: the D3Q27 RR or BGK step without LES or FDBC, FP32 or FP64 storage, "
                "lx a multiple of %u and ly of %u", kT2X, kT2Y);
  if (p.pass_planes < 0) return fail(nullptr, RRLBM_E_INVALID, "pass_planes must be >= 0");
  if (p.f16_rounding != RRLBM_F16_STOCHASTIC && p.f16_rounding != RRLBM_F16_NEAREST)
    return fail(nullptr, RRLBM_E_INVALID, "f16_rounding must be 0 (stochastic) or 1 (nearest)");
  if (!(p.f16_scale > 0.0)) return fail(nullptr, RRLBM_E_INVALID, "f16_scale must be positive");
  const uint64_t per_pop = n64 * storage_size(p.storage);
  if (per_pop >= (1ull << 32)) return fail(nullptr, RRLBM_E_INVALID, "one population array exceeds 4 GiB");
  PopLayout lay;
  {
    int rc = resolve_layout(p, &lay);
    if (rc) return rc;
  }

  rrlbm_ctx* c = new rrlbm_ctx();
  c->p = p;
  c->n = (uint32_t)n64;
  c->q = p.npop;
  c->grid.lx = p.lx; c->grid.ly = p.ly; c->grid.lz = p.lz; c->grid.n = c->n;
  c->grid.div_lx = make_fastdiv(p.lx);
  c->grid.div_lxly = make_fastdiv((uint32_t)p.lx * p.ly);
  c->grid.walls = (p.wall_x ? 1u : 0u) | (p.wall_y ? 2u : 0u) | (p.wall_z ? 4u : 0u);
  c->grid.lid_face = p.lid_face;
  c->grid.lid_x = p.lid_u[0]; c->grid.lid_y = p.lid_u[1]; c->grid.lid_z = p.lid_u[2];
  c->grid.rr6 = is_rr6(p) ? 1 : 0;
  c->nblocks = (c->n + kBlock - 1) / kBlock;
  c->lay = lay;

  auto bail = [&](int rc) {
    g_create_error = c->err;
    rrlbm_destroy(c);
    return rc;
  };
  hipError_t e = hipGetDevice(&c->device);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, c->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e != hipSuccess) return bail(fail(c, RRLBM_E_HIP, "HIP setup failed: %s", hipGetErrorString(e)));
  if (p.steps_per_pass == 2) c->pass = pass_geometry(p);

  const size_t n = c->n, qn = (size_t)c->q * n;
  int rc = RRLBM_OK;
  void** slots[] = {(void**)&c->ref.f, (void**)&c->f_temp};
  for (void** s : slots)
    if ((rc = dalloc(c, s, qn * sizeof(double)))) return bail(rc);
  void** fields[] = {(void**)&c->ref.rho, (void**)&c->ref.ux, (void**)&c->ref.uy, (void**)&c->ref.uz,
                     (void**)&c->ref.fx,  (void**)&c->ref.fy, (void**)&c->ref.fz, (void**)&c->ref.taut};
  for (void** s : fields)
    if ((rc = dalloc(c, s, n * sizeof(double)))) return bail(rc);
  c->pop_bytes = c->pop_cap = lay.bytes;
  if ((rc = dalloc(c, &c->pop_a, c->pop_bytes))) return bail(rc);
  if ((rc = dalloc(c, &c->pop_b, c->pop_bytes))) return bail(rc);
  if (p.wall_rr == 2)
    if ((rc = dalloc(c, &c->u_wall, 12 * (size_t)p.lx * p.lz * compute_size(p.storage)))) return bail(rc);
  if (p.use_les) {
    if ((rc = dalloc(c, &c->u_a, 3 * n * compute_size(p.storage)))) return bail(rc);
    if ((rc = dalloc(c, &c->u_b, 3 * n * compute_size(p.storage)))) return bail(rc);
  }
  if ((rc = dalloc(c, (void**)&c->partials, (size_t)c->nblocks * kDiagWidth * sizeof(double)))) return bail(rc);
  if ((rc = dalloc(c, (void**)&c->diag_out, kDiagWidth * sizeof(double)))) return bail(rc);
  if ((rc = dalloc(c, (void**)&c->d_step, 256))) return bail(rc);
  if ((rc = dalloc(c, (void**)&c->profiles, (size_t)p.ly * 10 * sizeof(double)))) return bail(rc);
  // rho = 1 and the uniform force, as initialize() leaves them
  e = ref_forcing(c->ref, c->grid, p.force[0], p.force[1], p.force[2], c->stream);
  if (e == hipSuccess) e = ref_init_fields(c->ref, c->grid, InitSpec{kInitRest, 0, 0, 0, 0, 0, 0}, c->stream);
  if (e == hipSuccess) e = ref_init_equilibrium(c->q, c->ref, c->grid, c->stream, is_bgk(c->p));
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return bail(fail(c, RRLBM_E_HIP, "initial fill failed: %s", hipGetErrorString(e)));
  *out = c;
  return RRLBM_OK;
}

int rrlbm_destroy(rrlbm_ctx* c) {
  if (!c) return RRLBM_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void* ptrs[] = {c->ref.f, c->f_temp, c->ref.rho, c->ref.ux, c->ref.uy, c->ref.uz, c->ref.fx, c->ref.fy,
                  c->ref.fz, c->ref.taut, c->pop_a, c->pop_b, c->u_a, c->u_b, c->u_wall, c->partials,
                  c->diag_out, c->profiles, c->d_step};
  for (void* p : ptrs) dfree(p);
  if (c->exec) (void)hipGraphExecDestroy(c->exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return RRLBM_OK;
}

const char* rrlbm_last_error(const rrlbm_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int rrlbm_get_params(const rrlbm_ctx* c, rrlbm_params* out) {
  if (!c || !out) return fail(nullptr, RRLBM_E_INVALID, "null argument");
  *out = c->p;
  return RRLBM_OK;
}

int rrlbm_synchronize(rrlbm_ctx* c) {
  NEED(c);
  CK(c, hipStreamSynchronize(c->stream));
  return RRLBM_OK;
}

int rrlbm_memory_bytes(const rrlbm_ctx* c, size_t* bytes) {
  if (!c || !bytes) return fail(nullptr, RRLBM_E_INVALID, "null argument");
  *bytes = c->bytes;
  return RRLBM_OK;
}

int rrlbm_release_reference(rrlbm_ctx* c) {
  NEED(c);
  if (!c->ref_ok) return RRLBM_OK;
  CK(c, hipStreamSynchronize(c->stream));
  void* ptrs[] = {c->ref.f, c->f_temp, c->ref.rho, c->ref.ux, c->ref.uy, c->ref.uz,
                  c->ref.fx, c->ref.fy, c->ref.fz, c->ref.taut};
  for (void* p : ptrs) dfree(p);
  c->bytes -= ((size_t)2 * c->q + 8) * c->n * sizeof(double);
  c->ref = RefFields{};
  c->f_temp = nullptr;
  c->ref_ok = false;
  return RRLBM_OK;
}

int rrlbm_upload(rrlbm_ctx* c, rrlbm_field f, const double* host, size_t count) {
  NEED_REF(c);
  size_t want = 0;
  double* d = field_ptr(c, f, &want);
  if (!d) return fail(c, RRLBM_E_INVALID, "unknown field %d", (int)f);
  if (!host || count != want) return fail(c, RRLBM_E_INVALID, "field %d holds %zu values, got %zu", (int)f, want, count);
  CK(c, hipMemcpyAsync(d, host, count * sizeof(double), hipMemcpyHostToDevice, c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  // uploaded force fields are the caller's: no forcing_GP state goes with them
  if (f == RRLBM_FIELD_FORCEX || f == RRLBM_FIELD_FORCEY || f == RRLBM_FIELD_FORCEZ)
    c->force_fields = rrlbm_ctx::Forcing{};
  return RRLBM_OK;
}

int rrlbm_download(rrlbm_ctx* c, rrlbm_field f, double* host, size_t count) {
  NEED_REF(c);
  size_t want = 0;
  double* d = field_ptr(c, f, &want);
  if (!d) return fail(c, RRLBM_E_INVALID, "unknown field %d", (int)f);
  if (!host || count != want) return fail(c, RRLBM_E_INVALID, "field %d holds %zu values, got %zu", (int)f, want, count);
  CK(c, hipMemcpyAsync(host, d, count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  return RRLBM_OK;
}

int rrlbm_init_arrays(rrlbm_ctx* c, double ustar, double ystar) {
  NEED_REF(c);
  if (!(ystar > 0.0)) return fail(c, RRLBM_E_INVALID, "ystar must be positive");
  CK(c, ref_init_arrays(c->q, c->ref, c->grid, ustar, ystar, c->p.force[0], c->stream));
  c->packed = false;
  return RRLBM_OK;
}

int rrlbm_forcing(rrlbm_ctx* c) {
  NEED_REF(c);
  CK(c, ref_forcing(c->ref, c->grid, c->p.force[0], c->p.force[1], c->p.force[2], c->stream));
  c->force_fields = rrlbm_ctx::Forcing{};
  return RRLBM_OK;
}

int rrlbm_forcing_gp(rrlbm_ctx* c, int istep, double rand_phase) {
  NEED_REF(c);
  CK(c, ref_forcing_gp(c->ref, c->grid, c->p.force[0], istep, rand_phase, c->stream));
  c->force_fields = rrlbm_ctx::Forcing{1, istep, rand_phase};
  return RRLBM_OK;
}

int rrlbm_les_wale(rrlbm_ctx* c) {
  NEED_REF(c);
  CK(c, ref_les_wale(c->ref, c->grid, c->p.cw, c->stream));
  return RRLBM_OK;
}

int rrlbm_collision_rr(rrlbm_ctx* c) {
  NEED_REF(c);
  CK(c, ref_collision(c->q, c->ref, c->grid, c->p.tau, c->p.use_les, c->stream));
  return RRLBM_OK;
}

int rrlbm_streaming(rrlbm_ctx* c) {
  NEED_REF(c);
  CK(c, ref_streaming(c->q, c->ref.f, c->f_temp, c->grid, c->stream));
  return RRLBM_OK;
}

int rrlbm_bb(rrlbm_ctx* c) {
  NEED_REF(c);
  if (c->grid.walls) CK(c, ref_bb(c->q, c->ref.f, c->f_temp, c->grid, c->stream));
  return RRLBM_OK;
}

int rrlbm_swap_f(rrlbm_ctx* c) {
  NEED_REF(c);
  std::swap(c->ref.f, c->f_temp);
  return RRLBM_OK;
}

int rrlbm_macrovar(rrlbm_ctx* c) {
  NEED_REF(c);
  CK(c, ref_macrovar(c->q, c->ref, c->grid, 0, c->stream, is_bgk(c->p)));
  return RRLBM_OK;
}

int rrlbm_collision_bgk(rrlbm_ctx* c) {
  NEED_REF(c);
  if (!is_bgk(c->p)) return fail(c, RRLBM_E_STATE, "collision_bgk needs params.collision_model = RRLBM_COLLISION_BGK");
  CK(c, ref_collision_bgk(c->ref, c->grid, c->p.tau, c->stream));
  return RRLBM_OK;
}

int rrlbm_rrbc(rrlbm_ctx* c) {
  NEED_REF(c);
  if (c->grid.walls) CK(c, ref_rrbc(c->q, c->ref, c->f_temp, c->grid, c->stream));
  return RRLBM_OK;
}

int rrlbm_fdbc(rrlbm_ctx* c) {
  NEED_REF(c);
  if (c->q != 27 || !c->p.wall_y || c->p.wall_x || c->p.wall_z || c->p.lid_face >= 0 || c->p.ly < 4 ||
      is_bgk(c->p) || is_rr6(c->p))
    return fail(c, RRLBM_E_UNSUPPORTED,
                "FDBC is the D3Q27 third-order RR channel's y-wall treatment (y walls only, ly >= 4)");
  CK(c, ref_fdbc(c->ref, c->grid, c->p.tau, c->p.use_les, c->stream));
  return RRLBM_OK;
}

int rrlbm_macrovar_bn(rrlbm_ctx* c) {
  NEED_REF(c);
  if (c->grid.walls) CK(c, ref_macrovar(c->q, c->ref, c->grid, 1, c->stream, is_bgk(c->p)));
  return RRLBM_OK;
}

int rrlbm_step_reference(rrlbm_ctx* c, int nsteps) {
  NEED_REF(c);
  if (nsteps < 0) return fail(c, RRLBM_E_INVALID, "nsteps must be >= 0");
  for (int s = 0; s < nsteps; ++s) {
    int rc;
    if (c->p.use_les && (rc = rrlbm_les_wale(c))) return rc;
    // BGK/main.cu:49-70: collision_BGK -> streaming -> streaming2 (BB) -> copy -> macrovar
    if ((rc = is_bgk(c->p) ? rrlbm_collision_bgk(c) : rrlbm_collision_rr(c))) return rc;
    if ((rc = rrlbm_streaming(c))) return rc;
    if ((rc = rrlbm_bb(c))) return rc;
    if ((rc = rrlbm_swap_f(c))) return rc;
    if ((rc = rrlbm_macrovar(c))) return rc;
    if (c->grid.walls && c->p.wall_rr) {
      // main.cu:74-77: RRBC, or FDBC in its place (wall_rr = 2)
      if ((rc = c->p.wall_rr == 2 ? rrlbm_fdbc(c) : rrlbm_rrbc(c))) return rc;
      if ((rc = rrlbm_macrovar_bn(c))) return rc;
    }
  }
  c->packed = false;
  return RRLBM_OK;
}

static int init_common(rrlbm_ctx* c, const InitSpec& s) {
  CK(c, ref_init_fields(c->ref, c->grid, s, c->stream));
  CK(c, ref_init_equilibrium(c->q, c->ref, c->grid, c->stream, is_bgk(c->p)));
  CK(c, ref_forcing(c->ref, c->grid, c->p.force[0], c->p.force[1], c->p.force[2], c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  c->force_fields = rrlbm_ctx::Forcing{};  // the force fields hold the uniform force again
  c->packed = false;
  return RRLBM_OK;
}

int rrlbm_init_rest(rrlbm_ctx* c) {
  NEED_REF(c);
  return init_common(c, InitSpec{kInitRest, 0, 0, 0, 0, 0, 0});
}

int rrlbm_init_taylor_green_2d(rrlbm_ctx* c, double u0) {
  NEED_REF(c);
  return init_common(c, InitSpec{kInitTaylorGreen2D, u0, 0, 0, 0, 0, 0});
}

int rrlbm_init_taylor_green_3d(rrlbm_ctx* c, double u0) {
  NEED_REF(c);
  return init_common(c, InitSpec{kInitTaylorGreen3D, u0, 0, 0, 0, 0, 0});
}

int rrlbm_init_channel(rrlbm_ctx* c, double u_tau, double visc, uint64_t seed, double amp_streak, double amp_noise) {
  NEED_REF(c);
  if (!(u_tau > 0.0) || !(visc > 0.0)) return fail(c, RRLBM_E_INVALID, "u_tau and visc must be positive");
  // bulk velocity of the mean profile (host, FP64)
  const double ystar = visc / u_tau;
  double ub = 0.0;
  const int ly = c->p.ly;
  for (int iy = 0; iy < ly; ++iy) {
    const double d = iy < ly / 2 ? iy + 0.5 : ly - iy - 0.5;
    const double yp = d / ystar;
    ub += u_tau * (yp <= 10.8 ? yp : log(yp) / 0.41 + 5.0);
  }
  ub /= ly;
  InitSpec s{kInitChannel, u_tau, visc, amp_streak, amp_noise, ub, seed};
  return init_common(c, s);
}

int rrlbm_init_equilibrium(rrlbm_ctx* c) {
  NEED_REF(c);
  CK(c, ref_init_equilibrium(c->q, c->ref, c->grid, c->stream, is_bgk(c->p)));
  c->packed = false;
  return RRLBM_OK;
}

int rrlbm_fused_pack(rrlbm_ctx* c) {
  NEED_REF(c);
  if (c->p.wall_rr == 2 && c->p.use_les)
    return fail(c, RRLBM_E_UNSUPPORTED, "the fused path runs FDBC without LES (the drop-in kernels take both)");
  if (c->p.use_les) CK(c, ref_les_wale(c->ref, c->grid, c->p.cw, c->stream));
  // the first collision takes forcing_GP's fields when they hold it
  // (main.cu:33-59: forcing_GP, then collision_RR)
  const int gp = c->force_fields.gp;
  CK(c, ref_pack(c->p.storage, c->q, c->ref, c->grid, c->pop_a, c->lay, c->p.tau, c->p.use_les, c->p.force[0],
                 c->p.force[1], c->p.force[2], c->p.f16_scale, c->stream, is_bgk(c->p), gp));
  c->force_cur = gp ? c->force_fields : rrlbm_ctx::Forcing{};
  if (c->p.use_les) CK(c, ref_copy_u(c->p.storage, c->ref, c->grid, c->u_a, c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  c->packed = true;
  return RRLBM_OK;
}

int rrlbm_fused_step(rrlbm_ctx* c, int nsteps) {
  NEED(c);
  if (!c->packed) return fail(c, RRLBM_E_STATE, "rrlbm_fused_pack must run before fused steps");
  if (nsteps < 0) return fail(c, RRLBM_E_INVALID, "nsteps must be >= 0");
  return fused_steps(c, nsteps);
}

int rrlbm_fused_step_gp(rrlbm_ctx* c, int istep, double rand_phase) {
  NEED(c);
  if (!c->packed) return fail(c, RRLBM_E_STATE, "rrlbm_fused_pack must run before fused steps");
  if (c->q != 27 || is_bgk(c->p) || is_rr6(c->p) || c->p.wall_rr == 2 || c->p.force[1] != 0.0 ||
      c->p.force[2] != 0.0)
    return fail(c, RRLBM_E_UNSUPPORTED,
                "forcing_GP in the fused step: the D3Q27 RR channel with force (force_in_x, 0, 0), RRBC or bounce-back walls");
  const rrlbm_ctx::Forcing next{1, istep, rand_phase};
  return fused_steps(c, 1, &next);
}

int rrlbm_fused_step_timed(rrlbm_ctx* c, int nsteps, double* ms) {
  NEED(c);
  if (!ms) return fail(c, RRLBM_E_INVALID, "null output");
  if (!c->packed) return fail(c, RRLBM_E_STATE, "rrlbm_fused_pack must run before fused steps");
  if (nsteps < 0) return fail(c, RRLBM_E_INVALID, "nsteps must be >= 0");
  CK(c, hipEventRecord(c->ev0, c->stream));
  int rc = fused_steps(c, nsteps);
  if (rc) return rc;
  CK(c, hipEventRecord(c->ev1, c->stream));
  CK(c, hipEventSynchronize(c->ev1));
  float t = 0.0f;
  CK(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
  *ms = (double)t;
  return RRLBM_OK;
}

int rrlbm_graph_capture(rrlbm_ctx* c, int nsteps) {
  NEED(c);
  if (!c->packed) return fail(c, RRLBM_E_STATE, "rrlbm_fused_pack must run before capturing fused steps");
  if (c->force_cur.gp)
    return fail(c, RRLBM_E_STATE, "the fused state's last collision took forcing_GP: run a plain fused step first");
  if (nsteps < 1 || nsteps > 100000) return fail(c, RRLBM_E_INVALID, "nsteps must be in [1, 100000]");
  CK(c, hipStreamSynchronize(c->stream));
  drop_graph(c);
  // the launches record the buffer swaps and the dither counter on the host;
  // nothing runs during capture, so restore them afterwards.  FP16 with
  // stochastic rounding: the captured steps take their step number as the
  // device counter d_step plus their place in the sequence, and the graph
  // ends by advancing d_step, so every replay rounds with new dither.
  void* const a0 = c->pop_a;
  void* const b0 = c->pop_b;
  void* const ua0 = c->u_a;
  void* const ub0 = c->u_b;
  const uint32_t sr0 = c->sr_step;
  const bool dev_step = f16_dither(c->p);
  if (dev_step) {
    c->capture_dev_step = true;
    c->sr_step = 0;
  }
  CK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = fused_steps(c, nsteps);
  hipError_t e = hipSuccess;
  if (!rc && dev_step) e = launch_counter_add(c->d_step, (uint32_t)nsteps, c->stream);
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(c->stream, &g);
  if (e == hipSuccess) e = e2;
  // the buffer swaps of the captured launches (one per launch: a two-step
  // launch advances two steps with one swap)
  const bool swap_pop = c->pop_a != a0, swap_u = c->u_a != ua0;
  c->pop_a = a0; c->pop_b = b0; c->u_a = ua0; c->u_b = ub0; c->sr_step = sr0;
  c->capture_dev_step = false;
  if (rc || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    if (rc) return rc;
    return fail(c, RRLBM_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
  }
  c->graph = g;
  CK(c, hipGraphInstantiate(&c->exec, c->graph, nullptr, nullptr, 0));
  c->graph_steps = nsteps;
  c->graph_pop_a = a0;
  c->graph_dev_step = dev_step;
  c->graph_swap_pop = swap_pop;
  c->graph_swap_u = swap_u;
  return RRLBM_OK;
}

int rrlbm_graph_launch(rrlbm_ctx* c, int times) {
  NEED(c);
  if (!c->exec) return fail(c, RRLBM_E_STATE, "no captured graph");
  if (times < 0) return fail(c, RRLBM_E_INVALID, "times must be >= 0");
  if (!c->packed) return fail(c, RRLBM_E_STATE, "no fused state (rrlbm_fused_pack must run before graph replays)");
  // the captured steps complete a step collided under the uniform force
  if (c->force_cur.gp)
    return fail(c, RRLBM_E_STATE, "the fused state's last collision took forcing_GP: run a plain fused step first");
  if (times > 0 && c->graph_dev_step) {
    hipError_t e = launch_counter_set(c->d_step, c->sr_step, c->stream);
    if (e != hipSuccess) return fail(c, RRLBM_E_HIP, "step counter launch failed: %s", hipGetErrorString(e));
  }
  for (int t = 0; t < times; ++t) {
    if (c->pop_a != c->graph_pop_a)
      return fail(c, RRLBM_E_STATE, "the state sits in the other buffer than at capture (an odd number of launches was captured); capture again");
    CK(c, hipGraphLaunch(c->exec, c->stream));
    if (c->graph_swap_pop) std::swap(c->pop_a, c->pop_b);
    if (c->graph_swap_u) std::swap(c->u_a, c->u_b);
    c->sr_step += (uint32_t)c->graph_steps;
    c->force_cur = rrlbm_ctx::Forcing{};
  }
  return RRLBM_OK;
}

int rrlbm_graph_launch_timed(rrlbm_ctx* c, int times, double* ms) {
  NEED(c);
  if (!ms) return fail(c, RRLBM_E_INVALID, "null output");
  CK(c, hipEventRecord(c->ev0, c->stream));
  int rc = rrlbm_graph_launch(c, times);
  if (rc) return rc;
  CK(c, hipEventRecord(c->ev1, c->stream));
  CK(c, hipEventSynchronize(c->ev1));
  float t = 0.0f;
  CK(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
  *ms = (double)t;
  return RRLBM_OK;
}

int rrlbm_fused_unpack(rrlbm_ctx* c) {
  NEED_REF(c);
  if (!c->packed) return fail(c, RRLBM_E_STATE, "no fused state to unpack");
  int rc = launch(c, kFinish, fused_args(c));
  if (rc) return rc;
  CK(c, hipStreamSynchronize(c->stream));
  return RRLBM_OK;
}

int rrlbm_fused_diag(rrlbm_ctx* c, rrlbm_diag* out) {
  NEED(c);
  if (!out) return fail(c, RRLBM_E_INVALID, "null output");
  if (!c->packed) return fail(c, RRLBM_E_STATE, "no fused state");
  int rc = launch(c, kDiag, fused_args(c));
  if (rc) return rc;
  return finalize_diag(c, c->nblocks, out);
}

int rrlbm_diag_flow(rrlbm_ctx* c, rrlbm_diag* out) {
  NEED_REF(c);
  if (!out) return fail(c, RRLBM_E_INVALID, "null output");
  CK(c, ref_diag(c->ref, c->grid, c->partials, c->stream));
  return finalize_diag(c, c->nblocks, out);
}

}  // extern "C"

// --------------------------------------------- allocation-free diagnostics
// rrlbm_diag_flow_noalloc keeps its per-workgroup partial sums (one row of
// kDiagWidth per block of 256 cells) and the kDiagWidth totals in `scratch`,
// a fixed array that every GPU the library runs on holds until the process
// ends, instead of memory the library allocates.  Host code never passes the
// array's name (a host-side stand-in): it takes each device's address once,
// with hipGetSymbolAddress, and uses that.
namespace {
constexpr uint32_t kNoAllocParts = RRLBM_NOALLOC_MAX_CELLS / kBlock;
static_assert(RRLBM_NOALLOC_MAX_CELLS % kBlock == 0, "whole blocks");
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) double scratch[(size_t)kNoAllocParts * kDiagWidth + kDiagWidth];

namespace {
// Address of `scratch` on the current device (taken once per device).
int scratch_of_current_device(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(nullptr, RRLBM_E_HIP, "hipGetDevice failed: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(nullptr, RRLBM_E_UNSUPPORTED, "device %d: at most %d devices", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(nullptr, RRLBM_E_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return RRLBM_OK;
}
}  // namespace

extern "C" {

int rrlbm_diag_flow_noalloc(const double* rho, const double* ux, const double* uy, const double* uz, int64_t n,
                            rrlbm_diag* out, void* stream) {
  if (!rho || !ux || !uy || !uz || !out) return fail(nullptr, RRLBM_E_INVALID, "null argument");
  if (n < 1 || n > (int64_t)RRLBM_NOALLOC_MAX_CELLS)
    return fail(nullptr, RRLBM_E_INVALID,
                "n = %lld cells: rrlbm_diag_flow_noalloc takes 1 to %d cells (the size of its fixed scratch array)",
                (long long)n, RRLBM_NOALLOC_MAX_CELLS);
  double* sc = nullptr;
  int rc = scratch_of_current_device(&sc);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  Grid g{};
  g.lx = (uint32_t)n; g.ly = 1; g.lz = 1; g.n = (uint32_t)n;
  RefFields r{};
  r.rho = const_cast<double*>(rho);
  r.ux = const_cast<double*>(ux);
  r.uy = const_cast<double*>(uy);
  r.uz = const_cast<double*>(uz);
  const uint32_t nparts = (g.n + kBlock - 1) / kBlock;
  double* totals = sc + (size_t)kNoAllocParts * kDiagWidth;
  // every partial row the finalisation reads is written by this call's first
  // kernel, and the totals by the second: nothing has to start at zero
  hipError_t e = ref_diag(r, g, sc, s);
  if (e == hipSuccess) e = launch_diag_finalize(sc, nparts, totals, s);
  double h[kDiagWidth];
  if (e == hipSuccess) e = hipMemcpyAsync(h, totals, sizeof h, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(nullptr, RRLBM_E_HIP, "diagnostics kernels failed: %s", hipGetErrorString(e));
  diag_from_sums(h, (double)n, out);
  if (h[8] != 0.0) return fail(nullptr, RRLBM_E_NUMERIC, "%.0f cells hold non-finite values", h[8]);
  return RRLBM_OK;
}

int rrlbm_plane_profiles(rrlbm_ctx* c, double* out, size_t count) {
  NEED_REF(c);
  if (!out || count != (size_t)c->p.ly * 10) return fail(c, RRLBM_E_INVALID, "profiles need 10*LY = %d values", c->p.ly * 10);
  CK(c, ref_plane_profiles(c->ref, c->grid, c->p.use_les, c->profiles, c->stream));
  CK(c, hipMemcpyAsync(out, c->profiles, count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  return RRLBM_OK;
}

int rrlbm_write_flow(rrlbm_ctx* c, const char* dir, int istep) {
  NEED_REF(c);
  if (!dir) return fail(c, RRLBM_E_INVALID, "null directory");
  int rc = make_dir(c, dir);
  if (rc) return rc;
  char path[1024];
  snprintf(path, sizeof path, "%s/%09d.dat", dir, istep);
  double* f[4] = {c->ref.rho, c->ref.ux, c->ref.uy, c->ref.uz};
  return write_fields(c, path, f, 4, c->n);
}

int rrlbm_write_f(rrlbm_ctx* c, const char* dir, int istep) {
  NEED_REF(c);
  if (!dir) return fail(c, RRLBM_E_INVALID, "null directory");
  int rc = make_dir(c, dir);
  if (rc) return rc;
  char path[1024];
  snprintf(path, sizeof path, "%s/f%09d.dat", dir, istep);
  double* f[1] = {c->ref.f};
  return write_fields(c, path, f, 1, (size_t)c->q * c->n);
}

int rrlbm_read_restart(rrlbm_ctx* c, const char* dir, int istep) {
  NEED_REF(c);
  if (!dir) return fail(c, RRLBM_E_INVALID, "null directory");
  char path[1024];
  snprintf(path, sizeof path, "%s/%09d.dat", dir, istep);
  double* m[4] = {c->ref.rho, c->ref.ux, c->ref.uy, c->ref.uz};
  int rc = read_fields(c, path, m, 4, c->n);
  if (rc) return rc;
  snprintf(path, sizeof path, "%s/f%09d.dat", dir, istep);
  double* f[1] = {c->ref.f};
  rc = read_fields(c, path, f, 1, (size_t)c->q * c->n);
  if (rc) return rc;
  // initial.cu:262: the force fields are re-imposed after a restart
  CK(c, ref_forcing(c->ref, c->grid, c->p.force[0], c->p.force[1], c->p.force[2], c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  c->force_fields = rrlbm_ctx::Forcing{};
  c->packed = false;
  return RRLBM_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ autotuner
// Picks, per configuration, the population layout (rows per group, padding)
// and the launch shape (cells per thread, XCD-remapped block order,
// persistent grid) of the fused step by timing trial steps on the context's
// own buffers.  Coordinate search: layouts first (default launch shape), then
// the launch shape with the best layout; every candidate is timed in three
// interleaved rounds and scored by its fastest round.  Choices are cached per
// configuration for the rest of the process.
namespace {

std::mutex g_tune_mu;
std::map<std::string, rrlbm_tuning> g_tune_cache;

int force_mode_key(const rrlbm_params& p) {
  return (p.force[0] == 0.0 && p.force[1] == 0.0 && p.force[2] == 0.0) ? 0 : (p.force[1] == 0.0 && p.force[2] == 0.0) ? 1 : 2;
}

std::string tune_key(const rrlbm_params& p) {
  char buf[256];
  snprintf(buf, sizeof buf, "%dx%dx%d q%d s%d w%d%d%d rr%d lid%d les%d/%d f%d m%d o%d p%d/%d", p.lx, p.ly, p.lz,
           p.npop, p.storage, p.wall_x, p.wall_y, p.wall_z, p.wall_rr, p.lid_face >= 0, p.use_les, p.les_lagged,
           force_mode_key(p), p.collision_model, p.rr_order, p.steps_per_pass, p.pass_planes);
  return buf;
}

int realloc_pops(rrlbm_ctx* c, size_t bytes) {
  CK(c, hipStreamSynchronize(c->stream));
  drop_graph(c);
  dfree(c->pop_a);
  dfree(c->pop_b);
  c->pop_a = c->pop_b = nullptr;
  c->bytes -= 2 * c->pop_cap;
  c->pop_cap = 0;
  int rc = dalloc(c, &c->pop_a, bytes);
  if (!rc) rc = dalloc(c, &c->pop_b, bytes);
  if (rc) return rc;
  c->pop_cap = bytes;
  return RRLBM_OK;
}

// Switch ctx to the layout (group_rows, pad), growing the buffers if needed.
int apply_layout(rrlbm_ctx* c, int group_rows, int pad) {
  rrlbm_params p = c->p;
  p.pop_group_rows = group_rows;
  p.pop_pad_bytes = pad;
  PopLayout l;
  if (resolve_layout(p, &l)) return fail(c, RRLBM_E_INVALID, "%s", g_create_error.c_str());
  if (l.bytes > c->pop_cap) {
    int rc = realloc_pops(c, l.bytes);
    if (rc) return rc;
  }
  if (memcmp(&l, &c->lay, sizeof l) != 0) drop_graph(c);  // captured launches carry the old layout
  c->lay = l;
  c->pop_bytes = l.bytes;
  c->p.pop_group_rows = p.pop_group_rows;
  c->p.pop_pad_bytes = p.pop_pad_bytes;
  return RRLBM_OK;
}

struct Candidate {
  int group_rows, pad, cpt, remap, pers;
  int spp = 1, planes = 0;  // steps per pass, z planes per chunk of the two-step launch
};

int apply_candidate(rrlbm_ctx* c, const Candidate& k) {
  int rc = apply_layout(c, k.group_rows, k.pad);
  if (rc) return rc;
  c->p.cells_per_thread = k.cpt;
  c->p.xcd_remap = k.remap;
  c->p.persistent_wg_per_cu = k.pers;
  if (c->p.steps_per_pass != k.spp || c->p.pass_planes != k.planes) drop_graph(c);
  c->p.steps_per_pass = k.spp;
  c->p.pass_planes = k.planes;
  if (k.spp == 2) c->pass = pass_geometry(c->p);
  return RRLBM_OK;
}

// Milliseconds per fused step over `steps` steps after two warm-up steps.
int trial(rrlbm_ctx* c, int steps, double* ms) {
  int rc = fused_steps(c, 2);
  if (rc) return rc;
  CK(c, hipEventRecord(c->ev0, c->stream));
  rc = fused_steps(c, steps);
  if (rc) return rc;
  CK(c, hipEventRecord(c->ev1, c->stream));
  CK(c, hipEventSynchronize(c->ev1));
  float t = 0.0f;
  CK(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
  *ms = (double)t / steps;
  return RRLBM_OK;
}

// Best of three interleaved rounds for each candidate; returns the index of the fastest.
int time_candidates(rrlbm_ctx* c, const std::vector<Candidate>& cand, int steps, std::vector<double>* best_ms,
                    int* best) {
  best_ms->assign(cand.size(), 1e300);
  for (int round = 0; round < 3; ++round) {
    for (size_t k = 0; k < cand.size(); ++k) {
      int rc = apply_candidate(c, cand[k]);
      if (rc) return rc;
      double ms = 0.0;
      rc = trial(c, steps, &ms);
      if (rc) return rc;
      (*best_ms)[k] = std::min((*best_ms)[k], ms);
    }
  }
  *best = 0;
  for (size_t k = 1; k < cand.size(); ++k)
    if ((*best_ms)[k] < (*best_ms)[*best]) *best = (int)k;
  return RRLBM_OK;
}

}  // namespace

extern "C" {

int rrlbm_autotune(rrlbm_ctx* c, int trial_steps, rrlbm_tuning* out) {
  NEED(c);
  if (!out) return fail(c, RRLBM_E_INVALID, "null output");
  if (trial_steps < 1 || trial_steps > 10000) return fail(c, RRLBM_E_INVALID, "trial_steps must be in [1, 10000]");
  if (c->packed)
    return fail(c, RRLBM_E_STATE, "rrlbm_autotune runs before rrlbm_fused_pack (it overwrites the fused buffers)");
  if (c->p.wall_rr == 2 && c->p.use_les)
    return fail(c, RRLBM_E_UNSUPPORTED, "the fused path runs FDBC without LES (the drop-in kernels take both)");
  memset(out, 0, sizeof *out);
  const std::string key = tune_key(c->p);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      *out = it->second;
      out->candidates = 0;
      return apply_candidate(c, Candidate{out->pop_group_rows, out->pop_pad_bytes, out->cells_per_thread,
                                          out->xcd_remap, out->persistent_wg_per_cu, out->steps_per_pass,
                                          out->pass_planes});
    }
  }
  const uint32_t sr0 = c->sr_step;
  const uint32_t rows = (uint32_t)c->p.ly * (uint32_t)c->p.lz;
  const bool rows64 = c->p.lx % 64 == 0;
  // phase 1: layouts with the default launch shape (the library default first)
  std::vector<Candidate> lays = {{0, -1, 1, 0, 0}, {-1, 0, 1, 0, 0}, {-1, 2048, 1, 0, 0}};
  const int ks_rows[] = {1, 4, 16, 64};
  const int ks_lane[] = {16};
  const int* ks = rows64 ? ks_rows : ks_lane;
  const int nks = rows64 ? 4 : 1;
  for (int j = 0; j < nks; ++j)
    if ((uint32_t)ks[j] < rows)
      for (int pad : {0, 256}) lays.push_back({ks[j], pad, 1, 0, 0});
  std::vector<double> ms1;
  int b1 = 0;
  int rc = time_candidates(c, lays, trial_steps, &ms1, &b1);
  if (rc) return rc;
  // phase 2: launch shapes with the best layout (resolved values, so the key is exact)
  Candidate base = lays[b1];
  rc = apply_candidate(c, base);
  if (rc) return rc;
  base.group_rows = c->p.pop_group_rows;
  base.pad = c->p.pop_pad_bytes;
  std::vector<Candidate> shapes;
  // two cells per thread only where use_pair() would launch the pair kernel
  const bool pair_ok = c->q == 27 && !c->p.use_les && !is_bgk(c->p) && !is_rr6(c->p) && c->p.wall_rr != 2;
  for (int cpt : {1, 2})
    for (int remap : {0, 1, 4, 16, 32, 64})
      for (int pers : {0, 4, 8}) {
        if (cpt == 2 && !pair_ok) continue;
        shapes.push_back({base.group_rows, base.pad, cpt, remap, pers});
      }
  // two steps per launch, z chunks of 8 to 32 planes, with the layout the
  // one-step trials chose and with the library default layout: the two kernels
  // do not always prefer the same one (structure of arrays can win the one-step
  // trials and run the two-step kernel 8 % slower than 16-row groups)
  if (two_step_ok(c->p)) {
    const Candidate lay_default{0, -1, 1, 0, 0};
    for (const Candidate& l : {base, lay_default})
      for (int planes : {8, 12, 16, 24, 32})
        if (planes <= c->p.lz) shapes.push_back({l.group_rows, l.pad, 1, base.remap, 0, 2, planes});
  }
  std::vector<double> ms2;
  int b2 = 0;
  rc = time_candidates(c, shapes, trial_steps, &ms2, &b2);
  if (rc) return rc;
  // confirmation: the winner against the library default (its layout resolved)
  // in trials three times as long; a winner that is not faster there is not
  // kept, since a short trial can favour a candidate by chance
  Candidate dflt{0, -1, 1, c->nblocks >= 2048 ? 32 : 0, 0, 1, 0};
  rc = apply_candidate(c, dflt);
  if (rc) return rc;
  dflt.group_rows = c->p.pop_group_rows;
  dflt.pad = c->p.pop_pad_bytes;
  const std::vector<Candidate> fin = {shapes[b2], dflt};
  std::vector<double> ms3;
  int b3 = 0;
  rc = time_candidates(c, fin, 3 * trial_steps, &ms3, &b3);
  if (rc) return rc;
  const Candidate best = fin[b3];
  rc = apply_candidate(c, best);
  if (rc) return rc;
  if (c->pop_cap > c->pop_bytes) {  // shrink to the chosen layout
    rc = realloc_pops(c, c->pop_bytes);
    if (rc) return rc;
  }
  c->sr_step = sr0;
  out->pop_group_rows = c->p.pop_group_rows;  // resolved (a candidate may hold the defaults 0 / -1)
  out->pop_pad_bytes = c->p.pop_pad_bytes;
  out->cells_per_thread = best.cpt;
  out->xcd_remap = best.remap;
  out->persistent_wg_per_cu = best.pers;
  out->steps_per_pass = best.spp;
  out->pass_planes = best.planes;
  out->candidates = (int32_t)(lays.size() + shapes.size() + fin.size());
  out->ms_per_step = ms3[b3];
  out->ms_per_step_default = ms3[1];
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = *out;
  }
  return RRLBM_OK;
}

}  // extern "C"

// ------------------------------------------------------------ batched ensembles
struct rrlbm_batch {
  int storage = RRLBM_STORE_F32;
  bool keep_reference = true;
  std::vector<rrlbm_ctx*> inst;
  std::vector<double> u0;  // initial Taylor-Green amplitude per instance
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  FusedArgs* d_table = nullptr;
  uint2* d_work = nullptr;      // work list: (instance, block) per tile, instance-major
  uint32_t total_blocks = 0;    // tiles in the work list
  int wg_per_cu = 0;            // 0: one workgroup per tile; k: persistent grid of k per CU
  int cus = 256;
  uint32_t parity = 0;  // 1: every instance's current state sits in its table dst buffer
  uint32_t sr_step = 0;  // batched / sequential steps taken (FP16 rounding dither)
  uint32_t* d_step = nullptr;  // device step counter of captured FP16 steps
  bool ready = false;   // instances initialised and packed, table uploaded
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_steps = 0;
  bool graph_dev_step = false;  // the captured steps read d_step (advanced by every replay)
  bool graph_swap_pop = false;  // a replay leaves the state in the other population buffer
  bool graph_swap_u = false;    // ... and the lagged LES velocity in the other velocity buffer
  uint32_t graph_parity = 0;
  int64_t cells = 0;
  std::string err;
};

namespace {

int bfail(rrlbm_batch* b, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (b)
    b->err = buf;
  else
    g_create_error = buf;
  return code;
}

#define BCK(b, call)                                                                                      \
  do {                                                                                                    \
    hipError_t e_ = (call);                                                                               \
    if (e_ != hipSuccess) return bfail((b), RRLBM_E_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define NEED_BATCH(b)                                                                         \
  do {                                                                                        \
    if (!(b)) return bfail(nullptr, RRLBM_E_INVALID, "null batch");                          \
    if (!(b)->ready) return bfail((b), RRLBM_E_STATE, "initialise the batch before stepping"); \
  } while (0)

// After a step of every instance each context's current buffer is the other one.
void batch_swap_all(rrlbm_batch* b) {
  for (rrlbm_ctx* c : b->inst) std::swap(c->pop_a, c->pop_b);
  b->parity ^= 1u;
}

// Device table: instance k's FusedArgs with src = the buffer holding its state
// when the batch parity is 0.
int batch_upload_table(rrlbm_batch* b) {
  std::vector<FusedArgs> t(b->inst.size());
  for (size_t k = 0; k < b->inst.size(); ++k) {
    t[k] = fused_args(b->inst[k]);
    t[k].sr_salt = instance_salt((uint32_t)k);
    if (b->parity) {
      const void* cur = t[k].src;
      t[k].src = t[k].dst;
      t[k].dst = const_cast<void*>(cur);
    }
  }
  BCK(b, hipMemcpyAsync(b->d_table, t.data(), t.size() * sizeof(FusedArgs), hipMemcpyHostToDevice, b->stream));
  BCK(b, hipStreamSynchronize(b->stream));
  return RRLBM_OK;
}

uint32_t batch_grid(const rrlbm_batch* b) {
  if (b->wg_per_cu <= 0) return b->total_blocks;
  return std::max(1u, std::min<uint32_t>(b->total_blocks, (uint32_t)b->wg_per_cu * (uint32_t)b->cus));
}

int batch_launch(rrlbm_batch* b) {
  hipError_t e = launch_batched(b->storage, b->d_table, b->d_work, b->total_blocks, batch_grid(b), b->parity,
                                b->sr_step, nullptr, b->stream);
  if (e != hipSuccess) return bfail(b, RRLBM_E_HIP, "batched launch failed: %s", hipGetErrorString(e));
  ++b->sr_step;
  batch_swap_all(b);
  return RRLBM_OK;
}

int batch_sequential(rrlbm_batch* b, int nsteps) {
  for (int s = 0; s < nsteps; ++s) {
    for (size_t k = 0; k < b->inst.size(); ++k) {
      rrlbm_ctx* c = b->inst[k];
      FusedArgs a = fused_args(c);
      a.sr_step = b->sr_step;
      a.sr_salt = instance_salt((uint32_t)k);
      hipError_t e = launch_fused(c->p.storage, 27, false, false, kStep, a, fused_grid(c, a), b->stream);
      if (e != hipSuccess) return bfail(b, RRLBM_E_HIP, "instance launch failed: %s", hipGetErrorString(e));
    }
    ++b->sr_step;
    batch_swap_all(b);
  }
  return RRLBM_OK;
}

int batch_graph_launch(rrlbm_batch* b, int times) {
  if (!b->exec) return bfail(b, RRLBM_E_STATE, "no captured graph");
  if (times > 0 && b->graph_dev_step) BCK(b, launch_counter_set(b->d_step, b->sr_step, b->stream));
  for (int t = 0; t < times; ++t) {
    if (b->parity != b->graph_parity)
      return bfail(b, RRLBM_E_STATE, "the graph was captured at the other buffer parity; capture again");
    BCK(b, hipGraphLaunch(b->exec, b->stream));
    for (int s = 0; s < b->graph_steps; ++s) batch_swap_all(b);
    b->sr_step += (uint32_t)b->graph_steps;
  }
  return RRLBM_OK;
}

}  // namespace

extern "C" {

int rrlbm_batch_create(const rrlbm_instance* inst, int count, int storage, int keep_reference, rrlbm_batch** out) {
  if (!inst || !out || count < 1) return bfail(nullptr, RRLBM_E_INVALID, "need at least one instance");
  *out = nullptr;
  rrlbm_batch* b = new rrlbm_batch();
  b->storage = storage;
  b->keep_reference = keep_reference != 0;
  auto bail = [&](int rc) {
    g_create_error = b->err;
    rrlbm_batch_destroy(b);
    return rc;
  };
  std::vector<uint2> work;
  for (int k = 0; k < count; ++k) {
    rrlbm_params p;
    rrlbm_params_default(&p);
    p.lx = inst[k].lx; p.ly = inst[k].ly; p.lz = inst[k].lz;
    p.npop = 27; p.use_les = 0; p.storage = storage;
    p.wall_x = p.wall_y = p.wall_z = 0; p.wall_rr = 0; p.lid_face = -1;
    p.tau = inst[k].tau;
    p.force[0] = p.force[1] = p.force[2] = 0.0;
    rrlbm_ctx* c = nullptr;
    int rc = rrlbm_create(&p, &c);
    if (rc) {
      b->err = std::string("instance ") + std::to_string(k) + ": " + g_create_error;
      return bail(rc);
    }
    b->inst.push_back(c);
    for (uint32_t g = 0; g < c->nblocks; ++g) work.push_back(make_uint2((uint32_t)k, g));
    b->cells += c->n;
  }
  for (int k = 0; k < count; ++k) b->u0.push_back(inst[k].u0);
  b->total_blocks = (uint32_t)work.size();
  b->cus = b->inst[0]->cus;
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  if (e == hipSuccess) e = hipMalloc(&b->d_table, count * sizeof(FusedArgs));
  if (e == hipSuccess) e = hipMalloc(&b->d_work, work.size() * sizeof(uint2));
  if (e == hipSuccess) e = hipMalloc(&b->d_step, 256);
  if (e == hipSuccess) e = hipMemsetAsync(b->d_step, 0, 256, b->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(b->d_work, work.data(), work.size() * sizeof(uint2), hipMemcpyHostToDevice, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return bail(bfail(b, RRLBM_E_HIP, "batch setup failed: %s", hipGetErrorString(e)));
  *out = b;
  return RRLBM_OK;
}

int rrlbm_batch_destroy(rrlbm_batch* b) {
  if (!b) return RRLBM_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (rrlbm_ctx* c : b->inst) rrlbm_destroy(c);
  dfree(b->d_table);
  dfree(b->d_work);
  dfree(b->d_step);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return RRLBM_OK;
}

const char* rrlbm_batch_last_error(const rrlbm_batch* b) { return b ? b->err.c_str() : g_create_error.c_str(); }

int rrlbm_batch_cells(const rrlbm_batch* b, int64_t* cells) {
  if (!b || !cells) return bfail(nullptr, RRLBM_E_INVALID, "null argument");
  *cells = b->cells;
  return RRLBM_OK;
}

int rrlbm_batch_init_taylor_green_3d(rrlbm_batch* b) {
  if (!b) return bfail(nullptr, RRLBM_E_INVALID, "null batch");
  BCK(b, hipStreamSynchronize(b->stream));
  for (size_t k = 0; k < b->inst.size(); ++k) {
    rrlbm_ctx* c = b->inst[k];
    if (!c->ref_ok) return bfail(b, RRLBM_E_STATE, "instance %zu: reference fields were released", k);
    int rc = rrlbm_init_taylor_green_3d(c, b->u0[k]);
    if (!rc) rc = rrlbm_fused_pack(c);
    if (!rc && !b->keep_reference) rc = rrlbm_release_reference(c);
    if (rc) return bfail(b, rc, "instance %zu: %s", k, c->err.c_str());
  }
  b->parity = 0;
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  int rc = batch_upload_table(b);
  if (rc) return rc;
  b->ready = true;
  return RRLBM_OK;
}

int rrlbm_batch_step(rrlbm_batch* b, int nsteps) {
  NEED_BATCH(b);
  if (nsteps < 0) return bfail(b, RRLBM_E_INVALID, "nsteps must be >= 0");
  for (int s = 0; s < nsteps; ++s) {
    int rc = batch_launch(b);
    if (rc) return rc;
  }
  return RRLBM_OK;
}

int rrlbm_batch_step_sequential(rrlbm_batch* b, int nsteps) {
  NEED_BATCH(b);
  if (nsteps < 0) return bfail(b, RRLBM_E_INVALID, "nsteps must be >= 0");
  return batch_sequential(b, nsteps);
}

int rrlbm_batch_graph_capture(rrlbm_batch* b, int nsteps) {
  NEED_BATCH(b);
  if (nsteps < 1) return bfail(b, RRLBM_E_INVALID, "nsteps must be >= 1");
  BCK(b, hipStreamSynchronize(b->stream));
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  const uint32_t start = b->parity;
  // FP16 (stochastic rounding): step numbers relative to the device counter,
  // which the graph advances at its end, so that replays take new dither
  const bool dev_step = b->storage == RRLBM_STORE_F16;
  BCK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  uint32_t par = start;
  hipError_t e = hipSuccess;
  for (int s = 0; s < nsteps && e == hipSuccess; ++s) {
    e = launch_batched(b->storage, b->d_table, b->d_work, b->total_blocks, batch_grid(b), par,
                       dev_step ? (uint32_t)s : b->sr_step + (uint32_t)s, dev_step ? b->d_step : nullptr, b->stream);
    par ^= 1u;
  }
  if (e == hipSuccess && dev_step) e = launch_counter_add(b->d_step, (uint32_t)nsteps, b->stream);
  if (e != hipSuccess) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(b->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    return bfail(b, RRLBM_E_HIP, "launch during capture failed: %s", hipGetErrorString(e));
  }
  BCK(b, hipStreamEndCapture(b->stream, &b->graph));
  BCK(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->graph_steps = nsteps;
  b->graph_parity = start;
  b->graph_dev_step = dev_step;
  return RRLBM_OK;
}

int rrlbm_batch_graph_launch(rrlbm_batch* b, int times) {
  NEED_BATCH(b);
  if (times < 0) return bfail(b, RRLBM_E_INVALID, "times must be >= 0");
  return batch_graph_launch(b, times);
}

int rrlbm_batch_time(rrlbm_batch* b, int mode, int nsteps, double* ms) {
  NEED_BATCH(b);
  if (!ms || nsteps < 0) return bfail(b, RRLBM_E_INVALID, "bad argument");
  BCK(b, hipStreamSynchronize(b->stream));
  BCK(b, hipEventRecord(b->ev0, b->stream));
  int rc = RRLBM_OK;
  if (mode == 0) rc = rrlbm_batch_step(b, nsteps);
  else if (mode == 1) rc = batch_sequential(b, nsteps);
  else if (mode == 2) rc = batch_graph_launch(b, nsteps);
  else return bfail(b, RRLBM_E_INVALID, "mode must be 0, 1 or 2");
  if (rc) return rc;
  BCK(b, hipEventRecord(b->ev1, b->stream));
  BCK(b, hipEventSynchronize(b->ev1));
  float t = 0.0f;
  BCK(b, hipEventElapsedTime(&t, b->ev0, b->ev1));
  *ms = (double)t;
  return RRLBM_OK;
}

int rrlbm_batch_diag(rrlbm_batch* b, int instance, rrlbm_diag* out) {
  NEED_BATCH(b);
  if (instance < 0 || instance >= (int)b->inst.size()) return bfail(b, RRLBM_E_INVALID, "no instance %d", instance);
  BCK(b, hipStreamSynchronize(b->stream));
  rrlbm_ctx* c = b->inst[instance];
  int rc = rrlbm_fused_diag(c, out);
  if (rc) return bfail(b, rc, "instance %d: %s", instance, c->err.c_str());
  return RRLBM_OK;
}

int rrlbm_batch_set_grid(rrlbm_batch* b, int wg_per_cu) {
  if (!b) return bfail(nullptr, RRLBM_E_INVALID, "null batch");
  if (wg_per_cu < 0 || wg_per_cu > 64) return bfail(b, RRLBM_E_INVALID, "wg_per_cu must be in [0, 64]");
  b->wg_per_cu = wg_per_cu;
  return RRLBM_OK;
}

int rrlbm_batch_autotune(rrlbm_batch* b, int trial_steps, int* wg_per_cu, double* ms_per_step) {
  if (!b) return bfail(nullptr, RRLBM_E_INVALID, "null batch");
  if (trial_steps < 1 || trial_steps > 10000) return bfail(b, RRLBM_E_INVALID, "trial_steps must be in [1, 10000]");
  if (b->ready)
    return bfail(b, RRLBM_E_STATE, "rrlbm_batch_autotune runs before initialisation (it steps the fused buffers)");
  for (rrlbm_ctx* c : b->inst)
    if (c->packed) return bfail(b, RRLBM_E_STATE, "an instance is already packed");
  int rc = batch_upload_table(b);
  if (rc) return rc;
  const uint32_t sr0 = b->sr_step;
  const int cand[] = {0, 2, 4, 8, 16};
  double best_ms[5];
  for (double& v : best_ms) v = 1e300;
  for (int round = 0; round < 3; ++round) {
    for (int k = 0; k < 5; ++k) {
      b->wg_per_cu = cand[k];
      for (int w = 0; w < 2; ++w)
        if ((rc = batch_launch(b))) return rc;
      BCK(b, hipEventRecord(b->ev0, b->stream));
      for (int st = 0; st < trial_steps; ++st)
        if ((rc = batch_launch(b))) return rc;
      BCK(b, hipEventRecord(b->ev1, b->stream));
      BCK(b, hipEventSynchronize(b->ev1));
      float t = 0.0f;
      BCK(b, hipEventElapsedTime(&t, b->ev0, b->ev1));
      best_ms[k] = std::min(best_ms[k], (double)t / trial_steps);
    }
  }
  int best = 0;
  for (int k = 1; k < 5; ++k)
    if (best_ms[k] < best_ms[best]) best = k;
  b->wg_per_cu = cand[best];
  b->sr_step = sr0;
  if (b->parity) batch_swap_all(b);  // back to parity 0 (the buffers hold no state yet)
  if (wg_per_cu) *wg_per_cu = cand[best];
  if (ms_per_step) *ms_per_step = best_ms[best];
  return RRLBM_OK;
}

int rrlbm_batch_context(rrlbm_batch* b, int instance, rrlbm_ctx** ctx) {
  if (!b || !ctx) return bfail(nullptr, RRLBM_E_INVALID, "null argument");
  if (instance < 0 || instance >= (int)b->inst.size()) return bfail(b, RRLBM_E_INVALID, "no instance %d", instance);
  BCK(b, hipStreamSynchronize(b->stream));
  *ctx = b->inst[instance];
  return RRLBM_OK;
}

}  // extern "C"
