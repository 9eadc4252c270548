// api.hip -- the extern "C" boundary of libpicard_hip.so (include/picard_hip.h).
//
// A picard_ctx owns one domain's device-resident state.  Each entry point
// validates its arguments on the host, launches on the context's stream and
// returns a status; kernels never see a shape that was not checked here.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>

#include "picard_internal.h"
#include "sched.h"

namespace picard {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

namespace {

template <typename T>
int dalloc(T **p, int64_t count) {
  *p = nullptr;
  if (count <= 0) count = 1;
  PICARD_HIP(hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * (size_t)count));
  return PICARD_OK;
}

template <typename T>
void dfree(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

int check_geometry(const picard_geometry *g) {
  if (!g) {
    set_error("geometry is NULL");
    return PICARD_ERR_ARG;
  }
  if (g->nx < 2 || g->ny < 2 || g->nz < 2) {
    set_error("Nx_local, Ny_local, Nz_local must each be >= 2");
    return PICARD_ERR_ARG;
  }
  if (g->species < 1 || g->species > 8) {
    set_error("species must be in 1..8");
    return PICARD_ERR_ARG;
  }
  if (g->max_particles < 1 || g->max_particles >= (int64_t(1) << 31)) {
    set_error("max_particles must be in 1..2^31-1");
    return PICARD_ERR_ARG;
  }
  for (int c = 0; c < 3; ++c) {
    if (!(g->dxyz[c] > 0.0) || !(g->gmax[c] > g->gmin[c]) || !(g->lmax[c] > g->lmin[c])) {
      set_error("cell sizes and box extents must be positive");
      return PICARD_ERR_ARG;
    }
  }
  if (g->precision != PICARD_FP64 && g->precision != PICARD_FP32_SPLIT) {
    set_error("precision must be PICARD_FP64 or PICARD_FP32_SPLIT");
    return PICARD_ERR_ARG;
  }
  if (g->deposit != PICARD_DEPOSIT_REFERENCE_ORDER && g->deposit != PICARD_DEPOSIT_FIXED_POINT) {
    set_error("deposit must be PICARD_DEPOSIT_REFERENCE_ORDER or PICARD_DEPOSIT_FIXED_POINT");
    return PICARD_ERR_ARG;
  }
  if (g->precision == PICARD_FP32_SPLIT && (g->nx > 16000 || g->ny > 16000 || g->nz > 16000)) {
    set_error("the FP32-split form keeps 16-bit cell coordinates (N <= 16000)");
    return PICARD_ERR_ARG;
  }
  const int64_t K = (int64_t)(g->nx + 1) * (g->ny + 1) * (g->nz + 1);
  if ((int64_t)g->species * K >= (int64_t(1) << 31)) {
    set_error("grid too large for 32-bit deposit keys");
    return PICARD_ERR_ARG;
  }
  return PICARD_OK;
}

int upload_species(picard_ctx *c) {
  double h[16];
  for (int s = 0; s < 8; ++s) {
    h[s] = c->spc.n2p[s];
    h[8 + s] = c->spc.charge[s];
  }
  PICARD_HIP(hipMemcpyAsync(c->consts, h, sizeof(h), hipMemcpyHostToDevice, c->stream));
  PICARD_HIP(hipStreamSynchronize(c->stream));
  return PICARD_OK;
}

void free_split(ParticlesSplit &p) {
  for (int d = 0; d < 3; ++d) {
    dfree(p.c[d]);
    dfree(p.f[d]);
    dfree(p.v[d]);
  }
  dfree(p.sp);
  dfree(p.id);
}

int alloc_split(ParticlesSplit &p, int64_t cap) {
  for (int d = 0; d < 3; ++d) {
    PICARD_TRY(dalloc(&p.c[d], cap));
    PICARD_TRY(dalloc(&p.f[d], cap));
    PICARD_TRY(dalloc(&p.v[d], cap));
  }
  PICARD_TRY(dalloc(&p.sp, cap));
  PICARD_TRY(dalloc(&p.id, cap));
  return PICARD_OK;
}

void release(picard_ctx *c) {
  free_split(c->ps);
  free_split(c->ps2);
  dfree(c->Nfix);
  dfree(c->E32);
  dfree(c->r32);
  dfree(c->Uw);
  for (int q = 0; q < 2; ++q) dfree(c->e32[q]);
  for (int q = 0; q < 4; ++q) dfree(c->eo32[q]);
  for (int q = 0; q < 8; ++q) dfree(c->p.r[q]);
  dfree(c->p.sp);
  dfree(c->U);
  dfree(c->kst);
  dfree(c->E);
  dfree(c->N);
  dfree(c->Nraw);
  dfree(c->C);
  dfree(c->D);
  bcpot_fft_release(c);
  dfree(c->haser_off);
  dfree(c->F);
  dfree(c->Fraw);
  dfree(c->Ffix);
  for (int q = 0; q < 3; ++q) {
    dfree(c->jb_mem[q]);
    c->jb[q] = nullptr;
  }
  dfree(c->b_mem);
  c->b = nullptr;
  dfree(c->consts);
  dfree(c->js);
  dfree(c->s.flags);
  dfree(c->s.scan);
  for (int q = 0; q < 6; ++q) dfree(c->s.work[q]);
  dfree(c->s.scan_tmp);
  dfree(c->s.counters);
  dfree(c->s.key);
  dfree(c->s.perm);
  dfree(c->s.hist);
  dfree(c->s.cursor);
  dfree(c->s.stage);
  dfree(c->s.stage_sp);
  dfree(c->s.partials);
  for (hipEvent_t e : c->tev) (void)hipEventDestroy(e);
  c->tev.clear();
  for (hipEvent_t &e : c->solve_ev)
    if (e) (void)hipEventDestroy(e), e = nullptr;
  if (c->js_host) (void)hipHostFree(c->js_host);
  if (c->host_counters) (void)hipHostFree(c->host_counters);
  c->js_host = nullptr;
  c->host_counters = nullptr;
  if (c->stream) (void)hipStreamDestroy(c->stream);
  c->stream = nullptr;
}

int create_impl(const picard_geometry *geo, const picard_species *sp, picard_ctx *c) {
  c->geo = *geo;
  c->spc = *sp;
  c->device = geo->device;
  PICARD_HIP(hipSetDevice(c->device));
  PICARD_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  GridDims &g = c->gd;
  g.nx = geo->nx;
  g.ny = geo->ny;
  g.nz = geo->nz;
  g.X = g.nx + 2;
  g.Y = g.ny + 2;
  g.Z = g.nz + 2;
  g.G = (int64_t)g.X * g.Y * g.Z;
  g.S = geo->species;
  for (int q = 0; q < 3; ++q) {
    g.flow[q] = geo->flow[q] ? 1 : 0;
    g.lmin[q] = geo->lmin[q];
    g.lmax[q] = geo->lmax[q];
    g.gmin[q] = geo->gmin[q];
    g.gmax[q] = geo->gmax[q];
    g.dxyz[q] = geo->dxyz[q];
  }
  g.dV = g.dxyz[0] * g.dxyz[1] * g.dxyz[2];
  c->cap = geo->max_particles;
  c->K = (int64_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  const int64_t cap = c->cap;
  c->split = geo->precision == PICARD_FP32_SPLIT;
  if (c->split) {
    PICARD_TRY(alloc_split(c->ps, cap));
    PICARD_TRY(alloc_split(c->ps2, cap));
    PICARD_TRY(dalloc(&c->Nfix, g.S * g.G));
    PICARD_TRY(dalloc(&c->E32, 4 * g.G));
    // the solve's FP32 arrays: interior cells, rows padded to 4 cells (zeros);
    // the correction buffers carry one more row of zeros after the last plane
    const int64_t NX32 = (g.nx + 3) & ~3;
    const int64_t nc32 = NX32 * g.ny * g.nz;
    PICARD_TRY(dalloc(&c->r32, nc32));
    PICARD_TRY(dalloc(&c->Uw, g.G));
    for (int q = 0; q < 2; ++q) PICARD_TRY(dalloc(&c->e32[q], nc32 + NX32));
    for (int q = 0; q < 4; ++q) PICARD_TRY(dalloc(&c->eo32[q], nc32));
    PICARD_HIP(hipMemsetAsync(c->r32, 0, sizeof(float) * nc32, c->stream));
    for (int q = 0; q < 2; ++q)
      PICARD_HIP(hipMemsetAsync(c->e32[q], 0, sizeof(float) * (nc32 + NX32), c->stream));
    for (int q = 0; q < 4; ++q) PICARD_HIP(hipMemsetAsync(c->eo32[q], 0, sizeof(float) * nc32, c->stream));
    PICARD_HIP(hipMemsetAsync(c->E32, 0, sizeof(float) * 4 * g.G, c->stream));
    const char *pk = std::getenv("PICARD_PAIR_KERNEL");
    c->pair_lds = !(pk && std::strcmp(pk, "regs") == 0);
  } else {
    for (int q = 0; q < 8; ++q) PICARD_TRY(dalloc(&c->p.r[q], cap));
    PICARD_TRY(dalloc(&c->p.sp, cap));
    c->order_free = geo->deposit == PICARD_DEPOSIT_FIXED_POINT;
    if (c->order_free) PICARD_TRY(dalloc(&c->Nfix, g.S * g.G));
  }
  PICARD_TRY(dalloc(&c->U, g.G));
  PICARD_TRY(dalloc(&c->kst, (int64_t)picard_ctx::kStampWords * picard_ctx::kStamps));
  PICARD_TRY(dalloc(&c->E, 3 * g.G));
  PICARD_TRY(dalloc(&c->N, g.S * g.G));
  if (!c->split) PICARD_TRY(dalloc(&c->Nraw, g.S * g.G));
  PICARD_TRY(dalloc(&c->C, g.G));
  const int64_t DN = (2 * (int64_t)g.nx + 1) * (2 * (int64_t)g.ny + 1) * (2 * (int64_t)g.nz + 1);
  PICARD_TRY(dalloc(&c->D, DN));
  // the Jacobi iterates and b start one element into their allocations, so
  // that node pairs (i, i+1) with i even are 16-byte aligned when nx is even
  if (!c->split)
    for (int q = 0; q < 3; ++q) {
      PICARD_TRY(dalloc(&c->jb_mem[q], g.G + 2));
      c->jb[q] = c->jb_mem[q] + 1;
    }
  PICARD_TRY(dalloc(&c->b_mem, g.G + 2));
  c->b = c->b_mem + 1;
  PICARD_TRY(dalloc(&c->consts, 16));
  PICARD_TRY(dalloc(&c->js, 1));
  if (!c->split) {
    PICARD_TRY(dalloc(&c->s.flags, cap));
    PICARD_TRY(dalloc(&c->s.scan, cap + 1));
    for (int q = 0; q < 6; ++q) PICARD_TRY(dalloc(&c->s.work[q], cap + 1));
  }
  // buckets: (species, cell) keys of both deposits and of the split form's
  // sort, plus the sort's out-of-grid bucket
  const int64_t nb = (int64_t)g.S * c->K + 1;
  const int64_t scan_n = cap > nb ? cap : nb;
  c->s.scan_tmp_len = scan_tmp_required(scan_n);
  PICARD_TRY(dalloc(&c->s.scan_tmp, c->s.scan_tmp_len));
  PICARD_TRY(dalloc(&c->s.counters, 16));
  PICARD_TRY(dalloc(&c->s.key, cap));
  if (!c->split) PICARD_TRY(dalloc(&c->s.perm, cap));
  PICARD_TRY(dalloc(&c->s.hist, nb + 1));
  PICARD_TRY(dalloc(&c->s.cursor, nb));
  const int64_t ncell = (int64_t)g.nx * g.ny * g.nz;
  int64_t pb = grid_blocks(g.G > ncell ? g.G : ncell, 256);
  const int64_t pb3 = interior_launch(g.nx, g.ny, g.nz).blocks();
  if (pb3 > pb) pb = pb3;
  if (make_sched(g.nx, g.ny, g.nz).B > pb) pb = make_sched(g.nx, g.ny, g.nz).B;
  if (pb < 8 * 256) pb = 8 * 256;                 // the Jacobi pass kernels: at most 256 blocks per XCD
  c->s.partials_len = pb;
  PICARD_TRY(dalloc(&c->s.partials, 4 * pb));     // two launches' worth (read + write) and scratch
  PICARD_HIP(hipHostMalloc(reinterpret_cast<void **>(&c->js_host), sizeof(JacobiState)));
  PICARD_HIP(hipHostMalloc(reinterpret_cast<void **>(&c->host_counters), 16 * sizeof(int32_t)));
  // the reference zeroes U, N, E and the particle arrays at start (picard.f90:434-439)
  PICARD_HIP(hipMemsetAsync(c->U, 0, sizeof(double) * g.G, c->stream));
  PICARD_HIP(hipMemsetAsync(c->E, 0, sizeof(double) * 3 * g.G, c->stream));
  PICARD_HIP(hipMemsetAsync(c->N, 0, sizeof(double) * g.S * g.G, c->stream));
  PICARD_HIP(hipMemsetAsync(c->C, 0, sizeof(double) * g.G, c->stream));
  PICARD_HIP(hipMemsetAsync(c->D, 0, sizeof(double) * DN, c->stream));
  PICARD_TRY(upload_species(c));
  return PICARD_OK;
}

int64_t field_len(const picard_ctx *c, int which) {
  const GridDims &g = c->gd;
  switch (which) {
    case PICARD_FIELD_U: return g.G;
    case PICARD_FIELD_E: return 3 * g.G;
    case PICARD_FIELD_N: return (int64_t)g.S * g.G;
    case PICARD_FIELD_C: return g.G;
    case PICARD_FIELD_D:
      return (2 * (int64_t)g.nx + 1) * (2 * (int64_t)g.ny + 1) * (2 * (int64_t)g.nz + 1);
    case PICARD_FIELD_F: return c->F ? 3 * (int64_t)g.S * g.G : -1;
    default: return -1;
  }
}

double *field_ptr(picard_ctx *c, int which) {
  switch (which) {
    case PICARD_FIELD_U: return c->U;
    case PICARD_FIELD_E: return c->E;
    case PICARD_FIELD_N: return c->N;
    case PICARD_FIELD_C: return c->C;
    case PICARD_FIELD_D: return c->D;
    case PICARD_FIELD_F: return c->F;
    default: return nullptr;
  }
}

#define CTX_CHECK(c)                          \
  do {                                        \
    if (!(c)) {                               \
      ::picard::set_error("context is NULL"); \
      return PICARD_ERR_ARG;                  \
    }                                         \
    PICARD_HIP(hipSetDevice((c)->device));    \
  } while (0)

}  // namespace

int ensure_flux(picard_ctx *c) {
  const GridDims &g = c->gd;
  const int64_t len = 3 * (int64_t)g.S * g.G;
  if (!c->F) PICARD_TRY(dalloc(&c->F, len));
  if (c->split) {
    if (!c->Ffix) PICARD_TRY(dalloc(&c->Ffix, len));
  } else {
    if (!c->Fraw) PICARD_TRY(dalloc(&c->Fraw, len));
  }
  return PICARD_OK;
}

}  // namespace picard

using namespace picard;

extern "C" {

const char *picard_last_error(void) { return g_last_error.c_str(); }

int picard_create(const picard_geometry *geo, const picard_species *sp, picard_ctx **out) {
  if (!out || !sp) {
    set_error("picard_create: NULL argument");
    return PICARD_ERR_ARG;
  }
  *out = nullptr;
  PICARD_TRY(check_geometry(geo));
  picard_ctx *c = new picard_ctx();
  const int rc = create_impl(geo, sp, c);
  if (rc != PICARD_OK) {
    release(c);
    delete c;
    return rc;
  }
  *out = c;
  return PICARD_OK;
}

int picard_destroy(picard_ctx *c) {
  if (!c) return PICARD_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  release(c);
  delete c;
  return PICARD_OK;
}

int picard_set_species(picard_ctx *c, const picard_species *sp) {
  CTX_CHECK(c);
  if (!sp) {
    set_error("species is NULL");
    return PICARD_ERR_ARG;
  }
  c->spc = *sp;
  return upload_species(c);
}

int picard_synchronize(picard_ctx *c) {
  CTX_CHECK(c);
  PICARD_HIP(hipStreamSynchronize(c->stream));
  return PICARD_OK;
}

int picard_particles_append(picard_ctx *c, const double *rp, const int32_t *sp, int64_t count) {
  CTX_CHECK(c);
  if (count < 0 || (count > 0 && (!rp || !sp))) {
    set_error("particles_append: bad arguments");
    return PICARD_ERR_ARG;
  }
  for (int64_t i = 0; i < count; ++i) {
    if (sp[i] < 1 || sp[i] > c->gd.S) {
      set_error("particles_append: species index outside 1..species");
      return PICARD_ERR_ARG;
    }
  }
  return c->split ? sp_append(c, rp, sp, count) : launch_append(c, rp, sp, count);
}

int picard_particles_read(picard_ctx *c, int64_t first, int64_t count, double *rp, int32_t *sp) {
  CTX_CHECK(c);
  if (count < 0 || (count > 0 && !rp)) {
    set_error("particles_read: bad arguments");
    return PICARD_ERR_ARG;
  }
  return c->split ? sp_read(c, first, count, rp, sp) : launch_read(c, first, count, rp, sp);
}

int picard_particles_count(picard_ctx *c, int64_t *n) {
  CTX_CHECK(c);
  if (!n) {
    set_error("particles_count: NULL");
    return PICARD_ERR_ARG;
  }
  *n = c->n;
  return PICARD_OK;
}

static int check_plasma(const picard_ctx *c, const picard_plasma *pl) {
  if (!pl) {
    set_error("plasma is NULL");
    return PICARD_ERR_ARG;
  }
  for (int s = 0; s < 8; ++s) {
    if (pl->ipp[s] < 0 || pl->epp[s] < 0 || pl->hpp[s] < 0 || pl->ipp[s] > 4096 ||
        pl->epp[s] > 64 || pl->hpp[s] > 64) {
      set_error("plasma: per-cell counts out of range");
      return PICARD_ERR_ARG;
    }
  }
  (void)c;
  return PICARD_OK;
}

int picard_generate_initial(picard_ctx *c, const picard_plasma *pl, uint64_t seed) {
  CTX_CHECK(c);
  PICARD_TRY(check_plasma(c, pl));
  return generate_initial(c, pl, seed);
}

int picard_generate_inflow(picard_ctx *c, const picard_plasma *pl, uint64_t seed, int64_t step) {
  CTX_CHECK(c);
  PICARD_TRY(check_plasma(c, pl));
  return generate_inflow(c, pl, seed, step);
}

int picard_particles_produce(picard_ctx *c, const picard_plasma *pl, const picard_haser *hs,
                             uint64_t seed, int64_t step, double dt, int64_t *produced) {
  CTX_CHECK(c);
  PICARD_TRY(check_plasma(c, pl));
  if (!hs || !(hs->rate_scale >= 0.0) || !(dt >= 0.0)) {
    set_error("particles_produce: haser parameters missing, or rate_scale/dt negative");
    return PICARD_ERR_ARG;
  }
  return produce_haser(c, pl, hs, seed, step, dt, produced);
}

int picard_set_production(picard_ctx *c, const picard_plasma *pl, const picard_haser *hs,
                          uint64_t seed) {
  CTX_CHECK(c);
  if (!hs) {
    c->prod_on = false;
    return PICARD_OK;
  }
  PICARD_TRY(check_plasma(c, pl));
  if (!(hs->rate_scale >= 0.0)) {
    set_error("set_production: rate_scale must be >= 0");
    return PICARD_ERR_ARG;
  }
  c->prod_plasma = *pl;
  c->prod_haser = *hs;
  c->prod_seed = seed;
  c->prod_on = true;
  return PICARD_OK;
}

int picard_advance_velocity_imp(picard_ctx *c, double deltat, double dt) {
  CTX_CHECK(c);
  if (!(dt != 0.0)) {
    set_error("advance_velocity_imp: dt must be nonzero");
    return PICARD_ERR_ARG;
  }
  return c->split ? sp_boris(c, dt, 1, deltat) : launch_boris(c, dt, 1, deltat);
}

int picard_advance_velocity_exp(picard_ctx *c, double deltat, double dt) {
  CTX_CHECK(c);
  if (!(dt != 0.0)) {
    set_error("advance_velocity_exp: dt must be nonzero");
    return PICARD_ERR_ARG;
  }
  return c->split ? sp_boris(c, dt, 2, deltat) : launch_boris(c, dt, 2, deltat);
}

int picard_diagnostics(picard_ctx *c, const double *mass, picard_diag *out) {
  CTX_CHECK(c);
  if (!mass || !out) {
    set_error("diagnostics: mass and out must be non-null");
    return PICARD_ERR_ARG;
  }
  return diagnostics(c, mass, out);
}

int picard_advance_position(picard_ctx *c, double dt) {
  CTX_CHECK(c);
  if (c->split) return sp_push(c, dt);
  return launch_push_x(c, 0, c->n, dt);
}

int picard_bc_outflow(picard_ctx *c, int64_t *num_local, int64_t *removed) {
  CTX_CHECK(c);
  int64_t nn = 0, rem = 0;
  if (c->split) {
    // the split form compacts by re-sorting on the cell (particle order changes)
    const int64_t before = c->n;
    PICARD_TRY(sp_sort(c, 0));
    nn = c->n;
    rem = before - c->n;
  } else {
    PICARD_TRY(launch_outflow(c, &nn, &rem));
  }
  if (num_local) *num_local = nn;
  if (removed) *removed = rem;
  return PICARD_OK;
}

int picard_particles_transfer(picard_ctx *c) {
  CTX_CHECK(c);
  if (c->split) return sp_sort(c, 1);
  return launch_wrap(c);
}

int picard_particles_gridify_all(picard_ctx *c) {
  CTX_CHECK(c);
  return c->split ? sp_gridify(c, true) : launch_gridify(c, true);
}

int picard_particles_gridify(picard_ctx *c) {
  CTX_CHECK(c);
  return c->split ? sp_gridify(c) : launch_gridify(c);
}

int picard_calc_distance(picard_ctx *c) {
  CTX_CHECK(c);
  return launch_calc_distance(c);
}

int picard_bc_potentials(picard_ctx *c) {
  CTX_CHECK(c);
  return (c->split || c->order_free) ? launch_bc_potentials_fft(c) : launch_bc_potentials(c);
}

int picard_calc_epotential(picard_ctx *c, double maxerr, int32_t maxit, int64_t *sweeps) {
  CTX_CHECK(c);
  if (maxit < 0 || !(maxerr >= 0.0)) {
    set_error("calc_epotential: maxit must be >= 0 and maxerr >= 0");
    return PICARD_ERR_ARG;
  }
  int64_t sw = 0;
  const int rc = c->split ? sp_epotential(c, maxerr, maxit, &sw)
                          : launch_calc_epotential(c, maxerr, maxit, &sw);
  if (sweeps) *sweeps = sw;
  return rc;
}

int picard_calc_efield(picard_ctx *c) {
  CTX_CHECK(c);
  return c->split ? sp_efield(c) : launch_calc_efield(c);
}

int picard_bc_fields(picard_ctx *c) {
  CTX_CHECK(c);
  if (c->split) return PICARD_OK;   // flow guards of E32 are written as zero by sp_efield
  return launch_bc_fields(c);
}

int picard_advance_velocity(picard_ctx *c, double dt) {
  CTX_CHECK(c);
  return c->split ? sp_boris(c, dt) : launch_boris(c, dt);
}

int picard_field_read(picard_ctx *c, int32_t which, double *host, int64_t count) {
  CTX_CHECK(c);
  const int64_t len = field_len(c, which);
  if (len < 0 || !host || count != len) {
    set_error("field_read: unknown field or wrong element count");
    return PICARD_ERR_ARG;
  }
  if (c->split && which == PICARD_FIELD_E) PICARD_TRY(sp_widen_e(c));
  PICARD_HIP(hipMemcpyAsync(host, field_ptr(c, which), sizeof(double) * len,
                            hipMemcpyDeviceToHost, c->stream));
  PICARD_HIP(hipStreamSynchronize(c->stream));
  return PICARD_OK;
}

int picard_field_write(picard_ctx *c, int32_t which, const double *host, int64_t count) {
  CTX_CHECK(c);
  const int64_t len = field_len(c, which);
  if (len < 0 || !host || count != len) {
    set_error("field_write: unknown field or wrong element count");
    return PICARD_ERR_ARG;
  }
  PICARD_HIP(hipMemcpyAsync(field_ptr(c, which), host, sizeof(double) * len,
                            hipMemcpyHostToDevice, c->stream));
  PICARD_HIP(hipStreamSynchronize(c->stream));
  if (which == PICARD_FIELD_D) {
    c->have_D = true;
    c->fft_ready = false;
  }
  return PICARD_OK;
}

int picard_timing_enable(picard_ctx *c, int32_t enable) {
  CTX_CHECK(c);
  c->timing = enable != 0;
  c->t_ms = 0.0;
  c->t_launches = 0;
  c->t_sweeps = 0;
  c->tk_ms = 0.0;
  c->tk_launches = 0;
  c->solve_ms = 0.0;
  c->solves = 0;
  return PICARD_OK;
}

int picard_timing_solve(picard_ctx *c, double *solve_ms, int64_t *solves) {
  CTX_CHECK(c);
  if (solve_ms) *solve_ms = c->solve_ms;
  if (solves) *solves = c->solves;
  return PICARD_OK;
}

int picard_timing_read(picard_ctx *c, double *jacobi_ms, int64_t *jacobi_launches) {
  CTX_CHECK(c);
  if (jacobi_ms) *jacobi_ms = c->t_ms;
  if (jacobi_launches) *jacobi_launches = c->t_launches;
  return PICARD_OK;
}

int picard_timing_sweeps(picard_ctx *c, int64_t *jacobi_sweeps) {
  CTX_CHECK(c);
  if (jacobi_sweeps) *jacobi_sweeps = c->t_sweeps;
  return PICARD_OK;
}

int picard_timing_kernel_clock(picard_ctx *c, double *jacobi_ms, int64_t *jacobi_launches) {
  CTX_CHECK(c);
  if (jacobi_ms) *jacobi_ms = c->tk_ms;
  if (jacobi_launches) *jacobi_launches = c->tk_launches;
  return PICARD_OK;
}

int picard_split_solver(picard_ctx *c, int32_t inner) {
  CTX_CHECK(c);
  if (inner < 2 || inner > kMaxSplitInner) {
    set_error("picard_split_solver: inner sweeps per refinement must be in [2, " +
              std::to_string(kMaxSplitInner) + "], got " +
              std::to_string(inner));
    return PICARD_ERR_ARG;
  }
  c->mx_inner = inner;
  return PICARD_OK;
}

int picard_split_sweeps_per_launch(picard_ctx *c, int32_t sweeps) {
  CTX_CHECK(c);
  if (sweeps < 1 || sweeps > 3) {
    set_error("picard_split_sweeps_per_launch: 1, 2 or 3 sweeps per launch, got " +
              std::to_string(sweeps));
    return PICARD_ERR_ARG;
  }
  c->mx_pairs = sweeps >= 2;
  c->mx_triples = sweeps == 3;
  return PICARD_OK;
}

// While timing is on, events around the solve phase of picard_step: at
// `end`, waits for the closing event (the solve has already read its sweep
// count back, so the stream is idle there) and adds the elapsed time.
static int solve_timer(picard_ctx *c, int end) {
  if (!c->timing) return PICARD_OK;
  for (hipEvent_t &e : c->solve_ev)
    if (!e) PICARD_HIP(hipEventCreate(&e));
  PICARD_HIP(hipEventRecord(c->solve_ev[end], c->stream));
  if (end) {
    float ms = 0.0f;
    PICARD_HIP(hipEventSynchronize(c->solve_ev[1]));
    PICARD_HIP(hipEventElapsedTime(&ms, c->solve_ev[0], c->solve_ev[1]));
    c->solve_ms += ms;
    ++c->solves;
  }
  return PICARD_OK;
}

int picard_step(picard_ctx *c, double dt, double maxerr, int32_t maxit, const double *produced,
                const int32_t *produced_species, int64_t nproduced, picard_step_stats *stats) {
  CTX_CHECK(c);
  picard_step_stats st{};
  int rc = PICARD_OK;
  if (c->split) {
    PICARD_TRY(sp_push(c, dt));                                          // picard.f90:564
    if (nproduced > 0) PICARD_TRY(picard_particles_append(c, produced, produced_species, nproduced));
    if (c->prod_on)                                                      // :566
      PICARD_TRY(produce_haser(c, &c->prod_plasma, &c->prod_haser, c->prod_seed, c->steps_done + 1,
                               dt, nullptr));
    const int64_t before = c->n;
    PICARD_TRY(sp_sort(c, 1));                                           // :569, :572
    st.num_local = c->n;
    st.removed = before - c->n;
    PICARD_TRY(sp_gridify(c));                                           // :577
    PICARD_TRY(launch_bc_potentials_fft(c));                             // :581
    PICARD_TRY(solve_timer(c, 0));
    rc = sp_epotential(c, maxerr, maxit, &st.sweeps);                    // :584
    if (rc < 0) return rc;
    PICARD_TRY(solve_timer(c, 1));
    PICARD_TRY(sp_efield(c));                                            // :587-589
    st.pushed = c->n;
    PICARD_TRY(sp_boris(c, dt));                                         // :630
  } else {
    PICARD_TRY(launch_push_x(c, 0, c->n, dt));                           // picard.f90:564
    if (nproduced > 0) PICARD_TRY(picard_particles_append(c, produced, produced_species, nproduced));
    if (c->prod_on)                                                      // :566
      PICARD_TRY(produce_haser(c, &c->prod_plasma, &c->prod_haser, c->prod_seed, c->steps_done + 1,
                               dt, nullptr));
    PICARD_TRY(launch_outflow(c, &st.num_local, &st.removed));          // :569
    PICARD_TRY(launch_wrap(c));                                         // :572
    PICARD_TRY(launch_gridify(c));                                      // :577
    PICARD_TRY(c->order_free ? launch_bc_potentials_fft(c) : launch_bc_potentials(c));  // :581
    PICARD_TRY(solve_timer(c, 0));
    rc = launch_calc_epotential(c, maxerr, maxit, &st.sweeps);          // :584
    if (rc < 0) return rc;
    PICARD_TRY(solve_timer(c, 1));
    PICARD_TRY(launch_calc_efield(c));                                  // :587
    PICARD_TRY(launch_bc_fields(c));                                    // :589
    st.pushed = c->n;
    PICARD_TRY(launch_boris(c, dt));                                    // :630
  }
  ++c->steps_done;
  if (stats) *stats = st;
  return rc;
}

}  // extern "C"
