// sources.hip -- device-side particle sources: the initial quasi-neutral load
// (ParticlesInitialize.f90:54-119) and the per-step guard-cell inflow
// (BCinflow.f90), generated in place on the GPU.
//
// Random numbers come from Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11), keyed by the seed
// and counted by (particle slot, draw index, step), so a run is reproducible
// and independent of the launch shape.  The streams differ from the Fortran
// runtime's random_number; the distributions are the reference's: uniform
// positions in the cell, drifting Maxwellian velocities by Box-Muller
// (Tools.f90:2-48), and the dt/2 rotation with norm restore.
#include "picard_internal.h"
#include "rng.h"

namespace picard {
namespace {

constexpr int kT = 256;

// Cells are given as a box of 1-based cell indices [lo, hi] (inclusive);
// one thread per (cell, ion): it writes the ion and its electrons/protons at
// the slots the reference's loop order gives them (cell-major, species, ion).
struct CellBox {
  int lo[3], hi[3];
  // the dimension whose lower cell edge is L_max itself (the +x, +y, +z
  // inflow faces: x0 = Lx_max, BCinflow.f90:59, y0 = Ly_max, :201, z0 =
  // Lz_max, :342), or -1: elsewhere the edge is Lmin + dxyz*(i-2)
  int upper = -1;
};

// Output of the sources: FP64 structure of arrays (reference form) ...
struct SinkF64 {
  ParticlesF64 p;
  __device__ void put(const GridDims &, int64_t q, uint32_t id, int sp, double x, double y, double z,
                      const double v[3]) const {
    p.r[0][q] = x; p.r[1][q] = y; p.r[2][q] = z;
    p.r[3][q] = v[0]; p.r[4][q] = v[1]; p.r[5][q] = v[2];
    p.r[6][q] = 0.0;
    p.r[7][q] = (double)id;
    p.sp[q] = sp;
  }
};

// ... or the FP32-split form (xi = (x - Lmin)/dx + 1/2 split into floor and offset).
struct SinkSplit {
  ParticlesSplit p;
  __device__ void put(const GridDims &g, int64_t q, uint32_t id, int sp, double x, double y, double z,
                      const double v[3]) const {
    const double r[3] = {x, y, z};
    for (int d = 0; d < 3; ++d) {
      const double xi = (r[d] - g.lmin[d]) / g.dxyz[d] + 0.5;
      double fl = floor(xi);
      float f = (float)(xi - fl);
      if (f >= 1.0f) {
        f = 0.0f;
        fl += 1.0;
      }
      p.c[d][q] = (int16_t)fl;
      p.f[d][q] = f;
      p.v[d][q] = (float)v[d];
    }
    p.sp[q] = (uint8_t)sp;
    p.id[q] = id;
  }
};

template <typename Sink>
__global__ void __launch_bounds__(kT) k_emit(Sink sink, int64_t base, uint32_t id_base, CellBox box,
                                              GridDims g, PlasmaConst pc, uint64_t key,
                                              uint32_t tag) {
  const int ex = box.hi[0] - box.lo[0] + 1, ey = box.hi[1] - box.lo[1] + 1;
  const int ez = box.hi[2] - box.lo[2] + 1;
  int ions_per_cell = 0;
  for (int s = pc.first_species - 1; s < pc.species; ++s) ions_per_cell += pc.ipp[s];
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  const int64_t total = (int64_t)ex * ey * ez * ions_per_cell;
  if (t >= total) return;
  const int64_t cell = t / ions_per_cell;
  int ion = (int)(t % ions_per_cell);
  const int ci = box.lo[0] + (int)(cell % ex);
  const int cj = box.lo[1] + (int)((cell / ex) % ey);
  const int ck = box.lo[2] + (int)(cell / ((int64_t)ex * ey));
  int s = pc.first_species - 1;
  int64_t slot = base + cell * pc.per_cell;
  while (ion >= pc.ipp[s]) {
    slot += (int64_t)pc.ipp[s] * pc.group[s];
    ion -= pc.ipp[s];
    ++s;
  }
  slot += (int64_t)ion * pc.group[s];
  // counter = index within this source call: identical draws for any num_local
  const uint64_t local = (uint64_t)(slot - base);
  Stream st{key, (uint32_t)local, (uint32_t)(local >> 32), tag};
  const double x0 = box.upper == 0 ? g.lmax[0] : g.lmin[0] + g.dxyz[0] * (ci - 2);
  const double y0 = box.upper == 1 ? g.lmax[1] : g.lmin[1] + g.dxyz[1] * (cj - 2);
  const double z0 = box.upper == 2 ? g.lmax[2] : g.lmin[2] + g.dxyz[2] * (ck - 2);
  const double px = x0 + g.dxyz[0] * st.uniform();
  const double py = y0 + g.dxyz[1] * st.uniform();
  const double pz = z0 + g.dxyz[2] * st.uniform();
  double v[3];
  maxwellian(st, pc, s, s, v);
  half_rotate(pc, s, v);
  auto put = [&](int64_t q, int sp_, const double vv[3]) {
    sink.put(g, q, id_base + (uint32_t)(q - base), sp_, px, py, pz, vv);
  };
  put(slot, s + 1, v);
  for (int e = 0; e < pc.epp[s]; ++e) {
    double ve_[3];
    maxwellian(st, pc, 0, s, ve_);
    half_rotate(pc, 0, ve_);
    put(slot + 1 + e, 1, ve_);
  }
  for (int h = 0; h < pc.hpp[s]; ++h) {
    double vh[3];
    maxwellian(st, pc, 1, s, vh);
    half_rotate(pc, 1, vh);
    put(slot + 1 + pc.epp[s] + h, 2, vh);
  }
}

int64_t box_cells(const CellBox &b) {
  return (int64_t)(b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) * (b.hi[2] - b.lo[2] + 1);
}

PlasmaConst make_pc(const picard_ctx *c, const picard_plasma *pl, int first_species) {
  PlasmaConst pc{};
  pc.species = c->gd.S;
  pc.first_species = first_species;
  pc.per_cell = 0;
  for (int s = 0; s < 8; ++s) {
    pc.ipp[s] = pl->ipp[s];
    pc.epp[s] = pl->epp[s];
    pc.hpp[s] = pl->hpp[s];
    pc.group[s] = 1 + pl->epp[s] + pl->hpp[s];
    pc.mass[s] = pl->mass[s];
    pc.kelvin[s] = pl->kelvin[s];
    for (int q = 0; q < 3; ++q) {
      pc.v0[s][q] = pl->v0[4 * s + q];
      pc.tB[s][q] = c->spc.tB[4 * s + q];
    }
  }
  for (int s = first_species - 1; s < pc.species; ++s) pc.per_cell += pc.ipp[s] * pc.group[s];
  for (int q = 0; q < 3; ++q) pc.ve[q] = c->spc.ve[q];
  return pc;
}

int emit_box(picard_ctx *c, const picard_plasma *pl, const CellBox &box, int first_species,
             uint64_t seed, uint32_t tag) {
  const PlasmaConst pc = make_pc(c, pl, first_species);
  const int64_t cells = box_cells(box);
  if (cells <= 0 || pc.per_cell == 0) return PICARD_OK;
  const int64_t count = cells * pc.per_cell;
  if (c->n + count > c->cap) {
    set_error("particle sources: capacity max_particles exceeded");
    return PICARD_ERR_CAPACITY;
  }
  int ions = 0;
  for (int s = first_species - 1; s < pc.species; ++s) ions += pc.ipp[s];
  const int64_t threads = cells * ions;
  if (c->split)
    k_emit<<<grid_blocks(threads, kT), kT, 0, c->stream>>>(SinkSplit{c->ps}, c->n,
                                                           (uint32_t)c->next_id, box, c->gd, pc,
                                                           seed, tag);
  else
    k_emit<<<grid_blocks(threads, kT), kT, 0, c->stream>>>(SinkF64{c->p}, c->n, (uint32_t)c->next_id,
                                                           box, c->gd, pc, seed, tag);
  PICARD_HIP(hipGetLastError());
  c->n += count;
  c->next_id += count;
  return PICARD_OK;
}

// ---------------------------------------------------------------------------
// Haser-model production (ParticlesProduce.f90:1-113).  Per interior cell and
// species with a neutral outflow (vn > 0): neutral density nden at the cell
// centre (:41-59, with the 8-point average distance when (4r)^3 < dV),
// num_create = floor(nu_i * nden * dV / n2p * rate_scale + rnd) ions (:68-71),
// each at a uniform point of the cell with velocity vn * r/|r| + ui * d/|d|,
// d uniform in [0,1)^3 as the reference draws it, moved by v*dt*rnd7
// (:73-81), followed by its epp electrons and hpp protons (:82-98).  Slots
// follow the reference's loop order: cells (x fastest), species, ions.

struct HaserConst {
  int32_t Sp;                // producing species
  int32_t sidx[8];           // their 0-based species indices
  int32_t epp[8], hpp[8], group[8];
  double Qn[8], vn[8], nu_i[8], nu_d[8], n2p[8];
  double ui[8], ue[8], uh[8];
  double dt, rate_scale;
};

__device__ __forceinline__ double sq(double a) { return a * a; }
__device__ __forceinline__ double nu_i_nden_dv(double nu_i, double nden, double dV) {
  return nu_i * nden * dV;
}

// Expected ion count of (cell, producing species k), ParticlesProduce.f90:41-70.
__device__ double haser_lambda(const GridDims &g, const HaserConst &hc, int ci, int cj, int ck, int k) {
  const double dx = g.dxyz[0], dy = g.dxyz[1], dz = g.dxyz[2];
  const double dV = dx * dy * dz;
  const double x0 = g.lmin[0] + dx * (ci - 2);
  const double y0 = g.lmin[1] + dy * (cj - 2);
  const double z0 = g.lmin[2] + dz * (ck - 2);
  double r2 = sq(x0 + 0.5 * dx) + sq(y0 + 0.5 * dy) + sq(z0 + 0.5 * dz);
  double r1 = sqrt(r2);
  if ((4.0 * r1) * (4.0 * r1) * (4.0 * r1) < dV) {
    double acc = 0.0;
    for (int q = 0; q < 8; ++q) {
      const double fx = (q & 4) ? 0.75 : 0.25, fy = (q & 2) ? 0.75 : 0.25, fz = (q & 1) ? 0.75 : 0.25;
      acc += sqrt(sq(x0 + fx * dx) + sq(y0 + fy * dy) + sq(z0 + fz * dz));
    }
    r1 = acc * 0.125;
    r2 = r1 * r1;
  }
  const int s = hc.sidx[k];
  const double nden = hc.Qn[s] / (2.0 * 6.2831853071795864769252867665590 * r2 * hc.vn[s]) *
                      exp(-hc.nu_d[s] * r1 / hc.vn[s]);
  return nu_i_nden_dv(hc.nu_i[s], nden, dV) / hc.n2p[s] * hc.rate_scale;
}

// slots[cs] = num_create * group for cs = cell * Sp + k
__global__ void __launch_bounds__(kT) k_haser_count(GridDims g, HaserConst hc, uint64_t key, uint32_t tag,
                                                    int32_t *__restrict__ slots) {
  const int64_t cs = (int64_t)blockIdx.x * kT + threadIdx.x;
  const int64_t total = (int64_t)g.nx * g.ny * g.nz * hc.Sp;
  if (cs >= total) return;
  const int k = (int)(cs % hc.Sp);
  const int64_t cell = cs / hc.Sp;
  const int ci = 2 + (int)(cell % g.nx), cj = 2 + (int)((cell / g.nx) % g.ny);
  const int ck = 2 + (int)(cell / ((int64_t)g.nx * g.ny));
  Stream st{key, (uint32_t)cs, (uint32_t)((uint64_t)cs >> 32), tag};
  const double rnd = st.uniform();
  const double v = haser_lambda(g, hc, ci, cj, ck, k) + rnd;
  const int32_t n = v < 2147483647.0 ? (int32_t)v : 2147483647;
  slots[cs] = n * hc.group[hc.sidx[k]];
}

// One thread per appended slot; the thread of an ion's first slot writes the
// ion and its electrons and protons.  off = exclusive scan of slots.
template <typename Sink>
__global__ void __launch_bounds__(kT) k_haser_emit(Sink sink, int64_t base, uint32_t id_base, GridDims g,
                                                   HaserConst hc, uint64_t key, uint32_t tag,
                                                   const int32_t *__restrict__ off, int64_t ncs) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= (int64_t)off[ncs]) return;
  // cs = last entry with off[cs] <= t
  int64_t lo = 0, hi = ncs - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if ((int64_t)off[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  const int64_t cs = lo;
  const int k = (int)(cs % hc.Sp);
  const int s = hc.sidx[k];
  const int64_t local = t - off[cs];
  if (local % hc.group[s] != 0) return;
  const int64_t pp = local / hc.group[s];
  const int64_t cell = cs / hc.Sp;
  const int ci = 2 + (int)(cell % g.nx), cj = 2 + (int)((cell / g.nx) % g.ny);
  const int ck = 2 + (int)(cell / ((int64_t)g.nx * g.ny));
  Stream st{key, (uint32_t)cs, (uint32_t)pp, tag};
  double rp[7];
  for (int q = 0; q < 7; ++q) rp[q] = st.uniform();
  const double x0 = g.lmin[0] + g.dxyz[0] * (ci - 2);
  const double y0 = g.lmin[1] + g.dxyz[1] * (cj - 2);
  const double z0 = g.lmin[2] + g.dxyz[2] * (ck - 2);
  double x[3] = {x0 + g.dxyz[0] * rp[0], y0 + g.dxyz[1] * rp[1], z0 + g.dxyz[2] * rp[2]};
  const double rn = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  const double dn = sqrt(rp[3] * rp[3] + rp[4] * rp[4] + rp[5] * rp[5]);
  double v[3];
  for (int c = 0; c < 3; ++c) v[c] = hc.vn[s] * x[c] / rn + hc.ui[s] * rp[3 + c] / dn;
  for (int c = 0; c < 3; ++c) x[c] = x[c] + v[c] * hc.dt * rp[6];
  const int64_t q0 = base + t;
  sink.put(g, q0, id_base + (uint32_t)t, s + 1, x[0], x[1], x[2], v);
  auto companion = [&](int64_t q, int sp_, double u) {
    double w[3], y[3];
    for (int c = 0; c < 3; ++c) w[c] = v[c] - u * rp[3 + c] / dn;
    for (int c = 0; c < 3; ++c) y[c] = x[c] - v[c] * hc.dt * rp[6] + w[c] * hc.dt * rp[6];
    sink.put(g, q, id_base + (uint32_t)(q - base), sp_, y[0], y[1], y[2], w);
  };
  for (int e = 0; e < hc.epp[s]; ++e) companion(q0 + 1 + e, 1, hc.ue[s] + hc.ui[s]);
  for (int h = 0; h < hc.hpp[s]; ++h) companion(q0 + 1 + hc.epp[s] + h, 2, hc.uh[s] + hc.ui[s]);
}

// Per-species speeds of ParticlesProduce.f90:26-36.
HaserConst make_hc(const picard_ctx *c, const picard_plasma *pl, const picard_haser *hs, double dt) {
  HaserConst hc{};
  const GridDims &g = c->gd;
  const double *m = pl->mass;
  for (int s = 0; s < g.S; ++s) {
    hc.epp[s] = pl->epp[s];
    hc.hpp[s] = pl->hpp[s];
    hc.group[s] = 1 + pl->epp[s] + pl->hpp[s];
    hc.Qn[s] = hs->Qn[s];
    hc.vn[s] = hs->vn[s];
    hc.nu_d[s] = hs->nu_d[s];
    const double n2p = c->spc.n2p[s];
    hc.nu_i[s] = hs->nu_i[s];
    hc.n2p[s] = n2p;
    const double Ek = hs->Ek[s];
    hc.ue[s] = sqrt(2.0 * m[s] * Ek * (1.0 / (m[0] * (m[s] + m[0]))));
    hc.uh[s] = sqrt(2.0 * m[s] * Ek * (1.0 / (m[1] * (m[s] + m[1]))));
    if (pl->epp[s] > 0) hc.ui[s] = sqrt(2.0 * m[0] * Ek * (1.0 / (m[s] * (m[s] + m[0]))));
    else if (pl->hpp[s] > 0) hc.ui[s] = sqrt(2.0 * m[1] * Ek * (1.0 / (m[s] * (m[s] + m[1]))));
    else hc.ui[s] = 0.0;
    if (hs->vn[s] > 0.0 && n2p > 0.0 && hs->nu_i[s] > 0.0 && hs->Qn[s] > 0.0)
      hc.sidx[hc.Sp++] = s;
  }
  hc.dt = dt;
  hc.rate_scale = hs->rate_scale;
  return hc;
}

}  // namespace

int produce_haser(picard_ctx *c, const picard_plasma *pl, const picard_haser *hs, uint64_t seed,
                  int64_t step, double dt, int64_t *produced) {
  const GridDims &g = c->gd;
  if (produced) *produced = 0;
  const HaserConst hc = make_hc(c, pl, hs, dt);
  if (hc.Sp == 0) return PICARD_OK;
  const int64_t ncs = (int64_t)g.nx * g.ny * g.nz * hc.Sp;
  if (c->haser_len < ncs + 1) {
    if (c->haser_off) PICARD_HIP(hipFree(c->haser_off));
    c->haser_off = nullptr;
    c->haser_len = 0;
    PICARD_HIP(hipMalloc(reinterpret_cast<void **>(&c->haser_off), sizeof(int32_t) * (ncs + 1)));
    c->haser_len = ncs + 1;
  }
  if (c->s.scan_tmp_len < scan_tmp_required(ncs)) {
    set_error("ParticlesProduce: scan workspace too small for the cell count");
    return PICARD_ERR_CAPACITY;
  }
  // an independent Philox key per source kind; counter tag by step
  const uint64_t key = seed ^ 0x5A5E12D4C0FFEE11ull;
  const uint32_t tag = 2u * (uint32_t)step;
  k_haser_count<<<grid_blocks(ncs, kT), kT, 0, c->stream>>>(g, hc, key, tag, c->haser_off);
  PICARD_HIP(hipGetLastError());
  PICARD_TRY(scan_exclusive_i32(c, c->haser_off, c->haser_off, ncs));
  int32_t total = 0;
  PICARD_HIP(hipMemcpyAsync(&total, c->haser_off + ncs, sizeof(int32_t), hipMemcpyDeviceToHost,
                            c->stream));
  PICARD_HIP(hipStreamSynchronize(c->stream));
  if (total < 0 || c->n + total > c->cap) {
    set_error("ParticlesProduce: capacity max_particles exceeded");
    return PICARD_ERR_CAPACITY;
  }
  if (total > 0) {
    if (c->split)
      k_haser_emit<<<grid_blocks(total, kT), kT, 0, c->stream>>>(
          SinkSplit{c->ps}, c->n, (uint32_t)c->next_id, g, hc, key, tag + 1u, c->haser_off, ncs);
    else
      k_haser_emit<<<grid_blocks(total, kT), kT, 0, c->stream>>>(
          SinkF64{c->p}, c->n, (uint32_t)c->next_id, g, hc, key, tag + 1u, c->haser_off, ncs);
    PICARD_HIP(hipGetLastError());
  }
  c->n += total;
  c->next_id += total;
  if (produced) *produced = total;
  return PICARD_OK;
}

int generate_initial(picard_ctx *c, const picard_plasma *pl, uint64_t seed) {
  const GridDims &g = c->gd;
  CellBox box{{2, 2, 2}, {g.nx + 1, g.ny + 1, g.nz + 1}};
  return emit_box(c, pl, box, 1, seed, 0xA11u);
}

// BCinflow.f90: +x face (x0 = Lx_max, :53-121), -x face (x0 = Lx_min - dx,
// :124-191), then the y faces (:194-333) and z faces (:334-470), species from
// 2, over the same cell ranges as the reference.
int generate_inflow(picard_ctx *c, const picard_plasma *pl, uint64_t seed, int64_t step) {
  const GridDims &g = c->gd;
  const int yn = g.flow[1] ? 1 : 0, zn = g.flow[2] ? 1 : 0;
  const uint32_t base_tag = 0x1000u + 8u * (uint32_t)step;
  if (g.flow[0]) {
    PICARD_TRY(emit_box(c, pl,
                        CellBox{{g.nx + 2, 2 - yn, 2 - zn}, {g.nx + 2, g.ny + 1 + yn, g.nz + 1 + zn}, 0},
                        1, seed, base_tag + 0));
    PICARD_TRY(emit_box(c, pl, CellBox{{1, 2 - yn, 2 - zn}, {1, g.ny + 1 + yn, g.nz + 1 + zn}}, 1,
                        seed, base_tag + 1));
  }
  if (g.flow[1]) {
    PICARD_TRY(emit_box(c, pl, CellBox{{2, g.ny + 2, 2}, {g.nx + 1, g.ny + 2, g.nz + 1 + zn}, 1}, 2,
                        seed, base_tag + 2));
    PICARD_TRY(emit_box(c, pl, CellBox{{2, 1, 2 - zn}, {g.nx + 1, 1, g.nz + 1}}, 2, seed,
                        base_tag + 3));
  }
  if (g.flow[2]) {   // +z takes the -y guard row, -z the +y one (BCinflow.f90:337, :407)
    PICARD_TRY(emit_box(c, pl, CellBox{{2, 2 - yn, g.nz + 2}, {g.nx + 1, g.ny + 1, g.nz + 2}, 2}, 2,
                        seed, base_tag + 4));
    PICARD_TRY(emit_box(c, pl, CellBox{{2, 2, 1}, {g.nx + 1, g.ny + 1 + yn, 1}}, 2, seed,
                        base_tag + 5));
  }
  return PICARD_OK;
}

}  // namespace picard
