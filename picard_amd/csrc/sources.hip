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

namespace picard {
namespace {

constexpr int kT = 256;

struct Philox {
  uint32_t c0, c1, c2, c3;
};

__device__ __forceinline__ void philox_round(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3,
                                             uint32_t k0, uint32_t k1) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  const uint32_t n0 = hi1 ^ c1 ^ k0;
  const uint32_t n2 = hi0 ^ c3 ^ k1;
  c0 = n0;
  c1 = lo1;
  c2 = n2;
  c3 = lo0;
}

// Philox4x32-10 of counter (c0..c3) with a 64-bit key.
__device__ __forceinline__ Philox philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                         uint64_t key) {
  uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    philox_round(c0, c1, c2, c3, k0, k1);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return {c0, c1, c2, c3};
}

// uniform double in [0, 1) from 53 random bits
__device__ __forceinline__ double u01(uint32_t a, uint32_t b) {
  const uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  return (double)m * (1.0 / 9007199254740992.0);
}

// A per-particle generator: successive 4-word blocks of one counter stream.
struct Stream {
  uint64_t key;
  uint32_t slot_lo, slot_hi, tag;
  uint32_t block = 0;
  uint32_t w[4];
  int used = 4;
  __device__ uint32_t next() {
    if (used == 4) {
      const Philox r = philox(slot_lo, slot_hi, tag, block++, key);
      w[0] = r.c0; w[1] = r.c1; w[2] = r.c2; w[3] = r.c3;
      used = 0;
    }
    return w[used++];
  }
  __device__ double uniform() {
    const uint32_t a = next();
    const uint32_t b = next();
    return u01(a, b);
  }
  // Box-Muller as Tools.f90:30-48, r2 > 0
  __device__ double normal() {
    const double r1 = uniform();
    double r2 = uniform();
    while (!(r2 > 0.0)) r2 = uniform();
    return sqrt(-2.0 * log(r2)) * sin(6.2831853071795864769252867665590 * r1);
  }
};

struct PlasmaConst {
  int32_t species;
  int32_t ipp[8], epp[8], hpp[8];
  int32_t group[8];            // 1 + epp + hpp
  int32_t per_cell;            // particles per cell over species first..S
  int32_t first_species;
  double mass[8], kelvin[8], v0[8][3];
  double ve[3], tB[8][3];
};

__device__ void maxwellian(Stream &st, const PlasmaConst &pc, int s_vel, int s_bulk, double v[3]) {
  const double T = pc.kelvin[s_vel], m = pc.mass[s_vel];
  if (T > 0.0 && m > 0.0) {
    const double a = m / (2.0 * 1.380650e-23 * T);
    const double pw = sqrt(1.0 / a);
    for (int c = 0; c < 3; ++c) v[c] = st.normal();
    for (int c = 0; c < 3; ++c) v[c] = v[c] * pw / sqrt(2.0) + pc.v0[s_bulk][c];
  } else {
    for (int c = 0; c < 3; ++c) v[c] = pc.v0[s_bulk][c];
  }
}

__device__ void half_rotate(const PlasmaConst &pc, int s, double v[3]) {
  const double *t = pc.tB[s];
  const double m1 = v[0] - pc.ve[0], m2 = v[1] - pc.ve[1], m3 = v[2] - pc.ve[2];
  double d1 = m1 + m2 * t[2] - m3 * t[1];
  double d2 = m2 + m3 * t[0] - m1 * t[2];
  double d3 = m3 + m1 * t[1] - m2 * t[0];
  const double sd = d1 * d1 + d2 * d2 + d3 * d3;
  if (sd > 0.0) {
    const double f = sqrt((m1 * m1 + m2 * m2 + m3 * m3) / sd);
    d1 *= f; d2 *= f; d3 *= f;
  }
  v[0] = d1 + pc.ve[0];
  v[1] = d2 + pc.ve[1];
  v[2] = d3 + pc.ve[2];
}

// Cells are given as a box of 1-based cell indices [lo, hi] (inclusive);
// one thread per (cell, ion): it writes the ion and its electrons/protons at
// the slots the reference's loop order gives them (cell-major, species, ion).
struct CellBox {
  int lo[3], hi[3];
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
  const double x0 = g.lmin[0] + g.dxyz[0] * (ci - 2);
  const double y0 = g.lmin[1] + g.dxyz[1] * (cj - 2);
  const double z0 = g.lmin[2] + g.dxyz[2] * (ck - 2);
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

}  // namespace

int generate_initial(picard_ctx *c, const picard_plasma *pl, uint64_t seed) {
  const GridDims &g = c->gd;
  CellBox box{{2, 2, 2}, {g.nx + 1, g.ny + 1, g.nz + 1}};
  return emit_box(c, pl, box, 1, seed, 0xA11u);
}

// BCinflow.f90: +x face (x0 = Lx_max), -x face (x0 = Lx_min - dx), then the y
// and z faces (species from 2), over the same cell ranges as the reference.
int generate_inflow(picard_ctx *c, const picard_plasma *pl, uint64_t seed, int64_t step) {
  const GridDims &g = c->gd;
  const int yn = g.flow[1] ? 1 : 0, zn = g.flow[2] ? 1 : 0;
  const uint32_t base_tag = 0x1000u + 8u * (uint32_t)step;
  if (g.flow[0]) {
    PICARD_TRY(emit_box(c, pl, CellBox{{g.nx + 2, 2 - yn, 2 - zn}, {g.nx + 2, g.ny + 1 + yn, g.nz + 1 + zn}},
                        1, seed, base_tag + 0));
    PICARD_TRY(emit_box(c, pl, CellBox{{1, 2 - yn, 2 - zn}, {1, g.ny + 1 + yn, g.nz + 1 + zn}}, 1, seed,
                        base_tag + 1));
  }
  if (g.flow[1]) {
    PICARD_TRY(emit_box(c, pl, CellBox{{2, g.ny + 2, 2}, {g.nx + 1, g.ny + 2, g.nz + 1 + zn}}, 2, seed,
                        base_tag + 2));
    PICARD_TRY(emit_box(c, pl, CellBox{{2, 1, 2 - zn}, {g.nx + 1, 1, g.nz + 1}}, 2, seed, base_tag + 3));
  }
  if (g.flow[2]) {
    PICARD_TRY(emit_box(c, pl, CellBox{{2, 2, g.nz + 2}, {g.nx + 1, g.ny + 1, g.nz + 2}}, 2, seed,
                        base_tag + 4));
    PICARD_TRY(emit_box(c, pl, CellBox{{2, 2, 1}, {g.nx + 1, g.ny + 1, 1}}, 2, seed, base_tag + 5));
  }
  return PICARD_OK;
}

}  // namespace picard
