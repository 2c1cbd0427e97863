// bmx_internal.h -- types shared by the kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bmx {

// Record field indices (src/chemistry/bmx_chem.H:38-118).
namespace rec {
enum : int {
  radius = 0, c_length = 1, theta = 2, phi = 3, area = 4, vol = 5,
  velx = 6, vely = 7, velz = 8, wx = 9, wy = 10, wz = 11, gx = 18, gy = 19,
  gz = 20, dadt = 21, dvdt = 22, bond_scale = 24,
  first_data = 25  // species A, B, C, their exchange increments, then the
                   // amounts taken from the fluid
};
enum : int {
  cell_type = 5, n_bnds = 6, seg1_id1 = 7, seg1_id2 = 11, site1 = 15,
  position = 19, fuse_flag = 20, split_flag = 21, new_flag = 22,
  fuse_tip = 25, dltd_site1 = 26, dltd_cpu2 = 31, fix_site = 32, id = 33,
  cpu = 34
};
enum : int { YEAST = 0, FUNGI = 1 };
enum : int { TIP = 0, SECOND_1 = 1, SECOND_2 = 2, INTERIOR = 3 };
}  // namespace rec

// Error bits raised by kernels (read back once per evolve).
enum : int32_t {
  ERRBIT_NAN = 1,
  ERRBIT_OVERFLOW = 2,
  ERRBIT_SKIN = 4,
  ERRBIT_BOND = 8,
  ERRBIT_MESH = 16,  // particle outside the mesh, or an over-full cell
  ERRBIT_CHEM = 32,  // negative concentration in the particle chemistry
};

// One periodic box with its bins.
struct BoxGeom {
  double plo[3];
  double phi[3];
  double L[3];      // phi - plo
  double halfL[3];  // 0.5 * L
  double dxi[3];    // 1 / ((phi - plo) / nbin)
  int32_t periodic[3];
  int32_t nbin[3];
  int32_t bin_offset;  // first global bin id of this box
  int32_t pad;
};

struct ForceParams {
  double WA, mu, zw, WZ, xi, grav, kb, bcut, visc;
  double mob;  // vel = mob * fc
  double dt;
};

// Bond topology of a segment (the integer fields the force pass reads from a
// neighbor, plus its own mutable state).
struct Topo {
  int32_t n_bnds;
  int32_t seg_id[4];
  int32_t seg_cpu[4];
  int32_t site[4];
  int32_t position;
  int32_t fuse_tip;
  int32_t fix_site;
  int32_t self_id;
  int32_t self_cpu;
  int32_t dltd;  // 1 once the force pass has reset the dltd_* fields
  int32_t pad;
};
static_assert(sizeof(Topo) == 80, "Topo layout");

// Per-particle arrays in sorted (binned) order.
struct StaticArrays {  // permuted at rebuild only
  int32_t *oid;        // original index
  int32_t *type;       // cell type
  int32_t *col;        // colony (multi-box contexts)
  double *bscale;      // bond_scale
  int32_t *key;        // global bin id at the last rebuild
};
struct DynArrays {     // rewritten every step (ping-pong)
  double4 *pos;        // x, y, z, radius
  double4 *rod;        // c_length, theta, phi, area
  double4 *trig;       // cos(theta), sin(theta), cos(phi), sin(phi)
  Topo *topo;
};

// The level mesh of the particle <-> mesh coupling: n cells per axis,
// dx = (phi - plo) / n and dxi = 1 / dx as AMReX's Geometry forms them.
struct MeshGeom {
  double plo[3];
  double dxi[3];
  double grid_vol;      // dx[0] * dx[1] * dx[2]
  double inv_cell_vol;  // 1 / grid_vol
  int32_t n[3];
  int32_t periodic[3];
};

struct AosLayout {
  int32_t stride, pos, rdata, id, cpu, idata;
};

}  // namespace bmx
