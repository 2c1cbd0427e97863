// bmx_device.h -- device-side force laws and geometry helpers.
//
// Every expression keeps the association order of the reference's C++
// (src/chemistry/bmx_cell_interaction_K.H, src/des/bmx_pc_interaction.cpp);
// the library is compiled with -ffp-contract=off, so each operation is
// rounded as written and sphere results are bit-identical to a CPU build of
// the same expressions.
#pragma once

#include "bmx_internal.h"
#include "bmx_math.h"

namespace bmx {

constexpr double kPi = 3.14159265358979323846;

// Word-wide maxima and flags that every block of a launch updates at its
// end.  The word is read first and the atomic taken only when it would
// change it, so the blocks do not queue on one address (most blocks find
// the maximum already there).  A stale read only costs an atomic.
__device__ __forceinline__ void max_u64_to(unsigned long long *w,
                                           unsigned long long v) {
  if (v > __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    atomicMax(w, v);
}
__device__ __forceinline__ void or_u64_to(unsigned long long *w,
                                          unsigned long long bits) {
  if ((__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & bits) !=
      bits)
    atomicOr(w, bits);
}

// Image of coordinate xj nearest to xi along one axis, then xi - image
// (the separation the reference forms against a periodic ghost copy,
// src/des/bmx_pc_interaction.cpp:191-193).
__device__ __forceinline__ double sep_axis(double xi, double xj, double L,
                                           double halfL, int periodic) {
  double xg = xj;
  if (periodic) {
    double t = xi - xj;
    if (t > halfL)
      xg = xj + L;
    else if (t < -halfL)
      xg = xj - L;
  }
  return xi - xg;
}

// Periodic enforcement of one coordinate into [lo, hi).
__device__ __forceinline__ double wrap_axis(double x, double lo, double hi,
                                            double L, int periodic) {
  if (periodic) {
    if (x >= hi) x -= L;
    if (x < lo) x += L;
    if (x >= hi) x = lo;
  }
  return x;
}

// floor((x - plo) * dxi) clamped to [0, nbin-1]
// (src/interpolation/bmx_interp_K.H:55-56).
__device__ __forceinline__ int bin_axis(double x, double plo, double dxi,
                                        int nbin) {
  double f = floor((x - plo) * dxi);
  if (!(f >= 0.0)) return 0;  // also NaN
  if (f > (double)(nbin - 1)) return nbin - 1;
  return (int)f;
}

__device__ __forceinline__ int bin_of(const double4 &p, const BoxGeom &g) {
  int bx = bin_axis(p.x, g.plo[0], g.dxi[0], g.nbin[0]);
  int by = bin_axis(p.y, g.plo[1], g.dxi[1], g.nbin[1]);
  int bz = bin_axis(p.z, g.plo[2], g.dxi[2], g.nbin[2]);
  return g.bin_offset + (bz * g.nbin[1] + by) * g.nbin[0] + bx;
}

// Growth-surface law (K.H:1019-1030 for spheres, 1099-1121 per rod site).
__device__ __forceinline__ double wall_law(double z, double rS,
                                           const ForceParams &fp) {
  double dz = z - fp.zw;
  double rA = rS + fp.WZ;
  double v;
  if (dz > 0.0 && dz < rA)
    v = -fp.xi * (dz - rA) * (dz - rA) * (dz - rS);
  else if (dz <= 0.0)
    v = fp.xi * rA * rA * rS;
  else
    v = 0.0;
  v -= fp.grav * fp.xi * rA * rA * rS;
  return v;
}

// Sphere-sphere cubic law, i-side output (K.H:243-270); adds to a[0..2].
__device__ __forceinline__ void sphere_pair(double d0, double d1, double d2,
                                            double r2, double rS,
                                            const ForceParams &fp,
                                            double &a0, double &a1,
                                            double &a2) {
  double rA = rS + fp.WA;
  double rn = sqrt(r2);
  double ux = d0 / rn, uy = d1 / rn, uz = d2 / rn;
  double F = (rn < rA) ? -fp.mu * (rn - rA) * (rn - rA) * (rn - rS) : 0.0;
  a0 += F * ux;
  a1 += F * uy;
  a2 += F * uz;
}

// ---------------------------------------------------------------- rods ----

__device__ __forceinline__ bool in01(double t) { return t >= 0.0 && t <= 1.0; }

// Closest points of two segments (K.H:51-221).
__device__ inline void min_distance(double xi1, double yi1, double zi1,
                                    double xi2, double yi2, double zi2,
                                    double xj1, double yj1, double zj1,
                                    double xj2, double yj2, double zj2,
                                    double &tau_i, double &tau_j) {
  double ri0 = xi2 - xi1, ri1 = yi2 - yi1, ri2 = zi2 - zi1;
  double rj0 = xj2 - xj1, rj1 = yj2 - yj1, rj2 = zj2 - zj1;
  double rirj = 0.0;
  rirj += ri0 * rj0;
  rirj += ri1 * rj1;
  rirj += ri2 * rj2;
  double Aii = ri0 * ri0 + ri1 * ri1 + ri2 * ri2;
  double Aij = -rirj;
  double Ajj = rj0 * rj0 + rj1 * rj1 + rj2 * rj2;
  double det = Aii * Ajj - Aij * Aij;
  double ex = xi1 - xj1, ey = yi1 - yj1, ez = zi1 - zj1;
  double Bi = -(ex * ri0 + ey * ri1 + ez * ri2);
  double Bj = ex * rj0 + ey * rj1 + ez * rj2;
  bool search = false;
  if (Aij == 0.0) {
    tau_i = Bi / Aii;
    tau_j = Bj / Ajj;
  } else if (det == 0.0) {
    double tj0 = Bj / Ajj;
    if (in01(tj0)) {
      tau_i = 0.0;
      tau_j = tj0;
    } else {
      double tj1 = (Bj + rirj) / Ajj;
      if (in01(tj1)) {
        tau_i = 1.0;
        tau_j = tj1;
      } else {
        search = true;
      }
    }
  } else {
    tau_i = (Ajj * Bi - Aij * Bj) / det;
    tau_j = (Aii * Bj - Aij * Bi) / det;
    if (in01(tau_i)) {
      if (tau_j < 0.0 || tau_j > 1.0) {
        double ti0 = Bi / Aii;
        double ti1 = (Bi + rirj) / Aii;
        bool ok0 = in01(ti0), ok1 = in01(ti1);
        if (ok0 && ok1) {
          double rx = xi1 + ti0 * ri0, ry = yi1 + ti0 * ri1,
                 rz = zi1 + ti0 * ri2;
          double s0 = (rx - xj1) * (rx - xj1) + (ry - yj1) * (ry - yj1) +
                      (rz - zj1) * (rz - zj1);
          rx = xi1 + ti1 * ri0;
          ry = yi1 + ti1 * ri1;
          rz = zi1 + ti1 * ri2;
          double s1 = (rx - xj2) * (rx - xj2) + (ry - yj2) * (ry - yj2) +
                      (rz - zj2) * (rz - zj2);
          if (s0 < s1) {
            tau_j = 0.0;
            tau_i = ti0;
          } else {
            tau_j = 1.0;
            tau_i = ti1;
          }
        } else if (ok0) {
          tau_j = 0.0;
          tau_i = ti0;
        } else if (ok1) {
          tau_j = 1.0;
          tau_i = ti1;
        } else {
          search = true;
        }
      }
    } else if (in01(tau_j)) {
      double tj0 = Bj / Ajj;
      double tj1 = (Bj + rirj) / Ajj;
      bool ok0 = in01(tj0), ok1 = in01(tj1);
      if (ok0 && ok1) {
        double rx = xj1 + tj0 * rj0, ry = yj1 + tj0 * rj1, rz = zj1 + tj0 * rj2;
        double s0 = (rx - xi1) * (rx - xi1) + (ry - yi1) * (ry - yi1) +
                    (rz - zi1) * (rz - zi1);
        rx = xj1 + tj1 * rj0;
        ry = yj1 + tj1 * rj1;
        rz = zj1 + tj1 * rj2;
        double s1 = (rx - xi2) * (rx - xi2) + (ry - yi2) * (ry - yi2) +
                    (rz - zi2) * (rz - zi2);
        if (s0 < s1) {
          tau_i = 0.0;
          tau_j = tj0;
        } else {
          tau_i = 1.0;
          tau_j = tj1;
        }
      } else if (ok0) {
        tau_i = 0.0;
        tau_j = tj0;
      } else if (ok1) {
        tau_i = 1.0;
        tau_j = tj1;
      } else {
        search = true;
      }
    } else {
      search = true;
    }
  }
  if (search) {
    double m0 = (xi1 - xj1) * (xi1 - xj1) + (yi1 - yj1) * (yi1 - yj1) +
                (zi1 - zj1) * (zi1 - zj1);
    double m1 = (xi1 - xj2) * (xi1 - xj2) + (yi1 - yj2) * (yi1 - yj2) +
                (zi1 - zj2) * (zi1 - zj2);
    double m2 = (xi2 - xj1) * (xi2 - xj1) + (yi2 - yj1) * (yi2 - yj1) +
                (zi2 - zj1) * (zi2 - zj1);
    double m3 = (xi2 - xj2) * (xi2 - xj2) + (yi2 - yj2) * (yi2 - yj2) +
                (zi2 - zj2) * (zi2 - zj2);
    int im = 0;
    double mn = m0;
    if (m1 < mn) { im = 1; mn = m1; }
    if (m2 < mn) { im = 2; mn = m2; }
    if (m3 < mn) { im = 3; mn = m3; }
    tau_i = (im >= 2) ? 1.0 : 0.0;
    tau_j = (im & 1) ? 1.0 : 0.0;
  }
}

// Per-segment constants of the pair loop: orientation and the lab-frame
// two-site mobility (K.H:555-583, 812-882).  The reference recomputes these
// for every pair from the same inputs; computing them once gives the same
// values.
struct RodSelf {
  double L, rad, bs, area;
  double nx, ny, nz;
  double dii[3][3], dij[3][3];
};

__device__ inline void rod_rotation(const double4 &tr, double R[3][3]) {
  double ct = tr.x, st = tr.y, cp = tr.z, sp = tr.w;
  R[0][0] = cp * st; R[0][1] = -sp; R[0][2] = -cp * ct;
  R[1][0] = sp * st; R[1][1] = cp;  R[1][2] = -sp * ct;
  R[2][0] = ct;      R[2][1] = 0.0; R[2][2] = st;
}

__device__ inline void rod_mobility(double area, double c_length, double visc,
                                    const double R[3][3], double dii[3][3],
                                    double dij[3][3]) {
  area /= 2.0;
  double arad = sqrt(area / (4.0 * kPi));
  double d0 = 1.0 / (6.0 * kPi * arad * visc);
  double e0, e1;
  if (c_length > 2.0 * arad) {
    double l3 = bmx_cube(c_length);
    e0 = (c_length * c_length - 2.0 * arad * arad / 3.0) /
         (4.0 * kPi * l3 * visc);
    e1 = (c_length * c_length + 2.0 * arad * arad / 3.0) /
         (8.0 * kPi * l3 * visc);
  } else {
    e0 = (1.0 - 3.0 * c_length / (16.0 * arad)) / (6.0 * kPi * arad * visc);
    e1 = (1.0 - 9.0 * c_length / (32.0 * arad)) / (6.0 * kPi * arad * visc);
  }
  // D <- R D R^T with D diagonal, in the reference's summation order:
  // tmp[a][b] = sum_k D[a][k] R[b][k]; out[a][b] = sum_k R[a][k] tmp[k][b].
  double D0[3] = {d0, d0, d0}, D1[3] = {e0, e1, e1};
  double t0[3][3], t1[3][3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      double s0 = 0.0, s1 = 0.0;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        double m0 = (a == k) ? D0[a] : 0.0;
        double m1 = (a == k) ? D1[a] : 0.0;
        s0 += m0 * R[b][k];
        s1 += m1 * R[b][k];
      }
      t0[a][b] = s0;
      t1[a][b] = s1;
    }
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      double s0 = 0.0, s1 = 0.0;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        s0 += R[a][k] * t0[k][b];
        s1 += R[a][k] * t1[k][b];
      }
      dii[a][b] = s0;
      dij[a][b] = s1;
    }
}

// The geometric part of RodSelf (all rod_pair_forces needs).
__device__ inline void rod_self_geom(RodSelf &S, const double4 &rod,
                                     const double4 &tr, double rad, double bs) {
  S.L = rod.x;
  S.area = rod.w;
  S.rad = rad;
  S.bs = bs;
  S.nx = tr.y * tr.z;
  S.ny = tr.y * tr.w;
  S.nz = tr.x;
}

// The mobility part of RodSelf (after rod_self_geom).
__device__ inline void rod_self_mob(RodSelf &S, const double4 &tr,
                                    double visc) {
  double R[3][3];
  rod_rotation(tr, R);
  rod_mobility(S.area, S.L, visc, R, S.dii, S.dij);
}

__device__ inline void rod_self_init(RodSelf &S, const double4 &rod,
                                     const double4 &tr, double rad, double bs,
                                     double visc) {
  rod_self_geom(S, rod, tr, rad, bs);
  rod_self_mob(S, tr, visc);
}

__device__ __forceinline__ void remove_axial(double f1[3], double f2[3],
                                             double rx, double ry, double rz) {
  double r = sqrt(rx * rx + ry * ry + rz * rz);
  double nx = rx / r, ny = ry / r, nz = rz / r;
  double fx = f1[0] - f2[0], fy = f1[1] - f2[1], fz = f1[2] - f2[2];
  double fn = fx * nx + fy * ny + fz * nz;
  f1[0] -= 0.5 * fn * nx;
  f1[1] -= 0.5 * fn * ny;
  f1[2] -= 0.5 * fn * nz;
  f2[0] += 0.5 * fn * nx;
  f2[1] += 0.5 * fn * ny;
  f2[2] += 0.5 * fn * nz;
}

// Segment-segment interaction, i-side outputs only (K.H:277-993): the
// reference discards the j-side outputs (PCI:251-261 are disabled) and its
// writes to the neighbor's record land in a per-pair copy (PCI:188).
// ti is i's own evolving record; tj is the neighbor's record at step start.
// rod_pair_forces forms the site forces f1, f2 (after the axial removal)
// and the i axis r_i; it reads only the geometric part of S.
// rod_pair_mobility applies the two-site mobility and adds the pair's
// velocity and angular terms to acc.  rod_pair_i is the two in sequence.
__device__ inline void rod_pair_forces(const double d[3], const RodSelf &S,
                                       Topo &ti, const Topo &tj, double Lj,
                                       double radj, double bsj,
                                       const double4 &trj,
                                       const ForceParams &fp, double f1[3],
                                       double f2[3], double ri[3]) {
  double rS = S.rad + radj;
  double rA = rS + fp.WA;
  double scale = S.bs * bsj;
  double kb = fp.kb;
  int j_id[4], j_cpu[4], j_site[4];
#pragma unroll
  for (int b = 0; b < 4; b++) {
    j_id[b] = tj.seg_id[b];
    j_cpu[b] = tj.seg_cpu[b];
    j_site[b] = tj.site[b];
  }
  int j_bnds = tj.n_bnds;
  int iid = ti.self_id, icpu = ti.self_cpu;
  int jid = tj.self_id, jcpu = tj.self_cpu;

  // Bond tables hold at most 4 entries (n_bnds is checked to 0..4 at bind
  // and a repair stops at 4).  Every lookup below runs over the 4 slots
  // with constant indices, so the tables stay in registers; a search for
  // the first match keeps the lowest matching slot.
  // K.H:338-362
  if (ti.position == rec::SECOND_1 || ti.position == rec::SECOND_2) {
    int site = (ti.position == rec::SECOND_1) ? 1 : 2;
    bool hit = false;
#pragma unroll
    for (int b = 0; b < 4; b++)
      hit = hit || (b < ti.n_bnds && ti.seg_id[b] == jid &&
                    ti.seg_cpu[b] == jcpu && ti.site[b] == site);
    if (hit && tj.position != rec::TIP) ti.position = rec::INTERIOR;
  }
  ti.dltd = 1;  // K.H:364-369
  // K.H:372-545
  bool i_to_j = false, j_to_i = false;
#pragma unroll
  for (int b = 0; b < 4; b++) {
    i_to_j = i_to_j || (b < ti.n_bnds && ti.seg_id[b] == jid && ti.seg_cpu[b] == jcpu);
    j_to_i = j_to_i || (b < j_bnds && j_id[b] == iid && j_cpu[b] == icpu);
  }
  if (!i_to_j && j_to_i) {
    int isite = (ti.fix_site == 0) ? ((ti.site[0] == 1) ? 2 : 1) : ti.fix_site;
    int nb = ti.n_bnds;
    if (nb >= 0 && nb <= 3) {
#pragma unroll
      for (int b = 0; b < 4; b++)
        if (b == nb) {
          ti.seg_id[b] = jid;
          ti.seg_cpu[b] = jcpu;
          ti.site[b] = isite;
        }
      ti.n_bnds = nb + 1;
    }
  } else if (i_to_j && !j_to_i) {
    int jsite = (tj.fix_site == 0) ? ((tj.site[0] == 1) ? 2 : 1) : tj.fix_site;
    if (j_bnds >= 0 && j_bnds <= 3) {
#pragma unroll
      for (int b = 0; b < 4; b++)
        if (b == j_bnds) {
          j_id[b] = iid;
          j_cpu[b] = icpu;
          j_site[b] = jsite;
        }
      j_bnds++;
    }
  }

  // K.H:553-612
  double xi1 = d[0] - 0.5 * S.nx * S.L, yi1 = d[1] - 0.5 * S.ny * S.L,
         zi1 = d[2] - 0.5 * S.nz * S.L;
  double xi2 = d[0] + 0.5 * S.nx * S.L, yi2 = d[1] + 0.5 * S.ny * S.L,
         zi2 = d[2] + 0.5 * S.nz * S.L;
  double nxj = trj.y * trj.z, nyj = trj.y * trj.w, nzj = trj.x;
  double xj1 = -0.5 * nxj * Lj, yj1 = -0.5 * nyj * Lj, zj1 = -0.5 * nzj * Lj;
  double xj2 = 0.5 * nxj * Lj, yj2 = 0.5 * nyj * Lj, zj2 = 0.5 * nzj * Lj;

#pragma unroll
  for (int a = 0; a < 3; a++) {
    f1[a] = 0.0;
    f2[a] = 0.0;
  }
  bool bonded = false;
  // K.H:615-698
  bool ihit = false, jhit = false;
  int si = 0, sj = 0;
#pragma unroll
  for (int b = 3; b >= 0; b--) {  // descending: the lowest match is kept
    if (b < ti.n_bnds && ti.seg_id[b] == jid && ti.seg_cpu[b] == jcpu) {
      ihit = true;
      si = ti.site[b];
    }
    if (b < j_bnds && j_id[b] == iid && j_cpu[b] == icpu) {
      jhit = true;
      sj = j_site[b];
    }
  }
  if (ihit && jhit) {
    int itip = ti.fuse_tip, jtip = tj.fuse_tip;
    if ((itip > 0 && itip == si) || (jtip > 0 && jtip == sj)) kb = scale * kb;
    if ((si == 1 || si == 2) && (sj == 1 || sj == 2)) {
      double ex = (si == 1) ? xi1 : xi2, ey = (si == 1) ? yi1 : yi2,
             ez = (si == 1) ? zi1 : zi2;
      double rx = ex - ((sj == 1) ? xj1 : xj2);
      double ry = ey - ((sj == 1) ? yj1 : yj2);
      double rz = ez - ((sj == 1) ? zj1 : zj2);
      if (si == 1) {
        f1[0] -= kb * rx;
        f1[1] -= kb * ry;
        f1[2] -= kb * rz;
      } else {
        f2[0] -= kb * rx;
        f2[1] -= kb * ry;
        f2[2] -= kb * rz;
      }
      bonded = true;
    }
  }
  // K.H:701-772
  if (!bonded) {
    double ti_, tj_;
    min_distance(xi1, yi1, zi1, xi2, yi2, zi2, xj1, yj1, zj1, xj2, yj2, zj2,
                 ti_, tj_);
    double ri0 = xi2 - xi1, ri1 = yi2 - yi1, ri2 = zi2 - zi1;
    double rj0 = xj2 - xj1, rj1 = yj2 - yj1, rj2 = zj2 - zj1;
    double rx = (xi1 + ti_ * ri0) - (xj1 + tj_ * rj0);
    double ry = (yi1 + ti_ * ri1) - (yj1 + tj_ * rj1);
    double rz = (zi1 + ti_ * ri2) - (zj1 + tj_ * rj2);
    double r = sqrt(rx * rx + ry * ry + rz * rz);
    if (r > 1.0e-8) {
      rx /= r;
      ry /= r;
      rz /= r;
    }
    double F = (r < rA) ? -fp.mu * (r - rA) * (r - rA) * (r - rS) : 0.0;
    if (F != 0.0) {
      f1[0] += F * rx * (1.0 - ti_);
      f1[1] += F * ry * (1.0 - ti_);
      f1[2] += F * rz * (1.0 - ti_);
      f2[0] += F * rx * ti_;
      f2[1] += F * ry * ti_;
      f2[2] += F * rz * ti_;
    }
  }
  // K.H:776-792
  ri[0] = xi1 - xi2;
  ri[1] = yi1 - yi2;
  ri[2] = zi1 - zi2;
  remove_axial(f1, f2, ri[0], ri[1], ri[2]);
}

__device__ inline void rod_pair_mobility(const RodSelf &S, const double f1[3],
                                         const double f2[3],
                                         const double ri[3], double acc[6]) {
  const double rxi = ri[0], ryi = ri[1], rzi = ri[2];
  // K.H:884-893
  double v1[3], v2[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    double s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int b = 0; b < 3; b++) {
      s1 += S.dii[a][b] * f1[b] + S.dij[a][b] * f2[b];
      s2 += S.dii[a][b] * f2[b] + S.dij[a][b] * f1[b];
    }
    v1[a] = s1;
    v2[a] = s2;
  }
  // K.H:973-985
  acc[0] += 0.5 * (v1[0] + v2[0]);
  acc[1] += 0.5 * (v1[1] + v2[1]);
  acc[2] += 0.5 * (v1[2] + v2[2]);
  double w0 = 0.5 * (ryi * v1[2] - rzi * v1[1]);
  double w1 = 0.5 * (rzi * v1[0] - rxi * v1[2]);
  double w2 = 0.5 * (rxi * v1[1] - ryi * v1[0]);
  w0 += -0.5 * (ryi * v2[2] - rzi * v2[1]);
  w1 += -0.5 * (rzi * v2[0] - rxi * v2[2]);
  w2 += -0.5 * (rxi * v2[1] - ryi * v2[0]);
  acc[3] += w0;
  acc[4] += w1;
  acc[5] += w2;
}

__device__ inline void rod_pair_i(const double d[3], const RodSelf &S,
                                  Topo &ti, const Topo &tj, double Lj,
                                  double radj, double bsj, const double4 &trj,
                                  const ForceParams &fp, double acc[6]) {
  double f1[3], f2[3], ri[3];
  rod_pair_forces(d, S, ti, tj, Lj, radj, bsj, trj, fp, f1, f2, ri);
  rod_pair_mobility(S, f1, f2, ri, acc);
}

// Rod growth-surface term (K.H:1031-1282; the mobility velocities of
// K.H:1176-1265 are overwritten by K.H:1266-1281, so only the
// force-as-velocity result is formed).
__device__ inline void rod_surface(const double4 &p, const RodSelf &S,
                                   const ForceParams &fp, double vel[3],
                                   double rot[3]) {
  double x1 = p.x - 0.5 * S.nx * S.L, y1 = p.y - 0.5 * S.ny * S.L,
         z1 = p.z - 0.5 * S.nz * S.L;
  double x2 = p.x + 0.5 * S.nx * S.L, y2 = p.y + 0.5 * S.ny * S.L,
         z2 = p.z + 0.5 * S.nz * S.L;
  double f1[3] = {0.0, 0.0, 0.0}, f2[3] = {0.0, 0.0, 0.0};
  f1[2] = wall_law(z1, S.rad, fp);
  f2[2] = wall_law(z2, S.rad, fp);
  double rx = x1 - x2, ry = y1 - y2, rz = z1 - z2;
  remove_axial(f1, f2, rx, ry, rz);
  vel[0] = 0.5 * (f1[0] + f2[0]);
  vel[1] = 0.5 * (f1[1] + f2[1]);
  vel[2] = 0.5 * (f1[2] + f2[2]);
#pragma unroll
  for (int a = 0; a < 3; a++) {
    f1[a] -= 0.5 * vel[a];
    f2[a] -= 0.5 * vel[a];
  }
  rot[0] = 0.5 * (ry * f1[2] - rz * f1[1]);
  rot[1] = 0.5 * (rz * f1[0] - rx * f1[2]);
  rot[2] = 0.5 * (rx * f1[1] - ry * f1[0]);
  rot[0] += -0.5 * (ry * f2[2] - rz * f2[1]);
  rot[1] += -0.5 * (rz * f2[0] - rx * f2[2]);
  rot[2] += -0.5 * (rx * f2[1] - ry * f2[0]);
}

// Orientation update of a segment from its angular term (PCI:334-432).
__device__ inline void rod_rotate(double &theta, double &phi,
                                  const double4 &tr, double clen,
                                  const double w[3], double subdt) {
  double R[3][3];
  rod_rotation(tr, R);
  double om[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < 3; b++) s += R[b][a] * w[b];
    om[a] = s;
  }
  om[0] = 0.0;
  om[1] *= 2.0 / (clen * clen);
  om[2] *= 2.0 / (clen * clen);
  double on = sqrt(om[1] * om[1] + om[2] * om[2]);
  if (on > 0.0) {
    om[1] /= on;
    om[2] /= on;
  }
  double ct = om[2], st = om[1];
  double orot[3][3] = {{1.0, 0.0, 0.0}, {0.0, ct, -st}, {0.0, st, ct}};
  double dth = on * subdt;
  double nd[3] = {bmx_cos(dth), bmx_sin(dth), 0.0};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < 3; b++) s += orot[b][a] * nd[b];
    om[a] = s;
  }
#pragma unroll
  for (int a = 0; a < 3; a++) {
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < 3; b++) s += R[a][b] * om[b];
    nd[a] = s;
  }
  theta = bmx_acos(nd[2]);
  on = sqrt(nd[0] * nd[0] + nd[1] * nd[1]);
  if (on > 0.0) nd[0] /= on;
  phi = bmx_acos(nd[0]);
  if (nd[1] < 0.0) phi = 2.0 * kPi - phi;
}


// ---------------------------------------------- wave-level helpers ---

// Inclusive prefix sum over segments of 'width' lanes (a power of two up
// to the wave size).
__device__ __forceinline__ int wave_incl_scan(int v, int width = 64) {
  const int lane = threadIdx.x & (width - 1);
  for (int o = 1; o < width; o <<= 1) {
    const int w = __shfl_up(v, o, width);
    if (lane >= o) v += w;
  }
  return v;
}

// DPP lane moves of GFX9 (rows of 16 lanes): row_shr:n reads lane i - n of
// the same row, row_bcast:15 gives rows 1 and 3 (row mask 0xa) the last lane
// of the row below, row_bcast:31 gives rows 2 and 3 (row mask 0xc) lane 31.
// A lane whose source is outside its row, or whose row is masked off,
// receives 'old'.  These move registers directly, where a shuffle is a
// round trip through the LDS crossbar.
constexpr int kDppRowShr = 0x110, kDppBcast15 = 0x142, kDppBcast31 = 0x143;

// Inclusive prefix sum over the 64 lanes of a wave (every lane active).
__device__ __forceinline__ int wave_incl_scan64(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr + 1, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr + 2, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr + 4, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr + 8, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kDppBcast15, 0xa, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kDppBcast31, 0xc, 0xf, false);
  return v;
}

// A double moved by one DPP pattern, as its two 32-bit halves; lanes
// without a source keep their own value.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = (int)(unsigned)b, hi = (int)(unsigned)(b >> 32);
  const unsigned l2 = (unsigned)__builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWS, 0xf, false);
  const unsigned h2 = (unsigned)__builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWS, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)h2 << 32) | l2));
}

// Largest value over the 64 lanes of a wave (every lane active), in every
// lane: running maxima within the rows, carried to lane 63 by the two
// broadcasts, then read from lane 63.
__device__ __forceinline__ double wave_max_f64(double v) {
  double w;
  w = dpp_f64<kDppRowShr + 1, 0xf>(v); v = w > v ? w : v;
  w = dpp_f64<kDppRowShr + 2, 0xf>(v); v = w > v ? w : v;
  w = dpp_f64<kDppRowShr + 4, 0xf>(v); v = w > v ? w : v;
  w = dpp_f64<kDppRowShr + 8, 0xf>(v); v = w > v ? w : v;
  w = dpp_f64<kDppBcast15, 0xa>(v); v = w > v ? w : v;
  w = dpp_f64<kDppBcast31, 0xc>(v); v = w > v ? w : v;
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), 63);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Orders the wave's LDS writes before its later LDS reads (lanes of one
// wave only; waves of a block run different trip counts).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Interacting entries of the window before entry w0 + j (0 <= j <= 128).
__device__ __forceinline__ int packed_before(int j, unsigned long long b0,
                                             unsigned long long b1, int n0) {
  if (j <= 64)
    return __popcll(j == 64 ? b0 : (b0 & ((1ull << j) - 1ull)));
  j -= 64;
  return n0 + __popcll(j == 64 ? b1 : (b1 & ((1ull << j) - 1ull)));
}


}  // namespace bmx
