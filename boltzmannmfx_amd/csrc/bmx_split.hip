// bmx_split.hip -- division of yeast cells and branching of hyphal segments
// (BMXParticleContainer::split_particles, src/des/bmx_split_particles.cpp:
// 16-166; checkSplit, setNewCell, setNewSegment, setNewSegments,
// src/chemistry/bmx_chem_K.H:25-69, 383-442, 457-1011, 2129-2493).
//
// Works on the device record array (original order): one pass decides how
// many particles each record creates, a scan places them after the existing
// records (in record order, ids next_id + offset), a second pass writes the
// parent and its children.  Random numbers come from bmx_uniform of the
// parent's (id, cpu): draw 0 in the check, draws 1.. for the new geometry,
// so the result does not depend on thread order.
#include "bmx_kernels.h"
#include "bmx_math.h"
#include "bmx_rng.h"

namespace bmx {

namespace {

struct Rec {
  double *pos;
  double *rd;
  int32_t *ip;
  int32_t *id;
  int32_t *cpu;
};

__device__ __forceinline__ Rec rec_at(uint8_t *aos, const AosLayout &lay,
                                      int64_t o) {
  uint8_t *r = aos + o * (int64_t)lay.stride;
  return Rec{reinterpret_cast<double *>(r + lay.pos),
             reinterpret_cast<double *>(r + lay.rdata),
             reinterpret_cast<int32_t *>(r + lay.idata),
             reinterpret_cast<int32_t *>(r + lay.id),
             reinterpret_cast<int32_t *>(r + lay.cpu)};
}

// child = parent (copyChildParameters copies every field), own id and cpu
__device__ void copy_child(uint8_t *aos, const AosLayout &lay, int64_t from,
                           int64_t to, int32_t id, int32_t cpu) {
  const uint64_t *s =
      reinterpret_cast<const uint64_t *>(aos + from * (int64_t)lay.stride);
  uint64_t *d = reinterpret_cast<uint64_t *>(aos + to * (int64_t)lay.stride);
  for (int k = 0; k < lay.stride / 8; k++) d[k] = s[k];
  Rec c = rec_at(aos, lay, to);
  *c.id = id;
  *c.cpu = cpu;
}

struct Draws {
  uint64_t seed;
  uint32_t call;
  int32_t id, cpu;
  uint32_t k;
  __device__ double next() { return bmx_uniform(seed, call, id, cpu, k++); }
};

struct SplitParams {
  double max_vol, max_len, max_rad, brnch_prob, split_prob, overlap,
      split_len;
};

enum : int { FUSE_ID = 23, FUSE_CPU = 24 };

__device__ int check_split(const double *rd, const int32_t *ip,
                           const SplitParams &P, Draws &r) {
  const int ct = ip[rec::cell_type];
  if (ct == rec::YEAST) return rd[rec::vol] > P.max_vol ? 1 : 0;
  if (ct != rec::FUNGI) return 0;
  const int position = ip[rec::position];
  if (position == rec::TIP) {
    if (rd[rec::radius] >= P.max_rad) {
      const double x = r.next();
      if (rd[rec::c_length] >= P.max_len)
        return (x > P.split_prob || ip[rec::n_bnds] == 0) ? 1 : 2;
    }
  } else if (position == rec::SECOND_1 || position == rec::SECOND_2) {
    const double x = r.next();
    if (x < P.brnch_prob) return 1;
  }
  return 0;
}

// fusion and deletion bookkeeping of a fresh segment
__device__ __forceinline__ void fresh(int32_t *ip, bool dltd) {
  ip[rec::fuse_flag] = 0;
  ip[rec::split_flag] = 0;
  ip[rec::new_flag] = 0;
  ip[FUSE_ID] = -1;
  ip[FUSE_CPU] = -1;
  if (dltd)
    for (int k = rec::dltd_site1; k <= rec::dltd_cpu2; k++) ip[k] = -1;
}

// polar and azimuthal angle of a unit vector, as the reference inverts them
__device__ __forceinline__ void to_angles(const double g[3], double &th,
                                          double &ph) {
  th = bmx_acos(g[2]);
  const double st = bmx_sin(th);
  ph = 0.0;
  if (st != 0.0) {
    ph = bmx_acos(g[0] / st);
    if (g[1] < 0.0) ph = 2.0 * BMX_PI - ph;
  }
}

// the unit vector (gx, gy, gz) tilted by 0.1 u1 rad at azimuth 2 pi u2
__device__ void tilt(double gx, double gy, double gz, Draws &r, double out[3]) {
  const double gtheta = bmx_acos(gz);
  const double gst = bmx_sin(gtheta);
  const double gct = gz;
  double gphi = 0.0;
  if (gst != 0.0) {
    gphi = bmx_acos(gx / gst);
    if (gy < 0.0) gphi = 2.0 * BMX_PI - gphi;
  }
  const double gcp = bmx_cos(gphi), gsp = bmx_sin(gphi);
  const double m00 = gcp * gst, m01 = -gsp, m02 = -gcp * gct;
  const double m10 = gsp * gst, m11 = gcp, m12 = -gsp * gct;
  const double m20 = gct, m21 = 0.0, m22 = gst;
  const double rtheta = 0.1 * r.next();
  const double rphi = 2.0 * BMX_PI * r.next();
  const double rct = bmx_cos(rtheta), rst = bmx_sin(rtheta);
  const double rcp = bmx_cos(rphi), rsp = bmx_sin(rphi);
  const double rx = rct, ry = rst * rsp, rz = rst * rcp;
  out[0] = m00 * rx + m01 * ry + m02 * rz;
  out[1] = m10 * rx + m11 * ry + m12 * rz;
  out[2] = m20 * rx + m21 * ry + m22 * rz;
}

__device__ __forceinline__ void segment_volume(double *rd) {
  const double clen = rd[rec::c_length], rad = rd[rec::radius];
  rd[rec::vol] = BMX_PI * rad * rad * clen;
  rd[rec::area] = 2.0 * BMX_PI * rad * clen + 2.0 * BMX_PI * rad * rad;
}

__device__ void new_cell(Rec &po, Rec &pn, double overlap, Draws &r) {
  double *a = po.rd, *b = pn.rd;
  const double x = po.pos[0], y = po.pos[1], z = po.pos[2];
  double radius = b[rec::radius];
  double theta = 2.0 * r.next() - 1.0;
  theta = bmx_acos(theta);
  const double phi = 2.0 * BMX_PI * r.next();
  const double nx = bmx_sin(theta) * bmx_cos(phi);
  const double ny = bmx_sin(theta) * bmx_sin(phi);
  const double nz = bmx_cos(theta);
  const double scale = 1.0 - overlap;
  pn.pos[0] = x + scale * 0.5 * nx * radius;
  pn.pos[1] = y + scale * 0.5 * ny * radius;
  pn.pos[2] = z + scale * 0.5 * nz * radius;
  po.pos[0] = x - scale * 0.5 * nx * radius;
  po.pos[1] = y - scale * 0.5 * ny * radius;
  po.pos[2] = z - scale * 0.5 * nz * radius;
  const double volume = a[rec::vol] / 2.0;
  radius = bmx_pow((3.0 * volume / (4.0 * BMX_PI)), 1.0 / 3.0);
  const double area = 4.0 * BMX_PI * radius * radius;
  const double dadt = 2.0 * a[rec::dvdt] / radius;
  a[rec::vol] = b[rec::vol] = volume;
  a[rec::area] = b[rec::area] = area;
  a[rec::dadt] = b[rec::dadt] = dadt;
  a[rec::radius] = b[rec::radius] = radius;
  b[rec::bond_scale] = 1.0;
  pn.ip[rec::fuse_tip] = 0;
  for (int k = rec::first_data + 6; k < rec::first_data + 9; k++) {
    a[k] = 0.5 * a[k];
    b[k] = 0.5 * b[k];
  }
}

// bond slot k (0..3) of a segment
__device__ __forceinline__ void set_bond(int32_t *ip, int k, int32_t id,
                                         int32_t cpu) {
  ip[rec::seg1_id1 + k] = id;
  ip[rec::seg1_id2 + k] = cpu;
}

__device__ void new_segment(Rec &po, Rec &pn, double split_len, Draws &r) {
  double *a = po.rd, *b = pn.rd;
  int32_t *ia = po.ip, *ib = pn.ip;
  const int32_t id = *pn.id, cpu = *pn.cpu;
  const int nbnd = ia[rec::n_bnds];
  const int position = ia[rec::position];
  const double ct = bmx_cos(a[rec::theta]), st = bmx_sin(a[rec::theta]);
  const double cp = bmx_cos(a[rec::phi]), sp = bmx_sin(a[rec::phi]);
  const double nx = st * cp, ny = st * sp, nz = ct;
  if (nbnd == 0) {
    // an unbonded segment halves: two tips bonded to each other
    const double x = po.pos[0], y = po.pos[1], z = po.pos[2];
    const double L = a[rec::c_length];
    po.pos[0] = x + 0.25 * L * nx;
    po.pos[1] = y + 0.25 * L * ny;
    po.pos[2] = z + 0.25 * L * nz;
    pn.pos[0] = x - 0.25 * L * nx;
    pn.pos[1] = y - 0.25 * L * ny;
    pn.pos[2] = z - 0.25 * L * nz;
    a[rec::c_length] = 0.5 * L;
    b[rec::c_length] = 0.5 * L;
    ia[rec::n_bnds] = 1;
    set_bond(ia, 0, id, cpu);
    ia[rec::site1] = 1;
    ia[rec::position] = rec::TIP;
    fresh(ia, true);
    ib[rec::n_bnds] = 1;
    set_bond(ib, 0, ia[rec::id], ia[rec::cpu]);
    ib[rec::site1] = 2;
    ib[rec::position] = rec::TIP;
    fresh(ib, true);
    ib[rec::fix_site] = 0;
    for (int k = rec::first_data + 6; k < rec::first_data + 9; k++) {
      a[k] = 0.5 * a[k];
      b[k] = 0.5 * b[k];
    }
  } else if (position == rec::TIP) {
    // the tip keeps split_len from its bonded end and a new tip continues
    // from its free end, tilted by up to 0.1 rad
    const int bsite = ia[rec::site1];
    double g[3];
    tilt(nx, ny, nz, r, g);
    const double x = po.pos[0], y = po.pos[1], z = po.pos[2];
    const double L = a[rec::c_length];
    const double seg_len = L - split_len;
    if (bsite == 1) {
      const double x1 = x - 0.5 * L * nx, y1 = y - 0.5 * L * ny, z1 = z - 0.5 * L * nz;
      const double x2 = x1 + split_len * nx, y2 = y1 + split_len * ny,
                   z2 = z1 + split_len * nz;
      po.pos[0] = 0.5 * (x1 + x2);
      po.pos[1] = 0.5 * (y1 + y2);
      po.pos[2] = 0.5 * (z1 + z2);
      a[rec::c_length] = split_len;
      ia[nbnd == 1 ? rec::site1 + 1 : rec::site1 + 2] = 2;
      ia[rec::position] = rec::SECOND_2;
      const double ax = x2 + g[0] * seg_len, ay = y2 + g[1] * seg_len,
                   az = z2 + g[2] * seg_len;
      pn.pos[0] = 0.5 * (x2 + ax);
      pn.pos[1] = 0.5 * (y2 + ay);
      pn.pos[2] = 0.5 * (z2 + az);
      ib[rec::site1] = 1;
    } else {
      const double x2 = x + 0.5 * L * nx, y2 = y + 0.5 * L * ny, z2 = z + 0.5 * L * nz;
      const double x1 = x2 - split_len * nx, y1 = y2 - split_len * ny,
                   z1 = z2 - split_len * nz;
      po.pos[0] = 0.5 * (x1 + x2);
      po.pos[1] = 0.5 * (y1 + y2);
      po.pos[2] = 0.5 * (z1 + z2);
      a[rec::c_length] = split_len;
      ia[nbnd == 1 ? rec::site1 + 1 : rec::site1 + 2] = 1;
      ia[rec::position] = rec::SECOND_1;
      const double ax = x1 - g[0] * seg_len, ay = y1 - g[1] * seg_len,
                   az = z1 - g[2] * seg_len;
      pn.pos[0] = 0.5 * (x1 + ax);
      pn.pos[1] = 0.5 * (y1 + ay);
      pn.pos[2] = 0.5 * (z1 + az);
      ib[rec::site1] = 2;
    }
    b[rec::c_length] = seg_len;
    to_angles(g, b[rec::theta], b[rec::phi]);
    ib[rec::position] = rec::TIP;
    fresh(ib, bsite == 1);  // the reference resets the deletion fields on one side only
    ib[rec::fix_site] = 0;
    ia[rec::n_bnds] = nbnd == 1 ? 2 : 3;
    set_bond(ia, nbnd == 1 ? 1 : 2, id, cpu);
    ib[rec::n_bnds] = 1;
    ib[rec::id] = id;
    ib[rec::cpu] = cpu;
    set_bond(ib, 0, ia[rec::id], ia[rec::cpu]);
    for (int k = rec::first_data + 6; k < rec::first_data + 9; k++) b[k] = 0.0;
  } else if (position == rec::SECOND_1 || position == rec::SECOND_2) {
    // a side branch of length 5e-4 at +-45 degrees in the horizontal plane
    double da = r.next();
    da = da > 0.5 ? 0.25 * BMX_PI : -0.25 * BMX_PI;
    const double tphi = a[rec::phi] + da;
    double g[3];
    tilt(bmx_cos(tphi), bmx_sin(tphi), 0.0, r, g);
    double gth, gph;
    to_angles(g, gth, gph);
    const double x = po.pos[0], y = po.pos[1], z = po.pos[2];
    const double L = a[rec::c_length];
    const double x1 = x - 0.5 * L * nx, y1 = y - 0.5 * L * ny, z1 = z - 0.5 * L * nz;
    const double x2 = x + 0.5 * L * nx, y2 = y + 0.5 * L * ny, z2 = z + 0.5 * L * nz;
    const double seg_len = 0.0005;
    if (position == rec::SECOND_1) {
      if (nbnd == 2) ia[rec::site1 + 2] = 1;
      else if (nbnd == 3) ia[rec::site1 + 3] = 1;
      const double ax = x1 - g[0] * seg_len, ay = y1 - g[1] * seg_len,
                   az = z1 - g[2] * seg_len;
      pn.pos[0] = 0.5 * (x1 + ax);
      pn.pos[1] = 0.5 * (y1 + ay);
      pn.pos[2] = 0.5 * (z1 + az);
      ib[rec::site1] = 2;
      ib[rec::site1 + 1] = 2;
    } else {
      if (nbnd == 2) ia[rec::site1 + 2] = 2;
      else if (nbnd == 3) ia[rec::site1 + 3] = 2;
      const double ax = x2 + g[0] * seg_len, ay = y2 + g[1] * seg_len,
                   az = z2 + g[2] * seg_len;
      pn.pos[0] = 0.5 * (x2 + ax);
      pn.pos[1] = 0.5 * (y2 + ay);
      pn.pos[2] = 0.5 * (z2 + az);
      ib[rec::site1] = 1;
      ib[rec::site1 + 1] = 1;
    }
    const double r0 = a[rec::radius];
    const double old_vol = L * r0 * r0 * BMX_PI;
    b[rec::theta] = gth;
    b[rec::phi] = gph;
    double radius = old_vol / (BMX_PI * (L + b[rec::c_length]));
    radius = sqrt(radius);
    a[rec::radius] = radius;
    b[rec::radius] = radius;
    b[rec::c_length] = seg_len;
    ia[rec::n_bnds] = nbnd == 2 ? 3 : 4;
    set_bond(ia, nbnd == 2 ? 2 : 3, id, cpu);
    ib[rec::n_bnds] = 2;
    set_bond(ib, 0, ia[rec::id], ia[rec::cpu]);
    const int from = nbnd == 2 ? 1 : 2;
    set_bond(ib, 1, ia[rec::seg1_id1 + from], ia[rec::seg1_id2 + from]);
    ib[rec::position] = rec::TIP;
    fresh(ib, true);
    ib[rec::fix_site] = 0;
    ia[rec::position] = rec::INTERIOR;
    for (int k = rec::first_data + 6; k < rec::first_data + 9; k++) b[k] = 0.0;
  }
  ib[rec::id] = id;
  ib[rec::cpu] = cpu;
  segment_volume(a);
  segment_volume(b);
  b[rec::bond_scale] = 1.0;
  ib[rec::fuse_tip] = 0;
}

// a tip splits into two new tips at +-45 degrees in the horizontal plane
__device__ void new_segments(Rec &po, Rec &p1, Rec &p2, double split_len,
                             Draws &r) {
  double *a = po.rd;
  int32_t *ia = po.ip;
  const int32_t id1 = *p1.id, cpu1 = *p1.cpu, id2 = *p2.id, cpu2 = *p2.cpu;
  const double phi = a[rec::phi];
  const double ct = bmx_cos(a[rec::theta]), st = bmx_sin(a[rec::theta]);
  const double cp = bmx_cos(phi), sp = bmx_sin(phi);
  const double nx = st * cp, ny = st * sp, nz = ct;
  const double da = 0.25 * BMX_PI;
  const double tphi1 = phi + da, tphi2 = phi - da;
  double g[2][3];
  tilt(bmx_cos(tphi1), bmx_sin(tphi1), 0.0, r, g[0]);
  tilt(bmx_cos(tphi2), bmx_sin(tphi2), 0.0, r, g[1]);
  const double x = po.pos[0], y = po.pos[1], z = po.pos[2];
  const double L = a[rec::c_length];
  const double x1 = x - 0.5 * L * nx, y1 = y - 0.5 * L * ny, z1 = z - 0.5 * L * nz;
  const double x2 = x + 0.5 * L * nx, y2 = y + 0.5 * L * ny, z2 = z + 0.5 * L * nz;
  const double seg_len = 0.5 * (L - split_len);
  const double new_len = L - 2.0 * seg_len;
  const int s = ia[rec::site1] == 1 ? 1 : 2;  // the bonded end
  double nwx, nwy, nwz;
  if (s == 1) {
    nwx = x1 + new_len * nx;
    nwy = y1 + new_len * ny;
    nwz = z1 + new_len * nz;
    po.pos[0] = 0.5 * (x1 + nwx);
    po.pos[1] = 0.5 * (y1 + nwy);
    po.pos[2] = 0.5 * (z1 + nwz);
  } else {
    nwx = x2 - new_len * nx;
    nwy = y2 - new_len * ny;
    nwz = z2 - new_len * nz;
    po.pos[0] = 0.5 * (x2 + nwx);
    po.pos[1] = 0.5 * (y2 + nwy);
    po.pos[2] = 0.5 * (z2 + nwz);
  }
  Rec *kid[2] = {&p1, &p2};
  const int32_t other_id[2] = {id2, id1}, other_cpu[2] = {cpu2, cpu1};
  for (int q = 0; q < 2; q++) {
    double ax, ay, az;
    if (s == 1) {
      ax = nwx + g[q][0] * seg_len;
      ay = nwy + g[q][1] * seg_len;
      az = nwz + g[q][2] * seg_len;
    } else {
      ax = nwx - g[q][0] * seg_len;
      ay = nwy - g[q][1] * seg_len;
      az = nwz - g[q][2] * seg_len;
    }
    Rec &k = *kid[q];
    k.pos[0] = 0.5 * (nwx + ax);
    k.pos[1] = 0.5 * (nwy + ay);
    k.pos[2] = 0.5 * (nwz + az);
    k.ip[rec::site1] = s;
    set_bond(k.ip, 0, ia[rec::id], ia[rec::cpu]);
    k.ip[rec::site1 + 1] = s;
    set_bond(k.ip, 1, other_id[q], other_cpu[q]);
    k.ip[rec::n_bnds] = 2;
  }
  const int far = 3 - s;  // the free end, where the branches attach
  const int b0 = ia[rec::n_bnds] == 1 ? 1 : 2;
  ia[rec::site1 + b0] = far;
  set_bond(ia, b0, id1, cpu1);
  ia[rec::site1 + b0 + 1] = far;
  set_bond(ia, b0 + 1, id2, cpu2);
  ia[rec::n_bnds] = b0 + 2;
  a[rec::c_length] = new_len;
  for (int q = 0; q < 2; q++) {
    Rec &k = *kid[q];
    k.rd[rec::c_length] = seg_len;
    to_angles(g[q], k.rd[rec::theta], k.rd[rec::phi]);
    k.ip[rec::position] = rec::TIP;
    k.ip[rec::fuse_flag] = 0;
    k.ip[rec::split_flag] = 0;
    k.ip[FUSE_ID] = -1;
    k.ip[FUSE_CPU] = -1;
    for (int j = rec::dltd_site1; j <= rec::dltd_cpu2; j++) k.ip[j] = -1;
    k.ip[rec::fix_site] = 0;
  }
  ia[rec::position] = rec::INTERIOR;
  p1.ip[rec::id] = id1;
  p1.ip[rec::cpu] = cpu1;
  p2.ip[rec::id] = id2;
  p2.ip[rec::cpu] = cpu2;
  segment_volume(a);
  for (int q = 0; q < 2; q++) {
    segment_volume(kid[q]->rd);
    kid[q]->rd[rec::bond_scale] = 1.0;
    kid[q]->ip[rec::fuse_tip] = 0;
  }
}

__global__ void k_split_count(uint8_t *__restrict__ aos, AosLayout lay,
                              int64_t n, SplitParams P, uint64_t seed,
                              uint32_t call, int32_t *__restrict__ cnt) {
  int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n) return;
  Rec p = rec_at(aos, lay, o);
  Draws r{seed, call, *p.id, *p.cpu, 0};
  cnt[o] = check_split(p.rd, p.ip, P, r);
}

__global__ void k_split_apply(uint8_t *__restrict__ aos, AosLayout lay,
                              int64_t n, const int32_t *__restrict__ start,
                              SplitParams P, uint64_t seed, uint32_t call,
                              int32_t next_id, int32_t cpu) {
  int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n) return;
  const int32_t off = start[o], k = start[o + 1] - off;
  if (k == 0) return;
  Rec p = rec_at(aos, lay, o);
  Draws r{seed, call, *p.id, *p.cpu, 1};
  copy_child(aos, lay, o, n + off, next_id + off, cpu);
  Rec a = rec_at(aos, lay, n + off);
  if (k == 1) {
    if (p.ip[rec::cell_type] == rec::YEAST)
      new_cell(p, a, P.overlap, r);
    else
      new_segment(p, a, P.split_len, r);
    a.ip[rec::id] = *a.id;
    a.ip[rec::cpu] = *a.cpu;
  } else {
    copy_child(aos, lay, o, n + off + 1, next_id + off + 1, cpu);
    Rec b = rec_at(aos, lay, n + off + 1);
    new_segments(p, a, b, P.split_len, r);
  }
}

int grid_of(int64_t n) { return (int)((n + 255) / 256 > 0 ? (n + 255) / 256 : 1); }

}  // namespace

hipError_t launch_split_count(uint8_t *aos, AosLayout lay, int64_t n,
                              const double sp[7], uint64_t seed, uint32_t call,
                              int32_t *cnt, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  SplitParams P{sp[0], sp[1], sp[2], sp[3], sp[4], sp[5], sp[6]};
  hipLaunchKernelGGL(k_split_count, dim3(grid_of(n)), dim3(256), 0, s, aos, lay,
                     n, P, seed, call, cnt);
  return hipGetLastError();
}

hipError_t launch_split_apply(uint8_t *aos, AosLayout lay, int64_t n,
                              const int32_t *start, const double sp[7],
                              uint64_t seed, uint32_t call, int32_t next_id,
                              int32_t cpu, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  SplitParams P{sp[0], sp[1], sp[2], sp[3], sp[4], sp[5], sp[6]};
  hipLaunchKernelGGL(k_split_apply, dim3(grid_of(n)), dim3(256), 0, s, aos, lay,
                     n, start, P, seed, call, next_id, cpu);
  return hipGetLastError();
}

}  // namespace bmx
