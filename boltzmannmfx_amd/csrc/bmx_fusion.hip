// bmx_fusion.hip -- fusion of hyphal tips into the sides of segments
// (BMXParticleContainer::EvaluateTipFusion, EvaluateInteriorFusion,
// CleanupFusion, src/des/bmx_pc_fusion.cpp:16-493; checkTipFusion,
// checkInteriorFusion, setSplitSegment, cleanupFusionBond,
// incrementBondScale, src/chemistry/bmx_chem_K.H:80-347, 1012-1357,
// 2494-2508).
//
// Each pass is one thread per particle over the rows of the last rebuild,
// reading and writing the device record array (original order).  The tip
// and interior passes record their decisions first and apply them after, so
// every particle sees its neighbours as they stood when the pass began (a
// tip that fuses turns interior, and would otherwise become a target of
// another tip partway through the pass); the cleanup pass writes bond lists and reads only ids, flags and
// deletion records.  Random numbers: bmx_uniform of the tip's (id, cpu),
// one draw per candidate that passes the geometric tests, in row order.
#include "bmx_kernels.h"
#include "bmx_device.h"
#include "bmx_rng.h"

namespace bmx {

namespace {

enum : int { FUSE_ID = 23, FUSE_CPU = 24, TAU_SPLIT = 23 /* rdata */ };

struct RecP {
  double *pos;
  double *rd;
  int32_t *ip;
};

__device__ __forceinline__ RecP recp(uint8_t *aos, const AosLayout &lay,
                                     int64_t o) {
  uint8_t *r = aos + o * (int64_t)lay.stride;
  return RecP{reinterpret_cast<double *>(r + lay.pos),
              reinterpret_cast<double *>(r + lay.rdata),
              reinterpret_cast<int32_t *>(r + lay.idata)};
}

struct FusionParams {
  double prob, max_sep, scale_inc, length_max;
};

// r_lm of the pair (maxInteractionDistance, K.H:23-40)
__device__ __forceinline__ double pair_reach(const double *ri, const int32_t *ii,
                                             const double *rj, const int32_t *ij,
                                             double wa) {
  const int ti = ii[rec::cell_type], tj = ij[rec::cell_type];
  if (ti == rec::YEAST && tj == rec::YEAST)
    return ri[rec::radius] + rj[rec::radius] + wa;
  if (ti == rec::FUNGI && tj == rec::FUNGI)
    return 1.5 * (ri[rec::c_length] + rj[rec::c_length]) + wa;
  return 0.0;
}

// i and j have a bonded neighbour in common (the reference's "bonded")
__device__ __forceinline__ bool share_neighbour(const int32_t *a,
                                                const int32_t *b) {
  for (int i = 0; i < a[rec::n_bnds]; i++)
    for (int j = 0; j < b[rec::n_bnds]; j++)
      if (a[rec::seg1_id1 + i] == b[rec::seg1_id1 + j] &&
          a[rec::seg1_id2 + i] == b[rec::seg1_id2 + j])
        return true;
  return false;
}

__device__ __forceinline__ void separation(const double4 &a, const double4 &b,
                                           const BoxGeom &g, double d[3]) {
  d[0] = sep_axis(a.x, b.x, g.L[0], g.halfL[0], g.periodic[0]);
  d[1] = sep_axis(a.y, b.y, g.L[1], g.halfL[1], g.periodic[1]);
  d[2] = sep_axis(a.z, b.z, g.L[2], g.halfL[2], g.periodic[2]);
}

// checkTipFusion's tests: true when tip i fuses into j, with the contact
// parameter along j in tau_out.  Reads both records, writes neither.
__device__ bool tip_test(const double rij[3], const double *ir,
                         const int32_t *ii, const double *jr, const int32_t *jj,
                         const FusionParams &P, uint64_t seed, uint32_t call,
                         int32_t id, int32_t cpu, uint32_t &draw,
                         double &tau_out) {
  if (!(ii[rec::cell_type] == rec::FUNGI && ii[rec::position] == rec::TIP &&
        ii[rec::fuse_flag] == 0 && jj[rec::cell_type] == rec::FUNGI &&
        jj[rec::position] != rec::TIP))
    return false;
  if (share_neighbour(ii, jj)) return false;
  const double c1 = ir[rec::c_length], c2 = jr[rec::c_length];
  double ct = bmx_cos(ir[rec::theta]), st = bmx_sin(ir[rec::theta]);
  double cp = bmx_cos(ir[rec::phi]), sp = bmx_sin(ir[rec::phi]);
  double nx = st * cp, ny = st * sp, nz = ct;
  const double xi1 = rij[0] - 0.5 * nx * c1, yi1 = rij[1] - 0.5 * ny * c1,
               zi1 = rij[2] - 0.5 * nz * c1;
  const double xi2 = rij[0] + 0.5 * nx * c1, yi2 = rij[1] + 0.5 * ny * c1,
               zi2 = rij[2] + 0.5 * nz * c1;
  ct = bmx_cos(jr[rec::theta]);
  st = bmx_sin(jr[rec::theta]);
  cp = bmx_cos(jr[rec::phi]);
  sp = bmx_sin(jr[rec::phi]);
  nx = st * cp;
  ny = st * sp;
  nz = ct;
  const double xj1 = -0.5 * nx * c2, yj1 = -0.5 * ny * c2, zj1 = -0.5 * nz * c2;
  const double xj2 = 0.5 * nx * c2, yj2 = 0.5 * ny * c2, zj2 = 0.5 * nz * c2;
  double tau_i, tau_j;
  min_distance(xi1, yi1, zi1, xi2, yi2, zi2, xj1, yj1, zj1, xj2, yj2, zj2,
               tau_i, tau_j);
  if (!(tau_j > 0.0 && tau_j < 1.0)) return false;
  // the closest point of the tip must be its free end
  if (!((ii[rec::site1] == 1 && tau_i == 1.0) ||
        (ii[rec::site1] == 2 && tau_i == 0.0)))
    return false;
  const double ri0 = xi1 + tau_i * (xi2 - xi1), ri1 = yi1 + tau_i * (yi2 - yi1),
               ri2 = zi1 + tau_i * (zi2 - zi1);
  const double rj0 = xj1 + tau_j * (xj2 - xj1), rj1 = yj1 + tau_j * (yj2 - yj1),
               rj2 = zj1 + tau_j * (zj2 - zj1);
  const double rx = ri0 - rj0, ry = ri1 - rj1, rz = ri2 - rj2;
  const double r = sqrt(rx * rx + ry * ry + rz * rz);
  if (!(r < P.max_sep)) return false;
  if (!(bmx_uniform(seed, call, id, cpu, draw++) < P.prob)) return false;
  tau_out = tau_j;
  return true;
}

// checkTipFusion's update of tip i after it fused into j at tau_j.
__device__ void tip_apply(double *ir, int32_t *ii, int32_t j_id, int32_t j_cpu,
                          double tau_j) {
  const int nb = ii[rec::n_bnds];
  const int site = ii[rec::site1] == 1 ? 2 : 1;
  ii[rec::fuse_flag] = 1;
  ii[FUSE_ID] = j_id;
  ii[FUSE_CPU] = j_cpu;
  ii[rec::position] = rec::INTERIOR;
  ir[TAU_SPLIT] = tau_j;
  if (nb == 1 || nb == 2) {
    ii[rec::site1 + nb] = site;
    ii[rec::seg1_id1 + nb] = j_id;
    ii[rec::seg1_id2 + nb] = j_cpu;
  }
  ii[rec::n_bnds] = nb + 1;
  ir[rec::bond_scale] = 0.0;
  ii[rec::fuse_tip] = site;
}

// Decisions of every tip against the records as they stand when the pass
// begins (no record is written here): tgt[t] = original index of the
// segment tip t fuses into (its first success in row order), or -1.  A tip
// that fuses stops testing, as the reference's tip turns interior on its
// first success and fails every later test.
__global__ void k_tip_decide(const uint8_t *__restrict__ aos, AosLayout lay,
                             int64_t n, const double4 *__restrict__ pos,
                             const int32_t *__restrict__ oid,
                             const int32_t *__restrict__ nbr,
                             const int32_t *__restrict__ ncnt, int64_t ncap,
                             BoxGeom g, double wa, FusionParams P,
                             uint64_t seed, uint32_t call,
                             int32_t *__restrict__ tgt,
                             double *__restrict__ ttau,
                             int32_t *__restrict__ fused) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  tgt[t] = -1;
  const int64_t o = oid[t];
  uint8_t *ab = const_cast<uint8_t *>(aos);
  RecP a = recp(ab, lay, o);
  if (!(a.ip[rec::cell_type] == rec::FUNGI && a.ip[rec::position] == rec::TIP))
    return;
  const double4 pi = pos[t];
  // the draws are keyed by the record's own id and cpu words
  const uint8_t *base = aos + o * (int64_t)lay.stride;
  const int32_t id = *reinterpret_cast<const int32_t *>(base + lay.id);
  const int32_t cpu = *reinterpret_cast<const int32_t *>(base + lay.cpu);
  uint32_t draw = 0;
  const int c = ncnt[t] & ((1 << 28) - 1);
  for (int k = 0; k < c; k++) {
    const int u = nbr[(int64_t)k * ncap + t];
    RecP b = recp(ab, lay, oid[u]);
    double d[3];
    separation(pi, pos[u], g, d);
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double rlm = pair_reach(a.rd, a.ip, b.rd, b.ip, wa);
    double tau;
    if (r2 <= (rlm - 1.0e-15) * (rlm - 1.0e-15) &&
        tip_test(d, a.rd, a.ip, b.rd, b.ip, P, seed, call, id, cpu, draw, tau)) {
      tgt[t] = oid[u];
      ttau[t] = tau;
      atomicOr(fused, 1);
      return;
    }
  }
}

// Applies the decisions: every fused tip updates its own record.
__global__ void k_tip_apply(uint8_t *__restrict__ aos, AosLayout lay, int64_t n,
                            const int32_t *__restrict__ oid,
                            const int32_t *__restrict__ tgt,
                            const double *__restrict__ ttau) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int32_t j = tgt[t];
  if (j < 0) return;
  RecP a = recp(aos, lay, oid[t]);
  // the target's id words are not changed by this pass
  RecP b = recp(aos, lay, j);
  tip_apply(a.rd, a.ip, b.ip[rec::id], b.ip[rec::cpu], ttau[t]);
}

// checkInteriorFusion, recorded: cnt[o] candidates, the last one's tip
__global__ void k_interior_decide(uint8_t *__restrict__ aos, AosLayout lay,
                                  int64_t n, const double4 *__restrict__ pos,
                                  const int32_t *__restrict__ oid,
                                  const int32_t *__restrict__ nbr,
                                  const int32_t *__restrict__ ncnt,
                                  int64_t ncap, BoxGeom g, double wa,
                                  int32_t *__restrict__ cnt,
                                  int4 *__restrict__ dec,
                                  double *__restrict__ dtau) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t o = oid[t];
  RecP a = recp(aos, lay, o);
  int found = 0;
  int4 last = make_int4(0, 0, 0, 0);
  double tau = 0.0;
  if (a.ip[rec::position] != rec::TIP) {
    const double4 pi = pos[t];
    const int c = ncnt[t] & ((1 << 28) - 1);
    for (int k = 0; k < c; k++) {
      const int u = nbr[(int64_t)k * ncap + t];
      RecP b = recp(aos, lay, oid[u]);
      if (!(b.ip[rec::fuse_flag] == 1)) continue;
      double d[3];
      separation(pi, pos[u], g, d);
      const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      const double rlm = pair_reach(a.rd, a.ip, b.rd, b.ip, wa);
      if (!(r2 <= (rlm - 1.0e-15) * (rlm - 1.0e-15))) continue;
      if (!(a.ip[rec::id] == b.ip[FUSE_ID] && a.ip[rec::cpu] == b.ip[FUSE_CPU]))
        continue;
      if (share_neighbour(a.ip, b.ip)) continue;
      found++;
      last = make_int4(b.ip[rec::id], b.ip[rec::cpu], 0, 0);
      tau = b.rd[TAU_SPLIT];
    }
  }
  cnt[o] = found;
  dec[o] = last;
  dtau[o] = tau;
}

__global__ void k_interior_apply(uint8_t *__restrict__ aos, AosLayout lay,
                                 int64_t n, const int32_t *__restrict__ cnt,
                                 const int4 *__restrict__ dec,
                                 const double *__restrict__ dtau,
                                 int32_t *__restrict__ one) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n) return;
  const int k = cnt[o];
  one[o] = k == 1 ? 1 : 0;
  if (k == 0) return;
  RecP a = recp(aos, lay, o);
  a.ip[rec::split_flag] = 1;
  a.ip[FUSE_ID] = dec[o].x;
  a.ip[FUSE_CPU] = dec[o].y;
  a.rd[TAU_SPLIT] = dtau[o];
}

__device__ __forceinline__ void seg_volume(double *rd) {
  const double clen = rd[rec::c_length], rad = rd[rec::radius];
  rd[rec::vol] = BMX_PI * rad * rad * clen;
  rd[rec::area] = 2.0 * BMX_PI * rad * clen + 2.0 * BMX_PI * rad * rad;
}

// setSplitSegment: o splits at tau_split, the far part becomes the new
// segment c (already a copy of o with its own id and cpu)
__device__ void split_segment(RecP &po, RecP &pc, int32_t id, int32_t cpu) {
  double *a = po.rd, *b = pc.rd;
  int32_t *io = po.ip, *ic = pc.ip;
  int32_t s_id[4], s_cpu[4], s_site[4];
  for (int k = 0; k < 4; k++) {
    s_id[k] = io[rec::seg1_id1 + k];
    s_cpu[k] = io[rec::seg1_id2 + k];
    s_site[k] = io[rec::site1 + k];
  }
  const double ct = bmx_cos(a[rec::theta]), st = bmx_sin(a[rec::theta]);
  const double cp = bmx_cos(a[rec::phi]), sp = bmx_sin(a[rec::phi]);
  const double nx = st * cp, ny = st * sp, nz = ct;
  const double L = a[rec::c_length];
  const double ox1 = po.pos[0] - 0.5 * nx * L, oy1 = po.pos[1] - 0.5 * ny * L,
               oz1 = po.pos[2] - 0.5 * nz * L;
  const double ex = po.pos[0] + 0.5 * nx * L, ey = po.pos[1] + 0.5 * ny * L,
               ez = po.pos[2] + 0.5 * nz * L;
  const double tau = a[TAU_SPLIT];
  const double ox2 = ox1 + tau * nx * L, oy2 = oy1 + tau * ny * L,
               oz2 = oz1 + tau * nz * L;
  po.pos[0] = 0.5 * (ox1 + ox2);
  po.pos[1] = 0.5 * (oy1 + oy2);
  po.pos[2] = 0.5 * (oz1 + oz2);
  pc.pos[0] = 0.5 * (ox2 + ex);
  pc.pos[1] = 0.5 * (oy2 + ey);
  pc.pos[2] = 0.5 * (oz2 + ez);
  ic[rec::site1] = 1;
  ic[rec::seg1_id1] = io[rec::id];
  ic[rec::seg1_id2] = io[rec::cpu];
  ic[rec::site1 + 1] = 1;
  ic[rec::seg1_id1 + 1] = io[FUSE_ID];
  ic[rec::seg1_id2 + 1] = io[FUSE_CPU];
  const int old_bnds = io[rec::n_bnds];
  int32_t n_id[2] = {0, 0}, n_cpu[2] = {0, 0};
  // bonds at the far end move to the new segment and are recorded as
  // deleted from o
  int nb = 0;
  for (int i = 0; i < old_bnds; i++) {
    if (s_site[i] != 2) continue;
    n_id[nb] = s_id[i];
    n_cpu[nb] = s_cpu[i];
    nb++;
    if (nb == 1) {
      io[rec::dltd_site1] = 2;
      io[rec::dltd_site1 + 2] = n_id[0];  // dltd_id1
      io[rec::dltd_site1 + 4] = n_cpu[0]; // dltd_cpu1
    } else {
      io[rec::dltd_site1 + 1] = 2;        // dltd_site2
      io[rec::dltd_site1 + 3] = n_id[1];  // dltd_id2
      io[rec::dltd_cpu2] = n_cpu[1];
      break;
    }
  }
  nb += 2;
  ic[rec::site1 + 2] = 2;
  ic[rec::seg1_id1 + 2] = n_id[0];
  ic[rec::seg1_id2 + 2] = n_cpu[0];
  if (nb > 3) {
    ic[rec::site1 + 3] = 2;
    ic[rec::seg1_id1 + 3] = n_id[1];
    ic[rec::seg1_id2 + 3] = n_cpu[1];
  }
  ic[rec::n_bnds] = nb;
  // o keeps its near-end bonds and meets the new segment and the tip
  nb = 0;
  for (int i = 0; i < old_bnds; i++) {
    if (s_site[i] != 1) continue;
    n_id[nb] = s_id[i];
    n_cpu[nb] = s_cpu[i];
    nb++;
    if (nb == 2) break;
  }
  io[rec::site1] = 1;
  io[rec::seg1_id1] = n_id[0];
  io[rec::seg1_id2] = n_cpu[0];
  if (nb > 1) {
    io[rec::site1 + 1] = 1;
    io[rec::seg1_id1 + 1] = n_id[1];
    io[rec::seg1_id2 + 1] = n_cpu[1];
  }
  io[rec::n_bnds] = nb + 2;
  const int s0 = nb == 1 ? 1 : 2;  // first of the two far-end slots
  io[rec::site1 + s0] = 2;
  io[rec::seg1_id1 + s0] = id;
  io[rec::seg1_id2 + s0] = cpu;
  io[rec::site1 + s0 + 1] = 2;
  io[rec::seg1_id1 + s0 + 1] = io[FUSE_ID];
  io[rec::seg1_id2 + s0 + 1] = io[FUSE_CPU];
  a[rec::c_length] = tau * L;
  b[rec::c_length] = (1.0 - tau) * L;
  io[rec::position] = rec::INTERIOR;
  ic[rec::position] = rec::INTERIOR;
  ic[rec::id] = id;
  ic[rec::cpu] = cpu;
  io[rec::fuse_flag] = 0;
  io[rec::split_flag] = 0;
  io[FUSE_ID] = -1;
  io[FUSE_CPU] = -1;
  ic[rec::fuse_flag] = 0;
  ic[rec::split_flag] = 0;
  ic[FUSE_ID] = -1;
  ic[FUSE_CPU] = -1;
  for (int k = rec::dltd_site1; k <= rec::dltd_cpu2; k++) ic[k] = -1;
  ic[rec::fix_site] = 2;
  ic[rec::new_flag] = 1;
  seg_volume(a);
  seg_volume(b);
  b[rec::bond_scale] = 1.0;
  ic[rec::fuse_tip] = 0;
  for (int k = rec::first_data + 6; k < rec::first_data + 9; k++) {
    a[k] = tau * a[k];
    b[k] = (1.0 - tau) * b[k];
  }
}

__global__ void k_fusion_split(uint8_t *__restrict__ aos, AosLayout lay,
                               int64_t n, const int32_t *__restrict__ cnt,
                               const int32_t *__restrict__ start,
                               int32_t next_id, int32_t cpu) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n) return;
  RecP p = recp(aos, lay, o);
  const int k = cnt[o];
  if (k == 0) {
    p.ip[rec::split_flag] = 0;
    return;
  }
  if (k != 1) return;  // two tips at once: left as the reference leaves it
  const int32_t off = start[o];
  const int64_t c = n + off;
  const uint64_t *s = reinterpret_cast<const uint64_t *>(aos + o * (int64_t)lay.stride);
  uint64_t *d = reinterpret_cast<uint64_t *>(aos + c * (int64_t)lay.stride);
  for (int w = 0; w < lay.stride / 8; w++) d[w] = s[w];
  const int32_t nid = next_id + off;
  *reinterpret_cast<int32_t *>(aos + c * (int64_t)lay.stride + lay.id) = nid;
  *reinterpret_cast<int32_t *>(aos + c * (int64_t)lay.stride + lay.cpu) = cpu;
  RecP q = recp(aos, lay, c);
  split_segment(p, q, nid, cpu);
  q.ip[rec::id] = nid;
  q.ip[rec::cpu] = cpu;
  p.ip[rec::fuse_flag] = 0;
  p.ip[rec::split_flag] = 1;
  p.ip[rec::new_flag] = 0;
  p.ip[FUSE_ID] = -1;
  p.ip[FUSE_CPU] = -1;
  q.ip[rec::fuse_flag] = 0;
  q.ip[rec::split_flag] = 0;
  q.ip[rec::new_flag] = 1;
  q.ip[FUSE_ID] = -1;
  q.ip[FUSE_CPU] = -1;
}

// cleanupFusionBond: i drops its bond to j when j split and recorded i
__device__ void cleanup_bond(int32_t *ip, const int32_t *jp) {
  int nb = ip[rec::n_bnds];
  const int32_t i_id = ip[rec::id], i_cpu = ip[rec::cpu];
  int32_t iid[4], icpu[4];
  for (int k = 0; k < 4; k++) {
    iid[k] = ip[rec::seg1_id1 + k];
    icpu[k] = ip[rec::seg1_id2 + k];
  }
  const int32_t j_id = jp[rec::id], j_cpu = jp[rec::cpu];
  for (int i = 0; i < nb; i++) {
    if (!(iid[i] == j_id && icpu[i] == j_cpu)) continue;
    if (!(ip[rec::new_flag] == 0 && ip[rec::fuse_flag] == 0 &&
          jp[rec::split_flag] == 1))
      continue;
    const bool deleted =
        (jp[rec::dltd_site1 + 2] == i_id && jp[rec::dltd_site1 + 4] == i_cpu) ||
        (jp[rec::dltd_site1 + 3] == i_id && jp[rec::dltd_cpu2] == i_cpu);
    if (!deleted) continue;
    int site = 0;
    for (int k = i; k < 3; k++) {
      if (site == 0) site = ip[rec::site1 + k];
      ip[rec::seg1_id1 + k] = ip[rec::seg1_id1 + k + 1];
      ip[rec::seg1_id2 + k] = ip[rec::seg1_id2 + k + 1];
      ip[rec::site1 + k] = ip[rec::site1 + k + 1];
    }
    if (site == 0) site = ip[rec::site1 + 3];
    ip[rec::seg1_id1 + 3] = 0;
    ip[rec::seg1_id2 + 3] = 0;
    ip[rec::site1 + 3] = 0;
    nb--;
    ip[rec::n_bnds] = nb;
    ip[rec::fix_site] = site;
  }
}

__global__ void k_fusion_cleanup(uint8_t *__restrict__ aos, AosLayout lay,
                                 int64_t n, const double4 *__restrict__ pos,
                                 const int32_t *__restrict__ oid,
                                 const int32_t *__restrict__ nbr,
                                 const int32_t *__restrict__ ncnt,
                                 int64_t ncap, BoxGeom g, double scale_inc,
                                 double reach) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  RecP a = recp(aos, lay, oid[t]);
  // incrementBondScale
  if (a.rd[rec::bond_scale] < 1.0) {
    a.rd[rec::bond_scale] += scale_inc;
    if (a.rd[rec::bond_scale] > 1.0) a.rd[rec::bond_scale] = 1.0;
  }
  if (a.rd[rec::bond_scale] >= 1.0) a.ip[rec::fuse_tip] = 0;
  const double4 pi = pos[t];
  const int c = ncnt[t] & ((1 << 28) - 1);
  for (int k = 0; k < c; k++) {
    const int u = nbr[(int64_t)k * ncap + t];
    double d[3];
    separation(pi, pos[u], g, d);
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (r2 <= (reach - 1.0e-15) * (reach - 1.0e-15))
      cleanup_bond(a.ip, recp(aos, lay, oid[u]).ip);
  }
}

int grid_of(int64_t n) { return (int)((n + 255) / 256 > 0 ? (n + 255) / 256 : 1); }

}  // namespace

hipError_t launch_tip_fusion(uint8_t *aos, AosLayout lay, int64_t n,
                             const double4 *pos, const int32_t *oid,
                             const int32_t *nbr, const int32_t *ncnt,
                             int64_t ncap, const BoxGeom &g, double wa,
                             const double fus[4], uint64_t seed, uint32_t call,
                             int32_t *tgt, double *ttau, int32_t *fused,
                             hipStream_t s) {
  if (n <= 0) return hipSuccess;
  FusionParams P{fus[0], fus[1], fus[2], fus[3]};
  hipLaunchKernelGGL(k_tip_decide, dim3(grid_of(n)), dim3(256), 0, s, aos, lay, n,
                     pos, oid, nbr, ncnt, ncap, g, wa, P, seed, call, tgt, ttau,
                     fused);
  hipLaunchKernelGGL(k_tip_apply, dim3(grid_of(n)), dim3(256), 0, s, aos, lay, n,
                     oid, tgt, ttau);
  return hipGetLastError();
}

hipError_t launch_interior_fusion(uint8_t *aos, AosLayout lay, int64_t n,
                                  const double4 *pos, const int32_t *oid,
                                  const int32_t *nbr, const int32_t *ncnt,
                                  int64_t ncap, const BoxGeom &g, double wa,
                                  int32_t *cnt, int4 *dec, double *dtau,
                                  int32_t *one, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_interior_decide, dim3(grid_of(n)), dim3(256), 0, s, aos,
                     lay, n, pos, oid, nbr, ncnt, ncap, g, wa, cnt, dec, dtau);
  hipLaunchKernelGGL(k_interior_apply, dim3(grid_of(n)), dim3(256), 0, s, aos,
                     lay, n, cnt, dec, dtau, one);
  return hipGetLastError();
}

hipError_t launch_fusion_split(uint8_t *aos, AosLayout lay, int64_t n,
                               const int32_t *cnt, const int32_t *start,
                               int32_t next_id, int32_t cpu, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fusion_split, dim3(grid_of(n)), dim3(256), 0, s, aos, lay,
                     n, cnt, start, next_id, cpu);
  return hipGetLastError();
}

hipError_t launch_fusion_cleanup(uint8_t *aos, AosLayout lay, int64_t n,
                                 const double4 *pos, const int32_t *oid,
                                 const int32_t *nbr, const int32_t *ncnt,
                                 int64_t ncap, const BoxGeom &g,
                                 double scale_inc, double reach,
                                 hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fusion_cleanup, dim3(grid_of(n)), dim3(256), 0, s, aos,
                     lay, n, pos, oid, nbr, ncnt, ncap, g, scale_inc, reach);
  return hipGetLastError();
}

}  // namespace bmx
