// bmx_exchange.hip -- ParticleExchange: transfer of the species A, B, C
// between bonded segments over the neighbour rows
// (src/des/bmx_pc_xchng.cpp:13-257; initExchange / evaluateExchange /
// applyIncrements, src/chemistry/bmx_chem_K.H:1951-2112).
//
// The concentrations and volumes live in original particle order
// (chem = {A, B, C, vol}, inc = {dA, dB, dC, -}, the record's rdata[25..27],
// rdata[5] and rdata[28..30]); the rows, bond tables and radii are the
// sorted arrays of the last rebuild.  Each row adds only into its own
// increments, in row order: the neighbour is seen through a copy in the
// reference, so its side of the flux is dropped there too.
#include "bmx_kernels.h"

namespace bmx {

namespace {

constexpr double kPi = 3.14159265358979323846;  // M_PI of <math.h>

__global__ void k_chem_unpack(const uint8_t *__restrict__ aos, AosLayout lay,
                              int64_t n, int has_rel,
                              const int32_t *__restrict__ oid,
                              double *__restrict__ state) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double *rd = reinterpret_cast<const double *>(
      aos + (int64_t)oid[t] * lay.stride + lay.rdata);
  double4 *q = reinterpret_cast<double4 *>(state + t * chem::stride);
  q[0] = make_double4(rd[rec::first_data], rd[rec::first_data + 1],
                      rd[rec::first_data + 2], rd[rec::vol]);
  q[1] = make_double4(rd[rec::first_data + 3], rd[rec::first_data + 4],
                      rd[rec::first_data + 5], rd[rec::area]);
  q[2] = make_double4(rd[rec::dadt], rd[rec::dvdt], rd[rec::gx], rd[rec::gy]);
  q[3] = has_rel ? make_double4(rd[rec::gz], rd[rec::first_data + 6],
                                rd[rec::first_data + 7], rd[rec::first_data + 8])
                 : make_double4(rd[rec::gz], 0.0, 0.0, 0.0);
}

// Walks the sorted arrays, so each record also gets its radius and, for a
// segment, its length from where the mechanics keep them.
__global__ void k_chem_pack(uint8_t *__restrict__ aos, AosLayout lay, int64_t n,
                            int has_rel, const double *__restrict__ state,
                            const int32_t *__restrict__ oid,
                            const int32_t *__restrict__ type,
                            const double4 *__restrict__ pos,
                            const double4 *__restrict__ rod) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t o = oid[t];
  double *rd = reinterpret_cast<double *>(aos + o * (int64_t)lay.stride + lay.rdata);
  const double4 *q = reinterpret_cast<const double4 *>(state + t * chem::stride);
  const double4 a = q[0], b = q[1], c = q[2], d = q[3];
  rd[rec::radius] = pos[t].w;
  if (type[t] == rec::FUNGI) rd[rec::c_length] = rod[t].x;
  rd[rec::area] = b.w;
  rd[rec::vol] = a.w;
  rd[rec::gx] = c.z;
  rd[rec::gy] = c.w;
  rd[rec::gz] = d.x;
  rd[rec::dadt] = c.x;
  rd[rec::dvdt] = c.y;
  rd[rec::first_data] = a.x;
  rd[rec::first_data + 1] = a.y;
  rd[rec::first_data + 2] = a.z;
  rd[rec::first_data + 3] = b.x;
  rd[rec::first_data + 4] = b.y;
  rd[rec::first_data + 5] = b.z;
  if (has_rel) {
    rd[rec::first_data + 6] = d.y;
    rd[rec::first_data + 7] = d.z;
    rd[rec::first_data + 8] = d.w;
  }
}

// The flux of one species from i to j (evaluateExchange, bmx_chem_K.H:
// 2041-2068): the area-weighted difference, clamped so that neither side
// gives more than its share per bond.
__device__ __forceinline__ double species_flux(double c1, double c2, double area,
                                               double mt, double dt, double v1,
                                               double v2, int nb1, int nb2) {
  double d = (c1 - c2) * area * mt * dt;
  if (d / v1 > c1 / (double)nb1) d = c1 * v1 / (double)nb1;
  if (-d / v2 > c2 / (double)nb2) d = -c2 * v2 / (double)nb2;
  return d;
}

__global__ void k_exchange_rows(const double4 *__restrict__ pos,
                                const int32_t *__restrict__ type,
                                const Topo *__restrict__ topo,
                                const int32_t *__restrict__ nbr,
                                const int32_t *__restrict__ ncnt, int64_t ncap,
                                int64_t n, double *__restrict__ state,
                                double mtA, double mtB, double mtC, double dt) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || type[t] != rec::FUNGI) return;
  const Topo ti = topo[t];
  const double r1 = pos[t].w;
  const double4 c1 = *reinterpret_cast<const double4 *>(state + t * chem::stride);
  double ia = 0.0, ib = 0.0, ic = 0.0;  // initExchange
  const int c = ncnt[t] & ((1 << 28) - 1);  // bits 28..30: image axes
  const int32_t *rp = nbr + t;
  for (int k = 0; k < c; k++, rp += ncap) {
    const int u = *rp;
    if (type[u] != rec::FUNGI) continue;
    const Topo tj = topo[u];
    int bi, bj;
    for (bi = 0; bi < ti.n_bnds; bi++)
      if (ti.seg_id[bi & 3] == tj.self_id && ti.seg_cpu[bi & 3] == tj.self_cpu) break;
    for (bj = 0; bj < tj.n_bnds; bj++)
      if (tj.seg_id[bj & 3] == ti.self_id && tj.seg_cpu[bj & 3] == ti.self_cpu) break;
    if (!(bi < ti.n_bnds && bj < tj.n_bnds)) continue;
    const double r2 = pos[u].w;
    const double4 c2 = *reinterpret_cast<const double4 *>(state + (int64_t)u * chem::stride);
    const double area1 = kPi * r1 * r1, area2 = kPi * r2 * r2;
    const double area = 0.5 * (area1 + area2);
    const double dA = species_flux(c1.x, c2.x, area, mtA, dt, c1.w, c2.w, ti.n_bnds, tj.n_bnds);
    const double dB = species_flux(c1.y, c2.y, area, mtB, dt, c1.w, c2.w, ti.n_bnds, tj.n_bnds);
    const double dC = species_flux(c1.z, c2.z, area, mtC, dt, c1.w, c2.w, ti.n_bnds, tj.n_bnds);
    ia -= dA / c1.w;
    ib -= dB / c1.w;
    ic -= dC / c1.w;
  }
  double *inc = state + t * chem::stride;
  inc[chem::dA] = ia;
  inc[chem::dB] = ib;
  inc[chem::dC] = ic;
}

// applyIncrements (bmx_chem_K.H:2092-2112)
__global__ void k_exchange_apply(const int32_t *__restrict__ type, int64_t n,
                                 double *__restrict__ state) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || type[t] != rec::FUNGI) return;
  double *q = state + t * chem::stride;
  q[chem::A] += q[chem::dA];
  q[chem::B] += q[chem::dB];
  q[chem::C] += q[chem::dC];
}

int grid256(int64_t n) { return (int)((n + 255) / 256 > 0 ? (n + 255) / 256 : 1); }

}  // namespace

hipError_t launch_chem_unpack(const uint8_t *aos, AosLayout lay, int64_t n,
                              int has_rel, const int32_t *oid, double *state,
                              hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_chem_unpack, dim3(grid256(n)), dim3(256), 0, s, aos, lay,
                     n, has_rel, oid, state);
  return hipGetLastError();
}

hipError_t launch_chem_pack(uint8_t *aos, AosLayout lay, int64_t n,
                            int has_rel, const double *state,
                            const int32_t *oid, const int32_t *type,
                            const double4 *pos, const double4 *rod,
                            hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_chem_pack, dim3(grid256(n)), dim3(256), 0, s, aos, lay,
                     n, has_rel, state, oid, type, pos, rod);
  return hipGetLastError();
}

hipError_t launch_exchange(const double4 *pos, const int32_t *type,
                           const Topo *topo,
                           const int32_t *nbr, const int32_t *ncnt,
                           int64_t ncap, int64_t n, double *state,
                           const double xpar[3], double dt, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_exchange_rows, dim3(grid256(n)), dim3(256), 0, s, pos,
                     type, topo, nbr, ncnt, ncap, n, state, xpar[0],
                     xpar[1], xpar[2], dt);
  hipLaunchKernelGGL(k_exchange_apply, dim3(grid256(n)), dim3(256), 0, s, type,
                     n, state);
  return hipGetLastError();
}

}  // namespace bmx
