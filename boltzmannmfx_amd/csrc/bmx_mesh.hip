// bmx_mesh.hip -- particle <-> mesh coupling on one level mesh.
//
// Particle -> mesh: ParticleToMesh with OneToOneDeposition or
// TrilinearDeposition (src/deposition/bmx_deposition_K.H:64-151) as called
// by bmx_calc_volume_fraction (src/des/bmx_calc_volume_fraction.cpp:14-86)
// and bmx_calc_txfr_fluid (src/des/bmx_calc_txfr.cpp:22-66), and the
// particle count per cell (pc->Increment, bmx_calc_txfr.cpp:96-98).
// Mesh -> particle: bmx_calc_txfr_particle (bmx_calc_txfr.cpp:72-396):
// one_to_one_interp / trilinear_interp (src/interpolation/bmx_interp_K.H),
// the species-A gradient (compute_grad_X, src/bmx.cpp:171-195), and
// xferMeshToParticleAndUpdateChem (src/chemistry/bmx_chem_K.H:1358-1940).
//
// Deposition is a gather, not a scatter: the particles are binned by the
// cell they deposit from (their own cell, or for trilinear the base cell
// floor(l + 1/2)), each bin is put in ascending particle order, and one
// thread per mesh cell sums the contributions of its source bins in that
// order.  Every cell therefore gets the same sum, in the same order, as a
// serial loop over the particles -- no floating-point atomics.
#include "bmx_device.h"
#include "bmx_kernels.h"
#include "bmx_math.h"

namespace bmx {

namespace {

constexpr int kMaxCellSort = 4096;  // particles one bin may hold

// floor((x - plo) * dxi) per axis (bmx_deposition_K.H:78-85); false when
// the particle lies outside the mesh
__device__ __forceinline__ bool mesh_cell(const double *x, const MeshGeom &g,
                                          int c[3]) {
  bool in = true;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    double f = floor((x[d] - g.plo[d]) * g.dxi[d]);
    in = in && (f >= 0.0 && f < (double)g.n[d]);
    c[d] = in ? (int)f : 0;
  }
  return in;
}

__device__ __forceinline__ int64_t ghost_lin(const MeshGeom &g, int i, int j,
                                             int k) {
  return ((int64_t)(k + 1) * (g.n[1] + 2) + (j + 1)) * (g.n[0] + 2) + (i + 1);
}

// Bin of each particle (sorted order t): its cell (TRI = 0) or its
// trilinear base cell floor((x - plo) * dxi + 1/2), which is n on the high
// face and wraps to 0 there on a periodic axis.  Neighbouring sorted
// particles mostly share a bin, so each run of equal bins in a wave takes
// its slots with one atomic (the run head's).
template <int TRI>
__global__ __launch_bounds__(256) void k_mesh_key(
    const double4 *__restrict__ pos, int64_t n, MeshGeom g,
    int32_t *__restrict__ key, int32_t *__restrict__ rank,
    int32_t *__restrict__ count, int32_t *__restrict__ err) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int32_t k = -1;
  if (t < n) {
    const double4 p = pos[t];
    const double x[3] = {p.x, p.y, p.z};
    int c[3];
    if (mesh_cell(x, g, c)) {
      int nb0 = g.n[0], nb1 = g.n[1];
      if (TRI) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
          int b = (int)floor((x[d] - g.plo[d]) * g.dxi[d] + 0.5);
          if (g.periodic[d] && b == g.n[d]) b = 0;
          c[d] = b;
        }
        nb0 = g.periodic[0] ? g.n[0] : g.n[0] + 1;
        nb1 = g.periodic[1] ? g.n[1] : g.n[1] + 1;
      }
      k = (c[2] * nb1 + c[1]) * nb0 + c[0];
    } else {
      atomicOr(err, ERRBIT_MESH);
    }
    key[t] = k;
  }
  const bool active = k >= 0;
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(k, 1, 64);
  const bool head = active && (lane == 0 || prev != k);
  const unsigned long long heads = __ballot(head);
  const unsigned long long act = __ballot(active);
  // this lane's run: from the last head at or below it to the next head or
  // the next inactive lane
  const unsigned long long below = heads & ((2ull << lane) - 1ull);
  const int first = 63 - __builtin_clzll(below | 1ull);
  const unsigned long long stops = (heads | ~act) & ~((2ull << lane) - 1ull);
  const int stop = stops ? __builtin_ctzll(stops) : 64;
  int base = 0;
  if (head) base = atomicAdd(&count[k], stop - lane);
  base = __shfl(base, first, 64);
  if (active) rank[t] = base + (lane - first);
}

// (original index << 32 | sorted index) of each binned particle at its slot
__global__ void k_mesh_scatter(const int32_t *__restrict__ key,
                               const int32_t *__restrict__ rank,
                               const int32_t *__restrict__ start,
                               const int32_t *__restrict__ oid, int64_t n,
                               unsigned long long *__restrict__ perm) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int32_t k = key[t];
  if (k >= 0)
    perm[start[k] + rank[t]] =
        ((unsigned long long)(uint32_t)oid[t] << 32) | (uint32_t)t;
}

__device__ __forceinline__ void cswap(unsigned long long &a,
                                      unsigned long long &b) {
  const unsigned long long lo = a < b ? a : b, hi = a < b ? b : a;
  a = lo;
  b = hi;
}

// Batcher's odd-even merge sort network on N register values (all indices
// compile-time constants once unrolled).
template <int N>
__device__ __forceinline__ void sort_net(unsigned long long (&v)[N]) {
#pragma unroll
  for (int p = 1; p < N; p <<= 1)
#pragma unroll
    for (int k = p; k >= 1; k >>= 1)
#pragma unroll
      for (int j = k % p; j + k < N; j += 2 * k)
#pragma unroll
        for (int i = 0; i < k; i++)
          if (i + j + k < N && (i + j) / (2 * p) == (i + j + k) / (2 * p))
            cswap(v[i + j], v[i + j + k]);
}

constexpr int kNetSort = 16;

// Each bin in ascending original index: up to 16 entries in registers by
// the network, longer bins by insertion sort in place (bounded).
__global__ void k_mesh_sort(const int32_t *__restrict__ start, int64_t nb,
                            unsigned long long *__restrict__ perm,
                            int32_t *__restrict__ err) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int32_t a = start[b], e = start[b + 1];
  const int32_t m = e - a;
  if (m <= 1) return;
  if (m <= kNetSort) {
    unsigned long long v[kNetSort];
#pragma unroll
    for (int i = 0; i < kNetSort; i++) v[i] = i < m ? perm[a + i] : ~0ull;
    sort_net(v);
#pragma unroll
    for (int i = 0; i < kNetSort; i++)
      if (i < m) perm[a + i] = v[i];
    return;
  }
  if (m > kMaxCellSort) {
    atomicOr(err, ERRBIT_MESH);
    return;
  }
  for (int32_t i = a + 1; i < e; i++) {
    const unsigned long long v = perm[i];
    int32_t j = i - 1;
    while (j >= a && perm[j] > v) {
      perm[j + 1] = perm[j];
      j--;
    }
    perm[j + 1] = v;
  }
}

// out-mode of a deposited cell value
enum : int { OUT_SUM = 0, OUT_VF = 1, OUT_COUNT = 2 };

__device__ __forceinline__ double finish(double s, int mode, double inv_vol) {
  if (mode != OUT_VF) return s;
  // mult(1/vol) (ParticleToMesh's vol_weight), mult(-1), plus(1), max 0.1
  double v = s * inv_vol;
  v = v * -1.0;
  v = v + 1.0;
  return v < 0.1 ? 0.1 : v;
}

// One thread per cell: the cell's particles in ascending order.
__global__ void k_deposit_one(const uint8_t *__restrict__ aos, AosLayout lay,
                              MeshGeom g, const int32_t *__restrict__ start,
                              const unsigned long long *__restrict__ perm,
                              int comp0, int ncomp, int mode,
                              double *__restrict__ out, int by_sorted) {
  const int64_t nc = (int64_t)g.n[0] * g.n[1] * g.n[2];
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const int32_t a = start[c], e = start[c + 1];
  if (mode == OUT_COUNT) {
    out[c] = (double)(e - a);
    return;
  }
  double sum[3] = {0.0, 0.0, 0.0};
  for (int32_t q = a; q < e; q++) {
    const unsigned long long e = perm[q];
    const int64_t src = by_sorted ? (int64_t)(uint32_t)e : (int64_t)(e >> 32);
    const double *rd = reinterpret_cast<const double *>(
        aos + src * lay.stride + lay.rdata) + comp0;
    for (int m = 0; m < ncomp; m++) sum[m] += rd[m];
  }
  for (int m = 0; m < ncomp; m++) out[m * nc + c] = finish(sum[m], mode, g.inv_cell_vol);
}

// Weight of corner (ii, jj, kk) of a particle's trilinear stencil
// (bmx_deposition_K.H:104-136): products normalised by their sum.
__device__ __forceinline__ double tri_weight(const double4 &p, const MeshGeom &g,
                                             int ii, int jj, int kk) {
  const double x[3] = {p.x, p.y, p.z};
  double w[3][2];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    double l = (x[d] - g.plo[d]) * g.dxi[d] + 0.5;
    double f = l - (int)floor(l);
    w[d][0] = 1. - f;
    w[d][1] = f;
  }
  double total = 0.0;
  for (int a = 0; a <= 1; ++a)
    for (int b = 0; b <= 1; ++b)
      for (int c = 0; c <= 1; ++c) total += w[0][a] * w[1][b] * w[2][c];
  return w[0][ii] * w[1][jj] * w[2][kk] / total;
}

// One thread per cell, merging its 8 source bins by particle index.  Source
// bin (i + 1 - ii, ...) holds the particles whose corner (ii, jj, kk) is
// this cell.
__global__ void k_deposit_tri(const uint8_t *__restrict__ aos, AosLayout lay,
                              MeshGeom g, const int32_t *__restrict__ start,
                              const unsigned long long *__restrict__ perm,
                              const double4 *__restrict__ pos, int comp0,
                              int ncomp, int mode, double *__restrict__ out,
                              int by_sorted) {
  const int64_t nc = (int64_t)g.n[0] * g.n[1] * g.n[2];
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const int i = (int)(c % g.n[0]);
  const int j = (int)((c / g.n[0]) % g.n[1]);
  const int k = (int)(c / ((int64_t)g.n[0] * g.n[1]));
  const int nb0 = g.periodic[0] ? g.n[0] : g.n[0] + 1;
  const int nb1 = g.periodic[1] ? g.n[1] : g.n[1] + 1;
  int32_t head[8], end[8];
#pragma unroll
  for (int s = 0; s < 8; s++) {
    const int ii = s & 1, jj = (s >> 1) & 1, kk = s >> 2;
    int b0 = i + 1 - ii, b1 = j + 1 - jj, b2 = k + 1 - kk;
    if (g.periodic[0] && b0 == g.n[0]) b0 = 0;
    if (g.periodic[1] && b1 == g.n[1]) b1 = 0;
    if (g.periodic[2] && b2 == g.n[2]) b2 = 0;
    const int64_t b = ((int64_t)b2 * nb1 + b1) * nb0 + b0;
    head[s] = start[b];
    end[s] = start[b + 1];
  }
  // (original index, sorted index) at each list head, ~0 once a list is done
  unsigned long long hv[8];
#pragma unroll
  for (int s = 0; s < 8; s++) hv[s] = head[s] < end[s] ? perm[head[s]] : ~0ull;
  double sum[3] = {0.0, 0.0, 0.0};
  for (;;) {
    int best = 0;
#pragma unroll
    for (int s = 1; s < 8; s++)
      if (hv[s] < hv[best]) best = s;
    if (hv[best] == ~0ull) break;
    const int64_t bo = (int64_t)(hv[best] >> 32);
    const int64_t bt = (int64_t)(uint32_t)hv[best];
    head[best]++;
    hv[best] = head[best] < end[best] ? perm[head[best]] : ~0ull;
    const int ii = best & 1, jj = (best >> 1) & 1, kk = best >> 2;
    const double w = tri_weight(pos[bt], g, ii, jj, kk);
    const double *rd = reinterpret_cast<const double *>(
        aos + (by_sorted ? bt : bo) * lay.stride + lay.rdata);
    for (int m = 0; m < ncomp; m++) sum[m] += w * rd[comp0 + m];
  }
  for (int m = 0; m < ncomp; m++) out[m * nc + c] = finish(sum[m], mode, g.inv_cell_vol);
}

// ------------------------------------------------------- mesh -> particle

// trilinear_interp (bmx_interp_K.H:17-44) of one component of a ghosted field
__device__ __forceinline__ double trilinear(const double *x, const double *f,
                                            const MeshGeom &g) {
  double hi[3], lo[3];
  int c[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    double l = (x[d] - g.plo[d]) * g.dxi[d] + 0.5;
    c[d] = (int)floor(l);
    hi[d] = l - c[d];
    lo[d] = 1 - hi[d];
  }
  const int i = c[0], j = c[1], k = c[2];
  return lo[0] * lo[1] * lo[2] * f[ghost_lin(g, i - 1, j - 1, k - 1)] +
         lo[0] * lo[1] * hi[2] * f[ghost_lin(g, i - 1, j - 1, k)] +
         lo[0] * hi[1] * lo[2] * f[ghost_lin(g, i - 1, j, k - 1)] +
         lo[0] * hi[1] * hi[2] * f[ghost_lin(g, i - 1, j, k)] +
         hi[0] * lo[1] * lo[2] * f[ghost_lin(g, i, j - 1, k - 1)] +
         hi[0] * lo[1] * hi[2] * f[ghost_lin(g, i, j - 1, k)] +
         hi[0] * hi[1] * lo[2] * f[ghost_lin(g, i, j, k - 1)] +
         hi[0] * hi[1] * hi[2] * f[ghost_lin(g, i, j, k)];
}

// Uptake half-step of one species: the cell takes d from the fluid, redone
// if either side would go negative.  'second' selects the fluid-empty form
// of the second half-step (bmx_chem_K.H:1780-1851), which differs from the
// first's (:1503-1583).
__device__ __forceinline__ void uptake(double &c, double &f, double &d,
                                       double vol, double fvol, bool second) {
  c += d / vol;
  double df = d / fvol;
  f -= df;
  if (c < 0.0) {
    c -= d / vol;
    f += df;
    d = -c * vol;
    c = 0.0;
    df = d / fvol;
    f -= df;
  }
  if (f < 0.0) {
    c -= d / vol;
    f += df;
    if (second) {
      d = -f * fvol;
      df = -f;
    } else {
      d = f * fvol;
      df = d / fvol;
    }
    f = 0.0;
    c += d / vol;
  }
}

struct ChemParams {
  double k1, k2, k3, kr1, kr2, kr3, kg, kv, rmax, lmax;
};

// The size fields of a particle the chemistry updates (rdata 0, 1, 4, 5,
// 21, 22), held in registers over the substeps.
struct Sizes {
  double radius, c_length, area, vol, dadt, dvdt;
};

// xferMeshToParticleAndUpdateChem with the checks of the reference's CPU
// build.  c: species A, B, C in, out; rel: amounts taken from the fluid.
// Returns ERRBIT_CHEM on the conditions the reference aborts on.
__device__ int chem_update(double grid_vol, double npart, Sizes &z, double c[3],
                           double rel[3], const double fm[3], int cell_type,
                           bool is_tip, int n_bnds, double dt, int nloop,
                           const ChemParams &P) {
  const double fluid_vol = grid_vol / npart;
  double kg = P.kg, kv = P.kv;
  if (cell_type == rec::FUNGI && z.radius >= P.rmax && z.c_length >= P.lmax) {
    kg = 0.0;
    kv = 0.0;
  }
  double orig_vol = z.vol;
  double f[3];
  for (int s = 0; s < 3; s++) {
    f[s] = fm[s];
    if (f[s] < 0.0 && fabs(f[s]) < 1.0e-15) f[s] = 0.0;
    rel[s] = 0.0;
  }
  if (f[0] < 0. || f[1] < 0. || f[2] < 0.) return ERRBIT_CHEM;
  const double dtp = dt / (double)nloop;
  double new_vol = z.vol;
  double new_area = z.area;
  bool negative = false;
  for (int it = 0; it < nloop; it++) {
    double d1[3];
    d1[0] = 0.5 * dtp * new_area * (P.k1 * f[0] - P.kr1 * c[0]);
    d1[1] = 0.0;
    d1[2] = 0.5 * dtp * new_area * (P.k3 * f[2] - P.kr3 * c[2]);
    for (int s = 0; s < 3; s++) {
      uptake(c[s], f[s], d1[s], orig_vol, fluid_vol, false);
      if (c[s] < 0.0 && fabs(c[s]) < 1.0e-15) c[s] = 0.0;
      negative = negative || c[s] < 0.0;
    }
    if (f[0] < 0. || f[1] < 0. || f[2] < 0.) return ERRBIT_CHEM;

    // A <-> B + C
    const double rA = -P.k2 * c[0] + P.kr2 * c[1] * c[2];
    const double rBC = P.k2 * c[0] - P.kr2 * c[1] * c[2];
    double dA = dtp * rA, dB = dtp * rBC, dC = dtp * rBC;
    c[0] += dA;
    c[1] += dB;
    c[2] += dC;
    if (c[0] < 0.0) {
      c[0] -= dA;
      dA = -c[0];
      dB = -dA;
      dC = -dA;
      c[0] += dA;
      c[1] += dB;
      c[2] += dC;
    }
    if (c[1] < 0.0) {
      c[1] -= dB;
      dB = -c[1];
      dA = -dB;
      dC = dB;
      c[0] += dA;
      c[1] += dB;
      c[2] += dC;
    }
    if (c[2] < 0.0) {
      c[2] -= dC;
      dC = -c[2];
      dA = -dC;
      dB = dC;
      c[0] += dA;
      c[1] += dB;
      c[2] += dC;
    }

    // growth consumes B
    const double rV = kv * c[1] * orig_vol;
    const double rB = -kg * c[1];
    dB = dtp * rB;
    c[1] += dB;
    double dV = dtp * rV;
    if (c[1] < 0.0) {
      c[1] -= dB;
      const double t = -c[1] / rB;
      dV = t * rV;
      c[1] = 0.0;
    }
    negative = negative || c[0] < 0.0 || c[1] < 0.0 || c[2] < 0.0;

    // sizes
    const double r0 = z.radius, l0 = z.c_length;
    if (cell_type == rec::YEAST) {
      z.dvdt = rV;
      new_vol += dV;
      z.vol = new_vol;
      const double r = bmx_pow((3.0 * new_vol / (4.0 * BMX_PI)), 1.0 / 3.0);
      z.area = 4.0 * BMX_PI * r * r;
      new_area = z.area;
      z.dadt = 2.0 * rV / r;
      z.radius = r;
    } else if (cell_type == rec::FUNGI && is_tip) {
      z.dvdt = rV;
      new_vol += dV;
      z.vol = new_vol;
      double r, l;
      if (r0 < P.rmax && l0 < P.lmax) {
        const double tau = bmx_pow(new_vol / orig_vol, 1.0 / 3.0);
        r = tau * r0;
        l = tau * l0;
        const double dtdt = 1.0 / (3.0 * bmx_pow(new_vol, 2.0 / 3.0)) * rV /
                            bmx_pow(orig_vol, 1.0 / 3.0);
        const double drdt = dtdt * r0, dldt = dtdt * l0;
        z.dadt = 2.0 * BMX_PI * (drdt * (r + l) + r * (drdt + dldt));
        z.area = n_bnds == 0 ? 2.0 * BMX_PI * r * (r + l) : BMX_PI * r * (r + 2.0 * l);
      } else if (r0 < P.rmax) {
        const double tau = sqrt(new_vol / orig_vol);
        r = tau * r0;
        l = l0;
        const double dtdt = 1.0 / (2.0 * sqrt(new_vol)) * rV / sqrt(orig_vol);
        const double drdt = dtdt * r0;
        z.dadt = 2.0 * BMX_PI * (2.0 * r + l) * drdt;
      } else {
        const double tau = new_vol / orig_vol;
        r = r0;
        l = tau * l0;
        const double dldt = rV * l0 / orig_vol;
        z.dadt = 2.0 * BMX_PI * r * dldt;
      }
      z.radius = r;
      z.c_length = l;
    } else if (cell_type == rec::FUNGI) {
      if (r0 < P.rmax) {
        z.dvdt = rV;
        new_vol += dV;
        z.vol = new_vol;
        const double tau = sqrt(new_vol / orig_vol);
        const double r = tau * r0;
        const double dtdt = 1.0 / (2.0 * sqrt(new_vol)) * rV / sqrt(orig_vol);
        const double drdt = dtdt * r0;
        z.radius = r;
        z.area = 2.0 * BMX_PI * r * l0;
        z.dadt = 2.0 * BMX_PI * (2.0 * r + l0) * drdt;
      } else {
        z.dvdt = 0.0;
        z.vol = orig_vol;
        z.dadt = 0.0;
      }
    }

    const double ratio = orig_vol / new_vol;
    double d2[3];
    for (int s = 0; s < 3; s++) c[s] *= ratio;
    d2[0] = 0.5 * dtp * new_area * (P.k1 * f[0] - P.kr1 * c[0]);
    d2[1] = 0.0;
    d2[2] = 0.5 * dtp * new_area * (P.k3 * f[2] - P.kr3 * c[2]);
    for (int s = 0; s < 3; s++) {
      uptake(c[s], f[s], d2[s], new_vol, fluid_vol, true);
      if (c[s] < 0.0 && fabs(c[s]) < 1.0e-15) c[s] = 0.0;
      negative = negative || c[s] < 0.0;
    }
    if (f[0] < 0. || f[1] < 0. || f[2] < 0.) return ERRBIT_CHEM;
    for (int s = 0; s < 3; s++) rel[s] += -(d1[s] + d2[s]);
    orig_vol = new_vol;
  }
  return negative ? ERRBIT_CHEM : 0;
}

// Largest value over the block (non-negative doubles), in thread 0.
__device__ __forceinline__ double block_max_nonneg(double v, double *sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 1; k < (int)(blockDim.x >> 6); k++) v = fmax(v, sh[k]);
  return v;
}

// One thread per particle in sorted order: interpolation at its position,
// the chemistry, the record fields written back, and the new radius, length
// and area into the sorted arrays with their maxima (at most one atomic per
// block and word).  Bounds of 4 waves per SIMD: 128 VGPRs, no spill (135
// and 3 waves unbounded).
__global__ __launch_bounds__(256, 4) void k_mesh_to_particle(
    double *__restrict__ state, int64_t n, MeshGeom g,
    const int32_t *__restrict__ start,
    const int32_t *__restrict__ type, const Topo *__restrict__ topo,
    double4 *__restrict__ pos, double4 *__restrict__ rod,
    const double *__restrict__ X, const double *__restrict__ vf, int cnc_tri,
    int vf_tri, ChemParams P, double dt, int nloop,
    unsigned long long *__restrict__ stats, int32_t *__restrict__ err) {
  __shared__ double sh[2][4];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rmax = 0.0, lmax = 0.0;
  int e = 0;
  if (t < n) {
    double4 p = pos[t];
    double4 rr = rod[t];
    const int ty = type[t];
    const double x[3] = {p.x, p.y, p.z};
    int c[3];
    if (mesh_cell(x, g, c)) {  // else reported by the binning pass
      // the particle's line of the chemistry state, all four parts at once
      double4 *q = reinterpret_cast<double4 *>(state + t * chem::stride);
      const double4 q0 = q[0], q1 = q[1], q2 = q[2];
      bool tip = false;
      int nb = 0;
      if (ty == rec::FUNGI) {
        const Topo tp = topo[t];
        tip = tp.position == rec::TIP;
        nb = tp.n_bnds;
      }
      const int64_t ng = (int64_t)(g.n[0] + 2) * (g.n[1] + 2) * (g.n[2] + 2);
      const int64_t cell = ((int64_t)c[2] * g.n[1] + c[1]) * g.n[0] + c[0];
      const int64_t gc = ghost_lin(g, c[0], c[1], c[2]);
      double fm[3];
      for (int m = 0; m < 3; m++)
        fm[m] = cnc_tri ? trilinear(x, X + m * ng, g) : X[m * ng + gc];
      const double v = vf_tri ? trilinear(x, vf, g) : vf[gc];
      const double np = (double)(start[cell + 1] - start[cell]);
      // gradient of species A at the cell, centred differences
      const int64_t sy = g.n[0] + 2, sz = sy * (g.n[1] + 2);
      const double gx = (X[gc + 1] - X[gc - 1]) * 0.5 * g.dxi[0];
      const double gy = (X[gc + sy] - X[gc - sy]) * 0.5 * g.dxi[1];
      const double gz = (X[gc + sz] - X[gc - sz]) * 0.5 * g.dxi[2];
      Sizes z{p.w, rr.x, q1.w, q0.w, q2.x, q2.y};
      double cs[3] = {q0.x, q0.y, q0.z};
      double rel[3];
      e = chem_update(g.grid_vol * v, np, z, cs, rel, fm, ty, tip, nb, dt, nloop, P);
      q[0] = make_double4(cs[0], cs[1], cs[2], z.vol);
      q[1] = make_double4(cs[0], cs[1], cs[2], z.area);
      q[2] = make_double4(z.dadt, z.dvdt, gx, gy);
      q[3] = make_double4(gz, rel[0], rel[1], rel[2]);
      p.w = z.radius;
      if (ty == rec::FUNGI) rr.x = z.c_length;
      rr.w = z.area;
      pos[t].w = z.radius;
      rod[t] = rr;
    }
    if (p.w >= 0.0) rmax = p.w; else e |= ERRBIT_CHEM;
    if (ty == rec::FUNGI) {
      if (rr.x > 0.0) lmax = rr.x; else e |= ERRBIT_CHEM;
    }
  }
  rmax = block_max_nonneg(rmax, sh[0]);
  lmax = block_max_nonneg(lmax, sh[1]);
  if (e) atomicOr(err, e);
  if (threadIdx.x == 0) {  // non-negative doubles order as their bits
    max_u64_to(&stats[0], (unsigned long long)__double_as_longlong(rmax));
    max_u64_to(&stats[1], (unsigned long long)__double_as_longlong(lmax));
  }
}

int grid_of(int64_t n) { return (int)((n + 255) / 256 > 0 ? (n + 255) / 256 : 1); }

}  // namespace

hipError_t launch_mesh_bin(const double4 *pos, const int32_t *oid, int64_t n,
                           const MeshGeom &g, int tri, int32_t *key,
                           int32_t *rank, int32_t *count, int32_t *start,
                           int32_t *scan_scratch, unsigned long long *perm,
                           int sorted, int32_t *err, hipStream_t s) {
  int64_t nb = 1;
  for (int d = 0; d < 3; d++)
    nb *= (tri && !g.periodic[d]) ? g.n[d] + 1 : g.n[d];
  hipError_t e = launch_fill_i32(count, nb, 0, s);
  if (e != hipSuccess) return e;
  if (n > 0) {
    if (tri)
      hipLaunchKernelGGL(k_mesh_key<1>, dim3(grid_of(n)), dim3(256), 0, s, pos,
                         n, g, key, rank, count, err);
    else
      hipLaunchKernelGGL(k_mesh_key<0>, dim3(grid_of(n)), dim3(256), 0, s, pos,
                         n, g, key, rank, count, err);
  }
  e = launch_scan(count, start, nb, scan_scratch, s);
  if (e != hipSuccess) return e;
  if (!sorted || n <= 0) return hipGetLastError();
  hipLaunchKernelGGL(k_mesh_scatter, dim3(grid_of(n)), dim3(256), 0, s, key,
                     rank, start, oid, n, perm);
  hipLaunchKernelGGL(k_mesh_sort, dim3(grid_of(nb)), dim3(256), 0, s, start,
                     nb, perm, err);
  return hipGetLastError();
}

hipError_t launch_deposit(const uint8_t *aos, AosLayout lay, const MeshGeom &g,
                          int tri, const int32_t *start,
                          const unsigned long long *perm, const double4 *pos,
                          int comp0, int ncomp, int mode, double *out,
                          hipStream_t s, int by_sorted) {
  const int64_t nc = (int64_t)g.n[0] * g.n[1] * g.n[2];
  if (tri)
    hipLaunchKernelGGL(k_deposit_tri, dim3(grid_of(nc)), dim3(256), 0, s, aos,
                       lay, g, start, perm, pos, comp0, ncomp, mode, out,
                       by_sorted);
  else
    hipLaunchKernelGGL(k_deposit_one, dim3(grid_of(nc)), dim3(256), 0, s, aos,
                       lay, g, start, perm, comp0, ncomp, mode, out, by_sorted);
  return hipGetLastError();
}

hipError_t launch_mesh_to_particle(double *state, int64_t n,
                                   const MeshGeom &g, const int32_t *start,
                                   const int32_t *type,
                                   const Topo *topo, double4 *pos, double4 *rod,
                                   const double *X, const double *vf,
                                   int cnc_tri, int vf_tri,
                                   const double chempar[10], double dt,
                                   int nloop, unsigned long long *stats,
                                   int32_t *err, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  ChemParams P{chempar[0], chempar[1], chempar[2], chempar[3], chempar[4],
               chempar[5], chempar[6], chempar[7], chempar[8], chempar[9]};
  // one wave per block: each wave runs a long dependent chain per particle
  // at three waves per SIMD, and one-wave blocks let the dispatcher refill
  // each SIMD wave by wave
  const unsigned nblk = (unsigned)((n + 63) / 64);
  hipLaunchKernelGGL(k_mesh_to_particle, dim3(nblk), dim3(64), 0, s,
                     state, n, g, start, type, topo, pos, rod, X, vf,
                     cnc_tri, vf_tri, P, dt, nloop, stats, err);
  return hipGetLastError();
}

}  // namespace bmx
