// bmx_adjoint.hip -- reverse (adjoint) kernels of the sphere step.
//
// Forward map per step: x' = wrap(x + dt m F(x)), with the pair law
// F(r) = -mu (r - R_A)^2 (r - R_S) (K.H:243-263) and the surface law
// (K.H:1009-1030).  With A = dF/dd = (F/r) I + (F' - F/r) u u^T, which is
// symmetric and even in d, and interaction sets that are symmetric, the
// transpose Jacobian is a gather over the same rows (the flat rows of the
// force pass):
//   lambda_i <- lambda_i + dt m [ sum_j A_ij (lambda_i - lambda_j)
//                                 + W'(z_i) lambda_i,z e_z ].
// Parameter terms of one step (summed per block into fixed slots, so the
// totals do not depend on scheduling):
//   dL/dmu  += dt m lambda_i . sum_j (F/mu) u_ij
//   dL/dW_A += dt m lambda_i . sum_j 2 mu (r - R_A)(r - R_S) u_ij
//   dL/dm   += dt lambda_i . F_i
//   dL/dr_i += dt m [ sum_j (-F'(r_ij)) u_ij . (lambda_i - lambda_j)
//                     + lambda_i,z dW/dr_i ]
// (F depends on r - R_S only, so dF/dR_S = -F'.)
#include "bmx_device.h"
#include "bmx_kernels.h"

namespace bmx {

namespace {

constexpr int kAdjBlock = 256;

__device__ __forceinline__ double img_axis(double d, const BoxGeom &g, int a) {
  if (g.periodic[a]) {
    if (d > g.halfL[a])
      d -= g.L[a];
    else if (d < -g.halfL[a])
      d += g.L[a];
  }
  return d;
}

// lambda_T = dL/dx_T for L = (1/N) sum |image(x_T - target)|^2; one
// partial of the sum per block.
__global__ __launch_bounds__(kAdjBlock) void k_adj_loss(
    const double4 *__restrict__ pos, const int32_t *__restrict__ oid,
    const double *__restrict__ target, int64_t n, BoxGeom g, double inv_n,
    double4 *__restrict__ lam, double *__restrict__ lpart) {
  __shared__ double sh[kAdjBlock];
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (t < n) {
    double4 p = pos[t];
    int64_t o = oid[t];
    double dx = img_axis(p.x - target[3 * o], g, 0);
    double dy = img_axis(p.y - target[3 * o + 1], g, 1);
    double dz = img_axis(p.z - target[3 * o + 2], g, 2);
    s = dx * dx + dy * dy + dz * dz;
    lam[t] = make_double4(2.0 * inv_n * dx, 2.0 * inv_n * dy, 2.0 * inv_n * dz, 0.0);
  }
  // fixed-shape tree: the same order every run
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kAdjBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) lpart[blockIdx.x] = sh[0];
}

// One adjoint step over the flat rows (tiles of 64 particles, one per
// wave; thread i of block b owns particle 256 b + i as in a row-by-row
// pass).  Per window of 128 entries the lanes apply the pair test
// (r_lm - small_number, then r < R_A); the pairs that pass are packed in
// entry order and their nine terms formed 64 at a time into LDS; every
// owner adds its own terms in row order.  The owner sums, the surface term
// and the per-block partials are those of a loop over the row.
constexpr int kAdjWin = 128;
constexpr int kAdjPass = 64;
struct AdjWave {
  double4 p[64];            // owners' positions
  double4 l[64];            // owners' adjoints
  int32_t pu[kAdjWin];      // packed pairs of the window: neighbor
  int32_t po[kAdjWin];      //   and owner lane
  double term[9][kAdjPass]; // terms of one pass
};
union AdjLds {
  AdjWave w[kAdjBlock / 64];
  double sh[3][kAdjBlock];
};

__global__ __launch_bounds__(kAdjBlock) void k_adj_step(
    const double4 *__restrict__ pos, const int32_t *__restrict__ cidx,
    const uint8_t *__restrict__ cown, const int32_t *__restrict__ wtot,
    const int32_t *__restrict__ ncnt, int32_t maxn,
    const int32_t *__restrict__ oid, int64_t n, BoxGeom g, ForceParams fp,
    const double4 *__restrict__ lam_in, double4 *__restrict__ lam_out,
    double *__restrict__ grad_r, double *__restrict__ acc) {
  __shared__ AdjLds lds;
  const int lane = threadIdx.x & 63;
  AdjWave &L = lds.w[threadIdx.x / 64];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = t < n;
  // Every load that does not depend on another is issued here, before
  // anything waits, from indexes valid for every lane (the last particle
  // for lanes past n, tile 0 for a wave past n): the row length, the wave's
  // entry count, the particle, its adjoint and the first window of entries
  // (a tile's region, and the padding after the last one, is always
  // readable; entries past the count are masked below).
  const bool live = t - lane < n;  // the wave's tile holds at least one particle
  const int64_t B = live ? (t / 64) * (int64_t)64 * maxn : 0;
  const int64_t tc = active ? t : n - 1;
  const int32_t raw = ncnt[tc];
  const int32_t T_l = wtot[live ? t / 64 : 0];
  const double4 p_l = pos[tc];
  const double4 li_l = lam_in[tc];
  int fu[2], fo[2];
#pragma unroll
  for (int h = 0; h < 2; h++) {
    fu[h] = cidx[B + h * 64 + lane];
    fo[h] = cown[B + h * 64 + lane];
  }
  const int T = live ? T_l : 0;
  // component by component: a select of whole vectors became a select of
  // stack addresses
  const double4 p = make_double4(active ? p_l.x : 0.0, active ? p_l.y : 0.0,
                                 active ? p_l.z : 0.0, active ? p_l.w : 0.0);
  const double4 li = make_double4(active ? li_l.x : 0.0, active ? li_l.y : 0.0,
                                  active ? li_l.z : 0.0, active ? li_l.w : 0.0);
  const int c = active ? (raw & ((1 << 28) - 1)) : 0;  // bits 28..30: image axes
  const int P = wave_incl_scan64(c) - c;
  L.p[lane] = p;
  L.l[lane] = li;
  wave_lds_sync();
  const unsigned long long below = (1ull << lane) - 1ull;
  double jl0 = 0.0, jl1 = 0.0, jl2 = 0.0;
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  double dmu = 0.0, dwa = 0.0, dr = 0.0;
  // The first window is entered even when T is 0 (then nothing passes the
  // test and nothing is added): with no branch between them and the loads
  // above, the compiler keeps those loads at the top.
  for (int w0 = 0; w0 == 0 || w0 < T; w0 += kAdjWin) {
    bool in[2];
    int uu[2], oo[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int e = w0 + h * 64 + lane;
      const bool v = e < T;
      int ue = fu[h], oe = fo[h];
      if (w0 > 0 && v) {
        ue = cidx[B + e];
        oe = cown[B + e];
      }
      uu[h] = v ? ue : 0;
      oo[h] = v ? oe : 0;
      const double4 pi = L.p[oo[h]], o = pos[uu[h]];
      double d0 = sep_axis(pi.x, o.x, g.L[0], g.halfL[0], g.periodic[0]);
      double d1 = sep_axis(pi.y, o.y, g.L[1], g.halfL[1], g.periodic[1]);
      double d2 = sep_axis(pi.z, o.z, g.L[2], g.halfL[2], g.periodic[2]);
      double r2 = d0 * d0 + d1 * d1 + d2 * d2;
      double rS = pi.w + o.w, rA = rS + fp.WA;
      double lim = rS + fp.WA - 1.0e-15;
      in[h] = v && (r2 <= lim * lim) && (sqrt(r2) < rA);
    }
    const unsigned long long b0 = __ballot(in[0]), b1 = __ballot(in[1]);
    const int n0 = __popcll(b0), ni = n0 + __popcll(b1);
    if (in[0]) {
      const int k = __popcll(b0 & below);
      L.pu[k] = uu[0];
      L.po[k] = oo[0];
    }
    if (in[1]) {
      const int k = n0 + __popcll(b1 & below);
      L.pu[k] = uu[1];
      L.po[k] = oo[1];
    }
    wave_lds_sync();
    const int lo = P > w0 ? P : w0;
    const int hi = P + c < w0 + kAdjWin ? P + c : w0 + kAdjWin;
    const int klo = lo < hi ? packed_before(lo - w0, b0, b1, n0) : 0;
    const int khi = lo < hi ? packed_before(hi - w0, b0, b1, n0) : 0;
    for (int k0 = 0; k0 < ni; k0 += kAdjPass) {
      const int k = k0 + lane;
      if (k < ni) {
        const int u = L.pu[k];
        const double4 pi = L.p[L.po[k]], lii = L.l[L.po[k]], o = pos[u];
        double d0 = sep_axis(pi.x, o.x, g.L[0], g.halfL[0], g.periodic[0]);
        double d1 = sep_axis(pi.y, o.y, g.L[1], g.halfL[1], g.periodic[1]);
        double d2 = sep_axis(pi.z, o.z, g.L[2], g.halfL[2], g.periodic[2]);
        double r2 = d0 * d0 + d1 * d1 + d2 * d2;
        double rS = pi.w + o.w, rA = rS + fp.WA;
        double r = sqrt(r2);
        double u0 = d0 / r, u1 = d1 / r, u2 = d2 / r;
        double F = -fp.mu * (r - rA) * (r - rA) * (r - rS);
        double Fp = -fp.mu * (r - rA) * (3.0 * r - 2.0 * rS - rA);
        double4 lj = lam_in[u];
        double l0 = lii.x - lj.x, l1 = lii.y - lj.y, l2 = lii.z - lj.z;
        double udl = u0 * l0 + u1 * l1 + u2 * l2;
        double uli = u0 * lii.x + u1 * lii.y + u2 * lii.z;
        double a = F / r, b = (Fp - F / r) * udl;
        L.term[0][lane] = a * l0 + b * u0;
        L.term[1][lane] = a * l1 + b * u1;
        L.term[2][lane] = a * l2 + b * u2;
        L.term[3][lane] = F * u0;
        L.term[4][lane] = F * u1;
        L.term[5][lane] = F * u2;
        L.term[6][lane] = -(r - rA) * (r - rA) * (r - rS) * uli;
        L.term[7][lane] = 2.0 * fp.mu * (r - rA) * (r - rS) * uli;
        L.term[8][lane] = -Fp * udl;
      }
      wave_lds_sync();
      const int a0 = klo > k0 ? klo : k0;
      const int a1 = khi < k0 + kAdjPass ? khi : k0 + kAdjPass;
      for (int k2 = a0; k2 < a1; k2++) {
        const int q = k2 - k0;
        jl0 += L.term[0][q];
        jl1 += L.term[1][q];
        jl2 += L.term[2][q];
        f0 += L.term[3][q];
        f1 += L.term[4][q];
        f2 += L.term[5][q];
        dmu += L.term[6][q];
        dwa += L.term[7][q];
        dr += L.term[8][q];
      }
      wave_lds_sync();
    }
  }
  double pmu = 0.0, pwa = 0.0, pm = 0.0;
  if (active) {
    const double ri = p.w;
    // growth surface
    double dz = p.z - fp.zw, rAw = ri + fp.WZ;
    double w, wp, wr;
    if (dz > 0.0 && dz < rAw) {
      w = -fp.xi * (dz - rAw) * (dz - rAw) * (dz - ri);
      wp = -fp.xi * (dz - rAw) * (3.0 * dz - 2.0 * ri - rAw);
      wr = -wp;
    } else if (dz <= 0.0) {
      w = fp.xi * rAw * rAw * ri;
      wp = 0.0;
      wr = fp.xi * (2.0 * rAw * ri + rAw * rAw);
    } else {
      w = 0.0;
      wp = 0.0;
      wr = 0.0;
    }
    w -= fp.grav * fp.xi * rAw * rAw * ri;
    wr -= fp.grav * fp.xi * (2.0 * rAw * ri + rAw * rAw);
    f2 += w;
    const double dtm = fp.dt * fp.mob;
    lam_out[t] = make_double4(li.x + dtm * jl0, li.y + dtm * jl1,
                              li.z + dtm * (jl2 + wp * li.z), 0.0);
    grad_r[oid[t]] += dtm * (dr + wr * li.z);
    pmu = dtm * dmu;
    pwa = dtm * dwa;
    pm = fp.dt * (li.x * f0 + li.y * f1 + li.z * f2);
  }
  __syncthreads();  // the partials reuse the window buffers
  lds.sh[0][threadIdx.x] = pmu;
  lds.sh[1][threadIdx.x] = pwa;
  lds.sh[2][threadIdx.x] = pm;
  __syncthreads();
  for (int w = kAdjBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      lds.sh[0][threadIdx.x] += lds.sh[0][threadIdx.x + w];
      lds.sh[1][threadIdx.x] += lds.sh[1][threadIdx.x + w];
      lds.sh[2][threadIdx.x] += lds.sh[2][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x < 3) acc[3 * blockIdx.x + threadIdx.x] += lds.sh[threadIdx.x][0];
}

// out[oid[t]] = in[t] (sorted -> original order) and its inverse.
__global__ void k_to_orig4(const double4 *__restrict__ in,
                           const int32_t *__restrict__ oid, int64_t n,
                           double4 *__restrict__ out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[oid[t]] = in[t];
}
__global__ void k_from_orig4(const double4 *__restrict__ in,
                             const int32_t *__restrict__ oid, int64_t n,
                             double4 *__restrict__ out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = in[oid[t]];
}

// Fixed-order sum of per-block partials (one block, fixed tree).
__global__ __launch_bounds__(256) void k_sum_partials(
    const double *__restrict__ part, int64_t nblk, int stride,
    double *__restrict__ out) {
  __shared__ double sh[256];
  for (int c = 0; c < stride; c++) {
    double s = 0.0;
    for (int64_t b = threadIdx.x; b < nblk; b += 256) s += part[b * stride + c];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = sh[0];
    __syncthreads();
  }
}

int blocks(int64_t n, int b) {
  int64_t g = (n + b - 1) / b;
  return (int)(g < 1 ? 1 : g);
}

}  // namespace

int32_t adj_blocks(int64_t n) { return blocks(n, kAdjBlock); }

hipError_t launch_adj_loss(const double4 *pos, const int32_t *oid,
                           const double *target, int64_t n, BoxGeom g,
                           double4 *lam, double *lpart, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_loss, dim3(blocks(n, kAdjBlock)), dim3(kAdjBlock), 0,
                     s, pos, oid, target, n, g, 1.0 / (double)n, lam, lpart);
  return hipGetLastError();
}

hipError_t launch_adj_step(const double4 *pos, const int32_t *cidx,
                           const uint8_t *cown, const int32_t *wtot,
                           const int32_t *ncnt, int32_t maxn,
                           const int32_t *oid, int64_t n, BoxGeom g,
                           ForceParams fp, const double4 *lam_in,
                           double4 *lam_out, double *grad_r, double *acc,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_step, dim3(blocks(n, kAdjBlock)), dim3(kAdjBlock), 0,
                     s, pos, cidx, cown, wtot, ncnt, maxn, oid, n, g, fp,
                     lam_in, lam_out, grad_r, acc);
  return hipGetLastError();
}

hipError_t launch_to_orig4(const double4 *in, const int32_t *oid, int64_t n,
                           double4 *out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_orig4, dim3(blocks(n, 256)), dim3(256), 0, s, in, oid,
                     n, out);
  return hipGetLastError();
}

hipError_t launch_from_orig4(const double4 *in, const int32_t *oid, int64_t n,
                             double4 *out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_from_orig4, dim3(blocks(n, 256)), dim3(256), 0, s, in,
                     oid, n, out);
  return hipGetLastError();
}

hipError_t launch_sum_partials(const double *part, int64_t nblk, int stride,
                               double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, part, nblk,
                     stride, out);
  return hipGetLastError();
}

}  // namespace bmx
