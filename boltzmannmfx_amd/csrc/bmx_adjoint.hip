// bmx_adjoint.hip -- reverse (adjoint) kernels of the sphere step.
//
// Forward map per step: x' = wrap(x + dt m F(x)), with the pair law
// F(r) = -mu (r - R_A)^2 (r - R_S) (K.H:243-263) and the surface law
// (K.H:1009-1030).  With A = dF/dd = (F/r) I + (F' - F/r) u u^T, which is
// symmetric and even in d, and interaction sets that are symmetric, the
// transpose Jacobian is a gather over the same rows:
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

__global__ __launch_bounds__(kAdjBlock) void k_adj_step(
    const double4 *__restrict__ pos, const int32_t *__restrict__ nbr,
    const int32_t *__restrict__ ncnt, int64_t ncap,
    const int32_t *__restrict__ oid, int64_t n, BoxGeom g, ForceParams fp,
    const double4 *__restrict__ lam_in, double4 *__restrict__ lam_out,
    double *__restrict__ grad_r, double *__restrict__ acc) {
  __shared__ double sh[3][kAdjBlock];
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double pmu = 0.0, pwa = 0.0, pm = 0.0;
  if (t < n) {
    const double4 p = pos[t];
    const double4 li = lam_in[t];
    const double ri = p.w;
    const int c = ncnt[t] & ((1 << 28) - 1);  // bits 28..30: image axes
    double jl0 = 0.0, jl1 = 0.0, jl2 = 0.0;
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    double dmu = 0.0, dwa = 0.0, dr = 0.0;
    for (int k = 0; k < c; k++) {
      int u = nbr[(int64_t)k * ncap + t];
      double4 o = pos[u];
      double d0 = sep_axis(p.x, o.x, g.L[0], g.halfL[0], g.periodic[0]);
      double d1 = sep_axis(p.y, o.y, g.L[1], g.halfL[1], g.periodic[1]);
      double d2 = sep_axis(p.z, o.z, g.L[2], g.halfL[2], g.periodic[2]);
      double r2 = d0 * d0 + d1 * d1 + d2 * d2;
      double rS = ri + o.w, rA = rS + fp.WA;
      double lim = rS + fp.WA - 1.0e-15;
      if (!(r2 <= lim * lim)) continue;
      double r = sqrt(r2);
      if (!(r < rA)) continue;
      double u0 = d0 / r, u1 = d1 / r, u2 = d2 / r;
      double F = -fp.mu * (r - rA) * (r - rA) * (r - rS);
      double Fp = -fp.mu * (r - rA) * (3.0 * r - 2.0 * rS - rA);
      double4 lj = lam_in[u];
      double l0 = li.x - lj.x, l1 = li.y - lj.y, l2 = li.z - lj.z;
      double udl = u0 * l0 + u1 * l1 + u2 * l2;
      double uli = u0 * li.x + u1 * li.y + u2 * li.z;
      double a = F / r, b = (Fp - F / r) * udl;
      jl0 += a * l0 + b * u0;
      jl1 += a * l1 + b * u1;
      jl2 += a * l2 + b * u2;
      f0 += F * u0;
      f1 += F * u1;
      f2 += F * u2;
      dmu += -(r - rA) * (r - rA) * (r - rS) * uli;
      dwa += 2.0 * fp.mu * (r - rA) * (r - rS) * uli;
      dr += -Fp * udl;
    }
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
  sh[0][threadIdx.x] = pmu;
  sh[1][threadIdx.x] = pwa;
  sh[2][threadIdx.x] = pm;
  __syncthreads();
  for (int w = kAdjBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
      sh[2][threadIdx.x] += sh[2][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x < 3) acc[3 * blockIdx.x + threadIdx.x] += sh[threadIdx.x][0];
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

hipError_t launch_adj_step(const double4 *pos, const int32_t *nbr,
                           const int32_t *ncnt, int64_t ncap,
                           const int32_t *oid, int64_t n, BoxGeom g,
                           ForceParams fp, const double4 *lam_in,
                           double4 *lam_out, double *grad_r, double *acc,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_step, dim3(blocks(n, kAdjBlock)), dim3(kAdjBlock), 0,
                     s, pos, nbr, ncnt, ncap, oid, n, g, fp, lam_in, lam_out,
                     grad_r, acc);
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
