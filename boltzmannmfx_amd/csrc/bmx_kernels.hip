// bmx_kernels.hip -- gfx950 kernels of the particle-mechanics path.
//
// Data live in bin-sorted order.  A rebuild bins every particle
// (floor((x - plo) * dxi), src/interpolation/bmx_interp_K.H:55-56), counts
// bins with atomics, scans the counts, scatters (original index, slot)
// pairs, orders each bin by original index, gathers every per-particle array
// into the new order and builds column-major neighbor rows over the +-1 bin
// stencil (spheres: then copied into a flat layout, each wave's rows back to
// back).  Rows are visited in a fixed order (stencil offsets dz, dy, dx
// lexicographic, ascending original index inside a bin), so the force pass
// sums each particle's pair terms in an order that does not depend on
// thread scheduling.  The force kernel is a gather: particle i reads its row
// and adds only into its own accumulators, as the reference's force lambda
// does (src/des/bmx_pc_interaction.cpp:170-279), then applies the overdamped
// update (src/des/bmx_pc_interaction.cpp:311-449) in the same launch,
// writing the next state into the other half of a ping-pong pair.
#include <type_traits>

#include "bmx_device.h"
#include "bmx_kernels.h"

namespace bmx {

namespace {

constexpr int kWave = 64;

__device__ __forceinline__ unsigned long long dbits(double x) {
  return (unsigned long long)__double_as_longlong(x);
}

// Hardware dispatches workgroup b to XCD b % 8.  With remap, XCD x instead
// works on a contiguous range of tiles, so neighbor gathers of adjacent
// tiles share that XCD's L2.
__device__ __forceinline__ int64_t tile_of(int remap) {
  int64_t b = blockIdx.x, nb = gridDim.x;
  if (!remap) return b;
  int64_t x = b & 7, r = b >> 3, per = nb >> 3, extra = nb & 7;
  return x * per + (x < extra ? x : extra) + r;
}

// Block-wide max / sum through LDS (sh: one slot per wave); every thread
// must call it; the result is valid in thread 0.
template <typename T>
__device__ __forceinline__ T block_max(T v, T *sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    T w = __shfl_xor(v, o, kWave);
    v = w > v ? w : v;
  }
  int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < (int)(blockDim.x / kWave); w++) v = sh[w] > v ? sh[w] : v;
  return v;
}
template <typename T>
__device__ __forceinline__ T block_sum(T v, T *sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < (int)(blockDim.x / kWave); w++) v += sh[w];
  return v;
}

template <typename T>
__global__ void k_fill(T *p, int64_t n, T v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

int grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  return (int)g;
}
int grid_capped(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (int)g;
}

// ------------------------------------------------------------- AoS I/O ---

__global__ void k_aos_unpack(const uint8_t *__restrict__ aos, AosLayout lay,
                             int64_t n, const int32_t *__restrict__ col,
                             const BoxGeom *__restrict__ boxes, StaticArrays st,
                             DynArrays dy, unsigned long long *stats) {
  __shared__ double shd[16];
  __shared__ unsigned long long shu[16];
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rmax = 0.0, lmax = 0.0;
  unsigned long long flags = 0, nbad = 0;
  if (t < n) {
  const uint8_t *r = aos + t * (int64_t)lay.stride;
  const double *pos = reinterpret_cast<const double *>(r + lay.pos);
  const double *rd = reinterpret_cast<const double *>(r + lay.rdata);
  const int32_t *id = reinterpret_cast<const int32_t *>(r + lay.idata);
  const BoxGeom &g = boxes[col ? col[t] : 0];
  double x = wrap_axis(pos[0], g.plo[0], g.phi[0], g.L[0], g.periodic[0]);
  double y = wrap_axis(pos[1], g.plo[1], g.phi[1], g.L[1], g.periodic[1]);
  double z = wrap_axis(pos[2], g.plo[2], g.phi[2], g.L[2], g.periodic[2]);
  double rad = rd[rec::radius];
  dy.pos[t] = make_double4(x, y, z, rad);
  st.oid[t] = (int32_t)t;
  int ty = id[rec::cell_type];
  st.type[t] = ty;
  if (st.col) st.col[t] = col[t];
  bool bad = !(rad >= 0.0) || !(x == x) || !(y == y) || !(z == z);
  if (ty == rec::YEAST) flags |= 1ull;
  if (ty == rec::FUNGI) flags |= 2ull;
  if (rad >= 0.0) rmax = rad;
  if (dy.rod) {
    double L = rd[rec::c_length], th = rd[rec::theta], ph = rd[rec::phi];
    dy.rod[t] = make_double4(L, th, ph, rd[rec::area]);
    dy.trig[t] = make_double4(bmx_cos(th), bmx_sin(th), bmx_cos(ph), bmx_sin(ph));
    st.bscale[t] = rd[rec::bond_scale];
    Topo tp;
    tp.n_bnds = id[rec::n_bnds];
    for (int b = 0; b < 4; b++) {
      tp.seg_id[b] = id[rec::seg1_id1 + b];
      tp.seg_cpu[b] = id[rec::seg1_id2 + b];
      tp.site[b] = id[rec::site1 + b];
    }
    tp.position = id[rec::position];
    tp.fuse_tip = id[rec::fuse_tip];
    tp.fix_site = id[rec::fix_site];
    tp.self_id = id[rec::id];
    tp.self_cpu = id[rec::cpu];
    tp.dltd = 0;
    tp.pad = 0;
    dy.topo[t] = tp;
    if (ty == rec::FUNGI) {
      if (L >= 0.0) lmax = L;
      if (tp.n_bnds < 0 || tp.n_bnds > 4 || !(L > 0.0)) bad = true;
    }
  }
  if (bad) nbad = 1;
  }
  rmax = block_max(rmax, shd);
  __syncthreads();
  lmax = block_max(lmax, shd);
  __syncthreads();
  nbad = block_sum(nbad, shu);
  __syncthreads();
  unsigned long long fy = block_max(flags & 1ull, shu);
  __syncthreads();
  unsigned long long ff = block_max(flags & 2ull, shu);
  if (threadIdx.x == 0) {
    if (rmax > 0.0) max_u64_to(&stats[0], dbits(rmax));
    if (lmax > 0.0) max_u64_to(&stats[1], dbits(lmax));
    if (fy) or_u64_to(&stats[2], 1ull);
    if (ff) or_u64_to(&stats[3], 1ull);
    if (nbad) atomicAdd(&stats[4], nbad);
  }
}

// Re-reads the selected fields of every bound particle from its record
// into the current sorted arrays (slot t holds original index st.oid[t]),
// the inverse of k_aos_pack for those fields; the sorted order, the static
// arrays and everything not selected are kept.  stats as in k_aos_unpack
// (maxima over the particles after the update, bad records).
__global__ void k_aos_refresh(const uint8_t *__restrict__ aos, AosLayout lay,
                              int64_t n, const BoxGeom *__restrict__ boxes,
                              StaticArrays st, DynArrays dy, uint32_t fields,
                              unsigned long long *stats) {
  __shared__ double shd[16];
  __shared__ unsigned long long shu[16];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rmax = 0.0, lmax = 0.0;
  unsigned long long nbad = 0;
  if (t < n) {
    const int64_t o = st.oid[t];
    const uint8_t *r = aos + o * (int64_t)lay.stride;
    const double *pos = reinterpret_cast<const double *>(r + lay.pos);
    const double *rd = reinterpret_cast<const double *>(r + lay.rdata);
    const int32_t *id = reinterpret_cast<const int32_t *>(r + lay.idata);
    const BoxGeom &g = boxes[st.col ? st.col[t] : 0];
    double4 p = dy.pos[t];
    if (fields & kFieldPositions) {
      p.x = wrap_axis(pos[0], g.plo[0], g.phi[0], g.L[0], g.periodic[0]);
      p.y = wrap_axis(pos[1], g.plo[1], g.phi[1], g.L[1], g.periodic[1]);
      p.z = wrap_axis(pos[2], g.plo[2], g.phi[2], g.L[2], g.periodic[2]);
    }
    if (fields & kFieldSizes) p.w = rd[rec::radius];
    dy.pos[t] = p;
    bool bad = !(p.w >= 0.0) || !(p.x == p.x) || !(p.y == p.y) || !(p.z == p.z);
    if (p.w >= 0.0) rmax = p.w;
    const int ty = st.type[t];
    if (dy.rod) {
      double4 rr = dy.rod[t];
      if (fields & kFieldSizes) {
        rr.x = rd[rec::c_length];
        rr.w = rd[rec::area];
        st.bscale[t] = rd[rec::bond_scale];
      }
      if (fields & kFieldOrientations) {
        rr.y = rd[rec::theta];
        rr.z = rd[rec::phi];
        dy.trig[t] = make_double4(bmx_cos(rr.y), bmx_sin(rr.y), bmx_cos(rr.z),
                                  bmx_sin(rr.z));
      }
      dy.rod[t] = rr;
      if (fields & kFieldBonds) {
        Topo tp;
        tp.n_bnds = id[rec::n_bnds];
        for (int b = 0; b < 4; b++) {
          tp.seg_id[b] = id[rec::seg1_id1 + b];
          tp.seg_cpu[b] = id[rec::seg1_id2 + b];
          tp.site[b] = id[rec::site1 + b];
        }
        tp.position = id[rec::position];
        tp.fuse_tip = id[rec::fuse_tip];
        tp.fix_site = id[rec::fix_site];
        tp.self_id = id[rec::id];
        tp.self_cpu = id[rec::cpu];
        tp.dltd = 0;
        tp.pad = 0;
        dy.topo[t] = tp;
      }
      if (ty == rec::FUNGI) {
        if (rr.x >= 0.0) lmax = rr.x;
        const int nb = dy.topo[t].n_bnds;
        if (nb < 0 || nb > 4 || !(rr.x > 0.0)) bad = true;
      }
    }
    if (bad) nbad = 1;
  }
  rmax = block_max(rmax, shd);
  __syncthreads();
  lmax = block_max(lmax, shd);
  __syncthreads();
  nbad = block_sum(nbad, shu);
  if (threadIdx.x == 0) {
    if (rmax > 0.0) max_u64_to(&stats[0], dbits(rmax));
    if (lmax > 0.0) max_u64_to(&stats[1], dbits(lmax));
    if (nbad) atomicAdd(&stats[4], nbad);
  }
}

// fc of the last step is kept in original particle order, six doubles per
// particle (force, then torque) in one 48-byte run: a step that writes it
// stores three 16-byte pieces per particle where the component-major
// layout of the export would scatter six 8-byte words.
__device__ __forceinline__ void store_fc(double *__restrict__ fc, int64_t o,
                                         double f0, double f1, double f2,
                                         double f3, double f4, double f5) {
  double2 *q = reinterpret_cast<double2 *>(fc + o * 6);
  q[0] = make_double2(f0, f1);
  q[1] = make_double2(f2, f3);
  q[2] = make_double2(f4, f5);
}

__global__ void k_aos_pack(uint8_t *__restrict__ aos, AosLayout lay, int64_t n,
                           StaticArrays st, DynArrays dy,
                           const double *__restrict__ fc, double mob,
                           int wrote_step, int any_fungi) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int64_t o = st.oid[t];
  uint8_t *r = aos + o * (int64_t)lay.stride;
  double *pos = reinterpret_cast<double *>(r + lay.pos);
  double *rd = reinterpret_cast<double *>(r + lay.rdata);
  int32_t *id = reinterpret_cast<int32_t *>(r + lay.idata);
  double4 p = dy.pos[t];
  pos[0] = p.x;
  pos[1] = p.y;
  pos[2] = p.z;
  if (wrote_step) {
    const double *f = fc + o * 6;  // original order, six per particle
    rd[rec::velx] = mob * f[0];
    rd[rec::vely] = mob * f[1];
    rd[rec::velz] = mob * f[2];
    rd[rec::wx] = f[3];
    rd[rec::wy] = f[4];
    rd[rec::wz] = f[5];
    id[rec::fuse_flag] = 0;
    id[rec::split_flag] = 0;
    id[rec::new_flag] = 0;
  }
  if (any_fungi && st.type[t] == rec::FUNGI) {
    double4 rr = dy.rod[t];
    rd[rec::theta] = rr.y;
    rd[rec::phi] = rr.z;
    Topo tp = dy.topo[t];
    id[rec::n_bnds] = tp.n_bnds;
    for (int b = 0; b < 4; b++) {
      id[rec::seg1_id1 + b] = tp.seg_id[b];
      id[rec::seg1_id2 + b] = tp.seg_cpu[b];
      id[rec::site1 + b] = tp.site[b];
    }
    id[rec::position] = tp.position;
    id[rec::fix_site] = tp.fix_site;
    if (tp.dltd)
      for (int k = rec::dltd_site1; k <= rec::dltd_cpu2; k++) id[k] = -1;
  }
}

// ------------------------------------------------------------- binning ---

// blocks of the bin count that take part in a folded displacement check
constexpr int kSkinBlocks = 64;
constexpr int kSkinGroup = 16;  // slots reduced together
template <bool MULTI>
__global__ void k_bin_count(const double4 *__restrict__ pos,
                            const int32_t *__restrict__ col,
                            const BoxGeom *__restrict__ boxes, BoxGeom g0,
                            int64_t n, int32_t *__restrict__ key,
                            int32_t *__restrict__ rank,
                            int32_t *__restrict__ count, SkinCheck sk) {
  // The input is in the order of the previous rebuild, so neighboring lanes
  // mostly share a bin: each run of equal bins in the wave takes its slots
  // with one atomic (the head's) and numbers its lanes from there.  Only the
  // slot set of a bin matters: the bins are sorted by original index next.
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = s < n;
  int b = -1;
  if (active) {
    const BoxGeom &g = MULTI ? boxes[col[s]] : g0;
    b = bin_of(pos[s], g);
    key[s] = b;
  }
  const int lane = threadIdx.x & (kWave - 1);
  const int prev = __shfl_up(b, 1, kWave);
  const bool head = active && (lane == 0 || prev != b);
  const unsigned long long heads = __ballot(head);
  const unsigned long long act = __ballot(active);
  // run of this lane: from the last head at or below it to the next head
  const unsigned long long below = heads & ((2ull << lane) - 1ull);
  const int first = 63 - __builtin_clzll(below | 1ull);
  const unsigned long long above = heads & ~((2ull << lane) - 1ull);
  const int stop = above ? __builtin_ctzll(above) : 64 - __builtin_clzll(act);
  int base = 0;
  if (head) base = atomicAdd(&count[b], stop - lane);
  base = __shfl(base, first, kWave);
  if (active) rank[s] = base + (lane - first);
  // the first kSkinBlocks blocks also reduce the displacement partials of
  // the period this rebuild ends, one share of every slot each, and clear
  // them (block-uniform branch: the block reductions synchronise).  Up to
  // kSkinGroup slots at a time: every load of the group is issued before
  // any reduction, and the group's block maxima share one barrier.
  const int nsb = gridDim.x < kSkinBlocks ? (int)gridDim.x : kSkinBlocks;
  if (sk.vpart && (int)blockIdx.x < nsb) {
    // blocks of 256 (launch_bin_count): four waves
    __shared__ double shd[kSkinGroup][256 / kWave];
    const int64_t chunk = (sk.stride + nsb - 1) / nsb;
    const int64_t c0 = blockIdx.x * chunk;
    const int64_t c1 = c0 + chunk < sk.stride ? c0 + chunk : sk.stride;
    const int wid = threadIdx.x / kWave;
    for (int g = 0; g < sk.nslots; g += kSkinGroup) {
      double m[kSkinGroup];
#pragma unroll
      for (int j = 0; j < kSkinGroup; j++) m[j] = 0.0;
      for (int64_t i = c0 + threadIdx.x; i < c1; i += blockDim.x) {
#pragma unroll
        for (int j = 0; j < kSkinGroup; j++) {
          if (g + j < sk.nslots) {
            double *v = sk.vpart + (int64_t)(g + j) * sk.stride + i;
            const double x = *v;
            m[j] = x > m[j] ? x : m[j];
            *v = 0.0;
          }
        }
      }
#pragma unroll
      for (int j = 0; j < kSkinGroup; j++) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
          const double w = __shfl_xor(m[j], o, kWave);
          m[j] = w > m[j] ? w : m[j];
        }
        if (lane == 0) shd[j][wid] = m[j];
      }
      __syncthreads();
      const int j = threadIdx.x;
      if (j < kSkinGroup && g + j < sk.nslots) {
        double mm = shd[j][0];
        for (int w = 1; w < (int)(blockDim.x / kWave); w++)
          mm = shd[j][w] > mm ? shd[j][w] : mm;
        if (mm > 0.0) atomicMax(&sk.slotmax[g + j], dbits(mm));
      }
      __syncthreads();
    }
  }
}

// Exclusive scan: 256 threads x 8 items per block, wave prefix by shuffles.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
static_assert(kScanItems == 8, "k_scan_local moves a thread's counts as two int4");
constexpr int kScanTile = kScanThreads * kScanItems;

template <typename T>
__device__ __forceinline__ T block_excl_scan(T s, T *wsum, T &total) {
  int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  T incl = s;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    T y = __shfl_up(incl, o, kWave);
    if (lane >= o) incl += y;
  }
  if (lane == kWave - 1) wsum[wid] = incl;
  __syncthreads();
  T pre = 0, tot = 0;
  int nw = blockDim.x / kWave;
  for (int w = 0; w < nw; w++) {
    if (w < wid) pre += wsum[w];
    tot += wsum[w];
  }
  total = tot;
  return pre + incl - s;
}

// The counts are consumed here: each is reset to zero once read, which
// leaves the count array cleared for the next binning.
__global__ __launch_bounds__(kScanThreads) void k_scan_local(
    int32_t *__restrict__ in, int32_t *__restrict__ out,
    int32_t *__restrict__ bsum, int64_t n, SkinCheck sk) {
  __shared__ int32_t wsum[kScanThreads / kWave];
  if (sk.vpart && blockIdx.x == 0 && threadIdx.x == 0) {
    // the folded displacement check: the slot maxima the bin count left,
    // summed in slot order (as k_skin_final)
    double total = 0.0;
    for (int sl = 0; sl < sk.nslots; sl++) {
      total += __longlong_as_double((long long)sk.slotmax[sl]);
      sk.slotmax[sl] = 0ull;
    }
    if (total > sk.limit) atomicOr(sk.err, ERRBIT_SKIN);
  }
  int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  // a thread's 8 counts are 32 aligned bytes (the arrays are allocation
  // bases): two 16-byte loads and stores, element by element at the tail
  const bool whole = base + kScanItems <= n;
  int32_t x[kScanItems];
  if (whole) {
    int4 *vin = reinterpret_cast<int4 *>(in + base);
    const int4 a = vin[0], b = vin[1];
    x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w;
    x[4] = b.x; x[5] = b.y; x[6] = b.z; x[7] = b.w;
    vin[0] = make_int4(0, 0, 0, 0);
    vin[1] = make_int4(0, 0, 0, 0);
  } else {
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
      x[i] = 0;
      if (base + i < n) {
        x[i] = in[base + i];
        in[base + i] = 0;
      }
    }
  }
  int32_t v[kScanItems];
  int32_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    v[i] = s;
    s += x[i];
  }
  int32_t total;
  int32_t ex = block_excl_scan(s, wsum, total);
  if (whole) {
    int4 *vout = reinterpret_cast<int4 *>(out + base);
    vout[0] = make_int4(v[0] + ex, v[1] + ex, v[2] + ex, v[3] + ex);
    vout[1] = make_int4(v[4] + ex, v[5] + ex, v[6] + ex, v[7] + ex);
  } else {
#pragma unroll
    for (int i = 0; i < kScanItems; i++)
      if (base + i < n) out[base + i] = v[i] + ex;
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Adds each tile's offset, the sum of the tile sums before it, which every
// block forms itself from the (L2-resident) tile sums, so no separate pass
// scans them; the last block also writes the grand total to out[n].
__global__ __launch_bounds__(kScanThreads) void k_scan_add(
    int32_t *__restrict__ out, const int32_t *__restrict__ bsum, int64_t nt,
    int64_t n) {
  __shared__ int32_t wsum[kScanThreads / kWave];
  int32_t part = 0;
  for (int64_t i = threadIdx.x; i < (int64_t)blockIdx.x; i += kScanThreads)
    part += bsum[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) wsum[threadIdx.x / kWave] = part;
  __syncthreads();
  int32_t add = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / kWave; w++) add += wsum[w];
  if (threadIdx.x == 0 && (int64_t)blockIdx.x == nt - 1) out[n] = add + bsum[blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  if (base + kScanItems <= n) {  // 32 aligned bytes, as in k_scan_local
    int4 *v = reinterpret_cast<int4 *>(out + base);
    int4 a = v[0], b = v[1];
    a.x += add; a.y += add; a.z += add; a.w += add;
    b.x += add; b.y += add; b.z += add; b.w += add;
    v[0] = a;
    v[1] = b;
  } else {
    for (int i = 0; i < kScanItems; i++)
      if (base + i < n) out[base + i] += add;
  }
}

__global__ void k_bin_scatter(const int32_t *__restrict__ key,
                              const int32_t *__restrict__ rank,
                              const int32_t *__restrict__ start,
                              const int32_t *__restrict__ oid, int64_t n,
                              unsigned long long *__restrict__ permk) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int64_t dst = (int64_t)start[key[s]] + rank[s];
  permk[dst] = ((unsigned long long)(uint32_t)oid[s] << 32) |
               (unsigned long long)(uint32_t)s;
}

// Orders each bin by original index and gathers the per-particle arrays
// into that order in one pass.  Thread j takes slot j of the scattered
// (bin-grouped, not yet ordered) entries (original index << 32 | old
// position): its place in its bin is the number of the bin's entries below
// it (the original indexes are distinct), and it copies its particle from
// the old position to that place.  A bin of more than kMaxBinSort
// particles is left in scatter order and reported as an overflow (its rows
// could not fit anyway), which bounds the work of every thread; after a
// non-finite force nothing is ordered.
constexpr int64_t kMaxBinSort = 1024;
__global__ void k_sort_gather(const unsigned long long *__restrict__ permk,
                              const int32_t *__restrict__ start, int64_t n,
                              StaticArrays ss, StaticArrays ds, DynArrays sd,
                              DynArrays dd, const int32_t *__restrict__ key,
                              int multi, int rods, int32_t *__restrict__ err,
                              const double *__restrict__ chem_src,
                              double *__restrict__ chem_dst) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const unsigned long long e = permk[j];
  const int64_t s = (int64_t)(uint32_t)(e & 0xffffffffull);
  const int32_t b = key[s];
  const int64_t a = start[b], z = start[b + 1];
  int64_t t = j;
  if (z - a > kMaxBinSort) {
    if (j == a) atomicOr(err, ERRBIT_OVERFLOW);
  } else if (!(*err & ERRBIT_NAN)) {
    t = a;
    for (int64_t i = a; i < z; i++) t += permk[i] < e;
  }
  dd.pos[t] = sd.pos[s];
  ds.oid[t] = ss.oid[s];
  ds.type[t] = ss.type[s];
  ds.key[t] = b;
  if (multi) ds.col[t] = ss.col[s];
  if (rods) {
    ds.bscale[t] = ss.bscale[s];
    dd.rod[t] = sd.rod[s];
    dd.trig[t] = sd.trig[s];
    dd.topo[t] = sd.topo[s];
  }
  if (chem_src) {  // the chemistry state's line (chem::stride doubles)
    const double4 *a = reinterpret_cast<const double4 *>(chem_src + s * chem::stride);
    double4 *b = reinterpret_cast<double4 *>(chem_dst + t * chem::stride);
    const double4 x0 = a[0], x1 = a[1], x2 = a[2], x3 = a[3];
    b[0] = x0;
    b[1] = x1;
    b[2] = x2;
    b[3] = x3;
  }
}

// -------------------------------------------------------- neighbor rows ---

// Interaction reach of a particle: r_lm = reach_i + reach_j + W_A for equal
// cell types (K.H:23-40).
template <bool RODS>
__device__ __forceinline__ double reach_of(const double4 &p, int ty,
                                           const double4 *rod, int64_t u) {
  if (RODS && ty == rec::FUNGI) return 1.5 * rod[u].x;
  return p.w;
}

// Separation along axis A: the periodic image test of sep_axis when bit A of
// M is set, the plain difference otherwise (used only where no image can be
// involved on that axis, which gives the same bits).
template <int M, int A>
__device__ __forceinline__ double sep_m(double xi, double xj, const BoxGeom &g) {
  if (M & (1 << A)) return sep_axis(xi, xj, g.L[A], g.halfL[A], g.periodic[A]);
  return xi - xj;
}

// Scans one contiguous candidate range [ua, ub) of the sorted arrays and
// appends the accepted candidates to the row (out walks the column-major
// row, one stride of ncap per entry).  The candidate loop is branch-free
// apart from the append: loads past the range end repeat the last entry and
// are masked out with the self pair.  Bit a of M asks for the periodic image
// test along axis a; an axis left out is one along which no candidate of
// the wave can need an image.  was = W_A + skin.
template <bool RODS, int M>
__device__ __forceinline__ void scan_range(
    int ua, int ub, int t, const double4 &p, double reach_i,
    const BoxGeom &g, const double4 *__restrict__ pos,
    const int32_t *__restrict__ type, const double4 *__restrict__ rod,
    int mode, double crit, double was, int32_t *&out, int64_t ncap,
    int32_t maxn, int &cnt) {
  for (int u0 = ua; u0 < ub; u0++) {
    double4 o[1];
    o[0] = pos[u0];
    {
      const int j = 0, u = u0;
      const double d0 = sep_m<M, 0>(p.x, o[j].x, g);
      const double d1 = sep_m<M, 1>(p.y, o[j].y, g);
      const double d2 = sep_m<M, 2>(p.z, o[j].z, g);
      double r2 = d0 * d0 + d1 * d1 + d2 * d2;
      bool ok;
      if (mode == 1) {
        ok = r2 <= crit;  // BMXCheckPair (src/des/bmx_pc.H:76)
      } else {
        double rj = RODS ? reach_of<true>(o[j], type[min(u, ub - 1)], rod,
                                          min(u, ub - 1))
                         : o[j].w;
        double cut = reach_i + rj + was;
        ok = r2 <= cut * cut;
      }
      if (ok && u < ub && u != t) {
        if (cnt < maxn) {
          *out = u;
          out += ncap;
        }
        cnt++;
      }
    }
  }
}


// A particle's bin stencil as memory ranges of the sorted arrays, in the
// canonical order (oz, then oy, then ox).  Each (oz, oy) row of bins is one
// or two ranges: the x window [bx - 1, bx + 1] is contiguous in memory
// (bins are x-fastest) unless it wraps a periodic x face, where it splits
// into the bins before the face and the bins after it (piece 0 holds the
// bins of ox = -1 first).  Rows outside a non-periodic box are empty.
// Everything that does not change from range to range is formed once per
// particle, with the row offsets already multiplied out, so that moving to
// the next range costs a few full-rate selects and adds and the two loads
// of its bounds (no integer multiply, which runs at a quarter of the rate).
struct Stencil {
  // bin id of the first bin of z row cz (base + cz * nby * nbx) for
  // oz = -1, 0, 1, and cy * nbx for oy = -1, 0, 1; -1 outside the box
  // (named fields, not arrays: a select between array elements becomes an
  // indexed scratch access)
  int z0, z1, z2, y0, y1, y2;
  int xa0, xb0, xa1, xb1;  // x pieces [xa, xb) of every row
  int two;                 // 1 when a row has two pieces
};

__device__ __forceinline__ Stencil make_stencil(int bx, int by, int bz,
                                                const BoxGeom &g) {
  Stencil st;
  const int nbx = g.nbin[0], nby = g.nbin[1], nbz = g.nbin[2];
  auto zrow = [&](int cz) {
    if (g.periodic[2]) cz = cz < 0 ? cz + nbz : (cz >= nbz ? cz - nbz : cz);
    return (cz >= 0 && cz < nbz) ? g.bin_offset + cz * nby * nbx : -1;
  };
  auto ycol = [&](int cy) {
    if (g.periodic[1]) cy = cy < 0 ? cy + nby : (cy >= nby ? cy - nby : cy);
    return (cy >= 0 && cy < nby) ? cy * nbx : -1;
  };
  st.z0 = zrow(bz - 1);
  st.z1 = zrow(bz);
  st.z2 = zrow(bz + 1);
  st.y0 = ycol(by - 1);
  st.y1 = ycol(by);
  st.y2 = ycol(by + 1);
  const int lo = bx - 1, hi = bx + 2;  // [lo, hi) before any wrap
  st.xa1 = st.xb1 = 0;
  if (lo < 0) {
    if (g.periodic[0]) {
      st.xa0 = nbx - 1;
      st.xb0 = nbx;
      st.xb1 = hi < nbx ? hi : nbx;
    } else {
      st.xa0 = 0;
      st.xb0 = hi < nbx ? hi : nbx;
    }
  } else if (hi > nbx) {
    st.xa0 = lo;
    st.xb0 = nbx;
    if (g.periodic[0]) st.xb1 = hi - nbx;
  } else {
    st.xa0 = lo;
    st.xb0 = hi;
  }
  st.two = st.xb1 > st.xa1;
  return st;
}

// Memory range of range k of the stencil (k < 9 << two): row r = k >> two
// (oz = r / 3, oy = r % 3), piece k & two.  Fields are read into locals
// before they are selected, so that the selects are between values.
__device__ __forceinline__ int2 stencil_range(const Stencil st, int k,
                                              const int32_t *__restrict__ start) {
  const int z0 = st.z0, z1 = st.z1, z2 = st.z2;
  const int y0 = st.y0, y1 = st.y1, y2 = st.y2;
  const int r = k >> st.two, piece = k & st.two;
  // r / 3 as 11 r / 32 (exact for r < 9), with 24-bit multiplies, which
  // run at the full rate
  const int iz = (int)(__umul24((unsigned)r, 11u) >> 5);
  const int iy = r - (int)__umul24((unsigned)iz, 3u);
  const int zz = iz == 0 ? z0 : (iz == 1 ? z1 : z2);
  const int yy = iy == 0 ? y0 : (iy == 1 ? y1 : y2);
  if ((zz | yy) < 0) return make_int2(0, 0);
  const int xa0 = st.xa0, xa1 = st.xa1, xb0 = st.xb0, xb1 = st.xb1;
  const int row = zz + yy;
  const int xa = piece ? xa1 : xa0, xb = piece ? xb1 : xb0;
  return make_int2(start[row + xa], start[row + xb]);
}

// Where a sphere build puts a particle's row while it grows: the first
// kStage entries in LDS (entry k at stage[k * stride], one word per thread
// and entry, so the lanes of a wave write neighbouring banks), the rest in
// the column-major row (entry k at col[k * ncap]).  With cols, every entry
// also goes to the column-major row, for the passes that read rows that way
// (exchange, fusion).  The flat rows are copied from here at the end.
constexpr int kStage = 8;
struct RowSink {
  int32_t *stage;
  int stride;
  int32_t *col;
  int64_t ncap;
  int cols;
};
__device__ __forceinline__ void row_put(const RowSink &sk, int k, int32_t u) {
  if (k < kStage) {
    sk.stage[k * sk.stride] = u;
    if (sk.cols) sk.col[(int64_t)k * sk.ncap] = u;
  } else {
    sk.col[(int64_t)k * sk.ncap] = u;
  }
}
__device__ __forceinline__ int32_t row_get(const RowSink &sk, int k) {
  return k < kStage ? sk.stage[k * sk.stride] : sk.col[(int64_t)k * sk.ncap];
}

// Scans a particle's stencil as one sequence of candidates and appends the
// accepted ones to its row (through the row sink).  Each lane moves on to its
// next range as soon as its current one ends, so a wave takes as many
// iterations as its longest stencil rather than the sum over ranges of the
// longest range; the next range's bounds are loaded one range ahead.  An
// iteration tests NC candidates of the range: two for one box (three or
// four spill at 64 VGPRs), one for several boxes (two spill there).
// Candidates visited are capped at scan_cap per particle (a range
// that would pass it is dropped and the row reported as overflowing).  Bit
// a of M asks for the periodic image
// test along axis a; an axis left out is one along which no candidate of
// the wave can need an image.  CRIT: the stencil policy's test of the
// squared distance against crit; otherwise the Verlet test against the
// reaches plus was = W_A + skin (a sphere's reach is its radius).
template <int M, bool CRIT, int NC>
__device__ __forceinline__ void scan_stencil(
    const Stencil sc, const BoxGeom &g, const int32_t *__restrict__ start,
    int64_t scan_cap, int t, const double4 &p,
    const double4 *__restrict__ pos, double crit, double was,
    const RowSink &sk, int32_t maxn, int &cnt) {
  const int nq = 9 << sc.two;
  // 32-bit count of the candidates taken; once a range would pass the cap,
  // it and every later range are dropped (scanned sticks above the cap)
  const int cap = scan_cap < (1 << 30) ? (int)scan_cap : (1 << 30);
  int scanned = 0;
  int q = 0;
  const int2 first = stencil_range(sc, 0, start);
  int2 next = stencil_range(sc, 1, start);
  int u = first.x, e = first.y;
  if (e - u > cap) {
    cnt = maxn + 1;
    e = u;
    scanned = cap + 1;
  } else {
    scanned = e - u;
  }
  for (;;) {
    while (u >= e && q + 1 < nq) {
      q++;
      u = next.x;
      e = next.y;
      const int len = e - u;
      if (len > cap - scanned) {
        cnt = maxn + 1;  // reported as overflow
        e = u;
        scanned = cap + 1;
      } else {
        scanned += len;
      }
      if (q + 1 < nq) next = stencil_range(sc, q + 1, start);
    }
    if (u >= e) break;
    // NC = 2: two candidates of the range per iteration (the second only
    // when the range holds it), appended in order
    const bool two = NC == 2 && u + 1 < e;
    const int u1 = two ? u + 1 : u;
    const double4 o = pos[u];
    bool ok, ok1 = false;
    {
      const double d0 = sep_m<M, 0>(p.x, o.x, g);
      const double d1 = sep_m<M, 1>(p.y, o.y, g);
      const double d2 = sep_m<M, 2>(p.z, o.z, g);
      const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
      if (CRIT) {
        ok = r2 <= crit;  // BMXCheckPair (src/des/bmx_pc.H:76)
      } else {
        const double cut = p.w + o.w + was;
        ok = r2 <= cut * cut;
      }
    }
    if (NC == 2) {
      const double4 o1 = pos[u1];
      const double d0 = sep_m<M, 0>(p.x, o1.x, g);
      const double d1 = sep_m<M, 1>(p.y, o1.y, g);
      const double d2 = sep_m<M, 2>(p.z, o1.z, g);
      const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
      if (CRIT) {
        ok1 = r2 <= crit;
      } else {
        const double cut = p.w + o1.w + was;
        ok1 = r2 <= cut * cut;
      }
    }
    if (ok && u != t) {
      if (cnt < maxn) row_put(sk, cnt, u);
      cnt++;
    }
    if (two && ok1 && u1 != t) {
      if (cnt < maxn) row_put(sk, cnt, u1);
      cnt++;
    }
    u += two ? 2 : 1;
  }
}

// Bits 28..30 of a row count mark the axes (x, y, z) along which a neighbor
// of the particle can need a periodic image before the next rebuild: the
// axis is periodic and the particle's bin is within two bins of a face (or
// the box has fewer than six bins along it).  Force kernels test images
// only along the axes marked somewhere in the wave.
constexpr int kImgShift = 28;
constexpr int32_t kRowMask = (1 << kImgShift) - 1;
__device__ __forceinline__ int row_len(int32_t raw) { return raw & kRowMask; }
__device__ __forceinline__ int row_img(int32_t raw) { return (raw >> kImgShift) & 7; }

// OR of a 3-bit axis mask over the wave.
__device__ __forceinline__ int wave_or3(int m) {
  return (__any(m & 1) ? 1 : 0) | (__any(m & 2) ? 2 : 0) | (__any(m & 4) ? 4 : 0);
}


template <bool MULTI, bool RODS>
__global__ __launch_bounds__(512, 8) void k_build_nbr(
    const double4 *__restrict__ pos, const int32_t *__restrict__ col,
    const int32_t *__restrict__ type, const double4 *__restrict__ rod,
    const BoxGeom *__restrict__ boxes, BoxGeom g0,
    const int32_t *__restrict__ start, const int32_t *__restrict__ key,
    int64_t n, int mode, double crit, double wa, double skin,
    int32_t *__restrict__ nbr, int32_t *__restrict__ ncnt, int64_t ncap,
    int32_t maxn, int32_t *__restrict__ err, int32_t *__restrict__ bstats,
    FlatRows flat, int cols, int remap) {
  const int64_t t = tile_of(remap) * blockDim.x + threadIdx.x;
  const int ti = (int)t;  // n < 2^31
  const double was = wa + skin;
  int32_t *out = nbr + t;
  // sphere builds stage the first entries of each row in LDS (kStage words
  // per thread, launched with that much dynamic LDS)
  extern __shared__ int32_t srow[];
  const RowSink sk{srow + threadIdx.x, (int)blockDim.x, nbr + t, ncap,
                   (cols || !flat.cidx) ? 1 : 0};
  int cnt = 0;
  // Candidates visited per particle are capped (a pile of particles in a few
  // bins would otherwise make this quadratic); hitting the cap is reported
  // as an overflow.  After a non-finite force no rows are built.
  const int64_t scan_cap = 64 * (int64_t)maxn + 4096;
  const bool active = t < n && !(*err & ERRBIT_NAN);
  // The row build over box g.  Several boxes: a wave whose particles all
  // lie in one box (all but the waves at colony boundaries) reads that box
  // once into scalar registers; the others read each particle's box.
  auto build = [&](const BoxGeom &g) {
    const int nbx = g.nbin[0], nby = g.nbin[1], nbz = g.nbin[2];
    int bx = 0, by = 0, bz = 0;
    double4 p = make_double4(0.0, 0.0, 0.0, 0.0);
    double reach_i = 0.0;
    // axes along which a candidate of this list build (scan) or a neighbor
    // before the next rebuild (force) can need a periodic image
    int scan_img = 0, force_img = 0;
    if (active) {
      p = pos[t];
      reach_i = RODS ? reach_of<true>(p, type[t], rod, t) : p.w;
      int lb = key[t] - g.bin_offset;
      bx = lb % nbx;
      int rr = lb / nbx;
      by = rr % nby;
      bz = rr / nby;
      const int b3[3] = {bx, by, bz};
      for (int a = 0; a < 3; a++) {
        if (!g.periodic[a]) continue;
        int nb = g.nbin[a];
        if (!(nb >= 4 && b3[a] >= 1 && b3[a] + 1 < nb)) scan_img |= 1 << a;
        if (!(nb >= 6 && b3[a] >= 2 && b3[a] + 2 < nb)) force_img |= 1 << a;
      }
    }
    const int wave_img = wave_or3(scan_img);
    if (active) {
      if (RODS) {
        // Segment builds (sparse mesh-cell bins): range by range, all lanes
        // together, which measured faster there than the flattened scan.
        int64_t scanned = 0;
        for (int oz = -1; oz <= 1; oz++) {
          int cz = bz + oz;
          if (g.periodic[2]) {
            cz = (cz + nbz) % nbz;
          } else if (cz < 0 || cz >= nbz) {
            continue;
          }
          for (int oy = -1; oy <= 1; oy++) {
            int cy = by + oy;
            if (g.periodic[1]) {
              cy = (cy + nby) % nby;
            } else if (cy < 0 || cy >= nby) {
              continue;
            }
            int row = g.bin_offset + (cz * nby + cy) * nbx;
            // up to three x-ranges; one contiguous range when no x wrap
            int ranges[3][2];
            int nr = 0;
            if (bx >= 1 && bx + 1 < nbx) {
              ranges[0][0] = start[row + bx - 1];
              ranges[0][1] = start[row + bx + 2];
              nr = 1;
            } else {
              for (int ox = -1; ox <= 1; ox++) {
                int cx = bx + ox;
                if (g.periodic[0]) {
                  cx = (cx + nbx) % nbx;
                } else if (cx < 0 || cx >= nbx) {
                  continue;
                }
                ranges[nr][0] = start[row + cx];
                ranges[nr][1] = start[row + cx + 1];
                nr++;
              }
            }
            for (int q = 0; q < nr; q++) {
              int64_t len = ranges[q][1] - ranges[q][0];
              if (scanned + len > scan_cap) {
                cnt = maxn + 1;  // reported as overflow below
                ranges[q][1] = ranges[q][0];
              }
              scanned += len;
              if (wave_img == 0)
                scan_range<RODS, 0>(ranges[q][0], ranges[q][1], ti, p, reach_i, g,
                                    pos, type, rod, mode, crit, was, out, ncap,
                                    maxn, cnt);
              else if (wave_img == 1)
                scan_range<RODS, 1>(ranges[q][0], ranges[q][1], ti, p, reach_i, g,
                                    pos, type, rod, mode, crit, was, out, ncap,
                                    maxn, cnt);
              else
                scan_range<RODS, 7>(ranges[q][0], ranges[q][1], ti, p, reach_i, g,
                                    pos, type, rod, mode, crit, was, out, ncap,
                                    maxn, cnt);
            }
          }
        }
      } else {
        const Stencil sc = make_stencil(bx, by, bz, g);
        auto run = [&](auto crit_tag) {
          constexpr bool C = decltype(crit_tag)::value;
          constexpr int NC = MULTI ? 1 : 2;
          if (wave_img == 0)
            scan_stencil<0, C, NC>(sc, g, start, scan_cap, ti, p, pos, crit, was, sk,
                                   maxn, cnt);
          else if (wave_img == 1)
            scan_stencil<1, C, NC>(sc, g, start, scan_cap, ti, p, pos, crit, was, sk,
                                   maxn, cnt);
          else
            scan_stencil<7, C, NC>(sc, g, start, scan_cap, ti, p, pos, crit, was, sk,
                                   maxn, cnt);
        };
        if (mode == 1)
          run(std::true_type{});
        else
          run(std::false_type{});
      }
      ncnt[t] = (cnt < maxn ? cnt : maxn) | (force_img << kImgShift);
      if (cnt > maxn) atomicOr(err, ERRBIT_OVERFLOW);
    } else if (t < n) {
      ncnt[t] = 0;
    }
  };
  if (MULTI) {
    const int32_t cl = col[t < n ? t : n - 1];
    const int cw = __builtin_amdgcn_readfirstlane(cl);
    if (__all(t >= n || cl == cw)) {
      const BoxGeom gw = boxes[cw];
      build(gw);
    } else {
      build(boxes[active ? col[t] : 0]);
    }
  } else {
    build(g0);
  }
  if (flat.cidx) {
    // Flat row layout: the tile's rows back to back from the start of the
    // tile's own region (tile * maxn entries per tile).
    const int tile = flat.tile, place = threadIdx.x & (tile - 1);
    const int kept = cnt < maxn ? cnt : maxn;
    const int incl = wave_incl_scan(kept, tile);
    const int tot = __shfl(incl, tile - 1, tile);
    if (place == 0 && t < n) flat.tot[t / tile] = tot;
    const int64_t b = (t / tile) * (int64_t)tile * maxn + (incl - kept);
    for (int k = 0; k < kept; k++) {
      flat.cidx[b + k] = RODS ? nbr[t + (int64_t)k * ncap] : row_get(sk, k);
      flat.cown[b + k] = (uint8_t)place;
    }
  }
  __shared__ int32_t shi[16];
  int m = block_max(cnt, shi);
  __syncthreads();
  int sm = block_sum(cnt, shi);
  if (threadIdx.x == 0) {
    bstats[2 * blockIdx.x] = m;
    bstats[2 * blockIdx.x + 1] = sm;
  }
}



// Reduces the per-block row statistics of the last build.
__global__ __launch_bounds__(1024) void k_reduce_rowstats(
    const int32_t *__restrict__ bstats, int32_t nblk,
    int32_t *__restrict__ stats, unsigned long long *__restrict__ total) {
  __shared__ int32_t shi[16];
  __shared__ unsigned long long shu[16];
  int m = 0;
  unsigned long long sm = 0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
    int v = bstats[2 * b];
    m = v > m ? v : m;
    sm += (unsigned long long)bstats[2 * b + 1];
  }
  m = block_max(m, shi);
  __syncthreads();
  sm = block_sum(sm, shu);
  if (threadIdx.x == 0) {
    stats[0] = m;
    *total = sm;
  }
}

// ------------------------------------------------------- force + update ---

// Largest step displacement dt*|vel| of the block, one store per block.
__device__ __forceinline__ void block_vmax(double v2, double dt, double *vpart) {
  __shared__ double shd[16];
  double m = block_max(v2, shd);
  if (threadIdx.x == 0) vpart[blockIdx.x] = dt * sqrt(m);
}

// Separation of two spheres along the axes, with the image test of sep_m.
template <int M>
__device__ __forceinline__ void sphere_sep(const double4 &p, const double4 &o,
                                           const BoxGeom &g, double &d0,
                                           double &d1, double &d2) {
  d0 = sep_m<M, 0>(p.x, o.x, g);
  d1 = sep_m<M, 1>(p.y, o.y, g);
  d2 = sep_m<M, 2>(p.z, o.z, g);
}

// ------------------------------------------------- flat sphere force ---
//
// The sphere step (force lambda PCI:170-279 with the sphere law K.H:243-276
// and the wall K.H:1009-1030, then the update PCI:311-449) over the flat row
// layout.  Lane l of a wave takes list entries l and l + 64 of each window
// of 128: it loads the neighbor index and its owner lane, gathers the
// neighbor, reads the owner's position from LDS and applies the interaction
// test (K.H:23-40 with PCI:212).  Interacting pairs are packed into LDS
// slots by ballot, entry 0 of every lane first, so slot order is entry
// order; the cubic law then runs once per 64 packed pairs and leaves each
// term in its slot.  An owner's interacting entries occupy a contiguous run
// of slots, which it adds in row order: every particle sums its pair terms
// in its row's order, as a loop over the row would.
constexpr int kFlatWin = 2 * kWave;
constexpr size_t kFlatWaveBytes =
    kWave * sizeof(double4) + 4 * kFlatWin * sizeof(double);

// Largest step displacement dt*|vel| of the wave, one store per wave at its
// wave index (the partial arrays hold one slot per 64 particles), with no
// block barrier: waves that finish early do not wait for the block's
// slowest.  (One atomic maximum per wave into a few shared words instead
// measured 2.4 times slower on the 1M-sphere step: the atomics serialize.)
__device__ __forceinline__ void wave_vmax(double v2, double dt, double *vpart,
                                          int64_t t, int lane, int64_t n) {
  const double m = wave_max_f64(v2);
  if (lane == 0 && t < n) vpart[t / kWave] = dt * sqrt(m);
}



template <bool MULTI, int M>
__device__ __forceinline__ void flat_rows(
    int T, int P, int c, int64_t B, int lane, int f0u, int f1u, int f0o, int f1o,
    const int32_t *__restrict__ cidx, const uint8_t *__restrict__ cown,
    const double4 *__restrict__ pos_in, const double4 *sp, const int32_t *scol,
    double *q0s, double *q1s, double *q2s, double *qrs,
    const BoxGeom *__restrict__ boxes, const BoxGeom &g0, const ForceParams &fp,
    double &a0, double &a1, double &a2) {
  const unsigned long long below = (1ull << lane) - 1ull;
  // f0u, f1u, f0o, f1o: with one box, the first window's entries, loaded
  // by the caller (entries past the count are masked below)
  for (int w0 = 0; w0 < T; w0 += kFlatWin) {
    const int e0 = w0 + lane, e1 = e0 + kWave;
    const bool v0 = e0 < T, v1 = e1 < T;
    int u0 = 0, u1 = 0, o0 = 0, o1 = 0;
    if (!MULTI && w0 == 0) {
      u0 = v0 ? f0u : 0;
      u1 = v1 ? f1u : 0;
      o0 = v0 ? f0o : 0;
      o1 = v1 ? f1o : 0;
    } else {
      if (v0) {
        u0 = cidx[B + e0];
        o0 = cown[B + e0];
      }
      if (v1) {
        u1 = cidx[B + e1];
        o1 = cown[B + e1];
      }
    }
    const double4 q0 = pos_in[u0], q1 = pos_in[u1];
    const double4 p0 = sp[o0], p1 = sp[o1];
    const BoxGeom &ga = MULTI ? boxes[scol[o0]] : g0;
    const BoxGeom &gb = MULTI ? boxes[scol[o1]] : g0;
    double x0, y0, z0, x1, y1, z1;
    sphere_sep<M>(p0, q0, ga, x0, y0, z0);
    sphere_sep<M>(p1, q1, gb, x1, y1, z1);
    const double r20 = x0 * x0 + y0 * y0 + z0 * z0;
    const double r21 = x1 * x1 + y1 * y1 + z1 * z1;
    const double lim0 = p0.w + q0.w + fp.WA - 1.0e-15;  // r_lm - small_number
    const double lim1 = p1.w + q1.w + fp.WA - 1.0e-15;
    const bool i0 = v0 && r20 <= lim0 * lim0;
    const bool i1 = v1 && r21 <= lim1 * lim1;
    const unsigned long long b0 = __ballot(i0), b1 = __ballot(i1);
    const int n0 = __popcll(b0), ni = n0 + __popcll(b1);
    if (i0) {
      const int k = __popcll(b0 & below);
      q0s[k] = x0;
      q1s[k] = y0;
      q2s[k] = z0;
      qrs[k] = p0.w + q0.w;
    }
    if (i1) {
      const int k = n0 + __popcll(b1 & below);
      q0s[k] = x1;
      q1s[k] = y1;
      q2s[k] = z1;
      qrs[k] = p1.w + q1.w;
    }
    wave_lds_sync();
    for (int k = lane; k < ni; k += kWave) {  // the cubic law, packed
      const double d0 = q0s[k], d1 = q1s[k], d2 = q2s[k];
      const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
      double fx = 0.0, fy = 0.0, fz = 0.0;
      sphere_pair(d0, d1, d2, r2, qrs[k], fp, fx, fy, fz);
      q0s[k] = fx;
      q1s[k] = fy;
      q2s[k] = fz;
    }
    wave_lds_sync();
    const int lo = P > w0 ? P : w0;
    const int hi = P + c < w0 + kFlatWin ? P + c : w0 + kFlatWin;
    if (lo < hi) {
      const int k1 = packed_before(hi - w0, b0, b1, n0);
      for (int k = packed_before(lo - w0, b0, b1, n0); k < k1; k++) {
        a0 += q0s[k];
        a1 += q1s[k];
        a2 += q2s[k];
      }
    }
    wave_lds_sync();
  }
}

template <bool MULTI>
__global__ __launch_bounds__(512, 6) void k_force_sphere_flat(
    const double4 *__restrict__ pos_in, double4 *__restrict__ pos_out,
    const int32_t *__restrict__ cidx, const uint8_t *__restrict__ cown,
    const int32_t *__restrict__ wtot, const int32_t *__restrict__ ncnt,
    int64_t n, int32_t maxn, const int32_t *__restrict__ col,
    const BoxGeom *__restrict__ boxes, BoxGeom g0, ForceParams fp,
    double *__restrict__ fc, const int32_t *__restrict__ fc_oid,
    double *__restrict__ vmax, int32_t *__restrict__ err, int remap) {
  extern __shared__ __align__(16) unsigned char flat_lds[];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = threadIdx.x / kWave, nwv = blockDim.x / kWave;
  double4 *sp = reinterpret_cast<double4 *>(flat_lds) + wv * kWave;
  double *q0s = reinterpret_cast<double *>(flat_lds + nwv * kWave * sizeof(double4)) +
                wv * 4 * kFlatWin;
  double *q1s = q0s + kFlatWin, *q2s = q1s + kFlatWin, *qrs = q2s + kFlatWin;
  int32_t *scol = reinterpret_cast<int32_t *>(flat_lds + nwv * kFlatWaveBytes) +
                  wv * kWave;
  const int64_t t = tile_of(remap) * blockDim.x + threadIdx.x;
  const bool active = t < n;
  // The wave's entries start at its region (64 * maxn entries per wave).
  // Every load that does not depend on another is issued here, before
  // anything waits, from indexes valid for every lane (the last particle
  // for lanes past n, tile 0 for a wave past n): the row length, the wave's
  // entry count, the particle, its box and the first window of entries (a
  // tile's region, and the padding after the last one, is always
  // readable).  The one-box path takes the first window from these; the
  // per-entry-box path, which would spill with them live, reloads it.
  const bool live = t - lane < n;  // the wave holds at least one particle
  const int64_t B = live ? (t / kWave) * (int64_t)kWave * maxn : 0;
  const int64_t tc = active ? t : n - 1;
  const int32_t raw_l = ncnt[tc];
  const int32_t T_l = wtot[live ? t / kWave : 0];
  const double4 p_l = pos_in[tc];
  const int32_t col_l = MULTI ? col[tc] : 0;
  const int f0u = cidx[B + lane], f1u = cidx[B + kWave + lane];
  const int f0o = cown[B + lane], f1o = cown[B + kWave + lane];
  const int T = live ? T_l : 0;
  const int32_t raw = active ? raw_l : 0;
  const int wave_img = wave_or3(row_img(raw));
  const int c = row_len(raw);
  const int P = wave_incl_scan64(c) - c;
  const double4 p = active ? p_l : make_double4(0.0, 0.0, 0.0, 0.0);
  sp[lane] = p;
  if (MULTI) scol[lane] = active ? col_l : 0;
  // Several boxes: a wave whose particles all lie in one box (all but the
  // waves at colony boundaries) takes the one-box path with that box's
  // geometry, read once into scalar registers; the others read each
  // owner's box per entry.
  bool one_box = true;
  BoxGeom gw = g0;
  if (MULTI) {
    const int cw = __builtin_amdgcn_readfirstlane(col_l);  // lane 0 is active
    one_box = __all(!active || col_l == cw);
    if (one_box) gw = boxes[cw];
  }
  wave_lds_sync();
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  if (one_box) {
    if (wave_img == 0)
      flat_rows<false, 0>(T, P, c, B, lane, f0u, f1u, f0o, f1o, cidx, cown,
                          pos_in, sp, scol, q0s, q1s, q2s, qrs, boxes, gw, fp,
                          a0, a1, a2);
    else if (wave_img == 1)
      flat_rows<false, 1>(T, P, c, B, lane, f0u, f1u, f0o, f1o, cidx, cown,
                          pos_in, sp, scol, q0s, q1s, q2s, qrs, boxes, gw, fp,
                          a0, a1, a2);
    else
      flat_rows<false, 7>(T, P, c, B, lane, f0u, f1u, f0o, f1o, cidx, cown,
                          pos_in, sp, scol, q0s, q1s, q2s, qrs, boxes, gw, fp,
                          a0, a1, a2);
  } else if (MULTI) {
    if (wave_img == 0)
      flat_rows<true, 0>(T, P, c, B, lane, f0u, f1u, f0o, f1o, cidx, cown,
                         pos_in, sp, scol, q0s, q1s, q2s, qrs, boxes, g0, fp,
                         a0, a1, a2);
    else if (wave_img == 1)
      flat_rows<true, 1>(T, P, c, B, lane, f0u, f1u, f0o, f1o, cidx, cown,
                         pos_in, sp, scol, q0s, q1s, q2s, qrs, boxes, g0, fp,
                         a0, a1, a2);
    else
      flat_rows<true, 7>(T, P, c, B, lane, f0u, f1u, f0o, f1o, cidx, cown,
                         pos_in, sp, scol, q0s, q1s, q2s, qrs, boxes, g0, fp,
                         a0, a1, a2);
  }
  double v2 = 0.0;
  if (active) {
    const BoxGeom &g = MULTI ? boxes[col_l] : g0;
    // growth surface (K.H:1009-1030); vcom x/y are zero
    a0 += 0.0;
    a1 += 0.0;
    a2 += wall_law(p.z, p.w, fp);
    double vx = fp.mob * a0, vy = fp.mob * a1, vz = fp.mob * a2;
    double x = p.x + fp.dt * vx, y = p.y + fp.dt * vy, z = p.z + fp.dt * vz;
    x = wrap_axis(x, g.plo[0], g.phi[0], g.L[0], g.periodic[0]);
    y = wrap_axis(y, g.plo[1], g.phi[1], g.L[1], g.periodic[1]);
    z = wrap_axis(z, g.plo[2], g.phi[2], g.L[2], g.periodic[2]);
    pos_out[t] = make_double4(x, y, z, p.w);
    if (fc) {  // original order, so a later rebuild does not reorder it
      store_fc(fc, fc_oid[t], a0, a1, a2, 0.0, 0.0, 0.0);
    }
    if (!(isfinite(a0) && isfinite(a1) && isfinite(a2))) atomicOr(err, ERRBIT_NAN);
    v2 = vx * vx + vy * vy + vz * vz;
  }
  if (vmax) wave_vmax(v2, fp.dt, vmax, t, lane, n);
}

// Serial pass over one particle's column-major row (the reference's
// per-particle neighbor loop): sphere pairs for yeast, segment pairs for
// hyphae, summed into acc in row order; ti is the particle's evolving
// record (rod_pair_i may repair a bond or promote a second segment).
template <bool MULTI>
__device__ void general_rows(int64_t t, const BoxGeom &g, const DynArrays &in,
                             const StaticArrays &st,
                             const int32_t *__restrict__ nbr,
                             const int32_t *__restrict__ ncnt, int64_t ncap,
                             const ForceParams &fp, double acc[6], Topo &ti) {
  const int ty = st.type[t];
  const double4 p = in.pos[t];
  const double4 rod_i = in.rod[t];
  const int c = row_len(ncnt[t]);
  const int32_t *row = nbr + t;
  if (ty == rec::YEAST) {
    for (int k = 0; k < c; k++) {
      int u = row[(int64_t)k * ncap];
      if (st.type[u] != rec::YEAST) continue;  // r_lm = 0 (K.H:39)
      double4 o = in.pos[u];
      double d0 = sep_axis(p.x, o.x, g.L[0], g.halfL[0], g.periodic[0]);
      double d1 = sep_axis(p.y, o.y, g.L[1], g.halfL[1], g.periodic[1]);
      double d2 = sep_axis(p.z, o.z, g.L[2], g.halfL[2], g.periodic[2]);
      double r2 = d0 * d0 + d1 * d1 + d2 * d2;
      double rS = p.w + o.w;
      double lim = rS + fp.WA - 1.0e-15;
      if (r2 <= lim * lim)
        sphere_pair(d0, d1, d2, r2, rS, fp, acc[0], acc[1], acc[2]);
    }
  } else if (ty == rec::FUNGI) {
    RodSelf S;
    rod_self_init(S, rod_i, in.trig[t], p.w, st.bscale[t], fp.visc);
    for (int k = 0; k < c; k++) {
      int u = row[(int64_t)k * ncap];
      if (st.type[u] != rec::FUNGI) continue;
      double4 o = in.pos[u];
      double4 rod_j = in.rod[u];
      double d[3];
      d[0] = sep_axis(p.x, o.x, g.L[0], g.halfL[0], g.periodic[0]);
      d[1] = sep_axis(p.y, o.y, g.L[1], g.halfL[1], g.periodic[1]);
      d[2] = sep_axis(p.z, o.z, g.L[2], g.halfL[2], g.periodic[2]);
      double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      double rlm = 1.5 * (rod_i.x + rod_j.x) + fp.WA;  // K.H:33-37
      double lim = rlm - 1.0e-15;
      if (r2 <= lim * lim) {
        Topo tj = in.topo[u];
        rod_pair_i(d, S, ti, tj, rod_j.x, o.w, st.bscale[u], in.trig[u], fp,
                   acc);
      }
    }
  }
}

// Surface term, overdamped update and outputs of one particle after its
// pair terms (K.H:1009-1282, PCI:311-449); returns |vel|^2.
__device__ double general_finish(int64_t t, const BoxGeom &g,
                                 const DynArrays &in, const DynArrays &out,
                                 const StaticArrays &st, const ForceParams &fp,
                                 double acc[6], Topo &ti,
                                 double *__restrict__ fc,
                                 int32_t *__restrict__ err) {
  const int ty = st.type[t];
  const double4 p = in.pos[t];
  const double4 rod_i = in.rod[t];
  const double4 tr_i = in.trig[t];
  if (ty == rec::YEAST) {
    acc[0] += 0.0;
    acc[1] += 0.0;
    acc[2] += wall_law(p.z, p.w, fp);
    acc[3] += 0.0;
    acc[4] += 0.0;
    acc[5] += 0.0;
  } else if (ty == rec::FUNGI) {
    RodSelf S;
    rod_self_init(S, rod_i, tr_i, p.w, st.bscale[t], fp.visc);
    double vs[3], ws[3];
    rod_surface(p, S, fp, vs, ws);
    acc[0] += vs[0];
    acc[1] += vs[1];
    acc[2] += vs[2];
    acc[3] += ws[0];
    acc[4] += ws[1];
    acc[5] += ws[2];
    ti.fix_site = 0;  // K.H:1060
  }
  double vx = fp.mob * acc[0], vy = fp.mob * acc[1], vz = fp.mob * acc[2];
  double x = p.x + fp.dt * vx, y = p.y + fp.dt * vy, z = p.z + fp.dt * vz;
  x = wrap_axis(x, g.plo[0], g.phi[0], g.L[0], g.periodic[0]);
  y = wrap_axis(y, g.plo[1], g.phi[1], g.L[1], g.periodic[1]);
  z = wrap_axis(z, g.plo[2], g.phi[2], g.L[2], g.periodic[2]);
  out.pos[t] = make_double4(x, y, z, p.w);
  if (ty == rec::FUNGI) {
    double th = rod_i.y, ph = rod_i.z;
    double w[3] = {acc[3], acc[4], acc[5]};
    rod_rotate(th, ph, tr_i, rod_i.x, w, fp.dt);
    out.rod[t] = make_double4(rod_i.x, th, ph, rod_i.w);
    out.trig[t] = make_double4(bmx_cos(th), bmx_sin(th), bmx_cos(ph), bmx_sin(ph));
  } else {
    out.rod[t] = rod_i;
    out.trig[t] = tr_i;
  }
  out.topo[t] = ti;
  if (fc) {  // original order
    store_fc(fc, st.oid[t], acc[0], acc[1], acc[2], acc[3], acc[4], acc[5]);
  }
  bool fin = true;
#pragma unroll
  for (int k = 0; k < 6; k++) fin = fin && isfinite(acc[k]);
  if (!fin) atomicOr(err, ERRBIT_NAN);
  return vx * vx + vy * vy + vz * vz;
}

// ------------------------------------------------ flat general force ---
//
// Segments (and mixed colonies) over the flat row layout with tiles of
// kGTile particles, one tile per wave.  Per window of 128 entries a light
// pass (neighbor position, length and type) applies the interaction test;
// the interacting entries are packed by ballot, in entry order, and only
// they load the heavy data and run the pair law (rod_pair_i, sphere_pair),
// each term into its LDS slot.  The tile's particles (lanes 0..kGTile-1)
// then add their runs of terms in row order.  The pair law is evaluated
// with the particle's record as it stands before its row; that is what the
// row loop sees unless an earlier pair of the row repaired a bond.  A tile
// where some pair repairs a bond is marked and recomputed row by row by
// k_force_general_redo (general_rows), so every result is the row loop's;
// keeping the row loop out of this kernel keeps its register budget low.  Promotion of a second segment and
// the dltd reset do not depend on the order of the pairs.
constexpr int kGTile = 32;
constexpr int kGWin = 2 * kWave;
constexpr int kGPass = kWave;  // packed pairs per pass
// LDS per wave: the terms of one pass (6 x 64), the tile's mobility
// matrices (18 x 32, formed once per segment) and the window's packed
// pairs (neighbor, owner, promotion flag: 3 x 128)
constexpr size_t kGenWaveBytes = 6 * kGPass * sizeof(double) +
                                 18 * kGTile * sizeof(double) +
                                 3 * kGWin * sizeof(int32_t);

template <bool MULTI>
__device__ __forceinline__ bool general_test(
    int64_t ti_, int64_t u, bool valid, const DynArrays &in,
    const StaticArrays &st, const BoxGeom *__restrict__ boxes,
    const BoxGeom &g0, const ForceParams &fp) {
  if (!valid) return false;
  const int ty = st.type[ti_], tyj = st.type[u];
  if (ty != tyj || (ty != rec::YEAST && ty != rec::FUNGI)) return false;
  const BoxGeom &g = MULTI ? boxes[st.col[ti_]] : g0;
  const double4 p = in.pos[ti_], o = in.pos[u];
  const double d0 = sep_axis(p.x, o.x, g.L[0], g.halfL[0], g.periodic[0]);
  const double d1 = sep_axis(p.y, o.y, g.L[1], g.halfL[1], g.periodic[1]);
  const double d2 = sep_axis(p.z, o.z, g.L[2], g.halfL[2], g.periodic[2]);
  const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
  double lim;
  if (ty == rec::YEAST) {
    lim = p.w + o.w + fp.WA - 1.0e-15;
  } else {
    const double rlm = 1.5 * (in.rod[ti_].x + in.rod[u].x) + fp.WA;  // K.H:33-37
    lim = rlm - 1.0e-15;
  }
  return r2 <= lim * lim;
}

template <bool MULTI>
__global__ __launch_bounds__(256, 4) void k_force_general_flat(
    DynArrays in, DynArrays out, StaticArrays st,
    const int32_t *__restrict__ cidx, const uint8_t *__restrict__ cown,
    const int32_t *__restrict__ wtot, int8_t *__restrict__ redo_tile,
    const int32_t *__restrict__ ncnt, int64_t n, int32_t maxn,
    const BoxGeom *__restrict__ boxes, BoxGeom g0,
    ForceParams fp, double *__restrict__ fc, double *__restrict__ vmax,
    int32_t *__restrict__ err, int remap) {
  extern __shared__ __align__(16) unsigned char gen_lds[];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = threadIdx.x / kWave, nwv = blockDim.x / kWave;
  double *res = reinterpret_cast<double *>(gen_lds + wv * kGenWaveBytes);
  double *mob = res + 6 * kGPass;  // mob[m * kGTile + owner]: dii (m < 9), dij
  int32_t *pu = reinterpret_cast<int32_t *>(mob + 18 * kGTile);
  int32_t *po = pu + kGWin;
  int32_t *pf = po + kGWin;
  const int64_t tile = tile_of(remap) * nwv + wv;
  const int64_t t0 = tile * kGTile;
  const int64_t t = t0 + lane;
  const bool active = lane < kGTile && t < n;
  int64_t B = 0;
  int T = 0;
  if (t0 < n) {
    B = tile * (int64_t)kGTile * maxn;
    T = wtot[tile];
  }
  const int c = active ? row_len(ncnt[t]) : 0;
  const int P = wave_incl_scan64(c) - c;
  if (active && st.type[t] == rec::FUNGI && c > 0) {
    // the two-site mobility of the segment, once (the pairs read it)
    RodSelf S;
    rod_self_geom(S, in.rod[t], in.trig[t], in.pos[t].w, st.bscale[t]);
    rod_self_mob(S, in.trig[t], fp.visc);
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) {
        mob[(3 * a + b) * kGTile + lane] = S.dii[a][b];
        mob[(9 + 3 * a + b) * kGTile + lane] = S.dij[a][b];
      }
  }
  wave_lds_sync();
  const unsigned long long below = (1ull << lane) - 1ull;
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  bool paired = false, promoted = false, repair = false;
  for (int w0 = 0; w0 < T; w0 += kGWin) {
    const int e0 = w0 + lane, e1 = e0 + kWave;
    const bool v0 = e0 < T, v1 = e1 < T;
    int u0 = 0, u1 = 0, o0 = 0, o1 = 0;
    if (v0) {
      u0 = cidx[B + e0];
      o0 = cown[B + e0];
    }
    if (v1) {
      u1 = cidx[B + e1];
      o1 = cown[B + e1];
    }
    const bool i0 = general_test<MULTI>(t0 + o0, u0, v0, in, st, boxes, g0, fp);
    const bool i1 = general_test<MULTI>(t0 + o1, u1, v1, in, st, boxes, g0, fp);
    const unsigned long long b0 = __ballot(i0), b1 = __ballot(i1);
    const int n0 = __popcll(b0), ni = n0 + __popcll(b1);
    if (i0) {
      const int k = __popcll(b0 & below);
      pu[k] = u0;
      po[k] = o0;
    }
    if (i1) {
      const int k = n0 + __popcll(b1 & below);
      pu[k] = u1;
      po[k] = o1;
    }
    wave_lds_sync();
    const int lo = P > w0 ? P : w0;
    const int hi = P + c < w0 + kGWin ? P + c : w0 + kGWin;
    const int klo = lo < hi ? packed_before(lo - w0, b0, b1, n0) : 0;
    const int khi = lo < hi ? packed_before(hi - w0, b0, b1, n0) : 0;
    for (int k0 = 0; k0 < ni; k0 += kGPass) {  // pair laws, packed
      const int k = k0 + lane;
      if (k < ni) {
        const int64_t u = pu[k], ti_ = t0 + po[k];
        const BoxGeom &g = MULTI ? boxes[st.col[ti_]] : g0;
        const double4 p = in.pos[ti_], o = in.pos[u];
        double d[3];
        d[0] = sep_axis(p.x, o.x, g.L[0], g.halfL[0], g.periodic[0]);
        d[1] = sep_axis(p.y, o.y, g.L[1], g.halfL[1], g.periodic[1]);
        d[2] = sep_axis(p.z, o.z, g.L[2], g.halfL[2], g.periodic[2]);
        double f[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        int flag = 0;
        if (st.type[ti_] == rec::YEAST) {
          const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
          sphere_pair(d[0], d[1], d[2], r2, p.w + o.w, fp, f[0], f[1], f[2]);
        } else {
          RodSelf S;
          rod_self_geom(S, in.rod[ti_], in.trig[ti_], p.w, st.bscale[ti_]);
          Topo tio = in.topo[ti_];
          const int pos0 = tio.position, nb0 = tio.n_bnds;
          double f1[3], f2[3], ri[3];
          {
            const Topo tj = in.topo[u];
            rod_pair_forces(d, S, tio, tj, in.rod[u].x, o.w, st.bscale[u],
                            in.trig[u], fp, f1, f2, ri);
          }
          if (tio.n_bnds != nb0) repair = true;
          if (tio.position != pos0) flag = 1;
          const int ow = po[k];
#pragma unroll
          for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) {
              S.dii[a][b] = mob[(3 * a + b) * kGTile + ow];
              S.dij[a][b] = mob[(9 + 3 * a + b) * kGTile + ow];
            }
          rod_pair_mobility(S, f1, f2, ri, f);
        }
#pragma unroll
        for (int q = 0; q < 6; q++) res[q * kGPass + lane] = f[q];
        pf[k] = flag;
      }
      wave_lds_sync();
      const int a0 = klo > k0 ? klo : k0;
      const int a1 = khi < k0 + kGPass ? khi : k0 + kGPass;
      for (int k2 = a0; k2 < a1; k2++) {
#pragma unroll
        for (int q = 0; q < 6; q++) acc[q] += res[q * kGPass + (k2 - k0)];
        promoted = promoted || pf[k2] != 0;
        paired = true;
      }
      wave_lds_sync();
    }
    wave_lds_sync();
  }
  // A tile with a bond repair is left to k_force_general_redo.
  const bool redo = __any(repair);
  if (lane == 0 && t0 < n) redo_tile[tile] = redo ? 1 : 0;
  double v2 = 0.0;
  if (active && !redo) {
    const BoxGeom &g = MULTI ? boxes[st.col[t]] : g0;
    Topo ti = in.topo[t];
    if (st.type[t] == rec::FUNGI && paired) {
      ti.dltd = 1;  // K.H:364-369
      if (promoted) ti.position = rec::INTERIOR;
    }
    v2 = general_finish(t, g, in, out, st, fp, acc, ti, fc, err);
  }
  if (vmax) block_vmax(v2, fp.dt, vmax);
}

// The row loop for the particles of the tiles that the flat pass marked
// (a pair of the tile repairs a bond).  Runs after the flat pass on the
// same stream; its displacement maxima join the step's partials through an
// atomic max (non-negative doubles order as their bit patterns).
template <bool MULTI>
__global__ __launch_bounds__(256) void k_force_general_redo(
    DynArrays in, DynArrays out, StaticArrays st,
    const int8_t *__restrict__ redo_tile, const int32_t *__restrict__ nbr,
    const int32_t *__restrict__ ncnt, int64_t ncap, int64_t n,
    const BoxGeom *__restrict__ boxes, BoxGeom g0, ForceParams fp,
    double *__restrict__ fc, double *__restrict__ vmax,
    int32_t *__restrict__ err) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || !redo_tile[t / kGTile]) return;
  const BoxGeom &g = MULTI ? boxes[st.col[t]] : g0;
  Topo ti = in.topo[t];
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  general_rows<MULTI>(t, g, in, st, nbr, ncnt, ncap, fp, acc, ti);
  const double v2 = general_finish(t, g, in, out, st, fp, acc, ti, fc, err);
  if (vmax)
    atomicMax(reinterpret_cast<unsigned long long *>(vmax), dbits(fp.dt * sqrt(v2)));
}

// InitBonds: per segment, scan the row in order and append a bond for the
// first site pair closer than rlim (K.H:1285-1444 with rlim = 1e-7, PCI:508).
template <bool MULTI>
__global__ void k_init_bonds(DynArrays dy, StaticArrays st,
                             const int32_t *__restrict__ nbr,
                             const int32_t *__restrict__ ncnt, int64_t ncap,
                             int64_t n, const BoxGeom *__restrict__ boxes,
                             BoxGeom g0, ForceParams fp,
                             int32_t *__restrict__ err) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || st.type[t] != rec::FUNGI) return;
  const BoxGeom &g = MULTI ? boxes[st.col[t]] : g0;
  const double rlim = 1.0e-7;
  double4 p = dy.pos[t];
  double4 rod_i = dy.rod[t];
  double4 tr = dy.trig[t];
  double Li = rod_i.x;
  double nxi = tr.y * tr.z, nyi = tr.y * tr.w, nzi = tr.x;
  Topo ti = dy.topo[t];
  int c = row_len(ncnt[t]);
  for (int k = 0; k < c; k++) {
    int u = nbr[(int64_t)k * ncap + t];
    if (st.type[u] != rec::FUNGI) continue;
    double4 o = dy.pos[u];
    double4 rod_j = dy.rod[u];
    double d[3];
    d[0] = sep_axis(p.x, o.x, g.L[0], g.halfL[0], g.periodic[0]);
    d[1] = sep_axis(p.y, o.y, g.L[1], g.halfL[1], g.periodic[1]);
    d[2] = sep_axis(p.z, o.z, g.L[2], g.halfL[2], g.periodic[2]);
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double lim = 1.5 * (Li + rod_j.x) + fp.WA - 1.0e-15;
    if (!(r2 <= lim * lim)) continue;
    double4 trj = dy.trig[u];
    double Lj = rod_j.x;
    double nxj = trj.y * trj.z, nyj = trj.y * trj.w, nzj = trj.x;
    double xi1 = d[0] - 0.5 * nxi * Li, yi1 = d[1] - 0.5 * nyi * Li,
           zi1 = d[2] - 0.5 * nzi * Li;
    double xi2 = d[0] + 0.5 * nxi * Li, yi2 = d[1] + 0.5 * nyi * Li,
           zi2 = d[2] + 0.5 * nzi * Li;
    double xj1 = -0.5 * nxj * Lj, yj1 = -0.5 * nyj * Lj, zj1 = -0.5 * nzj * Lj;
    double xj2 = 0.5 * nxj * Lj, yj2 = 0.5 * nyj * Lj, zj2 = 0.5 * nzj * Lj;
    // site pairs in the reference's order 1-1, 2-1, 1-2, 2-2
    for (int cse = 0; cse < 4; cse++) {
      bool i2 = (cse & 1) != 0, j2 = (cse & 2) != 0;
      double rx = (i2 ? xi2 : xi1) - (j2 ? xj2 : xj1);
      double ry = (i2 ? yi2 : yi1) - (j2 ? yj2 : yj1);
      double rz = (i2 ? zi2 : zi1) - (j2 ? zj2 : zj1);
      double r = sqrt(rx * rx + ry * ry + rz * rz);
      if (r < rlim) {
        int nb = ti.n_bnds;
        if (nb < 0 || nb > 3) {
          atomicOr(err, ERRBIT_BOND);
        } else {
          ti.seg_id[nb] = dy.topo[u].self_id;
          ti.seg_cpu[nb] = dy.topo[u].self_cpu;
          ti.site[nb] = i2 ? 2 : 1;
          ti.n_bnds = nb + 1;
        }
        break;
      }
    }
  }
  dy.topo[t].n_bnds = ti.n_bnds;
  for (int b = 0; b < 4; b++) {
    dy.topo[t].seg_id[b] = ti.seg_id[b];
    dy.topo[t].seg_cpu[b] = ti.seg_cpu[b];
    dy.topo[t].site[b] = ti.site[b];
  }
}

// Sum over the steps since the last rebuild of the largest displacement in
// any block bounds every particle's displacement; more than half the skin
// means a pair may have entered interaction range unlisted.  Stage 1: each
// block takes the per-slot max over a chunk of force blocks into slotmax
// (one atomic per slot per block, values >= 0 compare as integers).
// Block (x, y) takes slot y over a chunk of kSkinChunk force blocks.
constexpr int kSkinChunk = 4096;
__global__ __launch_bounds__(256) void k_skin_partial(
    double *__restrict__ vpart, int32_t stride, int clear,
    unsigned long long *__restrict__ slotmax) {
  __shared__ double shd[4];
  const int sl = blockIdx.y;
  const int64_t b0 = (int64_t)blockIdx.x * kSkinChunk;
  const int64_t b1 = b0 + kSkinChunk < stride ? b0 + kSkinChunk : stride;
  double m = 0.0;
  for (int64_t b = b0 + threadIdx.x; b < b1; b += blockDim.x) {
    double v = vpart[(int64_t)sl * stride + b];
    m = v > m ? v : m;
    if (clear) vpart[(int64_t)sl * stride + b] = 0.0;
  }
  m = block_max(m, shd);
  if (threadIdx.x == 0 && m > 0.0) atomicMax(&slotmax[sl], dbits(m));
}

// Stage 2: sum the slot maxima in slot order, compare, reset the slot
// maxima.
__global__ void k_skin_final(unsigned long long *__restrict__ slotmax,
                             int32_t nslots, double limit,
                             int32_t *__restrict__ err) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double total = 0.0;
  for (int sl = 0; sl < nslots; sl++) {
    total += __longlong_as_double((long long)slotmax[sl]);
    slotmax[sl] = 0ull;
  }
  if (total > limit) atomicOr(err, ERRBIT_SKIN);
}

// ------------------------------------------------------------- exports ---

// fc from its kept layout (original order, six per particle) to the
// reference's component-major fc[i + c * n]
__global__ void k_export_fc(const double *__restrict__ fc, int64_t n,
                            double *__restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < 6; c++) out[c * n + i] = fc[i * 6 + c];
}
__global__ void k_export_pos(const double4 *__restrict__ pos,
                             const int32_t *__restrict__ oid, int64_t n,
                             double *__restrict__ out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int64_t o = oid[t];
  double4 p = pos[t];
  out[3 * o] = p.x;
  out[3 * o + 1] = p.y;
  out[3 * o + 2] = p.z;
}
__global__ void k_export_orient(const double4 *__restrict__ rod,
                                const int32_t *__restrict__ oid, int64_t n,
                                double *__restrict__ out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int64_t o = oid[t];
  double4 r = rod[t];
  out[2 * o] = r.y;
  out[2 * o + 1] = r.z;
}
__global__ void k_export_i32(const int32_t *__restrict__ v,
                             const int32_t *__restrict__ oid, int64_t n,
                             int32_t *__restrict__ out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  out[oid[t]] = v[t];
}
__global__ void k_export_counts(const int32_t *__restrict__ ncnt,
                                const int32_t *__restrict__ oid, int64_t n,
                                int64_t *__restrict__ out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  out[oid[t]] = row_len(ncnt[t]);
}
// Rows from the flat layout: particle t's row starts after the rows of the
// particles before it in its tile.
__global__ void k_export_nbr(FlatRows flat, int32_t maxn,
                             const int32_t *__restrict__ ncnt,
                             const int32_t *__restrict__ oid, int64_t n,
                             const int64_t *__restrict__ row_off,
                             int32_t *__restrict__ out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t t0 = t / flat.tile * flat.tile;
  int64_t b = t0 * maxn;
  for (int64_t u = t0; u < t; u++) b += row_len(ncnt[u]);
  int64_t base = row_off[oid[t]];
  int c = row_len(ncnt[t]);
  for (int k = 0; k < c; k++) out[base + k] = oid[flat.cidx[b + k]];
}

}  // namespace

// ----------------------------------------------------------- launchers ---

hipError_t launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill<int32_t>, dim3(grid_capped(n, 256)), dim3(256), 0,
                     s, p, n, v);
  return hipGetLastError();
}
hipError_t launch_fill_u64(unsigned long long *p, int64_t n,
                           unsigned long long v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(grid_capped(n, 256)),
                     dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_aos_unpack(const uint8_t *aos, AosLayout lay, int64_t n,
                             const int32_t *col, const BoxGeom *boxes,
                             StaticArrays st, DynArrays dy,
                             unsigned long long *stats, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_aos_unpack, dim3(grid_for(n, 256)), dim3(256), 0, s,
                     aos, lay, n, col, boxes, st, dy, stats);
  return hipGetLastError();
}
hipError_t launch_aos_refresh(const uint8_t *aos, AosLayout lay, int64_t n,
                              const BoxGeom *boxes, StaticArrays st,
                              DynArrays dy, uint32_t fields,
                              unsigned long long *stats, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_aos_refresh, dim3(grid_for(n, 256)), dim3(256), 0, s,
                     aos, lay, n, boxes, st, dy, fields, stats);
  return hipGetLastError();
}

hipError_t launch_aos_pack(uint8_t *aos, AosLayout lay, int64_t n,
                           StaticArrays st, DynArrays dy, const double *fc,
                           double mob, int wrote_step, int any_fungi,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_aos_pack, dim3(grid_for(n, 256)), dim3(256), 0, s, aos,
                     lay, n, st, dy, fc, mob, wrote_step, any_fungi);
  return hipGetLastError();
}

hipError_t launch_bin_count(const double4 *pos, const int32_t *col,
                            const BoxGeom *boxes, BoxGeom g0, int multi,
                            int64_t n, int32_t *key, int32_t *rank,
                            int32_t *count, hipStream_t s, SkinCheck sk) {
  if (n <= 0) return hipSuccess;
  if (multi)
    hipLaunchKernelGGL(k_bin_count<true>, dim3(grid_for(n, 256)), dim3(256), 0,
                       s, pos, col, boxes, g0, n, key, rank, count, sk);
  else
    hipLaunchKernelGGL(k_bin_count<false>, dim3(grid_for(n, 256)), dim3(256),
                       0, s, pos, col, boxes, g0, n, key, rank, count, sk);
  return hipGetLastError();
}

int64_t scan_scratch_elems(int64_t nb) {
  return (nb + kScanTile - 1) / kScanTile + 1;
}

hipError_t launch_scan(int32_t *count, int32_t *start, int64_t nb,
                       int32_t *scratch, hipStream_t s, SkinCheck sk) {
  int64_t nt = (nb + kScanTile - 1) / kScanTile;
  if (nt < 1) nt = 1;
  if (nt > 1024 * 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nt), dim3(kScanThreads), 0,
                     s, count, start, scratch, nb, sk);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nt), dim3(kScanThreads), 0, s,
                     start, scratch, nt, nb);
  return hipGetLastError();
}

hipError_t launch_bin_scatter(const int32_t *key, const int32_t *rank,
                              const int32_t *start, const int32_t *oid,
                              int64_t n, unsigned long long *permk,
                              hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bin_scatter, dim3(grid_for(n, 256)), dim3(256), 0, s,
                     key, rank, start, oid, n, permk);
  return hipGetLastError();
}
hipError_t launch_sort_gather(const unsigned long long *permk,
                              const int32_t *start, int64_t n,
                              StaticArrays src_s, StaticArrays dst_s,
                              DynArrays src_d, DynArrays dst_d,
                              const int32_t *key, int multi, int rods,
                              int32_t *err, hipStream_t s,
                              const double *chem_src, double *chem_dst) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sort_gather, dim3(grid_for(n, 256)), dim3(256), 0, s,
                     permk, start, n, src_s, dst_s, src_d, dst_d, key, multi,
                     rods, err, chem_src, chem_dst);
  return hipGetLastError();
}

hipError_t launch_build_nbr(const double4 *pos, const int32_t *col,
                            const int32_t *type, const double4 *rod,
                            const BoxGeom *boxes, BoxGeom g0, int multi,
                            int rods, const int32_t *start, const int32_t *key,
                            int64_t n, int mode, double crit, double wa,
                            double skin, int32_t *nbr, int32_t *ncnt,
                            int64_t ncap, int32_t maxn, int32_t *err,
                            int32_t *bstats, FlatRows flat, int cols,
                            LaunchShape ls, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  int b = ls.block;
  const size_t stage = rods ? 0 : sizeof(int32_t) * kStage * (size_t)b;
#define BMX_BUILD(M, RD)                                                      \
  hipLaunchKernelGGL((k_build_nbr<M, RD>), dim3(grid_for(n, b)), dim3(b),     \
                     stage, s, pos, col, type, rod, boxes, g0, start, key, n, \
                     mode, crit, wa, skin, nbr, ncnt, ncap, maxn, err, bstats, \
                     flat, cols, ls.remap)
  if (multi && rods)
    BMX_BUILD(true, true);
  else if (multi)
    BMX_BUILD(true, false);
  else if (rods)
    BMX_BUILD(false, true);
  else
    BMX_BUILD(false, false);
#undef BMX_BUILD
  return hipGetLastError();
}

int32_t row_blocks(int64_t n, int block) { return grid_for(n, block); }

// The tail of a call in one launch: the row statistics of the last build
// (nblk > 0), then the error word (OR-ed in: calls enqueued back to back
// add up until the host reads and clears it), the longest row and the entry
// total written straight into host memory mapped for the device (out: 4
// words), and the device's error word cleared.
__global__ __launch_bounds__(1024) void k_flags_out(
    const int32_t *__restrict__ bstats, int32_t nblk,
    int32_t *__restrict__ stats, unsigned long long *__restrict__ total,
    int32_t *__restrict__ err, int32_t *__restrict__ out) {
  __shared__ int32_t shi[16];
  __shared__ unsigned long long shu[16];
  if (nblk > 0) {
    int m = 0;
    unsigned long long sm = 0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
      int v = bstats[2 * b];
      m = v > m ? v : m;
      sm += (unsigned long long)bstats[2 * b + 1];
    }
    m = block_max(m, shi);
    __syncthreads();
    sm = block_sum(sm, shu);
    if (threadIdx.x == 0) {
      stats[0] = m;
      *total = sm;
    }
  }
  if (threadIdx.x == 0) {
    const unsigned long long tot = *total;
    out[0] |= *err;  // the host clears the word when it reads it
    out[1] = stats[0];
    out[2] = (int32_t)(uint32_t)(tot & 0xffffffffull);
    out[3] = (int32_t)(uint32_t)(tot >> 32);
    *err = 0;
  }
}

hipError_t launch_flags_out(const int32_t *bstats, int32_t nblk, int32_t *stats,
                            unsigned long long *total, int32_t *err,
                            int32_t *host_out, hipStream_t s) {
  hipLaunchKernelGGL(k_flags_out, dim3(1), dim3(1024), 0, s, bstats, nblk,
                     stats, total, err, host_out);
  return hipGetLastError();
}

hipError_t launch_reduce_rowstats(const int32_t *bstats, int32_t nblk,
                                  int32_t *stats, unsigned long long *total,
                                  hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_rowstats, dim3(1), dim3(1024), 0, s, bstats,
                     nblk, stats, total);
  return hipGetLastError();
}

size_t flat_lds_bytes(int block, int multi) {
  return (size_t)(block / kWave) * (kFlatWaveBytes + (multi ? kWave * sizeof(int32_t) : 0));
}

hipError_t launch_force_sphere_flat(const double4 *pos_in, double4 *pos_out,
                                    const int32_t *cidx, const uint8_t *cown,
                                    const int32_t *wtot, const int32_t *ncnt,
                                    int64_t n, int32_t maxn,
                                    const int32_t *col,
                                    const BoxGeom *boxes, BoxGeom g0, int multi,
                                    ForceParams fp, double *fc,
                                    const int32_t *fc_oid, double *vmax,
                                    int32_t *err, LaunchShape ls,
                                    hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int b = ls.block;
  const size_t lds = flat_lds_bytes(b, multi);
  if (multi)
    hipLaunchKernelGGL(k_force_sphere_flat<true>, dim3(grid_for(n, b)), dim3(b),
                       lds, s, pos_in, pos_out, cidx, cown, wtot, ncnt, n, maxn,
                       col, boxes, g0, fp, fc, fc_oid, vmax, err, ls.remap);
  else
    hipLaunchKernelGGL(k_force_sphere_flat<false>, dim3(grid_for(n, b)), dim3(b),
                       lds, s, pos_in, pos_out, cidx, cown, wtot, ncnt, n, maxn,
                       col, boxes, g0, fp, fc, fc_oid, vmax, err, ls.remap);
  return hipGetLastError();
}

hipError_t launch_force_general(DynArrays in, DynArrays out, StaticArrays st,
                                const int32_t *cidx, const uint8_t *cown,
                                const int32_t *wtot, int8_t *redo_tile,
                                const int32_t *nbr, const int32_t *ncnt,
                                int64_t ncap, int64_t n, int32_t maxn,
                                const BoxGeom *boxes, BoxGeom g0, int multi,
                                ForceParams fp, double *fc, double *vmax,
                                int32_t *err, LaunchShape ls, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int b = ls.block > 256 ? 256 : ls.block;
  const int64_t tiles = (n + kGTile - 1) / kGTile;
  const int grid = grid_for(tiles, b / kWave);
  const size_t lds = (size_t)(b / kWave) * kGenWaveBytes;
  if (multi) {
    hipLaunchKernelGGL(k_force_general_flat<true>, dim3(grid), dim3(b), lds, s,
                       in, out, st, cidx, cown, wtot, redo_tile, ncnt, n, maxn,
                       boxes, g0, fp, fc, vmax, err, ls.remap);
    hipLaunchKernelGGL(k_force_general_redo<true>, dim3(grid_for(n, 256)),
                       dim3(256), 0, s, in, out, st, redo_tile, nbr, ncnt,
                       ncap, n, boxes, g0, fp, fc, vmax, err);
  } else {
    hipLaunchKernelGGL(k_force_general_flat<false>, dim3(grid), dim3(b), lds, s,
                       in, out, st, cidx, cown, wtot, redo_tile, ncnt, n, maxn,
                       boxes, g0, fp, fc, vmax, err, ls.remap);
    hipLaunchKernelGGL(k_force_general_redo<false>, dim3(grid_for(n, 256)),
                       dim3(256), 0, s, in, out, st, redo_tile, nbr, ncnt,
                       ncap, n, boxes, g0, fp, fc, vmax, err);
  }
  return hipGetLastError();
}

int flat_tile(int segments) { return segments ? kGTile : kWave; }
int64_t flat_entries(int64_t ncap, int tile, int32_t maxn) {
  // + one window: the force pass reads a window before it knows the count
  return (ncap + tile - 1) / tile * (int64_t)tile * maxn + kFlatWin;
}

hipError_t launch_init_bonds(DynArrays dy, StaticArrays st, const int32_t *nbr,
                             const int32_t *ncnt, int64_t ncap, int64_t n,
                             const BoxGeom *boxes, BoxGeom g0, int multi,
                             ForceParams fp, int32_t *err, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (multi)
    hipLaunchKernelGGL(k_init_bonds<true>, dim3(grid_for(n, 128)), dim3(128),
                       0, s, dy, st, nbr, ncnt, ncap, n, boxes, g0, fp, err);
  else
    hipLaunchKernelGGL(k_init_bonds<false>, dim3(grid_for(n, 128)), dim3(128),
                       0, s, dy, st, nbr, ncnt, ncap, n, boxes, g0, fp, err);
  return hipGetLastError();
}

hipError_t launch_skin_check(double *vpart, int32_t nslots, int32_t stride,
                             double limit, int clear,
                             unsigned long long *slotmax, int32_t *err,
                             hipStream_t s) {
  int chunks = (stride + kSkinChunk - 1) / kSkinChunk;
  if (chunks < 1) chunks = 1;
  hipLaunchKernelGGL(k_skin_partial, dim3(chunks, nslots > 0 ? nslots : 1),
                     dim3(256), 0, s, vpart, stride, clear, slotmax);
  hipLaunchKernelGGL(k_skin_final, dim3(1), dim3(64), 0, s, slotmax, nslots,
                     limit, err);
  return hipGetLastError();
}

hipError_t launch_export_fc(const double *fc, int64_t n,
                            double *out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_export_fc, dim3(grid_for(n, 256)), dim3(256), 0, s, fc, n,
                     out);
  return hipGetLastError();
}
hipError_t launch_export_pos(const double4 *pos, const int32_t *oid, int64_t n,
                             double *out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_export_pos, dim3(grid_for(n, 256)), dim3(256), 0, s,
                     pos, oid, n, out);
  return hipGetLastError();
}
hipError_t launch_export_orient(const double4 *rod, const int32_t *oid,
                                int64_t n, double *out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_export_orient, dim3(grid_for(n, 256)), dim3(256), 0, s,
                     rod, oid, n, out);
  return hipGetLastError();
}
hipError_t launch_export_i32(const int32_t *v, const int32_t *oid, int64_t n,
                             int32_t *out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_export_i32, dim3(grid_for(n, 256)), dim3(256), 0, s, v,
                     oid, n, out);
  return hipGetLastError();
}
hipError_t launch_export_counts(const int32_t *ncnt, const int32_t *oid,
                                int64_t n, int64_t *out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_export_counts, dim3(grid_for(n, 256)), dim3(256), 0, s,
                     ncnt, oid, n, out);
  return hipGetLastError();
}
hipError_t launch_export_nbr(FlatRows flat, int32_t maxn, const int32_t *ncnt,
                             const int32_t *oid, int64_t n,
                             const int64_t *row_off, int32_t *out,
                             hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_export_nbr, dim3(grid_for(n, 256)), dim3(256), 0, s,
                     flat, maxn, ncnt, oid, n, row_off, out);
  return hipGetLastError();
}

}  // namespace bmx
