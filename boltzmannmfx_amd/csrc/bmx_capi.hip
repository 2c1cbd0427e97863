// bmx_capi.hip -- the C ABI (include/bmx_mi355x.h) over the gfx950 kernels.
//
// One context = one HIP stream, one set of device buffers, one or more
// periodic boxes.  Per step the host enqueues (every rebuild_every steps) the
// binning and list build, then one fused force+update launch; nothing is
// copied back until the caller asks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bmx_mi355x.h"
#include "bmx_kernels.h"

using namespace bmx;

namespace {

thread_local std::string g_last_error;

bmx_status fail(bmx_status st, const std::string &msg) {
  g_last_error = msg;
  return st;
}

#define HIPCHK(expr)                                                        \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess)                                                   \
      return fail(BMX_ERR_HIP, std::string(#expr) + ": " +                  \
                                   hipGetErrorString(e_));                  \
  } while (0)

template <typename T>
struct DevBuf;

// releases a call-local device buffer on every way out of the call
template <typename T>
struct Releaser {
  DevBuf<T> &b;
  ~Releaser() { b.release(); }
};

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  // fixed: p is a slice of memory the buffer does not own (the static
  // scratch of the allocation-free entry point); it never grows or frees
  bool fixed = false;
  hipError_t alloc(size_t count) {
    if (count <= n && p) return hipSuccess;
    if (fixed) return hipErrorOutOfMemory;
    release();
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc((void **)&p, bytes);
    if (e == hipSuccess) n = count;
    return e;
  }
  void release() {
    if (fixed) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

}  // namespace

struct bmx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;      // the stream every call enqueues on
  hipStream_t own_stream = nullptr;  // the context's own (stream when unset)
  // bmx_evolve_async: steps in flight, done_ev recorded after them
  bool pending = false;
  hipEvent_t done_ev = nullptr;

  double fpar[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double mobility = 1.0;
  bool have_fpar = false;
  bmx_geometry geom{};
  bool have_geom = false;
  bmx_neighbor_params nbp{};
  bmx_launch_params lp{64, 1};  // one wave per block: the fastest shape measured

  // particles
  int64_t n = 0, ncap = 0;
  int32_t ncol = 1;
  std::vector<bmx_geometry> colony_geom;
  std::vector<int64_t> colony_start;
  std::vector<BoxGeom> boxes, boxes_next;
  DevBuf<BoxGeom> d_boxes;
  DevBuf<int32_t> d_col_in;
  bool bound = false;
  bool any_fungi = false, any_yeast = false;
  double rmax = 0.0, lmax = 0.0;

  // sorted arrays
  DevBuf<int32_t> oid[2], type[2], col[2], key[2];
  DevBuf<double> bscale[2];
  DevBuf<double4> pos[2], rod[2], trig[2];
  DevBuf<Topo> topo[2];
  int cs = 0;  // current static buffer
  int cd = 0;  // current dynamic buffer

  // record array: recs is the device record array the kernels read and
  // write, either the context's own copy (d_aos) or, bound in place, the
  // caller's device records
  DevBuf<uint8_t> d_aos;
  uint8_t *recs = nullptr;
  bool in_place = false;
  void *user_aos = nullptr;
  bool aos_on_device = false;
  // tiles bound as one domain: sync copies each tile's records back
  std::vector<void *> tiles;
  std::vector<int64_t> tile_start;
  AosLayout lay{};

  // binning and rows
  int32_t nbins = 0;
  double crit = 0.0;
  double verlet_radius = 0.0;
  DevBuf<int32_t> d_keytmp, d_rank, d_count, d_start, d_scan;
  bool count_clear = false;  // d_count holds zeros (the scan resets it)
  DevBuf<unsigned long long> d_permk;
  int32_t maxn = 96;
  DevBuf<int32_t> d_nbr, d_ncnt;
  // flat row layout of the force pass: per-tile entry totals and bases,
  // entries back to back, the particle's place in its tile per entry
  int flat_tile = 64;
  DevBuf<int32_t> d_wtot, d_cidx;
  DevBuf<uint8_t> d_cown;
  DevBuf<int8_t> d_redo;  // segment tiles left to the row loop this step
  // species A, B, C and volume, and their exchange increments, in original
  // order (loaded from the records at the first exchange)
  // Adjoint working memory, kept across calls: checkpoints, loss target and
  // partials, multipliers, and two history sets (positions of every step of
  // a segment, and the rows, row counts and order of each of its rebuilds)
  // so that one segment is recomputed while the previous one is reversed.
  struct AdjHist {
    DevBuf<double4> pos;
    DevBuf<int32_t> idx, tot, cnt, oid;
    DevBuf<uint8_t> own;
  };
  struct AdjBufs {
    DevBuf<double> target, gr, acc, lpart, out;
    DevBuf<double4> ckpt, lam0, lam1, lamo, fin;
    AdjHist h[2];
    hipStream_t rev = nullptr;  // the reverse steps (default priority)
    hipEvent_t rdone[2] = {nullptr, nullptr}, adone[2] = {nullptr, nullptr};
    hipEvent_t join = nullptr;
  } adj;
  // chemistry state (chem:: layout, sorted order: d_chemst[cc] is current,
  // the rebuild gathers it into the other buffer), loaded from the records
  // at the first exchange or mesh -> particle call after a bind and written
  // back when the records are packed
  DevBuf<double> d_chemst[2];
  int cc = 0;
  bool have_chem = false;
  // particle <-> mesh coupling: bins over the level mesh, staged fields
  DevBuf<int32_t> d_mkey, d_mrank, d_mcount, d_mstart, d_mscan;
  DevBuf<unsigned long long> d_mperm;
  DevBuf<double> d_mout, d_min;
  int32_t n_real = 0, n_int = 0;
  bool have_list = false;
  int64_t step = 0, last_build = -1;

  // outputs and flags
  DevBuf<double> d_fc;  // last step's fc: original order, 6 per particle
  bool wrote_step = false;
  DevBuf<int32_t> d_err, d_stats, d_bstats;
  DevBuf<unsigned long long> d_total, d_ustats;
  DevBuf<double> d_vpart;      // [slot][vstride] max dt*|vel| per wave/block
  DevBuf<unsigned long long> d_slotmax;  // per-slot maxima (skin check)
  int32_t vmax_slots = 1;
  int32_t vstride = 1;         // waves of ncap particles (blocks of 64)
  int32_t row_nblk = 0;        // blocks of the last row build
  // pinned: err, max row, total (2 words), ..., guard periods' errors
  // (8, 9), the chemistry's size maxima (16..19), the adjoint forward's
  // error word (20)
  int32_t *h_flags = nullptr;
  int32_t *h_flags_dev = nullptr;  // h_flags as the device addresses it

  // timing: events around every rebuild and around every run of force
  // launches between two rebuilds (or a rebuild and the end of a call)
  bool timing = false;
  hipEvent_t run_e0 = nullptr;  // start of the open run of force launches
  int32_t run_n = 0;            // launches in the open run
  struct Run {
    hipEvent_t e0, e1;
    int32_t n;
  };
  std::vector<Run> ev_force;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_rebuild;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  double force_ms = 0.0, rebuild_ms = 0.0;
  int64_t force_launches = 0, rebuilds = 0;
  int64_t last_entries = 0;
  int32_t last_max_row = 0;
  int64_t rebuilds_total = 0;  // list builds since bind (graph replays too)

  // guarded Verlet steps: each rebuild period is checked as it ends and a
  // period that moved a particle more than skin/2 is taken again with lists
  // rebuilt every step (two snapshots: one period in flight while the one
  // before it is checked)
  bool guard = false;
  bool guarding = false;  // inside a guarded evolve: rebuild() checks nothing
  struct Snap {
    DevBuf<double4> pos, rod, trig;
    DevBuf<Topo> topo;
    DevBuf<int32_t> oid, type, key, col;
    DevBuf<double> bscale;
    hipEvent_t done = nullptr;
  } gsnap[2];
  int64_t guard_redone = 0;  // periods taken again since bind

  // captured step sequences
  hipGraphExec_t graph_exec = nullptr;
  double graph_dt = 0.0;
  int32_t graph_nsteps = 0;
  int64_t graph_step0 = -1;
  int graph_cs = -1, graph_cd = -1;
  bool graph_have_list = false;
};

namespace {

LaunchShape shape(const bmx_ctx *c) {
  LaunchShape s;
  s.block = c->lp.block_size;
  s.remap = c->lp.xcd_remap;
  return s;
}

StaticArrays static_arrays(bmx_ctx *c, int k) {
  StaticArrays s;
  s.oid = c->oid[k].p;
  s.type = c->type[k].p;
  s.col = c->ncol > 1 ? c->col[k].p : nullptr;
  s.bscale = c->bscale[k].p;
  s.key = c->key[k].p;
  return s;
}
DynArrays dyn_arrays(bmx_ctx *c, int k) {
  DynArrays d;
  d.pos = c->pos[k].p;
  d.rod = c->rod[k].p;
  d.trig = c->trig[k].p;
  d.topo = c->topo[k].p;
  return d;
}

ForceParams force_params(const bmx_ctx *c, double dt) {
  ForceParams fp;
  fp.WA = c->fpar[0];
  fp.mu = c->fpar[1];
  fp.zw = c->fpar[2];
  fp.WZ = c->fpar[3];
  fp.xi = c->fpar[4];
  fp.grav = c->fpar[5];
  fp.kb = c->fpar[6];
  fp.bcut = c->fpar[7];
  fp.visc = c->fpar[8];
  fp.mob = c->mobility;
  fp.dt = dt;
  return fp;
}

hipEvent_t next_event(bmx_ctx *c) {
  if (c->ev_used == c->ev_pool.size()) {
    hipEvent_t e;
    // timing only: read after the stream is synchronized, so no system
    // fence (cache write-back) is needed when the event is recorded
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess)
      return nullptr;
    c->ev_pool.push_back(e);
  }
  return c->ev_pool[c->ev_used++];
}

// Collect finished event pairs into the running totals.
bmx_status drain_timing(bmx_ctx *c) {
  for (auto &r : c->ev_force) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, r.e0, r.e1));
    c->force_ms += ms;
    c->force_launches += r.n;
  }
  for (auto &pr : c->ev_rebuild) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, pr.first, pr.second));
    c->rebuild_ms += ms;
    c->rebuilds++;
  }
  c->ev_force.clear();
  c->ev_rebuild.clear();
  c->ev_used = 0;
  return BMX_OK;
}

// Box geometry with bins chosen for the neighbor policy.
bmx_status make_box(const bmx_ctx *c, const bmx_geometry &gm, BoxGeom &b,
                    int32_t bin_offset) {
  for (int d = 0; d < 3; d++) {
    if (!(gm.prob_hi[d] > gm.prob_lo[d]))
      return fail(BMX_ERR_BAD_ARG, "prob_hi must exceed prob_lo");
    b.plo[d] = gm.prob_lo[d];
    b.phi[d] = gm.prob_hi[d];
    b.L[d] = gm.prob_hi[d] - gm.prob_lo[d];
    b.halfL[d] = 0.5 * b.L[d];
    b.periodic[d] = gm.periodic[d] ? 1 : 0;
    int nb;
    if (c->nbp.mode == BMX_NBR_STENCIL) {
      nb = gm.n_cell[d];
      if (nb < 1) return fail(BMX_ERR_BAD_ARG, "n_cell must be >= 1");
      if (b.periodic[d] && nb < 4)
        return fail(BMX_ERR_BAD_ARG,
                    "stencil lists need >= 4 cells along a periodic axis");
    } else {
      double pct = c->nbp.bin_factor_pct > 0 ? c->nbp.bin_factor_pct : 100;
      double edge = c->verlet_radius * pct / 100.0;
      double f = std::floor(b.L[d] / edge);
      nb = f < 1.0 ? 1 : (f > 1.0e6 ? 1000000 : (int)f);
      if (b.periodic[d] && nb < 3)
        return fail(BMX_ERR_BAD_ARG,
                    "periodic box shorter than 3 interaction radii");
    }
    b.nbin[d] = nb;
    double dx = b.L[d] / (double)nb;
    b.dxi[d] = 1.0 / dx;
  }
  b.bin_offset = bin_offset;
  b.pad = 0;
  int64_t tot = (int64_t)b.nbin[0] * b.nbin[1] * b.nbin[2];
  if (tot > (int64_t)1 << 30) return fail(BMX_ERR_BAD_ARG, "too many bins");
  return BMX_OK;
}

// Bins of every box for the current force and neighbor parameters.  The
// boxes are formed first and committed only when all of them are valid, so
// a rejected parameter change leaves the previous binning in place.
bmx_status setup_boxes(bmx_ctx *c) {
  // Verlet radius: the largest r_lm any pair can have, plus the skin
  // (K.H:23-40).
  double rl = 0.0;
  if (c->any_yeast) rl = std::max(rl, 2.0 * c->rmax + c->fpar[0]);
  if (c->any_fungi) rl = std::max(rl, 1.5 * (2.0 * c->lmax) + c->fpar[0]);
  const double vr = rl + c->nbp.skin;
  double crit;
  if (c->nbp.mode == BMX_NBR_VERLET) {
    if (!(vr > 0.0))
      return fail(BMX_ERR_BAD_ARG, "interaction radius must be positive");
    crit = vr * vr;
  } else {
    crit = c->nbp.neighborhood;
  }
  const double vr_old = c->verlet_radius;
  c->verlet_radius = vr;  // make_box reads it
  std::vector<BoxGeom> &boxes = c->boxes_next;
  boxes.resize((size_t)c->ncol);
  int64_t off = 0;
  for (int k = 0; k < c->ncol; k++) {
    bmx_status st = make_box(c, c->colony_geom[k], boxes[k], (int32_t)off);
    if (st == BMX_OK) {
      off += (int64_t)boxes[k].nbin[0] * boxes[k].nbin[1] * boxes[k].nbin[2];
      if (off > (int64_t)1 << 30) st = fail(BMX_ERR_BAD_ARG, "too many bins");
    }
    if (st != BMX_OK) {
      c->verlet_radius = vr_old;
      return st;
    }
  }
  // committed: the rows of the previous bins are stale from here on
  c->have_list = false;
  c->crit = crit;
  c->boxes.swap(boxes);
  c->nbins = (int32_t)off;
  c->count_clear = false;
  c->graph_step0 = -1;
  if (c->d_count.fixed && (size_t)c->nbins + 1 > c->d_count.n) {
    c->bound = false;
    return fail(BMX_ERR_BAD_ARG, "the binning needs more bins (" +
                                     std::to_string(c->nbins) +
                                     ") than the fixed scratch holds");
  }
  hipError_t e = c->d_boxes.alloc(c->ncol);
  if (e == hipSuccess)
    e = hipMemcpyAsync(c->d_boxes.p, c->boxes.data(), sizeof(BoxGeom) * c->ncol,
                       hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = c->d_count.alloc(c->nbins + 1);
  if (e == hipSuccess) e = c->d_start.alloc(c->nbins + 1);
  if (e == hipSuccess) e = c->d_scan.alloc(scan_scratch_elems(c->nbins) + 1);
  if (e != hipSuccess) {
    // the bin arrays no longer match the boxes: nothing may run on them
    c->bound = false;
    return fail(BMX_ERR_HIP, std::string("bin arrays: ") + hipGetErrorString(e));
  }
  return BMX_OK;
}

// Row arrays for ncap particles and maxn entries per row.  A failure leaves
// the context unbound (the rows would be smaller than the kernels assume).
bmx_status ensure_rows(bmx_ctx *c) {
  c->have_list = false;
  size_t need = (size_t)c->maxn * (size_t)c->ncap;
  c->flat_tile = flat_tile(c->any_fungi ? 1 : 0);
  const int64_t nw = (c->ncap + c->flat_tile - 1) / c->flat_tile;
  const int64_t nflat = flat_entries(c->ncap, c->flat_tile, c->maxn);
  hipError_t e = c->d_nbr.alloc(need);
  if (e == hipSuccess) e = c->d_ncnt.alloc(c->ncap);
  if (e == hipSuccess) e = c->d_wtot.alloc(nw + 1);
  if (e == hipSuccess) e = c->d_redo.alloc(nw + 1);
  if (e == hipSuccess) e = c->d_cidx.alloc(nflat);
  if (e == hipSuccess) e = c->d_cown.alloc(nflat);
  if (e != hipSuccess) {
    c->bound = false;
    return fail(BMX_ERR_HIP, std::string("row arrays: ") + hipGetErrorString(e));
  }
  return BMX_OK;
}

// Per-block partial arrays sized for the smallest block size; the skin
// check reads every slot, so unused tails stay zero.
bmx_status alloc_partials(bmx_ctx *c) {
  c->vmax_slots = std::max(1, c->nbp.rebuild_every);
  c->vstride = std::max(1, row_blocks(c->ncap, 64));
  hipError_t e = c->d_vpart.alloc((size_t)c->vmax_slots * c->vstride);
  if (e == hipSuccess)
    e = launch_fill_u64(reinterpret_cast<unsigned long long *>(c->d_vpart.p),
                        (int64_t)c->vmax_slots * c->vstride, 0ull, c->stream);
  if (e == hipSuccess) e = c->d_bstats.alloc(2 * (size_t)c->vstride);
  if (e == hipSuccess) e = c->d_slotmax.alloc(c->vmax_slots);
  if (e == hipSuccess)
    e = launch_fill_u64(c->d_slotmax.p, c->vmax_slots, 0ull, c->stream);
  if (e != hipSuccess) {
    c->bound = false;
    return fail(BMX_ERR_HIP, std::string("partials: ") + hipGetErrorString(e));
  }
  return BMX_OK;
}

FlatRows flat_rows(bmx_ctx *c) {
  return FlatRows{c->flat_tile, c->d_cidx.p, c->d_cown.p, c->d_wtot.p};
}

bool verlet_checked(const bmx_ctx *c) {
  return c->nbp.mode == BMX_NBR_VERLET && c->nbp.rebuild_every > 1;
}

// Ends the open run of timed force launches (before a rebuild, at the end
// of a call).
bmx_status close_run(bmx_ctx *c) {
  if (!c->run_e0) return BMX_OK;
  hipEvent_t e1 = next_event(c);
  if (!e1) return fail(BMX_ERR_HIP, "event allocation failed");
  HIPCHK(hipEventRecord(e1, c->stream));
  c->ev_force.push_back({c->run_e0, e1, c->run_n});
  c->run_e0 = nullptr;
  c->run_n = 0;
  return BMX_OK;
}

// Binning + gather into the other buffers + list build.  With a Verlet list
// in force, the displacements since that list was built are checked against
// skin/2 first and their partials cleared, so every caller that rebuilds
// (steps, exchange, fusion, InitBonds) starts the next period from zero.
// check_skin = false is for passes that repeat checked arithmetic.  The
// force and adjoint passes read the flat rows only; cols = true also writes
// the column-major rows of a sphere context, for the passes that walk rows
// that way (segment contexts always have them).
bmx_status rebuild(bmx_ctx *c, bool check_skin = true, bool cols = false) {
  hipStream_t s = c->stream;
  if (c->timing) {
    bmx_status st = close_run(c);
    if (st != BMX_OK) return st;
  }
  // the displacement check of the period this rebuild ends rides on the
  // bin count and the scan
  SkinCheck skc{};
  if (check_skin && !c->guarding && c->have_list && verlet_checked(c))
    skc = SkinCheck{c->d_vpart.p, c->vmax_slots, c->vstride, c->d_slotmax.p,
                    0.5 * c->nbp.skin, c->d_err.p};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c->timing) {
    e0 = next_event(c);
    e1 = next_event(c);
    if (!e0 || !e1) return fail(BMX_ERR_HIP, "event allocation failed");
    HIPCHK(hipEventRecord(e0, s));
  }
  int multi = c->ncol > 1;
  const BoxGeom g0 = c->boxes[0];
  StaticArrays ss = static_arrays(c, c->cs), ds = static_arrays(c, 1 - c->cs);
  DynArrays sd = dyn_arrays(c, c->cd), dd = dyn_arrays(c, 1 - c->cd);
  // the scan leaves the counts at zero; fill only a fresh array
  if (!c->count_clear) HIPCHK(launch_fill_i32(c->d_count.p, c->nbins + 1, 0, s));
  HIPCHK(launch_bin_count(sd.pos, ss.col, c->d_boxes.p, g0, multi, c->n,
                          c->d_keytmp.p, c->d_rank.p, c->d_count.p, s, skc));
  HIPCHK(launch_scan(c->d_count.p, c->d_start.p, c->nbins, c->d_scan.p, s, skc));
  c->count_clear = true;
  HIPCHK(launch_bin_scatter(c->d_keytmp.p, c->d_rank.p, c->d_start.p, ss.oid,
                            c->n, c->d_permk.p, s));
  const double *chem_src = c->have_chem ? c->d_chemst[c->cc].p : nullptr;
  double *chem_dst = c->have_chem ? c->d_chemst[1 - c->cc].p : nullptr;
  HIPCHK(launch_sort_gather(c->d_permk.p, c->d_start.p, c->n, ss, ds, sd, dd,
                            c->d_keytmp.p, multi, c->any_fungi ? 1 : 0,
                            c->d_err.p, s, chem_src, chem_dst));
  c->cs = 1 - c->cs;
  c->cd = 1 - c->cd;
  if (c->have_chem) c->cc = 1 - c->cc;
  StaticArrays st = static_arrays(c, c->cs);
  DynArrays dy = dyn_arrays(c, c->cd);
  HIPCHK(launch_build_nbr(dy.pos, st.col, st.type, dy.rod, c->d_boxes.p, g0,
                          multi, c->any_fungi ? 1 : 0, c->d_start.p, st.key,
                          c->n, c->nbp.mode == BMX_NBR_STENCIL ? 1 : 0,
                          c->crit, c->fpar[0], c->nbp.skin, c->d_nbr.p,
                          c->d_ncnt.p, c->ncap, c->maxn, c->d_err.p,
                          c->d_bstats.p, flat_rows(c), cols ? 1 : 0, shape(c),
                          s));
  c->row_nblk = row_blocks(c->n, c->lp.block_size);
  if (c->timing) {
    HIPCHK(hipEventRecord(e1, s));
    c->ev_rebuild.push_back({e0, e1});
  }
  c->have_list = true;
  c->last_build = c->step;
  c->rebuilds_total++;
  return BMX_OK;
}


bmx_status enqueue_step(bmx_ctx *c, double dt, bool write_fc) {
  hipStream_t s = c->stream;
  bool need = !c->have_list ||
              (c->step - c->last_build) >= (int64_t)c->nbp.rebuild_every;
  if (need) {
    bmx_status st = rebuild(c);
    if (st != BMX_OK) return st;
  }
  int slot = (int)(c->step - c->last_build);
  double *vm = verlet_checked(c)
                   ? c->d_vpart.p + (int64_t)(slot % c->vmax_slots) * c->vstride
                   : nullptr;
  if (c->timing && !c->run_e0) {
    c->run_e0 = next_event(c);
    if (!c->run_e0) return fail(BMX_ERR_HIP, "event allocation failed");
    HIPCHK(hipEventRecord(c->run_e0, s));
  }
  ForceParams fp = force_params(c, dt);
  int multi = c->ncol > 1;
  double *fc = write_fc ? c->d_fc.p : nullptr;
  if (!c->any_fungi) {
    HIPCHK(launch_force_sphere_flat(
        c->pos[c->cd].p, c->pos[1 - c->cd].p, c->d_cidx.p, c->d_cown.p,
        c->d_wtot.p, c->d_ncnt.p, c->n, c->maxn,
        multi ? c->col[c->cs].p : nullptr, c->d_boxes.p, c->boxes[0], multi,
        fp, fc, c->oid[c->cs].p, vm, c->d_err.p, shape(c), s));
  } else {
    HIPCHK(launch_force_general(
        dyn_arrays(c, c->cd), dyn_arrays(c, 1 - c->cd), static_arrays(c, c->cs),
        c->d_cidx.p, c->d_cown.p, c->d_wtot.p, c->d_redo.p, c->d_nbr.p,
        c->d_ncnt.p, c->ncap, c->n, c->maxn, c->d_boxes.p, c->boxes[0], multi,
        fp, fc, vm, c->d_err.p, shape(c), s));
  }
  if (c->timing) c->run_n++;
  c->cd = 1 - c->cd;
  c->step++;
  if (write_fc) c->wrote_step = true;
  return BMX_OK;
}

// One force launch without bookkeeping of rebuilds (the adjoint drives its
// own rebuild schedule).  slot >= 0 records the step's displacement
// partials for the skin check in that slot.
bmx_status enqueue_force_only(bmx_ctx *c, double dt, int slot) {
  ForceParams fp = force_params(c, dt);
  double *vm = (slot >= 0 && verlet_checked(c))
                   ? c->d_vpart.p + (int64_t)(slot % c->vmax_slots) * c->vstride
                   : nullptr;
  HIPCHK(launch_force_sphere_flat(
      c->pos[c->cd].p, c->pos[1 - c->cd].p, c->d_cidx.p, c->d_cown.p,
      c->d_wtot.p, c->d_ncnt.p, c->n, c->maxn, nullptr, c->d_boxes.p,
      c->boxes[0], 0, fp, nullptr, nullptr, vm, c->d_err.p, shape(c),
      c->stream));
  c->cd = 1 - c->cd;
  c->step++;
  return BMX_OK;
}

// Reads the error word and list statistics; clears the error word.
// Enqueues the read-back of the error word and list statistics (and clears
// the error word) on the context's stream.
bmx_status enqueue_flags(bmx_ctx *c) {
  hipStream_t s = c->stream;
  if (c->timing) {
    bmx_status st = close_run(c);
    if (st != BMX_OK) return st;
  }
  if (c->h_flags_dev) {
    // one launch writes the words into host memory the device maps
    HIPCHK(launch_flags_out(c->d_bstats.p, c->row_nblk, c->d_stats.p,
                            c->d_total.p, c->d_err.p, c->h_flags_dev, s));
    return BMX_OK;
  }
  if (c->row_nblk > 0)
    HIPCHK(launch_reduce_rowstats(c->d_bstats.p, c->row_nblk, c->d_stats.p,
                                  c->d_total.p, s));
  HIPCHK(hipMemcpyAsync(c->h_flags, c->d_err.p, sizeof(int32_t),
                        hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(c->h_flags + 1, c->d_stats.p, sizeof(int32_t),
                        hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(c->h_flags + 2, c->d_total.p, sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, s));
  HIPCHK(launch_fill_i32(c->d_err.p, 1, 0, s));
  return BMX_OK;
}

// Interprets the words enqueue_flags brought back, once they have landed.
bmx_status read_flags(bmx_ctx *c) {
  if (c->timing) {
    bmx_status st = drain_timing(c);
    if (st != BMX_OK) return st;
  }
  int32_t err = c->h_flags[0];
  c->h_flags[0] = 0;
  c->last_max_row = c->h_flags[1];
  unsigned long long tot;
  std::memcpy(&tot, c->h_flags + 2, sizeof(tot));
  c->last_entries = (int64_t)tot;
  if (err & ERRBIT_OVERFLOW) {
    char buf[160];
    std::snprintf(buf, sizeof buf,
                  "neighbor row of %d entries exceeds max_neighbors=%d",
                  c->last_max_row, c->maxn);
    c->have_list = false;
    return fail(BMX_ERR_NBR_OVERFLOW, buf);
  }
  if (err & ERRBIT_BOND)
    return fail(BMX_ERR_BOND_OVERFLOW, "a segment would get a fifth bond");
  if (err & ERRBIT_SKIN)
    return fail(BMX_ERR_SKIN,
                "displacement since the last rebuild exceeded skin/2");
  if (err & ERRBIT_NAN)
    return fail(BMX_ERR_NAN, "non-finite force (coincident particles?)");
  return BMX_OK;
}

// Reads the error word and list statistics; clears the error word.
bmx_status check_flags(bmx_ctx *c) {
  bmx_status st = enqueue_flags(c);
  if (st != BMX_OK) return st;
  HIPCHK(hipStreamSynchronize(c->stream));
  return read_flags(c);
}

// Completes the steps a bmx_evolve_async call left in flight: waits for them
// and returns their status.
bmx_status settle(bmx_ctx *c) {
  if (!c->pending) return BMX_OK;
  c->pending = false;
  HIPCHK(hipEventSynchronize(c->done_ev));
  return read_flags(c);
}

// Every call that needs bound particles comes here first; steps still in
// flight from bmx_evolve_async are completed, and their status returned.
bmx_status require_ready(bmx_ctx *c) {
  if (!c) return fail(BMX_ERR_BAD_ARG, "null context");
  bmx_status st = settle(c);
  if (st != BMX_OK) return st;
  if (!c->bound) return fail(BMX_ERR_STATE, "no particles bound");
  return BMX_OK;
}

void reset_graph(bmx_ctx *c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  c->graph_exec = nullptr;
  c->graph_step0 = -1;
}

// Writes the state the force pass changes into the device record copy.
// The layout holds the amounts taken from the fluid (rdata[first_data + 6..8]).
int chem_rel(const bmx_ctx *c) { return c->n_real >= rec::first_data + 9 ? 1 : 0; }

// The chemistry state from the records, unless it is loaded.  A captured
// step sequence does not move it, so the graph is captured again.
bmx_status load_chem(bmx_ctx *c) {
  if (c->have_chem) return BMX_OK;
  for (auto &b : c->d_chemst) HIPCHK(b.alloc((size_t)c->ncap * chem::stride));
  c->cc = 0;
  HIPCHK(launch_chem_unpack(c->recs, c->lay, c->n, chem_rel(c), c->oid[c->cs].p,
                            c->d_chemst[0].p, c->stream));
  c->have_chem = true;
  reset_graph(c);
  return BMX_OK;
}

// Writes the chemistry state back into the records and drops it (reloaded
// at the next use), for passes that keep their own copies of the sorted
// arrays (the guard's snapshots, the autotuner's trials, the adjoint).
bmx_status drop_chem(bmx_ctx *c) {
  if (!c->have_chem) return BMX_OK;
  HIPCHK(launch_chem_pack(c->recs, c->lay, c->n, chem_rel(c),
                          c->d_chemst[c->cc].p, c->oid[c->cs].p,
                          c->type[c->cs].p, c->pos[c->cd].p, c->rod[c->cd].p,
                          c->stream));
  c->have_chem = false;
  reset_graph(c);
  return BMX_OK;
}

bmx_status pack_records(bmx_ctx *c) {
  HIPCHK(launch_aos_pack(c->recs, c->lay, c->n, static_arrays(c, c->cs),
                         dyn_arrays(c, c->cd), c->d_fc.p, c->mobility,
                         c->wrote_step ? 1 : 0, c->any_fungi ? 1 : 0, c->stream));
  if (c->have_chem)
    HIPCHK(launch_chem_pack(c->recs, c->lay, c->n, chem_rel(c),
                            c->d_chemst[c->cc].p, c->oid[c->cs].p,
                            c->type[c->cs].p, c->pos[c->cd].p, c->rod[c->cd].p,
                            c->stream));
  return BMX_OK;
}

// A bind replaces the previous one whatever its outcome: the context is
// unbound from the first line, and only a bind that completes binds it
// again, so no kernel can run on arrays sized for other particles.
bmx_status bind_common(bmx_ctx *c, const void *aos, const int64_t *cstart,
                       int32_t ncol, const bmx_geometry *geoms,
                       const bmx_layout *layout, int32_t on_device,
                       bool in_place = false) {
  reset_graph(c);
  c->bound = false;
  c->have_list = false;
  c->row_nblk = 0;
  if (!aos && cstart[ncol] > 0)
    return fail(BMX_ERR_BAD_ARG, "null record array");
  if (!c->have_fpar) return fail(BMX_ERR_STATE, "force params not set");
  bmx_layout L = layout ? *layout : bmx_default_layout();
  if (L.stride_bytes <= 0 || L.pos_offset % 8 || L.rdata_offset % 8 ||
      L.idata_offset % 4 || L.stride_bytes % 8 || L.n_real < 26 ||
      L.n_int < 35 || L.pos_offset + 24 > L.stride_bytes ||
      L.rdata_offset + 8 * L.n_real > L.stride_bytes ||
      L.idata_offset + 4 * L.n_int > L.stride_bytes)
    return fail(BMX_ERR_BAD_ARG, "invalid record layout");
  int64_t n = cstart[ncol];
  if (n < 0 || n > (int64_t)INT32_MAX / 2)
    return fail(BMX_ERR_BAD_ARG, "particle count out of range");
  for (int k = 0; k < ncol; k++)
    if (cstart[k + 1] < cstart[k])
      return fail(BMX_ERR_BAD_ARG, "colony offsets must be non-decreasing");
  for (int k = 0; k < ncol; k++)
    for (int d = 0; d < 3; d++)
      if (!(std::isfinite(geoms[k].prob_lo[d]) &&
            std::isfinite(geoms[k].prob_hi[d]) &&
            geoms[k].prob_hi[d] > geoms[k].prob_lo[d]))
        return fail(BMX_ERR_BAD_ARG, "prob_hi must exceed prob_lo");
  hipStream_t s = c->stream;
  c->n = n;
  c->ncap = std::max<int64_t>(((n + 255) / 256) * 256, 256);
  c->ncol = ncol;
  c->colony_geom.assign(geoms, geoms + ncol);
  c->colony_start.assign(cstart, cstart + ncol + 1);
  c->lay.stride = L.stride_bytes;
  c->lay.pos = L.pos_offset;
  c->lay.rdata = L.rdata_offset;
  c->lay.id = L.id_offset;
  c->lay.cpu = L.cpu_offset;
  c->lay.idata = L.idata_offset;
  size_t cap = (size_t)c->ncap;
  for (int k = 0; k < 2; k++) {
    HIPCHK(c->oid[k].alloc(cap));
    HIPCHK(c->type[k].alloc(cap));
    HIPCHK(c->key[k].alloc(cap));
    HIPCHK(c->bscale[k].alloc(cap));
    HIPCHK(c->pos[k].alloc(cap));
    HIPCHK(c->rod[k].alloc(cap));
    HIPCHK(c->trig[k].alloc(cap));
    HIPCHK(c->topo[k].alloc(cap));
    if (ncol > 1) HIPCHK(c->col[k].alloc(cap));
  }
  HIPCHK(c->d_keytmp.alloc(cap));
  HIPCHK(c->d_rank.alloc(cap));
  HIPCHK(c->d_permk.alloc(cap));
  HIPCHK(c->d_fc.alloc(6 * cap));
  HIPCHK(c->d_ustats.alloc(8));
  size_t bytes = (size_t)n * (size_t)L.stride_bytes;
  c->aos_on_device = on_device != 0;
  c->user_aos = const_cast<void *>(aos);
  c->tiles.clear();
  c->tile_start.clear();
  c->in_place = in_place;
  if (in_place) {
    // the caller's device records are the context's record array
    c->recs = static_cast<uint8_t *>(const_cast<void *>(aos));
  } else {
    HIPCHK(c->d_aos.alloc(bytes));
    if (bytes)
      HIPCHK(hipMemcpyAsync(c->d_aos.p, aos, bytes,
                            on_device ? hipMemcpyDeviceToDevice
                                      : hipMemcpyHostToDevice,
                            s));
    c->recs = c->d_aos.p;
  }
  // Boxes before unpack (wrap needs them); bins are set after the stats.
  c->boxes.resize(ncol);
  for (int k = 0; k < ncol; k++) {
    BoxGeom &b = c->boxes[k];
    std::memset(&b, 0, sizeof b);
    for (int d = 0; d < 3; d++) {
      b.plo[d] = geoms[k].prob_lo[d];
      b.phi[d] = geoms[k].prob_hi[d];
      b.L[d] = b.phi[d] - b.plo[d];
      b.halfL[d] = 0.5 * b.L[d];
      b.periodic[d] = geoms[k].periodic[d] ? 1 : 0;
      b.nbin[d] = 1;
      b.dxi[d] = 1.0;
    }
  }
  HIPCHK(c->d_boxes.alloc(ncol));
  HIPCHK(hipMemcpyAsync(c->d_boxes.p, c->boxes.data(), sizeof(BoxGeom) * ncol,
                        hipMemcpyHostToDevice, s));
  const int32_t *dcol = nullptr;
  if (ncol > 1) {
    std::vector<int32_t> hc((size_t)n);
    for (int k = 0; k < ncol; k++)
      for (int64_t i = cstart[k]; i < cstart[k + 1]; i++) hc[i] = k;
    HIPCHK(c->d_col_in.alloc(cap));
    HIPCHK(hipMemcpyAsync(c->d_col_in.p, hc.data(), sizeof(int32_t) * n,
                          hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    dcol = c->d_col_in.p;
  }
  c->cs = 0;
  c->cd = 0;
  HIPCHK(launch_fill_u64(c->d_ustats.p, 8, 0ull, s));
  StaticArrays st = static_arrays(c, 0);
  DynArrays dy = dyn_arrays(c, 0);
  HIPCHK(launch_aos_unpack(c->recs, c->lay, n, dcol, c->d_boxes.p, st, dy,
                           c->d_ustats.p, s));
  unsigned long long hs[8];
  HIPCHK(hipMemcpyAsync(hs, c->d_ustats.p, sizeof hs, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (hs[4] != 0)
    return fail(BMX_ERR_BAD_ARG,
                "particle records with negative radius, non-finite position, "
                "or (segments) n_bnds outside 0..4 or c_length <= 0");
  std::memcpy(&c->rmax, &hs[0], sizeof(double));
  std::memcpy(&c->lmax, &hs[1], sizeof(double));
  c->any_yeast = hs[2] != 0;
  c->any_fungi = hs[3] != 0;
  bmx_status bs = setup_boxes(c);
  if (bs != BMX_OK) return bs;
  bs = ensure_rows(c);
  if (bs != BMX_OK) return bs;
  HIPCHK(launch_fill_i32(c->d_err.p, 1, 0, s));
  bs = alloc_partials(c);
  if (bs != BMX_OK) return bs;
  c->row_nblk = 0;
  HIPCHK(hipStreamSynchronize(s));
  c->step = 0;
  c->last_build = -1;
  c->rebuilds_total = 0;
  c->guard_redone = 0;
  c->have_list = false;
  c->wrote_step = false;
  c->have_chem = false;
  c->n_real = L.n_real;
  c->n_int = L.n_int;
  c->bound = true;
  return BMX_OK;
}

}  // namespace

// ================================================================ C ABI ===

extern "C" {

int32_t bmx_abi_version(void) { return BMX_ABI_VERSION; }

void bmx_io_set_error(const char *msg) { g_last_error = msg ? msg : ""; }

const char *bmx_last_error(void) { return g_last_error.c_str(); }


bmx_status bmx_ctx_create(bmx_ctx **out, int32_t device) {
  if (!out) return fail(BMX_ERR_BAD_ARG, "null output pointer");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    return fail(BMX_ERR_NO_DEVICE, "no HIP device available");
  if (device < 0 || device >= count)
    return fail(BMX_ERR_BAD_ARG, "device ordinal out of range");
  HIPCHK(hipSetDevice(device));
  bmx_ctx *c = new (std::nothrow) bmx_ctx();
  if (!c) return fail(BMX_ERR_OOM, "context allocation failed");
  c->device = device;
  c->nbp.mode = BMX_NBR_VERLET;
  c->nbp.rebuild_every = 10;
  c->nbp.max_neighbors = 0;
  c->nbp.bin_factor_pct = 100;
  c->nbp.skin = 0.0;
  c->nbp.neighborhood = 1.0;
  hipError_t e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
  c->stream = c->own_stream;
  if (e == hipSuccess)
    e = hipEventCreateWithFlags(&c->done_ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipHostMalloc((void **)&c->h_flags, 128);
  if (e == hipSuccess) e = c->d_err.alloc(1) ;
  if (e == hipSuccess) e = c->d_stats.alloc(4);
  if (e == hipSuccess) e = c->d_total.alloc(1);
  if (e != hipSuccess) {
    bmx_ctx_destroy(c);
    return fail(BMX_ERR_HIP, std::string("context setup: ") +
                                 hipGetErrorString(e));
  }
  std::memset(c->h_flags, 0, 128);
  if (hipHostGetDevicePointer((void **)&c->h_flags_dev, c->h_flags, 0) != hipSuccess) {
    (void)hipGetLastError();
    c->h_flags_dev = nullptr;  // read back by copies instead
  }
  *out = c;
  return BMX_OK;
}

void bmx_ctx_destroy(bmx_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->own_stream) (void)hipStreamSynchronize(c->own_stream);
  reset_graph(c);
  for (auto e : c->ev_pool) (void)hipEventDestroy(e);
  for (int k = 0; k < 2; k++) {
    c->oid[k].release();
    c->type[k].release();
    c->col[k].release();
    c->key[k].release();
    c->bscale[k].release();
    c->pos[k].release();
    c->rod[k].release();
    c->trig[k].release();
    c->topo[k].release();
  }
  c->d_boxes.release();
  c->d_col_in.release();
  c->d_aos.release();
  c->d_keytmp.release();
  c->d_rank.release();
  c->d_count.release();
  c->d_start.release();
  c->d_scan.release();
  c->d_permk.release();
  c->d_nbr.release();
  c->d_ncnt.release();
  c->d_wtot.release();
  c->d_redo.release();
  c->d_cidx.release();
  c->d_cown.release();
  c->d_fc.release();
  c->d_err.release();
  c->d_stats.release();
  c->d_bstats.release();
  c->d_total.release();
  c->d_vpart.release();
  c->d_slotmax.release();
  c->d_ustats.release();
  for (auto &b : c->d_chemst) b.release();
  c->d_mkey.release();
  c->d_mrank.release();
  c->d_mcount.release();
  c->d_mstart.release();
  c->d_mscan.release();
  c->d_mperm.release();
  c->d_mout.release();
  c->d_min.release();
  for (auto &z : c->gsnap) {
    z.pos.release();
    z.rod.release();
    z.trig.release();
    z.topo.release();
    z.oid.release();
    z.type.release();
    z.key.release();
    z.col.release();
    z.bscale.release();
    if (z.done) (void)hipEventDestroy(z.done);
  }
  {
    bmx_ctx::AdjBufs &a = c->adj;
    if (a.rev) (void)hipStreamSynchronize(a.rev);
    for (DevBuf<double> *b : {&a.target, &a.gr, &a.acc, &a.lpart, &a.out}) b->release();
    for (DevBuf<double4> *b : {&a.ckpt, &a.lam0, &a.lam1, &a.lamo, &a.fin}) b->release();
    for (auto &h : a.h) {
      h.pos.release();
      h.idx.release();
      h.tot.release();
      h.cnt.release();
      h.oid.release();
      h.own.release();
    }
    for (hipEvent_t e : {a.rdone[0], a.rdone[1], a.adone[0], a.adone[1], a.join})
      if (e) (void)hipEventDestroy(e);
    if (a.rev) (void)hipStreamDestroy(a.rev);
  }
  if (c->h_flags) (void)hipHostFree(c->h_flags);
  if (c->done_ev) (void)hipEventDestroy(c->done_ev);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
}

bmx_status bmx_set_force_params(bmx_ctx *c, const double fpar[9]) {
  if (!c || !fpar) return fail(BMX_ERR_BAD_ARG, "null argument");
  for (int k = 0; k < 9; k++)
    if (!std::isfinite(fpar[k]))
      return fail(BMX_ERR_BAD_ARG, "force parameters must be finite");
  double old[9];
  std::memcpy(old, c->fpar, sizeof old);
  std::memcpy(c->fpar, fpar, sizeof c->fpar);
  reset_graph(c);
  if (c->bound) {
    (void)hipSetDevice(c->device);
    // the interaction radius depends on W_A; a rejected change keeps the
    // previous parameters and bins
    bmx_status st = setup_boxes(c);
    if (st != BMX_OK) {
      std::memcpy(c->fpar, old, sizeof old);
      return st;
    }
  }
  c->have_fpar = true;
  return BMX_OK;
}

bmx_status bmx_set_mobility(bmx_ctx *c, double m) {
  if (!c || !std::isfinite(m)) return fail(BMX_ERR_BAD_ARG, "bad mobility");
  c->mobility = m;
  reset_graph(c);
  return BMX_OK;
}

bmx_status bmx_set_geometry(bmx_ctx *c, const bmx_geometry *g) {
  if (!c || !g) return fail(BMX_ERR_BAD_ARG, "null argument");
  if (c->bound) return fail(BMX_ERR_STATE, "set geometry before binding");
  c->geom = *g;
  c->have_geom = true;
  return BMX_OK;
}

bmx_status bmx_set_neighbor_params(bmx_ctx *c, const bmx_neighbor_params *p) {
  if (!c || !p) return fail(BMX_ERR_BAD_ARG, "null argument");
  if (p->mode != BMX_NBR_VERLET && p->mode != BMX_NBR_STENCIL)
    return fail(BMX_ERR_BAD_ARG, "unknown neighbor mode");
  if (p->rebuild_every < 1 || p->max_neighbors < 0 ||
      p->max_neighbors > (1 << 20) || !(p->skin >= 0.0) ||
      p->bin_factor_pct < 0)
    return fail(BMX_ERR_BAD_ARG, "bad neighbor parameters");
  if (p->mode == BMX_NBR_STENCIL && p->rebuild_every != 1)
    return fail(BMX_ERR_BAD_ARG, "stencil lists are rebuilt every step");
  const bmx_neighbor_params old = c->nbp;
  const int32_t old_maxn = c->maxn;
  c->nbp = *p;
  if (c->nbp.bin_factor_pct == 0) c->nbp.bin_factor_pct = 100;
  c->maxn = p->max_neighbors > 0 ? p->max_neighbors : 96;
  reset_graph(c);
  if (c->bound) {
    (void)hipSetDevice(c->device);
    bmx_status st = setup_boxes(c);
    if (st != BMX_OK) {
      if (c->bound) {  // rejected before any change: keep the old policy
        c->nbp = old;
        c->maxn = old_maxn;
      }
      return st;
    }
    st = ensure_rows(c);
    if (st != BMX_OK) return st;
    st = alloc_partials(c);
    if (st != BMX_OK) return st;
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return BMX_OK;
}

bmx_status bmx_set_launch_params(bmx_ctx *c, const bmx_launch_params *p) {
  if (!c || !p) return fail(BMX_ERR_BAD_ARG, "null argument");
  int b = p->block_size;
  if (b != 64 && b != 128 && b != 256 && b != 512)
    return fail(BMX_ERR_BAD_ARG, "block_size must be 64, 128, 256 or 512");
  c->lp = *p;
  reset_graph(c);
  return BMX_OK;
}

bmx_status bmx_bind_particles_aos(bmx_ctx *c, const void *aos, int64_t n,
                                  const bmx_layout *layout,
                                  int32_t on_device) {
  if (!c) return fail(BMX_ERR_BAD_ARG, "null context");
  if (!c->have_geom) return fail(BMX_ERR_STATE, "geometry not set");
  if (on_device < BMX_AOS_HOST || on_device > BMX_AOS_IN_PLACE)
    return fail(BMX_ERR_BAD_ARG, "aos_on_device must be 0, 1 or 2");
  (void)hipSetDevice(c->device);
  if (on_device == BMX_AOS_IN_PLACE && n > 0) {
    // in place: the records must be device memory of this context's device
    hipPointerAttribute_t at;
    if (!aos || hipPointerGetAttributes(&at, aos) != hipSuccess ||
        at.type != hipMemoryTypeDevice || at.device != c->device) {
      (void)hipGetLastError();
      c->bound = false;
      return fail(BMX_ERR_BAD_ARG,
                  "records bound in place must be device memory of the "
                  "context's device");
    }
  }
  int64_t cs[2] = {0, n};
  return bind_common(c, aos, cs, 1, &c->geom, layout, on_device ? 1 : 0,
                     on_device == BMX_AOS_IN_PLACE);
}

bmx_status bmx_bind_tiles_aos(bmx_ctx *c, const void *const *tiles,
                              const int64_t *counts, int32_t ntiles,
                              const bmx_layout *layout, int32_t on_device) {
  if (!c) return fail(BMX_ERR_BAD_ARG, "null context");
  if (!c->have_geom) return fail(BMX_ERR_STATE, "geometry not set");
  if (ntiles < 1 || !tiles || !counts || (on_device != 0 && on_device != 1))
    return fail(BMX_ERR_BAD_ARG, "bad tile arguments");
  bmx_layout L = layout ? *layout : bmx_default_layout();
  if (L.stride_bytes <= 0) return fail(BMX_ERR_BAD_ARG, "invalid record layout");
  (void)hipSetDevice(c->device);
  std::vector<int64_t> start((size_t)ntiles + 1, 0);
  for (int32_t k = 0; k < ntiles; k++) {
    if (counts[k] < 0 || (counts[k] > 0 && !tiles[k]))
      return fail(BMX_ERR_BAD_ARG, "bad tile count or pointer");
    start[k + 1] = start[k] + counts[k];
  }
  const int64_t n = start[ntiles];
  if (n > (int64_t)INT32_MAX / 2)
    return fail(BMX_ERR_BAD_ARG, "particle count out of range");
  // the tiles, one after another, form the context's record array
  c->bound = false;
  const size_t stride = (size_t)L.stride_bytes;
  HIPCHK(c->d_aos.alloc((size_t)n * stride));
  for (int32_t k = 0; k < ntiles; k++)
    if (counts[k] > 0)
      HIPCHK(hipMemcpyAsync(c->d_aos.p + (size_t)start[k] * stride, tiles[k],
                            (size_t)counts[k] * stride,
                            on_device ? hipMemcpyDeviceToDevice
                                      : hipMemcpyHostToDevice,
                            c->stream));
  int64_t cs[2] = {0, n};
  bmx_status st = bind_common(c, c->d_aos.p, cs, 1, &c->geom, &L, 1, true);
  if (st != BMX_OK) return st;
  // bound over the context's own copy, synced back tile by tile
  c->in_place = false;
  c->aos_on_device = on_device != 0;
  c->user_aos = nullptr;
  c->tiles.assign(const_cast<void **>(tiles), const_cast<void **>(tiles) + ntiles);
  c->tile_start.swap(start);
  return BMX_OK;
}

bmx_status bmx_bind_colonies_aos(bmx_ctx *c, const void *aos,
                                 const int64_t *colony_start, int32_t ncol,
                                 const bmx_geometry *geoms,
                                 const bmx_layout *layout,
                                 int32_t on_device) {
  if (!c || !colony_start || !geoms || ncol < 1)
    return fail(BMX_ERR_BAD_ARG, "bad colony arguments");
  if (colony_start[0] != 0)
    return fail(BMX_ERR_BAD_ARG, "colony_start[0] must be 0");
  (void)hipSetDevice(c->device);
  c->geom = geoms[0];
  c->have_geom = true;
  return bind_common(c, aos, colony_start, ncol, geoms, layout, on_device);
}

int64_t bmx_num_particles(const bmx_ctx *c) { return c ? c->n : 0; }

bmx_status bmx_sync_to_aos(bmx_ctx *c, void *aos) {
  bmx_status st = require_ready(c);
  if (st != BMX_OK) return st;
  (void)hipSetDevice(c->device);
  st = pack_records(c);
  if (st != BMX_OK) return st;
  const hipMemcpyKind kind =
      c->aos_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  const size_t stride = (size_t)c->lay.stride;
  if (!aos && !c->tiles.empty()) {
    // a level bound as tiles: each tile gets its own records back
    for (size_t k = 0; k < c->tiles.size(); k++) {
      const size_t b = (size_t)(c->tile_start[k + 1] - c->tile_start[k]) * stride;
      if (b)
        HIPCHK(hipMemcpyAsync(c->tiles[k], c->recs + (size_t)c->tile_start[k] * stride,
                              b, kind, c->stream));
    }
  } else {
    void *dst = aos ? aos : c->user_aos;
    if (!dst)
      return fail(BMX_ERR_STATE,
                  "the particle count has changed since bind: pass a record "
                  "buffer with room for bmx_num_particles() records");
    size_t bytes = (size_t)c->n * stride;
    // bound in place, the records were packed where they live
    if (dst != c->recs && bytes)
      HIPCHK(hipMemcpyAsync(dst, c->recs, bytes, kind, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return BMX_OK;
}

bmx_status bmx_init_bonds(bmx_ctx *c) {
  bmx_status st = require_ready(c);
  if (st != BMX_OK) return st;
  (void)hipSetDevice(c->device);
  reset_graph(c);
  st = rebuild(c);
  if (st != BMX_OK) return st;
  if (c->any_fungi)
    HIPCHK(launch_init_bonds(dyn_arrays(c, c->cd), static_arrays(c, c->cs),
                             c->d_nbr.p, c->d_ncnt.p, c->ncap, c->n,
                             c->d_boxes.p, c->boxes[0], c->ncol > 1,
                             force_params(c, 0.0), c->d_err.p, c->stream));
  // the first evolve step rebuilds, as EvolveParticles does after InitBonds
  c->have_list = false;
  return check_flags(c);
}

}  // extern "C"
namespace {
// The steps of bmx_evolve, enqueued, with the closing skin check (unless
// close_check is false: the rebuild points still check) and the read-back
// of the error words; nothing waits for them.
bmx_status enqueue_evolve(bmx_ctx *c, double dt, int32_t nsteps,
                          bool close_check = true) {
  bmx_status st;
  for (int32_t k = 0; k < nsteps; k++) {
    st = enqueue_step(c, dt, k == nsteps - 1);
    if (st != BMX_OK) return st;
  }
  if (c->timing) {  // the run ends before the closing skin check
    st = close_run(c);
    if (st != BMX_OK) return st;
  }
  if (close_check && verlet_checked(c) && nsteps > 0 && c->have_list)
    HIPCHK(launch_skin_check(c->d_vpart.p, c->vmax_slots, c->vstride,
                             0.5 * c->nbp.skin, 0, c->d_slotmax.p, c->d_err.p,
                             c->stream));
  return enqueue_flags(c);
}
}  // namespace
extern "C" {

}  // extern "C"
namespace {

// Host bookkeeping of a guarded period's start.
struct GuardMark {
  int slot;
  int32_t steps;       // steps of the period in this call
  int32_t done0;       // steps of the call before it
  int64_t step, last_build;
  int cs, cd;
  bool have_list, wrote_step;
};

template <typename T>
hipError_t dev_copy(T *dst, const T *src, int64_t n, hipStream_t s) {
  return hipMemcpyAsync(dst, src, sizeof(T) * (size_t)n, hipMemcpyDeviceToDevice, s);
}

// Snapshot of the state a period starts from: the current dynamic and
// static arrays (the ones the next step reads).
bmx_status guard_save(bmx_ctx *c, int slot) {
  bmx_ctx::Snap &z = c->gsnap[slot];
  const int64_t cap = c->ncap;
  hipStream_t s = c->stream;
  HIPCHK(z.pos.alloc(cap));
  HIPCHK(z.oid.alloc(cap));
  HIPCHK(z.type.alloc(cap));
  HIPCHK(z.key.alloc(cap));
  HIPCHK(dev_copy(z.pos.p, c->pos[c->cd].p, cap, s));
  HIPCHK(dev_copy(z.oid.p, c->oid[c->cs].p, cap, s));
  HIPCHK(dev_copy(z.type.p, c->type[c->cs].p, cap, s));
  HIPCHK(dev_copy(z.key.p, c->key[c->cs].p, cap, s));
  if (c->ncol > 1) {
    HIPCHK(z.col.alloc(cap));
    HIPCHK(dev_copy(z.col.p, c->col[c->cs].p, cap, s));
  }
  if (c->any_fungi) {
    HIPCHK(z.rod.alloc(cap));
    HIPCHK(z.trig.alloc(cap));
    HIPCHK(z.topo.alloc(cap));
    HIPCHK(z.bscale.alloc(cap));
    HIPCHK(dev_copy(z.rod.p, c->rod[c->cd].p, cap, s));
    HIPCHK(dev_copy(z.trig.p, c->trig[c->cd].p, cap, s));
    HIPCHK(dev_copy(z.topo.p, c->topo[c->cd].p, cap, s));
    HIPCHK(dev_copy(z.bscale.p, c->bscale[c->cs].p, cap, s));
  }
  return BMX_OK;
}

bmx_status guard_restore(bmx_ctx *c, const GuardMark &m) {
  bmx_ctx::Snap &z = c->gsnap[m.slot];
  const int64_t cap = c->ncap;
  hipStream_t s = c->stream;
  HIPCHK(dev_copy(c->pos[m.cd].p, z.pos.p, cap, s));
  HIPCHK(dev_copy(c->oid[m.cs].p, z.oid.p, cap, s));
  HIPCHK(dev_copy(c->type[m.cs].p, z.type.p, cap, s));
  HIPCHK(dev_copy(c->key[m.cs].p, z.key.p, cap, s));
  if (c->ncol > 1) HIPCHK(dev_copy(c->col[m.cs].p, z.col.p, cap, s));
  if (c->any_fungi) {
    HIPCHK(dev_copy(c->rod[m.cd].p, z.rod.p, cap, s));
    HIPCHK(dev_copy(c->trig[m.cd].p, z.trig.p, cap, s));
    HIPCHK(dev_copy(c->topo[m.cd].p, z.topo.p, cap, s));
    HIPCHK(dev_copy(c->bscale[m.cs].p, z.bscale.p, cap, s));
  }
  c->step = m.step;
  c->last_build = m.last_build;
  c->cs = m.cs;
  c->cd = m.cd;
  c->have_list = m.have_list;
  c->wrote_step = m.wrote_step;
  return BMX_OK;
}

// bmx_evolve with the guard on: the call's steps in periods that end at
// the Verlet list's rebuild points (or at the end of the call).  Each
// period starts from a snapshot and ends with the skin check, its error
// word copied to its own host word; the next period is enqueued before the
// previous one's word is read, so the device does not wait on the host.  A
// period whose displacement bound exceeded skin/2 is restored from its
// snapshot and taken again with lists rebuilt before every step (a fresh
// list never misses a pair); the period after it is discarded and
// enqueued again.
bmx_status evolve_guarded(bmx_ctx *c, double dt, int32_t nsteps) {
  const int K = c->nbp.rebuild_every;
  hipStream_t s = c->stream;
  {
    // the snapshots hold the particle arrays only
    bmx_status sc = drop_chem(c);
    if (sc != BMX_OK) return sc;
  }
  for (auto &z : c->gsnap)
    if (!z.done) HIPCHK(hipEventCreateWithFlags(&z.done, hipEventDisableTiming));
  GuardMark q[2];
  int nq = 0, next_slot = 0;
  int32_t done = 0;
  bmx_status st = BMX_OK;
  c->guarding = true;
  auto finish = [&](bmx_status r) {
    c->guarding = false;
    return r;
  };
  while (done < nsteps || nq > 0) {
    if (done < nsteps && nq < 2) {
      // the period: up to the next rebuild point
      const bool due = !c->have_list || (c->step - c->last_build) >= K;
      const int32_t left = due ? K : (int32_t)(K - (c->step - c->last_build));
      const int32_t len = std::min<int32_t>(left, nsteps - done);
      GuardMark m{next_slot, len, done, c->step, c->last_build, c->cs, c->cd,
                  c->have_list, c->wrote_step};
      next_slot ^= 1;
      st = guard_save(c, m.slot);
      if (st != BMX_OK) return finish(st);
      for (int32_t k = 0; k < len; k++) {
        st = enqueue_step(c, dt, done + k == nsteps - 1);
        if (st != BMX_OK) return finish(st);
      }
      // the period's bound, cleared when the period is complete
      const bool ends = (c->step - c->last_build) >= K;
      HIPCHK(launch_skin_check(c->d_vpart.p, c->vmax_slots, c->vstride,
                               0.5 * c->nbp.skin, ends ? 1 : 0, c->d_slotmax.p,
                               c->d_err.p, s));
      HIPCHK(hipMemcpyAsync(c->h_flags + 8 + m.slot, c->d_err.p, sizeof(int32_t),
                            hipMemcpyDeviceToHost, s));
      HIPCHK(launch_fill_i32(c->d_err.p, 1, 0, s));
      HIPCHK(hipEventRecord(c->gsnap[m.slot].done, s));
      q[nq++] = m;
      done += len;
      continue;
    }
    // the oldest period in flight
    const GuardMark m = q[0];
    HIPCHK(hipEventSynchronize(c->gsnap[m.slot].done));
    const int32_t err = c->h_flags[8 + m.slot];
    if (err & ERRBIT_SKIN) {
      // drain the period after it, then take this one again
      HIPCHK(hipStreamSynchronize(s));
      st = guard_restore(c, m);
      if (st != BMX_OK) return finish(st);
      // nothing of the old period remains in the displacement record; each
      // step taken again writes slot 0 of its own fresh period, so the last
      // one's displacement starts the next period's record
      HIPCHK(launch_fill_u64(reinterpret_cast<unsigned long long *>(c->d_vpart.p),
                             (int64_t)c->vmax_slots * c->vstride, 0ull, s));
      for (int32_t k = 0; k < m.steps; k++) {
        c->have_list = false;  // a fresh list before every step
        st = enqueue_step(c, dt, m.done0 + k == nsteps - 1);
        if (st != BMX_OK) return finish(st);
      }
      c->guard_redone++;
      done = m.done0 + m.steps;
      nq = 0;
      continue;
    }
    if (err) {  // another error: reported by the flags below
      HIPCHK(hipStreamSynchronize(s));
      HIPCHK(launch_fill_i32(c->d_err.p, 1, err, s));
      nq = 0;
      break;
    }
    q[0] = q[1];
    nq--;
  }
  c->guarding = false;
  return enqueue_flags(c);
}

}  // namespace
extern "C" {

bmx_status bmx_set_guard(bmx_ctx *c, int32_t on) {
  if (!c) return fail(BMX_ERR_BAD_ARG, "null context");
  c->guard = on != 0;
  reset_graph(c);
  return BMX_OK;
}

bmx_status bmx_evolve(bmx_ctx *c, double dt, int32_t nsteps) {
  bmx_status st = require_ready(c);
  if (st != BMX_OK) return st;
  if (nsteps < 0 || !std::isfinite(dt))
    return fail(BMX_ERR_BAD_ARG, "bad dt or nsteps");
  (void)hipSetDevice(c->device);
  if (c->guard && verlet_checked(c) && nsteps > 0) {
    st = evolve_guarded(c, dt, nsteps);
    if (st != BMX_OK) return st;
    HIPCHK(hipStreamSynchronize(c->stream));
    return read_flags(c);
  }
  st = enqueue_evolve(c, dt, nsteps);
  if (st != BMX_OK) return st;
  HIPCHK(hipStreamSynchronize(c->stream));
  return read_flags(c);
}

bmx_status bmx_evolve_async(bmx_ctx *c, double dt, int32_t nsteps) {
  if (!c) return fail(BMX_ERR_BAD_ARG, "null context");
  // asynchronous calls chain: the error words of all of them add up in the
  // mapped host word until bmx_wait (with copied words, each waits for the
  // one before)
  bmx_status st = c->h_flags_dev ? BMX_OK : settle(c);
  if (st != BMX_OK) return st;
  if (!c->bound) return fail(BMX_ERR_STATE, "no particles bound");
  if (nsteps < 0 || !std::isfinite(dt))
    return fail(BMX_ERR_BAD_ARG, "bad dt or nsteps");
  if (c->guard && verlet_checked(c))
    return fail(BMX_ERR_STATE, "guarded steps are taken by bmx_evolve");
  (void)hipSetDevice(c->device);
  st = enqueue_evolve(c, dt, nsteps, false);
  if (st != BMX_OK) return st;
  HIPCHK(hipEventRecord(c->done_ev, c->stream));
  c->pending = true;
  return BMX_OK;
}

bmx_status bmx_wait(bmx_ctx *c) {
  if (!c) return fail(BMX_ERR_BAD_ARG, "null context");
  (void)hipSetDevice(c->device);
  return settle(c);
}

bmx_status bmx_query(bmx_ctx *c, int32_t *done) {
  if (!c || !done) return fail(BMX_ERR_BAD_ARG, "null argument");
  (void)hipSetDevice(c->device);
  *done = 1;
  if (!c->pending) return BMX_OK;
  hipError_t e = hipEventQuery(c->done_ev);
  if (e == hipErrorNotReady) {
    *done = 0;
    return BMX_OK;
  }
  HIPCHK(e);
  return BMX_OK;
}

bmx_status bmx_refresh_from_aos(bmx_ctx *c, const void *aos, uint32_t fields) {
  bmx_status st = require_ready(c);
  if (st != BMX_OK) return st;
  const uint32_t all = kFieldPositions | kFieldSizes | kFieldOrientations |
                       kFieldBonds | kFieldSpecies;
  if (fields & ~all) return fail(BMX_ERR_BAD_ARG, "unknown field bits");
  if (c->ncol != 1)
    return fail(BMX_ERR_STATE, "refresh needs a single-colony context");
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  const size_t stride = (size_t)c->lay.stride;
  const hipMemcpyKind kind =
      c->aos_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  // the caller's records into the record array the kernels read
  if (c->in_place) {
    if (aos && aos != c->recs)
      return fail(BMX_ERR_BAD_ARG, "records bound in place are re-read in place");
  } else if (!aos && !c->tiles.empty()) {
    for (size_t k = 0; k < c->tiles.size(); k++) {
      const size_t b = (size_t)(c->tile_start[k + 1] - c->tile_start[k]) * stride;
      if (b)
        HIPCHK(hipMemcpyAsync(c->recs + (size_t)c->tile_start[k] * stride,
                              c->tiles[k], b, kind, s));
    }
  } else {
    const void *src = aos ? aos : c->user_aos;
    if (!src)
      return fail(BMX_ERR_STATE, "no record array to re-read: pass one");
    if (c->n)
      HIPCHK(hipMemcpyAsync(c->recs, src, (size_t)c->n * stride, kind, s));
  }
  HIPCHK(launch_fill_u64(c->d_ustats.p, 8, 0ull, s));
  HIPCHK(launch_aos_refresh(c->recs, c->lay, c->n, c->d_boxes.p,
                            static_arrays(c, c->cs), dyn_arrays(c, c->cd),
                            fields, c->d_ustats.p, s));
  unsigned long long hs[8];
  HIPCHK(hipMemcpyAsync(hs, c->d_ustats.p, sizeof hs, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (hs[4] != 0) {
    c->bound = false;
    return fail(BMX_ERR_BAD_ARG,
                "re-read records with negative radius, non-finite position, "
                "or (segments) n_bnds outside 0..4 or c_length <= 0; "
                "the context is unbound");
  }
  // the chemistry state holds species and sizes: reloaded at next use
  if (fields & (kFieldSpecies | kFieldSizes)) c->have_chem = false;
  if (fields & (kFieldPositions | kFieldSizes)) {
    // the lists are stale: the next step rebuilds them, with bins for the
    // new reach when sizes grew or shrank
    double rmax, lmax;
    std::memcpy(&rmax, &hs[0], sizeof rmax);
    std::memcpy(&lmax, &hs[1], sizeof lmax);
    c->have_list = false;
    if (rmax != c->rmax || lmax != c->lmax) {
      c->rmax = rmax;
      c->lmax = lmax;
      st = setup_boxes(c);
      if (st != BMX_OK) {
        c->bound = false;
        return st;
      }
    }
  }
  reset_graph(c);
  return BMX_OK;
}

bmx_status bmx_set_stream(bmx_ctx *c, void *stream) {
  if (!c) return fail(BMX_ERR_BAD_ARG, "null context");
  (void)hipSetDevice(c->device);
  bmx_status st = settle(c);
  if (st != BMX_OK) return st;
  // work already enqueued on the old stream completes first
  HIPCHK(hipStreamSynchronize(c->stream));
  c->stream = stream ? static_cast<hipStream_t>(stream) : c->own_stream;
  return BMX_OK;
}

bmx_status bmx_evolve_graph(bmx_ctx *c, double dt, int32_t nsteps) {
  bmx_status st = require_ready(c);
  if (st != BMX_OK) return st;
  if (nsteps <= 0 || !std::isfinite(dt))
    return fail(BMX_ERR_BAD_ARG, "bad dt or nsteps");
  if (c->timing)
    return fail(BMX_ERR_STATE, "disable event timing for graph replay");
  if (c->guard && verlet_checked(c))  // periods may be taken again: eager
    return bmx_evolve(c, dt, nsteps);
  (void)hipSetDevice(c->device);
  // A captured sequence is valid for the same dt, step count, list phase and
  // buffer parity; rebuild points are part of the graph.
  int64_t phase = c->have_list ? (c->step - c->last_build) : -1;
  bool reuse = c->graph_exec && c->graph_dt == dt &&
               c->graph_nsteps == nsteps && c->graph_cs == c->cs &&
               c->graph_cd == c->cd && c->graph_step0 == phase &&
               c->graph_have_list == c->have_list;
  int cs0 = c->cs, cd0 = c->cd;
  bool hl0 = c->have_list;
  if (!reuse) {
    reset_graph(c);
    int64_t save_step = c->step, save_last = c->last_build;
    bool save_wrote = c->wrote_step;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    bmx_status cst = BMX_OK;
    for (int32_t k = 0; k < nsteps && cst == BMX_OK; k++)
      cst = enqueue_step(c, dt, k == nsteps - 1);
    if (cst == BMX_OK && verlet_checked(c))
      (void)launch_skin_check(c->d_vpart.p, c->vmax_slots, c->vstride,
                              0.5 * c->nbp.skin, 0, c->d_slotmax.p,
                              c->d_err.p, c->stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (cst != BMX_OK) {
      if (g) (void)hipGraphDestroy(g);
      return cst;
    }
    if (e != hipSuccess)
      return fail(BMX_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess)
      return fail(BMX_ERR_HIP,
                  std::string("instantiate: ") + hipGetErrorString(e));
    c->graph_dt = dt;
    c->graph_nsteps = nsteps;
    c->graph_cs = cs0;
    c->graph_cd = cd0;
    c->graph_step0 = phase;
    c->graph_have_list = hl0;
    // capture recorded the launches without running them: restore the
    // bookkeeping to the pre-capture state, then replay below
    c->step = save_step;
    c->last_build = save_last;
    c->wrote_step = save_wrote;
    c->cs = cs0;
    c->cd = cd0;
    c->have_list = hl0;
  }
  HIPCHK(hipGraphLaunch(c->graph_exec, c->stream));
  // advance bookkeeping exactly as the eager loop would
  for (int32_t k = 0; k < nsteps; k++) {
    bool need = !c->have_list ||
                (c->step - c->last_build) >= (int64_t)c->nbp.rebuild_every;
    if (need) {
      c->cs = 1 - c->cs;
      c->cd = 1 - c->cd;
      if (c->have_chem) c->cc = 1 - c->cc;
      c->have_list = true;
      c->last_build = c->step;
      c->rebuilds_total++;
    }
    c->cd = 1 - c->cd;
    c->step++;
  }
  c->wrote_step = true;
  return check_flags(c);
}


bmx_status bmx_adjoint(bmx_ctx *c, double dt, int32_t nsteps,
                       int32_t ckpt_every, const double *target, double *loss,
                       double grad3[3], double *grad_r,
                       bmx_adjoint_timing *timing) {
  bmx_status st = require_ready(c);
  if (st != BMX_OK) return st;
  if (c->any_fungi || c->ncol != 1)
    return fail(BMX_ERR_BAD_ARG, "the adjoint covers spheres in one box");
  const int K = c->nbp.rebuild_every;
  if (nsteps <= 0 || ckpt_every <= 0 || ckpt_every % K != 0 || !target ||
      !loss || !grad3 || !grad_r || !std::isfinite(dt))
    return fail(BMX_ERR_BAD_ARG,
                "bad adjoint arguments (checkpoint_every must be a multiple "
                "of rebuild_every)");
  (void)hipSetDevice(c->device);
  reset_graph(c);
  st = drop_chem(c);  // the checkpoints hold positions only
  if (st != BMX_OK) return st;
  hipStream_t s = c->stream;
  const int64_t n = c->n, cap = c->ncap;
  const int C = ckpt_every;
  const int S = (nsteps + C - 1) / C;
  const int R = (C + K - 1) / K;  // rebuilds per segment
  // rows of each rebuild of a segment, in the flat layout of the force pass
  const size_t rowbuf = (size_t)flat_entries(cap, c->flat_tile, c->maxn);
  const size_t tilebuf = (size_t)((cap + c->flat_tile - 1) / c->flat_tile + 1);
  bmx_ctx::AdjBufs &A = c->adj;
  if (!A.rev) {
    HIPCHK(hipStreamCreateWithFlags(&A.rev, hipStreamNonBlocking));
    for (hipEvent_t *e : {&A.rdone[0], &A.rdone[1], &A.adone[0], &A.adone[1], &A.join})
      HIPCHK(hipEventCreateWithFlags(e, hipEventDisableTiming));
  }
  hipStream_t s2 = A.rev;
  // work enqueued on the context stream before this call comes first
  HIPCHK(hipEventRecord(A.join, s));
  HIPCHK(hipStreamWaitEvent(s2, A.join, 0));
  HIPCHK(A.target.alloc(3 * (size_t)std::max<int64_t>(n, 1)));
  HIPCHK(A.gr.alloc((size_t)std::max<int64_t>(n, 1)));
  HIPCHK(A.ckpt.alloc((size_t)S * cap));
  for (auto &h : A.h) {
    HIPCHK(h.pos.alloc((size_t)C * cap));
    HIPCHK(h.idx.alloc((size_t)R * rowbuf));
    HIPCHK(h.own.alloc((size_t)R * rowbuf));
    HIPCHK(h.tot.alloc((size_t)R * tilebuf));
    HIPCHK(h.cnt.alloc((size_t)R * cap));
    HIPCHK(h.oid.alloc((size_t)R * cap));
  }
  HIPCHK(A.lam0.alloc(cap));
  HIPCHK(A.lam1.alloc(cap));
  HIPCHK(A.lamo.alloc(cap));
  HIPCHK(A.fin.alloc(cap));
  const int32_t nblk = adj_blocks(n);
  HIPCHK(A.acc.alloc(3 * (size_t)nblk));
  HIPCHK(A.lpart.alloc((size_t)nblk));
  HIPCHK(A.out.alloc(4));
  double *d_gr = A.gr.p, *d_acc = A.acc.p, *d_lpart = A.lpart.p, *d_out = A.out.p;
  double4 *d_ckpt = A.ckpt.p, *d_lamo = A.lamo.p, *d_final = A.fin.p;
  if (n)
    HIPCHK(hipMemcpyAsync(A.target.p, target, sizeof(double) * 3 * n,
                          hipMemcpyHostToDevice, s));
  HIPCHK(launch_fill_u64(reinterpret_cast<unsigned long long *>(d_gr), n, 0ull, s));
  HIPCHK(launch_fill_u64(reinterpret_cast<unsigned long long *>(d_acc),
                         3 * (int64_t)nblk, 0ull, s));
  hipEvent_t e0, e1, e2;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventCreate(&e2));
  HIPCHK(hipEventRecord(e0, s));
  const int64_t step0 = c->step;
  const ForceParams fp = force_params(c, dt);
  // ---- forward with checkpoints (rebuilds at every multiple of K)
  // The skin check runs on this pass only: the recomputed segments repeat
  // its arithmetic exactly.
  c->have_list = false;
  for (int t = 0; t < nsteps; t++) {
    if (t % C == 0)
      HIPCHK(launch_to_orig4(c->pos[c->cd].p, c->oid[c->cs].p, n,
                             d_ckpt + (size_t)(t / C) * cap, s));
    if (t % K == 0) {
      st = rebuild(c);
      if (st != BMX_OK) return st;
    }
    st = enqueue_force_only(c, dt, t % K);
    if (st != BMX_OK) return st;
  }
  if (verlet_checked(c))
    HIPCHK(launch_skin_check(c->d_vpart.p, c->vmax_slots, c->vstride,
                             0.5 * c->nbp.skin, 1, c->d_slotmax.p, c->d_err.p,
                             s));
  HIPCHK(launch_to_orig4(c->pos[c->cd].p, c->oid[c->cs].p, n, d_final, s));
  HIPCHK(launch_adj_loss(c->pos[c->cd].p, c->oid[c->cs].p, A.target.p, n,
                         c->boxes[0], A.lam0.p, d_lpart, s));
  HIPCHK(launch_to_orig4(A.lam0.p, c->oid[c->cs].p, n, d_lamo, s));
  HIPCHK(hipEventRecord(e1, s));
  {
    // a stale list (skin exceeded) or a non-finite force ends the call here
    HIPCHK(hipMemcpyAsync(c->h_flags + 20, c->d_err.p, sizeof(int32_t),
                          hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (c->h_flags[20]) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      (void)hipEventDestroy(e2);
      c->have_list = false;
      return check_flags(c);
    }
  }
  // ---- reverse.  Segment sg is recomputed on the context stream into
  // history set sg % 2 while the reverse steps of segment sg + 1 run on the
  // second stream from the other set; a set is written again only after
  // the reverse steps that read it (two segments later) are done.
  double4 *lam_in = A.lam0.p, *lam_out = A.lam1.p;
  for (int sg = S - 1; sg >= 0; sg--) {
    const int t0 = sg * C;
    const int len = std::min(C, nsteps - t0);
    const int hs = sg & 1;
    bmx_ctx::AdjHist &H = A.h[hs];
    if (sg + 2 <= S - 1) HIPCHK(hipStreamWaitEvent(s, A.adone[hs], 0));
    HIPCHK(launch_from_orig4(d_ckpt + (size_t)sg * cap, c->oid[c->cs].p, n,
                             c->pos[c->cd].p, s));
    c->have_list = false;
    for (int t = 0; t < len; t++) {
      // Each step writes the next state straight into the history; the
      // context's buffers carry the state only through a rebuild.
      if (t % K == 0) {
        if (t > 0)
          HIPCHK(hipMemcpyAsync(c->pos[c->cd].p, H.pos.p + (size_t)t * cap,
                                sizeof(double4) * cap, hipMemcpyDeviceToDevice, s));
        st = rebuild(c, false);
        if (st != BMX_OK) return st;
        HIPCHK(hipMemcpyAsync(H.pos.p + (size_t)t * cap, c->pos[c->cd].p,
                              sizeof(double4) * cap, hipMemcpyDeviceToDevice, s));
        const int r = t / K;
        HIPCHK(hipMemcpyAsync(H.idx.p + (size_t)r * rowbuf, c->d_cidx.p,
                              sizeof(int32_t) * rowbuf, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(H.own.p + (size_t)r * rowbuf, c->d_cown.p, rowbuf,
                              hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(H.tot.p + (size_t)r * tilebuf, c->d_wtot.p,
                              sizeof(int32_t) * tilebuf, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(H.cnt.p + (size_t)r * cap, c->d_ncnt.p,
                              sizeof(int32_t) * cap, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(H.oid.p + (size_t)r * cap, c->oid[c->cs].p,
                              sizeof(int32_t) * cap, hipMemcpyDeviceToDevice, s));
      }
      if (t + 1 < len)  // the last state of the segment is not needed
        HIPCHK(launch_force_sphere_flat(
            H.pos.p + (size_t)t * cap, H.pos.p + (size_t)(t + 1) * cap,
            c->d_cidx.p, c->d_cown.p, c->d_wtot.p, c->d_ncnt.p, c->n,
            c->maxn, nullptr, c->d_boxes.p, c->boxes[0], 0, fp, nullptr,
            nullptr, nullptr, c->d_err.p, shape(c), s));
    }
    HIPCHK(hipEventRecord(A.rdone[hs], s));
    HIPCHK(hipStreamWaitEvent(s2, A.rdone[hs], 0));
    // lambda of the segment's last step, in that step's order
    int r = (len - 1) / K;
    HIPCHK(launch_from_orig4(d_lamo, H.oid.p + (size_t)r * cap, n, lam_in, s2));
    for (int t = len - 1; t >= 0; t--) {
      r = t / K;
      const int32_t *oid = H.oid.p + (size_t)r * cap;
      HIPCHK(launch_adj_step(H.pos.p + (size_t)t * cap, H.idx.p + (size_t)r * rowbuf,
                             H.own.p + (size_t)r * rowbuf,
                             H.tot.p + (size_t)r * tilebuf,
                             H.cnt.p + (size_t)r * cap, c->maxn, oid, n,
                             c->boxes[0], fp, lam_in, lam_out, d_gr, d_acc, s2));
      std::swap(lam_in, lam_out);
      if (t % K == 0 && t > 0) {  // back across a rebuild: reorder lambda
        HIPCHK(launch_to_orig4(lam_in, oid, n, d_lamo, s2));
        HIPCHK(launch_from_orig4(d_lamo, H.oid.p + (size_t)(r - 1) * cap, n,
                                 lam_in, s2));
      }
    }
    HIPCHK(launch_to_orig4(lam_in, H.oid.p, n, d_lamo, s2));
    HIPCHK(hipEventRecord(A.adone[hs], s2));
  }
  // the reverse steps end before the sums
  HIPCHK(hipEventRecord(A.join, s2));
  HIPCHK(hipStreamWaitEvent(s, A.join, 0));
  HIPCHK(launch_sum_partials(d_acc, nblk, 3, d_out, s));
  HIPCHK(launch_sum_partials(d_lpart, nblk, 1, d_out + 3, s));
  HIPCHK(hipEventRecord(e2, s));
  // leave the context in the final forward state
  HIPCHK(launch_from_orig4(d_final, c->oid[c->cs].p, n, c->pos[c->cd].p, s));
  c->have_list = false;
  c->step = step0 + nsteps;
  double out[4];
  HIPCHK(hipMemcpyAsync(out, d_out, sizeof out, hipMemcpyDeviceToHost, s));
  if (n)
    HIPCHK(hipMemcpyAsync(grad_r, d_gr, sizeof(double) * n,
                          hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  float fms = 0.f, bms = 0.f;
  HIPCHK(hipEventElapsedTime(&fms, e0, e1));
  HIPCHK(hipEventElapsedTime(&bms, e1, e2));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  const double m = c->mobility;
  grad3[0] = out[0];
  grad3[1] = out[1];
  grad3[2] = -m * m * out[2];
  *loss = out[3] / (double)std::max<int64_t>(n, 1);
  if (timing) {
    timing->forward_ms = fms;
    timing->backward_ms = bms;
    timing->segments = S;
    timing->pad = 0;
  }
  c->wrote_step = false;
  return check_flags(c);
}

bmx_status bmx_get_fc(bmx_ctx *c, double *out) {
  bmx_status st = require_ready(c);
  if (st != BMX_OK) return st;
  if (!out && c->n > 0) return fail(BMX_ERR_BAD_ARG, "null output");
  if (!c->wrote_step) return fail(BMX_ERR_STATE, "no step has run");
  (void)hipSetDevice(c->device);
  DevBuf<double> tmp;
  HIPCHK(tmp.alloc(6 * (size_t)std::max<int64_t>(c->n, 1)));
  HIPCHK(launch_export_fc(c->d_fc.p, c->n, tmp.p, c->stream));
  if (c->n)
    HIPCHK(hipMemcpyAsync(out, tmp.p, sizeof(double) * 6 * c->n,
                          hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  tmp.release();
  return BMX_OK;
}

bmx_status bmx_get_positions(bmx_ctx *c, double *out) {
  bmx_status st = require_ready(c);
  if (st != BMX_OK) return st;
  if (!out && c->n > 0) return fail(BMX_ERR_BAD_ARG, "null output");
  (void)hipSetDevice(c->device);
  DevBuf<double> tmp;
  HIPCHK(tmp.alloc(3 * (size_t)std::max<int64_t>(c->n, 1)));
  HIPCHK(launch_export_pos(c->pos[c->cd].p, c->oid[c->cs].p, c->n, tmp.p,
                           c->stream));
  if (c->n)
    HIPCHK(hipMemcpyAsync(out, tmp.p, sizeof(double) * 3 * c->n,
                          hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  tmp.release();
  return BMX_OK;
}

bmx_status bmx_get_orientations(bmx_ctx *c, double *out) {
  bmx_status st = require_ready(c);
  if (st != BMX_OK) return st;
  if (!out && c->n > 0) return fail(BMX_ERR_BAD_ARG, "null output");
  (void)hipSetDevice(c->device);
  DevBuf<double> tmp;
  HIPCHK(tmp.alloc(2 * (size_t)std::max<int64_t>(c->n, 1)));
  HIPCHK(launch_export_orient(c->rod[c->cd].p, c->oid[c->cs].p, c->n, tmp.p,
                              c->stream));
  if (c->n)
    HIPCHK(hipMemcpyAsync(out, tmp.p, sizeof(double) * 2 * c->n,
                          hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  tmp.release();
  return BMX_OK;
}

bmx_status bmx_get_int_fields(bmx_ctx *c, int32_t *out) {
  bmx_status st = require_ready(c);
  if (st != BMX_OK) return st;
  if (!out && c->n > 0) return fail(BMX_ERR_BAD_ARG, "null output");
  (void)hipSetDevice(c->device);
  st = pack_records(c);
  if (st != BMX_OK) return st;
  std::vector<uint8_t> host((size_t)c->n * c->lay.stride);
  if (!host.empty())
    HIPCHK(hipMemcpyAsync(host.data(), c->recs, host.size(),
                          hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int64_t i = 0; i < c->n; i++)
    std::memcpy(out + 35 * i, host.data() + i * c->lay.stride + c->lay.idata,
                35 * sizeof(int32_t));
  return BMX_OK;
}

bmx_status bmx_get_neighbors(bmx_ctx *c, int64_t *offsets, int32_t *indices,
                             int64_t capacity, int64_t *total) {
  bmx_status st = require_ready(c);
  if (st != BMX_OK) return st;
  if (!offsets || !total) return fail(BMX_ERR_BAD_ARG, "null output");
  if (!c->have_list) return fail(BMX_ERR_STATE, "no current neighbor list");
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  DevBuf<int64_t> cnt;
  HIPCHK(cnt.alloc((size_t)c->n + 1));
  HIPCHK(launch_export_counts(c->d_ncnt.p, c->oid[c->cs].p, c->n, cnt.p, s));
  std::vector<int64_t> hc((size_t)c->n + 1, 0);
  if (c->n)
    HIPCHK(hipMemcpyAsync(hc.data(), cnt.p, sizeof(int64_t) * c->n,
                          hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  offsets[0] = 0;
  for (int64_t i = 0; i < c->n; i++) offsets[i + 1] = offsets[i] + hc[i];
  *total = offsets[c->n];
  if (indices) {
    if (capacity < *total) return fail(BMX_ERR_BAD_ARG, "capacity too small");
    DevBuf<int64_t> off;
    DevBuf<int32_t> out;
    HIPCHK(off.alloc((size_t)c->n + 1));
    HIPCHK(out.alloc((size_t)std::max<int64_t>(*total, 1)));
    HIPCHK(hipMemcpyAsync(off.p, offsets, sizeof(int64_t) * (c->n + 1),
                          hipMemcpyHostToDevice, s));
    HIPCHK(launch_export_nbr(flat_rows(c), c->maxn, c->d_ncnt.p,
                             c->oid[c->cs].p, c->n, off.p, out.p, s));
    if (*total)
      HIPCHK(hipMemcpyAsync(indices, out.p, sizeof(int32_t) * (*total),
                            hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    off.release();
    out.release();
  }
  cnt.release();
  return BMX_OK;
}

bmx_status bmx_get_bins(bmx_ctx *c, int32_t *out) {
  bmx_status st = require_ready(c);
  if (st != BMX_OK) return st;
  if (!out && c->n > 0) return fail(BMX_ERR_BAD_ARG, "null output");
  if (!c->have_list) return fail(BMX_ERR_STATE, "no current binning");
  (void)hipSetDevice(c->device);
  DevBuf<int32_t> tmp;
  HIPCHK(tmp.alloc((size_t)std::max<int64_t>(c->n, 1)));
  HIPCHK(launch_export_i32(c->key[c->cs].p, c->oid[c->cs].p, c->n, tmp.p,
                           c->stream));
  if (c->n)
    HIPCHK(hipMemcpyAsync(out, tmp.p, sizeof(int32_t) * c->n,
                          hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  tmp.release();
  return BMX_OK;
}

bmx_status bmx_get_bin_dims(bmx_ctx *c, int32_t colony, int32_t out3[3]) {
  bmx_status st = require_ready(c);
  if (st != BMX_OK) return st;
  if (!out3 || colony < 0 || colony >= c->ncol)
    return fail(BMX_ERR_BAD_ARG, "bad colony");
  for (int d = 0; d < 3; d++) out3[d] = c->boxes[colony].nbin[d];
  return BMX_OK;
}

bmx_status bmx_set_timing(bmx_ctx *c, int32_t enable) {
  if (!c) return fail(BMX_ERR_BAD_ARG, "null context");
  if (enable < 0) return fail(BMX_ERR_BAD_ARG, "negative timing switch");
  if (!enable && c->run_e0) {
    // an open run is dropped, not timed
    c->run_e0 = nullptr;
    c->run_n = 0;
  }
  c->timing = enable != 0;
  return BMX_OK;
}

bmx_status bmx_get_timing(bmx_ctx *c, bmx_timing *out) {
  if (!c || !out) return fail(BMX_ERR_BAD_ARG, "null argument");
  out->force_ms = c->force_ms;
  out->force_launches = c->force_launches;
  out->rebuild_ms = c->rebuild_ms;
  out->rebuilds = c->rebuilds;
  out->neighbor_entries = c->last_entries;
  out->max_row = c->last_max_row;
  out->nbins = c->nbins;
  out->rebuilds_total = c->rebuilds_total;
  out->steps_total = c->step;
  out->guard_redone = c->guard_redone;
  return BMX_OK;
}

bmx_status bmx_reset_timing(bmx_ctx *c) {
  if (!c) return fail(BMX_ERR_BAD_ARG, "null context");
  c->force_ms = c->rebuild_ms = 0.0;
  c->force_launches = c->rebuilds = 0;
  return BMX_OK;
}

bmx_status bmx_synchronize(bmx_ctx *c) {
  if (!c) return fail(BMX_ERR_BAD_ARG, "null context");
  (void)hipSetDevice(c->device);
  HIPCHK(hipStreamSynchronize(c->stream));
  return BMX_OK;
}

}  // extern "C"

// ----------------------------------------------------------------- exchange
extern "C" bmx_status bmx_exchange(bmx_ctx *c, const double xpar[3], double dt) {
  bmx_status st = require_ready(c);
  if (st != BMX_OK) return st;
  if (!xpar || !std::isfinite(dt) || !std::isfinite(xpar[0]) ||
      !std::isfinite(xpar[1]) || !std::isfinite(xpar[2]))
    return fail(BMX_ERR_BAD_ARG, "bad exchange arguments");
  if (c->n_real < rec::first_data + 6)
    return fail(BMX_ERR_BAD_ARG,
                "the record layout has no species fields (rdata[25..30])");
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  st = load_chem(c);
  if (st != BMX_OK) return st;
  // the rows are rebuilt at the current positions (xchng.cpp:68-77)
  st = rebuild(c, true, true);
  if (st != BMX_OK) return st;
  HIPCHK(launch_exchange(c->pos[c->cd].p, c->type[c->cs].p, c->topo[c->cd].p,
                         c->d_nbr.p, c->d_ncnt.p, c->ncap, c->n,
                         c->d_chemst[c->cc].p, xpar, dt, s));
  return check_flags(c);
}

// ---------------------------------------------------------------- autotuner
namespace {

// Copy of the particle arrays a step reads, so trials can start from the
// same state.
struct StateSnap {
  DevBuf<int32_t> oid, type, col, key;
  DevBuf<double> bscale;
  DevBuf<double4> pos, rod, trig;
  DevBuf<Topo> topo;
  int cs = 0, cd = 0;
  int64_t step = 0;
};

template <typename T>
hipError_t dcopy(T *dst, const T *src, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  return hipMemcpyAsync(dst, src, sizeof(T) * (size_t)n,
                        hipMemcpyDeviceToDevice, s);
}

bmx_status snap_take(bmx_ctx *c, StateSnap &z) {
  const int64_t cap = c->ncap;
  hipStream_t s = c->stream;
  HIPCHK(z.oid.alloc(cap));
  HIPCHK(z.type.alloc(cap));
  HIPCHK(z.key.alloc(cap));
  HIPCHK(z.bscale.alloc(cap));
  HIPCHK(z.pos.alloc(cap));
  HIPCHK(z.rod.alloc(cap));
  HIPCHK(z.trig.alloc(cap));
  HIPCHK(z.topo.alloc(cap));
  HIPCHK(dcopy(z.oid.p, c->oid[c->cs].p, cap, s));
  HIPCHK(dcopy(z.type.p, c->type[c->cs].p, cap, s));
  HIPCHK(dcopy(z.key.p, c->key[c->cs].p, cap, s));
  HIPCHK(dcopy(z.bscale.p, c->bscale[c->cs].p, cap, s));
  HIPCHK(dcopy(z.pos.p, c->pos[c->cd].p, cap, s));
  HIPCHK(dcopy(z.rod.p, c->rod[c->cd].p, cap, s));
  HIPCHK(dcopy(z.trig.p, c->trig[c->cd].p, cap, s));
  HIPCHK(dcopy(z.topo.p, c->topo[c->cd].p, cap, s));
  if (c->ncol > 1) {
    HIPCHK(z.col.alloc(cap));
    HIPCHK(dcopy(z.col.p, c->col[c->cs].p, cap, s));
  }
  z.cs = c->cs;
  z.cd = c->cd;
  z.step = c->step;
  return BMX_OK;
}

// Puts the saved state back; the next step rebuilds the lists.
bmx_status snap_restore(bmx_ctx *c, const StateSnap &z) {
  const int64_t cap = c->ncap;
  hipStream_t s = c->stream;
  HIPCHK(dcopy(c->oid[z.cs].p, z.oid.p, cap, s));
  HIPCHK(dcopy(c->type[z.cs].p, z.type.p, cap, s));
  HIPCHK(dcopy(c->key[z.cs].p, z.key.p, cap, s));
  HIPCHK(dcopy(c->bscale[z.cs].p, z.bscale.p, cap, s));
  HIPCHK(dcopy(c->pos[z.cd].p, z.pos.p, cap, s));
  HIPCHK(dcopy(c->rod[z.cd].p, z.rod.p, cap, s));
  HIPCHK(dcopy(c->trig[z.cd].p, z.trig.p, cap, s));
  HIPCHK(dcopy(c->topo[z.cd].p, z.topo.p, cap, s));
  if (c->ncol > 1) HIPCHK(dcopy(c->col[z.cs].p, z.col.p, cap, s));
  c->cs = z.cs;
  c->cd = z.cd;
  c->step = z.step;
  c->have_list = false;
  c->last_build = -1;
  c->wrote_step = false;
  HIPCHK(launch_fill_u64(reinterpret_cast<unsigned long long *>(c->d_vpart.p),
                         (int64_t)c->vmax_slots * c->vstride, 0ull, s));
  HIPCHK(launch_fill_u64(c->d_slotmax.p, c->vmax_slots, 0ull, s));
  return BMX_OK;
}

std::string tune_key(const bmx_ctx *c) {
  hipDeviceProp_t prop;
  std::string arch = "unknown";
  if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) {
    arch = prop.gcnArchName;
    arch = arch.substr(0, arch.find(':'));
  }
  double vol = 0.0;
  for (const bmx_geometry &g : c->colony_geom)
    vol += (g.prob_hi[0] - g.prob_lo[0]) * (g.prob_hi[1] - g.prob_lo[1]) *
           (g.prob_hi[2] - g.prob_lo[2]);
  // packing from the largest radius: enough to tell workloads apart
  double phi = vol > 0.0 ? (double)c->n * 4.18879020478639 * c->rmax *
                               c->rmax * c->rmax / vol
                         : 0.0;
  char buf[256];
  std::snprintf(buf, sizeof buf,
                "%s mode=%d K=%d ncol=%d n=%.2g phi=%.2g rods=%d", arch.c_str(),
                c->nbp.mode, c->nbp.rebuild_every, c->ncol, (double)c->n, phi,
                c->any_fungi ? 1 : 0);
  return buf;
}

bool json_int(const std::string &line, const char *name, int32_t *v) {
  std::string pat = std::string("\"") + name + "\":";
  size_t at = line.find(pat);
  return at != std::string::npos &&
         std::sscanf(line.c_str() + at + pat.size(), " %d", v) == 1;
}

// Last cache line for key; false when none.
bool read_cache(const char *path, const std::string &key, bmx_tuning *t) {
  FILE *f = std::fopen(path, "r");
  if (!f) return false;
  std::string want = "\"key\": \"" + key + "\"";
  bool found = false;
  char buf[1024];
  while (std::fgets(buf, sizeof buf, f)) {
    std::string line(buf);
    if (line.find(want) == std::string::npos) continue;
    bmx_tuning r{};
    if (json_int(line, "block_size", &r.block_size) &&
        json_int(line, "xcd_remap", &r.xcd_remap) &&
        json_int(line, "bin_factor_pct", &r.bin_factor_pct)) {
      *t = r;
      found = true;
    }
  }
  std::fclose(f);
  return found;
}

}  // namespace

extern "C" bmx_status bmx_autotune(bmx_ctx *c, double dt, int32_t trial_steps,
                                   const char *cache_path, bmx_tuning *out) {
  bmx_status st = require_ready(c);
  if (st != BMX_OK) return st;
  if (!out || !std::isfinite(dt) || trial_steps < 0)
    return fail(BMX_ERR_BAD_ARG, "bad autotune arguments");
  (void)hipSetDevice(c->device);
  std::memset(out, 0, sizeof *out);
  const std::string key = tune_key(c);
  bmx_neighbor_params nb = c->nbp;
  bmx_launch_params lp = c->lp;
  bmx_tuning hit{};
  if (cache_path && read_cache(cache_path, key, &hit)) {
    nb.bin_factor_pct = hit.bin_factor_pct;
    lp.block_size = hit.block_size;
    lp.xcd_remap = hit.xcd_remap;
    st = bmx_set_neighbor_params(c, &nb);
    if (st == BMX_OK) st = bmx_set_launch_params(c, &lp);
    if (st != BMX_OK) return st;
    *out = hit;
    out->from_cache = 1;
    return BMX_OK;
  }
  const bool was_timing = c->timing;
  c->timing = false;
  const int K = c->nbp.rebuild_every;
  const int steps = std::max(K, ((trial_steps + K - 1) / K) * K);
  st = drop_chem(c);  // the trials' snapshot holds the particle arrays only
  if (st != BMX_OK) return st;
  StateSnap z;
  st = snap_take(c, z);
  if (st != BMX_OK) return st;
  std::vector<int32_t> pcts;
  if (c->nbp.mode == BMX_NBR_VERLET)
    pcts = {100, 150, 200};
  else
    pcts = {c->nbp.bin_factor_pct};
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  bmx_tuning best{};
  best.us_per_step = 1.0e300;
  bmx_status err = BMX_OK;
  for (int32_t pct : pcts) {
    nb.bin_factor_pct = pct;
    st = bmx_set_neighbor_params(c, &nb);
    if (st == BMX_ERR_BAD_ARG) continue;  // too few bins at this edge
    if (st != BMX_OK) { err = st; break; }
    for (int32_t block : {64, 128, 256, 512}) {
      for (int32_t remap : {0, 1}) {
        lp.block_size = block;
        lp.xcd_remap = remap;
        st = bmx_set_launch_params(c, &lp);
        if (st == BMX_OK) st = snap_restore(c, z);
        if (st == BMX_OK) st = bmx_evolve_graph(c, dt, steps);  // capture
        float bestc = 1.0e30f;
        for (int rep = 0; rep < 3 && st == BMX_OK; rep++) {
          st = snap_restore(c, z);
          if (st != BMX_OK) break;
          if (hipEventRecord(e0, c->stream) != hipSuccess) { st = BMX_ERR_HIP; break; }
          st = bmx_evolve_graph(c, dt, steps);
          if (st != BMX_OK) break;
          float ms = 0.f;
          if (hipEventRecord(e1, c->stream) != hipSuccess ||
              hipEventSynchronize(e1) != hipSuccess ||
              hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            st = fail(BMX_ERR_HIP, "autotune timing failed");
            break;
          }
          bestc = std::min(bestc, ms);
        }
        if (st != BMX_OK) { err = st; break; }
        out->candidates++;
        double us = 1000.0 * bestc / steps;
        if (us < best.us_per_step) {
          best.us_per_step = us;
          best.block_size = block;
          best.xcd_remap = remap;
          best.bin_factor_pct = pct;
        }
      }
      if (err != BMX_OK) break;
    }
    if (err != BMX_OK) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->timing = was_timing;
  if (err != BMX_OK) return err;
  if (out->candidates == 0)
    return fail(BMX_ERR_BAD_ARG, "no launchable configuration");
  nb.bin_factor_pct = best.bin_factor_pct;
  lp.block_size = best.block_size;
  lp.xcd_remap = best.xcd_remap;
  st = bmx_set_neighbor_params(c, &nb);
  if (st == BMX_OK) st = bmx_set_launch_params(c, &lp);
  if (st == BMX_OK) st = snap_restore(c, z);
  if (st != BMX_OK) return st;
  HIPCHK(hipStreamSynchronize(c->stream));
  int32_t ncand = out->candidates;
  *out = best;
  out->candidates = ncand;
  out->from_cache = 0;
  if (cache_path) {
    FILE *f = std::fopen(cache_path, "a");
    if (!f) return fail(BMX_ERR_IO, std::string("cannot append to ") + cache_path);
    std::fprintf(f,
                 "{\"key\": \"%s\", \"block_size\": %d, \"xcd_remap\": %d, "
                 "\"bin_factor_pct\": %d, \"us_per_step\": %.3f}\n",
                 key.c_str(), best.block_size, best.xcd_remap,
                 best.bin_factor_pct, best.us_per_step);
    std::fclose(f);
  }
  return BMX_OK;
}

// ------------------------------------------------------ particle <-> mesh
namespace {

bmx_status mesh_geom(bmx_ctx *c, MeshGeom &g) {
  if (c->ncol != 1)
    return fail(BMX_ERR_STATE, "mesh coupling needs a single-colony context");
  const bmx_geometry &G = c->colony_geom[0];
  double dx[3];
  int64_t ng = 1;
  for (int d = 0; d < 3; d++) {
    if (G.n_cell[d] <= 0)
      return fail(BMX_ERR_BAD_ARG, "the geometry has no level mesh (n_cell)");
    g.plo[d] = G.prob_lo[d];
    dx[d] = (G.prob_hi[d] - G.prob_lo[d]) / (double)G.n_cell[d];
    g.dxi[d] = 1.0 / dx[d];
    g.n[d] = G.n_cell[d];
    g.periodic[d] = G.periodic[d] ? 1 : 0;
    ng *= (int64_t)G.n_cell[d] + 2;
  }
  if (ng > (int64_t)1 << 30)
    return fail(BMX_ERR_BAD_ARG, "level mesh too large (> 2^30 cells)");
  g.grid_vol = dx[0] * dx[1] * dx[2];
  g.inv_cell_vol = 1.0 / (dx[0] * dx[1] * dx[2]);
  return BMX_OK;
}

int64_t mesh_cells(const MeshGeom &g, int ghost) {
  return ((int64_t)g.n[0] + 2 * ghost) * (g.n[1] + 2 * ghost) *
         (g.n[2] + 2 * ghost);
}

// Bins the particles (current sorted arrays) over the level mesh.  The
// records are not needed: the positions are the sorted ones and the
// deposited fields (vol, rdata[31..33]) are only changed in the records.
bmx_status mesh_bin(bmx_ctx *c, const MeshGeom &g, int tri, int sorted) {
  (void)hipSetDevice(c->device);
  int64_t nb = 1;
  for (int d = 0; d < 3; d++)
    nb *= (tri && !g.periodic[d]) ? (int64_t)g.n[d] + 1 : g.n[d];
  size_t cap = (size_t)c->ncap;
  HIPCHK(c->d_mkey.alloc(cap));
  HIPCHK(c->d_mrank.alloc(cap));
  HIPCHK(c->d_mperm.alloc(cap));
  HIPCHK(c->d_mcount.alloc((size_t)nb + 1));
  HIPCHK(c->d_mstart.alloc((size_t)nb + 1));
  HIPCHK(c->d_mscan.alloc((size_t)scan_scratch_elems(nb) + 1));
  HIPCHK(launch_mesh_bin(c->pos[c->cd].p, c->oid[c->cs].p, c->n, g, tri,
                         c->d_mkey.p, c->d_mrank.p, c->d_mcount.p,
                         c->d_mstart.p, c->d_mscan.p, c->d_mperm.p, sorted,
                         c->d_err.p, c->stream));
  return BMX_OK;
}

// Reads and clears the error word after the mesh kernels.
bmx_status mesh_flags(bmx_ctx *c) {
  hipStream_t s = c->stream;
  HIPCHK(hipMemcpyAsync(c->h_flags, c->d_err.p, sizeof(int32_t),
                        hipMemcpyDeviceToHost, s));
  HIPCHK(launch_fill_i32(c->d_err.p, 1, 0, s));
  HIPCHK(hipStreamSynchronize(s));
  const int32_t err = c->h_flags[0];
  if (err & ERRBIT_MESH)
    return fail(BMX_ERR_MESH,
                "a particle lies outside the level mesh, or one cell holds "
                "more than 4096 particles");
  if (err & ERRBIT_CHEM)
    return fail(BMX_ERR_CHEM,
                "negative concentration in the particle chemistry, or growth "
                "left a negative radius or a segment length <= 0");
  if (err)
    return fail(BMX_ERR_STATE, "unexpected error bits before the mesh call");
  return BMX_OK;
}

// One deposited field (ncomp components) into out (host or device).
bmx_status mesh_deposit(bmx_ctx *c, int scheme, int comp0, int ncomp, int mode,
                        double *out, int32_t on_device) {
  bmx_status st = require_ready(c);
  if (st != BMX_OK) return st;
  if (!out) return fail(BMX_ERR_BAD_ARG, "null output mesh");
  if (scheme != BMX_DEPOSIT_ONE_TO_ONE && scheme != BMX_DEPOSIT_TRILINEAR)
    return fail(BMX_ERR_BAD_ARG, "unknown deposition scheme");
  MeshGeom g;
  st = mesh_geom(c, g);
  if (st != BMX_OK) return st;
  const int tri = scheme == BMX_DEPOSIT_TRILINEAR ? 1 : 0;
  st = mesh_bin(c, g, tri, mode != 2);
  if (st != BMX_OK) return st;
  const int64_t nc = mesh_cells(g, 0);
  double *dst = out;
  if (!on_device) {
    HIPCHK(c->d_mout.alloc((size_t)(nc * ncomp)));
    dst = c->d_mout.p;
  }
  // the volume and the amounts taken from the fluid are read from the
  // chemistry state while it holds them
  const uint8_t *src = c->recs;
  AosLayout lay = c->lay;
  int comp = comp0, by_sorted = 0;
  if (c->have_chem && (comp0 == rec::vol || comp0 == rec::first_data + 6)) {
    src = reinterpret_cast<const uint8_t *>(c->d_chemst[c->cc].p);
    lay.stride = chem::stride * (int32_t)sizeof(double);
    lay.rdata = 0;
    comp = comp0 == rec::vol ? chem::vol : chem::rel0;
    by_sorted = 1;
  }
  HIPCHK(launch_deposit(src, lay, g, tri, c->d_mstart.p, c->d_mperm.p,
                        c->pos[c->cd].p, comp, ncomp, mode, dst, c->stream,
                        by_sorted));
  if (!on_device)
    HIPCHK(hipMemcpyAsync(out, dst, sizeof(double) * (size_t)(nc * ncomp),
                          hipMemcpyDeviceToHost, c->stream));
  return mesh_flags(c);
}

}  // namespace

extern "C" bmx_status bmx_calc_volume_fraction(bmx_ctx *c, int32_t scheme,
                                               double *vf, int32_t on_device) {
  return mesh_deposit(c, scheme, rec::vol, 1, 1, vf, on_device);
}

extern "C" bmx_status bmx_calc_txfr_fluid(bmx_ctx *c, int32_t scheme,
                                          double *x_rhs, int32_t on_device) {
  if (c && c->bound && c->n_real < rec::first_data + 9)
    return fail(BMX_ERR_BAD_ARG,
                "the record layout has no fluid increments (rdata[31..33])");
  return mesh_deposit(c, scheme, rec::first_data + 6, 3, 0, x_rhs, on_device);
}

extern "C" bmx_status bmx_count_particles_per_cell(bmx_ctx *c, double *npart,
                                                   int32_t on_device) {
  return mesh_deposit(c, BMX_DEPOSIT_ONE_TO_ONE, 0, 1, 2, npart, on_device);
}

extern "C" bmx_status bmx_calc_txfr_particle(bmx_ctx *c, const double chempar[10],
                                             const double *x_k, const double *vf,
                                             double dt, int32_t nloop,
                                             int32_t cnc_scheme,
                                             int32_t vf_scheme,
                                             int32_t on_device) {
  bmx_status st = require_ready(c);
  if (st != BMX_OK) return st;
  if (!chempar || !x_k || !vf) return fail(BMX_ERR_BAD_ARG, "null argument");
  if (nloop < 1 || !std::isfinite(dt))
    return fail(BMX_ERR_BAD_ARG, "need nloop >= 1 and a finite dt");
  for (int k = 0; k < 10; k++)
    if (!std::isfinite(chempar[k]))
      return fail(BMX_ERR_BAD_ARG, "chemistry parameters must be finite");
  for (int32_t sc : {cnc_scheme, vf_scheme})
    if (sc != BMX_DEPOSIT_ONE_TO_ONE && sc != BMX_DEPOSIT_TRILINEAR)
      return fail(BMX_ERR_BAD_ARG, "unknown interpolation scheme");
  if (c->n_real < rec::first_data + 9)
    return fail(BMX_ERR_BAD_ARG,
                "the record layout has no species fields (rdata[25..33])");
  MeshGeom g;
  st = mesh_geom(c, g);
  if (st != BMX_OK) return st;
  st = mesh_bin(c, g, 0, 0);  // particles per cell
  if (st != BMX_OK) return st;
  hipStream_t s = c->stream;
  const int64_t ng = mesh_cells(g, 1);
  const double *X = x_k, *V = vf;
  if (!on_device) {
    HIPCHK(c->d_min.alloc((size_t)(4 * ng)));
    HIPCHK(hipMemcpyAsync(c->d_min.p, x_k, sizeof(double) * (size_t)(3 * ng),
                          hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->d_min.p + 3 * ng, vf, sizeof(double) * (size_t)ng,
                          hipMemcpyHostToDevice, s));
    X = c->d_min.p;
    V = c->d_min.p + 3 * ng;
  }
  // the chemistry updates the chemistry state (species, volume, area,
  // rates, gradient, amounts taken), the sorted sizes and their maxima; the
  // records get it when they are packed, and the rows are rebuilt at the
  // next step for the new reaches
  st = load_chem(c);
  if (st != BMX_OK) return st;
  HIPCHK(launch_fill_u64(c->d_ustats.p, 8, 0ull, s));
  HIPCHK(launch_mesh_to_particle(c->d_chemst[c->cc].p, c->n, g, c->d_mstart.p,
                                 c->type[c->cs].p,
                                 c->any_fungi ? c->topo[c->cd].p : nullptr,
                                 c->pos[c->cd].p, c->rod[c->cd].p, X, V,
                                 cnc_scheme, vf_scheme, chempar, dt, nloop,
                                 c->d_ustats.p, c->d_err.p, s));
  HIPCHK(hipMemcpyAsync(c->h_flags + 16, c->d_ustats.p, 2 * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, s));
  st = mesh_flags(c);  // synchronizes the stream
  if (st == BMX_ERR_HIP) return st;
  // the sorted sizes have changed either way: new bins, rows rebuilt
  std::memcpy(&c->rmax, c->h_flags + 16, sizeof(double));
  std::memcpy(&c->lmax, c->h_flags + 18, sizeof(double));
  reset_graph(c);
  bmx_status sb = setup_boxes(c);
  return st != BMX_OK ? st : sb;
}

// ------------------------------------------------------ division, branching
namespace {

// Binds the device records src[0, n) again (a copy is taken first): new
// sorted arrays and bins, the rows rebuilt at the next step.  The caller's
// record buffer and its host/device kind are kept.
bmx_status rebind_records(bmx_ctx *c, const uint8_t *src, int64_t n) {
  const size_t bytes = (size_t)n * (size_t)c->lay.stride;
  DevBuf<uint8_t> copy;
  Releaser<uint8_t> rc{copy};
  HIPCHK(copy.alloc(bytes));
  if (bytes)
    HIPCHK(hipMemcpyAsync(copy.p, src, bytes, hipMemcpyDeviceToDevice, c->stream));
  bmx_layout L;
  L.stride_bytes = c->lay.stride;
  L.pos_offset = c->lay.pos;
  L.rdata_offset = c->lay.rdata;
  L.id_offset = c->lay.id;
  L.cpu_offset = c->lay.cpu;
  L.idata_offset = c->lay.idata;
  L.n_real = c->n_real;
  L.n_int = c->n_int;
  const bool was_on_device = c->aos_on_device;
  void *user = c->user_aos;
  std::vector<void *> tiles;
  std::vector<int64_t> tile_start;
  tiles.swap(c->tiles);
  tile_start.swap(c->tile_start);
  const bmx_geometry g0 = c->colony_geom[0];
  int64_t cs[2] = {0, n};
  bmx_status st = bind_common(c, copy.p, cs, 1, &g0, &L, 1);
  c->aos_on_device = was_on_device;
  c->user_aos = user;
  if (!tile_start.empty() && tile_start.back() == n) {
    // same particles: the tiles still hold them
    c->tiles.swap(tiles);
    c->tile_start.swap(tile_start);
  }
  return st;
}

}  // namespace

extern "C" bmx_status bmx_split_particles(bmx_ctx *c, const bmx_split_params *sp,
                                          uint64_t seed, uint32_t call,
                                          int32_t next_id, int32_t cpu,
                                          int64_t *n_new) {
  bmx_status st = require_ready(c);
  if (st != BMX_OK) return st;
  if (!sp || !n_new) return fail(BMX_ERR_BAD_ARG, "null argument");
  if (c->ncol != 1)
    return fail(BMX_ERR_STATE, "splitting needs a single-colony context");
  if (c->n_real < rec::first_data + 9 || c->n_int < rec::cpu + 1)
    return fail(BMX_ERR_BAD_ARG,
                "the record layout lacks the species or topology fields");
  const double p[7] = {sp->max_vol,    sp->max_seg_length, sp->max_seg_radius,
                       sp->branching_probability, sp->splitting_probability,
                       sp->overlap,    sp->seg_split_length};
  for (double v : p)
    if (!std::isfinite(v)) return fail(BMX_ERR_BAD_ARG, "split parameters must be finite");
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  // the records take the device state first (positions, orientations,
  // bonds, species)
  st = pack_records(c);
  if (st != BMX_OK) return st;
  const int64_t n = c->n;
  DevBuf<int32_t> cnt, start, scratch;
  Releaser<int32_t> r1{cnt}, r2{start}, r3{scratch};
  HIPCHK(cnt.alloc((size_t)n + 1));
  HIPCHK(start.alloc((size_t)n + 1));
  HIPCHK(scratch.alloc((size_t)scan_scratch_elems(n) + 1));
  HIPCHK(launch_split_count(c->recs, c->lay, n, p, seed, call, cnt.p, s));
  HIPCHK(launch_scan(cnt.p, start.p, n, scratch.p, s));
  int32_t total = 0;
  HIPCHK(hipMemcpyAsync(&total, start.p + n, sizeof total,
                        hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  *n_new = total;
  if (total > 0) {
    if ((int64_t)next_id + total > (int64_t)INT32_MAX)
      return fail(BMX_ERR_BAD_ARG, "particle ids would overflow int32");
    const size_t stride = (size_t)c->lay.stride;
    DevBuf<uint8_t> grown;
    Releaser<uint8_t> r4{grown};
    HIPCHK(grown.alloc((size_t)(n + total) * stride));
    HIPCHK(hipMemcpyAsync(grown.p, c->recs, (size_t)n * stride,
                          hipMemcpyDeviceToDevice, s));
    HIPCHK(launch_split_apply(grown.p, c->lay, n, start.p, p, seed, call,
                              next_id, cpu, s));
    HIPCHK(hipStreamSynchronize(s));
    // new sorted arrays, rows rebuilt at the next step; the caller's buffer
    // no longer fits the records
    st = rebind_records(c, grown.p, n + total);
    c->user_aos = nullptr;
    if (st != BMX_OK) return st;
  }
  return BMX_OK;
}

// ------------------------------------------------------------------ fusion
extern "C" bmx_status bmx_fusion(bmx_ctx *c, const bmx_fusion_params *fp,
                                 uint64_t seed, uint32_t call, int32_t next_id,
                                 int32_t cpu, int32_t *fused, int64_t *n_new) {
  bmx_status st = require_ready(c);
  if (st != BMX_OK) return st;
  if (!fp || !fused || !n_new) return fail(BMX_ERR_BAD_ARG, "null argument");
  if (c->ncol != 1)
    return fail(BMX_ERR_STATE, "fusion needs a single-colony context");
  if (c->n_real < rec::first_data + 9 || c->n_int < rec::cpu + 1)
    return fail(BMX_ERR_BAD_ARG,
                "the record layout lacks the species or topology fields");
  const double fus[4] = {fp->fusion_probability, fp->max_fusion_separation,
                         fp->bond_scaling_increment, fp->max_seg_length};
  for (double v : fus)
    if (!std::isfinite(v)) return fail(BMX_ERR_BAD_ARG, "fusion parameters must be finite");
  *fused = 0;
  *n_new = 0;
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  reset_graph(c);
  // records and rows at the current positions (EvaluateTipFusion rebuilds)
  st = pack_records(c);
  if (st != BMX_OK) return st;
  st = rebuild(c, true, true);
  if (st != BMX_OK) return st;
  st = check_flags(c);
  if (st != BMX_OK) return st;
  const BoxGeom g0 = c->boxes[0];
  const double wa = c->fpar[0];
  const int64_t n = c->n;
  DevBuf<int32_t> flag, tgt;
  DevBuf<double> ttau;
  Releaser<int32_t> rf{flag}, rt{tgt};
  Releaser<double> rtt{ttau};
  HIPCHK(flag.alloc(1));
  HIPCHK(tgt.alloc((size_t)n + 1));
  HIPCHK(ttau.alloc((size_t)n + 1));
  HIPCHK(launch_fill_i32(flag.p, 1, 0, s));
  HIPCHK(launch_tip_fusion(c->recs, c->lay, n, c->pos[c->cd].p,
                           c->oid[c->cs].p, c->d_nbr.p, c->d_ncnt.p, c->ncap, g0,
                           wa, fus, seed, call, tgt.p, ttau.p, flag.p, s));
  int32_t any = 0;
  HIPCHK(hipMemcpyAsync(&any, flag.p, sizeof any, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (!any) return BMX_OK;  // no record changed
  *fused = 1;
  // interior fusion over the same rows, decisions then updates
  DevBuf<int32_t> cnt, one, start, scratch;
  DevBuf<int4> dec;
  DevBuf<double> dtau;
  Releaser<int32_t> r1{cnt}, r2{one}, r3{start}, r4{scratch};
  Releaser<int4> r5{dec};
  Releaser<double> r6{dtau};
  HIPCHK(cnt.alloc((size_t)n + 1));
  HIPCHK(one.alloc((size_t)n + 1));
  HIPCHK(start.alloc((size_t)n + 1));
  HIPCHK(scratch.alloc((size_t)scan_scratch_elems(n) + 1));
  HIPCHK(dec.alloc((size_t)n));
  HIPCHK(dtau.alloc((size_t)n));
  HIPCHK(launch_interior_fusion(c->recs, c->lay, n, c->pos[c->cd].p,
                                c->oid[c->cs].p, c->d_nbr.p, c->d_ncnt.p, c->ncap,
                                g0, wa, cnt.p, dec.p, dtau.p, one.p, s));
  HIPCHK(launch_scan(one.p, start.p, n, scratch.p, s));
  int32_t total = 0;
  HIPCHK(hipMemcpyAsync(&total, start.p + n, sizeof total, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if ((int64_t)next_id + total > (int64_t)INT32_MAX)
    return fail(BMX_ERR_BAD_ARG, "particle ids would overflow int32");
  const size_t stride = (size_t)c->lay.stride;
  DevBuf<uint8_t> grown;
  Releaser<uint8_t> r7{grown};
  HIPCHK(grown.alloc((size_t)(n + total) * stride));
  HIPCHK(hipMemcpyAsync(grown.p, c->recs, (size_t)n * stride,
                        hipMemcpyDeviceToDevice, s));
  HIPCHK(launch_fusion_split(grown.p, c->lay, n, cnt.p, start.p, next_id, cpu, s));
  HIPCHK(hipStreamSynchronize(s));
  *n_new = total;
  st = rebind_records(c, grown.p, n + total);
  if (total > 0) c->user_aos = nullptr;
  if (st != BMX_OK) return st;
  grown.release();
  // cleanup over rows rebuilt with the new segments
  st = rebuild(c, true, true);
  if (st != BMX_OK) return st;
  st = check_flags(c);
  if (st != BMX_OK) return st;
  HIPCHK(launch_fusion_cleanup(c->recs, c->lay, c->n, c->pos[c->cd].p,
                               c->oid[c->cs].p, c->d_nbr.p, c->d_ncnt.p, c->ncap,
                               c->boxes[0], fus[2], 1.5 * fus[3], s));
  HIPCHK(hipStreamSynchronize(s));
  void *user = c->user_aos;
  st = rebind_records(c, c->recs, c->n);
  c->user_aos = user;
  return st;
}

// ------------------------------------------------- allocation-free evolve
// bmx_evolve_aos_noalloc: bind (device records, in place) + evolve + sync
// for callers that cannot allocate device memory.  Every working array of
// the call lives in `scratch`, a file-scope device array of fixed size that
// each GPU holds from library load to process end.  In host code `scratch`
// names a host-side stand-in, not device memory: the host uses only the
// address hipGetSymbolAddress returns for the current device, kept per
// device.  A per-device context whose buffers are fixed slices of that
// memory runs the same binding, step and packing code as bmx_evolve; its
// buffers never grow, so a call that would need more than a slice fails
// with BMX_ERR_BAD_ARG instead of allocating.  All calls on one GPU share
// its scratch and context: calls that overlap on one GPU race.
namespace {

constexpr int64_t kNaParticles = BMX_NOALLOC_MAX_PARTICLES;
constexpr int32_t kNaNeighbors = BMX_NOALLOC_MAX_NEIGHBORS;
constexpr int64_t kNaBins = BMX_NOALLOC_MAX_BINS;
constexpr int32_t kNaRebuild = BMX_NOALLOC_MAX_REBUILD_EVERY;
constexpr int kNaDevices = 64;

constexpr size_t na_align(size_t b) { return (b + 255) / 256 * 256; }

// Bytes of each slice of the scratch, in carve order.
constexpr size_t kNaTiles = (size_t)(kNaParticles + 31) / 32 + 1;  // tiles of 32
constexpr size_t kNaFlat = (size_t)kNaTiles * 32 * kNaNeighbors + 2 * 64;
constexpr size_t kNaBytes =
    2 * 3 * na_align(4 * kNaParticles)           // oid, type, key
    + 2 * na_align(8 * kNaParticles)             // bscale
    + 2 * 3 * na_align(32 * kNaParticles)        // pos, rod, trig
    + 2 * na_align(sizeof(Topo) * kNaParticles)  // topo
    + 2 * na_align(4 * kNaParticles)             // keytmp, rank
    + na_align(8 * kNaParticles)                 // permk
    + na_align(8 * 6 * kNaParticles)             // fc
    + na_align(4 * kNaNeighbors * kNaParticles)  // nbr (column-major rows)
    + na_align(4 * kNaParticles)                 // ncnt
    + na_align(4 * kNaTiles) + na_align(kNaTiles)  // wtot, redo
    + na_align(4 * kNaFlat) + na_align(kNaFlat)  // cidx, cown
    + 2 * na_align(4 * (kNaBins + 1))            // count, start
    + na_align(4 * ((kNaBins + 2047) / 2048 + 2))  // scan tile sums
    + na_align(8 * (kNaParticles / 64 + 1) * kNaRebuild)  // displacement slots
    + na_align(8 * kNaRebuild)                   // slot maxima
    + na_align(4 * 2 * (kNaParticles / 64 + 1))  // row-build block stats
    + na_align(4) + na_align(16) + na_align(8) + na_align(64)  // err, stats, total, ustats
    + na_align(sizeof(BoxGeom));

static __device__ __align__(256) uint8_t scratch[kNaBytes];

struct NaDevice {
  bool ready = false;
  uint8_t *ptr = nullptr;  // this device's copy of scratch
  bmx_ctx ctx;
  int32_t flags[16];       // host words the error/statistics reads land in
};
NaDevice g_na[kNaDevices];

template <typename T>
void na_slice(DevBuf<T> &b, uint8_t *&cur, size_t count) {
  b.p = reinterpret_cast<T *>(cur);
  b.n = count;
  b.fixed = true;
  cur += na_align(sizeof(T) * count);
}

// Once per device, with the device current: the scratch address and the
// context's fixed buffers over it.
bmx_status na_setup(NaDevice &d, int dev) {
  uint8_t *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(BMX_ERR_HIP, std::string("scratch address: ") + hipGetErrorString(e));
  bmx_ctx &c = d.ctx;
  c.device = dev;
  uint8_t *cur = ptr;
  const size_t N = (size_t)kNaParticles;
  for (int k = 0; k < 2; k++) {
    na_slice(c.oid[k], cur, N);
    na_slice(c.type[k], cur, N);
    na_slice(c.key[k], cur, N);
    na_slice(c.bscale[k], cur, N);
    na_slice(c.pos[k], cur, N);
    na_slice(c.rod[k], cur, N);
    na_slice(c.trig[k], cur, N);
    na_slice(c.topo[k], cur, N);
    c.col[k].fixed = true;  // one box: never used
  }
  na_slice(c.d_keytmp, cur, N);
  na_slice(c.d_rank, cur, N);
  na_slice(c.d_permk, cur, N);
  na_slice(c.d_fc, cur, 6 * N);
  na_slice(c.d_nbr, cur, (size_t)kNaNeighbors * N);
  na_slice(c.d_ncnt, cur, N);
  na_slice(c.d_wtot, cur, kNaTiles);
  na_slice(c.d_redo, cur, kNaTiles);
  na_slice(c.d_cidx, cur, kNaFlat);
  na_slice(c.d_cown, cur, kNaFlat);
  na_slice(c.d_count, cur, (size_t)kNaBins + 1);
  na_slice(c.d_start, cur, (size_t)kNaBins + 1);
  na_slice(c.d_scan, cur, (size_t)(kNaBins + 2047) / 2048 + 2);
  na_slice(c.d_vpart, cur, (N / 64 + 1) * (size_t)kNaRebuild);
  na_slice(c.d_slotmax, cur, (size_t)kNaRebuild);
  na_slice(c.d_bstats, cur, 2 * (N / 64 + 1));
  na_slice(c.d_err, cur, 1);
  na_slice(c.d_stats, cur, 4);
  na_slice(c.d_total, cur, 1);
  na_slice(c.d_ustats, cur, 8);
  na_slice(c.d_boxes, cur, 1);
  if ((size_t)(cur - ptr) > kNaBytes)
    return fail(BMX_ERR_HIP, "scratch layout exceeds the scratch array");
  // the record array is the caller's; a colony array is never needed
  c.d_aos.fixed = true;  // never used: the records are bound in place
  c.d_col_in.fixed = true;
  c.h_flags = d.flags;
  // host vectors reach their sizes here, not in a call
  c.colony_geom.reserve(1);
  c.colony_start.reserve(2);
  c.boxes.reserve(1);
  c.boxes_next.reserve(1);
  d.ptr = ptr;
  d.ready = true;
  return BMX_OK;
}

}  // namespace

extern "C" bmx_status bmx_evolve_aos_noalloc(void *aos, int64_t n,
                                             const bmx_layout *layout,
                                             const bmx_geometry *geom,
                                             const double fpar[9],
                                             const bmx_neighbor_params *nbp,
                                             double dt, int32_t nsteps,
                                             void *stream) {
  if (!geom || !fpar || !nbp) return fail(BMX_ERR_BAD_ARG, "null argument");
  if (n < 0 || n > kNaParticles)
    return fail(BMX_ERR_BAD_ARG, "particle count outside 0.." +
                                     std::to_string(kNaParticles));
  if (n > 0 && !aos) return fail(BMX_ERR_BAD_ARG, "null record array");
  if (nsteps < 0 || !std::isfinite(dt))
    return fail(BMX_ERR_BAD_ARG, "bad dt or nsteps");
  const int32_t maxn = nbp->max_neighbors > 0 ? nbp->max_neighbors : kNaNeighbors;
  if (maxn > kNaNeighbors)
    return fail(BMX_ERR_BAD_ARG, "max_neighbors above " + std::to_string(kNaNeighbors));
  if (nbp->rebuild_every > kNaRebuild)
    return fail(BMX_ERR_BAD_ARG, "rebuild_every above " + std::to_string(kNaRebuild));
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNaDevices)
    return fail(BMX_ERR_NO_DEVICE, "no current HIP device");
  NaDevice &d = g_na[dev];
  if (!d.ready) {
    bmx_status st = na_setup(d, dev);
    if (st != BMX_OK) return st;
  }
  bmx_ctx *c = &d.ctx;
  c->stream = static_cast<hipStream_t>(stream);
  c->timing = false;
  // The previous call's binding is void; the configuration calls then only
  // record their parameters for the bind below, which sizes every buffer
  // within its fixed slice (growth fails instead of allocating).
  c->bound = false;
  bmx_status st = bmx_set_force_params(c, fpar);
  if (st != BMX_OK) return st;
  bmx_neighbor_params p = *nbp;
  p.max_neighbors = maxn;
  st = bmx_set_neighbor_params(c, &p);
  if (st != BMX_OK) return st;
  st = bmx_set_geometry(c, geom);
  if (st != BMX_OK) return st;
  int64_t cs[2] = {0, n};
  st = bind_common(c, aos, cs, 1, geom, layout, 1, true);
  if (st != BMX_OK) return st;
  st = bmx_evolve(c, dt, nsteps);
  if (st != BMX_OK) return st;
  st = pack_records(c);  // into the caller's records
  if (st != BMX_OK) return st;
  HIPCHK(hipStreamSynchronize(c->stream));
  return BMX_OK;
}
