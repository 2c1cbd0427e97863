// bmx_kernels.h -- host-side launchers for the gfx950 kernels.
#pragma once

#include "bmx_internal.h"

namespace bmx {

struct LaunchShape {
  int block;  // threads per block
  int remap;  // XCD-contiguous tile mapping
};

// Flat row layout written by the list build: each tile of 'tile' particles
// (64 for spheres, 32 for segment contexts) has its rows back to back in
// cidx from the start of its own region (tile index * tile * maxn), tot[tile]
// entries, with the particle's place in the tile beside every entry in
// cown.  cidx == nullptr builds no flat rows.
struct FlatRows {
  int tile;
  int32_t *cidx;
  uint8_t *cown;
  int32_t *tot;
};
// Entries to allocate for the flat rows of ncap particles.
int64_t flat_entries(int64_t ncap, int tile, int32_t maxn);

hipError_t launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long *p, int64_t n,
                           unsigned long long v, hipStream_t s);

// AoS -> sorted SoA (identity order), periodic wrap, per-box bounds stats.
// stats: [0] max radius bits, [1] max c_length bits, [2] any yeast,
// [3] any fungi, [4] bad record count
hipError_t launch_aos_unpack(const uint8_t *aos, AosLayout lay, int64_t n,
                             const int32_t *col, const BoxGeom *boxes,
                             StaticArrays st, DynArrays dy,
                             unsigned long long *stats, hipStream_t s);
// Fields bmx_refresh_from_aos re-reads (the BMX_FIELD_* values of the ABI).
enum : uint32_t {
  kFieldPositions = 1, kFieldSizes = 2, kFieldOrientations = 4,
  kFieldBonds = 8, kFieldSpecies = 16
};
// AoS -> the current sorted arrays, selected fields only, in place of the
// values there (stats as launch_aos_unpack: [0] max radius, [1] max
// c_length, [4] bad records)
hipError_t launch_aos_refresh(const uint8_t *aos, AosLayout lay, int64_t n,
                              const BoxGeom *boxes, StaticArrays st,
                              DynArrays dy, uint32_t fields,
                              unsigned long long *stats, hipStream_t s);
// sorted SoA -> AoS at original positions (fields the force pass writes)
hipError_t launch_aos_pack(uint8_t *aos, AosLayout lay, int64_t n,
                           StaticArrays st, DynArrays dy, const double *fc,
                           double mob, int wrote_step, int any_fungi,
                           hipStream_t s);

// The displacement check of a Verlet period folded into the first two
// binning kernels of the rebuild that ends it: the first blocks of the bin
// count reduce the per-step partials (as launch_skin_check with clear = 1,
// into slotmax), and the first block of the scan sums the slot maxima,
// compares them with limit and resets them.  vpart = nullptr: no check.
struct SkinCheck {
  double *vpart;
  int32_t nslots, stride;
  unsigned long long *slotmax;
  double limit;
  int32_t *err;
};

// ---- binning: key/rank, scan, scatter, per-bin order with the gather
hipError_t launch_bin_count(const double4 *pos, const int32_t *col,
                            const BoxGeom *boxes, BoxGeom g0, int multi,
                            int64_t n, int32_t *key, int32_t *rank,
                            int32_t *count, hipStream_t s,
                            SkinCheck sk = SkinCheck{});
// exclusive scan of count[0..nb) into start[0..nb]; scratch >= scan_scratch(nb)
int64_t scan_scratch_elems(int64_t nb);
hipError_t launch_scan(int32_t *count, int32_t *start, int64_t nb,
                       int32_t *scratch, hipStream_t s,
                       SkinCheck sk = SkinCheck{});
hipError_t launch_bin_scatter(const int32_t *key, const int32_t *rank,
                              const int32_t *start, const int32_t *oid,
                              int64_t n, unsigned long long *permk,
                              hipStream_t s);
// Orders each bin by original index and gathers the particles into that
// order (src -> dst buffers); key: each old position's bin.
hipError_t launch_sort_gather(const unsigned long long *permk,
                              const int32_t *start, int64_t n,
                              StaticArrays src_s, StaticArrays dst_s,
                              DynArrays src_d, DynArrays dst_d,
                              const int32_t *key, int multi, int rods,
                              int32_t *err, hipStream_t s,
                              const double *chem_src = nullptr,
                              double *chem_dst = nullptr);

// ---- neighbor list: flat rows (FlatRows) always; column-major rows
// nbr[k*ncap + i] for segment contexts, and for sphere contexts only when
// cols = 1 (the passes that walk rows that way: exchange, fusion); ncnt
// holds the row length, bits 28..30 the axes that may need a periodic image
// mode 0: r2 <= (reach_i + reach_j + wa + skin)^2 (reach = radius, or 1.5 *
// c_length for segments); mode 1: r2 <= crit (the neighborhood value)
hipError_t launch_build_nbr(const double4 *pos, const int32_t *col,
                            const int32_t *type, const double4 *rod,
                            const BoxGeom *boxes, BoxGeom g0, int multi,
                            int rods, const int32_t *start, const int32_t *key,
                            int64_t n, int mode, double crit, double wa,
                            double skin, int32_t *nbr, int32_t *ncnt,
                            int64_t ncap, int32_t maxn, int32_t *err,
                            int32_t *bstats, FlatRows flat, int cols,
                            LaunchShape ls, hipStream_t s);
// Particles per tile of the flat row layout: 64 for spheres, 32 for
// segment (and mixed) contexts.
int flat_tile(int segments);
// Sphere force + overdamped update, fused, over the flat row layout.  fc
// (6*ncap, original order: fc[oid + c*ncap]) is written when non-null,
// with fc_oid the original index of each sorted row; vmax[wave] receives
// the wave's max dt*|vel| when non-null.
hipError_t launch_force_sphere_flat(const double4 *pos_in, double4 *pos_out,
                                    const int32_t *cidx, const uint8_t *cown,
                                    const int32_t *wtot, const int32_t *ncnt,
                                    int64_t n, int32_t maxn,
                                    const int32_t *col,
                                    const BoxGeom *boxes, BoxGeom g0, int multi,
                                    ForceParams fp, double *fc,
                                    const int32_t *fc_oid, double *vmax,
                                    int32_t *err, LaunchShape ls,
                                    hipStream_t s);
// blocks of a row build / force launch over n particles (partial arrays)
int32_t row_blocks(int64_t n, int block);
// The end of a call in one launch: row statistics (nblk > 0), then err,
// longest row and entry total (low, high word) into host_out (host memory
// mapped for the device), err cleared.
hipError_t launch_flags_out(const int32_t *bstats, int32_t nblk, int32_t *stats,
                            unsigned long long *total, int32_t *err,
                            int32_t *host_out, hipStream_t s);
// stats[0] = longest row, *total = entries, from the per-block partials
hipError_t launch_reduce_rowstats(const int32_t *bstats, int32_t nblk,
                                  int32_t *stats, unsigned long long *total,
                                  hipStream_t s);

// ---- segment (and mixed) force + overdamped update, fused: fc and vmax
// as for launch_force_sphere_flat, vmax per block.
hipError_t launch_force_general(DynArrays in, DynArrays out, StaticArrays st,
                                const int32_t *cidx, const uint8_t *cown,
                                const int32_t *wtot, int8_t *redo_tile,
                                const int32_t *nbr, const int32_t *ncnt,
                                int64_t ncap, int64_t n, int32_t maxn,
                                const BoxGeom *boxes, BoxGeom g0, int multi,
                                ForceParams fp, double *fc, double *vmax,
                                int32_t *err, LaunchShape ls, hipStream_t s);
hipError_t launch_init_bonds(DynArrays dy, StaticArrays st, const int32_t *nbr,
                             const int32_t *ncnt, int64_t ncap, int64_t n,
                             const BoxGeom *boxes, BoxGeom g0, int multi,
                             ForceParams fp, int32_t *err, hipStream_t s);
// Displacement record of the Verlet skin check: one slot of `stride`
// doubles per step since the last rebuild; force kernels store each wave's
// (or block's) largest dt*|vel| at its index in their step's slot.
// sum over slots of max over the slot > limit -> err |= ERRBIT_SKIN;
// clear = 1 zeroes the partials (at a rebuild).  slotmax: nslots zeroed
// words of scratch (left zeroed).
hipError_t launch_skin_check(double *vpart, int32_t nslots, int32_t stride,
                             double limit, int clear,
                             unsigned long long *slotmax, int32_t *err,
                             hipStream_t s);

// ---- exports to original order
hipError_t launch_export_fc(const double *fc, int64_t n, double *out,
                            hipStream_t s);
hipError_t launch_export_pos(const double4 *pos, const int32_t *oid, int64_t n,
                             double *out, hipStream_t s);
hipError_t launch_export_orient(const double4 *rod, const int32_t *oid,
                                int64_t n, double *out, hipStream_t s);
hipError_t launch_export_i32(const int32_t *v, const int32_t *oid, int64_t n,
                             int32_t *out, hipStream_t s);
hipError_t launch_export_nbr(FlatRows flat, int32_t maxn, const int32_t *ncnt,
                             const int32_t *oid, int64_t n,
                             const int64_t *row_off /* by original index */,
                             int32_t *out, hipStream_t s);
hipError_t launch_export_counts(const int32_t *ncnt, const int32_t *oid,
                                int64_t n, int64_t *out_counts,
                                hipStream_t s);

// ---- adjoint of the sphere step (bmx_adjoint.hip)
int32_t adj_blocks(int64_t n);
hipError_t launch_adj_loss(const double4 *pos, const int32_t *oid,
                           const double *target, int64_t n, BoxGeom g,
                           double4 *lam, double *lpart, hipStream_t s);
hipError_t launch_adj_step(const double4 *pos, const int32_t *cidx,
                           const uint8_t *cown, const int32_t *wtot,
                           const int32_t *ncnt, int32_t maxn,
                           const int32_t *oid, int64_t n, BoxGeom g,
                           ForceParams fp, const double4 *lam_in,
                           double4 *lam_out, double *grad_r, double *acc,
                           hipStream_t s);
hipError_t launch_to_orig4(const double4 *in, const int32_t *oid, int64_t n,
                           double4 *out, hipStream_t s);
hipError_t launch_from_orig4(const double4 *in, const int32_t *oid, int64_t n,
                             double4 *out, hipStream_t s);
hipError_t launch_sum_partials(const double *part, int64_t nblk, int stride,
                               double *out, hipStream_t s);

// The device chemistry state: per particle, in sorted order (moved with the
// other sorted arrays at every rebuild), the 16 doubles of one 128-byte line
// that ParticleExchange, the mesh -> particle chemistry and the deposits read
// and write, so none of them touches the records until they are packed (at
// sync, and before a pass that works on the records).
namespace chem {
enum : int {
  A = 0, B = 1, C = 2, vol = 3,          // species, cell volume
  dA = 4, dB = 5, dC = 6, area = 7,      // species increments, area
  dadt = 8, dvdt = 9, gx = 10, gy = 11,  // growth rates, species-A gradient
  gz = 12, rel0 = 13, rel1 = 14, rel2 = 15,  // amounts taken from the fluid
  stride = 16
};
}
// records -> state (sorted position t from record oid[t]); has_rel: the
// layout holds rdata[first_data + 6..8]
hipError_t launch_chem_unpack(const uint8_t *aos, AosLayout lay, int64_t n,
                              int has_rel, const int32_t *oid, double *state,
                              hipStream_t s);
// state -> records, with each particle's radius (pos.w) and, for segments,
// its length (rod.x) from the sorted arrays
hipError_t launch_chem_pack(uint8_t *aos, AosLayout lay, int64_t n,
                            int has_rel, const double *state,
                            const int32_t *oid, const int32_t *type,
                            const double4 *pos, const double4 *rod,
                            hipStream_t s);
// ParticleExchange (bmx_exchange.hip) on the chemistry state
hipError_t launch_exchange(const double4 *pos, const int32_t *type,
                           const Topo *topo,
                           const int32_t *nbr, const int32_t *ncnt,
                           int64_t ncap, int64_t n, double *state,
                           const double xpar[3], double dt, hipStream_t s);

// ---- particle <-> mesh (bmx_mesh.hip), over the sorted arrays
// Bins the particles by mesh cell, or by trilinear base cell (tri);
// start[0..nb] from the counts; with 'sorted', perm lists each bin's
// particles as (original index << 32 | sorted index), ascending.
hipError_t launch_mesh_bin(const double4 *pos, const int32_t *oid, int64_t n,
                           const MeshGeom &g, int tri, int32_t *key,
                           int32_t *rank, int32_t *count, int32_t *start,
                           int32_t *scan_scratch, unsigned long long *perm,
                           int sorted, int32_t *err, hipStream_t s);
// mode 0: sums of rdata[comp0..comp0+ncomp); 1: volume fraction of the
// deposited volume; 2: particles per cell (one-to-one bins).  The source
// records are indexed by original index, or by sorted index with by_sorted
// (the chemistry state seen through a layout of stride 128 bytes)
hipError_t launch_deposit(const uint8_t *aos, AosLayout lay, const MeshGeom &g,
                          int tri, const int32_t *start,
                          const unsigned long long *perm, const double4 *pos,
                          int comp0, int ncomp, int mode, double *out,
                          hipStream_t s, int by_sorted = 0);
// interpolation + chemistry per particle on the chemistry state; segments'
// bond count and tip flag from topo (current);
// new sizes into pos.w / rod.x / rod.w; stats[0], stats[1]: max radius and
// segment length (bits)
hipError_t launch_mesh_to_particle(double *state, int64_t n,
                                   const MeshGeom &g, const int32_t *start,
                                   const int32_t *type,
                                   const Topo *topo, double4 *pos, double4 *rod,
                                   const double *X, const double *vf,
                                   int cnc_tri, int vf_tri,
                                   const double chempar[10], double dt,
                                   int nloop, unsigned long long *stats,
                                   int32_t *err, hipStream_t s);

// ---- division and branching (bmx_split.hip); sp = {max_vol, max_len,
// max_rad, brnch_prob, split_prob, overlap, split_len}
hipError_t launch_split_count(uint8_t *aos, AosLayout lay, int64_t n,
                              const double sp[7], uint64_t seed, uint32_t call,
                              int32_t *cnt, hipStream_t s);
// records [0, n) with room for the new ones at n + start[o]
hipError_t launch_split_apply(uint8_t *aos, AosLayout lay, int64_t n,
                              const int32_t *start, const double sp[7],
                              uint64_t seed, uint32_t call, int32_t next_id,
                              int32_t cpu, hipStream_t s);

// ---- fusion (bmx_fusion.hip); fus = {fusion_probability,
// max_fusion_separation, bond_scaling_increment, max_seg_length}; tip
// fusion decides every tip first (tgt/ttau: n scratch entries each), then
// applies the decisions
hipError_t launch_tip_fusion(uint8_t *aos, AosLayout lay, int64_t n,
                             const double4 *pos, const int32_t *oid,
                             const int32_t *nbr, const int32_t *ncnt,
                             int64_t ncap, const BoxGeom &g, double wa,
                             const double fus[4], uint64_t seed, uint32_t call,
                             int32_t *tgt, double *ttau, int32_t *fused,
                             hipStream_t s);
// cnt[o]: tips that fused into o; one[o] = (cnt[o] == 1), the new segments
hipError_t launch_interior_fusion(uint8_t *aos, AosLayout lay, int64_t n,
                                  const double4 *pos, const int32_t *oid,
                                  const int32_t *nbr, const int32_t *ncnt,
                                  int64_t ncap, const BoxGeom &g, double wa,
                                  int32_t *cnt, int4 *dec, double *dtau,
                                  int32_t *one, hipStream_t s);
hipError_t launch_fusion_split(uint8_t *aos, AosLayout lay, int64_t n,
                               const int32_t *cnt, const int32_t *start,
                               int32_t next_id, int32_t cpu, hipStream_t s);
hipError_t launch_fusion_cleanup(uint8_t *aos, AosLayout lay, int64_t n,
                                 const double4 *pos, const int32_t *oid,
                                 const int32_t *nbr, const int32_t *ncnt,
                                 int64_t ncap, const BoxGeom &g,
                                 double scale_inc, double reach,
                                 hipStream_t s);

}  // namespace bmx
