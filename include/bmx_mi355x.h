/*
 * bmx_mi355x.h -- C ABI of the MI355X particle-mechanics library.
 *
 * Replaces, for one AMReX level, the particle-mechanics path of BMX's
 * BMXParticleContainer (src/des/bmx_pc.H:86-174):
 *   - neighbor binning and list construction done by AMReX's
 *     NeighborParticleContainer (clearNeighbors / Redistribute /
 *     fillNeighbors / buildNeighborList, src/des/bmx_pc_interaction.cpp:97-106),
 *   - the force lambda (src/des/bmx_pc_interaction.cpp:170-279) over the
 *     force laws of src/chemistry/bmx_cell_interaction_K.H,
 *   - the overdamped update lambda (src/des/bmx_pc_interaction.cpp:311-449),
 *   - InitBonds (src/des/bmx_pc_interaction.cpp:500-668).
 *
 * Plain C types only.  Every call returns a bmx_status; bmx_last_error()
 * gives a message for the calling thread's last failure.  A context owns one
 * HIP stream (or uses the caller's, bmx_set_stream) and all device buffers;
 * calls on one context must not overlap.
 */
#ifndef BMX_MI355X_H
#define BMX_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BMX_ABI_VERSION 1

typedef struct bmx_ctx bmx_ctx;

typedef enum {
  BMX_OK = 0,
  BMX_ERR_BAD_ARG = 1,       /* invalid argument or geometry */
  BMX_ERR_OOM = 2,           /* device or host allocation failed */
  BMX_ERR_HIP = 3,           /* HIP runtime error */
  BMX_ERR_NAN = 4,           /* non-finite force (coincident particles) */
  BMX_ERR_NBR_OVERFLOW = 5,  /* a neighbor row exceeded max_neighbors */
  BMX_ERR_SKIN = 6,          /* displacement bound exceeded skin/2 between
                                rebuilds: the pair set may be incomplete */
  BMX_ERR_STATE = 7,         /* call order (e.g. evolve before bind) */
  BMX_ERR_IO = 8,            /* file parse / write failure */
  BMX_ERR_BOND_OVERFLOW = 9, /* InitBonds found a fifth bond */
  BMX_ERR_NO_DEVICE = 10,    /* no HIP device / library not usable */
  BMX_ERR_MESH = 11,         /* a particle outside the level mesh, or more
                                than 4096 particles in one mesh cell */
  BMX_ERR_CHEM = 12          /* negative concentration in the particle
                                chemistry (where the reference aborts) */
} bmx_status;

/* Where the fields of one particle record sit (bytes).  The default is
 * AMReX Particle<34,35> as BMX declares it (src/des/bmx_pc.H:86-96):
 * pos[3] f64 @0, rdata[34] f64 @24, id i32 @296, cpu i32 @300,
 * idata[35] i32 @304, stride 448. */
typedef struct {
  int32_t stride_bytes;
  int32_t pos_offset;
  int32_t rdata_offset;
  int32_t id_offset;
  int32_t cpu_offset;
  int32_t idata_offset;
  int32_t n_real;
  int32_t n_int;
} bmx_layout;

/* Level geometry (Geom(lev).ProbLo/ProbHi, geometry.is_periodic) and the
 * level mesh n_cell, which sets the bins in BMX_NBR_STENCIL mode. */
typedef struct {
  double prob_lo[3];
  double prob_hi[3];
  int32_t periodic[3];
  int32_t n_cell[3];
} bmx_geometry;

/* Neighbor-list policy.
 *   BMX_NBR_STENCIL: candidates are all particles in the +-1 bin stencil of
 *     the level mesh whose squared separation is <= 'neighborhood'
 *     (BMXCheckPair, src/des/bmx_pc.H:63-81); rebuilt every step as the
 *     reference does (src/des/bmx_pc_interaction.cpp:97-103).
 *   BMX_NBR_VERLET: candidates within max interaction distance + skin, bins
 *     no smaller than that radius, rebuilt every rebuild_every steps.  The set
 *     of interacting pairs is identical to the stencil list's as long as
 *     BMX_ERR_SKIN is not raised. */
enum { BMX_NBR_VERLET = 0, BMX_NBR_STENCIL = 1 };

typedef struct {
  int32_t mode;           /* BMX_NBR_VERLET or BMX_NBR_STENCIL */
  int32_t rebuild_every;  /* steps between rebuilds (>= 1) */
  int32_t max_neighbors;  /* row capacity (0 = default 96) */
  int32_t bin_factor_pct; /* VERLET: bin edge >= radius*pct/100 (0 = 100) */
  double skin;            /* VERLET: skin added to the max r_lm */
  double neighborhood;    /* STENCIL: DEM::neighborhood
                             (src/chemistry/bmx_chem.cpp:124) */
} bmx_neighbor_params;

/* Launch shape of the pair kernels (the autotuner's search space).
   A new context uses {64, 1}. */
typedef struct {
  int32_t block_size;     /* threads per block: 64, 128, 256, 512 */
  int32_t xcd_remap;      /* 1: give each of the 8 XCDs a contiguous range
                             of particle tiles */
} bmx_launch_params;

typedef struct {
  double force_ms;        /* summed duration of force+update launches */
  int64_t force_launches;
  double rebuild_ms;      /* summed duration of binning + list builds */
  int64_t rebuilds;
  int64_t neighbor_entries; /* total list entries at the last rebuild */
  int32_t max_row;          /* longest row at the last rebuild */
  int32_t nbins;            /* bins at the last rebuild */
  int64_t rebuilds_total;   /* list builds since bind, timed or not */
  int64_t steps_total;      /* steps since bind */
  int64_t guard_redone;     /* guarded periods taken again since bind */
} bmx_timing;

/* --- lifetime ---------------------------------------------------------- */
int32_t bmx_abi_version(void);
const char *bmx_last_error(void);
/* device = HIP ordinal */
bmx_status bmx_ctx_create(bmx_ctx **out, int32_t device);
void bmx_ctx_destroy(bmx_ctx *ctx);

/* --- configuration ----------------------------------------------------- */
/* fpar[9] = {boundary_width W_A, stiffness mu, z_wall, wall_boundary_width,
 *            wall_stiffness, gravity, bond_strength, bond_cutoff,
 *            viscous_drag}: BMXCellInteraction::getForceParams()
 *            (src/chemistry/bmx_cell_interaction.cpp:84-104). */
bmx_status bmx_set_force_params(bmx_ctx *ctx, const double fpar[9]);
/* Overdamped mobility m: vel = m * fc.  The reference has m = 1
 * (src/des/bmx_pc_interaction.cpp:321-323), which is the default. */
bmx_status bmx_set_mobility(bmx_ctx *ctx, double mobility);
bmx_status bmx_set_geometry(bmx_ctx *ctx, const bmx_geometry *geom);
bmx_status bmx_set_neighbor_params(bmx_ctx *ctx,
                                   const bmx_neighbor_params *p);
bmx_status bmx_set_launch_params(bmx_ctx *ctx, const bmx_launch_params *p);

/* Result of bmx_autotune. */
typedef struct {
  int32_t block_size;       /* chosen launch shape ... */
  int32_t xcd_remap;
  int32_t bin_factor_pct;   /* ... and Verlet bin edge, % of the radius */
  int32_t from_cache;       /* 1: read from the cache file, no trials run */
  double us_per_step;       /* best measured time per step (0 from cache) */
  int32_t candidates;       /* configurations timed */
  int32_t pad;
} bmx_tuning;
/* Autotuner (no counterpart in the reference, whose launch shape is fixed by
 * AMReX): times graph-replayed rebuild periods of the bound particles for
 * every block size, XCD mapping and (Verlet policy) bin edge of 100, 150
 * and 200% of the interaction radius, applies the fastest and returns it.
 * The particle state is restored afterwards, so tuning does not advance the
 * run; the next step rebuilds the lists.  When cache_path is not NULL the
 * choice is looked up in / appended to that file (one JSON object per line,
 * keyed by architecture, policy, rebuild interval, colony count, particle
 * count and packing).  trial_steps is rounded up to whole rebuild periods;
 * trials of several periods, taken after the run has passed its first
 * rebuild, rank the shapes as a long run does (two-period trials from
 * freshly bound records have favoured shapes that later ran slower). */
bmx_status bmx_autotune(bmx_ctx *ctx, double dt, int32_t trial_steps,
                        const char *cache_path, bmx_tuning *out);
bmx_layout bmx_default_layout(void);

/* --- particles --------------------------------------------------------- */
/* Where bound records live: copied from host memory, copied from device
 * memory, or device memory used in place (the kernels read and write the
 * caller's records directly: no copy at bind or sync; the caller keeps the
 * array alive, and its count and order unchanged, until the next bind). */
enum { BMX_AOS_HOST = 0, BMX_AOS_DEVICE = 1, BMX_AOS_IN_PLACE = 2 };
/* Binds n records (aos_on_device: a BMX_AOS_* value) and applies the
 * periodic wrap that Redistribute performs (src/des/bmx_pc_init.cpp:131).
 * The record array is remembered for bmx_sync_to_aos.  A bind replaces the
 * previous one: when it fails, the context is left with no particles bound
 * (BMX_ERR_STATE from every call that needs them until a bind succeeds).
 * The lists, the sorted order and the step count start afresh; between
 * binds the context keeps them, so a caller that binds once and then calls
 * bmx_evolve per step (the reference's EvolveParticles cadence,
 * src/timestepping/bmx_evolve.cpp:97) keeps its Verlet lists across calls. */
bmx_status bmx_bind_particles_aos(bmx_ctx *ctx, const void *aos, int64_t n,
                                  const bmx_layout *layout,
                                  int32_t aos_on_device);
/* A level held as several tiles (AMReX: one particle tile per box and tile,
 * ParIter over the level, src/des/bmx_pc_interaction.cpp:114), bound as ONE
 * neighbour domain with the context's geometry (the level's
 * Geom(lev).ProbLo/ProbHi and periodicity): pairs across tile and box
 * boundaries are found as AMReX finds them through its ghost copies
 * (fillNeighbors, :100), and periodic images apply only at the domain's
 * periodic faces.  Tile k holds counts[k] records (host memory, or device
 * memory when aos_on_device == 1); particle indices run through the tiles
 * in order.  bmx_sync_to_aos(ctx, NULL) writes each tile's records back to
 * it. */
bmx_status bmx_bind_tiles_aos(bmx_ctx *ctx, const void *const *tiles,
                              const int64_t *counts, int32_t ntiles,
                              const bmx_layout *layout, int32_t aos_on_device);
/* Several independent periodic boxes ("colonies") in one context: particles
 * [colony_start[c], colony_start[c+1]) live in geoms[c].  Each colony's
 * neighbor search, bins and wrap are its own. */
bmx_status bmx_bind_colonies_aos(bmx_ctx *ctx, const void *aos,
                                 const int64_t *colony_start, int32_t ncol,
                                 const bmx_geometry *geoms,
                                 const bmx_layout *layout,
                                 int32_t aos_on_device);
/* Writes positions, velocities, angular velocities, orientation angles and
 * the integer fields the force pass changes back into 'aos' (host or device,
 * as bound; NULL = the bound array, or each bound tile; records bound in
 * place are written where they are), and, once bmx_exchange or
 * bmx_calc_txfr_particle has run since the bind, the chemistry fields they
 * keep on the device: radius, c_length, area, vol, gx..gz, dadt, dvdt and
 * rdata[first_data .. first_data + 8]. */
bmx_status bmx_sync_to_aos(bmx_ctx *ctx, void *aos);
/* Fields bmx_refresh_from_aos re-reads. */
enum {
  BMX_FIELD_POSITIONS = 1,    /* pos[3] */
  BMX_FIELD_SIZES = 2,        /* radius; segments: c_length, area, bond_scale */
  BMX_FIELD_ORIENTATIONS = 4, /* segments: theta, phi */
  BMX_FIELD_BONDS = 8,        /* segments: n_bnds, bond ids/cpus/sites,
                                 position, fuse_tip, fix_site */
  BMX_FIELD_SPECIES = 16      /* rdata[25..33] (exchange and chemistry) */
};
/* The caller changed the given fields of the bound particles (same count,
 * same order): re-read just those, keeping the sorted order and, unless
 * positions or sizes changed, the neighbour lists.  aos = NULL re-reads the
 * bound array (or tiles); records bound in place are re-read in place.
 * Positions or sizes make the next step rebuild the lists (with new bins
 * when the largest reach changed); sizes or species make the next exchange
 * or chemistry call reload the device chemistry fields from the records
 * (so sync before editing records, as for every field).  A record that
 * fails the bind's checks unbinds the context (BMX_ERR_BAD_ARG). */
bmx_status bmx_refresh_from_aos(bmx_ctx *ctx, const void *aos, uint32_t fields);
int64_t bmx_num_particles(const bmx_ctx *ctx);

/* --- the hot path ------------------------------------------------------ */
/* InitBonds: bond detection between touching segment ends (rlim = 1e-7,
 * src/des/bmx_pc_interaction.cpp:508). */
bmx_status bmx_init_bonds(bmx_ctx *ctx);
/* EvolveParticles: nsteps explicit steps of size dt (nsubsteps == 1 per
 * step, src/des/bmx_pc_interaction.cpp:56-65).  Synchronizes before
 * returning and reports list overflow, skin violation or non-finite forces. */
bmx_status bmx_evolve(bmx_ctx *ctx, double dt, int32_t nsteps);
/* Same steps, replayed from one captured HIP graph (captured on first use
 * for this (dt, nsteps) and re-used while the configuration is unchanged). */
bmx_status bmx_evolve_graph(bmx_ctx *ctx, double dt, int32_t nsteps);
/* Guarded Verlet steps (on = 1; off by default): bmx_evolve checks every
 * rebuild period as it ends, and a period in which the summed per-step
 * displacement bound exceeded skin/2 (a pair may have come into range
 * unlisted) is restored from a snapshot taken at its start and taken again
 * with lists rebuilt before every step, so BMX_ERR_SKIN is never returned
 * and every force uses a list that holds all interacting pairs, as the
 * reference's per-step rebuild (src/des/bmx_pc_interaction.cpp:97-103)
 * does.  The next period is enqueued before the previous one's check is
 * read, so the device does not wait on it; the snapshot is one copy of the
 * particle arrays per period.  With the guard on, bmx_evolve_graph runs
 * these eager guarded steps.  bmx_timing.guard_redone counts the periods
 * taken again. */
bmx_status bmx_set_guard(bmx_ctx *ctx, int32_t on);
/* The steps of bmx_evolve, enqueued on the context's stream without waiting
 * for them; asynchronous calls may follow one another, the device running
 * them back to back.  bmx_wait completes them and returns their combined
 * status (list overflow, skin, non-finite force); any other call on the
 * context does so first too, so an error of asynchronous steps is reported
 * by bmx_wait or by the next call.  The skin check of asynchronous steps
 * runs at the rebuild points (not also at the end of each call).  Not with
 * the guard on.  bmx_query sets *done to 1 once they have finished. */
bmx_status bmx_evolve_async(bmx_ctx *ctx, double dt, int32_t nsteps);
bmx_status bmx_wait(bmx_ctx *ctx);
bmx_status bmx_query(bmx_ctx *ctx, int32_t *done);
/* Enqueue the context's work on the caller's stream (a hipStream_t; NULL =
 * the context's own stream, the default).  Work already enqueued completes
 * first.  Keep the stream alive while the context uses it. */
bmx_status bmx_set_stream(bmx_ctx *ctx, void *stream);

/* --- gradients (no counterpart in the reference) ------------------------ */
typedef struct {
  double forward_ms;   /* forward steps incl. checkpoints and the loss */
  double backward_ms;  /* recompute of each segment + reverse steps */
  int32_t segments;
  int32_t pad;
} bmx_adjoint_timing;
/* Discrete adjoint of nsteps sphere steps (spheres only, one box):
 * L = (1/N) sum_i |minimum image(x_T,i - target_i)|^2, target given as n*3
 * doubles in original order.  Positions are checkpointed every
 * checkpoint_every steps (a multiple of rebuild_every) and each segment is
 * recomputed before its reverse pass.  Outputs: *loss, grad3 = {dL/dmu
 * (stiffness), dL/dW_A (adhesion band), dL/dgamma with mobility 1/gamma} and
 * grad_radius (n doubles, original order).  The context is left in the
 * final forward state.  timing may be NULL. */
bmx_status bmx_adjoint(bmx_ctx *ctx, double dt, int32_t nsteps,
                       int32_t checkpoint_every, const double *target_xyz,
                       double *loss, double grad3[3], double *grad_radius,
                       bmx_adjoint_timing *timing);

/* ParticleExchange (src/des/bmx_pc_xchng.cpp:13-257, kernels
 * src/chemistry/bmx_chem_K.H:1951-2112): rebuilds the rows at the current
 * positions, then moves species A, B, C (rdata[25..27]) between bonded
 * segments at rates xpar = {mass_transfer_A, _B, _C}
 * (BMXChemistry::getExchangeParameters, bmx_chem.cpp:221-228) over dt, and
 * leaves the increments in rdata[28..30].  The records must carry at least
 * 31 reals.  bmx_sync_to_aos writes the species back. */
bmx_status bmx_exchange(bmx_ctx *ctx, const double xpar[3], double dt);

/* --- particle checkpoints ------------------------------------------------
 * The "particles" directory of a BMX checkpoint, AMReX's particle
 * checkpoint format as BMX writes it (pc->Checkpoint(name, "particles"),
 * src/io/bmx_chk.cpp:156-159) and reads it on restart (pc->Restart,
 * src/io/bmx_restart.cpp:104-107). */
typedef struct {
  int32_t version;      /* 20: Version_Two_Dot_Zero_double, 21: Two_Dot_One */
  int32_t n_real;       /* real components after x, y, z */
  int32_t n_int;        /* int components after the two id words */
  int32_t finest_level;
  int64_t n_particles;
  int64_t next_id;
} bmx_checkpoint_header;
bmx_status bmx_checkpoint_info(const char *dir, bmx_checkpoint_header *out);
/* Reads every particle of every level into records (NULL: only *n_out).
 * real_map[k] / int_map[k] name the file component that fills rdata[k] /
 * idata[k] (-1: left zero); NULL maps take component k for k (the file
 * must then hold all of the record's components).  Maps let checkpoints of
 * older record layouts restart. */
bmx_status bmx_read_checkpoint(const char *dir, void *records, int64_t capacity,
                               int64_t *n_out, const bmx_layout *layout,
                               const int32_t *real_map, int32_t n_real_map,
                               const int32_t *int_map, int32_t n_int_map);
/* Writes Version_Two_Dot_One_double with all n_real / n_int components, one
 * grid spanning geom->n_cell at the given level. */
bmx_status bmx_write_checkpoint(const char *dir, const void *records, int64_t n,
                                const bmx_layout *layout,
                                const bmx_geometry *geom, int32_t level,
                                int64_t next_id);

/* --- inspection (original particle order) ------------------------------ */
/* fc of the last step, 6*n doubles, component-major: the reference's
 * fc[i + c*ntot] (src/des/bmx_pc_interaction.cpp:244-249). */
/* --- particle <-> mesh coupling ------------------------------------------
 * On the level mesh of bmx_set_geometry (n_cell; dx = (hi - lo) / n_cell).
 * Meshes are arrays of doubles, x fastest: cell (i, j, k) of component m at
 * m*N + i + n0*(j + n1*k), N = n0*n1*n2.  Meshes passed IN carry one ghost
 * layer: (n0+2)(n1+2)(n2+2) values per component, cell (i, j, k) at
 * (i+1) + (n0+2)*((j+1) + (n1+2)*(k+1)), ghosts filled by the caller
 * (periodic images or boundary values, as FillBoundary and the species BCs
 * leave them).  on_device != 0: the pointers are device memory of the
 * context's device; otherwise host memory.  Single-colony contexts only.
 * Sums over the particles of a cell run in ascending particle index. */
enum { BMX_DEPOSIT_ONE_TO_ONE = 0, BMX_DEPOSIT_TRILINEAR = 1 };

/* Volume fraction of the fluid: vf = max(1 - sum(vol)/(dx dy dz), 0.1) with
 * vol = rdata[5] deposited one-to-one or trilinear (bmx::
 * bmx_calc_volume_fraction, src/des/bmx_calc_volume_fraction.cpp:14-86,
 * without the X_k rescaling).  vf: N values. */
bmx_status bmx_calc_volume_fraction(bmx_ctx *ctx, int32_t scheme, double *vf,
                                    int32_t on_device);
/* Species taken from the fluid, rdata[31..33], summed per cell
 * (bmx::bmx_calc_txfr_fluid, src/des/bmx_calc_txfr.cpp:22-66).  3N values. */
bmx_status bmx_calc_txfr_fluid(bmx_ctx *ctx, int32_t scheme, double *x_rhs,
                               int32_t on_device);
/* Particles per cell (pc->Increment, src/des/bmx_calc_txfr.cpp:96-98). */
bmx_status bmx_count_particles_per_cell(bmx_ctx *ctx, double *npart,
                                        int32_t on_device);
/* Mesh -> particle (bmx::bmx_calc_txfr_particle, src/des/bmx_calc_txfr.cpp:
 * 72-396): the species X_k (3 ghosted components) and vf (1 ghosted) are
 * interpolated to every particle (one-to-one or trilinear), the gradient of
 * species A at its cell goes to rdata[18..20], and
 * xferMeshToParticleAndUpdateChem (src/chemistry/bmx_chem_K.H:1358-1940)
 * integrates uptake, reaction and growth over nloop substeps ("substeps",
 * src/setup/bmx_init.cpp:85-86), updating radius, c_length, area, vol,
 * dadt, dvdt and rdata[25..33].  chempar = {k1, k2, k3, kr1, kr2, kr3, kg,
 * kv, radius_max, length_max} (BMXChemistry::getChemParams,
 * src/chemistry/bmx_chem.cpp:203-216).  The next step rebuilds the rows for
 * the new sizes.  BMX_ERR_CHEM where the reference's CPU build aborts
 * (negative concentrations); the records are then unspecified. */
bmx_status bmx_calc_txfr_particle(bmx_ctx *ctx, const double chempar[10],
                                  const double *x_k, const double *vf,
                                  double dt, int32_t nloop, int32_t cnc_scheme,
                                  int32_t vf_scheme, int32_t on_device);

/* --- division and branching ---------------------------------------------
 * split_particles (src/des/bmx_split_particles.cpp:16-166): yeast cells
 * above max_vol divide (setNewCell), hyphal tips at full radius and length
 * grow a new tip or split into two (setNewSegment / setNewSegments), and
 * segments behind a tip branch with branching_probability.  The new
 * particles follow the existing ones in record order with ids next_id + k
 * and the given cpu; *n_new returns their number (the caller advances its
 * NextID by it).  Random numbers are a pure function of (seed, call,
 * particle id, cpu): a run is reproducible and independent of the launch,
 * but the reference's AMReX engines cannot be matched draw for draw.  When
 * particles were added, the records are re-bound: the next bmx_sync_to_aos
 * needs a buffer of bmx_num_particles() records. */
typedef struct {
  double max_vol;               /* chem_species.max_vol */
  double max_seg_length;        /* chem_species.max_seg_length (3.5e-3) */
  double max_seg_radius;        /* chem_species.max_seg_radius (2.5e-4) */
  double branching_probability; /* chem_species.branching_probability (1e-4) */
  double splitting_probability; /* chem_species.splitting_probability (0.1) */
  double overlap;               /* BMXChemistry::p_overlap (0.2) */
  double seg_split_length;      /* chem_species.seg_split_length (3.0e-3) */
} bmx_split_params;
bmx_status bmx_split_particles(bmx_ctx *ctx, const bmx_split_params *params,
                               uint64_t seed, uint32_t call, int32_t next_id,
                               int32_t cpu, int64_t *n_new);

/* --- fusion ----------------------------------------------------------------
 * EvaluateTipFusion, then, if a tip fused, EvaluateInteriorFusion and
 * CleanupFusion (src/des/bmx_pc_fusion.cpp:16-493, called by bmx::Evolve
 * at src/timestepping/bmx_evolve.cpp:100-104) over rows rebuilt at the
 * current positions: a tip whose free end comes within
 * max_fusion_separation of the side of a segment fuses with
 * fusion_probability, the segment splits there (new segments get ids
 * next_id + k and the given cpu; *n_new returns their number), stale bonds
 * are removed and bond scales grow by bond_scaling_increment.  *fused tells
 * whether a tip fused.  Random numbers as bmx_split_particles.  When
 * segments were added the next bmx_sync_to_aos needs a buffer of
 * bmx_num_particles() records. */
typedef struct {
  double fusion_probability;     /* chem_species.fusion_probability (0) */
  double max_fusion_separation;  /* chem_species.max_fusion_separation (5e-4) */
  double bond_scaling_increment; /* chem_species.bond_scaling_increment (1e-3) */
  double max_seg_length;         /* chem_species.max_seg_length (3.5e-3) */
} bmx_fusion_params;
bmx_status bmx_fusion(bmx_ctx *ctx, const bmx_fusion_params *params,
                      uint64_t seed, uint32_t call, int32_t next_id, int32_t cpu,
                      int32_t *fused, int64_t *n_new);

/* --- allocation-free evolve ---------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): one call does what bmx_bind_particles_aos (device
 * records) + bmx_evolve(dt, nsteps) + bmx_sync_to_aos do on a context, on
 * the caller's device records in place, with every working array in a
 * static device scratch the library holds per GPU (about 40 MB, from load
 * to process end).  It calls no allocating routine.  Limits: n <=
 * BMX_NOALLOC_MAX_PARTICLES, max_neighbors <= BMX_NOALLOC_MAX_NEIGHBORS
 * (0 = that), rebuild_every <= BMX_NOALLOC_MAX_REBUILD_EVERY, at most
 * BMX_NOALLOC_MAX_BINS bins; beyond them BMX_ERR_BAD_ARG, before any GPU
 * work.  One box (no colonies).  stream: a hipStream_t, or NULL for the
 * null stream; the call synchronizes it before returning.  Calls share
 * their GPU's scratch, which keeps the last call's contents: calls that
 * overlap on one GPU race. */
#define BMX_NOALLOC_MAX_PARTICLES 32768
#define BMX_NOALLOC_MAX_NEIGHBORS 64
#define BMX_NOALLOC_MAX_REBUILD_EVERY 64
#define BMX_NOALLOC_MAX_BINS 524288
bmx_status bmx_evolve_aos_noalloc(void *aos_device, int64_t n,
                                  const bmx_layout *layout,
                                  const bmx_geometry *geom,
                                  const double fpar[9],
                                  const bmx_neighbor_params *nbp, double dt,
                                  int32_t nsteps, void *stream);

bmx_status bmx_get_fc(bmx_ctx *ctx, double *out6n);
/* positions, n*3 doubles (x,y,z per particle) */
bmx_status bmx_get_positions(bmx_ctx *ctx, double *out3n);
/* theta, phi per particle (2*n doubles, interleaved) */
bmx_status bmx_get_orientations(bmx_ctx *ctx, double *out2n);
/* integer record fields idata[0..34] per particle (35*n int32) */
bmx_status bmx_get_int_fields(bmx_ctx *ctx, int32_t *out35n);
/* Current list, rows in original order, entries as original indices in the
 * order the force pass sums them.  offsets has n+1 entries; *total receives
 * the entry count; indices may be NULL to query the size. */
bmx_status bmx_get_neighbors(bmx_ctx *ctx, int64_t *offsets,
                             int32_t *indices, int64_t capacity,
                             int64_t *total);
/* Linear bin id (bz*nby + by)*nbx + bx of each particle at the last rebuild,
 * bins from floor((pos - prob_lo) * dxi) (src/interpolation/bmx_interp_K.H:55-56). */
bmx_status bmx_get_bins(bmx_ctx *ctx, int32_t *out_n);
/* bins per dimension in use for colony c */
bmx_status bmx_get_bin_dims(bmx_ctx *ctx, int32_t colony, int32_t out3[3]);

/* --- measurement ------------------------------------------------------- */
/* enable = 0 turns timing off; enable > 0 brackets, with HIP events on the
 * context's stream, every rebuild and every run of force launches between
 * two rebuilds (or a rebuild and the end of a bmx_evolve call).  Events sit
 * only at the ends of runs, so force_ms / force_launches is the launches'
 * mean duration without per-launch event overhead, and rebuild_ms plus
 * force_ms covers the steps' device time.  bmx_evolve_graph refuses to run
 * while timing is on. */
bmx_status bmx_set_timing(bmx_ctx *ctx, int32_t enable);
bmx_status bmx_get_timing(bmx_ctx *ctx, bmx_timing *out);
bmx_status bmx_reset_timing(bmx_ctx *ctx);
bmx_status bmx_synchronize(bmx_ctx *ctx);

/* --- particle files and synthetic inputs (host only) ------------------- */
/* InitParticlesAscii format (src/des/bmx_pc_init.cpp:32-104): the count on
 * line 1, then per particle x y z radius c_length theta phi area vol velx
 * vely velz wx wy wz fx fy fz taux tauy tauz gx gy gz dadt dvdt cell_type
 * (extra trailing columns ignored).  Records are zeroed, then bond_scale = 1,
 * fuse_tip = 0, id = index + 1, cpu = 0 (mirrored into idata).  records ==
 * NULL queries the count. */
bmx_status bmx_read_particles_ascii(const char *path, void *records,
                                    int64_t capacity, int64_t *n_out,
                                    const bmx_layout *layout);
bmx_status bmx_write_particles_ascii(const char *path, const void *records,
                                     int64_t n, const bmx_layout *layout);
/* E[4/3 pi r^3] for r uniform in [rmin, rmax] */
double bmx_mean_sphere_volume(double rmin, double rmax);
/* n spheres, radius uniform in [rmin, rmax], in a box of the given aspect
 * sized to volume fraction phi (lo = 0); periodic axes are filled
 * uniformly, other axes keep spheres inside [r, L - r].  overlap_free = 1
 * uses random sequential insertion. */
bmx_status bmx_generate_spheres(void *records, int64_t n, double rmin,
                                double rmax, double phi, uint64_t seed,
                                const double aspect[3],
                                const int32_t periodic[3],
                                int32_t overlap_free, double out_lo[3],
                                double out_hi[3], const bmx_layout *layout);
/* n rod segments grown as branching chains from nroots roots lying in the
 * plane above z_surface (each tree at its own height in [r, 3r] above it):
 * every tip extends straight, and with probability branch_prob a segment
 * also starts a side branch at +-[angle_min, angle_max] degrees.  Chained
 * ends coincide, so InitBonds bonds them. */
bmx_status bmx_generate_hyphae(void *records, int64_t n, int32_t nroots,
                               double seg_len, double seg_rad,
                               double branch_prob, double angle_min_deg,
                               double angle_max_deg, uint64_t seed,
                               const double box_lo[3], const double box_hi[3],
                               double z_surface, const bmx_layout *layout);

#ifdef __cplusplus
}
#endif
#endif
