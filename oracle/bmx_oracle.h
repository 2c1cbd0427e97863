/*
 * bmx_oracle.h -- CPU restatement of the BMX particle-mechanics path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  It is the checker that the HIP path is compared against
 * (tests/, __graft_entry__.smoke(), and the cpu_baseline leg of bench.py).
 *
 * It restates, in plain C99 over the reference's own 448-byte AoS particle
 * record, the functions the reference runs on its CPU path:
 *
 *   orc_max_interaction_distance  src/chemistry/bmx_cell_interaction_K.H:23-40
 *   orc_min_distance              src/chemistry/bmx_cell_interaction_K.H:51-221
 *   orc_evaluate_force            src/chemistry/bmx_cell_interaction_K.H:237-994
 *   orc_surface_force             src/chemistry/bmx_cell_interaction_K.H:1005-1283
 *   orc_fix_bonds                 src/chemistry/bmx_cell_interaction_K.H:1285-1444
 *   orc_update_particle           src/des/bmx_pc_interaction.cpp:317-436
 *   orc_force_pass                src/des/bmx_pc_interaction.cpp:170-279
 *   orc_state_evolve              src/des/bmx_pc_interaction.cpp:13-495
 *   orc_state_init_bonds          src/des/bmx_pc_interaction.cpp:500-668
 *   orc_bin_index                 src/interpolation/bmx_interp_K.H:55-56
 *   orc_build_neighbors           AMReX NeighborList over the +-1 bin stencil
 *                                 (un-vendored; semantics defined here)
 *   orc_state_exchange            src/des/bmx_pc_xchng.cpp:13-257
 *   orc_deposit                   src/deposition/bmx_deposition_K.H:64-151
 *   orc_volume_fraction           src/des/bmx_calc_volume_fraction.cpp:14-86
 *   orc_chem_update               src/chemistry/bmx_chem_K.H:1358-1940
 *   orc_mesh_to_particle          src/des/bmx_calc_txfr.cpp:72-396
 *   orc_split_particles           src/des/bmx_split_particles.cpp:16-166
 *   orc_fusion                    src/des/bmx_pc_fusion.cpp:16-493
 *
 * Every function is compiled with -ffp-contract=off so that each floating
 * point operation is rounded exactly as written, in the reference's order.
 */
#ifndef BMX_ORACLE_H
#define BMX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Particle record: AMReX Particle<34,35> as used by BMXParticleContainer
 * (src/des/bmx_pc.H:86-96, counts src/chemistry/bmx_chem.H:16-17). */
#define ORC_NREAL 34
#define ORC_NINT 35

typedef struct {
  double pos[3];
  double rdata[ORC_NREAL];
  int32_t id;
  int32_t cpu;
  int32_t idata[ORC_NINT];
} orc_particle;

/* realIdx (src/chemistry/bmx_chem.H:38-69) */
enum {
  ORC_R_RADIUS = 0, ORC_R_CLENGTH = 1, ORC_R_THETA = 2, ORC_R_PHI = 3,
  ORC_R_AREA = 4, ORC_R_VOL = 5, ORC_R_VELX = 6, ORC_R_VELY = 7,
  ORC_R_VELZ = 8, ORC_R_WX = 9, ORC_R_WY = 10, ORC_R_WZ = 11,
  ORC_R_GX = 18, ORC_R_GY = 19, ORC_R_GZ = 20, ORC_R_DADT = 21,
  ORC_R_DVDT = 22, ORC_R_BOND_SCALE = 24, ORC_R_FIRST_DATA = 25
};
/* intIdx (src/chemistry/bmx_chem.H:77-118) */
enum {
  ORC_I_CELL_TYPE = 5, ORC_I_N_BNDS = 6,
  ORC_I_SEG1_ID1 = 7, ORC_I_SEG2_ID1 = 8, ORC_I_SEG3_ID1 = 9, ORC_I_SEG4_ID1 = 10,
  ORC_I_SEG1_ID2 = 11, ORC_I_SEG2_ID2 = 12, ORC_I_SEG3_ID2 = 13, ORC_I_SEG4_ID2 = 14,
  ORC_I_SITE1 = 15, ORC_I_SITE2 = 16, ORC_I_SITE3 = 17, ORC_I_SITE4 = 18,
  ORC_I_POSITION = 19, ORC_I_FUSE_FLAG = 20, ORC_I_SPLIT_FLAG = 21,
  ORC_I_NEW_FLAG = 22, ORC_I_FUSE_TIP = 25,
  ORC_I_DLTD_SITE1 = 26, ORC_I_DLTD_SITE2 = 27, ORC_I_DLTD_ID1 = 28,
  ORC_I_DLTD_ID2 = 29, ORC_I_DLTD_CPU1 = 30, ORC_I_DLTD_CPU2 = 31,
  ORC_I_FIX_SITE = 32, ORC_I_ID = 33, ORC_I_CPU = 34
};
/* cellType / siteLocation (src/chemistry/bmx_chem.H:120-136) */
enum { ORC_YEAST = 0, ORC_FUNGI = 1 };
enum { ORC_TIP = 0, ORC_SECOND_1 = 1, ORC_SECOND_2 = 2, ORC_INTERIOR = 3 };

/* fpar layout (src/chemistry/bmx_cell_interaction.cpp:84-93). */
enum {
  ORC_F_WA = 0, ORC_F_MU = 1, ORC_F_ZWALL = 2, ORC_F_WZ = 3, ORC_F_XI = 4,
  ORC_F_GRAV = 5, ORC_F_KBOND = 6, ORC_F_BCUT = 7, ORC_F_VISC = 8
};

/* ---- pair / particle kernels ------------------------------------------ */
double orc_max_interaction_distance(const double *p1, const double *p2,
                                    const int32_t *i1, const int32_t *i2,
                                    double bndry_width);
void orc_min_distance(double xi1, double yi1, double zi1, double xi2,
                      double yi2, double zi2, double xj1, double yj1,
                      double zj1, double xj2, double yj2, double zj2,
                      double *tau_i, double *tau_j);
void orc_evaluate_force(const double *r12, double *par1, double *par2,
                        int32_t *ipar1, int32_t *ipar2, double *vel1,
                        double *vel2, double *rotv1, double *rotv2,
                        const double *fpar);
void orc_surface_force(const double *pos, double *par, int32_t *ipar,
                       double *vel, double *rotv, const double *fpar);
/* returns 0, or 1 when a bond would be written past the 4 bond slots */
int orc_fix_bonds(const double *r12, const double *ripar, int32_t *ipar,
                  const double *rjpar, int32_t *jpar, double rlim);
void orc_update_particle(orc_particle *p, const double fc[6], double subdt);

/* ---- geometry / binning ------------------------------------------------ */
typedef struct {
  double plo[3];
  double phi[3];
  int32_t periodic[3];
  int32_t nbin[3];   /* bins per dimension */
} orc_geom;

/* Periodic enforcement applied by Redistribute at the start and end of each
 * step: one shift by the domain length into [plo, phi). */
void orc_wrap(orc_particle *p, int64_t n, const orc_geom *g);
/* floor((x - plo) * dxi) per dimension, clamped to [0, nbin-1]; linear index
 * (bz*nby + by)*nbx + bx.  dxi = nbin / (phi - plo). */
int64_t orc_bin_index(const double *x, const orc_geom *g, int32_t b3[3]);

/* Neighbor candidate predicate:
 *   ORC_NBR_RADIUS  : |pos_i - image(pos_j)|^2 <= radius^2
 *   ORC_NBR_PAIR    : per-pair cutoff, see below
 *   ORC_NBR_STENCIL : |pos_i - image(pos_j)|^2 <= radius  (BMXCheckPair,
 *                     src/des/bmx_pc.H:63-81: squared distance vs the
 *                     'neighborhood' value)
 * Candidates are drawn from the +-1 bin stencil.  Rows are in canonical
 * order: stencil offsets (dz, dy, dx) lexicographic in -1..1, then ascending
 * particle index within a bin. */
enum { ORC_NBR_RADIUS = 0, ORC_NBR_STENCIL = 1, ORC_NBR_PAIR = 2 };
/* ORC_NBR_PAIR: |d|^2 <= (reach_i + reach_j + W_A + radius)^2 with reach =
 * radius (spheres) or 1.5 c_length (segments), 'radius' the skin and W_A
 * set by orc_set_pair_wa (the driver sets it from fpar).  The bins must be
 * no smaller than the largest such cutoff. */
void orc_set_pair_wa(double wa);

/* returns 0, -1 on allocation failure, -2 on bad geometry */
int orc_build_neighbors(const orc_particle *p, int64_t n, const orc_geom *g,
                        int mode, double radius, int64_t *offsets /* n+1 */,
                        int32_t **indices /* malloc'd, caller frees */);
/* O(N^2) brute force over all minimum images, rows ascending by index. */
int orc_build_neighbors_brute(const orc_particle *p, int64_t n,
                              const orc_geom *g, int mode, double radius,
                              int64_t *offsets, int32_t **indices);
/* minimum-image separation pos_i - image(pos_j) as the force loop forms it */
void orc_pair_separation(const double *xi, const double *xj, const orc_geom *g,
                         double d[3]);

/* ---- driver ------------------------------------------------------------ */
typedef struct orc_state orc_state;

orc_state *orc_state_create(orc_particle *p, int64_t n, const orc_geom *g,
                            const double fpar[9], int nbr_mode,
                            double nbr_radius, int rebuild_every);
void orc_state_destroy(orc_state *s);
/* one PCI force pass + update per step; fc_out (6*n, component-major) gets
 * the last step's accumulated fc if non-NULL.  Returns 0 or an error. */
int orc_state_evolve(orc_state *s, double dt, int nsteps, double *fc_out);
/* t = 0 bond detection over the neighbor list (rlim = 1e-7) */
int orc_state_init_bonds(orc_state *s);
/* ParticleExchange (src/des/bmx_pc_xchng.cpp:13-257; kernels
 * src/chemistry/bmx_chem_K.H:1951-2112): rebuild the list at the current
 * positions, evaluate every row, apply the increments.  xpar = {mtA, mtB,
 * mtC} (BMXChemistry::getExchangeParameters, bmx_chem.cpp:221-228). */
int orc_state_exchange(orc_state *s, const double xpar[3], double dt);
void orc_exchange_init(double *rdata, const int32_t *idata);
void orc_exchange_pair(double *rdata1, double *rdata2, const int32_t *idata1,
                       const int32_t *idata2, const double xpar[3], double dt);
void orc_exchange_apply(double *rdata, const int32_t *idata);
/* neighbor list used by the last step (CSR over particle indices) */
int64_t orc_state_neighbor_count(const orc_state *s);
void orc_state_get_neighbors(const orc_state *s, int64_t *offsets,
                             int32_t *indices);
int64_t orc_state_step(const orc_state *s);
/* overdamped mobility m (vel = m fc); 1 by default, as the reference */
void orc_state_set_mobility(orc_state *s, double m);
/* Gradient of L = (1/N) sum |min image(x_T - target)|^2 after nsteps sphere
 * steps (Verlet pair lists, rebuild_every, skin) with respect to mu, W_A and
 * gamma = 1/m (grad[0..2]) and every radius (grad_r), by the discrete
 * adjoint; p is advanced to x_T.  All states are kept (no recompute). */
int orc_adjoint(orc_particle *p, int64_t n, const orc_geom *g,
                const double fpar[9], double mobility, double dt, int nsteps,
                int rebuild_every, double skin, const double *target,
                double *loss, double grad[3], double *grad_r);
/* 0: sin, 1: cos, 2: acos of the shared implementation (bmx_math.h) */
double orc_math(int which, double x);
/* ---- particle <-> mesh (one level; the orc_geom's nbin is the mesh) ----
 * Output meshes: valid cells, x fastest, components slowest.  Input meshes:
 * one ghost layer, (n+2) per axis.  Particles outside the mesh are counted
 * (return value) and left out.  scheme 0: one-to-one, 1: trilinear. */
int64_t orc_deposit(const orc_particle *p, int64_t n, const orc_geom *g,
                    int scheme, int comp0, int ncomp, double *out);
int64_t orc_volume_fraction(const orc_particle *p, int64_t n, const orc_geom *g,
                            int scheme, double *vf);
int64_t orc_count_per_cell(const orc_particle *p, int64_t n, const orc_geom *g,
                           double *out);
int orc_chem_update(double grid_vol, double npart, double *cell_par,
                    double *mesh_vals, double *p_vals, double dt, int nloop,
                    const double *chempar, const int32_t *cell_ipar);
int orc_mesh_to_particle(orc_particle *p, int64_t n, const orc_geom *g,
                         const double *X, const double *vf,
                         const double chempar[10], double dt, int nloop,
                         int cnc_scheme, int vf_scheme);
/* interpolation of ncomp components of a ghosted field at x */
void orc_interp(const double *x, const double *f, const orc_geom *g, int scheme,
                int ncomp, double *vals);
/* split_particles (src/des/bmx_split_particles.cpp:16-166): records
 * p[0..n) with room for 'capacity'; sp = {max_vol, max_len, max_rad,
 * brnch_prob, split_prob, overlap, split_len}; random numbers from
 * bmx_uniform(seed, call, id, cpu, draw).  Returns the new count or -1. */
int64_t orc_split_particles(orc_particle *p, int64_t n, int64_t capacity,
                            const orc_geom *g, const double sp[7],
                            uint64_t seed, uint32_t call, int32_t next_id,
                            int32_t cpu);
/* EvaluateTipFusion / EvaluateInteriorFusion / CleanupFusion
 * (src/des/bmx_pc_fusion.cpp:16-493) over neighbour lists built like the
 * state's (nbr_mode, nbr_radius); fus = {fusion_probability,
 * max_fusion_separation, bond_scaling_increment, max_seg_length}.  *fused
 * tells whether a tip fused.  Returns the new count, -1 on failure, -2 if
 * capacity is short. */
int64_t orc_fusion(orc_particle *p, int64_t n, int64_t capacity,
                   const orc_geom *g, const double fpar[9], int nbr_mode,
                   double nbr_radius, const double fus[4], uint64_t seed,
                   uint32_t call, int32_t next_id, int32_t cpu, int *fused);
/* x^y of the shared implementation (bmx_math.h) */
double orc_pow(double x, double y);
/* OpenMP threads the force pass uses (1 without OpenMP) */
int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
