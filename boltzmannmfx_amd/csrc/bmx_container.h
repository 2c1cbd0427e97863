// bmx_container.h -- C++ host layer with the reference's particle-container
// entry points, over the C ABI of include/bmx_mi355x.h.
//
//   BMXParticleContainer::InitParticlesAscii   src/des/bmx_pc_init.cpp:17-132
//   BMXParticleContainer::InitBonds            src/des/bmx_pc_interaction.cpp:500-668
//   BMXParticleContainer::EvolveParticles      src/des/bmx_pc_interaction.cpp:13-495
//   BMXParticleContainer::split_particles      src/des/bmx_split_particles.cpp:16-166
//   BMXParticleContainer::EvaluateTipFusion +
//     EvaluateInteriorFusion + CleanupFusion   src/des/bmx_pc_fusion.cpp:16-493
//   BMXParticleContainer::WriteAsciiFileForInit src/des/bmx_pc.H:154
//   Deck / read_deck                           the ParmParse keys the path reads
//                                              (src/chemistry/bmx_cell_interaction.cpp:54-93,
//                                               src/chemistry/bmx_chem.cpp:114-124,
//                                               src/setup/bmx_init.cpp:57-79)
//
// The load-balancing cost arguments of EvolveParticles/InitBonds
// (src/des/bmx_pc_interaction.cpp:460-475) have no counterpart: one context
// holds one level on one device.
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bmx_mi355x.h"

namespace bmx {

class Error : public std::runtime_error {
 public:
  Error(bmx_status code, const std::string &msg)
      : std::runtime_error(msg), code_(code) {}
  bmx_status code() const { return code_; }

 private:
  bmx_status code_;
};

// The input-deck values the particle path uses.
struct Deck {
  std::map<std::string, std::vector<std::string>> kv;  // every key = value
  bool has(const std::string &k) const { return kv.count(k) != 0; }
  double real(const std::string &k, int i = 0) const;
  double real_or(const std::string &k, double dflt, int i = 0) const;
  int integer(const std::string &k, int i = 0) const;
  int integer_or(const std::string &k, int dflt, int i = 0) const;
  std::string str(const std::string &k) const;

  // derived
  bmx_geometry finest_geometry() const;  // level mesh of amr.max_level
  // The region the finest level covers when refinement is tagged by the
  // deck's bmx.tag_region box (bmx::ErrorEst, src/setup/bmx_tagging.cpp):
  // at each level the cells whose centres lie in the box, grown by
  // amr.n_error_buf, refined by amr.ref_ratio and aligned outward to
  // amr.blocking_factor, within the level below.  A model of AMReX's grid
  // generation for one tagged box; the whole domain without tagging or
  // with amr.max_level = 0.
  void finest_region(double lo[3], double hi[3]) const;
  // Every level's box under the same model, coarsest first, each as
  // {lo[3], hi[3]}: level 0 is the domain; the list ends at the finest
  // level that tagging creates.
  std::vector<std::array<double, 6>> level_regions() const;
  // The mesh of one level over its box: the level's cells, periodic along
  // an axis where the domain is and the box spans it.
  bmx_geometry level_geometry(int level) const;
  void force_params(double fpar[9]) const;
  double neighborhood() const;           // DEM::neighborhood
  // chem_species.mass_transfer_A/B/C (src/chemistry/bmx_chem.cpp:100-105)
  void exchange_params(double xpar[3]) const;
  // fluid.init_conc_species (FLUID::init_conc), zeros where absent
  void init_conc(double conc[3]) const;
  // chem_species.max_vol, max_seg_length, max_seg_radius,
  // branching_probability, splitting_probability, seg_split_length with the
  // reference's defaults (src/mods/bmx_chem_species_parms.cpp:55-66) and
  // the overlap 0.2 (src/chemistry/bmx_chem.cpp:110); false without max_vol
  bool split_params(bmx_split_params &p) const;
  // chem_species.fusion_probability, max_fusion_separation,
  // bond_scaling_increment, max_seg_length with the reference's defaults
  // (src/chemistry/bmx_chem.cpp:106-112)
  void fusion_params(bmx_fusion_params &p) const;
};

// Parses "key = v1 v2 ..." lines; '#' starts a comment; quotes are dropped.
Deck read_deck(const std::string &path);

class BMXParticleContainer {
 public:
  BMXParticleContainer(const bmx_geometry &geom, const double fpar[9],
                       const bmx_neighbor_params &nbp, int device = 0);
  // Container configured from a deck: the level mesh of amr.max_level, the
  // cell_force parameters and the reference's per-step stencil list.
  explicit BMXParticleContainer(const Deck &deck, int device = 0);
  // Container for one AMR level of a deck: that level's mesh over its box
  // (Deck::level_geometry), with the deck's parameters.
  BMXParticleContainer(const Deck &deck, int level, int device);
  ~BMXParticleContainer();
  BMXParticleContainer(const BMXParticleContainer &) = delete;
  BMXParticleContainer &operator=(const BMXParticleContainer &) = delete;

  // Reads the 27-field file; the species start at init_conc (the deck's
  // fluid.init_conc_species for a deck-built container, else zero).
  void InitParticlesAscii(const std::string &file);
  // Records supplied by the caller (host memory), e.g. a restart.
  void InitParticles(const void *records, int64_t n);
  void InitBonds();
  // One step of size dt; nsubsteps is set to 1 as in the reference
  // (dtsolid = dt, src/des/bmx_pc_interaction.cpp:56-65).
  void EvolveParticles(double dt, int &nsubsteps);
  // ParticleExchange (src/des/bmx_pc_xchng.cpp:13-257) with the transfer
  // rates of the deck (or set_exchange_params).
  void ParticleExchange(double dt);
  void set_exchange_params(const double xpar[3]);
  // Division and branching with the deck's (or set_split_params) limits;
  // new particles get ids from NextID (one past the largest id read) and
  // cpu 0.  Random numbers: bmx_uniform(seed, call, id, cpu), the call
  // counting the split steps of this container (seed: bmx.split_seed).
  void split_particles(double time);
  void set_split_params(const bmx_split_params &p, uint64_t seed = 0);
  // The fusion steps of bmx::Evolve (src/timestepping/bmx_evolve.cpp:100-104):
  // EvaluateTipFusion, and when a tip fused EvaluateInteriorFusion and
  // CleanupFusion; returns whether a tip fused.
  bool Fusion();
  void set_fusion_params(const bmx_fusion_params &p);
  void set_init_conc(const double conc[3]);
  // n steps without copying back in between.
  void EvolveSteps(double dt, int nsteps);
  void WriteAsciiFileForInit(const std::string &file);
  // The same format for records held by the caller (default layout).
  static void WriteAscii(const std::string &file, const void *records, int64_t n);
  // The particle part of a BMX checkpoint (pc->Checkpoint(name, "particles"),
  // src/io/bmx_chk.cpp:156-159; pc->Restart, src/io/bmx_restart.cpp:104-107):
  // dir is the "particles" directory, written at this container's level.
  void Checkpoint(const std::string &dir, int level = 0);
  void Restart(const std::string &dir);

  int64_t NumberOfParticles() const { return n_; }
  // NextID: the id the next new particle gets (one past the largest id
  // bound, advanced by splits and fusions); containers of the levels of one
  // run share one sequence through these.
  int64_t NextID() const { return next_id_; }
  void SetNextID(int64_t id) { next_id_ = id; }
  // The call numbers of the next split_particles and Fusion draws (each
  // counts its own calls from 0); the levels of one run number them by step.
  void SetDrawCalls(uint32_t split_call, uint32_t fusion_call) {
    split_calls_ = split_call;
    fusion_calls_ = fusion_call;
  }
  // Host copy of the records, brought up to date.
  const std::vector<uint8_t> &Records();
  // Particles whose position lies outside [lo, hi).
  int64_t CountOutside(const double lo[3], const double hi[3]);
  bmx_ctx *context() { return ctx_; }

 private:
  void check(bmx_status st) const;
  bmx_ctx *ctx_ = nullptr;
  bmx_geometry geom_{};
  double xpar_[3] = {0.0, 0.0, 0.0};
  double conc_[3] = {0.0, 0.0, 0.0};
  bmx_split_params sp_{};
  bool have_sp_ = false;
  bmx_fusion_params fp_{0.0, 5.0e-4, 1.0e-3, 3.5e-3};
  uint32_t fusion_calls_ = 0;
  uint64_t seed_ = 0;
  uint32_t split_calls_ = 0;
  int64_t next_id_ = 1;
  std::vector<uint8_t> host_;
  int64_t n_ = 0;
  bool dirty_ = false;
};

}  // namespace bmx
