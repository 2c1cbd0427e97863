// bmx_run.cpp -- runs the particle-mechanics path of a BMX input deck.
//
//   bmx_run <deck> [max_step] [out.dat]
//
// Mirrors the DEM part of the reference's time loop (src/main.cpp:313-342,
// src/timestepping/bmx_evolve.cpp:93-99): read the deck, InitParticlesAscii
// from particles.input_file (relative to the deck's directory), InitBonds at
// time 0, then EvolveParticles, split_particles (when the deck sets
// chem_species.max_vol), ParticleExchange and the fusion steps every step of
// bmx.fixed_dt.  The fluid step of bmx::Evolve is not part of this path.
// Prints one line per step with the time per step and, at the end, writes
// the particles in the input format.
//
// Levels.  The reference keeps each particle on the finest AMR level whose
// boxes hold it and builds each level's list from that level's particles,
// with that level's cells as bins (src/des/bmx_pc_interaction.cpp:22-28,
// Redistribute at :98-99 and :485-486).  While every particle lies in the
// finest level's box (Deck::finest_region) the run is one container on the
// finest mesh.  Otherwise it is one container per level, each on its level's
// mesh over its box (Deck::level_regions): a particle is placed on the
// finest level whose box holds it, every level runs the step on its own
// particles, and all particles are placed again after the step, each level
// binding its particles in id order.  New particles take ids from one
// sequence, and every level numbers its split and fusion draws by the step.
// The output then lists the particles in id order.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "bmx_container.h"

namespace {

// The finest level whose box holds position x (level 0 is the domain).
int level_of(const double x[3], const std::vector<std::array<double, 6>> &regs) {
  for (int l = (int)regs.size() - 1; l > 0; l--) {
    bool in = true;
    for (int d = 0; d < 3; d++) in = in && x[d] >= regs[l][d] && x[d] < regs[l][3 + d];
    if (in) return l;
  }
  return 0;
}

// Copies of the given records in (cpu, id) order.
std::vector<uint8_t> sorted_copy(std::vector<const uint8_t *> all) {
  const bmx_layout L = bmx_default_layout();
  auto id_of = [&](const uint8_t *r) {
    int32_t id, cpu;
    std::memcpy(&id, r + L.id_offset, sizeof id);
    std::memcpy(&cpu, r + L.cpu_offset, sizeof cpu);
    return std::make_pair(cpu, id);
  };
  std::stable_sort(all.begin(), all.end(),
                   [&](const uint8_t *a, const uint8_t *b) { return id_of(a) < id_of(b); });
  std::vector<uint8_t> out(all.size() * (size_t)L.stride_bytes);
  for (size_t i = 0; i < all.size(); i++)
    std::memcpy(out.data() + i * L.stride_bytes, all[i], L.stride_bytes);
  return out;
}

std::vector<uint8_t> sorted_by_id(const std::vector<uint8_t> &recs, int64_t n) {
  const bmx_layout L = bmx_default_layout();
  std::vector<const uint8_t *> all;
  for (int64_t i = 0; i < n; i++) all.push_back(recs.data() + (size_t)i * L.stride_bytes);
  return sorted_copy(all);
}

// Records of all levels, in id order.
std::vector<uint8_t> gather(std::vector<std::unique_ptr<bmx::BMXParticleContainer>> &lev) {
  const bmx_layout L = bmx_default_layout();
  std::vector<const uint8_t *> all;
  for (auto &c : lev) {
    const std::vector<uint8_t> &r = c->Records();
    for (int64_t i = 0; i < c->NumberOfParticles(); i++)
      all.push_back(r.data() + (size_t)i * L.stride_bytes);
  }
  return sorted_copy(all);
}

// Binds each level's particles (records in id order) to its container.
void place(const std::vector<uint8_t> &recs,
           const std::vector<std::array<double, 6>> &regs,
           std::vector<std::unique_ptr<bmx::BMXParticleContainer>> &lev) {
  const bmx_layout L = bmx_default_layout();
  const size_t n = recs.size() / L.stride_bytes;
  std::vector<std::vector<uint8_t>> part(regs.size());
  for (size_t i = 0; i < n; i++) {
    const uint8_t *r = recs.data() + i * L.stride_bytes;
    double x[3];
    std::memcpy(x, r + L.pos_offset, sizeof x);
    std::vector<uint8_t> &p = part[level_of(x, regs)];
    p.insert(p.end(), r, r + L.stride_bytes);
  }
  for (size_t l = 0; l < regs.size(); l++)
    lev[l]->InitParticles(part[l].data(), (int64_t)(part[l].size() / L.stride_bytes));
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <deck> [max_step] [out.dat]\n", argv[0]);
    return 2;
  }
  try {
    std::string deck_path = argv[1];
    bmx::Deck deck = bmx::read_deck(deck_path);
    int max_step = argc > 2 ? std::stoi(argv[2]) : deck.integer_or("bmx.max_step", 1);
    double dt = deck.real("bmx.fixed_dt");
    std::string dir = ".";
    size_t slash = deck_path.find_last_of('/');
    if (slash != std::string::npos) dir = deck_path.substr(0, slash);
    std::string input = deck.str("particles.input_file");
    if (!input.empty() && input[0] != '/') input = dir + "/" + input;

    bmx::BMXParticleContainer pc(deck);
    bmx_split_params sp;
    const bool split = deck.split_params(sp);
    pc.InitParticlesAscii(input);
    std::printf("Read %lld particles from %s\n", (long long)pc.NumberOfParticles(),
                input.c_str());
    double rlo[3], rhi[3];
    deck.finest_region(rlo, rhi);
    const std::vector<std::array<double, 6>> regs = deck.level_regions();
    // one container per level once a particle lies outside the finest box
    std::vector<std::unique_ptr<bmx::BMXParticleContainer>> lev;
    int64_t next_id = 0;
    auto by_level = [&]() {
      for (int l = 0; l < (int)regs.size(); l++)
        lev.emplace_back(new bmx::BMXParticleContainer(deck, l, 0));
      next_id = pc.NextID();
      place(sorted_by_id(pc.Records(), pc.NumberOfParticles()), regs, lev);
      std::printf("Levels: %d (particles outside the finest level's box run on "
                  "coarser levels)\n", (int)regs.size());
    };
    if (pc.CountOutside(rlo, rhi) > 0) by_level();
    double time = 0.0;
    for (int step = 0; step < max_step; step++) {
      auto t0 = std::chrono::steady_clock::now();
      int64_t total = 0;
      if (lev.empty()) {
        if (time == 0.0) pc.InitBonds();
        int nsub = 0;
        pc.EvolveParticles(dt, nsub);
        if (split) pc.split_particles(time);  // bmx_evolve.cpp:98
        pc.ParticleExchange(dt);              // bmx_evolve.cpp:99
        pc.Fusion();                          // bmx_evolve.cpp:100-104
        total = pc.NumberOfParticles();
        if (pc.CountOutside(rlo, rhi) > 0) by_level();
      } else {
        for (auto &c : lev) {
          if (c->NumberOfParticles() == 0) continue;
          if (time == 0.0) c->InitBonds();
          int nsub = 0;
          c->EvolveParticles(dt, nsub);
          c->SetNextID(next_id);
          c->SetDrawCalls((uint32_t)step, (uint32_t)step);
          if (split) c->split_particles(time);
          c->ParticleExchange(dt);
          c->Fusion();
          next_id = c->NextID();
        }
        place(gather(lev), regs, lev);  // Redistribute
        for (auto &c : lev) total += c->NumberOfParticles();
      }
      auto t1 = std::chrono::steady_clock::now();
      time += dt;
      std::printf("Step %d time %g particles %lld Time per step %.6f\n", step + 1, time,
                  (long long)total, std::chrono::duration<double>(t1 - t0).count());
    }
    if (argc > 3) {
      if (lev.empty()) {
        pc.WriteAsciiFileForInit(argv[3]);
      } else {
        const std::vector<uint8_t> all = gather(lev);
        bmx::BMXParticleContainer::WriteAscii(
            argv[3], all.data(), (int64_t)(all.size() / bmx_default_layout().stride_bytes));
      }
    }
  } catch (const bmx::Error &e) {
    std::fprintf(stderr, "bmx_run: %s (status %d)\n", e.what(), (int)e.code());
    return 1;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "bmx_run: %s\n", e.what());
    return 1;
  }
  return 0;
}
