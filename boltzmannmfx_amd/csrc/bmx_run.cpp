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
// Prints
// one line per step with the time per step and, at the end, writes the
// particles in the input format.
#include <chrono>
#include <cstdio>
#include <string>

#include "bmx_container.h"

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
    // Every particle runs on the finest level here; the reference keeps the
    // particles outside the refined region on coarser levels, with their
    // own lists (src/des/bmx_pc_interaction.cpp:22-28).  Say so when any lie
    // there.
    double rlo[3], rhi[3];
    deck.finest_region(rlo, rhi);
    auto outside = [&](const char *when) {
      const int64_t k = pc.CountOutside(rlo, rhi);
      if (k > 0)
        std::printf("Warning: %lld particles %s lie outside the finest level's region "
                    "[%g %g %g] - [%g %g %g]; this driver keeps them on the finest "
                    "level\n", (long long)k, when, rlo[0], rlo[1], rlo[2], rhi[0],
                    rhi[1], rhi[2]);
    };
    outside("at the start");
    double time = 0.0;
    for (int step = 0; step < max_step; step++) {
      auto t0 = std::chrono::steady_clock::now();
      if (time == 0.0) pc.InitBonds();
      int nsub = 0;
      pc.EvolveParticles(dt, nsub);
      if (split) pc.split_particles(time);  // bmx_evolve.cpp:98
      pc.ParticleExchange(dt);              // bmx_evolve.cpp:99
      pc.Fusion();                          // bmx_evolve.cpp:100-104
      auto t1 = std::chrono::steady_clock::now();
      time += dt;
      std::printf("Step %d time %g particles %lld Time per step %.6f\n", step + 1, time,
                  (long long)pc.NumberOfParticles(),
                  std::chrono::duration<double>(t1 - t0).count());
    }
    outside("at the end");
    if (argc > 3) pc.WriteAsciiFileForInit(argv[3]);
  } catch (const bmx::Error &e) {
    std::fprintf(stderr, "bmx_run: %s (status %d)\n", e.what(), (int)e.code());
    return 1;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "bmx_run: %s\n", e.what());
    return 1;
  }
  return 0;
}
