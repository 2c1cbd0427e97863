// region.cpp -- prints Deck::finest_region and Deck::finest_geometry of a
// deck, or with "levels" every level's box and mesh (Deck::level_regions,
// Deck::level_geometry), one line each (host code only, no device call).
// Built and run by tests/test_deck_region.py.
#include <cstdio>
#include <cstring>

#include "bmx_container.h"

int main(int argc, char **argv) {
  if (argc != 2 && argc != 3) return 2;
  bmx::Deck deck = bmx::read_deck(argv[1]);
  if (argc == 3 && std::strcmp(argv[2], "levels") == 0) {
    const auto regs = deck.level_regions();
    for (int l = 0; l < (int)regs.size(); l++) {
      const bmx_geometry g = deck.level_geometry(l);
      std::printf("%.17g %.17g %.17g %.17g %.17g %.17g %d %d %d %d %d %d\n", g.prob_lo[0],
                  g.prob_lo[1], g.prob_lo[2], g.prob_hi[0], g.prob_hi[1], g.prob_hi[2],
                  g.n_cell[0], g.n_cell[1], g.n_cell[2], g.periodic[0], g.periodic[1],
                  g.periodic[2]);
    }
    return 0;
  }
  double lo[3], hi[3];
  deck.finest_region(lo, hi);
  const bmx_geometry g = deck.finest_geometry();
  std::printf("%.17g %.17g %.17g %.17g %.17g %.17g %d %d %d\n", lo[0], lo[1], lo[2],
              hi[0], hi[1], hi[2], g.n_cell[0], g.n_cell[1], g.n_cell[2]);
  return 0;
}
