// region.cpp -- prints Deck::finest_region and Deck::finest_geometry of a
// deck (host code only, no device call).  Built and run by
// tests/test_deck_region.py.
#include <cstdio>

#include "bmx_container.h"

int main(int argc, char **argv) {
  if (argc != 2) return 2;
  bmx::Deck deck = bmx::read_deck(argv[1]);
  double lo[3], hi[3];
  deck.finest_region(lo, hi);
  const bmx_geometry g = deck.finest_geometry();
  std::printf("%.17g %.17g %.17g %.17g %.17g %.17g %d %d %d\n", lo[0], lo[1], lo[2],
              hi[0], hi[1], hi[2], g.n_cell[0], g.n_cell[1], g.n_cell[2]);
  return 0;
}
