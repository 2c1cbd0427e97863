// Host-only driver for the two untrusted-input parsers (the InitParticlesAscii
// reader in bmx_io.cpp and the particle checkpoint reader in
// bmx_checkpoint.cpp), built with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_parsers_sanitized.py and run over
// the shipped inputs and corrupted copies of them.  A parse error is an
// ordinary outcome (printed, exit 0); a sanitizer report ends the process
// with a non-zero status.
//
//   parse_harness ascii <file> <scratch file>
//   parse_harness chk <dir> <scratch dir> [real_map int_map]   (maps: "3,-1,4")
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bmx_mi355x.h"

static std::string g_err;
extern "C" void bmx_io_set_error(const char *msg) { g_err = msg ? msg : ""; }

static std::vector<int32_t> parse_map(const char *s) {
  std::vector<int32_t> m;
  while (*s) {
    char *end = nullptr;
    m.push_back((int32_t)std::strtol(s, &end, 10));
    s = (*end == ',') ? end + 1 : end;
    if (end == s && *s) break;
  }
  return m;
}

static int report(const char *what, bmx_status st, int64_t n) {
  std::printf("%s status=%d n=%lld %s\n", what, (int)st, (long long)n,
              st == BMX_OK ? "" : g_err.c_str());
  return 0;
}

// capacity cap: a header that claims more particles than this is read as a
// size query only
static const int64_t kCap = 200000;

static int run_ascii(const char *path, const char *scratch) {
  const bmx_layout L = bmx_default_layout();
  int64_t n = 0;
  bmx_status st = bmx_read_particles_ascii(path, nullptr, 0, &n, &L);
  report("ascii-size", st, n);
  if (st != BMX_OK || n > kCap) return 0;
  std::vector<uint8_t> recs((size_t)(n + 1) * L.stride_bytes);
  st = bmx_read_particles_ascii(path, recs.data(), n, &n, &L);
  report("ascii-read", st, n);
  if (st != BMX_OK) return 0;
  st = bmx_write_particles_ascii(scratch, recs.data(), n, &L);
  report("ascii-write", st, n);
  if (st != BMX_OK) return 0;
  int64_t n2 = 0;
  std::vector<uint8_t> again((size_t)(n + 1) * L.stride_bytes);
  st = bmx_read_particles_ascii(scratch, again.data(), n, &n2, &L);
  report("ascii-reread", st, n2);
  return (st == BMX_OK && n2 == n) ? 0 : 3;
}

static int run_chk(const char *dir, const char *scratch, const char *rm,
                   const char *im) {
  const bmx_layout L = bmx_default_layout();
  bmx_checkpoint_header h;
  bmx_status st = bmx_checkpoint_info(dir, &h);
  report("chk-info", st, st == BMX_OK ? h.n_particles : 0);
  if (st != BMX_OK) return 0;
  std::vector<int32_t> rmap, imap;
  if (rm && im) {
    rmap = parse_map(rm);
    imap = parse_map(im);
  }
  int64_t n = 0;
  st = bmx_read_checkpoint(dir, nullptr, 0, &n, &L, rm ? rmap.data() : nullptr,
                           (int32_t)rmap.size(), im ? imap.data() : nullptr,
                           (int32_t)imap.size());
  report("chk-size", st, n);
  if (st != BMX_OK || n > kCap || n < 0) return 0;
  std::vector<uint8_t> recs((size_t)(n + 1) * L.stride_bytes);
  st = bmx_read_checkpoint(dir, recs.data(), n, &n, &L, rm ? rmap.data() : nullptr,
                           (int32_t)rmap.size(), im ? imap.data() : nullptr,
                           (int32_t)imap.size());
  report("chk-read", st, n);
  if (st != BMX_OK) return 0;
  bmx_geometry g;
  std::memset(&g, 0, sizeof g);
  for (int d = 0; d < 3; d++) {
    g.prob_lo[d] = 0.0;
    g.prob_hi[d] = 1.0;
    g.n_cell[d] = 8;
  }
  st = bmx_write_checkpoint(scratch, recs.data(), n, &L, &g, 0, n + 1);
  report("chk-write", st, n);
  if (st != BMX_OK) return 0;
  int64_t n2 = 0;
  std::vector<uint8_t> again((size_t)(n + 1) * L.stride_bytes);
  st = bmx_read_checkpoint(scratch, again.data(), n, &n2, &L, nullptr, 0, nullptr, 0);
  report("chk-reread", st, n2);
  if (st != BMX_OK || n2 != n) return 3;
  // the written file carries every component, so records come back whole
  for (int64_t i = 0; i < n; i++)
    if (std::memcmp(recs.data() + i * L.stride_bytes, again.data() + i * L.stride_bytes,
                    (size_t)L.stride_bytes) != 0)
      return 4;
  return 0;
}

int main(int argc, char **argv) {
  if (argc >= 4 && std::strcmp(argv[1], "ascii") == 0) return run_ascii(argv[2], argv[3]);
  if ((argc == 4 || argc == 6) && std::strcmp(argv[1], "chk") == 0)
    return run_chk(argv[2], argv[3], argc == 6 ? argv[4] : nullptr,
                   argc == 6 ? argv[5] : nullptr);
  std::fprintf(stderr, "usage: parse_harness ascii <file> <scratch> | chk <dir> <scratch> [rmap imap]\n");
  return 2;
}
