// bmx_io.cpp -- host-side particle I/O and synthetic colony generators.
//
//   bmx_read_particles_ascii   BMXParticleContainer::InitParticlesAscii
//                              (src/des/bmx_pc_init.cpp:17-132): count line,
//                              then 27 fields per particle
//   bmx_write_particles_ascii  the same format (the reference's
//                              WriteAsciiFileForInit, src/des/bmx_pc.H:154)
//   bmx_generate_spheres       seeded random sequential insertion
//   bmx_generate_hyphae        seeded branching segment chains
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/bmx_mi355x.h"
#include "bmx_io.h"

namespace {

thread_local std::string g_io_error;

bmx_status io_fail(bmx_status st, const std::string &m) {
  g_io_error = m;
  bmx_io_set_error(m.c_str());
  return st;
}

constexpr double kPi = 3.14159265358979323846;

// splitmix64 stream: portable, seeded, identical on every host
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

struct Rec {
  uint8_t *base;
  bmx_layout L;
  double *pos(int64_t i) {
    return reinterpret_cast<double *>(base + i * L.stride_bytes + L.pos_offset);
  }
  double *rd(int64_t i) {
    return reinterpret_cast<double *>(base + i * L.stride_bytes + L.rdata_offset);
  }
  int32_t *id(int64_t i) {
    return reinterpret_cast<int32_t *>(base + i * L.stride_bytes + L.id_offset);
  }
  int32_t *cpu(int64_t i) {
    return reinterpret_cast<int32_t *>(base + i * L.stride_bytes + L.cpu_offset);
  }
  int32_t *idata(int64_t i) {
    return reinterpret_cast<int32_t *>(base + i * L.stride_bytes + L.idata_offset);
  }
};

// Defaults set for every new particle after its 27 fields are read
// (src/des/bmx_pc_init.cpp:87-100): bond_scale = 1, fuse_tip = 0,
// id/cpu mirrored into idata.
// What InitParticlesAscii sets beyond the file's fields
// (src/des/bmx_pc_init.cpp:87-99): the chemistry model's integer header
// (BMXChemistry::setIntegers, bmx_chem.cpp:50-60 and 76-83: 3 species, 0
// extra ints, 9 reals, 0 ints, increments at offset 6), bond_scale = 1,
// fuse_tip = 0, and the id and cpu.  The species start at zero here; the
// deck's fluid.init_conc_species is applied by the container.
void finish_record(Rec &r, int64_t i, int32_t pid) {
  int32_t *hdr = r.idata(i);
  hdr[0] = 3;
  hdr[1] = 0;
  hdr[2] = 9;
  hdr[3] = 0;
  hdr[4] = 6;
  r.rd(i)[24] = 1.0;
  r.idata(i)[25] = 0;
  *r.id(i) = pid;
  *r.cpu(i) = 0;
  r.idata(i)[33] = pid;
  r.idata(i)[34] = 0;
}

}  // namespace

extern "C" {

bmx_layout bmx_default_layout(void) {
  bmx_layout L;
  L.stride_bytes = 448;
  L.pos_offset = 0;
  L.rdata_offset = 24;
  L.id_offset = 296;
  L.cpu_offset = 300;
  L.idata_offset = 304;
  L.n_real = 34;
  L.n_int = 35;
  return L;
}

bmx_status bmx_read_particles_ascii(const char *path, void *records,
                                    int64_t capacity, int64_t *n_out,
                                    const bmx_layout *layout) {
  if (!path || !n_out) return io_fail(BMX_ERR_BAD_ARG, "null argument");
  bmx_layout L = layout ? *layout : bmx_default_layout();
  std::ifstream ifs(path);
  if (!ifs.good()) return io_fail(BMX_ERR_IO, std::string("cannot open ") + path);
  long long np = -1;
  ifs >> np;
  if (!ifs.good() || np < 0)
    return io_fail(BMX_ERR_IO,
                   "cannot read number of particles from the first line");
  *n_out = np;
  if (!records) return BMX_OK;  // size query
  if (capacity < np) return io_fail(BMX_ERR_BAD_ARG, "capacity too small");
  Rec r{static_cast<uint8_t *>(records), L};
  std::memset(records, 0, (size_t)np * L.stride_bytes);
  // pos[3], rdata[radius .. dvdt] (23 values), cell_type
  for (long long i = 0; i < np; i++) {
    double *p = r.pos(i), *rd = r.rd(i);
    for (int k = 0; k < 3; k++) ifs >> p[k];
    for (int k = 0; k <= 22; k++) ifs >> rd[k];
    int ct = -1;
    ifs >> ct;
    if (ifs.fail()) {
      char buf[96];
      std::snprintf(buf, sizeof buf, "error reading particle %lld", i);
      return io_fail(BMX_ERR_IO, buf);
    }
    r.idata(i)[5] = ct;
    // the rest of the line (extra columns some decks carry) is ignored
    std::string rest;
    std::getline(ifs, rest);
    finish_record(r, i, (int32_t)(i + 1));
  }
  return BMX_OK;
}

bmx_status bmx_write_particles_ascii(const char *path, const void *records,
                                     int64_t n, const bmx_layout *layout) {
  if (!path || (!records && n > 0)) return io_fail(BMX_ERR_BAD_ARG, "null argument");
  bmx_layout L = layout ? *layout : bmx_default_layout();
  FILE *f = std::fopen(path, "w");
  if (!f) return io_fail(BMX_ERR_IO, std::string("cannot write ") + path);
  Rec r{static_cast<uint8_t *>(const_cast<void *>(records)), L};
  std::fprintf(f, "%lld\n", (long long)n);
  for (int64_t i = 0; i < n; i++) {
    const double *p = r.pos(i), *rd = r.rd(i);
    for (int k = 0; k < 3; k++) std::fprintf(f, "%.17g ", p[k]);
    for (int k = 0; k <= 22; k++) std::fprintf(f, "%.17g ", rd[k]);
    std::fprintf(f, "%d\n", r.idata(i)[5]);
  }
  bool ok = std::fclose(f) == 0;
  return ok ? BMX_OK : io_fail(BMX_ERR_IO, "write failed");
}

double bmx_mean_sphere_volume(double rmin, double rmax) {
  if (rmax <= rmin) return 4.0 / 3.0 * kPi * rmin * rmin * rmin;
  return kPi * (std::pow(rmax, 4) - std::pow(rmin, 4)) / (3.0 * (rmax - rmin));
}

bmx_status bmx_generate_spheres(void *records, int64_t n, double rmin,
                                double rmax, double phi, uint64_t seed,
                                const double aspect[3],
                                const int32_t periodic[3], int32_t overlap_free,
                                double out_lo[3], double out_hi[3],
                                const bmx_layout *layout) {
  if (!records || n <= 0 || !(rmin > 0.0) || !(rmax >= rmin) ||
      !(phi > 0.0 && phi < 0.6) || !aspect || !periodic || !out_lo || !out_hi)
    return io_fail(BMX_ERR_BAD_ARG, "bad generator arguments");
  bmx_layout L = layout ? *layout : bmx_default_layout();
  Rec r{static_cast<uint8_t *>(records), L};
  std::memset(records, 0, (size_t)n * L.stride_bytes);
  double vbox = (double)n * bmx_mean_sphere_volume(rmin, rmax) / phi;
  double ap = aspect[0] * aspect[1] * aspect[2];
  double h = std::cbrt(vbox / ap);
  double len[3];
  for (int d = 0; d < 3; d++) {
    len[d] = aspect[d] * h;
    out_lo[d] = 0.0;
    out_hi[d] = len[d];
  }
  Rng rng(seed);
  // cell grid of edge >= 2 rmax for overlap tests
  int nc[3];
  double ce[3];
  for (int d = 0; d < 3; d++) {
    nc[d] = std::max(1, (int)std::floor(len[d] / (2.0 * rmax)));
    if (nc[d] > 4096) nc[d] = 4096;
    ce[d] = len[d] / nc[d];
  }
  int64_t ncell = (int64_t)nc[0] * nc[1] * nc[2];
  std::vector<int32_t> head(overlap_free ? ncell : 0, -1), next;
  if (overlap_free) next.assign((size_t)n, -1);
  auto cell_of = [&](const double *x, int c3[3]) {
    for (int d = 0; d < 3; d++) {
      int c = (int)std::floor(x[d] / ce[d]);
      c3[d] = c < 0 ? 0 : (c >= nc[d] ? nc[d] - 1 : c);
    }
    return ((int64_t)c3[2] * nc[1] + c3[1]) * nc[0] + c3[0];
  };
  int64_t placed = 0;
  int64_t attempts = 0, max_attempts = 400 * n + 100000;
  while (placed < n) {
    if (++attempts > max_attempts)
      return io_fail(BMX_ERR_BAD_ARG,
                     "random sequential insertion did not converge");
    double rad = rng.uniform(rmin, rmax);
    double x[3];
    for (int d = 0; d < 3; d++) {
      if (periodic[d])
        x[d] = rng.uniform(0.0, len[d]);
      else
        x[d] = rng.uniform(rad, len[d] - rad);  // inside the walls
    }
    if (overlap_free) {
      int c3[3];
      cell_of(x, c3);
      bool ok = true;
      for (int oz = -1; oz <= 1 && ok; oz++)
        for (int oy = -1; oy <= 1 && ok; oy++)
          for (int ox = -1; ox <= 1 && ok; ox++) {
            int cc[3] = {c3[0] + ox, c3[1] + oy, c3[2] + oz};
            bool skip = false;
            for (int d = 0; d < 3; d++) {
              if (cc[d] < 0 || cc[d] >= nc[d]) {
                if (periodic[d])
                  cc[d] = (cc[d] + nc[d]) % nc[d];
                else
                  skip = true;
              }
            }
            if (skip) continue;
            int64_t cid = ((int64_t)cc[2] * nc[1] + cc[1]) * nc[0] + cc[0];
            for (int32_t j = head[cid]; j >= 0 && ok; j = next[j]) {
              const double *q = r.pos(j);
              double s2 = 0.0;
              for (int d = 0; d < 3; d++) {
                double t = x[d] - q[d];
                if (periodic[d]) {
                  if (t > 0.5 * len[d]) t -= len[d];
                  if (t < -0.5 * len[d]) t += len[d];
                }
                s2 += t * t;
              }
              double rr = rad + r.rd(j)[0];
              if (s2 < rr * rr) ok = false;
            }
          }
      if (!ok) continue;
      int64_t cid = cell_of(x, c3);
      next[placed] = head[cid];
      head[cid] = (int32_t)placed;
    }
    double *p = r.pos(placed), *rd = r.rd(placed);
    for (int d = 0; d < 3; d++) p[d] = x[d];
    rd[0] = rad;                                 // radius
    rd[1] = rad;                                 // c_length = radius
    rd[4] = 4.0 * kPi * rad * rad;               // area
    rd[5] = 4.0 / 3.0 * kPi * rad * rad * rad;   // vol
    r.idata(placed)[5] = 0;                      // YEAST
    finish_record(r, placed, (int32_t)(placed + 1));
    placed++;
  }
  return BMX_OK;
}

bmx_status bmx_generate_hyphae(void *records, int64_t n, int32_t nroots,
                               double seg_len, double seg_rad,
                               double branch_prob, double angle_min_deg,
                               double angle_max_deg, uint64_t seed,
                               const double box_lo[3], const double box_hi[3],
                               double z_surface, const bmx_layout *layout) {
  if (!records || n <= 0 || nroots <= 0 || nroots > n || !(seg_len > 0.0) ||
      !(seg_rad > 0.0) || !box_lo || !box_hi)
    return io_fail(BMX_ERR_BAD_ARG, "bad generator arguments");
  bmx_layout L = layout ? *layout : bmx_default_layout();
  Rec r{static_cast<uint8_t *>(records), L};
  std::memset(records, 0, (size_t)n * L.stride_bytes);
  Rng rng(seed);
  double lx = box_hi[0] - box_lo[0], ly = box_hi[1] - box_lo[1];
  // Growth tips: (parent index, direction angle in the surface plane)
  struct Tip {
    int64_t seg;
    double ang;
  };
  std::vector<Tip> tips;
  std::vector<int> nb((size_t)n, 0);
  std::vector<double> root_z;
  auto set_seg = [&](int64_t i, double cx, double cy, double ang, double z) {
    double *p = r.pos(i), *rd = r.rd(i);
    // wrap into the periodic surface box
    cx = box_lo[0] + std::fmod(std::fmod(cx - box_lo[0], lx) + lx, lx);
    cy = box_lo[1] + std::fmod(std::fmod(cy - box_lo[1], ly) + ly, ly);
    p[0] = cx;
    p[1] = cy;
    p[2] = z;
    rd[0] = seg_rad;
    rd[1] = seg_len;
    rd[2] = 0.5 * kPi;  // theta: lying in the surface plane
    double a = std::fmod(std::fmod(ang, 2.0 * kPi) + 2.0 * kPi, 2.0 * kPi);
    rd[3] = a;          // phi
    rd[4] = 2.0 * kPi * seg_rad * seg_len;           // lateral area
    rd[5] = kPi * seg_rad * seg_rad * seg_len;       // volume
    r.idata(i)[5] = 1;  // FUNGI
    finish_record(r, i, (int32_t)(i + 1));
    r.idata(i)[19] = 0;  // TIP until it gets a child
  };
  int64_t cnt = 0;
  for (int k = 0; k < nroots && cnt < n; k++) {
    double cx = rng.uniform(box_lo[0], box_hi[0]);
    double cy = rng.uniform(box_lo[1], box_hi[1]);
    double ang = rng.uniform(0.0, 2.0 * kPi);
    // each tree lies at its own height just above the surface
    double z = z_surface + seg_rad + rng.uniform(0.0, 2.0 * seg_rad);
    set_seg(cnt, cx, cy, ang, z);
    tips.push_back({cnt, ang});
    cnt++;
  }
  // grow breadth-first: each tip extends straight ahead; with probability
  // branch_prob a side branch starts at the same end at +-[amin, amax]
  size_t head = 0;
  auto grow = [&](int64_t parent, double ang) -> int64_t {
    if (cnt >= n) return -1;
    const double *pp = r.pos(parent);
    double pa = r.rd(parent)[3];
    // parent's site-2 end, child's site-1 end coincide
    double ex = pp[0] + 0.5 * seg_len * std::cos(pa);
    double ey = pp[1] + 0.5 * seg_len * std::sin(pa);
    double cx = ex + 0.5 * seg_len * std::cos(ang);
    double cy = ey + 0.5 * seg_len * std::sin(ang);
    int64_t c = cnt++;
    set_seg(c, cx, cy, ang, pp[2]);
    r.idata(parent)[19] = 3;  // INTERIOR once extended
    return c;
  };
  while (cnt < n && head < tips.size()) {
    Tip t = tips[head++];
    int64_t c = grow(t.seg, t.ang);
    if (c < 0) break;
    tips.push_back({c, t.ang});
    if (cnt < n && rng.uniform() < branch_prob && nb[t.seg] < 2) {
      double a = rng.uniform(angle_min_deg, angle_max_deg) * kPi / 180.0;
      if (rng.uniform() < 0.5) a = -a;
      int64_t b = grow(t.seg, t.ang + a);
      if (b >= 0) {
        tips.push_back({b, t.ang + a});
        nb[t.seg]++;
      }
    }
  }
  if (cnt < n) return io_fail(BMX_ERR_BAD_ARG, "hyphae growth stalled");
  return BMX_OK;
}

}  // extern "C"
