// bmx_checkpoint.cpp -- the particle part of a BMX checkpoint.
//
// BMX writes and reads its particles with AMReX's particle checkpoint
// (pc->Checkpoint(name, "particles"), src/io/bmx_chk.cpp:156-159;
// pc->Restart(name, "particles"), src/io/bmx_restart.cpp:104-107).  AMReX is
// not part of the reference tree, so the format is restated here from the
// checkpoints the reference ships (exec/rst_test/chk20000,
// exec/fungi/chk357900, exec/nlev_large_test/dbg/chk300000):
//
//   <dir>/Header (text)
//     version            Version_Two_Dot_Zero_double | Version_Two_Dot_One_double
//     3                  space dimension
//     nr, nr names       real components after x, y, z
//     ni, ni names       int components after the id words
//     1                  checkpoint flag (ids are stored)
//     N                  number of particles
//     next_id
//     finest_level
//     finest_level + 1 lines: grids per level
//     one line per grid, level by level: file_number count offset
//   <dir>/Level_<l>/Particle_H   the level's BoxArray
//   <dir>/Level_<l>/DATA_<file>  per grid, at offset: count x (2 + ni)
//                                int32, then count x (3 + nr) doubles
//
// The two id words are (id, cpu) in Two_Dot_Zero.  In Two_Dot_One they are
// the high and low halves of one 64-bit word holding (id + 2^39) << 24 | cpu
// (decoded from exec/fungi/chk357900: its words equal idata[id]/idata[cpu]
// of every record).  Data are little-endian.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/bmx_mi355x.h"
#include "bmx_io.h"

namespace {

bmx_status ck_fail(bmx_status st, const std::string &m) {
  bmx_io_set_error(m.c_str());
  return st;
}

struct Grid {
  int level, file;
  int64_t count, offset;
};

struct Header {
  int version = 0;  // 20 or 21
  int nreal = 0, nint = 0;
  int64_t n = 0, next_id = 0;
  int finest = 0;
  std::vector<Grid> grids;
};

bmx_status read_header(const std::string &dir, Header &h) {
  std::ifstream is(dir + "/Header");
  if (!is.good()) return ck_fail(BMX_ERR_IO, "cannot open " + dir + "/Header");
  std::string ver;
  is >> ver;
  if (ver == "Version_Two_Dot_Zero_double")
    h.version = 20;
  else if (ver == "Version_Two_Dot_One_double")
    h.version = 21;
  else
    return ck_fail(BMX_ERR_IO, "unsupported particle checkpoint version " + ver);
  int dim = 0;
  is >> dim >> h.nreal;
  if (!is.good() || dim != 3 || h.nreal < 0 || h.nreal > 4096)
    return ck_fail(BMX_ERR_IO, "bad dimension or real component count");
  std::string name;
  for (int k = 0; k < h.nreal; k++) is >> name;
  is >> h.nint;
  if (!is.good() || h.nint < 0 || h.nint > 4096)
    return ck_fail(BMX_ERR_IO, "bad int component count");
  for (int k = 0; k < h.nint; k++) is >> name;
  int is_chk = 0;
  is >> is_chk >> h.n >> h.next_id >> h.finest;
  if (!is.good() || is_chk != 1 || h.n < 0 || h.finest < 0 || h.finest > 64)
    return ck_fail(BMX_ERR_IO, "not a particle checkpoint (or a corrupt header)");
  std::vector<int> ngrids(h.finest + 1);
  for (int l = 0; l <= h.finest; l++) is >> ngrids[l];
  int64_t tot = 0;
  for (int l = 0; l <= h.finest; l++)
    for (int g = 0; g < ngrids[l]; g++) {
      Grid gr{l, 0, 0, 0};
      is >> gr.file >> gr.count >> gr.offset;
      if (!is.good() || gr.count < 0 || gr.offset < 0 || gr.file < 0)
        return ck_fail(BMX_ERR_IO, "bad grid line in the header");
      tot += gr.count;
      h.grids.push_back(gr);
    }
  if (tot != h.n)
    return ck_fail(BMX_ERR_IO, "grid counts do not add up to the particle count");
  return BMX_OK;
}

struct RecW {
  uint8_t *base;
  bmx_layout L;
  uint8_t *at(int64_t i) { return base + i * (int64_t)L.stride_bytes; }
  double *pos(int64_t i) { return reinterpret_cast<double *>(at(i) + L.pos_offset); }
  double *rd(int64_t i) { return reinterpret_cast<double *>(at(i) + L.rdata_offset); }
  int32_t *id(int64_t i) { return reinterpret_cast<int32_t *>(at(i) + L.id_offset); }
  int32_t *cpu(int64_t i) { return reinterpret_cast<int32_t *>(at(i) + L.cpu_offset); }
  int32_t *idata(int64_t i) { return reinterpret_cast<int32_t *>(at(i) + L.idata_offset); }
};

void decode_ids(int version, int32_t w0, int32_t w1, int32_t &id, int32_t &cpu) {
  if (version == 20) {
    id = w0;
    cpu = w1;
    return;
  }
  const uint64_t v = ((uint64_t)(uint32_t)w0 << 32) | (uint64_t)(uint32_t)w1;
  id = (int32_t)((int64_t)(v >> 24) - ((int64_t)1 << 39));
  cpu = (int32_t)(v & 0xFFFFFFu);
}

void encode_ids(int32_t id, int32_t cpu, int32_t &w0, int32_t &w1) {
  const uint64_t v = ((uint64_t)((int64_t)id + ((int64_t)1 << 39)) << 24) |
                     ((uint64_t)(uint32_t)cpu & 0xFFFFFFu);
  w0 = (int32_t)(uint32_t)(v >> 32);
  w1 = (int32_t)(uint32_t)(v & 0xFFFFFFFFu);
}

bool make_dir(const std::string &p) {
  if (mkdir(p.c_str(), 0755) == 0) return true;
  struct stat sb;
  return stat(p.c_str(), &sb) == 0 && S_ISDIR(sb.st_mode);
}

std::string data_path(const std::string &dir, int level, int file) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "/Level_%d/DATA_%05d", level, file);
  return dir + buf;
}

}  // namespace

extern "C" {

bmx_status bmx_checkpoint_info(const char *dir, bmx_checkpoint_header *out) {
  if (!dir || !out) return ck_fail(BMX_ERR_BAD_ARG, "null argument");
  Header h;
  bmx_status st = read_header(dir, h);
  if (st != BMX_OK) return st;
  out->version = h.version;
  out->n_real = h.nreal;
  out->n_int = h.nint;
  out->finest_level = h.finest;
  out->n_particles = h.n;
  out->next_id = h.next_id;
  return BMX_OK;
}

bmx_status bmx_read_checkpoint(const char *dir, void *records, int64_t capacity,
                               int64_t *n_out, const bmx_layout *layout,
                               const int32_t *real_map, int32_t n_real_map,
                               const int32_t *int_map, int32_t n_int_map) {
  if (!dir || !n_out) return ck_fail(BMX_ERR_BAD_ARG, "null argument");
  Header h;
  bmx_status st = read_header(dir, h);
  if (st != BMX_OK) return st;
  *n_out = h.n;
  if (!records) return BMX_OK;  // size query
  if (capacity < h.n) return ck_fail(BMX_ERR_BAD_ARG, "capacity too small");
  const bmx_layout L = layout ? *layout : bmx_default_layout();
  if ((real_map && n_real_map > L.n_real) || (int_map && n_int_map > L.n_int) ||
      n_real_map < 0 || n_int_map < 0)
    return ck_fail(BMX_ERR_BAD_ARG, "component map longer than the record");
  // identity maps need the file to hold every record component
  if (!real_map && h.nreal < L.n_real)
    return ck_fail(BMX_ERR_BAD_ARG,
                   "the file has fewer real components than the record: pass a map");
  if (!int_map && h.nint < L.n_int)
    return ck_fail(BMX_ERR_BAD_ARG,
                   "the file has fewer int components than the record: pass a map");
  const int nrm = real_map ? n_real_map : L.n_real;
  const int nim = int_map ? n_int_map : L.n_int;
  for (int k = 0; k < nrm; k++)
    if (real_map && real_map[k] >= h.nreal)
      return ck_fail(BMX_ERR_BAD_ARG, "real map entry beyond the file's components");
  for (int k = 0; k < nim; k++)
    if (int_map && int_map[k] >= h.nint)
      return ck_fail(BMX_ERR_BAD_ARG, "int map entry beyond the file's components");
  RecW r{static_cast<uint8_t *>(records), L};
  std::memset(records, 0, (size_t)h.n * (size_t)L.stride_bytes);
  const int iw = 2 + h.nint, rw = 3 + h.nreal;
  std::vector<int32_t> ib;
  std::vector<double> rb;
  int64_t at = 0;
  for (const Grid &g : h.grids) {
    if (g.count == 0) continue;
    const std::string path = data_path(dir, g.level, g.file);
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return ck_fail(BMX_ERR_IO, "cannot open " + path);
    ib.resize((size_t)g.count * iw);
    rb.resize((size_t)g.count * rw);
    bool ok = std::fseek(f, (long)g.offset, SEEK_SET) == 0 &&
              std::fread(ib.data(), sizeof(int32_t), ib.size(), f) == ib.size() &&
              std::fread(rb.data(), sizeof(double), rb.size(), f) == rb.size();
    std::fclose(f);
    if (!ok) return ck_fail(BMX_ERR_IO, "short read in " + path);
    for (int64_t p = 0; p < g.count; p++, at++) {
      const int32_t *iv = &ib[(size_t)p * iw];
      const double *rv = &rb[(size_t)p * rw];
      int32_t id = 0, cpu = 0;
      decode_ids(h.version, iv[0], iv[1], id, cpu);
      *r.id(at) = id;
      *r.cpu(at) = cpu;
      double *pos = r.pos(at);
      pos[0] = rv[0];
      pos[1] = rv[1];
      pos[2] = rv[2];
      double *rd = r.rd(at);
      for (int k = 0; k < nrm; k++) {
        const int src = real_map ? real_map[k] : k;
        if (src >= 0) rd[k] = rv[3 + src];
      }
      int32_t *idv = r.idata(at);
      for (int k = 0; k < nim; k++) {
        const int src = int_map ? int_map[k] : k;
        if (src >= 0) idv[k] = iv[2 + src];
      }
    }
  }
  return BMX_OK;
}

bmx_status bmx_write_checkpoint(const char *dir, const void *records, int64_t n,
                                const bmx_layout *layout,
                                const bmx_geometry *geom, int32_t level,
                                int64_t next_id) {
  if (!dir || !geom || (!records && n > 0) || n < 0 || level < 0 || level > 64)
    return ck_fail(BMX_ERR_BAD_ARG, "bad checkpoint arguments");
  const bmx_layout L = layout ? *layout : bmx_default_layout();
  const std::string d(dir);
  char lev[32];
  std::snprintf(lev, sizeof lev, "/Level_%d", level);
  if (!make_dir(d) || !make_dir(d + lev))
    return ck_fail(BMX_ERR_IO, "cannot create " + d + lev);
  {
    std::ofstream hs(d + "/Header");
    if (!hs.good()) return ck_fail(BMX_ERR_IO, "cannot write " + d + "/Header");
    hs << "Version_Two_Dot_One_double\n3\n" << L.n_real << "\n";
    for (int k = 0; k < L.n_real; k++) hs << "real_comp" << k << "\n";
    hs << L.n_int << "\n";
    for (int k = 0; k < L.n_int; k++) hs << "int_comp" << k << "\n";
    hs << 1 << "\n" << n << "\n" << next_id << "\n" << level << "\n";
    for (int l = 0; l <= level; l++) hs << 1 << "\n";
    for (int l = 0; l < level; l++) hs << "0 0 0\n";
    hs << "0 " << n << " 0\n";
    if (!hs.good()) return ck_fail(BMX_ERR_IO, "cannot write " + d + "/Header");
  }
  {
    std::ofstream bs(d + lev + "/Particle_H");
    bs << "(1 0\n((0,0,0) (" << geom->n_cell[0] - 1 << "," << geom->n_cell[1] - 1
       << "," << geom->n_cell[2] - 1 << ") (0,0,0))\n)\n";
    if (!bs.good()) return ck_fail(BMX_ERR_IO, "cannot write Particle_H");
  }
  RecW r{static_cast<uint8_t *>(const_cast<void *>(records)), L};
  const int iw = 2 + L.n_int, rw = 3 + L.n_real;
  std::vector<int32_t> ib((size_t)n * iw);
  std::vector<double> rb((size_t)n * rw);
  for (int64_t p = 0; p < n; p++) {
    int32_t *iv = &ib[(size_t)p * iw];
    encode_ids(*r.id(p), *r.cpu(p), iv[0], iv[1]);
    std::memcpy(iv + 2, r.idata(p), sizeof(int32_t) * L.n_int);
    double *rv = &rb[(size_t)p * rw];
    std::memcpy(rv, r.pos(p), sizeof(double) * 3);
    std::memcpy(rv + 3, r.rd(p), sizeof(double) * L.n_real);
  }
  const std::string path = data_path(d, level, 0);
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) return ck_fail(BMX_ERR_IO, "cannot write " + path);
  bool ok = std::fwrite(ib.data(), sizeof(int32_t), ib.size(), f) == ib.size() &&
            std::fwrite(rb.data(), sizeof(double), rb.size(), f) == rb.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return ck_fail(BMX_ERR_IO, "short write to " + path);
  return BMX_OK;
}

}  // extern "C"
