// bmx_container.cpp -- BMXParticleContainer over the C ABI, and the deck
// reader.
#include "bmx_container.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>

namespace bmx {

namespace {
constexpr double kPi = 3.14159265358979323846;

std::string trim(const std::string &s) {
  size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}
}  // namespace

Deck read_deck(const std::string &path) {
  std::ifstream ifs(path);
  if (!ifs.good()) throw Error(BMX_ERR_IO, "cannot open deck " + path);
  Deck d;
  std::string line;
  while (std::getline(ifs, line)) {
    size_t h = line.find('#');
    if (h != std::string::npos) line = line.substr(0, h);
    size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    std::string key = trim(line.substr(0, eq));
    std::string rest = line.substr(eq + 1);
    for (char &c : rest)
      if (c == '"') c = ' ';
    std::istringstream vs(rest);
    std::vector<std::string> vals;
    std::string v;
    while (vs >> v) vals.push_back(v);
    if (!key.empty()) d.kv[key] = vals;  // later entries override, as ParmParse::query
  }
  return d;
}

double Deck::real(const std::string &k, int i) const {
  auto it = kv.find(k);
  if (it == kv.end() || (int)it->second.size() <= i)
    throw Error(BMX_ERR_BAD_ARG, "deck key missing: " + k);
  return std::stod(it->second[i]);
}
double Deck::real_or(const std::string &k, double dflt, int i) const {
  return has(k) ? real(k, i) : dflt;
}
int Deck::integer(const std::string &k, int i) const {
  auto it = kv.find(k);
  if (it == kv.end() || (int)it->second.size() <= i)
    throw Error(BMX_ERR_BAD_ARG, "deck key missing: " + k);
  return std::stoi(it->second[i]);
}
int Deck::integer_or(const std::string &k, int dflt, int i) const {
  return has(k) ? integer(k, i) : dflt;
}
std::string Deck::str(const std::string &k) const {
  auto it = kv.find(k);
  if (it == kv.end() || it->second.empty())
    throw Error(BMX_ERR_BAD_ARG, "deck key missing: " + k);
  return it->second[0];
}

bmx_geometry Deck::finest_geometry() const {
  bmx_geometry g;
  int lev = integer_or("amr.max_level", 0);
  int ratio = integer_or("amr.ref_ratio", 2);
  for (int d = 0; d < 3; d++) {
    g.prob_lo[d] = real("geometry.prob_lo", d);
    g.prob_hi[d] = real("geometry.prob_hi", d);
    g.periodic[d] = integer_or("geometry.is_periodic", 0, d);
    int nc = integer("amr.n_cell", d);
    for (int l = 0; l < lev; l++) nc *= ratio;
    g.n_cell[d] = nc;
  }
  return g;
}

std::vector<std::array<double, 6>> Deck::level_regions() const {
  const int levels = integer_or("amr.max_level", 0);
  const int ratio = integer_or("amr.ref_ratio", 2);
  const int buf = integer_or("amr.n_error_buf", 1);
  const int bf = std::max(1, integer_or("amr.blocking_factor", 8));
  bool tag = false;
  if (has("bmx.tag_region")) {
    const std::string v = str("bmx.tag_region");
    tag = v == "true" || v == "1" || v == "True";
  }
  int64_t klo[3], khi[3], nc[3];
  double plo[3], phi[3];
  for (int d = 0; d < 3; d++) {
    plo[d] = real("geometry.prob_lo", d);
    phi[d] = real("geometry.prob_hi", d);
    nc[d] = integer("amr.n_cell", d);
    klo[d] = 0;
    khi[d] = nc[d] - 1;
  }
  auto box = [&]() {
    std::array<double, 6> r;
    for (int d = 0; d < 3; d++) {
      const double dx = (phi[d] - plo[d]) / (double)nc[d];
      r[d] = plo[d] + (double)klo[d] * dx;
      r[3 + d] = plo[d] + (double)(khi[d] + 1) * dx;
    }
    return r;
  };
  std::vector<std::array<double, 6>> out{box()};
  // without tagging every level covers the domain: the finest is the mesh of
  // amr.max_level over the whole domain (one level here)
  for (int l = 0; l < levels && tag; l++) {
    // tagged cells of level l, with ErrorEst's own test: the centre
    // (k + 0.5) dx (from 0, as bmx_tagging.cpp forms it) in [tlo, thi]
    int64_t a[3], b[3];
    bool any = true;
    for (int d = 0; d < 3; d++) {
      const double dx = (phi[d] - plo[d]) / (double)nc[d];
      const double tlo = real("bmx.tag_region_lo", d), thi = real("bmx.tag_region_hi", d);
      a[d] = khi[d] + 1;
      b[d] = klo[d] - 1;
      for (int64_t k = klo[d]; k <= khi[d]; k++) {
        const double c = ((double)k + 0.5) * dx;
        if (c >= tlo && c <= thi) {
          a[d] = std::min(a[d], k);
          b[d] = std::max(b[d], k);
        }
      }
      any = any && a[d] <= b[d];
    }
    if (!any) break;  // nothing tagged: the finest level is this one
    for (int d = 0; d < 3; d++) {
      // grown by the error buffer within level l, refined, aligned outward
      // to the blocking factor, within the domain
      const int64_t g0 = std::max(a[d] - buf, klo[d]);
      const int64_t g1 = std::min(b[d] + buf, khi[d]);
      nc[d] *= ratio;
      klo[d] = std::max<int64_t>((g0 * ratio) / bf * bf, 0);
      khi[d] = std::min<int64_t>(((g1 + 1) * ratio + bf - 1) / bf * bf - 1, nc[d] - 1);
    }
    out.push_back(box());
  }
  return out;
}

void Deck::finest_region(double lo[3], double hi[3]) const {
  const std::array<double, 6> r = level_regions().back();
  for (int d = 0; d < 3; d++) {
    lo[d] = r[d];
    hi[d] = r[3 + d];
  }
}

bmx_geometry Deck::level_geometry(int level) const {
  const std::vector<std::array<double, 6>> regs = level_regions();
  if (level < 0 || level >= (int)regs.size())
    throw Error(BMX_ERR_BAD_ARG, "no such level: " + std::to_string(level));
  const int ratio = integer_or("amr.ref_ratio", 2);
  bmx_geometry g;
  for (int d = 0; d < 3; d++) {
    const double plo = real("geometry.prob_lo", d), phi = real("geometry.prob_hi", d);
    int64_t nc = integer("amr.n_cell", d);
    for (int l = 0; l < level; l++) nc *= ratio;
    const double dx = (phi - plo) / (double)nc;
    g.prob_lo[d] = regs[level][d];
    g.prob_hi[d] = regs[level][3 + d];
    g.n_cell[d] = (int32_t)std::llround((g.prob_hi[d] - g.prob_lo[d]) / dx);
    // a level wraps across a periodic face only where it spans the domain
    g.periodic[d] = integer_or("geometry.is_periodic", 0, d) && (int64_t)g.n_cell[d] == nc;
  }
  return g;
}

// BMXCellInteraction::setParams (src/chemistry/bmx_cell_interaction.cpp:54-93)
void Deck::force_params(double fpar[9]) const {
  fpar[0] = real("cell_force.boundary_width");
  fpar[1] = real("cell_force.stiffness");
  fpar[2] = real_or("fluid.surface_location", 0.0);
  fpar[3] = real("cell_force.wall_boundary_width");
  fpar[4] = real("cell_force.wall_stiffness");
  fpar[5] = real_or("cell_force.gravity", 0.0);
  fpar[6] = real("cell_force.bond_strength");
  fpar[7] = real("cell_force.bond_cutoff");
  fpar[8] = real("cell_force.viscous_drag");
}

// DEM::neighborhood = 1.1 (2 r(max_vol) + neighbor_width)
// (src/chemistry/bmx_chem.cpp:114-124)
double Deck::neighborhood() const {
  double vol = real_or("chem_species.max_vol", 0.0);
  double radius = std::pow(3.0 * vol / (4.0 * kPi), 1.0 / 3.0);
  return 1.1 * (2.0 * radius + real("cell_force.neighbor_width"));
}

BMXParticleContainer::BMXParticleContainer(const bmx_geometry &geom,
                                           const double fpar[9],
                                           const bmx_neighbor_params &nbp,
                                           int device) {
  check(bmx_ctx_create(&ctx_, device));
  check(bmx_set_force_params(ctx_, fpar));
  check(bmx_set_geometry(ctx_, &geom));
  check(bmx_set_neighbor_params(ctx_, &nbp));
  geom_ = geom;
}

void BMXParticleContainer::Checkpoint(const std::string &dir, int level) {
  const std::vector<uint8_t> &recs = Records();
  bmx_layout L = bmx_default_layout();
  int64_t next_id = 1;
  for (int64_t i = 0; i < n_; i++) {
    int32_t id;
    std::memcpy(&id, recs.data() + (size_t)i * L.stride_bytes + L.id_offset, sizeof id);
    if ((int64_t)id + 1 > next_id) next_id = (int64_t)id + 1;
  }
  check(bmx_write_checkpoint(dir.c_str(), recs.data(), n_, &L, &geom_, level, next_id));
}

void BMXParticleContainer::Restart(const std::string &dir) {
  int64_t n = 0;
  bmx_layout L = bmx_default_layout();
  check(bmx_read_checkpoint(dir.c_str(), nullptr, 0, &n, &L, nullptr, 0, nullptr, 0));
  std::vector<uint8_t> recs((size_t)n * L.stride_bytes);
  check(bmx_read_checkpoint(dir.c_str(), recs.data(), n, &n, &L, nullptr, 0, nullptr, 0));
  InitParticles(recs.data(), n);
}

static bmx_neighbor_params stencil_params(const Deck &deck) {
  bmx_neighbor_params p;
  p.mode = BMX_NBR_STENCIL;
  p.rebuild_every = 1;
  p.max_neighbors = 256;
  p.bin_factor_pct = 100;
  p.skin = 0.0;
  p.neighborhood = deck.neighborhood();
  return p;
}

static std::vector<double> fpar_of(const Deck &deck) {
  std::vector<double> f(9);
  deck.force_params(f.data());
  return f;
}

BMXParticleContainer::BMXParticleContainer(const Deck &deck, int device)
    : BMXParticleContainer(deck.finest_geometry(), fpar_of(deck).data(),
                           stencil_params(deck), device) {
  deck.exchange_params(xpar_);
  deck.init_conc(conc_);
  have_sp_ = deck.split_params(sp_);
  deck.fusion_params(fp_);
  seed_ = (uint64_t)deck.integer_or("bmx.split_seed", 0);
}

BMXParticleContainer::BMXParticleContainer(const Deck &deck, int level, int device)
    : BMXParticleContainer(deck.level_geometry(level), fpar_of(deck).data(),
                           stencil_params(deck), device) {
  deck.exchange_params(xpar_);
  deck.init_conc(conc_);
  have_sp_ = deck.split_params(sp_);
  deck.fusion_params(fp_);
  seed_ = (uint64_t)deck.integer_or("bmx.split_seed", 0);
}

void Deck::fusion_params(bmx_fusion_params &p) const {
  p.fusion_probability = real_or("chem_species.fusion_probability", 0.0);
  p.max_fusion_separation = real_or("chem_species.max_fusion_separation", 5.0e-4);
  p.bond_scaling_increment = real_or("chem_species.bond_scaling_increment", 0.001);
  p.max_seg_length = real_or("chem_species.max_seg_length", 0.0035);
}

void BMXParticleContainer::set_fusion_params(const bmx_fusion_params &p) { fp_ = p; }

bool BMXParticleContainer::Fusion() {
  int32_t fused = 0;
  int64_t added = 0;
  // the fusion draws use calls of their own, after those of the splits
  check(bmx_fusion(ctx_, &fp_, seed_ ^ 0x66757369ull, fusion_calls_++, (int32_t)next_id_,
                   0, &fused, &added));
  if (added > 0) {
    next_id_ += added;
    n_ += added;
    host_.resize((size_t)n_ * (size_t)bmx_default_layout().stride_bytes);
  }
  if (fused) dirty_ = true;
  return fused != 0;
}

bool Deck::split_params(bmx_split_params &p) const {
  if (!has("chem_species.max_vol")) return false;
  p.max_vol = real("chem_species.max_vol");
  p.max_seg_length = real_or("chem_species.max_seg_length", 0.0035);
  p.seg_split_length = real_or("chem_species.seg_split_length", 0.0030);
  p.max_seg_radius = real_or("chem_species.max_seg_radius", 0.00025);
  p.branching_probability = real_or("chem_species.branching_probability", 0.0001);
  p.splitting_probability = real_or("chem_species.splitting_probability", 0.1);
  p.overlap = 0.2;
  return true;
}

void BMXParticleContainer::set_split_params(const bmx_split_params &p,
                                            uint64_t seed) {
  sp_ = p;
  have_sp_ = true;
  seed_ = seed;
}

void BMXParticleContainer::split_particles(double /*time*/) {
  if (!have_sp_) throw Error(BMX_ERR_STATE, "no split parameters (chem_species.max_vol)");
  int64_t added = 0;
  check(bmx_split_particles(ctx_, &sp_, seed_, split_calls_++, (int32_t)next_id_, 0,
                            &added));
  if (added > 0) {
    // NextID advances by the new particles (bmx_split_particles.cpp:72-81)
    next_id_ += added;
    n_ += added;
    host_.resize((size_t)n_ * (size_t)bmx_default_layout().stride_bytes);
    dirty_ = true;
  }
}

void Deck::exchange_params(double xpar[3]) const {
  xpar[0] = real_or("chem_species.mass_transfer_A", 0.0);
  xpar[1] = real_or("chem_species.mass_transfer_B", 0.0);
  xpar[2] = real_or("chem_species.mass_transfer_C", 0.0);
}

void Deck::init_conc(double conc[3]) const {
  for (int c = 0; c < 3; c++) {
    auto it = kv.find("fluid.init_conc_species");
    conc[c] = (it != kv.end() && (int)it->second.size() > c)
                  ? real("fluid.init_conc_species", c)
                  : 0.0;
  }
}

void BMXParticleContainer::set_exchange_params(const double xpar[3]) {
  for (int c = 0; c < 3; c++) xpar_[c] = xpar[c];
}

void BMXParticleContainer::set_init_conc(const double conc[3]) {
  for (int c = 0; c < 3; c++) conc_[c] = conc[c];
}

void BMXParticleContainer::ParticleExchange(double dt) {
  check(bmx_exchange(ctx_, xpar_, dt));
  dirty_ = true;
}

BMXParticleContainer::~BMXParticleContainer() {
  if (ctx_) bmx_ctx_destroy(ctx_);
}

void BMXParticleContainer::check(bmx_status st) const {
  if (st != BMX_OK) throw Error(st, bmx_last_error());
}

void BMXParticleContainer::InitParticlesAscii(const std::string &file) {
  int64_t n = 0;
  bmx_layout L = bmx_default_layout();
  check(bmx_read_particles_ascii(file.c_str(), nullptr, 0, &n, &L));
  std::vector<uint8_t> recs((size_t)n * L.stride_bytes);
  check(bmx_read_particles_ascii(file.c_str(), recs.data(), n, &n, &L));
  // rdata[first_data + c] = init_conc[c] (src/des/bmx_pc_init.cpp:87-91)
  for (int64_t i = 0; i < n; i++) {
    double *rd = reinterpret_cast<double *>(recs.data() + (size_t)i * L.stride_bytes +
                                            L.rdata_offset);
    for (int c = 0; c < 3; c++) rd[25 + c] = conc_[c];
  }
  InitParticles(recs.data(), n);
}

void BMXParticleContainer::InitParticles(const void *records, int64_t n) {
  bmx_layout L = bmx_default_layout();
  host_.assign(static_cast<const uint8_t *>(records),
               static_cast<const uint8_t *>(records) + (size_t)n * L.stride_bytes);
  n_ = n;
  check(bmx_bind_particles_aos(ctx_, host_.data(), n, &L, 0));
  dirty_ = false;
  next_id_ = 1;
  for (int64_t i = 0; i < n; i++) {
    int32_t id;
    std::memcpy(&id, host_.data() + (size_t)i * L.stride_bytes + L.id_offset, sizeof id);
    if ((int64_t)id + 1 > next_id_) next_id_ = (int64_t)id + 1;
  }
}

void BMXParticleContainer::InitBonds() {
  check(bmx_init_bonds(ctx_));
  dirty_ = true;
}

void BMXParticleContainer::EvolveParticles(double dt, int &nsubsteps) {
  nsubsteps = 1;
  check(bmx_evolve(ctx_, dt, 1));
  dirty_ = true;
}

void BMXParticleContainer::EvolveSteps(double dt, int nsteps) {
  check(bmx_evolve(ctx_, dt, nsteps));
  dirty_ = true;
}

const std::vector<uint8_t> &BMXParticleContainer::Records() {
  if (dirty_) {
    check(bmx_sync_to_aos(ctx_, host_.data()));
    dirty_ = false;
  }
  return host_;
}

int64_t BMXParticleContainer::CountOutside(const double lo[3], const double hi[3]) {
  const std::vector<uint8_t> &r = Records();
  const bmx_layout L = bmx_default_layout();
  int64_t out = 0;
  for (int64_t i = 0; i < n_; i++) {
    double x[3];
    std::memcpy(x, r.data() + i * L.stride_bytes + L.pos_offset, sizeof x);
    for (int d = 0; d < 3; d++)
      if (!(x[d] >= lo[d] && x[d] < hi[d])) {
        out++;
        break;
      }
  }
  return out;
}

void BMXParticleContainer::WriteAsciiFileForInit(const std::string &file) {
  const std::vector<uint8_t> &r = Records();
  bmx_layout L = bmx_default_layout();
  check(bmx_write_particles_ascii(file.c_str(), r.data(), n_, &L));
}

void BMXParticleContainer::WriteAscii(const std::string &file, const void *records,
                                      int64_t n) {
  bmx_layout L = bmx_default_layout();
  const bmx_status st = bmx_write_particles_ascii(file.c_str(), records, n, &L);
  if (st != BMX_OK) throw Error(st, bmx_last_error());
}

}  // namespace bmx
