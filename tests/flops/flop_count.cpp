// flop_count.cpp -- FP64 operation counts of one step of the segment path
// and of one mesh -> particle chemistry call, measured on the oracle's
// restatement of the reference's arithmetic.
//
// TEST INFRASTRUCTURE: built and run by tests/test_oracle_flops.py.  The
// oracle's source is compiled with every `double` replaced by a type that
// counts its +, -, *, / and sqrt (fma counts 2), so the counts follow the
// branches each particle and pair actually takes.
//
//   flop_count records n plo[3] phi[3] periodic[3] n_cell[3] fpar[9] hood
//     one segment step in the stencil policy: candidates and interacting
//     pairs of the step's lists, and the operations of each part of it
//   flop_count mesh records n plo[3] phi[3] periodic[3] n_cell[3] X vf
//              chempar[10] dt nloop cnc_scheme vf_scheme
//     one bmx_calc_txfr_particle over the records (X: 3 ghosted species
//     meshes, vf: the ghosted volume fraction, files of doubles)
//
// Records are Particle<34,35> (448 bytes) in a file.  Output: one JSON
// object.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

struct Flop {
  double v;
  static long long n;
  Flop() = default;
  Flop(double x) : v(x) {}
  explicit operator double() const { return v; }
  explicit operator int() const { return (int)v; }
  explicit operator long() const { return (long)v; }
  explicit operator long long() const { return (long long)v; }
  explicit operator unsigned long long() const { return (unsigned long long)v; }
};
long long Flop::n = 0;
#define BMX_FLOP_OP(o)                               \
  inline Flop operator o(Flop a, Flop b) {           \
    Flop::n++;                                       \
    return Flop(a.v o b.v);                          \
  }
BMX_FLOP_OP(+)
BMX_FLOP_OP(-)
BMX_FLOP_OP(*)
BMX_FLOP_OP(/)
#define BMX_FLOP_CMP(o) \
  inline bool operator o(Flop a, Flop b) { return a.v o b.v; }
BMX_FLOP_CMP(<)
BMX_FLOP_CMP(>)
BMX_FLOP_CMP(<=)
BMX_FLOP_CMP(>=)
BMX_FLOP_CMP(==)
BMX_FLOP_CMP(!=)
inline Flop operator-(Flop a) { return Flop(-a.v); }
inline Flop &operator+=(Flop &a, Flop b) { return a = a + b; }
inline Flop &operator-=(Flop &a, Flop b) { return a = a - b; }
inline Flop &operator*=(Flop &a, Flop b) { return a = a * b; }
inline Flop &operator/=(Flop &a, Flop b) { return a = a / b; }
inline Flop sqrt(Flop a) { Flop::n++; return Flop(std::sqrt(a.v)); }
inline Flop fabs(Flop a) { return Flop(std::fabs(a.v)); }
inline Flop floor(Flop a) { return Flop(std::floor(a.v)); }
inline Flop fma(Flop a, Flop b, Flop c) { Flop::n += 2; return Flop(std::fma(a.v, b.v, c.v)); }
inline Flop frexp(Flop a, int *e) { return Flop(std::frexp(a.v, e)); }
inline Flop ldexp(Flop a, int e) { return Flop(std::ldexp(a.v, e)); }
inline int isfinite(Flop a) { return std::isfinite(a.v); }
inline int isnan(Flop a) { return std::isnan(a.v); }

#define double Flop
#include "../../oracle/bmx_oracle.c"
#undef double

static Flop *read_doubles(const char *path, size_t count) {
  Flop *a = (Flop *)std::malloc(sizeof(Flop) * count);
  FILE *f = std::fopen(path, "rb");
  if (!a || !f) return nullptr;
  for (size_t i = 0; i < count; i++) {
    double v;
    if (std::fread(&v, sizeof v, 1, f) != 1) return nullptr;
    a[i] = Flop(v);
  }
  std::fclose(f);
  return a;
}

static int mesh_mode(int argc, char **argv) {
  if (argc != 32) {
    std::fprintf(stderr, "usage: flop_count mesh records n plo[3] phi[3] periodic[3] "
                         "n_cell[3] X vf chempar[10] dt nloop cnc vfs\n");
    return 2;
  }
  const long long n = std::atoll(argv[3]);
  orc_particle *p = (orc_particle *)std::malloc(sizeof(orc_particle) * (size_t)n);
  FILE *f = std::fopen(argv[2], "rb");
  if (!f || !p || std::fread(p, sizeof(orc_particle), (size_t)n, f) != (size_t)n) return 3;
  std::fclose(f);
  orc_geom g;
  int a = 4;
  for (int d = 0; d < 3; d++) g.plo[d] = std::atof(argv[a++]);
  for (int d = 0; d < 3; d++) g.phi[d] = std::atof(argv[a++]);
  for (int d = 0; d < 3; d++) g.periodic[d] = std::atoi(argv[a++]);
  for (int d = 0; d < 3; d++) g.nbin[d] = std::atoi(argv[a++]);
  const size_t ng = (size_t)(g.nbin[0] + 2) * (g.nbin[1] + 2) * (g.nbin[2] + 2);
  Flop *X = read_doubles(argv[a++], 3 * ng);
  Flop *vf = read_doubles(argv[a++], ng);
  if (!X || !vf) return 3;
  Flop chempar[10];
  for (int k = 0; k < 10; k++) chempar[k] = std::atof(argv[a++]);
  const Flop dt = std::atof(argv[a++]);
  const int nloop = std::atoi(argv[a++]);
  const int cnc = std::atoi(argv[a++]);
  const int vfs = std::atoi(argv[a++]);
  Flop::n = 0;
  const int err = orc_mesh_to_particle(p, n, &g, X, vf, chempar, dt, nloop, cnc, vfs);
  const long long f_call = Flop::n;
  Flop::n = 0;
  Flop c = bmx_pow(Flop(2.3e-13), Flop(1.0) / Flop(3.0));
  const long long f_pow = Flop::n - 1;
  // the general double-double path that the cube-root path falls back to
  Flop::n = 0;
  c = c + bmx_exp_dd(bmx_dd_mul_d(bmx_log_dd(Flop(2.3e-13)), Flop(1.0) / Flop(3.0)));
  (void)c;
  const long long f_pow_general = Flop::n - 2;
  std::printf("{\"particles\": %lld, \"error_bits\": %d, \"flop_mesh_to_particle\": %lld, "
              "\"flop_pow\": %lld, \"flop_pow_general\": %lld}\n", n, err, f_call, f_pow,
              f_pow_general);
  std::free(p);
  std::free(X);
  std::free(vf);
  return 0;
}

int main(int argc, char **argv) {
  if (argc > 1 && std::strcmp(argv[1], "mesh") == 0) return mesh_mode(argc, argv);
  if (argc != 25) {
    std::fprintf(stderr,
                 "usage: flop_count records n plo[3] phi[3] periodic[3] "
                 "n_cell[3] fpar[9] neighborhood\n");
    return 2;
  }
  const long long n = std::atoll(argv[2]);
  orc_particle *p = (orc_particle *)std::malloc(sizeof(orc_particle) * (size_t)n);
  FILE *f = std::fopen(argv[1], "rb");
  if (!f || !p || std::fread(p, sizeof(orc_particle), (size_t)n, f) != (size_t)n) return 3;
  std::fclose(f);
  orc_geom g;
  int a = 3;
  for (int d = 0; d < 3; d++) g.plo[d] = std::atof(argv[a++]);
  for (int d = 0; d < 3; d++) g.phi[d] = std::atof(argv[a++]);
  for (int d = 0; d < 3; d++) g.periodic[d] = std::atoi(argv[a++]);
  for (int d = 0; d < 3; d++) g.nbin[d] = std::atoi(argv[a++]);
  Flop fpar[9];
  for (int k = 0; k < 9; k++) fpar[k] = std::atof(argv[a++]);
  const double hood = std::atof(argv[a++]);
  orc_state *s = orc_state_create(p, n, &g, fpar, ORC_NBR_STENCIL, hood, 1);
  if (!s || orc_state_init_bonds(s) != 0) return 4;
  // the step's lists, as the step builds them
  Flop::n = 0;
  if (rebuild(s) != 0) return 5;
  const long long f_build = Flop::n;
  long long cand = 0, pairs = 0, f_accept = 0, f_pair = 0;
  long long f_wall = 0, f_update = 0;
  std::memcpy(s->snap, s->p, sizeof(orc_particle) * (size_t)n);
  Flop fc[6];
  for (long long i = 0; i < n; i++) {
    orc_particle *pi = &s->p[i];
    Flop acc[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t k = s->off[i]; k < s->off[i + 1]; k++) {
      const orc_particle *pj = &s->snap[s->idx[k]];
      cand++;
      long long m0 = Flop::n;
      Flop d[3];
      orc_pair_separation(pi->pos, pj->pos, &s->g, d);
      Flop r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      Flop r_lm = orc_max_interaction_distance(pi->rdata, pj->rdata, pi->idata,
                                               pj->idata, fpar[ORC_F_WA]);
      const bool in = r2 <= (r_lm - 1.0e-15) * (r_lm - 1.0e-15);
      f_accept += Flop::n - m0;
      if (in) {
        pairs++;
        Flop v1[3] = {0, 0, 0}, v2[3] = {0, 0, 0}, w1[3] = {0, 0, 0}, w2[3] = {0, 0, 0};
        orc_particle p2 = *pj;
        m0 = Flop::n;
        orc_evaluate_force(d, pi->rdata, p2.rdata, pi->idata, p2.idata, v1, v2, w1,
                           w2, fpar);
        for (int c = 0; c < 3; c++) {
          acc[c] += v1[c];
          acc[3 + c] += w1[c];
        }
        f_pair += Flop::n - m0;
      }
    }
    long long m0 = Flop::n;
    Flop vc[3] = {0, 0, 0}, vr[3] = {0, 0, 0};
    orc_surface_force(pi->pos, pi->rdata, pi->idata, vc, vr, fpar);
    for (int c = 0; c < 3; c++) {
      acc[c] += vc[c];
      acc[3 + c] += vr[c];
    }
    f_wall += Flop::n - m0;
    for (int c = 0; c < 6; c++) fc[c] = acc[c];
    m0 = Flop::n;
    orc_update_particle(pi, fc, Flop(1.0));
    f_update += Flop::n - m0;
  }
  // one sine and one cosine of a representative angle
  Flop::n = 0;
  Flop t = bmx_sin(Flop(1.1)) + bmx_cos(Flop(0.7));
  (void)t;
  const long long f_sincos = Flop::n - 1;
  // one segment's two-site mobility (rotation applied) and one pair of site
  // velocities from it: the pair law forms both segments' every time
  Flop R[3][3], dii[3][3], dij[3][3];
  seg_rotation(Flop(0.3), Flop(0.95), Flop(0.8), Flop(0.6), R);
  Flop::n = 0;
  seg_mobility(Flop(3.0e-6), Flop(3.0e-3), Flop(200.0), R, dii, dij);
  const long long f_mobility = Flop::n;
  Flop f1[3] = {1, 2, 3}, f2[3] = {3, 2, 1}, va[3], vb[3];
  Flop::n = 0;
  site_velocities(dii, dij, f1, f2, va, vb);
  const long long f_site = Flop::n;
  int fungi_pairs_trig = 8;  // cos and sin of both angles of both segments
  std::printf(
      "{\"particles\": %lld, \"candidates\": %lld, \"pairs\": %lld, "
      "\"flop_build\": %lld, \"flop_accept\": %lld, \"flop_pair\": %lld, "
      "\"flop_wall\": %lld, \"flop_update\": %lld, \"flop_sin_plus_cos\": %lld, "
      "\"trig_calls_per_pair\": %d, \"flop_mobility\": %lld, "
      "\"flop_site_velocities\": %lld}\n",
      n, cand, pairs, f_build, f_accept, f_pair, f_wall, f_update, f_sincos,
      fungi_pairs_trig, f_mobility, f_site);
  orc_state_destroy(s);
  std::free(p);
  return 0;
}
