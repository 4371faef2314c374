// Seeded synthetic block systems for the solve-phase workloads.
//
// The reference builds its block systems from BCM physics (Bcmlib, outside
// the solve phase); these generators produce systems of the same block shape
// and sparsity (block rows on a 2D/3D cell grid, one block per neighbour) so
// that the CPU and GPU paths see byte-identical inputs. Random numbers come
// from std::mt19937_64 with Box-Muller normals, drawn in a fixed order.
#include "bcm_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <random>

namespace bcm {

namespace {

struct Normal {
  std::mt19937_64 g;
  bool has = false;
  double spare = 0.0;
  std::vector<double>* record = nullptr;     // appends every draw when set
  const std::vector<double>* tape = nullptr;  // replays recorded draws when set
  size_t pos = 0;
  explicit Normal(uint64_t seed) : g(seed) {}
  double uniform() { return ((double)(g() >> 11) + 0.5) * 0x1.0p-53; }
  double draw() {
    if (has) { has = false; return spare; }
    const double u1 = uniform(), u2 = uniform();
    const double r = std::sqrt(-2.0 * std::log(u1));
    const double t = 6.283185307179586 * u2;
    spare = r * std::sin(t);
    has = true;
    return r * std::cos(t);
  }
  double operator()() {
    if (tape) return (*tape)[pos++];
    const double v = draw();
    if (record) record->push_back(v);
    return v;
  }
};

// The random template of an elasticity cell depends on (m, seed) only, so
// its normal draws are recorded once and replayed for every cell of the
// same template (families of cells differ only in their physics
// parameters). Bounded to kTapeBytes of host memory in total.
constexpr size_t kTapeBytes = size_t(2) << 30;
std::mutex g_tape_mu;
std::map<std::pair<int, uint64_t>, std::shared_ptr<const std::vector<double>>> g_tapes;
size_t g_tape_bytes = 0;

double harmonic_mean(double a, double b) { return 2.0 * a * b / (a + b); }

// Assembles block rows from (row, col) -> block maps into CSR form.
struct Assembler {
  int nsup, bla;
  std::vector<std::map<int, std::vector<double>>> rows;
  Assembler(int n, int b) : nsup(n), bla(b), rows(n) {}
  std::vector<double>& block(int i, int j) {
    auto& blk = rows[i][j];
    if (blk.empty()) blk.assign((size_t)bla * bla, 0.0);
    return blk;
  }
  // Adds diagonal blocks = scale * M M^T / bla + diag(rowabs + margin).
  void add_diagonals(Normal& rnd, const std::vector<double>& scale, const std::vector<double>& margin) {
    const int b = bla;
    std::vector<double> m((size_t)b * b), rowabs(b);
    for (int i = 0; i < nsup; ++i) {
      for (auto& v : m) v = rnd();
      std::fill(rowabs.begin(), rowabs.end(), 0.0);
      for (auto& kv : rows[i]) {
        if (kv.first == i) continue;
        for (int c = 0; c < b; ++c)
          for (int r = 0; r < b; ++r) rowabs[r] += std::fabs(kv.second[(size_t)c * b + r]);
      }
      auto& d = block(i, i);
      for (int c = 0; c < b; ++c)
        for (int r = 0; r < b; ++r) {
          double s = 0.0;
          for (int k = 0; k < b; ++k) s += m[(size_t)k * b + r] * m[(size_t)k * b + c];
          d[(size_t)c * b + r] = scale[i] * s / b;
        }
      for (int r = 0; r < b; ++r) d[(size_t)r * b + r] += rowabs[r] + margin[i];
    }
  }
  BlockCsr build(bool upper_only) const {
    BlockCsr a;
    a.nsup = nsup;
    a.sprnd.resize(nsup + 1);
    for (int i = 0; i <= nsup; ++i) a.sprnd[i] = i * bla;
    a.ia.assign(nsup + 1, 0);
    size_t nb = 0;
    for (int i = 0; i < nsup; ++i)
      for (auto& kv : rows[i])
        if (!upper_only || kv.first >= i) ++nb;
    a.ja.reserve(nb);
    a.bsa.reserve(nb);
    a.a.resize(nb * (size_t)bla * bla);
    int64_t off = 0;
    for (int i = 0; i < nsup; ++i) {
      for (auto& kv : rows[i]) {
        if (upper_only && kv.first < i) continue;
        a.ja.push_back(kv.first);
        a.bsa.push_back(off);
        std::memcpy(&a.a[off], kv.second.data(), sizeof(double) * bla * bla);
        off += (int64_t)bla * bla;
      }
      a.ia[i + 1] = (int)a.ja.size();
    }
    a.blamx = bla;
    return a;
  }
};

}  // namespace

void gen_heat2d(int nx, int ny, int bla, uint64_t seed, BlockCsr& upper, std::vector<double>& rhs) {
  if (nx <= 0 || ny <= 0 || bla <= 0) throw Error(kErrArg, "gen_heat2d: bad sizes");
  Normal rnd(seed);
  const int nsup = nx * ny;
  std::vector<double> k(nsup);
  for (auto& v : k) v = std::exp(rnd());
  Assembler as(nsup, bla);
  const int bb = bla * bla;
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      const int c = x + nx * y;
      const int nbr[2] = {x + 1 < nx ? c + 1 : -1, y + 1 < ny ? c + nx : -1};
      for (int t = 0; t < 2; ++t) {
        const int d = nbr[t];
        if (d < 0) continue;
        const double h = harmonic_mean(k[c], k[d]) / 3.0;
        auto& up = as.block(c, d);
        for (int e = 0; e < bb; ++e) up[e] = rnd() * h;
        auto& lo = as.block(d, c);
        for (int r = 0; r < bla; ++r)
          for (int q = 0; q < bla; ++q) lo[(size_t)q * bla + r] = up[(size_t)r * bla + q];
      }
    }
  std::vector<double> ones(nsup, 1.0);
  as.add_diagonals(rnd, ones, ones);
  upper = as.build(true);
  rhs.resize((size_t)nsup * bla);
  for (auto& v : rhs) v = rnd();
}

void gen_seepage3d(int nx, int ny, int nz, int bla, double skew, uint64_t seed, BlockCsr& a,
                   std::vector<double>& rhs) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || bla <= 0) throw Error(kErrArg, "gen_seepage3d: bad sizes");
  Normal rnd(seed);
  const int nsup = nx * ny * nz;
  const int bb = bla * bla;
  const double sc = 1.0 / std::sqrt((double)bla);
  Assembler as(nsup, bla);
  std::vector<double> s(bb), b1(bb), b2(bb);
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const int c = x + nx * (y + ny * z);
        const int nbr[3] = {x + 1 < nx ? c + 1 : -1, y + 1 < ny ? c + nx : -1,
                            z + 1 < nz ? c + nx * ny : -1};
        for (int t = 0; t < 3; ++t) {
          const int d = nbr[t];
          if (d < 0) continue;
          for (auto& v : s) v = rnd() * sc;
          for (auto& v : b1) v = rnd() * sc;
          for (auto& v : b2) v = rnd() * sc;
          auto& up = as.block(c, d);
          auto& lo = as.block(d, c);
          for (int col = 0; col < bla; ++col)
            for (int r = 0; r < bla; ++r) {
              const size_t e = (size_t)col * bla + r, et = (size_t)r * bla + col;
              up[e] = s[e] + skew * (b1[e] - b2[et]) * 0.5;
              lo[e] = s[et] + skew * (b2[e] - b1[et]) * 0.5;
            }
        }
      }
  std::vector<double> ones(nsup, 1.0);
  as.add_diagonals(rnd, ones, ones);
  a = as.build(false);
  rhs.resize((size_t)nsup * bla);
  for (auto& v : rhs) v = rnd();
}

void gen_elastic_cell(int m, const CellParams& cp, uint64_t template_seed, BlockCsr& upper,
                      std::vector<double>& rhs6) {
  if (m < 3) throw Error(kErrArg, "gen_elastic_cell: cell grid must be at least 3^3");
  const int bla = 24, nterm = 8, bb = bla * bla;
  const int nsup = m * m * m;
  Normal rnd(template_seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)m));
  std::shared_ptr<const std::vector<double>> tape;
  std::vector<double> recorded;
  {
    std::lock_guard<std::mutex> lk(g_tape_mu);
    auto it = g_tapes.find({m, template_seed});
    if (it != g_tapes.end()) tape = it->second;
  }
  if (tape) rnd.tape = tape.get();
  else rnd.record = &recorded;
  // Stiffness field: smoothed sphere of radius cp.radius around (0.5 + shift, 0.5, 0.5).
  std::vector<double> e(nsup);
  const double w = 0.5 / m;
  for (int z = 0; z < m; ++z)
    for (int y = 0; y < m; ++y)
      for (int x = 0; x < m; ++x) {
        auto per = [](double d) { d -= std::floor(d + 0.5); return d; };
        const double dx = per((x + 0.5) / m - 0.5 - cp.shift), dy = per((y + 0.5) / m - 0.5),
                     dz = per((z + 0.5) / m - 0.5);
        const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
        const double inside = 1.0 / (1.0 + std::exp(-(cp.radius - dist) / w));
        e[x + m * (y + m * z)] = 1.0 + (cp.contrast - 1.0) * inside;
      }
  const double nu = cp.poisson;
  const double cpl = 0.5 * nu / (1.0 - nu);
  // Stiffness assembled from one positive semidefinite element matrix per
  // cell face: K = h * (W_d o T T^T / bla) couples the two cells as
  // [K -K; -K K] (h: harmonic mean of the two stiffnesses, W_d: Poisson
  // coupling weights of the components for face direction d, T: template
  // block). The sum is a block graph Laplacian -- not diagonally dominant,
  // singular only along the cell-wise constant fields -- plus kMargin times
  // the local stiffness on the diagonal, so the system is SPD with the
  // conditioning of a discrete elliptic operator.
  constexpr double kMargin = 1e-3;
  Assembler as(nsup, bla);
  std::vector<double> t(bb), k(bb);
  const double sc = 1.0 / std::sqrt((double)bla);
  for (int z = 0; z < m; ++z)
    for (int y = 0; y < m; ++y)
      for (int x = 0; x < m; ++x) {
        const int c = x + m * (y + m * z);
        const int nbr[3] = {(x + 1) % m + m * (y + m * z), x + m * ((y + 1) % m + m * z),
                            x + m * (y + m * ((z + 1) % m))};
        for (int d = 0; d < 3; ++d) {
          const int o = nbr[d];
          for (auto& v : t) v = rnd() * sc;
          const double h = harmonic_mean(e[c], e[o]);
          for (int col = 0; col < bla; ++col)
            for (int r = col; r < bla; ++r) {
              double s2 = 0.0;
              for (int j = 0; j < bla; ++j) s2 += t[(size_t)j * bla + r] * t[(size_t)j * bla + col];
              const int ar = r / nterm, ac = col / nterm;
              const double wgt = ar == ac ? (ar == d ? 1.0 + 2.0 * cpl : 1.0) : cpl;
              k[(size_t)col * bla + r] = k[(size_t)r * bla + col] = h * wgt * s2 / bla;
            }
          auto& cc = as.block(c, c);
          auto& oo = as.block(o, o);
          auto& co = as.block(c, o);
          auto& oc = as.block(o, c);
          for (int i = 0; i < bb; ++i) {
            cc[i] += k[i];
            oo[i] += k[i];
            co[i] -= k[i];
            oc[i] -= k[i];
          }
        }
      }
  for (int c = 0; c < nsup; ++c) {
    auto& d = as.block(c, c);
    for (int r = 0; r < bla; ++r) d[(size_t)r * bla + r] += kMargin * e[c];
  }
  upper = as.build(true);
  if (!tape) {
    std::lock_guard<std::mutex> lk(g_tape_mu);
    const size_t bytes = recorded.size() * sizeof(double);
    if (g_tape_bytes + bytes <= kTapeBytes && !g_tapes.count({m, template_seed})) {
      g_tapes[{m, template_seed}] = std::make_shared<const std::vector<double>>(std::move(recorded));
      g_tape_bytes += bytes;
    }
  }
  // Macroscopic strains (Voigt order 11, 22, 33, 23, 13, 12): the load is the
  // jump of the stiffness across each cell times the strain's traction.
  static const double eps[6][3][3] = {
      {{1, 0, 0}, {0, 0, 0}, {0, 0, 0}}, {{0, 0, 0}, {0, 1, 0}, {0, 0, 0}},
      {{0, 0, 0}, {0, 0, 0}, {0, 0, 1}}, {{0, 0, 0}, {0, 0, .5}, {0, .5, 0}},
      {{0, 0, .5}, {0, 0, 0}, {.5, 0, 0}}, {{0, .5, 0}, {.5, 0, 0}, {0, 0, 0}}};
  rhs6.assign((size_t)6 * nsup * bla, 0.0);
  for (int z = 0; z < m; ++z)
    for (int y = 0; y < m; ++y)
      for (int x = 0; x < m; ++x) {
        const int c = x + m * (y + m * z);
        const int fwd[3] = {(x + 1) % m + m * (y + m * z), x + m * ((y + 1) % m + m * z),
                            x + m * (y + m * ((z + 1) % m))};
        const int bwd[3] = {(x + m - 1) % m + m * (y + m * z), x + m * ((y + m - 1) % m + m * z),
                            x + m * (y + m * ((z + m - 1) % m))};
        for (int k = 0; k < 6; ++k)
          for (int a = 0; a < 3; ++a) {
            double v = 0.0;
            for (int d = 0; d < 3; ++d) v += 0.5 * (e[fwd[d]] - e[bwd[d]]) * eps[k][d][a];
            for (int tt = 0; tt < nterm; ++tt)
              rhs6[(size_t)k * nsup * bla + (size_t)c * bla + a * nterm + tt] = v / (1.0 + tt);
          }
      }
}

}  // namespace bcm
