// host.cpp -- the host side above the device boundary (include/ems_host.h):
// coefficient builder, row pooling, CPML layer builder, Gaussian waveform
// and the end-criterion run loop, restated from the reference in C++ with
// the same f64 expressions (compiled without FP contraction).
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/ems_gpu.h"
#include "../../include/ems_host.h"

namespace {

// src/constants.rs
constexpr double C0 = 299792458.0;
constexpr double MU0 = 1.25663706127e-6;
const double EPS0 = 1.0 / (MU0 * C0 * C0);
const double Z0 = MU0 * C0;
// CoeffRounding::OpenEms vacuum constants (src/constants.rs:29, :37)
constexpr double OPENEMS_EPS0 = 8.85418781762e-12;
constexpr double OPENEMS_MU0 = 1.256637062e-6;

struct Grid {
  const double *l[3];
  int n[3];
  // Grid::delta (grid.rs:562-566)
  double delta(int a, int i) const { return l[a][i + 1] - l[a][i]; }
  // Grid::dual_delta (grid.rs:592-597)
  double dual_delta(int a, int i) const {
    double lo = i > 0 ? delta(a, i - 1) : 0.0;
    double hi = i + 1 < n[a] ? delta(a, i) : 0.0;
    return 0.5 * (lo + hi);
  }
  size_t index(int i, int j, int k) const { return ((size_t)i * n[1] + j) * n[2] + k; }
};

Grid make_grid(const double *x, int nx, const double *y, int ny, const double *z, int nz) {
  Grid g;
  g.l[0] = x, g.l[1] = y, g.l[2] = z;
  g.n[0] = nx, g.n[1] = ny, g.n[2] = nz;
  return g;
}

// Axis::others (src/axis.rs): (next, next of next).
inline int nxt(int a) { return (a + 1) % 3; }
inline int prv(int a) { return (a + 2) % 3; }

struct Mat {
  double eps_r[3], mu_r[3], kappa[3], sigma_m[3];
};

template <class F>
void parallel_for(int n, int threads, F f) {
  if (threads <= 1 || n < 2) {
    for (int i = 0; i < n; i++) f(i);
    return;
  }
  threads = std::min(threads, n);
  std::vector<std::thread> ts;
  for (int t = 0; t < threads; t++)
    ts.emplace_back([=, &f]() {
      for (int i = t; i < n; i += threads) f(i);
    });
  for (auto &t : ts) t.join();
}

}  // namespace

extern "C" {

double ems_courant_limit(const double *x, int nx, const double *y, int ny, const double *z,
                         int nz) {
  const double *L[3] = {x, y, z};
  const int n[3] = {nx, ny, nz};
  double s = -0.0;  // Iterator::sum folds from -0.0
  for (int a = 0; a < 3; a++) {
    double m = INFINITY;  // Grid::min_delta (grid.rs:600-606)
    for (int i = 0; i + 1 < n[a]; i++) m = std::fmin(m, L[a][i + 1] - L[a][i]);
    s += 1.0 / (m * m);
  }
  return 1.0 / (C0 * std::sqrt(s));
}

extern "C++" {
// EdgeMaterials::from_cells for the E edge along a at node p
// (edge_material.rs:396-401): the cells' capacitance and conductance
// around it (primary_edge_cg, coeffs.rs:264-295) back to eps_r, kappa.
template <class Cell>
void edge_eps(const Grid &g, const Cell &cell, int a, const int p[3], double *eps_e,
              double *kap_e) {
  const int b = nxt(a), c = prv(a);
  const double l = g.delta(a, p[a]);
  double cap = 0.0, cond = 0.0;
  for (int cb : {p[b] - 1, p[b]}) {
    if (cb < 0 || cb >= g.n[b] - 1) continue;
    for (int cc : {p[c] - 1, p[c]}) {
      if (cc < 0 || cc >= g.n[c] - 1) continue;
      int ce[3];
      ce[a] = p[a], ce[b] = cb, ce[c] = cc;
      const double area = 0.25 * g.delta(b, cb) * g.delta(c, cc);
      const Mat &m = cell(ce);
      cap += EPS0 * m.eps_r[a] * area;
      cond += m.kappa[a] * area;
    }
  }
  cap = cap / l, cond = cond / l;
  const double kk = g.delta(a, p[a]) / (g.dual_delta(b, p[b]) * g.dual_delta(c, p[c]));
  *eps_e = cap * kk / EPS0;
  *kap_e = cond * kk;
}

// EdgeMaterials::from_cells for the H dual edge along a at node p
// (edge_material.rs:404-410): the series inductance and resistance of its
// two half cells (dual_edge_lr, coeffs.rs:303-325) back to mu_r, sigma_m.
template <class Cell>
void edge_mu(const Grid &g, const Cell &cell, int a, const int p[3], double *mu_e,
             double *sig_e) {
  const int b = nxt(a), c = prv(a);
  const double area = g.delta(b, p[b]) * g.delta(c, p[c]);
  double inv_l = 0.0, r_over_l2 = 0.0;
  for (int ca : {p[a] - 1, p[a]}) {
    if (ca < 0 || ca >= g.n[a] - 1) continue;
    int ce[3] = {p[0], p[1], p[2]};
    ce[a] = ca;
    const double half = 0.5 * g.delta(a, ca);
    const Mat &m = cell(ce);
    const double lc = MU0 * m.mu_r[a] * area / half;
    const double rc = m.sigma_m[a] * area / half;
    inv_l += 1.0 / lc;
    r_over_l2 += rc / (lc * lc);
  }
  const double L = 1.0 / inv_l;
  const double R = L * L * r_over_l2;
  const double kk = g.dual_delta(a, p[a]) / (g.delta(b, p[b]) * g.delta(c, p[c]));
  *mu_e = L * kk / MU0;
  *sig_e = R * kk;
}
}  // extern "C++"

int ems_coefficients_from_cells(const double *x, int nx, const double *y, int ny,
                                const double *z, int nz, const uint16_t *cell_mat,
                                const double *mat_table, int n_mat, const uint8_t *pec,
                                const int face_pec[6], double dt, int rounding, int threads,
                                float *lanes) {
  if (nx < 2 || ny < 2 || nz < 2 || !mat_table || n_mat < 1 || !lanes || !face_pec)
    return EMS_E_INVALID;
  if (!(dt > 0.0) || !std::isfinite(dt)) return EMS_E_INVALID;
  const Grid g = make_grid(x, nx, y, ny, z, nz);
  const size_t N = (size_t)nx * ny * nz;
  const int cn[3] = {nx - 1, ny - 1, nz - 1};
  std::vector<Mat> mats(n_mat);
  for (int m = 0; m < n_mat; m++) {
    const double *t = mat_table + 12 * m;
    for (int a = 0; a < 3; a++) {
      mats[m].eps_r[a] = t[a], mats[m].mu_r[a] = t[3 + a];
      mats[m].kappa[a] = t[6 + a], mats[m].sigma_m[a] = t[9 + a];
    }
  }
  if (cell_mat) {
    const size_t nc = (size_t)cn[0] * cn[1] * cn[2];
    for (size_t c = 0; c < nc; c++)
      if (cell_mat[c] >= n_mat) return EMS_E_INVALID;
  }
  auto cell = [&](const int c[3]) -> const Mat & {
    if (!cell_mat) return mats[0];
    return mats[cell_mat[((size_t)c[0] * cn[1] + c[1]) * cn[2] + c[2]]];
  };
  // CoeffRounding::store / eps0 / mu0 (coeffs.rs:735-760)
  auto store = [&](double v) { return rounding ? (double)(float)v : v; };
  const double eps0 = rounding ? OPENEMS_EPS0 : EPS0;
  const double mu0 = rounding ? OPENEMS_MU0 : MU0;
  std::memset(lanes, 0, 12 * N * sizeof(float));
  bool pec_face[3][2];
  for (int a = 0; a < 3; a++) pec_face[a][0] = face_pec[2 * a], pec_face[a][1] = face_pec[2 * a + 1];

  parallel_for(nx, threads, [&](int i) {
    for (int a = 0; a < 3; a++) {
      const int b = nxt(a), c = prv(a);
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
          const int p[3] = {i, j, k};
          const size_t idx = g.index(i, j, k);
          if (p[a] + 1 < g.n[a]) {
            const bool on_face = (p[b] == 0 && pec_face[b][0]) ||
                                 (p[b] == g.n[b] - 1 && pec_face[b][1]) ||
                                 (p[c] == 0 && pec_face[c][0]) ||
                                 (p[c] == g.n[c] - 1 && pec_face[c][1]);
            const bool pm = pec && pec[(size_t)a * N + idx];
            if (!(on_face || pm)) {
              double eps_e, kap_e;
              edge_eps(g, cell, a, p, &eps_e, &kap_e);
              // e_edge_geometry + e_coefficients (coeffs.rs:766-774, 817-820, 930-933)
              const double geo = g.dual_delta(b, p[b]) * g.dual_delta(c, p[c]) / g.delta(a, p[a]);
              const double C = store(eps0 * eps_e * geo);
              const double G = store(kap_e * geo);
              const double den = C + 0.5 * G * dt;
              lanes[(size_t)a * N + idx] = (float)((C - 0.5 * G * dt) / den);
              lanes[(size_t)(3 + a) * N + idx] = (float)(dt / den);
            }
          }
          if (p[b] + 1 < g.n[b] && p[c] + 1 < g.n[c]) {
            double mu_e, sig_e;
            edge_mu(g, cell, a, p, &mu_e, &sig_e);
            // coeffs.rs:939-955
            const double geo = g.delta(b, p[b]) * g.delta(c, p[c]) / g.dual_delta(a, p[a]);
            const double Ls = store(mu0 * mu_e * geo);
            const double Rs = store(sig_e * geo);
            const double den = Ls + 0.5 * Rs * dt;
            lanes[(size_t)(6 + a) * N + idx] = (float)((Ls - 0.5 * Rs * dt) / den);
            lanes[(size_t)(9 + a) * N + idx] = (float)(dt / den);
          }
        }
    }
  });
  return EMS_OK;
}

int64_t ems_raster_cells(const double *x, int nx, const double *y, int ny, const double *z,
                         int nz, const ems_solid *solid, uint16_t id, uint16_t *cell_mat) {
  if (!x || !y || !z || !solid || !cell_mat || nx < 2 || ny < 2 || nz < 2) return EMS_E_INVALID;
  if (solid->kind < 0 || solid->kind > 2 || (solid->kind == 2 && (solid->axis < 0 ||
                                                                   solid->axis > 2)))
    return EMS_E_INVALID;
  for (int q = 0; q < 6; q++)
    if (std::isnan(solid->p[q])) return EMS_E_INVALID;
  const double *l[3] = {x, y, z};
  const int nc[3] = {nx - 1, ny - 1, nz - 1};
  std::vector<double> c[3];
  for (int a = 0; a < 3; a++) {
    c[a].resize(nc[a]);
    for (int i = 0; i < nc[a]; i++) c[a][i] = 0.5 * (l[a][i] + l[a][i + 1]);
  }
  // per axis the solid's extent [lo, hi] (closed for the round shapes'
  // bounding box: its test decides) and the cells whose centres lie in it
  const double *p = solid->p;
  double lo[3], hi[3];
  int rad[2] = {0, 0};
  if (solid->kind == 0) {
    for (int a = 0; a < 3; a++) lo[a] = p[a], hi[a] = p[3 + a];
  } else if (solid->kind == 1) {
    for (int a = 0; a < 3; a++) lo[a] = p[a] - p[3], hi[a] = p[a] + p[3];
  } else {
    const int ax = solid->axis;
    rad[0] = ax == 0 ? 1 : 0, rad[1] = ax == 2 ? 1 : 2;
    lo[rad[0]] = p[0] - p[2], hi[rad[0]] = p[0] + p[2];
    lo[rad[1]] = p[1] - p[2], hi[rad[1]] = p[1] + p[2];
    lo[ax] = p[3], hi[ax] = p[4];
  }
  int i0[3], i1[3];
  for (int a = 0; a < 3; a++) {
    i0[a] = (int)(std::lower_bound(c[a].begin(), c[a].end(), lo[a]) - c[a].begin());
    i1[a] = (int)(std::upper_bound(c[a].begin(), c[a].end(), hi[a]) - c[a].begin());
    if (i1[a] <= i0[a]) return 0;
  }
  int64_t count = 0;
  for (int i = i0[0]; i < i1[0]; i++)
    for (int j = i0[1]; j < i1[1]; j++) {
      uint16_t *row = cell_mat + ((size_t)i * nc[1] + j) * nc[2];
      for (int k = i0[2]; k < i1[2]; k++) {
        const double cc[3] = {c[0][i], c[1][j], c[2][k]};
        bool in;
        if (solid->kind == 0) {
          in = true;
          for (int a = 0; a < 3; a++) in = in && lo[a] <= cc[a] && cc[a] < hi[a];
        } else if (solid->kind == 1) {
          const double dx = cc[0] - p[0], dy = cc[1] - p[1], dz = cc[2] - p[2];
          in = dx * dx + dy * dy + dz * dz < p[3] * p[3];
        } else {
          const int ax = solid->axis;
          const double d0 = cc[rad[0]] - p[0], d1 = cc[rad[1]] - p[1];
          in = d0 * d0 + d1 * d1 < p[2] * p[2] && p[3] <= cc[ax] && cc[ax] < p[4];
        }
        if (in) row[k] = id, count++;
      }
    }
  return count;
}

int ems_edge_materials_from_cells(const double *x, int nx, const double *y, int ny,
                                  const double *z, int nz, const uint16_t *cell_mat,
                                  const double *mat_table, int n_mat, int threads,
                                  double *eps_r, double *mu_r) {
  if (nx < 2 || ny < 2 || nz < 2 || !mat_table || n_mat < 1 || !eps_r || !mu_r)
    return EMS_E_INVALID;
  const Grid g = make_grid(x, nx, y, ny, z, nz);
  const size_t N = (size_t)nx * ny * nz;
  const int cn[3] = {nx - 1, ny - 1, nz - 1};
  std::vector<Mat> mats(n_mat);
  for (int m = 0; m < n_mat; m++) {
    const double *t = mat_table + 12 * m;
    for (int a = 0; a < 3; a++) {
      mats[m].eps_r[a] = t[a], mats[m].mu_r[a] = t[3 + a];
      mats[m].kappa[a] = t[6 + a], mats[m].sigma_m[a] = t[9 + a];
    }
  }
  if (cell_mat) {
    const size_t nc = (size_t)cn[0] * cn[1] * cn[2];
    for (size_t c = 0; c < nc; c++)
      if (cell_mat[c] >= n_mat) return EMS_E_INVALID;
  }
  auto cell = [&](const int c[3]) -> const Mat & {
    if (!cell_mat) return mats[0];
    return mats[cell_mat[((size_t)c[0] * cn[1] + c[1]) * cn[2] + c[2]]];
  };
  std::memset(eps_r, 0, 3 * N * sizeof(double));
  std::memset(mu_r, 0, 3 * N * sizeof(double));
  parallel_for(nx, threads, [&](int i) {
    for (int a = 0; a < 3; a++) {
      const int b = nxt(a), c = prv(a);
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
          const int p[3] = {i, j, k};
          const size_t idx = g.index(i, j, k);
          double v, w;
          if (p[a] + 1 < g.n[a]) {
            edge_eps(g, cell, a, p, &v, &w);
            eps_r[(size_t)a * N + idx] = v;
          }
          if (p[b] + 1 < g.n[b] && p[c] + 1 < g.n[c]) {
            edge_mu(g, cell, a, p, &v, &w);
            mu_r[(size_t)a * N + idx] = v;
          }
        }
    }
  });
  return EMS_OK;
}

int ems_pool_lane(const float *dense, int nx, int ny, int nz, float *rows_out,
                  uint32_t *index_out, uint64_t *nrows_out) {
  if (!dense || nx < 1 || ny < 1 || nz < 1) return EMS_E_INVALID;
  const size_t count = (size_t)nx * ny;
  const size_t limit = std::max<size_t>(count / 4, 1);  // POOL_RATIO (coeffs.rs:418)
  // hash (coeffs.rs:425-431): FNV-1a over the bit patterns
  auto hash = [&](const float *r) {
    uint64_t h = 0xcbf29ce484222325ull;
    for (int k = 0; k < nz; k++) {
      uint32_t b;
      std::memcpy(&b, r + k, 4);
      h = (h ^ (uint64_t)b) * 0x00000100000001b3ull;
    }
    return h;
  };
  std::unordered_map<uint64_t, std::vector<uint32_t>> seen;
  std::vector<float> pool;
  for (size_t r = 0; r < count; r++) {
    const float *row = dense + r * nz;
    auto &cands = seen[hash(row)];
    int64_t hit = -1;
    for (uint32_t c : cands)
      if (std::memcmp(pool.data() + (size_t)c * nz, row, nz * sizeof(float)) == 0) {
        hit = c;
        break;
      }
    if (hit < 0) {
      const size_t c = pool.size() / nz;
      if (c >= limit) return 0;
      pool.insert(pool.end(), row, row + nz);
      cands.push_back((uint32_t)c);
      hit = (int64_t)c;
    }
    if (index_out) index_out[r] = (uint32_t)hit;
  }
  if (rows_out) std::memcpy(rows_out, pool.data(), pool.size() * sizeof(float));
  if (nrows_out) *nrows_out = pool.size() / nz;
  return 1;
}

void ems_cpml_coefficients(const ems_cpml_params *p, double rho, double delta, double dt,
                           double n, double out[7]) {
  const double m = p->grading_order;
  const double g = std::pow(rho, m);
  // Cpml::sigma_max (fdtd.rs:266-271) with eta = n*Z0
  const double eta = n * Z0;
  const double sigma = p->sigma_factor * 0.8 * (m + 1.0) / (eta * delta) * g;
  const double kappa = 1.0 + (p->kappa_max - 1.0) * g;
  const double alpha = p->alpha_max * (1.0 - rho);
  const double b = std::exp(-(sigma / kappa + alpha) * dt / EPS0);
  const double den = sigma * kappa + kappa * kappa * alpha;
  const double x = (sigma / kappa + alpha) * dt / EPS0;
  double a, shift;
  if (!(sigma > 0.0 && den > 0.0)) {
    a = 0.0, shift = 0.0;
  } else if (p->convolution == 1 && x <= 8.0) {  // LINEAR_UP_TO (cpml.rs:337)
    const double cdt = sigma * dt / (EPS0 * kappa * kappa);
    const double gg =
        x < 1e-3 ? 0.5 - x / 3.0 + x * x / 8.0 : (-std::expm1(-x) - x * b) / (x * x);
    const double a1 = -cdt * gg;
    const double a0 = cdt * std::expm1(-x) / x - a1;
    a = a0 + a1 / b, shift = a1 / b;
  } else {
    a = sigma * (b - 1.0) / den, shift = 0.0;
  }
  out[0] = rho, out[1] = sigma, out[2] = kappa, out[3] = alpha, out[4] = b, out[5] = a;
  out[6] = 1.0 / kappa - 1.0 - shift;
}

double ems_cpml_layer_index(const double *x, int nx, const double *y, int ny, const double *z,
                            int nz, const float *lanes, double dt, int face, int cells) {
  const Grid g = make_grid(x, nx, y, ny, z, nz);
  const size_t N = (size_t)nx * ny * nz;
  const int p = face / 2;
  const bool is_max = face % 2 == 1;
  const int np = g.n[p];
  const int inner = is_max ? np - 1 - cells : cells;
  const int l0 = is_max ? inner : 0, l1 = is_max ? np : inner + 1;      // lines
  const int c0 = is_max ? inner : 0, c1 = is_max ? np - 1 : inner;      // cells
  double we = 0.0, se = 0.0, wh = 0.0, sh = 0.0;
  for (int comp : {nxt(p), prv(p)}) {
    const int b = nxt(comp), c = prv(comp);
    const float *ca = lanes + (size_t)comp * N, *cb = lanes + (size_t)(3 + comp) * N;
    const float *da = lanes + (size_t)(6 + comp) * N, *db = lanes + (size_t)(9 + comp) * N;
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
          const int node[3] = {i, j, k};
          const int along = node[p];
          if (!((along >= l0 && along < l1) || (along >= c0 && along < c1))) continue;
          const size_t idx = g.index(i, j, k);
          if (along >= l0 && along < l1 && node[comp] + 1 < g.n[comp] && cb[idx] != 0.0f) {
            const double area = g.dual_delta(b, node[b]) * g.dual_delta(c, node[c]);
            const double l = g.delta(comp, node[comp]);
            const double cap = dt * (1.0 + (double)ca[idx]) / (2.0 * (double)cb[idx]);
            const double vol = area * l;
            we += vol;
            se += vol * cap * l / (EPS0 * area);
          }
          if (along >= c0 && along < c1 && node[b] + 1 < g.n[b] && node[c] + 1 < g.n[c] &&
              db[idx] != 0.0f) {
            const double area = g.delta(b, node[b]) * g.delta(c, node[c]);
            const double l = g.dual_delta(comp, node[comp]);
            const double ind = dt * (1.0 + (double)da[idx]) / (2.0 * (double)db[idx]);
            const double vol = area * l;
            wh += vol;
            sh += vol * ind * l / (MU0 * area);
          }
        }
  }
  const double eps = we > 0.0 ? se / we : 1.0;
  const double mu = wh > 0.0 ? sh / wh : 1.0;
  const double n_r = std::sqrt(eps * mu);
  if (std::fabs(n_r - 1.0) < 1e-6 || !std::isfinite(n_r) || n_r <= 0.0) return 1.0;
  return n_r;
}

// Layer::new (cpml.rs:1122-1207): depth positions and grading of the E
// side (the layer's lines, outer face included, interface excluded) and the
// H side (the midpoints of its cells), in increasing index order.
static int layer_profiles(const Grid &g, int face, const ems_cpml_params *prm, double dt,
                          double n_face, std::vector<int> &e_idx, std::vector<int> &h_idx,
                          std::vector<double> &ec, std::vector<double> &hc) {
  const int p = face / 2;
  const bool is_max = face % 2 == 1;
  const int np = g.n[p];
  const int c = prm->cells;
  if (c < 1 || c > np - 1) return EMS_E_INVALID;
  const int inner = is_max ? np - 1 - c : c;
  auto rho = [&](double v) {
    return std::clamp(std::fabs(v - (double)inner) / (double)c, 0.0, 1.0);
  };
  e_idx.clear(), h_idx.clear();
  if (is_max) {
    for (int k = inner + 1; k < np; k++) e_idx.push_back(k);
    for (int k = inner; k < np - 1; k++) h_idx.push_back(k);
  } else {
    for (int k = 0; k < inner; k++) e_idx.push_back(k), h_idx.push_back(k);
  }
  auto node_delta = [&](int k) {
    const bool lo = k > 0, hi = k + 1 < np;
    if (lo && hi) return 0.5 * (g.delta(p, k - 1) + g.delta(p, k));
    return lo ? g.delta(p, k - 1) : g.delta(p, k);
  };
  ec.assign(7 * e_idx.size(), 0.0);
  hc.assign(7 * h_idx.size(), 0.0);
  for (size_t q = 0; q < e_idx.size(); q++)
    ems_cpml_coefficients(prm, rho((double)e_idx[q]), node_delta(e_idx[q]), dt, n_face, &ec[7 * q]);
  for (size_t q = 0; q < h_idx.size(); q++)
    ems_cpml_coefficients(prm, rho((double)h_idx[q] + 0.5), g.delta(p, h_idx[q]), dt, n_face,
                          &hc[7 * q]);
  return EMS_OK;
}

int ems_cpml_profile(const double *x, int nx, const double *y, int ny, const double *z, int nz,
                     int face, const ems_cpml_params *prm, double dt, double n_face, int e_side,
                     double *out, int *count) {
  if (face < 0 || face > 5 || !prm) return EMS_E_INVALID;
  const Grid g = make_grid(x, nx, y, ny, z, nz);
  std::vector<int> e_idx, h_idx;
  std::vector<double> ec, hc;
  int rc = layer_profiles(g, face, prm, dt, n_face, e_idx, h_idx, ec, hc);
  if (rc) return rc;
  const std::vector<double> &v = e_side ? ec : hc;
  if (out) std::memcpy(out, v.data(), v.size() * sizeof(double));
  if (count) *count = (int)(v.size() / 7);
  return EMS_OK;
}

int ems_cpml_layer(const double *x, int nx, const double *y, int ny, const double *z, int nz,
                   int face, const ems_cpml_params *prm, double dt, double n_face,
                   ems_cpml_slab out[4], float *coeff) {
  if (face < 0 || face > 5 || !prm || !out || !coeff) return EMS_E_INVALID;
  const Grid g = make_grid(x, nx, y, ny, z, nz);
  const int p = face / 2;
  const int np = g.n[p];
  std::vector<int> e_idx, h_idx;
  std::vector<double> ec, hc;
  int rc = layer_profiles(g, face, prm, dt, n_face, e_idx, h_idx, ec, hc);
  if (rc) return rc;
  const int a1 = nxt(p), a2 = prv(p);
  auto fill = [&](int q, int e_side, int comp, const std::vector<int> &idx,
                  const std::vector<double> &co) {
    ems_cpml_slab &s = out[q];
    s.e_side = e_side, s.comp = comp, s.normal = p;
    for (int d = 0; d < 3; d++) s.lo[d] = 0, s.hi[d] = g.n[d];
    if (e_side) {
      s.hi[comp] = g.n[comp] - 1;  // no padding
    } else {
      const int o = comp == a1 ? a2 : a1;
      s.hi[o] = g.n[o] - 1;
    }
    s.lo[p] = idx.empty() ? 0 : idx.front();
    s.hi[p] = idx.empty() ? 0 : idx.back() + 1;
    float *B = coeff + (size_t)q * 3 * np, *A = B + np, *K = A + np;
    for (size_t d = 0; d < idx.size(); d++) {
      B[d] = (float)co[7 * d + 4];
      A[d] = (float)co[7 * d + 5];
      K[d] = (float)co[7 * d + 6];
    }
    s.b = B, s.a = A, s.k = K;
  };
  fill(0, 1, a1, e_idx, ec);
  fill(1, 1, a2, e_idx, ec);
  fill(2, 0, a1, h_idx, hc);
  fill(3, 0, a2, h_idx, hc);
  return EMS_OK;
}

// waveform.rs:240-248 (gaussian_tau, gaussian_delay)
static double gaussian_tau(double fc) { return std::sqrt(M_LN10) / (M_PI * fc); }
static double gaussian_delay(double fc) { return gaussian_tau(fc) * std::sqrt(6.0 * M_LN10); }

double ems_gaussian(double f0, double fc, double t) {
  const double delay = gaussian_delay(fc);
  if (!(t >= 0.0 && t <= 2.0 * delay)) return 0.0;
  const double x = (t - delay) / gaussian_tau(fc);
  return std::cos(2.0 * M_PI * f0 * (t - delay)) * std::exp(-x * x);
}

double ems_gaussian_duration(double fc) { return 2.0 * gaussian_delay(fc); }

void ems_gaussian_values(double f0, double fc, double dt, double offset, uint64_t count,
                         double *out) {
  for (uint64_t n = 0; n < count; n++) out[n] = ems_gaussian(f0, fc, ((double)n + offset) * dt);
}

void ems_gaussian_table(double f0, double fc, double dt, double offset, uint64_t count,
                        float *amp) {
  for (uint64_t n = 0; n < count; n++)
    amp[n] = (float)ems_gaussian(f0, fc, ((double)n + offset) * dt);
}

// Waveform::OpenEmsGaussian (src/waveform.rs:454-466; oe_tau / oe_delay
// :260-268): tau = 3/(2 pi fc), T = 3 tau. Sampled (dt > 0): sample index
// k = floor(t/dt + 1/4 + 1e-9) (either grid), non-zero for
// 1 <= k <= floor(2T/dt + 1e-9); continuous (dt = 0): 0 < t <= 2T.
double ems_openems_gaussian(double f0, double fc, double t, double dt) {
  const double tau = 3.0 / (2.0 * M_PI * fc), delay = 3.0 * tau;
  bool inside;
  if (dt > 0.0) {
    const double k = std::floor(t / dt + 0.25 + 1e-9);
    inside = k >= 1.0 && k <= std::floor(2.0 * delay / dt + 1e-9);
  } else {
    inside = t > 0.0 && t <= 2.0 * delay;
  }
  if (!inside) return 0.0;
  const double x = (t - delay) / tau;
  return std::cos(2.0 * M_PI * f0 * (t - delay)) * std::exp(-x * x);
}

double ems_openems_gaussian_duration(double fc) { return 6.0 * 3.0 / (2.0 * M_PI * fc); }

// openems_signal / openems_lag (src/sim/map.rs:278-305): the amplitude of a
// model excitation with openEMS timing, tabulated per step at the engine's
// sample time of the source kind (SoftE (n+1/2)dt, HardE (n+1)dt, SoftH
// n dt, HardH (n+1/2)dt): the openEMS Gaussian at tau = t - lag - D dt
// (lag dt/2 for soft, dt for hard kinds; D = floor(delay/dt)), 0 for
// tau < -dt/4; as f32.
int ems_openems_source_table(double f0, double fc, double delay, int kind, double dt,
                             uint64_t count, float *amp) {
  if (kind < 0 || kind > 3 || !(dt > 0.0) || (count && !amp)) return EMS_E_INVALID;
  static const double offset[4] = {0.5, 1.0, 0.0, 0.5};
  const bool hard = kind == 1 || kind == 3;
  const double lag = hard ? dt : 0.5 * dt;
  const double steps = delay > 0.0 ? std::floor(delay / dt) : 0.0;
  const double shift = lag + steps * dt;
  for (uint64_t n = 0; n < count; n++) {
    const double t = ((double)n + offset[kind]) * dt;
    const double tau = t - shift;
    amp[n] = tau < -0.25 * dt ? 0.0f : (float)ems_openems_gaussian(f0, fc, tau, dt);
  }
  return EMS_OK;
}

// fold_parallel (src/engine/lumped.rs:275-282): g and c in parallel with the
// edge's recovered C = dt(1 + ca)/(2cb) and G = (1 - ca)/cb.
int ems_fold_parallel(float ca, float cb, double dt, double g, double c, float *ca_out,
                      float *cb_out) {
  if (cb == 0.0f || !ca_out || !cb_out) return EMS_E_INVALID;
  const double a = (double)ca, b = (double)cb;
  const double cap = dt * (1.0 + a) / (2.0 * b) + c;
  const double cond = (1.0 - a) / b + g;
  const double den = cap + 0.5 * cond * dt;
  *ca_out = (float)((cap - 0.5 * cond * dt) / den);
  *cb_out = (float)(dt / den);
  return EMS_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ DFT --
namespace {

const double TAU = 6.283185307179586476925286766559;
const uint64_t ANCHOR_INTERVAL = 1024;  // dft.rs:44

struct Cx {
  double re, im;
};
// Complex64::cis / from_polar (post/complex.rs:34-43)
Cx cis(double theta) { return {1.0 * std::cos(theta), 1.0 * std::sin(theta)}; }
// Complex64 * Complex64 (complex.rs:153-161)
Cx mul(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
double fract(double x) { return x - std::trunc(x); }

// phase_cycles (dft.rs:204-212): f(t0 + n dt) mod 1 with the exact
// rounding error of (f dt) n recovered by a fused multiply-add.
double phase_cycles(double f, double t0, double dt, double n) {
  const double a = f * dt;
  const double p = a * n;
  const double e = std::fma(a, n, -p);
  const double c = f * t0;
  return fract(fract(p) + e + fract(c));
}

}  // namespace

extern "C" {

// dft (dft.rs:218-222) through RunningDft (dft.rs:75-190): the running
// phasor advanced by one complex product per sample and re-anchored from
// the exact phase every 1024 samples.
void ems_dft(double t0, double dt, const double *x, uint64_t n, const double *freqs, int nf,
             double *out) {
  for (int k = 0; k < nf; k++) {
    const double f = freqs[k];
    const Cx step = cis(-TAU * f * dt);
    Cx ph = cis(-TAU * phase_cycles(f, t0, dt, 0.0));
    Cx acc = {0.0, 0.0};
    for (uint64_t s = 0; s < n; s++) {
      acc.re = acc.re + ph.re * x[s];
      acc.im = acc.im + ph.im * x[s];
      ph = mul(ph, step);
      if ((s + 1) % ANCHOR_INTERVAL == 0)
        ph = cis(-TAU * phase_cycles(f, t0, dt, (double)(s + 1)));
    }
    out[2 * k] = acc.re * dt;
    out[2 * k + 1] = acc.im * dt;
  }
}

// RunningDft's phasors at sample n (anchor, dft.rs:116-122): e^{-j2 pi f t_n}
// from the exact phase, per frequency, as (re, im) pairs.
void ems_dft_phasors(const double *freqs, int nf, double t0, double dt, uint64_t n, double *out) {
  for (int k = 0; k < nf; k++) {
    const Cx p = cis(-TAU * phase_cycles(freqs[k], t0, dt, (double)n));
    out[2 * k] = p.re, out[2 * k + 1] = p.im;
  }
}

// RunningDft's per-sample rotation (dft.rs:93-97): e^{-j2 pi f dt}.
void ems_dft_steps(const double *freqs, int nf, double dt, double *out) {
  for (int k = 0; k < nf; k++) {
    const Cx p = cis(-TAU * freqs[k] * dt);
    out[2 * k] = p.re, out[2 * k + 1] = p.im;
  }
}

}  // extern "C"

// ------------------------------------------------------------- run loop --
namespace {

struct Stepper {
  virtual ~Stepper() = default;
  virtual uint64_t timestep() = 0;
  virtual double dt() = 0;
  virtual int step_with_energy(double *w) = 0;
  virtual int run(uint64_t n) = 0;
  // Stepper::energy_now (run.rs:295-299): the energy of the current state
  // in `measure` without stepping, where the stepper has it.
  virtual bool has_energy_now() { return false; }
  virtual int energy_now(int measure, double *w) {
    (void)measure, (void)w;
    return EMS_E_UNSUPPORTED;
  }
  // per probe: sample times and values (run.rs:302-306)
  virtual int signals(std::function<void(const std::vector<double> &, const std::vector<double> &)>) {
    return 0;
  }
};

// Settling (run.rs:309-381)
struct Settling {
  double freq, tol, window, max_ext;
  struct Sum {
    size_t done = 0;
    double re = 0, im = 0, abs = 0;
  };
  std::vector<Sum> sums;
  struct Checked {
    uint64_t steps;
    double w;
    std::vector<std::pair<double, double>> x;
  };
  std::vector<Checked> history;

  int update(Stepper &s, uint64_t steps, double w) {
    const double om = 2.0 * M_PI * freq;
    size_t q = 0;
    int rc = s.signals([&](const std::vector<double> &t, const std::vector<double> &v) {
      if (sums.size() == q) sums.push_back(Sum{});
      Sum &p = sums[q];
      const size_t n = std::min(t.size(), v.size());
      for (size_t k = p.done; k < n; k++) {
        const double sn = std::sin(om * t[k]), cs = std::cos(om * t[k]);
        p.re += v[k] * cs;
        p.im -= v[k] * sn;
        p.abs += std::fabs(v[k]);
      }
      p.done = n;
      q++;
    });
    Checked c{steps, w, {}};
    for (auto &p : sums) c.x.emplace_back(p.re, p.im);
    history.push_back(std::move(c));
    return rc;
  }

  bool settled() const {
    if (history.empty()) return true;
    const Checked &last = history.back();
    const uint64_t span = (uint64_t)std::ceil(window * (double)last.steps);
    const uint64_t from = last.steps > span ? last.steps - span : 0;
    size_t n = 0;
    while (n < history.size() && history[n].steps <= from) n++;
    if (n == 0) return false;
    const Checked &then = history[n - 1];
    if (last.w <= 0.0) return true;
    const double fall = std::log(then.w / last.w);
    if (std::isnan(fall) || fall <= 0.0) return false;
    for (size_t q = 0; q < sums.size(); q++) {
      const double re = q < then.x.size() ? then.x[q].first : 0.0;
      const double im = q < then.x.size() ? then.x[q].second : 0.0;
      const double moved = std::hypot(last.x[q].first - re, last.x[q].second - im);
      if (!(2.0 * moved <= fall * tol * sums[q].abs)) return false;
    }
    return true;
  }
};

constexpr double NEGATIVE_ENERGY = 1e-12;  // run.rs:419

// run::drive (run.rs:424-540), default schedule (energy as a by-product of
// the step after each check; invariant measure, nothing held).
int drive(Stepper &s, const ems_run_limits &lim, ems_run_outcome &out) {
  if (lim.has_criteria && !(lim.energy_ratio >= 0.0)) return EMS_E_INVALID;
  if (lim.has_excitation_end && std::isnan(lim.excitation_end)) return EMS_E_INVALID;
  if (lim.energy_measure < 0 || lim.energy_measure > 1) return EMS_E_INVALID;
  const uint64_t interval = lim.check_interval ? lim.check_interval : 50;
  const uint64_t max = lim.max_timesteps;
  const auto start = std::chrono::steady_clock::now();
  uint64_t steps = 0;
  double w = 0.0, w_max = 0.0, ratio_db = 0.0;
  bool met = false;
  uint64_t met_steps = 0;
  Settling settle{lim.settle_frequency, lim.settle_tolerance, lim.settle_window,
                  lim.settle_max_extension, {}, {}};
  // run.rs:436-442: openEMS's schedule evaluates the state reached after
  // the interval; another measure than the invariant needs the state's own
  // evaluation before the step.
  const int measure = lim.energy_measure;
  const bool supported = s.has_energy_now();
  const bool after = lim.check_after_interval && supported;
  const bool before = !after && measure != 0 && supported;
  int reason;
  for (;;) {
    if (steps >= max) {
      reason = 1;
      break;
    }
    uint64_t n;
    int rc;
    if (after) {
      // run.rs:456-470: the next regular check or the extra one before it
      const uint64_t regular = (steps / interval + 1) * interval;
      uint64_t next = regular;
      if (lim.has_extra_check && lim.extra_check > steps && lim.extra_check < regular)
        next = lim.extra_check;
      next = std::min(next, max);
      if ((rc = s.run(next - steps))) return rc;
      steps = next;
      if (steps % interval != 0 && !(lim.has_extra_check && lim.extra_check == steps)) continue;
      n = s.timestep();
      if ((rc = s.energy_now(measure, &w))) return rc;
    } else if (before) {
      n = s.timestep();
      if ((rc = s.energy_now(measure, &w))) return rc;
      if ((rc = s.run(1))) return rc;
      steps += 1;
    } else {
      n = s.timestep();
      if ((rc = s.step_with_energy(&w))) return rc;
      steps += 1;
    }
    w_max = std::fmax(w_max, w);
    const double ratio = w_max > 0.0 ? w / w_max : 1.0;
    ratio_db = ratio > 0.0 ? 10.0 * std::log10(ratio) : -INFINITY;
    const double t = (double)n * s.dt();
    if (!std::isfinite(w) || w < -NEGATIVE_ENERGY * w_max) {
      reason = 3;
      break;
    }
    if (lim.has_settle && lim.has_criteria)
      if ((rc = settle.update(s, steps, w))) return rc;
    if (lim.has_criteria && (!lim.has_excitation_end || t > lim.excitation_end) && w_max > 0.0 &&
        ratio < lim.energy_ratio) {
      if (!met) met = true, met_steps = steps;
      const bool done = !lim.has_settle ||
                        (double)steps >= (double)met_steps * (1.0 + lim.settle_max_extension) ||
                        settle.settled();
      if (done) {
        reason = 0;
        break;
      }
    }
    if (!after) {
      const uint64_t chunk = std::min(interval - 1, max - steps);
      if ((rc = s.run(chunk))) return rc;
      steps += chunk;
    }
  }
  out.reason = reason;
  out.ratio_db = ratio_db;
  out.steps = steps;
  out.timestep = s.timestep();
  out.time = (double)s.timestep() * s.dt();
  out.energy = w;
  out.max_energy = w_max;
  out.elapsed_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
  out.has_criterion_steps = met;
  out.criterion_steps = met_steps;
  return EMS_OK;
}

struct GpuStepper : Stepper {
  ems_gpu *h;
  double dtv;
  std::vector<int> probes;
  std::vector<std::vector<double>> vals;
  GpuStepper(ems_gpu *h_, double dt_, const int *p, int np) : h(h_), dtv(dt_), probes(p, p + np) {}
  uint64_t timestep() override { return ems_gpu_timestep(h); }
  double dt() override { return dtv; }
  int step_with_energy(double *w) override { return ems_gpu_step_with_energy(h, w); }
  int run(uint64_t n) override { return ems_gpu_run(h, n); }
  bool has_energy_now() override { return true; }
  int energy_now(int measure, double *w) override { return ems_gpu_energy_in(h, measure, w); }
  int signals(std::function<void(const std::vector<double> &, const std::vector<double> &)> f)
      override {
    for (int id : probes) {
      uint64_t count = 0;
      int rc = ems_gpu_probe_values(h, id, nullptr, 0, &count);
      if (rc) return rc;
      std::vector<double> v(count), t(count);
      if (count && (rc = ems_gpu_probe_values(h, id, v.data(), count, &count))) return rc;
      if (count && (rc = ems_gpu_probe_times(h, id, t.data(), count, &count))) return rc;
      f(t, v);
    }
    return 0;
  }
};

struct ScriptStepper : Stepper {
  const double *w;
  uint64_t nw, n = 0;
  const double *sig;
  uint64_t ns;
  uint64_t timestep() override { return n; }
  double dt() override { return 1.0; }
  int step_with_energy(double *out) override {
    *out = w[std::min<uint64_t>(n, nw - 1)];
    n += 1;
    return 0;
  }
  int run(uint64_t k) override {
    n += k;
    return 0;
  }
  int signals(std::function<void(const std::vector<double> &, const std::vector<double> &)> f)
      override {
    if (!sig) return 0;
    std::vector<double> t(n), v(n);
    for (uint64_t k = 0; k < n; k++) t[k] = (double)k, v[k] = k < ns ? sig[k] : 0.0;
    f(t, v);
    return 0;
  }
};

}  // namespace

extern "C" {

int ems_run_until(ems_gpu *h, const ems_run_limits *limits, const int *probes, int n_probes,
                  ems_run_outcome *out) {
  if (!h || !limits || !out) return EMS_E_INVALID;
  GpuStepper s(h, ems_gpu_dt(h), probes, n_probes);
  return drive(s, *limits, *out);
}

int ems_run_until_batch(ems_gpu_batch *b, int n, ems_gpu *const *models,
                        const ems_run_limits *lim, ems_run_outcome *out) {
  if (!b || n < 1 || !models || !lim || !out) return EMS_E_INVALID;
  if (lim->has_settle || lim->check_after_interval || lim->energy_measure) return EMS_E_UNSUPPORTED;
  if (lim->has_criteria && !(lim->energy_ratio >= 0.0)) return EMS_E_INVALID;
  if (lim->has_excitation_end && std::isnan(lim->excitation_end)) return EMS_E_INVALID;
  const uint64_t interval = lim->check_interval ? lim->check_interval : 50;
  const uint64_t max = lim->max_timesteps;
  const auto start = std::chrono::steady_clock::now();
  struct St {
    uint64_t steps = 0, met_steps = 0;
    double w = 0.0, w_max = 0.0, ratio_db = 0.0;
    bool met = false, done = false;
    int reason = 1;
  };
  std::vector<St> st(n);
  std::vector<uint8_t> act(n, 1);
  int rc;
  for (;;) {
    // (All running models have taken the same number of steps.)
    bool any = false;
    for (int q = 0; q < n; q++) {
      if (st[q].done) continue;
      if (st[q].steps >= max) {
        st[q].done = true, st[q].reason = 1;
        act[q] = 0;
        continue;
      }
      any = true;
    }
    if (!any) break;
    // The energy step of drive(): W^n of the state before the step.
    std::vector<uint64_t> t0(n);
    std::vector<double> wq(n, 0.0);
    for (int q = 0; q < n; q++)
      if (act[q]) t0[q] = ems_gpu_timestep(models[q]);
    if ((rc = ems_gpu_energy_many(models, n, act.data(), wq.data()))) return rc;
    for (int q = 0; q < n; q++)
      if (act[q]) st[q].w = wq[q];
    if ((rc = ems_gpu_batch_run(b, act.data(), 1))) return rc;
    uint64_t steps_now = 0;
    for (int q = 0; q < n; q++) {
      if (!act[q]) continue;
      St &s = st[q];
      s.steps += 1;
      steps_now = s.steps;
      s.w_max = std::fmax(s.w_max, s.w);
      const double ratio = s.w_max > 0.0 ? s.w / s.w_max : 1.0;
      s.ratio_db = ratio > 0.0 ? 10.0 * std::log10(ratio) : -INFINITY;
      const double t = (double)t0[q] * ems_gpu_dt(models[q]);
      if (!std::isfinite(s.w) || s.w < -NEGATIVE_ENERGY * s.w_max) {
        s.done = true, s.reason = 3, act[q] = 0;
        continue;
      }
      if (lim->has_criteria && (!lim->has_excitation_end || t > lim->excitation_end) &&
          s.w_max > 0.0 && ratio < lim->energy_ratio) {
        if (!s.met) s.met = true, s.met_steps = s.steps;
        s.done = true, s.reason = 0, act[q] = 0;
      }
    }
    bool left = false;
    for (int q = 0; q < n; q++) left = left || act[q];
    if (!left) break;
    const uint64_t chunk = std::min(interval - 1, max - steps_now);
    if ((rc = ems_gpu_batch_run(b, act.data(), chunk))) return rc;
    for (int q = 0; q < n; q++)
      if (act[q]) st[q].steps += chunk;
  }
  const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
  for (int q = 0; q < n; q++) {
    ems_run_outcome &o = out[q];
    o.reason = st[q].reason;
    o.ratio_db = st[q].ratio_db;
    o.steps = st[q].steps;
    o.timestep = ems_gpu_timestep(models[q]);
    o.time = (double)o.timestep * ems_gpu_dt(models[q]);
    o.energy = st[q].w;
    o.max_energy = st[q].w_max;
    o.elapsed_s = el;
    o.has_criterion_steps = st[q].met;
    o.criterion_steps = st[q].met_steps;
  }
  return EMS_OK;
}

int ems_drive_script(const double *w, uint64_t nw, const double *signal, uint64_t ns,
                     const ems_run_limits *limits, ems_run_outcome *out) {
  if (!w || nw == 0 || !limits || !out) return EMS_E_INVALID;
  ScriptStepper s;
  s.w = w, s.nw = nw, s.sig = signal, s.ns = ns;
  return drive(s, *limits, *out);
}

}  // extern "C"

// ---- conformal stability estimate (conformal.rs:2858-3055) ----

namespace {

struct Stab {
  size_t N;
  int n[3];
  std::vector<double> sq[3];
  const ems_stability_desc *d;
};

// apply_s (conformal.rs:2927-2995): w = S v.
void stab_apply(const Stab &st, const std::vector<double> (&v)[3], std::vector<double> (&w)[3]) {
  const ems_stability_desc &d = *st.d;
  const int *n = st.n;
  for (int a = 0; a < 3; a++) std::fill(w[a].begin(), w[a].end(), 0.0);
  const size_t stride[3] = {(size_t)n[1] * n[2], (size_t)n[2], 1};
  for (int a = 0; a < 3; a++) {
    const int b = (a + 1) % 3, c = (a + 2) % 3;
    const size_t sb = stride[b], sc = stride[c];
    const std::vector<double> &vb = v[b], &vc = v[c], &rb = st.sq[b], &rc = st.sq[c];
    std::vector<double> &wb = w[b], &wc = w[c];
    auto last = [&](int dd, int p) { return (b == dd || c == dd) && p + 1 >= n[dd]; };
    const int nk = (b == 2 || c == 2) ? std::max(n[2] - 1, 0) : n[2];
    const float *dbl = d.db[a];
    for (int i = 0; i < n[0]; i++) {
      if (last(0, i)) continue;
      for (int j = 0; j < n[1]; j++) {
        if (last(1, j)) continue;
        const size_t base = ((size_t)i * n[1] + j) * n[2];
        for (int k = 0; k < nk; k++) {
          const double dv = (double)dbl[base + k];
          if (dv == 0.0) continue;
          const size_t q = base + k;
          const double curl =
              vc[q + sb] * rc[q + sb] - vc[q] * rc[q] - vb[q + sc] * rb[q + sc] + vb[q] * rb[q];
          const double f = dv * curl;
          if (f == 0.0) continue;
          wc[q + sb] += f;
          wc[q] -= f;
          wb[q + sc] -= f;
          wb[q] += f;
        }
      }
    }
  }
  auto x = [&](int a, uint64_t k) { return v[a][k] * st.sq[a][k]; };
  for (int g = 0; g < d.n_groups; g++) {
    double curl = -0.0;
    for (uint32_t t = d.g_term_start[g]; t < d.g_term_start[g + 1]; t++)
      curl += d.g_term_w[t] * x(d.g_term_axis[t], d.g_term_idx[t]);
    const double f = d.g_dt_over_l[g] * curl;
    for (uint32_t t = d.g_term_start[g]; t < d.g_term_start[g + 1]; t++)
      w[d.g_term_axis[t]][d.g_term_idx[t]] += d.g_term_w[t] * f;
  }
  std::vector<double> y;
  for (int b = 0; b < d.n_blocks; b++) {
    const uint32_t r0 = d.blk_start[b], nr = d.blk_start[b + 1] - r0;
    y.assign(nr, 0.0);
    for (uint32_t i = 0; i < nr; i++) {
      double s = -0.0;
      for (uint32_t t = d.row_term_start[r0 + i]; t < d.row_term_start[r0 + i + 1]; t++)
        s += d.row_term_w[t] * x(d.row_term_axis[t], d.row_term_idx[t]);
      y[i] = s;
    }
    for (uint32_t i = 0; i < nr; i++) {  // Envelope::solve
      const uint32_t f = d.env_first[r0 + i];
      const double *row = d.env_val + d.env_off[r0 + i];
      double s = y[i];
      for (uint32_t jj = f; jj < i; jj++) s -= row[jj - f] * y[jj];
      y[i] = s / row[i - f];
    }
    for (uint32_t i = nr; i-- > 0;) {
      const uint32_t f = d.env_first[r0 + i];
      const double *row = d.env_val + d.env_off[r0 + i];
      y[i] /= row[i - f];
      const double yi = y[i];
      for (uint32_t jj = f; jj < i; jj++) y[jj] -= row[jj - f] * yi;
    }
    for (uint32_t i = 0; i < nr; i++)
      for (uint32_t t = d.row_term_start[r0 + i]; t < d.row_term_start[r0 + i + 1]; t++)
        w[d.row_term_axis[t]][d.row_term_idx[t]] += d.row_term_w[t] * d.blk_dt[b] * y[i];
  }
  for (int a = 0; a < 3; a++)
    for (size_t k = 0; k < st.N; k++) w[a][k] *= st.sq[a][k];
}

}  // namespace

extern "C" {

double ems_conformal_stability(const ems_stability_desc *d) {
  if (!d || d->nodes[0] < 2 || d->nodes[1] < 2 || d->nodes[2] < 2 || d->n_groups < 0 ||
      d->n_blocks < 0 || d->n_seeds < 0 || d->iterations < 0)
    return EMS_E_INVALID;
  for (int a = 0; a < 3; a++)
    if (!d->cb[a] || !d->db[a]) return EMS_E_INVALID;
  Stab st;
  st.d = d;
  for (int a = 0; a < 3; a++) st.n[a] = d->nodes[a];
  st.N = (size_t)st.n[0] * st.n[1] * st.n[2];
  // every index the loops follow
  const uint32_t ng = (uint32_t)d->n_groups, nb = (uint32_t)d->n_blocks;
  auto edge_ok = [&](int32_t a, uint64_t k) { return a >= 0 && a < 3 && k < st.N; };
  if (ng && (!d->g_dt_over_l || !d->g_term_start)) return EMS_E_INVALID;
  for (uint32_t g = 0; g < ng; g++)
    for (uint32_t t = d->g_term_start[g]; t < d->g_term_start[g + 1]; t++)
      if (!edge_ok(d->g_term_axis[t], d->g_term_idx[t])) return EMS_E_INVALID;
  if (nb && (!d->blk_start || !d->blk_dt || !d->row_term_start || !d->env_first || !d->env_off ||
             !d->env_val))
    return EMS_E_INVALID;
  for (uint32_t b = 0; b < nb; b++) {
    const uint32_t r0 = d->blk_start[b], r1 = d->blk_start[b + 1];
    if (r1 < r0) return EMS_E_INVALID;
    for (uint32_t r = r0; r < r1; r++) {
      if (d->env_first[r] > r - r0 ||
          d->env_off[r + 1] - d->env_off[r] != (uint64_t)(r - r0 - d->env_first[r] + 1))
        return EMS_E_INVALID;
      for (uint32_t t = d->row_term_start[r]; t < d->row_term_start[r + 1]; t++)
        if (!edge_ok(d->row_term_axis[t], d->row_term_idx[t])) return EMS_E_INVALID;
    }
  }
  for (int q = 0; q < d->n_seeds; q++)
    if (!edge_ok(d->seed_axis[q], d->seed_idx[q])) return EMS_E_INVALID;
  for (int a = 0; a < 3; a++) {
    st.sq[a].resize(st.N);
    for (size_t k = 0; k < st.N; k++) st.sq[a][k] = std::sqrt((double)d->cb[a][k]);
  }
  std::vector<double> v[3], w[3];
  for (int a = 0; a < 3; a++) {
    // xorshift: a fixed sequence per array
    uint64_t state = 0x9e3779b97f4a7c15ull ^ (uint64_t)(a + 1);
    v[a].resize(st.N);
    w[a].assign(st.N, 0.0);
    for (size_t k = 0; k < st.N; k++) {
      state ^= state << 13;
      state ^= state >> 7;
      state ^= state << 17;
      const double r = (double)(state >> 11) / (double)(1ull << 53);
      v[a][k] = st.sq[a][k] > 0.0 ? 1e-3 * (r - 0.5) : 0.0;
    }
  }
  for (int q = 0; q < d->n_seeds; q++)
    if (st.sq[d->seed_axis[q]][d->seed_idx[q]] > 0.0) v[d->seed_axis[q]][d->seed_idx[q]] += 1.0;
  double lambda = 0.0;
  const int it = std::max(d->iterations, 1);
  for (int s = 0; s < it; s++) {
    double nn = -0.0;
    for (int a = 0; a < 3; a++)
      for (size_t k = 0; k < st.N; k++) nn += v[a][k] * v[a][k];
    const double norm = std::sqrt(nn);
    if (norm == 0.0) return 0.0;
    for (int a = 0; a < 3; a++)
      for (size_t k = 0; k < st.N; k++) v[a][k] /= norm;
    stab_apply(st, v, w);
    double l = -0.0;
    for (int a = 0; a < 3; a++)
      for (size_t k = 0; k < st.N; k++) l += v[a][k] * w[a][k];
    lambda = l;
    for (int a = 0; a < 3; a++) std::swap(v[a], w[a]);
  }
  return std::sqrt(std::max(lambda, 0.0)) / 2.0;
}

}  // extern "C"

// ---- setup producers: dispersive media and lumped elements ----
//
// What the reference's compile step runs on the host before the engine
// steps: the edges a dispersive cell box reaches with their pole shares
// (Dispersive::fill_cells), the install plan of a dispersive medium
// (Dispersive::install) and of a lumped element (LumpedElement::install),
// producing the descriptions ems_gpu_add_dispersive and ems_gpu_add_lumped
// take.

namespace {

bool pole_valid(const ems_pole_entry &e) {
  if (e.kind < 0 || e.kind > 2) return false;
  const int np = e.kind == 0 ? 3 : 2;
  for (int q = 0; q < np; q++)
    if (!(std::isfinite(e.p[q]) && e.p[q] >= 0.0)) return false;
  return true;
}

// Pole::scaled (dispersive.rs:186-204): a share s of the pole's strength.
void pole_scale(ems_pole_entry &e, double s) {
  if (e.kind == 2)
    e.p[0] = e.p[0] * s;
  else
    e.p[0] = e.p[0] * std::sqrt(s);
}

// The pole as a series L, R and 1/C at edge capacitance (inductance) m
// (dispersive.rs:300-345): Lorentz L = 1/(m wp^2), R = gamma L,
// 1/C = w0^2 L; Drude the same without 1/C; Debye C = m d_eps, R = tau/C.
void pole_lrc(const ems_pole_entry &e, double m, double *l, double *r, double *inv_c) {
  if (e.kind == 0) {
    const double wp = e.p[0], w0 = e.p[1], g = e.p[2];
    *l = 1.0 / (m * wp * wp), *r = g * *l, *inv_c = w0 * w0 * *l;
  } else if (e.kind == 1) {
    const double wp = e.p[0], g = e.p[1];
    *l = 1.0 / (m * wp * wp), *r = g * *l, *inv_c = 0.0;
  } else {
    const double c = m * e.p[0];
    *l = 0.0, *r = e.p[1] / c, *inv_c = 1.0 / c;
  }
}

}  // namespace

extern "C" {

int ems_dispersive_fill_cells(const double *x, int nx, const double *y, int ny, const double *z,
                              int nz, const int32_t cells[6], int field, const ems_pole_entry *poles,
                              int n_poles, ems_pole_entry *out, uint64_t cap, uint64_t *count) {
  if (!cells || !count || n_poles < 0 || (n_poles && !poles) || field < 0 || field > 1)
    return EMS_E_INVALID;
  for (int q = 0; q < n_poles; q++)
    if (!pole_valid(poles[q])) return EMS_E_INVALID;
  const Grid g = make_grid(x, nx, y, ny, z, nz);
  const int lo[3] = {cells[0], cells[2], cells[4]}, hi[3] = {cells[1], cells[3], cells[5]};
  for (int a = 0; a < 3; a++)
    if (lo[a] < 0 || hi[a] > g.n[a] - 1 || lo[a] > hi[a]) return EMS_E_INVALID;
  uint64_t w_out = 0;
  *count = 0;
  if (lo[0] == hi[0] || lo[1] == hi[1] || lo[2] == hi[2]) return EMS_OK;
  auto inside = [&](int a, int c) { return lo[a] <= c && c < hi[a]; };
  for (int a = 0; a < 3; a++) {
    const int b = nxt(a), c = prv(a);
    // along the edge's axis the cells, across it the nodes that touch them
    auto lim = [&](int t, bool along) { return along ? hi[t] : std::min(hi[t] + 1, g.n[t]); };
    const bool e = field == 0;
    const int ea = lim(a, e), eb = lim(b, !e), ec = lim(c, !e);
    for (int pa = lo[a]; pa < ea; pa++)
      for (int pb = lo[b]; pb < eb; pb++)
        for (int pc = lo[c]; pc < ec; pc++) {
          double w_in = 0.0, w = 0.0;
          if (e) {
            for (int cb : {pb - 1, pb}) {
              if (cb < 0 || cb >= g.n[b] - 1) continue;
              for (int cc : {pc - 1, pc}) {
                if (cc < 0 || cc >= g.n[c] - 1) continue;
                const double area = g.delta(b, cb) * g.delta(c, cc);
                w += area;
                if (inside(b, cb) && inside(c, cc)) w_in += area;
              }
            }
          } else {
            for (int ca : {pa - 1, pa}) {
              if (ca < 0 || ca >= g.n[a] - 1) continue;
              const double ln = g.delta(a, ca);
              w += ln;
              if (inside(a, ca)) w_in += ln;
            }
          }
          const double share = w > 0.0 ? w_in / w : 0.0;
          if (!(share > 0.0)) continue;
          for (int q = 0; q < n_poles; q++) {
            ems_pole_entry en = poles[q];
            pole_scale(en, share);
            if (en.p[0] == 0.0) continue;  // null poles are not pushed
            en.field = field, en.axis = a;
            en.node[a] = pa, en.node[b] = pb, en.node[c] = pc;
            if (out && w_out < cap) out[w_out] = en;
            w_out++;
          }
        }
  }
  *count = w_out;
  return out && w_out > cap ? EMS_E_INVALID : EMS_OK;
}

int ems_dispersive_plan(const double *x, int nx, const double *y, int ny, const double *z, int nz,
                        double dt, const float *lanes, int scheme, const ems_pole_entry *entries,
                        int n, ems_pole_plan *out) {
  if (!lanes || !out || n < 0 || (n && !entries) || !(dt > 0.0) || scheme < 0 || scheme > 1)
    return EMS_E_INVALID;
  const Grid g = make_grid(x, nx, y, ny, z, nz);
  const size_t N = (size_t)nx * ny * nz;
  std::vector<int> ord;
  for (int q = 0; q < n; q++) {
    const ems_pole_entry &e = entries[q];
    if (!pole_valid(e) || e.field < 0 || e.field > 1 || e.axis < 0 || e.axis > 2)
      return EMS_E_INVALID;
    const int a = e.axis, b = nxt(a), c = prv(a);
    for (int t = 0; t < 3; t++)
      if (e.node[t] < 0 || e.node[t] >= g.n[t]) return EMS_E_INVALID;
    const bool fits = e.field == 0 ? e.node[a] + 1 < g.n[a]
                                   : (e.node[b] + 1 < g.n[b] && e.node[c] + 1 < g.n[c]);
    if (!fits) return EMS_E_INVALID;
    if (e.p[0] == 0.0) continue;
    ord.push_back(q);
  }
  // the edges in plane order (dispersive.rs:2090-2105): x, field, axis, y, z
  auto key = [&](int q) {
    const ems_pole_entry &e = entries[q];
    return std::array<int, 5>{e.node[0], e.field, e.axis, e.node[1], e.node[2]};
  };
  std::stable_sort(ord.begin(), ord.end(), [&](int p, int q) { return key(p) < key(q); });
  std::map<std::array<uint64_t, 5>, int> cmap;
  int32_t ne[2] = {0, 0};
  uint64_t np[2] = {0, 0};
  int32_t nc = 0, skipped = 0;
  bool have_last = false;
  std::array<int, 5> last{};
  for (int q : ord) {
    const ems_pole_entry &e = entries[q];
    const int a = e.axis, b = nxt(a), c = prv(a), f = e.field;
    const size_t idx = g.index(e.node[0], e.node[1], e.node[2]);
    double coef, m;
    if (f == 0) {
      coef = (double)lanes[(size_t)(3 + a) * N + idx];
      const double geo = g.dual_delta(b, e.node[b]) * g.dual_delta(c, e.node[c]) /
                         g.delta(a, e.node[a]);
      m = EPS0 * geo;
    } else {
      coef = (double)lanes[(size_t)(9 + a) * N + idx];
      const double geo = g.delta(b, e.node[b]) * g.delta(c, e.node[c]) / g.dual_delta(a, e.node[a]);
      m = MU0 * geo;
    }
    const std::array<int, 5> k = key(q);
    const bool new_edge = !have_last || k != last;
    have_last = true, last = k;
    if (coef == 0.0) {  // PEC or padding: no medium there (dispersive.rs:2110-2118)
      if (new_edge) skipped++;
      continue;
    }
    if (new_edge) {
      const int32_t s = ne[f]++;
      out->axis[f][s] = a;
      for (int t = 0; t < 3; t++) out->nodes[f][3 * s + t] = e.node[t];
      out->sum_inv_z[f][s] = 0.0;
      out->pole_end[f][s] = (uint32_t)np[f];
    }
    double l, r, inv_c, cst[4];
    int32_t kind;
    pole_lrc(e, m, &l, &r, &inv_c);
    if (scheme == 1 && e.kind != 2) {
      // leapfrog (dispersive.rs:375-393): k_a, 1/C, k_b, L/2
      const double d = l / dt + 0.5 * r;
      cst[0] = (l / dt - 0.5 * r) / d, cst[1] = inv_c, cst[2] = 1.0 / d, cst[3] = 0.5 * l;
      kind = 1;
    } else {
      // trapezoidal branch (dispersive.rs:348-371): 2L/dt, 1/C, 1/z, (2L/dt) dt/4
      const double a2 = 2.0 * l / dt;
      const double zz = a2 + r + 0.5 * dt * inv_c;
      cst[0] = a2, cst[1] = inv_c, cst[2] = 1.0 / zz, cst[3] = 0.25 * a2 * dt;
      kind = 0;
      out->sum_inv_z[f][ne[f] - 1] += cst[2];
    }
    std::array<uint64_t, 5> ck;
    for (int t = 0; t < 4; t++) std::memcpy(&ck[t], &cst[t], sizeof(double));
    ck[4] = (uint64_t)kind;
    auto it = cmap.find(ck);
    int ci;
    if (it == cmap.end()) {
      ci = nc++;
      cmap.emplace(ck, ci);
      out->a[ci] = cst[0], out->inv_c[ci] = cst[1], out->inv_z[ci] = cst[2];
      out->half_l[ci] = cst[3], out->kind[ci] = kind;
    } else {
      ci = it->second;
    }
    out->cidx[f][np[f]++] = (uint32_t)ci;
    out->pole_end[f][ne[f] - 1] = (uint32_t)np[f];
  }
  for (int f = 0; f < 2; f++) out->n_edges[f] = ne[f], out->n_poles[f] = np[f];
  out->n_consts = nc;
  out->skipped = skipped;
  return EMS_OK;
}

int ems_lumped_plan(const int32_t nodes[3], double dt, const ems_lumped_element *el, float *ca,
                    float *cb, ems_lumped_install *out) {
  if (!nodes || !el || !ca || !cb || !out || !(dt > 0.0) || el->axis < 0 || el->axis > 2 ||
      el->n_edges < 1 || !el->nodes || el->topology < 0 || el->topology > 1)
    return EMS_E_INVALID;
  const int ne = el->n_edges, a = el->axis;
  const bool box = el->columns > 0;
  if (!box && (!el->edge_r || !el->edge_l || !el->edge_c)) return EMS_E_INVALID;
  auto has = [](double v) { return !std::isnan(v); };
  auto ok = [](double v) { return std::isnan(v) || (std::isfinite(v) && v > 0.0); };
  // which of R, L, C the element has (every edge the same set)
  bool pr, pl, pc;
  if (box) {
    if (!ok(el->r) || !ok(el->l) || !ok(el->c) || el->series < 1) return EMS_E_INVALID;
    pr = has(el->r), pl = has(el->l), pc = has(el->c);
  } else {
    pr = has(el->edge_r[0]), pl = has(el->edge_l[0]), pc = has(el->edge_c[0]);
    for (int q = 0; q < ne; q++)
      if (!ok(el->edge_r[q]) || !ok(el->edge_l[q]) || !ok(el->edge_c[q]) ||
          has(el->edge_r[q]) != pr || has(el->edge_l[q]) != pl || has(el->edge_c[q]) != pc)
        return EMS_E_INVALID;
  }
  const bool parallel = el->topology == 0;
  const bool branch = parallel ? pl : (pl || pc);
  const size_t N = (size_t)nodes[0] * nodes[1] * nodes[2];
  for (int q = 0; q < ne; q++) {
    const int32_t *p = el->nodes + 3 * q;
    for (int t = 0; t < 3; t++)
      if (p[t] < 0 || p[t] >= nodes[t]) return EMS_E_INVALID;
    if (p[a] + 1 >= nodes[a]) return EMS_E_INVALID;
    const size_t idx = ((size_t)p[0] * nodes[1] + p[1]) * nodes[2] + p[2];
    if (idx >= N || cb[idx] == 0.0f) return EMS_E_INVALID;  // PEC edge
  }
  const double inf = INFINITY;
  for (int q = 0; q < ne; q++) {
    // the edge's share (lumped.rs:829-878): G, C in parallel, (R, L, C) of a branch
    double ge, ce, br = 0.0, bl = 0.0, bc = inf;
    bool eb = false;
    if (box) {
      const double m = (double)el->series, nn = (double)el->columns;
      const double r = el->r, l = el->l, c = el->c;
      if (parallel) {
        ge = pr ? m / (nn * r) : 0.0, ce = pc ? m * c / nn : 0.0;
        if (pl) eb = true, br = 0.0, bl = nn * l / m, bc = inf;
      } else if (!branch) {
        ge = pr ? m / (nn * r) : 0.0, ce = 0.0;
      } else {
        ge = 0.0, ce = 0.0, eb = true;
        br = pr ? nn * r / m : 0.0, bl = pl ? nn * l / m : 0.0, bc = pc ? m * c / nn : inf;
      }
    } else {
      const double r = el->edge_r[q], l = el->edge_l[q], c = el->edge_c[q];
      if (parallel) {
        ge = pr ? 1.0 / r : 0.0, ce = pc ? c : 0.0;
        if (pl) eb = true, br = 0.0, bl = l, bc = inf;
      } else if (!branch) {
        ge = pr ? 1.0 / r : 0.0, ce = 0.0;
      } else {
        ge = 0.0, ce = 0.0, eb = true;
        br = pr ? r : 0.0, bl = pl ? l : 0.0, bc = pc ? c : inf;
      }
    }
    const int32_t *p = el->nodes + 3 * q;
    const size_t idx = ((size_t)p[0] * nodes[1] + p[1]) * nodes[2] + p[2];
    if (ge != 0.0 || ce != 0.0) {
      float fa, fb;
      ems_fold_parallel(ca[idx], cb[idx], dt, ge, ce, &fa, &fb);
      ca[idx] = fa, cb[idx] = fb;
    }
    out->cb_e[q] = (double)cb[idx];
    out->rec_g[q] = ge, out->rec_c[q] = ce;
    out->g_src[q] = (branch && !parallel) ? 0.0 : ge;
    if (branch) {
      if (!eb) br = 0.0, bl = 0.0, bc = inf;
      const double ic = std::isfinite(bc) ? 1.0 / bc : 0.0;
      out->two_l_dt[q] = 2.0 * bl / dt;
      out->inv_c[q] = ic;
      out->z[q] = 2.0 * bl / dt + br + 0.5 * dt * ic;
    }
  }
  out->branch = branch ? ((el->explicit_inductor && parallel) ? 2 : 1) : 0;
  out->src_in_branch = branch && !parallel;
  return EMS_OK;
}

}  // extern "C"
