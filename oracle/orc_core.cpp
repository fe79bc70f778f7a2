// Oracle core implementation. TEST INFRASTRUCTURE ONLY (see orc_core.hpp).
#include "orc_core.hpp"

#include <algorithm>
#include <limits>

namespace orc {

Stencil make_stencil(int id, double s) {
  Stencil st;
  st.id = id;
  st.scaling = s;
  auto push = [&](double x, double y, double z) { st.dir.push_back({x * 1.0, y * 1.0, z * 1.0}); };
  if (id == ST_D2Q9) {
    // stencils/D2Q9.cpp:40-61
    st.D = 2;
    st.Q = 9;
    const double d[9][2] = {{0, 0}, {s, 0}, {0, s}, {-s, 0}, {0, -s}, {s, s}, {-s, s}, {-s, -s}, {s, -s}};
    for (auto& v : d) push(v[0], v[1], 0);
    st.w = {4. / 9., 1. / 9., 1. / 9., 1. / 9., 1. / 9., 1. / 36., 1. / 36., 1. / 36., 1. / 36.};
    st.cs2 = s * s / 3.;
  } else if (id == ST_D3Q19) {
    // stencils/D3Q19.cpp:37-60 (Mohamad ordering)
    st.D = 3;
    st.Q = 19;
    const double d[19][3] = {{0, 0, 0},  {s, 0, 0},   {0, 0, s},   {-s, 0, 0}, {0, 0, -s}, {0, -s, 0}, {0, s, 0},
                             {s, 0, s},  {-s, 0, s},  {-s, 0, -s}, {s, 0, -s}, {s, -s, 0}, {s, s, 0},  {-s, s, 0},
                             {-s, -s, 0}, {0, -s, s}, {0, s, s},   {0, s, -s}, {0, -s, -s}};
    for (auto& v : d) push(v[0], v[1], v[2]);
    st.w.push_back(1. / 3.);
    for (int i = 1; i <= 6; i++) st.w.push_back(1. / 18.);
    for (int i = 7; i <= 18; i++) st.w.push_back(1. / 36.);
    st.cs2 = s * s / 3.;
  } else if (id == ST_D3Q27) {
    // stencils/D3Q27.cpp:37-110
    st.D = 3;
    st.Q = 27;
    const double d[27][3] = {{0, 0, 0},   {s, 0, 0},    {-s, 0, 0},  {0, s, 0},   {0, -s, 0},   {0, 0, s},
                             {0, 0, -s},  {0, s, s},    {0, -s, -s}, {0, s, -s},  {0, -s, s},   {s, 0, s},
                             {-s, 0, -s}, {s, 0, -s},   {-s, 0, s},  {s, s, 0},   {-s, -s, 0},  {s, -s, 0},
                             {-s, s, 0},  {s, s, s},    {-s, -s, -s}, {s, s, -s}, {-s, -s, s},  {s, -s, s},
                             {-s, s, -s}, {s, -s, -s},  {-s, s, s}};
    for (auto& v : d) push(v[0], v[1], v[2]);
    st.w.push_back(8. / 27.);
    for (int i = 1; i <= 6; i++) st.w.push_back(2. / 27.);
    for (int i = 7; i <= 18; i++) st.w.push_back(1. / 54.);
    for (int i = 19; i <= 26; i++) st.w.push_back(1. / 216.);
    st.cs2 = s * s / 3.;
  } else if (id == ST_D2Q25H) {
    // stencils/D2Q25H.cpp:42-82
    st.D = 2;
    st.Q = 25;
    const double r = (std::sqrt(5.) - std::sqrt(2.)) / std::sqrt(3.);
    const double w0 = (-3 - 3 * r * r * r * r + 54 * r * r) / (75 * r * r);
    const double wm = (9 * r * r * r * r - 6 - 27 * r * r) / (300 * r * r * (r * r - 1));
    const double wn = (9 - 6 * r * r * r * r - 27 * r * r) / (300 * (1 - r * r));
    const double w0n = w0 * wn, w0m = w0 * wm, wmm = wm * wm, wmn = wm * wn, wnn = wn * wn;
    st.w = {w0 * w0, w0m, w0m, w0m, w0m, wmm, wmm, wmm, wmm, w0n, w0n, w0n, w0n,
            wnn,     wnn, wnn, wnn, wmn, wmn, wmn, wmn, wmn, wmn, wmn, wmn};
    const double cm = s * std::sqrt(5. - std::sqrt(10.)) / std::sqrt(3.);
    const double cn = std::sqrt(5. + std::sqrt(10.)) * s / std::sqrt(3.);
    const double d[25][2] = {{0, 0},   {cm, 0},   {0, cm},   {-cm, 0},   {0, -cm},  {cm, cm},  {-cm, cm},
                             {-cm, -cm}, {cm, -cm}, {cn, 0},   {0, cn},    {-cn, 0},  {0, -cn},  {cn, cn},
                             {-cn, cn}, {-cn, -cn}, {cn, -cn}, {cm, cn},   {cm, -cn}, {-cm, -cn}, {-cm, cn},
                             {cn, cm},  {cn, -cm}, {-cn, -cm}, {-cn, cm}};
    for (auto& v : d) push(v[0], v[1], 0);
    st.cs2 = s * s / (3);
  } else {
    throw std::runtime_error("unknown stencil id");
  }
  st.max_speed = s * std::sqrt(2);
  // opposite index: the direction with negated velocity
  st.opp.assign(st.Q, 0);
  for (int i = 0; i < st.Q; i++)
    for (int j = 0; j < st.Q; j++)
      if (st.dir[i][0] == -st.dir[j][0] && st.dir[i][1] == -st.dir[j][1] && st.dir[i][2] == -st.dir[j][2])
        st.opp[i] = j;
  return st;
}

// Legendre P_n and P_{n-1} at x (long double).
static void legendre(int n, long double x, long double& pn, long double& pn1) {
  long double p0 = 1, p1 = x;
  if (n == 0) {
    pn = 1;
    pn1 = 0;
    return;
  }
  for (int k = 2; k <= n; k++) {
    long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  pn = p1;
  pn1 = p0;
}

void gll_01(int p, std::vector<double>& x, std::vector<double>& w) {
  // Gauss-Lobatto on [-1,1]: endpoints and roots of P_p'. Newton on
  // (x P_p - P_{p-1}) from Chebyshev-Gauss-Lobatto guesses.
  const int n = p + 1;
  std::vector<long double> xs(n), ws(n);
  for (int j = 0; j < n; j++) {
    long double xx = -std::cos(3.14159265358979323846264338327950288L * j / p);
    if (j == 0) xx = -1;
    if (j == p) xx = 1;
    if (j > 0 && j < p) {
      for (int it = 0; it < 100; it++) {
        long double pn, pn1;
        legendre(p, xx, pn, pn1);
        long double dx = (xx * pn - pn1) / ((p + 1) * pn);
        xx -= dx;
        if (std::fabs((double)dx) < 1e-19) break;
      }
    }
    long double pn, pn1;
    legendre(p, xx, pn, pn1);
    xs[j] = xx;
    ws[j] = 2.0L / (p * (p + 1) * pn * pn);
  }
  x.resize(n);
  w.resize(n);
  for (int j = 0; j < n; j++) {
    // symmetric: enforce exact symmetry about 1/2
    long double xa = (xs[j] + 1) / 2;
    long double xb = (1 - xs[p - j]) / 2;
    x[j] = (double)((xa + xb) / 2);
    w[j] = (double)(ws[j] / 2);
  }
  x[0] = 0.0;
  x[p] = 1.0;
  if (p % 2 == 0) x[p / 2] = 0.5;
}

void Lagrange1D::init(const std::vector<double>& x) {
  n = (int)x.size();
  nodes = x;
  inv_den.assign(n, 1.0);
  for (int j = 0; j < n; j++) {
    double d = 1;
    for (int m = 0; m < n; m++)
      if (m != j) d *= (x[j] - x[m]);
    inv_den[j] = 1.0 / d;
  }
}

double Lagrange1D::value(int j, double xi) const {
  double v = inv_den[j];
  for (int m = 0; m < n; m++)
    if (m != j) v *= (xi - nodes[m]);
  return v;
}

double Lagrange1D::deriv(int j, double xi) const {
  double s = 0;
  for (int k = 0; k < n; k++) {
    if (k == j) continue;
    double t = 1;
    for (int m = 0; m < n; m++)
      if (m != j && m != k) t *= (xi - nodes[m]);
    s += t;
  }
  return s * inv_den[j];
}

void Mesh::map(int cell, const double xi[3], double x[3]) const {
  if (cartesian) {
    const auto& a = cell_vertex(cell, 0);
    const auto& b = cell_vertex(cell, n_cell_vertices() - 1);
    for (int d = 0; d < 3; d++) x[d] = d < D ? a[d] + xi[d] * (b[d] - a[d]) : 0.0;
    return;
  }
  for (int d = 0; d < 3; d++) x[d] = 0;
  for (int v = 0; v < n_cell_vertices(); v++) {
    double Nv = 1;
    for (int a = 0; a < D; a++) Nv *= ((v >> a) & 1) ? xi[a] : (1 - xi[a]);
    const auto& X = cell_vertex(cell, v);
    for (int d = 0; d < D; d++) x[d] += Nv * X[d];
  }
}

void Mesh::jacobian(int cell, const double xi[3], double J[3][3]) const {
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) J[a][b] = (a == b && a >= D) ? 1.0 : 0.0;
  if (cartesian) {
    const auto& a0 = cell_vertex(cell, 0);
    const auto& b0 = cell_vertex(cell, n_cell_vertices() - 1);
    for (int d = 0; d < D; d++) J[d][d] = b0[d] - a0[d];
    return;
  }
  for (int v = 0; v < n_cell_vertices(); v++) {
    const auto& X = cell_vertex(cell, v);
    for (int b = 0; b < D; b++) {
      double dN = 1;
      for (int a = 0; a < D; a++) {
        int bit = (v >> a) & 1;
        if (a == b)
          dN *= bit ? 1.0 : -1.0;
        else
          dN *= bit ? xi[a] : (1 - xi[a]);
      }
      for (int d = 0; d < D; d++) J[d][b] += dN * X[d];
    }
  }
}

static bool solve3(const double J[3][3], const double r[3], double s[3], int D) {
  if (D == 2) {
    double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    if (det == 0) return false;
    s[0] = (r[0] * J[1][1] - J[0][1] * r[1]) / det;
    s[1] = (J[0][0] * r[1] - r[0] * J[1][0]) / det;
    s[2] = 0;
    return true;
  }
  double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  if (det == 0) return false;
  double inv[3][3];
  inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
  inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
  inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
  inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
  inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
  inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
  inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
  inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
  inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
  for (int a = 0; a < 3; a++) s[a] = inv[a][0] * r[0] + inv[a][1] * r[1] + inv[a][2] * r[2];
  return true;
}

bool Mesh::inverse_map(int cell, const double x[3], double xi[3]) const {
  if (cartesian) {
    const auto& a = cell_vertex(cell, 0);
    const auto& b = cell_vertex(cell, n_cell_vertices() - 1);
    for (int d = 0; d < 3; d++) xi[d] = d < D ? (x[d] - a[d]) / (b[d] - a[d]) : 0.0;
    return true;
  }
  // Newton iteration on the multilinear map, started at the cell centre.
  for (int d = 0; d < 3; d++) xi[d] = d < D ? 0.5 : 0.0;
  double prev = 1e300;
  for (int it = 0; it < 100; it++) {
    double y[3], J[3][3], r[3], s[3];
    map(cell, xi, y);
    for (int d = 0; d < 3; d++) r[d] = d < D ? y[d] - x[d] : 0.0;
    jacobian(cell, xi, J);
    if (!solve3(J, r, s, D)) return false;
    double m = 0;
    for (int d = 0; d < D; d++) {
      xi[d] -= s[d];
      m = std::max(m, std::fabs(s[d]));
    }
    if (m < 1e-15) return true;
    // converged to the rounding level of the residual: the step no longer shrinks
    if (m < 1e-12 && m > 0.25 * prev) return true;
    prev = m;
    if (std::fabs(xi[0]) > 1e3) return false;
  }
  return true;
}

double Mesh::min_vertex_distance() const {
  double mind = 1e100;
  const int nvc = n_cell_vertices();
  for (int c = 0; c < n_cells(); c++) {
    for (int i = 0; i < nvc; i++)
      for (int j = i + 1; j < nvc; j++) {
        const auto& a = cell_vertex(c, i);
        const auto& b = cell_vertex(c, j);
        double s = 0;
        for (int d = 0; d < D; d++) s += (a[d] - b[d]) * (a[d] - b[d]);
        mind = std::min(mind, std::sqrt(s));
      }
  }
  return mind;
}

Mesh make_box_mesh(int D, const int N[3], const double lo[3], const double hi[3], bool cartesian) {
  Mesh m;
  m.D = D;
  for (int a = 0; a < 3; a++) m.N[a] = (a < D) ? N[a] : 1;
  m.cartesian = cartesian;
  const int nz = D == 3 ? m.N[2] + 1 : 1;
  m.vert.resize((size_t)m.nv(0) * m.nv(1) * nz);
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < m.nv(1); j++)
      for (int i = 0; i < m.nv(0); i++) {
        int idx[3] = {i, j, k};
        std::array<double, 3> x{0, 0, 0};
        for (int a = 0; a < D; a++) x[a] = lo[a] + (hi[a] - lo[a]) * idx[a] / m.N[a];
        m.vert[m.vidx(i, j, k)] = x;
      }
  return m;
}

}  // namespace orc
