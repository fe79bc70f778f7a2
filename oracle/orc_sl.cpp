// Oracle SL assembly. TEST INFRASTRUCTURE ONLY (see orc_sl.hpp).
#include "orc_sl.hpp"

#include <algorithm>
#include <set>
#include <cmath>
#include <stdexcept>

namespace orc {

void FEQSpace::init(const Mesh& m, int order) {
  mesh = m;
  p = order;
  gll_01(p, gll, gllw);
  lag.init(gll);
  for (int a = 0; a < 3; a++) n[a] = (a < mesh.D) ? mesh.N[a] * p + 1 : 1;
  const int64_t nd = n_dofs();
  x.resize(nd);
  owner.resize(nd);
  for (int k = 0; k < n[2]; k++)
    for (int j = 0; j < n[1]; j++)
      for (int i = 0; i < n[0]; i++) {
        int idx[3] = {i, j, k};
        int c[3] = {0, 0, 0}, mloc[3] = {0, 0, 0};
        for (int a = 0; a < mesh.D; a++) {
          // lowest-index cell containing the node (lexicographic cell order)
          c[a] = idx[a] == 0 ? 0 : (idx[a] - 1) / p;
          mloc[a] = idx[a] - c[a] * p;
        }
        int cell = mesh.cell_index(c);
        double xi[3] = {gll[mloc[0]], gll[mloc[1]], mesh.D == 3 ? gll[mloc[2]] : 0.0};
        double xx[3];
        mesh.map(cell, xi, xx);
        int64_t d = dof(i, j, k);
        x[d] = {xx[0], xx[1], xx[2]};
        owner[d] = cell;
      }
}

int64_t FEQSpace::cell_dof(int cell, int m0, int m1, int m2) const {
  int c[3];
  mesh.cell_coords(cell, c);
  return dof(c[0] * p + m0, c[1] * p + m1, mesh.D == 3 ? c[2] * p + m2 : 0);
}

void FEQSpace::shape_values(const double xi[3], double* out) const {
  const int np = p + 1;
  double vx[16], vy[16], vz[16];
  for (int m = 0; m < np; m++) {
    vx[m] = lag.value(m, xi[0]);
    vy[m] = lag.value(m, xi[1]);
    vz[m] = mesh.D == 3 ? lag.value(m, xi[2]) : 1.0;
  }
  const int nz = mesh.D == 3 ? np : 1;
  int o = 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < np; j++)
      for (int i = 0; i < np; i++) out[o++] = (mesh.D == 3 ? vx[i] * vy[j] * vz[k] : vx[i] * vy[j]);
}

int neighbor_cell(const Mesh& mesh, int cell, int f, bool* wrapped) {
  int c[3];
  mesh.cell_coords(cell, c);
  const int a = f / 2, side = f % 2;
  *wrapped = false;
  c[a] += side ? 1 : -1;
  if (c[a] < 0 || c[a] >= mesh.N[a]) {
    if (!mesh.periodic[a]) return -1;
    c[a] = c[a] < 0 ? mesh.N[a] - 1 : 0;
    *wrapped = true;
  }
  return mesh.cell_index(c);
}

void across_periodic(const Mesh& mesh, int cell, int f, const double p[3], double out[3]) {
  // offset = (vertex 0 of the lower periodic face) - (vertex 0 of the upper face),
  // both at this cell's transverse lattice position (PeriodicBoundary.cpp:261-270).
  int c[3];
  mesh.cell_coords(cell, c);
  const int a = f / 2;
  int lo[3] = {c[0], c[1], c[2]}, hi[3] = {c[0], c[1], c[2]};
  lo[a] = 0;
  hi[a] = mesh.N[a];
  const auto& vlo = mesh.vert[mesh.vidx(lo[0], lo[1], mesh.D == 3 ? lo[2] : 0)];
  const auto& vhi = mesh.vert[mesh.vidx(hi[0], hi[1], mesh.D == 3 ? hi[2] : 0)];
  for (int d = 0; d < 3; d++) {
    double off = d < mesh.D ? vlo[d] - vhi[d] : 0.0;
    if (f % 2 == 0)
      out[d] = p[d] - off;  // lower face: this cell is cell[0] of the pair
    else
      out[d] = p[d] + off;
  }
}

static void face_normal(const Mesh& mesh, int cell, int f, double pnt[3], double nrm[3]) {
  // some point on the face and some normal vector (SemiLagrangianTools.cpp:150-168)
  const int D = mesh.D;
  const int a = f / 2, side = f % 2;
  int fv[4], nfv = 0;
  for (int v = 0; v < mesh.n_cell_vertices(); v++)
    if (((v >> a) & 1) == side) fv[nfv++] = v;
  const auto& v0 = mesh.cell_vertex(cell, fv[0]);
  const auto& v1 = mesh.cell_vertex(cell, fv[1]);
  for (int d = 0; d < 3; d++) pnt[d] = v0[d];
  if (D == 2) {
    double t0 = v1[0] - v0[0], t1 = v1[1] - v0[1];
    nrm[0] = t1;
    nrm[1] = -t0;
    nrm[2] = 0;
  } else {
    const auto& v2 = mesh.cell_vertex(cell, fv[2]);
    double d1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    double d2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    nrm[0] = d1[1] * d2[2] - d1[2] * d2[1];
    nrm[1] = d1[2] * d2[0] - d1[0] * d2[2];
    nrm[2] = d1[0] * d2[1] - d1[1] * d2[0];
  }
}

static int face_crossed_first_without_mapping(const Mesh& mesh, int cell, const double pin[3], const double pout[3],
                                              double pb[3], double* lambda) {
  double lambda_min = 1;
  int index_min = 0;
  for (int f = 0; f < 2 * mesh.D; f++) {
    double pnt[3], nrm[3];
    face_normal(mesh, cell, f, pnt, nrm);
    double num = 0, den = 0;
    for (int d = 0; d < mesh.D; d++) {
      num += (pin[d] - pnt[d]) * nrm[d];
      den += (pin[d] - pout[d]) * nrm[d];
    }
    double lam = num / den;
    if (lam < 0) continue;
    if (lam < lambda_min) {
      lambda_min = lam;
      index_min = f;
    }
  }
  if (lambda_min == 1) return -1;
  *lambda = lambda_min;
  for (int d = 0; d < 3; d++) pb[d] = pin[d] + (pout[d] - pin[d]) * lambda_min;
  return index_min;
}

int face_crossed_first(const Mesh& mesh, int cell, const double pin[3], const double pout[3], double pb[3],
                       double* lambda, bool without_mapping) {
  if (without_mapping) return face_crossed_first_without_mapping(mesh, cell, pin, pout, pb, lambda);
  const int D = mesh.D;
  double pi[3], po[3];
  if (!mesh.inverse_map(cell, pin, pi)) throw std::runtime_error("inverse map failed");
  if (!mesh.inverse_map(cell, pout, po)) throw std::runtime_error("inverse map failed");
  for (int i = 0; i < D; i++) {
    clamp_unit(pi[i]);
    clamp_unit(po[i]);
  }
  int face = -1;
  *lambda = 100;
  // x faces first; later axes replace only on strictly smaller lambda
  for (int a = 0; a < D; a++) {
    double lam;
    int fid;
    if (po[a] < 0) {
      lam = (0 - pi[a]) / (po[a] - pi[a]);
      fid = 2 * a;
    } else if (po[a] > 1) {
      lam = (1 - pi[a]) / (po[a] - pi[a]);
      fid = 2 * a + 1;
    } else {
      continue;
    }
    if (a == 0 || lam < *lambda) {
      *lambda = lam;
      face = fid;
    }
  }
  if (face == -1) return -1;
  double h[3] = {0, 0, 0};
  for (int i = 0; i < D; i++) {
    h[i] = pi[i] + (po[i] - pi[i]) * (*lambda);
    clamp_unit(h[i]);
  }
  mesh.map(cell, h, pb);
  return face;
}

int locate_departure(const FEQSpace& V, const Stencil& st, double dt, int64_t dof, int alpha, double xi[3],
                     int* src_dir, std::vector<WallHit>* hits) {
  const Mesh& mesh = V.mesh;
  double cur[3] = {V.x[dof][0], V.x[dof][1], V.x[dof][2]};
  double dep[3];
  for (int d = 0; d < 3; d++) dep[d] = cur[d] + st.dir[alpha][d] * (-dt);
  int cell = V.owner[dof];
  int dir = alpha;
  if (src_dir) *src_dir = alpha;
  int life = 0;
  while (true) {
    life++;
    if (life > 50) return -1;
    double pb[3], lambda = 100;
    int face;
    try {
      face = face_crossed_first(mesh, cell, cur, dep, pb, &lambda, false);
    } catch (std::runtime_error&) {
      face = face_crossed_first(mesh, cell, cur, dep, pb, &lambda, true);
    }
    if (face == -1) break;
    bool wrapped;
    int nb = neighbor_cell(mesh, cell, face, &wrapped);
    if (nb < 0) {
      // Velocity-neq bounce-back wall (SemiLagrangian.cpp:358-385): the path turns
      // into the opposite direction at the hit point p_b and continues for the
      // remaining distance; the cell does not change.
      dir = st.opp[dir];
      double vel = 0, dist = 0;
      for (int d = 0; d < mesh.D; d++) {
        vel += st.dir[dir][d] * st.dir[dir][d];
        dist += (dep[d] - pb[d]) * (dep[d] - pb[d]);
      }
      vel = std::sqrt(vel);
      dist = std::sqrt(dist);
      for (int d = 0; d < mesh.D; d++) {
        cur[d] = pb[d];
        dep[d] = pb[d] - st.dir[dir][d] * dist / vel;
      }
      if (hits) hits->push_back({face, cell});
      continue;
    }
    if (wrapped) {
      double c2[3], d2[3];
      across_periodic(mesh, cell, face, pb, c2);
      across_periodic(mesh, cell, face, dep, d2);
      for (int d = 0; d < 3; d++) {
        cur[d] = c2[d];
        dep[d] = d2[d];
      }
    } else {
      for (int d = 0; d < 3; d++) cur[d] = pb[d];
    }
    cell = nb;
  }
  if (src_dir) *src_dir = dir;
  if (!mesh.inverse_map(cell, dep, xi)) throw std::runtime_error("inverse map failed at departure point");
  for (int d = 0; d < mesh.D; d++) {
    clamp_unit(xi[d]);
    if (xi[d] < 0 || xi[d] > 1) throw std::runtime_error("departure point outside its cell");
  }
  if (mesh.D == 2) xi[2] = 0;
  return cell;
}

void assemble_sl(const FEQSpace& V, const Stencil& st, double dt, SLOperator& op) {
  const int64_t n = V.n_dofs();
  const int nl = V.n_local();
  const int np = V.p + 1;
  op.Q = st.Q;
  op.n = n;
  op.block.assign(st.Q - 1, CSR());
  op.bvec.assign(st.Q - 1, std::vector<double>());
  op.n_hits = 0;
  std::set<int> hit_cells;
  int64_t n_ident = 0;
  for (int alpha = 1; alpha < st.Q; alpha++) {
    CSR& A = op.block[alpha - 1];
    std::vector<int> cellv(n), srcv(n);
    std::vector<std::array<double, 3>> xiv(n);
    std::vector<std::vector<WallHit>> hitv(n);
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : n_ident)
    for (int64_t i = 0; i < n; i++) {
      double xi[3];
      cellv[i] = locate_departure(V, st, dt, i, alpha, xi, &srcv[i], &hitv[i]);
      xiv[i] = {xi[0], xi[1], xi[2]};
      if (cellv[i] < 0) n_ident++;
    }
    // VelocityNeqBounceBack::calculateBoundaryValues: per hit, 2 w_alpha rho (e_alpha . u_w) / cs2
    // with rho = 1 and the destination direction alpha (scaled e and cs2)
    bool any_hit = false, any_flip = false;
    for (int64_t i = 0; i < n; i++) {
      if (!hitv[i].empty()) any_hit = true;
      if (cellv[i] >= 0 && srcv[i] != alpha) any_flip = true;
    }
    if (any_hit) {
      std::vector<double>& b = op.bvec[alpha - 1];
      b.assign(n, 0.0);
      for (int64_t i = 0; i < n; i++)
        for (const WallHit& h : hitv[i]) {
          double exu = 0;
          for (int d = 0; d < st.D; d++) exu += V.mesh.wall_u[h.face][d] * st.dir[alpha][d];
          b[i] = b[i] + 2 * st.w[alpha] * 1.0 * exu / st.cs2;
          op.n_hits++;
          hit_cells.insert(h.cell);
        }
    }
    A.rowptr.assign(n + 1, 0);
    std::vector<int> cnt(n, 0);
    std::vector<double> vals((size_t)n * nl);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      if (cellv[i] < 0) {
        cnt[i] = 1;
        continue;
      }
      double* v = &vals[(size_t)i * nl];
      V.shape_values(xiv[i].data(), v);
      int c = 0;
      for (int j = 0; j < nl; j++)
        if (!(std::fabs(v[j]) < 1e-10)) c++;
      cnt[i] = c;
    }
    for (int64_t i = 0; i < n; i++) A.rowptr[i + 1] = A.rowptr[i] + cnt[i];
    A.col.resize(A.rowptr[n]);
    A.val.resize(A.rowptr[n]);
    if (any_flip) A.src.assign(A.rowptr[n], (int8_t)alpha);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      int64_t o = A.rowptr[i];
      if (cellv[i] < 0) {
        // life-time limit: 1 on the diagonal of block (alpha, alpha)
        A.col[o] = i;
        A.val[o] = 1.0;
        continue;
      }
      const double* v = &vals[(size_t)i * nl];
      // entries sorted by global column
      std::vector<std::pair<int64_t, double>> e;
      e.reserve(nl);
      int t = 0;
      const int nz = V.mesh.D == 3 ? np : 1;
      for (int k = 0; k < nz; k++)
        for (int jj = 0; jj < np; jj++)
          for (int ii = 0; ii < np; ii++, t++)
            if (!(std::fabs(v[t]) < 1e-10)) e.push_back({V.cell_dof(cellv[i], ii, jj, k), v[t]});
      std::sort(e.begin(), e.end());
      for (auto& pr : e) {
        A.col[o] = pr.first;
        A.val[o] = pr.second;
        if (any_flip) A.src[o] = (int8_t)srcv[i];
        o++;
      }
    }
  }
  op.n_identity_rows = n_ident;
  op.n_hit_cells = (int64_t)hit_cells.size();
}

void sl_apply(const SLOperator& op, const double* x, double* y, bool with_boundary) {
  const int64_t n = op.n;
  for (int a = 1; a < op.Q; a++) {
    const CSR& A = op.block[a - 1];
    const std::vector<double>& b = op.bvec[a - 1];
    const bool flip = !A.src.empty();
    double* ya = y + (size_t)a * n;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      double s = 0;
      for (int64_t k = A.rowptr[i]; k < A.rowptr[i + 1]; k++) {
        const int src = flip ? A.src[k] : a;
        s += A.val[k] * x[(size_t)src * n + A.col[k]];
      }
      if (with_boundary && !b.empty()) s = s + b[i];
      ya[i] = s;
    }
  }
}

void csr_vmult(const CSR& A, const double* x, double* y, int64_t n) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    double s = 0;
    for (int64_t k = A.rowptr[i]; k < A.rowptr[i + 1]; k++) s += A.val[k] * x[A.col[k]];
    y[i] = s;
  }
}

}  // namespace orc
