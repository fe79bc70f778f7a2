// Oracle SEDG + RK5. TEST INFRASTRUCTURE ONLY (see orc_sedg.hpp).
#include "orc_sedg.hpp"

#include <algorithm>
#include <cmath>

namespace orc {

// Carpenter-Kennedy (1994) 2N-storage RK5 coefficients as written in
// timeintegration/RungeKutta5LowStorage.h:29-46
const double RK5_A[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                         -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
const double RK5_B[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                         1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                         2277821191437.0 / 14882151754819.0};

void DGSpace::init(const Mesh& m, int order) {
  mesh = m;
  p = order;
  gll_01(p, gll, gllw);
  lag.init(gll);
  const int np = p + 1;
  nl = mesh.D == 3 ? np * np * np : np * np;
  x.resize(n_dofs());
  for (int c = 0; c < mesh.n_cells(); c++)
    for (int l = 0; l < nl; l++) {
      int m0 = l % np, m1 = (l / np) % np, m2 = l / (np * np);
      double xi[3] = {gll[m0], gll[m1], mesh.D == 3 ? gll[m2] : 0.0};
      double xx[3];
      mesh.map(c, xi, xx);
      x[(int64_t)c * nl + l] = {xx[0], xx[1], xx[2]};
    }
}

static void inv_and_det(const double J[3][3], int D, double Ji[3][3], double* det) {
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) Ji[a][b] = 0;
  if (D == 2) {
    double d = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    Ji[0][0] = J[1][1] / d;
    Ji[0][1] = -J[0][1] / d;
    Ji[1][0] = -J[1][0] / d;
    Ji[1][1] = J[0][0] / d;
    *det = d;
    return;
  }
  double d = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
             J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / d;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / d;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / d;
  Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / d;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / d;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / d;
  Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / d;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / d;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / d;
  *det = d;
}

void assemble_sedg(const DGSpace& V, const Stencil& st, bool central_flux, SLOperator& op) {
  const Mesh& mesh = V.mesh;
  const int D = mesh.D, np = V.p + 1, nl = V.nl;
  const int64_t n = V.n_dofs();
  op.Q = st.Q;
  op.n = n;
  op.block.assign(st.Q - 1, CSR());
  auto lidx = [&](int m0, int m1, int m2) { return m0 + np * (m1 + np * m2); };
  // row length: nl volume columns + one neighbour column per face touching the node
  std::vector<int> rowlen(nl);
  for (int i = 0; i < nl; i++) {
    int mi[3] = {i % np, (i / np) % np, D == 3 ? i / (np * np) : 0};
    int nf = 0;
    for (int a = 0; a < D; a++) nf += (mi[a] == 0 || mi[a] == V.p) ? 1 : 0;
    rowlen[i] = nl + nf;
  }
  // walls on non-periodic axes: the blocks (alpha, opposite alpha) and the system vector
  bool walls = false;
  for (int a = 0; a < D; a++) walls = walls || !mesh.periodic[a];
  op.opp_block.assign(st.Q - 1, -1);
  op.opp_diag.assign(walls ? st.Q - 1 : 0, std::vector<double>());
  op.bvec.assign(st.Q - 1, std::vector<double>());
  for (int alpha = 1; alpha < st.Q; alpha++) {
    op.opp_block[alpha - 1] = st.opp[alpha] - 1;
    if (walls) {
      op.opp_diag[alpha - 1].assign(n, 0.0);
      op.bvec[alpha - 1].assign(n, 0.0);
    }
  }
  for (int alpha = 1; alpha < st.Q; alpha++) {
    CSR& A = op.block[alpha - 1];
    A.rowptr.assign(n + 1, 0);
    for (int64_t r = 0; r < n; r++) A.rowptr[r + 1] = A.rowptr[r] + rowlen[r % nl];
    A.col.resize(A.rowptr[n]);
    A.val.resize(A.rowptr[n]);
  }

#pragma omp parallel for schedule(dynamic, 4)
  for (int c = 0; c < mesh.n_cells(); c++) {
    std::vector<double> JxW(nl), invM(nl);
    std::vector<double> Dm[3];
    for (int k = 0; k < 3; k++) Dm[k].assign((size_t)nl * nl, 0.0);
    for (int i = 0; i < nl; i++) {
      int mi[3] = {i % np, (i / np) % np, D == 3 ? i / (np * np) : 0};
      double xi[3] = {V.gll[mi[0]], V.gll[mi[1]], D == 3 ? V.gll[mi[2]] : 0.0};
      double J[3][3], Ji[3][3], det;
      mesh.jacobian(c, xi, J);
      inv_and_det(J, D, Ji, &det);
      double wq = V.gllw[mi[0]] * V.gllw[mi[1]] * (D == 3 ? V.gllw[mi[2]] : 1.0);
      JxW[i] = det * wq;
      double ph = 1;
      for (int a = 0; a < D; a++) ph *= V.lag.value(mi[a], xi[a]);
      // mass (SEDGMinLee.cpp:321-334), derivative (SEDGMinLee.cpp:337-358)
      invM[i] = 1. / (ph * ph * JxW[i]);
      for (int j = 0; j < nl; j++) {
        int mj[3] = {j % np, (j / np) % np, D == 3 ? j / (np * np) : 0};
        double ug[3] = {0, 0, 0};
        for (int b = 0; b < D; b++) {
          double t = 1;
          for (int a = 0; a < D; a++) t *= (a == b) ? V.lag.deriv(mj[a], xi[a]) : V.lag.value(mj[a], xi[a]);
          ug[b] = t;
        }
        for (int k = 0; k < D; k++) {
          double g = 0;
          for (int b = 0; b < D; b++) g += Ji[b][k] * ug[b];
          Dm[k][(size_t)i * nl + j] += g * (ph * JxW[i]);
        }
      }
    }
    // face geometry per (face, q): neighbour DoF, local DoF, JxW*normal
    struct FaceQ {
      int li;
      int64_t gn;  // neighbour's coincident node, -1 on a wall
      int face;    // box face 2*axis+side (walls)
      double nrm[3];
      double jxw;
    };
    std::vector<FaceQ> fq;
    std::vector<int> fq_face;
    for (int f = 0; f < 2 * D; f++) {
      const int a = f / 2, side = f % 2;
      bool wrapped;
      int nb = neighbor_cell(mesh, c, f, &wrapped);  // -1: a wall of a non-periodic axis
      const int tb = (a == 0) ? 1 : 0;
      const int tc = (a == 2) ? 1 : 2;
      const int nq = D == 3 ? np * np : np;
      for (int q = 0; q < nq; q++) {
        int m[3] = {0, 0, 0};
        m[a] = side ? V.p : 0;
        m[tb] = q % np;
        if (D == 3) m[tc] = q / np;
        int mn[3] = {m[0], m[1], m[2]};
        mn[a] = side ? 0 : V.p;
        double xi[3] = {V.gll[m[0]], V.gll[m[1]], D == 3 ? V.gll[m[2]] : 0.0};
        double J[3][3], Ji[3][3], det;
        mesh.jacobian(c, xi, J);
        inv_and_det(J, D, Ji, &det);
        FaceQ F;
        F.li = lidx(m[0], m[1], m[2]);
        F.gn = nb < 0 ? -1 : (int64_t)nb * nl + lidx(mn[0], mn[1], mn[2]);
        F.face = f;
        double len = 0;
        const double sgn = side ? 1.0 : -1.0;
        for (int k = 0; k < 3; k++) F.nrm[k] = 0;
        for (int k = 0; k < D; k++) {
          F.nrm[k] = sgn * det * Ji[a][k];
          len += F.nrm[k] * F.nrm[k];
        }
        len = std::sqrt(len);
        for (int k = 0; k < D; k++) F.nrm[k] /= len;
        F.jxw = len * (V.gllw[m[tb]] * (D == 3 ? V.gllw[m[tc]] : 1.0));
        fq.push_back(F);
      }
    }
    std::vector<std::pair<int64_t, double>> row;
    for (int alpha = 1; alpha < st.Q; alpha++) {
      CSR& A = op.block[alpha - 1];
      const double* e = st.dir[alpha].data();
      for (int i = 0; i < nl; i++) {
        row.clear();
        const int64_t gi = (int64_t)c * nl + i;
        // cell block: -(e . D) M^-1 (SEDGMinLee.cpp:414-458)
        for (int j = 0; j < nl; j++) {
          size_t ij = (size_t)i * nl + j;
          double s = Dm[0][ij] * (-e[0]);
          if (D == 2)
            s += (-e[1]) * Dm[1][ij];
          else
            s += (-e[1]) * Dm[1][ij] + (-e[2]) * Dm[2][ij];
          row.push_back({(int64_t)c * nl + j, s * invM[i]});
        }
        // faces in face order (SEDGMinLee.cpp:461-536)
        for (const FaceQ& F : fq) {
          if (F.li != i) continue;
          double prefactor = F.jxw;
          double exn = 0.0;
          for (int k = 0; k < D; k++) exn += F.nrm[k] * e[k];
          prefactor *= exn;
          double cell_entry = 0.0, neighbor_entry = 0.0;
          if (F.gn < 0) {
            // velocity-neq bounce-back wall (VelocityNeqBounceBack.cpp:46-135): density 1, the
            // wall velocity of the face; the other side of the face is the opposite direction
            // at the same node, and the wall velocity enters the system vector
            double exu = 0.0;
            for (int k = 0; k < D; k++) exu += mesh.wall_u[F.face][k] * e[k];
            const double density = 1.0;
            double opp_entry = 0.0, vec_entry = 0.0;
            if (central_flux) {
              cell_entry = 0.5 * prefactor;
              opp_entry = 0.5 * prefactor;
              vec_entry = -prefactor * st.w[alpha] * density * exu / st.cs2;
            } else if (exn < 0) {
              cell_entry = prefactor;
              opp_entry = -prefactor;
              vec_entry = -2 * prefactor * st.w[alpha] * density * exu / st.cs2;
            }
            row[i].second += cell_entry * invM[i];
            op.opp_diag[alpha - 1][gi] += opp_entry * invM[i];
            op.bvec[alpha - 1][gi] += vec_entry * invM[i];
            continue;
          }
          if (central_flux) {
            cell_entry = 0.5 * prefactor;
            neighbor_entry = 0.5 * prefactor;
          } else if (exn < 0) {
            cell_entry = prefactor;
            neighbor_entry = -prefactor;
          }
          row[i].second += cell_entry * invM[i];
          row.push_back({F.gn, neighbor_entry * invM[i]});
        }
        std::sort(row.begin(), row.end(), [](const std::pair<int64_t, double>& x, const std::pair<int64_t, double>& y) {
          return x.first < y.first;
        });
        int64_t o = A.rowptr[gi];
        for (auto& kv : row) {
          A.col[o] = kv.first;
          A.val[o] = kv.second;
          o++;
        }
        for (; o < A.rowptr[gi + 1]; o++) {  // slots of wall faces (no neighbour entry)
          A.col[o] = gi;
          A.val[o] = 0.0;
        }
      }
    }
  }
}

void rk5_step(const SLOperator& op, double dt, double* const* f, std::vector<double>& work) {
  // RungeKutta5LowStorage.cpp:56-105, operation by operation; the product is the block
  // matrix's (diagonal blocks, plus the (alpha, opposite alpha) diagonals of wall faces)
  const int nb = op.Q - 1;
  const int64_t n = op.n;
  work.resize((size_t)2 * nb * n);
  double* Af = work.data();
  double* df = work.data() + (size_t)nb * n;
  const bool walls = !op.opp_diag.empty();
  for (int s = 0; s < 5; s++) {
    for (int b = 0; b < nb; b++) csr_vmult(op.block[b], f[b], Af + (size_t)b * n, n);
    if (walls) {  // blocks (alpha, opposite alpha) of the wall faces: diagonal
#pragma omp parallel for schedule(static)
      for (int64_t t = 0; t < (int64_t)nb * n; t++) {
        const int b = (int)(t / n);
        const int64_t i = t % n;
        Af[t] += op.opp_diag[b][i] * f[op.opp_block[b]][i];
      }
    }
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < (int64_t)nb * n; t++) {
      const int b = (int)(t / n);
      const int64_t i = t % n;
      double af = Af[t];
      if (walls) af += op.bvec[b][i];  // m_Af += systemVector (zero without walls)
      af *= dt;
      double d;
      if (s == 0) {
        d = af;
      } else {
        d = df[t];
        d *= RK5_A[s];
        d += af;
      }
      d *= RK5_B[s];
      f[b][i] += d;
      d *= 1.0 / RK5_B[s];  // operator/= of deal.II / Trilinos vectors: times the reciprocal
      df[t] = d;
    }
  }
}

}  // namespace orc
