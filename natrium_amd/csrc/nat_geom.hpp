// Multilinear (Q1) or axis-aligned cell geometry on the device.
#pragma once
#include "nat_internal.hpp"

namespace nat {

// Jacobian d x / d xi of cell c at unit coordinates xi (MappingQ1, or
// MappingCartesian from the first and last vertex when m.cartesian). D is the mesh
// dimension (m.D) as a constant, so J stays in registers.
template <int D>
__device__ __forceinline__ void cell_jac(const MeshDev& m, const int c[3], const double xi[3], double J[3][3]) {
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) J[a][b] = (a == b && a >= D) ? 1.0 : 0.0;
  constexpr int nv = D == 3 ? 8 : 4;
  auto V = [&](int v) {
    return m.vert + 3 * ((size_t)(c[0] + (v & 1)) +
                         (size_t)(m.N[0] + 1) * ((size_t)(c[1] + ((v >> 1) & 1)) +
                                                 (size_t)(m.N[1] + 1) * (D == 3 ? c[2] + ((v >> 2) & 1) : 0)));
  };
  if (m.cartesian) {
    const double* a0 = V(0);
    const double* b0 = V(nv - 1);
#pragma unroll
    for (int d = 0; d < D; d++) J[d][d] = b0[d] - a0[d];
    return;
  }
#pragma unroll
  for (int v = 0; v < nv; v++) {
    const double* X = V(v);
#pragma unroll
    for (int b = 0; b < D; b++) {
      double dN = 1;
#pragma unroll
      for (int a = 0; a < D; a++) {
        const int bit = (v >> a) & 1;
        dN *= (a == b) ? (bit ? 1.0 : -1.0) : (bit ? xi[a] : 1 - xi[a]);
      }
#pragma unroll
      for (int d = 0; d < D; d++) J[d][b] += dN * X[d];
    }
  }
}

// Inverse and determinant of a D x D Jacobian (upper-left block of J).
__device__ __forceinline__ double inv_det(int D, const double J[3][3], double Ji[3][3]) {
  if (D == 2) {
    const double d = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    const double r = 1.0 / d;
    Ji[0][0] = J[1][1] * r;
    Ji[0][1] = -J[0][1] * r;
    Ji[1][0] = -J[1][0] * r;
    Ji[1][1] = J[0][0] * r;
    return d;
  }
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const double d = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  const double r = 1.0 / d;
  Ji[0][0] = c00 * r;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r;
  Ji[1][0] = c01 * r;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * r;
  Ji[2][0] = c02 * r;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r;
  return d;
}

}  // namespace nat
