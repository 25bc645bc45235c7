// Reference-element basis of the reference's mixed element: transverse
// "P2 x grad P1" edge/face functions (12 per triangle) and P3 nodal functions
// for E_z (10 per triangle), plus the 16-point degree-8 triangle rule.
//
// Formulas follow basis_ls.f:15-89, basis_vec.f:17-71, phi1_2d_mat.f,
// phi2_2d_mat.f, phi3_2d_mat.f and quad_triangle.f:42-145 of the reference
// (backend/fortran/).  Shared by host and device code.
#pragma once
#include <hip/hip_runtime.h>

namespace emu {

#define EMU_HD __host__ __device__ __forceinline__

struct QuadRule {
  double x[16], y[16], w[16];
};

EMU_HD void quad16(int q, double &x, double &y, double &w) {
  const double wbar = 1.443156076777862e-1 / 2.0;
  const double w1 = 2.852749028018549e-1 / 6.0;
  const double w2 = 9.737549286959440e-2 / 6.0;
  const double w3 = 3.096521116041552e-1 / 6.0;
  const double w4 = 1.633818850466092e-1 / 1.2e1;
  const double cb = 1.0 / 3.0;
  const double a1 = 4.592925882927229e-1, b1 = 8.141482341455413e-2;
  const double a2 = 5.054722831703103e-2, b2 = 8.989055433659379e-1;
  const double a3 = 1.705693077517601e-1, b3 = 6.588613844964797e-1;
  const double a4 = 7.284923929554041e-1, b4 = 2.63112829634638689e-1, c4 = 8.394777409957211e-3;
  switch (q) {
    case 0: x = cb; y = cb; w = wbar; break;
    case 1: x = a1; y = a1; w = w1; break;
    case 2: x = a1; y = b1; w = w1; break;
    case 3: x = b1; y = a1; w = w1; break;
    case 4: x = a2; y = a2; w = w2; break;
    case 5: x = a2; y = b2; w = w2; break;
    case 6: x = b2; y = a2; w = w2; break;
    case 7: x = a3; y = a3; w = w3; break;
    case 8: x = a3; y = b3; w = w3; break;
    case 9: x = b3; y = a3; w = w3; break;
    case 10: x = a4; y = b4; w = w4; break;
    case 11: x = b4; y = a4; w = w4; break;
    case 12: x = b4; y = c4; w = w4; break;
    case 13: x = c4; y = b4; w = w4; break;
    case 14: x = c4; y = a4; w = w4; break;
    default: x = a4; y = c4; w = w4; break;
  }
}

// P1: value and reference gradient
EMU_HD void ref_p1(double x, double y, double p[3], double gx[3], double gy[3]) {
  p[0] = 1 - x - y; gx[0] = -1; gy[0] = -1;
  p[1] = x;         gx[1] = 1;  gy[1] = 0;
  p[2] = y;         gx[2] = 0;  gy[2] = 1;
}

EMU_HD void ref_p2(double x, double y, double p[6], double gx[6], double gy[6]) {
  p[0] = (-1 + x + y) * (-1 + 2 * x + 2 * y);
  p[1] = x * (-1 + 2 * x);
  p[2] = y * (-1 + 2 * y);
  p[3] = -4 * x * (-1 + x + y);
  p[4] = 4 * x * y;
  p[5] = -4 * y * (-1 + x + y);
  gx[0] = -3 + 4 * x + 4 * y; gy[0] = -3 + 4 * x + 4 * y;
  gx[1] = -1 + 4 * x;         gy[1] = 0;
  gx[2] = 0;                  gy[2] = -1 + 4 * y;
  gx[3] = -4 * (-1 + 2 * x + y); gy[3] = -4 * x;
  gx[4] = 4 * y;              gy[4] = 4 * x;
  gx[5] = -4 * y;             gy[5] = -4 * (-1 + x + 2 * y);
}

EMU_HD void ref_p3(double x, double y, double p[10], double gx[10], double gy[10]) {
  p[0] = -((-1 + x + y) * (-2 + 3 * x + 3 * y) * (-1 + 3 * x + 3 * y)) / 2.;
  p[1] = (x * (-2 + 3 * x) * (-1 + 3 * x)) / 2.;
  p[2] = (y * (-2 + 3 * y) * (-1 + 3 * y)) / 2.;
  p[3] = (9 * x * (-1 + x + y) * (-2 + 3 * x + 3 * y)) / 2.;
  p[4] = (-9 * x * (-1 + 3 * x) * (-1 + x + y)) / 2.;
  p[5] = (9 * x * (-1 + 3 * x) * y) / 2.;
  p[6] = (9 * x * y * (-1 + 3 * y)) / 2.;
  p[7] = (-9 * y * (-1 + x + y) * (-1 + 3 * y)) / 2.;
  p[8] = (9 * y * (-1 + x + y) * (-2 + 3 * x + 3 * y)) / 2.;
  p[9] = -27 * x * y * (-1 + x + y);
  double g1 = (-11 - 27 * x * x + x * (36 - 54 * y) + 36 * y - 27 * y * y) / 2.;
  gx[0] = g1;
  gx[1] = 1 - 9 * x + (27 * x * x) / 2.;
  gx[2] = 0;
  gx[3] = (9 * (2 + 9 * x * x - 5 * y + 3 * y * y + 2 * x * (-5 + 6 * y))) / 2.;
  gx[4] = (-9 * (1 + 9 * x * x - y + x * (-8 + 6 * y))) / 2.;
  gx[5] = (9 * (-1 + 6 * x) * y) / 2.;
  gx[6] = (9 * y * (-1 + 3 * y)) / 2.;
  gx[7] = (-9 * y * (-1 + 3 * y)) / 2.;
  gx[8] = (9 * y * (-5 + 6 * x + 6 * y)) / 2.;
  gx[9] = -27 * y * (-1 + 2 * x + y);
  gy[0] = g1;
  gy[1] = 0;
  gy[2] = 1 - 9 * y + (27 * y * y) / 2.;
  gy[3] = (9 * x * (-5 + 6 * x + 6 * y)) / 2.;
  gy[4] = (-9 * x * (-1 + 3 * x)) / 2.;
  gy[5] = (9 * x * (-1 + 3 * x)) / 2.;
  gy[6] = (9 * x * (-1 + 6 * y)) / 2.;
  gy[7] = (-9 * (1 - 8 * y + 9 * y * y + x * (-1 + 6 * y))) / 2.;
  gy[8] = (9 * (2 + 3 * x * x - 10 * y + 9 * y * y + x * (-5 + 12 * y))) / 2.;
  gy[9] = -27 * x * (-1 + x + 2 * y);
}

// Physical gradient: grad = T^T grad_ref with T = inverse of the affine map
// (jacobian_p1_2d.f, asmbly.f:194-199).  T is column-major: T00,T10,T01,T11.
EMU_HD void phys_grad(const double *T, double gx, double gy, double &ox, double &oy) {
  ox = T[0] * gx + T[1] * gy;
  oy = T[2] * gx + T[3] * gy;
}

// P2 node attached to each transverse function (basis_list(2,...)).
// orient[6]: per edge (1-2, 2-3, 1-3) the local vertices (j1, j2) ordered by
// global (periodic-representative) node number.
EMU_HD int transverse_node(int l, const signed char *orient) {
  if (l < 3) return l + 3;  // face functions: mid-edge nodes 4,5,6
  int e = (l - 3) / 3, q = (l - 3) % 3;
  if (q == 0) return orient[2 * e];
  if (q == 1) return orient[2 * e + 1];
  return e + 3;
}

// Transverse basis function l at reference point (x, y): vec (2) and
// curl_t (basis_vec.f:38-67).
EMU_HD void transverse_basis(int l, double x, double y, const double *T, const signed char *orient,
                             double &vx, double &vy, double &curl) {
  double p1[3], g1x[3], g1y[3], p2[6], g2x[6], g2y[6];
  ref_p1(x, y, p1, g1x, g1y);
  ref_p2(x, y, p2, g2x, g2y);
  int m, n1, n2 = -1;
  if (l < 3) {
    m = l + 3;
    const int opp[3] = {2, 0, 1};
    n1 = opp[l];
  } else {
    int e = (l - 3) / 3, q = (l - 3) % 3;
    int j1 = orient[2 * e], j2 = orient[2 * e + 1];
    if (q == 0) { m = j1; n1 = j2; }
    else if (q == 1) { m = j2; n1 = j1; }
    else { m = e + 3; n1 = j1; n2 = j2; }
  }
  double gpx, gpy, g2px, g2py;
  phys_grad(T, g1x[n1], g1y[n1], gpx, gpy);
  if (n2 >= 0) {
    double ax, ay;
    phys_grad(T, g1x[n2], g1y[n2], ax, ay);
    gpx -= ax;
    gpy -= ay;
  }
  phys_grad(T, g2x[m], g2y[m], g2px, g2py);
  vx = p2[m] * gpx;
  vy = p2[m] * gpy;
  curl = g2px * gpy - g2py * gpx;
}

#undef EMU_HD
}  // namespace emu
