// ke0.cpp — hex8 element stiffness Ke0 (24x24) by exact integration.
//
// The reference integrates B^T C B |J| with 2x2x2 Gauss points (element_kernel.cpp:155-179).
// Here the same integral is evaluated in closed form: for trilinear shape functions on a
// cube of edge h (a = h/2), with Lame constants lambda, mu,
//   K[3a+i, 3b+j] = a * ( lambda I(a,i;b,j) + mu I(a,j;b,i) + mu delta_ij sum_k I(a,k;b,k) )
//   I(a,p;b,q)    = int_{[-1,1]^3} dN_a/dxi_p dN_b/dxi_q dxi
// and I factorises per axis into 1D integrals of linear polynomials. Node order and the
// engineering-shear Voigt convention match element_kernel.cpp:36-50,132-151, so the result
// equals the reference's Ke0 to rounding (2-point Gauss is exact for this integrand).
#include <cmath>

namespace amt {

namespace {
// element_kernel.cpp:40-49: natural coordinates of the 8 nodes
const int kXi[8][3] = {{-1, -1, -1}, {+1, -1, -1}, {+1, +1, -1}, {-1, +1, -1},
                       {-1, -1, +1}, {+1, -1, +1}, {+1, +1, +1}, {-1, +1, +1}};

// 1D factor of int dN_a/dxi_p dN_b/dxi_q along axis t, with N = prod_t (1 + x_t xi_t)/2.
double axis_integral(int xa, int xb, bool da, bool db) {
  if (da && db) return xa * xb / 2.0;       // int (xa/2)(xb/2)
  if (da) return xa / 2.0;                  // int (xa/2)(1+xb s)/2
  if (db) return xb / 2.0;
  return (1.0 + xa * xb / 3.0) / 2.0;       // int (1+xa s)(1+xb s)/4
}

double I(int a, int p, int b, int q) {
  double v = 1.0;
  for (int t = 0; t < 3; ++t) v *= axis_integral(kXi[a][t], kXi[b][t], t == p, t == q);
  return v;
}
}  // namespace

void hex8_stiffness(double E, double nu, double h, double* Ke) {
  const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double mu = E / (2.0 * (1.0 + nu));
  const double a = h / 2.0;
  for (int na = 0; na < 8; ++na)
    for (int i = 0; i < 3; ++i)
      for (int nb = 0; nb < 8; ++nb)
        for (int j = 0; j < 3; ++j) {
          double v = lam * I(na, i, nb, j) + mu * I(na, j, nb, i);
          if (i == j) v += mu * (I(na, 0, nb, 0) + I(na, 1, nb, 1) + I(na, 2, nb, 2));
          Ke[(3 * na + i) * 24 + 3 * nb + j] = a * v;
        }
}

}  // namespace amt
