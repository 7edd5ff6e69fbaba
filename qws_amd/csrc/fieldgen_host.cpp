// fieldgen_host.cpp -- host-side constants of the clover construction.
#include <cmath>
#include <complex>
#include <stdexcept>

#include "fieldgen.h"

namespace qws {

// gamma matrices of QWS's projector tables (wilson_d.h:540-657): gamma_4 =
// diag(1,1,-1,-1), gamma_k = [[0, i sigma_k], [-i sigma_k, 0]].
static void gamma_matrices(std::complex<double> g[4][4][4]) {
  using C = std::complex<double>;
  const C I(0, 1);
  const C pauli[3][2][2] = {{{0, 1}, {1, 0}}, {{0, -I}, {I, 0}}, {{1, 0}, {0, -1}}};
  for (int m = 0; m < 4; ++m)
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) g[m][a][b] = 0;
  for (int k = 0; k < 3; ++k)
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) {
        g[k][a][b + 2] = I * pauli[k][a][b];
        g[k][a + 2][b] = -I * pauli[k][a][b];
      }
  g[3][0][0] = g[3][1][1] = 1;
  g[3][2][2] = g[3][3][3] = -1;
}

// sigma_{mu nu} = (i/2)[gamma_mu, gamma_nu] in the chirality basis of QWS's
// clover packing: e_{i,p} = (e_p +- e_{p+2})/sqrt2 for chirality i = 0 (+), 1 (-).
SigmaChiral sigma_chiral_blocks() {
  using C = std::complex<double>;
  C g[4][4][4];
  gamma_matrices(g);
  SigmaChiral out{};
  const double r = 1.0 / std::sqrt(2.0);
  int k = 0;
  for (int mu = 0; mu < 4; ++mu)
    for (int nu = mu + 1; nu < 4; ++nu, ++k) {
      C sig[4][4];
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
          C comm = 0;
          for (int c = 0; c < 4; ++c) comm += g[mu][a][c] * g[nu][c][b] - g[nu][a][c] * g[mu][c][b];
          sig[a][b] = C(0, 0.5) * comm;
        }
      double e[2][2][4] = {};  // chirality, pair, spin
      for (int i = 0; i < 2; ++i)
        for (int p = 0; p < 2; ++p) {
          e[i][p][p] = r;
          e[i][p][p + 2] = i == 0 ? r : -r;
        }
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
          for (int p = 0; p < 2; ++p)
            for (int q = 0; q < 2; ++q) {
              C v = 0;
              for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) v += e[i][p][a] * sig[a][b] * e[j][q][b];
              if (i == j) {
                out.v[k][i][p][q][0] = v.real();
                out.v[k][i][p][q][1] = v.imag();
              } else if (std::abs(v) > 1e-14) {
                throw std::logic_error("sigma_{mu nu} is not block diagonal in the chirality basis");
              }
            }
    }
  return out;
}

}  // namespace qws
