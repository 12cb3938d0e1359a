// Gate matrices (host, FP64), column-major vectorized: m[i + dim*j] = M(i, j).
//
// These are the textbook definitions of the standard gates in the form the
// reference evaluates them (src/framework/linalg/matrix_utils/
// vmatrix_defs.hpp:197-381 and matrix_defs.hpp:150-212); each entry is built
// from std::exp / std::cos / std::sin of the same arguments so that the FP64
// values agree bit for bit with the reference's.
#pragma once

#include <cmath>
#include <complex>
#include <vector>

namespace aerhip {
namespace gates {

using cd = std::complex<double>;
using cvec = std::vector<cd>;

inline cvec u3(double theta, double phi, double lambda) {
  const cd i(0., 1.);
  cvec m(4);
  m[0] = std::cos(0.5 * theta);
  m[2] = -std::exp(i * lambda) * std::sin(0.5 * theta);
  m[1] = std::exp(i * phi) * std::sin(0.5 * theta);
  m[3] = std::exp(i * (phi + lambda)) * std::cos(0.5 * theta);
  return m;
}

inline cvec u4(double theta, double phi, double lambda, double gamma) {
  const cd i(0., 1.);
  cvec m(4);
  m[0] = std::exp(i * gamma) * std::cos(0.5 * theta);
  m[2] = -std::exp(i * (lambda + gamma)) * std::sin(0.5 * theta);
  m[1] = std::exp(i * (phi + gamma)) * std::sin(0.5 * theta);
  m[3] = std::exp(i * (phi + lambda + gamma)) * std::cos(0.5 * theta);
  return m;
}

inline cvec u2(double phi, double lambda) {
  const cd i(0., 1.);
  const cd s(1. / std::sqrt(2), 0.);
  cvec m(4);
  m[0] = s;
  m[2] = -std::exp(i * lambda) * s;
  m[1] = std::exp(i * phi) * s;
  m[3] = std::exp(i * (phi + lambda)) * s;
  return m;
}

inline cvec r(double theta, double phi) {
  const cd i(0., 1.);
  cvec m(4);
  m[0] = std::cos(0.5 * theta);
  m[2] = -i * std::exp(-i * phi) * std::sin(0.5 * theta);
  m[1] = -i * std::exp(i * phi) * std::sin(0.5 * theta);
  m[3] = std::cos(0.5 * theta);
  return m;
}

inline cvec rx(double theta) {
  const cd i(0., 1.);
  cvec m(4);
  m[0] = std::cos(0.5 * theta);
  m[2] = -i * std::sin(0.5 * theta);
  m[1] = m[2];
  m[3] = m[0];
  return m;
}

inline cvec ry(double theta) {
  cvec m(4);
  m[0] = std::cos(0.5 * theta);
  m[2] = -1.0 * std::sin(0.5 * theta);
  m[1] = -m[2];
  m[3] = m[0];
  return m;
}

inline cvec rz(double theta) {
  const cd i(0., 1.);
  cvec m(4);
  m[0] = std::exp(-i * 0.5 * theta);
  m[3] = std::exp(i * 0.5 * theta);
  return m;
}

inline cvec rxx(double theta) {
  const cd i(0., 1.);
  const double c = std::cos(0.5 * theta), s = std::sin(0.5 * theta);
  cvec m(16);
  m[0 + 0 * 4] = c;
  m[0 + 3 * 4] = -i * s;
  m[1 + 1 * 4] = c;
  m[1 + 2 * 4] = -i * s;
  m[2 + 1 * 4] = -i * s;
  m[2 + 2 * 4] = c;
  m[3 + 0 * 4] = -i * s;
  m[3 + 3 * 4] = c;
  return m;
}

inline cvec ryy(double theta) {
  const cd i(0., 1.);
  const double c = std::cos(0.5 * theta), s = std::sin(0.5 * theta);
  cvec m(16);
  m[0 + 0 * 4] = c;
  m[0 + 3 * 4] = i * s;
  m[1 + 1 * 4] = c;
  m[1 + 2 * 4] = -i * s;
  m[2 + 1 * 4] = -i * s;
  m[2 + 2 * 4] = c;
  m[3 + 0 * 4] = i * s;
  m[3 + 3 * 4] = c;
  return m;
}

inline cvec rzz_diag(double theta) {
  const cd i(0., 1.);
  const cd ep = std::exp(i * 0.5 * theta), em = std::exp(-i * 0.5 * theta);
  return cvec({em, ep, ep, em});
}

inline cvec rzx(double theta) {
  const cd i(0., 1.);
  const double c = std::cos(0.5 * theta), s = std::sin(0.5 * theta);
  cvec m(16);
  m[0 + 0 * 4] = c;
  m[0 + 2 * 4] = -i * s;
  m[1 + 1 * 4] = c;
  m[1 + 3 * 4] = i * s;
  m[2 + 0 * 4] = -i * s;
  m[2 + 2 * 4] = c;
  m[3 + 1 * 4] = i * s;
  m[3 + 3 * 4] = c;
  return m;
}

// Fixed matrices (matrix_defs.hpp:150-212), column-major.
inline cvec sx() { return cvec({{0.5, 0.5}, {0.5, -0.5}, {0.5, -0.5}, {0.5, 0.5}}); }
inline cvec sxdg() { return cvec({{0.5, -0.5}, {0.5, 0.5}, {0.5, 0.5}, {0.5, -0.5}}); }
inline cvec ecr() {
  const double s = 1. / std::sqrt(2.);
  // rows: [0, s, 0, is], [s, 0, -is, 0], [0, is, 0, s], [-is, 0, s, 0]
  cvec m(16);
  const cd row[4][4] = {{0, s, 0, cd(0, s)},
                        {s, 0, cd(0, -s), 0},
                        {0, cd(0, s), 0, s},
                        {cd(0, -s), 0, s, 0}};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) m[i + 4 * j] = row[i][j];
  return m;
}

}  // namespace gates
}  // namespace aerhip
