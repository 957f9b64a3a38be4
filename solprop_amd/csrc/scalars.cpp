// Host scalar paths (see scalars.hpp). Compiled with -ffp-contract=off.
#include "scalars.hpp"

#include <stdexcept>

namespace sp {

void ScalarPath::assign(const sp_scalar_path& p) {
  kind = p.kind;
  value = p.value;
  degree = p.degree;
  edges.clear();
  coefficients.clear();
  if (kind == 2) {
    if (p.n_intervals < 1 || !p.edges || !p.coefficients || p.degree < 0)
      throw std::invalid_argument("a Chebyshev scalar path needs edges and coefficients");
    edges.assign(p.edges, p.edges + p.n_intervals + 1);
    coefficients.assign(p.coefficients,
                        p.coefficients + static_cast<std::size_t>(p.n_intervals) * (degree + 1));
  } else if (kind != 0 && kind != 1) {
    throw std::invalid_argument("a scalar path's kind is 0, 1 or 2");
  }
}

double ScalarPath::operator()(double z) const {
  if (kind == 0) return value;
  if (kind == 1) return z;
  const int n = static_cast<int>(edges.size()) - 1;
  int j = 0;
  while (j < n - 1 && z >= edges[j + 1]) ++j;
  const double a = edges[j], b = edges[j + 1];
  const double x = (2.0 * z - a - b) / (b - a);
  const double* c = coefficients.data() + static_cast<std::size_t>(j) * (degree + 1);
  double b1 = 0.0, b2 = 0.0;
  for (int k = degree; k > 0; --k) {
    const double t = 2.0 * x * b1 - b2 + c[k];
    b2 = b1;
    b1 = t;
  }
  return x * b1 - b2 + c[0];
}

double stage_position(double z, double c, double dz) { return z + c * dz; }

void lagrange_basis(const double* points, int n, double upper, bool integrate, double* rows,
                    double* weights) {
  std::vector<double> coefficients(n), product(n);
  if (!integrate)
    for (int i = 0; i < n * n; ++i) rows[i] = 0.0;
  for (int index = 0; index < n; ++index) {
    const double node = points[index];
    int length = 1;
    coefficients[0] = 1.0;
    double denominator = 1.0;
    for (int other = 0; other < n; ++other) {
      if (other == index) continue;
      const double root = points[other];
      product[0] = 0.0;
      for (int power = 0; power < length; ++power) product[power + 1] = coefficients[power];
      for (int power = 0; power < length; ++power) product[power] -= root * coefficients[power];
      length += 1;
      for (int power = 0; power < length; ++power) coefficients[power] = product[power];
      denominator *= node - root;
    }
    if (integrate) {
      double integral = 0.0;
      for (int power = length - 1; power >= 0; --power)
        integral = (integral + coefficients[power] / (power + 1)) * upper;
      weights[index] = integral / denominator;
    } else {
      for (int power = 0; power < length; ++power)
        rows[index * n + power] = coefficients[power] / denominator;
    }
  }
}

double path_increment(const ScalarPath& path, double z0, double z1) {
  return path(z1) - path(z0);
}

}  // namespace sp
