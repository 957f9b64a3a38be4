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

double path_increment(const ScalarPath& path, double z0, double z1) {
  return path(z1) - path(z0);
}

}  // namespace sp
