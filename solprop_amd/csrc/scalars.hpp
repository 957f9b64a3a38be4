// Host scalar paths of z and the stage positions, compiled without contraction so that
// they evaluate to the same floats as the host layer's (and the reference's) Python.
#pragma once

#include <vector>

extern "C" {
#include "../../include/solprop_mi355x.h"
}

namespace sp {

struct ScalarPath {
  int kind = 0;
  double value = 0.0;
  int degree = 0;
  std::vector<double> edges, coefficients;
  // Throws std::invalid_argument on a malformed path.
  void assign(const sp_scalar_path& p);
  double operator()(double z) const;
  bool constant() const { return kind == 0; }
};

// z + c dz, and W(z1) - W(z0), rounded once each.
double stage_position(double z, double c, double dz);
double path_increment(const ScalarPath& path, double z0, double z1);

// The Lagrange basis on n distinct points, as the reference forms it
// (src/solprop/stepper/multistep.py:515-564): each numerator built from its own roots,
// ascending; its monomial coefficients (`rows`, n x n, when !integrate) or its integrals
// from 0 to `upper` by Horner's rule (`weights`, n, when integrate).
void lagrange_basis(const double* points, int n, double upper, bool integrate, double* rows,
                    double* weights);

}  // namespace sp
