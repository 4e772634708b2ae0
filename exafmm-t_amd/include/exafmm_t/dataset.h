// Test-data generators with the reference's semantics (include/dataset.h:14-146):
// drand48 seeded with srand48(seed); cube in [0,1)^3, sphere = normalised
// cube samples in [-1,1)^3, Plummer with R < 100 truncation; source charges
// drand48() - 0.5 (complex: both parts).
#pragma once
#include <cstdlib>
#include "exafmm_t.h"

namespace exafmm_t {

template <typename T>
Bodies<T> init_targets(int n, const char* distribution, int seed) {
  Bodies<T> b(n);
  srand48(seed);
  const char d = distribution ? distribution[0] : 'c';
  if (d == 'c') {
    for (auto& x : b)
      for (int k = 0; k < 3; k++) x.X[k] = drand48();
  } else if (d == 's') {
    for (auto& x : b) {
      for (int k = 0; k < 3; k++) x.X[k] = drand48() * 2 - 1;
      x.X /= std::sqrt(norm(x.X));
    }
  } else if (d == 'p') {
    int i = 0, xmax = 0;
    while (i < n) {
      const real_t x1 = drand48(), x2 = drand48(), x3 = drand48();
      const real_t R = 1.0 / std::sqrt(std::pow(x1, -2.0 / 3.0) - 1.0);
      if (R < 100) {
        const real_t Z = (1.0 - 2.0 * x2) * R;
        const real_t Rxy = std::sqrt(R * R - Z * Z);
        b[i].X[0] = Rxy * std::cos(2.0 * PI * x3);
        b[i].X[1] = Rxy * std::sin(2.0 * PI * x3);
        b[i].X[2] = Z;
        for (int k = 0; k < 3; k++) xmax = xmax > std::fabs(b[i].X[k]) ? xmax : int(std::fabs(b[i].X[k]));
        i++;
      }
    }
    const real_t scale = 0.5 / (xmax + 1);
    for (auto& x : b)
      for (int k = 0; k < 3; k++) x.X[k] = x.X[k] * scale + 0.5;
  } else {
    throw std::runtime_error(std::string("unknown distribution ") + distribution);
  }
  for (int i = 0; i < n; i++) b[i].ibody = i;
  return b;
}

template <typename T>
Bodies<T> init_sources(int n, const char* distribution, int seed) {
  Bodies<T> b = init_targets<T>(n, distribution, seed);
  for (auto& x : b) {
    if constexpr (std::is_same<T, complex_t>::value) {
      const real_t re = drand48() - 0.5;
      x.q = complex_t(re, drand48() - 0.5);
    } else {
      x.q = drand48() - 0.5;
    }
  }
  return b;
}

}  // namespace exafmm_t
