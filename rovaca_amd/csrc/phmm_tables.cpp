#include "phmm_tables.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace rvc {
namespace {

// Jacobian log table: entry k = log10(1 + 10^(-k * 1e-4)), k = 0 .. 80000,
// stored in the working precision (ph/roDV_common/context.h:44-48).
constexpr int kJacEntries = 80001;
constexpr double kJacStep = 0.0001;
constexpr double kJacMaxDiff = 8.0;

template <typename T>
struct Log10Sum {
  std::vector<T> jac;
  Log10Sum() : jac(kJacEntries) {
    for (int k = 0; k < kJacEntries; ++k) jac[k] = (T)std::log10(1.0 + std::pow(10.0, -(double)k * kJacStep));
  }
  // log10(10^a + 10^b) by table lookup in precision T, rounding the table
  // index half away from zero (context.h:62-82).
  T operator()(T a, T b) const {
    T lo = a, hi = b;
    if (lo > hi) std::swap(lo, hi);
    if (std::isinf(lo) || std::isinf(hi)) return hi;
    const T diff = hi - lo;
    if (diff >= (T)kJacMaxDiff) return hi;
    const T scaled = (T)(diff * (T)(1.0 / kJacStep));
    const int idx = scaled > (T)0 ? (int)(scaled + (T)0.5) : (int)(scaled - (T)0.5);
    return hi + jac[idx];
  }
};

// Match-to-match probability for gap-open qualities (i, d), i >= d:
// 1 - 10^(-i/10) - 10^(-d/10) evaluated as 10^(log1p(-min(1, 10^S))/ln 10)
// with S the table log-sum in precision T (context.h:50-60).
template <typename T>
void fill_mm(T* mm) {
  Log10Sum<T> sum;
  const double inv_ln10 = 1.0 / std::log(10.0);
  for (int hi = 0, base = 0; hi < 128; base += ++hi) {
    for (int lo = 0; lo <= hi; ++lo) {
      const double s = (double)sum((T)(-0.1 * hi), (T)(-0.1 * lo));
      const double l = std::log1p(-std::min(1.0, std::pow(10.0, s))) * inv_ln10;
      mm[base + lo] = (T)std::pow(10.0, l);
    }
  }
}

}  // namespace

void build_tables(HostTables* t) {
  for (int q = 0; q < 128; ++q) {
    t->ph2pr_f[q] = powf(10.f, -((float)q) / 10.f);
    t->ph2pr_d[q] = std::pow(10.0, -((double)q) / 10.0);
  }
  fill_mm<float>(t->mm_f);
  fill_mm<double>(t->mm_d);
}

}  // namespace rvc
