// Statistics of per-frame observables: ccp's sostat (Program/src/sostat.cu,
// sostat.cuh), restated with its float/double rounding points so that the
// analysis files carry the same numbers.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace ccpmi {

struct SeriesStat {
  double avg = 0.0, var = 0.0, sem = 0.0;
  uint32_t blocks = 0;  // number of blocks where the blocking method stopped
  uint32_t cut = 0;     // first value kept (marginal standard error rule)
  bool settled = false; // the cut differs from the half-length one
};

// Mean, unbiased variance and independent-sample error (calc_stats on a
// float vector, sostat.cu:72-93). The squares are float products, summed in
// double, as in ccp.
inline SeriesStat independent_stats(const float *v, size_t n) {
  double m1 = 0.0, m2 = 0.0;
  for (size_t i = 0; i < n; ++i) {
    m1 += v[i];
    m2 += v[i] * v[i];
  }
  m1 /= n;
  m2 /= n;
  SeriesStat s;
  s.avg = m1;
  s.var = (m2 - m1 * m1) / (1.0 - 1.0 / n);
  s.sem = std::sqrt(s.var / n);
  return s;
}

// Error of the mean of a correlated series by the blocking method
// (Flyvbjerg-Petersen), stopping once the variance of the mean no longer
// rises beyond its own uncertainty (sostat.cu:96-138). ccp leaves sem and
// the block count unset for fewer than 4 values; here they keep the independent
// estimate and n.
inline SeriesStat blocking_stats(const float *v, size_t n) {
  std::vector<float> av(v, v + n);
  SeriesStat ind = independent_stats(av.data(), av.size());
  SeriesStat s;
  s.avg = ind.avg;
  s.var = ind.var;
  s.sem = ind.sem;
  s.blocks = (uint32_t)n;
  size_t len = av.size();
  double var_mean = ind.var / len;
  double bound = var_mean * (1.0 + std::sqrt(2.0 / (len - 1.0)));
  while (len >= 4) {
    size_t half = 0;
    for (; 2 * half + 1 < len; ++half) av[half] = 0.5 * (av[2 * half] + av[2 * half + 1]);
    len = half;
    av.resize(len);
    ind = independent_stats(av.data(), len);
    var_mean = ind.var / len;
    s.sem = std::sqrt(var_mean);
    s.blocks = (uint32_t)len;
    if (var_mean > bound)
      bound = var_mean * (1.0 + std::sqrt(2.0 / (len - 1.0)));
    else
      break;
  }
  return s;
}

// Time-series statistics (sostat.cu:141-183): the thermalisation cut is
// the one of n/2, n/4, ..., n/64 with the smallest marginal standard error
// var (m - 1) / m^2 of the remaining m values; the series counts as
// thermalised unless the cut stays at n/2. Statistics of the values after
// the cut by the blocking method. With no finite marginal error (fewer than
// two values after every cut) the cut stays at n/2, where ccp leaves it unset.
inline SeriesStat series_stats(const float *v, size_t n) {
  double best = INFINITY;
  uint32_t keep_from = (uint32_t)(n / 2);
  for (uint32_t d = 2; d < 128; d *= 2) {
    const uint32_t from = (uint32_t)(n / d);
    const size_t m = n - from;
    const SeriesStat ind = independent_stats(v + from, m);
    // m * m in 32-bit unsigned arithmetic, as ccp (wraps above 65535 values)
    const double mse = ind.var * (double)(uint32_t)(m - 1) / (double)((uint32_t)m * (uint32_t)m);
    if (mse < best) {
      best = mse;
      keep_from = from;
    }
  }
  SeriesStat s = blocking_stats(v + keep_from, n - keep_from);
  s.cut = keep_from;
  s.settled = keep_from != n / 2;
  return s;
}

// Inverse-variance weighted combination of per-simulation results
// (sostat.cu:186-211); a weight that is not finite counts as 1.
inline SeriesStat weighted_stats(const std::vector<SeriesStat> &v) {
  double m1 = 0.0, m2 = 0.0, w1 = 0.0, w2 = 0.0;
  for (const SeriesStat &e : v) {
    double w = 1.0 / (e.sem * e.sem);
    if (!std::isfinite(w)) w = 1.0;
    m1 += w * e.avg;
    m2 += w * e.avg * e.avg;
    w1 += w;
    w2 += w * w;
  }
  m1 /= w1;
  m2 /= w1;
  SeriesStat s;
  s.avg = m1;
  s.var = (m2 - m1 * m1) / (1.0 - w2 / (w1 * w1));
  s.sem = std::sqrt(s.var * w2 / (w1 * w1));
  return s;
}

// Plain combination of per-simulation averages from their sums m1 = sum avg
// and m2 = sum avg^2 over n simulations (sostat.cu:214-235).
inline SeriesStat plain_stats(double m1, double m2, size_t n) {
  m1 /= n;
  m2 /= n;
  SeriesStat s;
  s.avg = m1;
  s.var = (m2 - m1 * m1) / (1.0 - 1.0 / n);
  s.sem = std::sqrt(s.var / n);
  return s;
}

} // namespace ccpmi
