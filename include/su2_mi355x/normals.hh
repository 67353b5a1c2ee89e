// normals.hh -- the first n draws of std::normal_distribution<double>(0, 1)
// over a std::mt19937, bit for bit as libstdc++ produces them, in two passes.
//
// initnormal (adjointfield.hh:240-258) draws the momenta from a fresh
// normal_distribution over the trajectory's engine.  libstdc++'s polar method
// (bits/random.tcc, normal_distribution::operator()) takes pairs
// x, y = 2 u - 1 of generate_canonical<double, 53> uniforms until
// r2 = x^2 + y^2 is in (0, 1], returns y m and keeps x m for the next call,
// m = sqrt(-2 log(r2) / r2), each result scaled by stddev and shifted by mean.
// Pass 1 walks the engine in order and keeps the accepted (x, y, r2); pass 2,
// the logarithms and square roots, is independent per pair and runs on up to
// 16 threads.  The engine ends in the state the one-by-one draw leaves it in.
//
// Used by the C ABI (su2gpu_md_update with the host stream, su2gpu_initnormal)
// and by the C++ adapter's initnormal.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <random>
#include <thread>
#include <vector>

namespace su2_mi355x {
namespace detail {

// generate_canonical<double, 53>(mt19937) of libstdc++: two 32-bit words,
// sum = w0 + w1 2^32 accumulated in double, sum / 2^64, clamped below 1
// (spelled out: the library template recomputes its word count with
// long-double logarithms on every call)
inline double canonical53(std::mt19937 &e) {
  double sum = 0.0, tmp = 1.0;
  sum += double(e() - std::mt19937::min()) * tmp;
  tmp *= 4294967296.0;
  sum += double(e() - std::mt19937::min()) * tmp;
  tmp *= 4294967296.0;
  const double ret = sum / tmp;
  return ret >= 1.0 ? std::nextafter(1.0, 0.0) : ret;
}

// out[0 .. n): the normals; pairs: scratch, resized to 3 ceil(n/2)
inline void polar_normals(std::mt19937 &engine, double *out, size_t n, std::vector<double> &pairs) {
  const size_t npairs = (n + 1) / 2;
  pairs.resize(3 * npairs);
  for (size_t k = 0; k < npairs; ++k) {
    double x, y, r2;
    do {
      x = 2.0 * canonical53(engine) - 1.0;
      y = 2.0 * canonical53(engine) - 1.0;
      r2 = x * x + y * y;
    } while (r2 > 1.0 || r2 == 0.0);
    pairs[3 * k] = x;
    pairs[3 * k + 1] = y;
    pairs[3 * k + 2] = r2;
  }
  auto transform = [&](size_t k0, size_t k1) {
    for (size_t k = k0; k < k1; ++k) {
      const double x = pairs[3 * k], y = pairs[3 * k + 1], r2 = pairs[3 * k + 2];
      const double mult = std::sqrt(-2 * std::log(r2) / r2);
      out[2 * k] = (y * mult) * 1.0 + 0.0;  // ret * stddev + mean
      if (2 * k + 1 < n) out[2 * k + 1] = (x * mult) * 1.0 + 0.0;
    }
  };
  const unsigned hw = std::thread::hardware_concurrency();
  const size_t nt = npairs < (size_t(1) << 18) ? 1 : std::min<size_t>(16, hw ? hw : 1);
  std::vector<std::thread> pool;
  for (size_t t = 1; t < nt; ++t) pool.emplace_back(transform, npairs * t / nt, npairs * (t + 1) / nt);
  transform(0, npairs / nt);
  for (auto &th : pool) th.join();
}

}  // namespace detail
}  // namespace su2_mi355x
