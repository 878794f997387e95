// flopcount.h -- a float64 type that counts its arithmetic, for the flop-counting
// build of the oracle (oracle_flops.cpp).  TEST/MEASUREMENT INFRASTRUCTURE.
//
// Counting convention: +, -, *, / and sqrt = 1 op each; each transcendental
// (exp, sin, cos, tanh, atan, atan2, asin, acos) = one "trans" event (weighted 20 by the
// reader); negation, comparisons, floor, fabs and conversions are free.  A
// region marked ORC_PAUSE/ORC_RESUME (loop-invariant work) is not counted.
#pragma once
#include <cmath>
#include <cstdint>

namespace orcflops {

struct Counters {
  uint64_t arith = 0, trans = 0, draws = 0;
  int paused = 0;
};
extern thread_local Counters g;

struct CReal {
  double v;
  CReal() = default;
  constexpr CReal(double x) : v(x) {}
  constexpr CReal(float x) : v(x) {}
  constexpr CReal(int x) : v(x) {}
  constexpr CReal(long x) : v((double)x) {}
  constexpr CReal(long long x) : v((double)x) {}
  constexpr CReal(unsigned long x) : v((double)x) {}
  constexpr CReal(unsigned long long x) : v((double)x) {}
  explicit operator float() const { return (float)v; }
  explicit operator long() const { return (long)v; }
  explicit operator long long() const { return (long long)v; }
  explicit operator int() const { return (int)v; }
  explicit operator bool() const { return v != 0; }
  CReal operator-() const { return CReal(-v); }
  CReal operator+() const { return *this; }
  CReal& operator+=(CReal o) { if (!g.paused) ++g.arith; v += o.v; return *this; }
  CReal& operator-=(CReal o) { if (!g.paused) ++g.arith; v -= o.v; return *this; }
  CReal& operator*=(CReal o) { if (!g.paused) ++g.arith; v *= o.v; return *this; }
  CReal& operator/=(CReal o) { if (!g.paused) ++g.arith; v /= o.v; return *this; }
};

inline CReal operator+(CReal a, CReal b) { if (!g.paused) ++g.arith; return CReal(a.v + b.v); }
inline CReal operator-(CReal a, CReal b) { if (!g.paused) ++g.arith; return CReal(a.v - b.v); }
inline CReal operator*(CReal a, CReal b) { if (!g.paused) ++g.arith; return CReal(a.v * b.v); }
inline CReal operator/(CReal a, CReal b) { if (!g.paused) ++g.arith; return CReal(a.v / b.v); }
inline bool operator<(CReal a, CReal b) { return a.v < b.v; }
inline bool operator>(CReal a, CReal b) { return a.v > b.v; }
inline bool operator<=(CReal a, CReal b) { return a.v <= b.v; }
inline bool operator>=(CReal a, CReal b) { return a.v >= b.v; }
inline bool operator==(CReal a, CReal b) { return a.v == b.v; }
inline bool operator!=(CReal a, CReal b) { return a.v != b.v; }

inline CReal c_sqrt(CReal a) { if (!g.paused) ++g.arith; return CReal(std::sqrt(a.v)); }
inline CReal c_trans1(double r) { if (!g.paused) ++g.trans; return CReal(r); }

}  // namespace orcflops

namespace std {
inline orcflops::CReal sqrt(orcflops::CReal a) { return orcflops::c_sqrt(a); }
inline orcflops::CReal exp(orcflops::CReal a) { return orcflops::c_trans1(std::exp(a.v)); }
inline orcflops::CReal sin(orcflops::CReal a) { return orcflops::c_trans1(std::sin(a.v)); }
inline orcflops::CReal cos(orcflops::CReal a) { return orcflops::c_trans1(std::cos(a.v)); }
inline orcflops::CReal tanh(orcflops::CReal a) { return orcflops::c_trans1(std::tanh(a.v)); }
inline orcflops::CReal atan(orcflops::CReal a) { return orcflops::c_trans1(std::atan(a.v)); }
inline orcflops::CReal atan2(orcflops::CReal a, orcflops::CReal b) { return orcflops::c_trans1(std::atan2(a.v, b.v)); }
inline orcflops::CReal acos(orcflops::CReal a) { return orcflops::c_trans1(std::acos(a.v)); }
inline orcflops::CReal asin(orcflops::CReal a) { return orcflops::c_trans1(std::asin(a.v)); }
inline orcflops::CReal floor(orcflops::CReal a) { return orcflops::CReal(std::floor(a.v)); }
inline orcflops::CReal fabs(orcflops::CReal a) { return orcflops::CReal(std::fabs(a.v)); }
inline bool isnan(orcflops::CReal a) { return std::isnan(a.v); }
inline bool isfinite(orcflops::CReal a) { return std::isfinite(a.v); }
}  // namespace std

#define ORC_PAUSE (++orcflops::g.paused)
#define ORC_RESUME (--orcflops::g.paused)
#define ORC_DRAW (++orcflops::g.draws)
