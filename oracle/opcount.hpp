/* Operation counter for the oracle (test and measurement infrastructure,
 * never part of the product).
 *
 * oracle/Makefile compiles mhd_oracle.c a second time, as C++ with this
 * header forced in front of it and `double` defined as `fp64`: a double in
 * a struct whose arithmetic operators count what they do.  The counted
 * build computes exactly what the oracle computes (the same IEEE
 * operations in the same order), so its counts are the floating-point
 * operations of PLUTO's algorithm as the oracle restates it, per call of
 * its entry points.  Single-threaded (no OpenMP), so the counters are
 * plain globals.
 */
#ifndef PGPU_OPCOUNT_HPP
#define PGPU_OPCOUNT_HPP

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <cmath>
#include <type_traits>

struct OpCounts {
  unsigned long long add, mul, div, sqrt, neg, fabs, cmp;
};
extern "C" OpCounts g_opcounts;

struct fp64 {
  double v;
  fp64() = default;
  template <class T, typename std::enable_if<std::is_arithmetic<T>::value, int>::type = 0>
  constexpr fp64(T x) : v((double)x) {}
  template <class T, typename std::enable_if<std::is_arithmetic<T>::value, int>::type = 0>
  explicit operator T() const { return (T)v; }
};
static_assert(sizeof(fp64) == 8, "fp64 must keep the layout of a double");

inline fp64 operator+(fp64 a, fp64 b) { g_opcounts.add++; return fp64(a.v + b.v); }
inline fp64 operator-(fp64 a, fp64 b) { g_opcounts.add++; return fp64(a.v - b.v); }
inline fp64 operator*(fp64 a, fp64 b) { g_opcounts.mul++; return fp64(a.v * b.v); }
inline fp64 operator/(fp64 a, fp64 b) { g_opcounts.div++; return fp64(a.v / b.v); }
inline fp64 operator-(fp64 a) { g_opcounts.neg++; return fp64(-a.v); }
inline fp64 operator+(fp64 a) { return a; }
inline fp64& operator+=(fp64& a, fp64 b) { a = a + b; return a; }
inline fp64& operator-=(fp64& a, fp64 b) { a = a - b; return a; }
inline fp64& operator*=(fp64& a, fp64 b) { a = a * b; return a; }
inline fp64& operator/=(fp64& a, fp64 b) { a = a / b; return a; }
inline bool operator<(fp64 a, fp64 b) { g_opcounts.cmp++; return a.v < b.v; }
inline bool operator>(fp64 a, fp64 b) { g_opcounts.cmp++; return a.v > b.v; }
inline bool operator<=(fp64 a, fp64 b) { g_opcounts.cmp++; return a.v <= b.v; }
inline bool operator>=(fp64 a, fp64 b) { g_opcounts.cmp++; return a.v >= b.v; }
inline bool operator==(fp64 a, fp64 b) { g_opcounts.cmp++; return a.v == b.v; }
inline bool operator!=(fp64 a, fp64 b) { g_opcounts.cmp++; return a.v != b.v; }
inline fp64 sqrt(fp64 a) { g_opcounts.sqrt++; return fp64(std::sqrt(a.v)); }
inline fp64 fabs(fp64 a) { g_opcounts.fabs++; return fp64(std::fabs(a.v)); }

#define double fp64

#endif
