// Forward-mode (tangent) FP64 scalar for the oracle (test infrastructure
// only).
//
// Dual carries a value and its derivative along one direction; every
// arithmetic operation propagates the derivative by the chain rule, so the
// oracle compiled with `double` mapped onto Dual computes the directional
// derivative of its step along the seeded direction.  Values are computed
// with the plain double arithmetic, so they equal the ordinary build's.
// ORC_FROZEN drops the derivative of the values the discrete adjoint holds
// at their forward values; ORC_FROZEN_INCREMENT gives the curved path's
// WENO output the derivative of its input (its increment held fixed).
#pragma once
#include <math.h>
#include <stdlib.h>
#include <string.h>

using real_t = decltype(1.0);

struct Dual {
    real_t v, d;
    Dual() = default;
    constexpr Dual(real_t x) : v(x), d(0.0) {}
    constexpr Dual(float x) : v(x), d(0.0) {}
    constexpr Dual(int x) : v(x), d(0.0) {}
    constexpr Dual(real_t x, real_t dx) : v(x), d(dx) {}
    explicit operator real_t() const { return v; }
    explicit operator int() const { return (int)v; }
    Dual operator-() const { return Dual(-v, -d); }
    Dual &operator+=(Dual o) { v += o.v; d += o.d; return *this; }
    Dual &operator-=(Dual o) { v -= o.v; d -= o.d; return *this; }
    Dual &operator*=(Dual o) { d = d * o.v + v * o.d; v *= o.v; return *this; }
    Dual &operator/=(Dual o) { d = (d * o.v - v * o.d) / (o.v * o.v); v /= o.v; return *this; }
};

inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(Dual a, Dual b) { return Dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)); }
inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
inline bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
inline bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
inline bool operator==(Dual a, Dual b) { return a.v == b.v; }
inline bool operator!=(Dual a, Dual b) { return a.v != b.v; }
inline Dual sqrt(Dual a)
{
    const real_t s = ::sqrt(a.v);
    return Dual(s, s > 0.0 ? a.d / (2.0 * s) : 0.0);
}
inline Dual fabs(Dual a) { return a.v < 0.0 ? Dual(-a.v, -a.d) : (a.v > 0.0 ? a : Dual(a.v, 0.0)); }
inline Dual log10(Dual a) { return Dual(::log10(a.v), a.d / (a.v * ::log(10.0))); }

#define ORC_FROZEN(x) (Dual((x).v, 0.0))
#define ORC_FROZEN_INCREMENT(out, in) ((out).d = (in).d)
