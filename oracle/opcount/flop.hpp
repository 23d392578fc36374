// Op-counting FP64 scalar for the oracle (test infrastructure only).
//
// Flop wraps one IEEE double and counts every +, -, *, / and sqrt into the
// current phase (ORC_PHASE markers in the oracle sources); negation, fabs,
// comparisons and min/max are not counted.  The arithmetic is the plain
// double arithmetic, so the counting build computes the same values as the
// ordinary one.  Built by oracle/opcount/Makefile from the unchanged oracle
// sources with `double` mapped onto Flop after the system headers.
#pragma once
#include <math.h>
#include <stdlib.h>
#include <string.h>

using real_t = decltype(1.0);

extern unsigned long long g_ops[16];
extern int g_phase;
extern long long g_weno_cells;

#define ORC_PHASE(k) (g_phase = (k))
#define ORC_WENO_CELL() (++g_weno_cells)

struct Flop {
    real_t v;
    Flop() = default;
    constexpr Flop(real_t x) : v(x) {}
    constexpr Flop(float x) : v(x) {}
    constexpr Flop(int x) : v(x) {}
    explicit operator real_t() const { return v; }
    explicit operator int() const { return (int)v; }
    Flop operator-() const { return Flop(-v); }
    Flop &operator+=(Flop o) { ++g_ops[g_phase]; v += o.v; return *this; }
    Flop &operator-=(Flop o) { ++g_ops[g_phase]; v -= o.v; return *this; }
    Flop &operator*=(Flop o) { ++g_ops[g_phase]; v *= o.v; return *this; }
    Flop &operator/=(Flop o) { ++g_ops[g_phase]; v /= o.v; return *this; }
};
static_assert(sizeof(Flop) == sizeof(real_t), "Flop must keep the layout of a double");

inline Flop operator+(Flop a, Flop b) { ++g_ops[g_phase]; return Flop(a.v + b.v); }
inline Flop operator-(Flop a, Flop b) { ++g_ops[g_phase]; return Flop(a.v - b.v); }
inline Flop operator*(Flop a, Flop b) { ++g_ops[g_phase]; return Flop(a.v * b.v); }
inline Flop operator/(Flop a, Flop b) { ++g_ops[g_phase]; return Flop(a.v / b.v); }
inline bool operator<(Flop a, Flop b) { return a.v < b.v; }
inline bool operator>(Flop a, Flop b) { return a.v > b.v; }
inline bool operator<=(Flop a, Flop b) { return a.v <= b.v; }
inline bool operator>=(Flop a, Flop b) { return a.v >= b.v; }
inline bool operator==(Flop a, Flop b) { return a.v == b.v; }
inline bool operator!=(Flop a, Flop b) { return a.v != b.v; }
inline Flop sqrt(Flop a) { ++g_ops[g_phase]; return Flop(::sqrt(a.v)); }
inline Flop fabs(Flop a) { return Flop(::fabs(a.v)); }
inline Flop log10(Flop a) { ++g_ops[g_phase]; return Flop(::log10(a.v)); }
