// Counting scalar for the oracle (test infrastructure).
//
// Instantiating the oracle templates with CReal counts every +, -, *, / on
// floating-point values (1 flop each) and, separately, sqrt and sin/cos
// calls.  This gives the algorithmic flop count of the restated algorithm on
// the exact inputs of a workload; it is what a roofline "achieved TFLOP/s"
// figure divides by kernel time.
#pragma once
#include <cmath>
#include <cstdint>

namespace oracle {

struct FlopTally {
    uint64_t addsub = 0, mul = 0, div = 0, sqrt = 0, trig = 0;
    uint64_t flops() const { return addsub + mul + div; }
};
inline FlopTally& tally() { static thread_local FlopTally t; return t; }

struct CReal {
    double v;
    CReal() : v(0) {}
    CReal(double x) : v(x) {}                 // implicit from double
    explicit operator double() const { return v; }
};
inline CReal operator+(CReal a, CReal b) { ++tally().addsub; return CReal(a.v + b.v); }
inline CReal operator-(CReal a, CReal b) { ++tally().addsub; return CReal(a.v - b.v); }
inline CReal operator*(CReal a, CReal b) { ++tally().mul;    return CReal(a.v * b.v); }
inline CReal operator/(CReal a, CReal b) { ++tally().div;    return CReal(a.v / b.v); }
inline CReal operator-(CReal a) { return CReal(-a.v); }      // sign flip: not counted
inline bool operator>(CReal a, CReal b)  { return a.v > b.v; }
inline bool operator<(CReal a, CReal b)  { return a.v < b.v; }
inline bool operator>=(CReal a, CReal b) { return a.v >= b.v; }
inline bool operator<=(CReal a, CReal b) { return a.v <= b.v; }
inline bool operator==(CReal a, CReal b) { return a.v == b.v; }
inline CReal sqrt(CReal a) { ++tally().sqrt; return CReal(std::sqrt(a.v)); }
inline CReal sin(CReal a)  { ++tally().trig; return CReal(std::sin(a.v)); }
inline CReal cos(CReal a)  { ++tally().trig; return CReal(std::cos(a.v)); }

} // namespace oracle
