// Interpolation operators: families, orders, the rules that tie them to N
// and G, and exact rational Lagrange weights rounded once to double.
//
// Restates src/operators.jl:314-442 (Operators, boundedmode), :531-595
// (check_operators), :607-682 (lagrange_weights, unit_lagrange_weights)
// and src/schedule.jl:908-915,1097-1120 (interpolation_weights and the
// conservative prolongation weights).
#pragma once

#include <array>
#include <cstdint>
#include <vector>

#include "forest.hpp"

namespace tah {

// An exact rational with a 128-bit numerator and denominator, always
// reduced, denominator positive. Every weight the schedule asks for is a
// Lagrange weight of at most a dozen integer nodes at a quarter-integer
// target, far inside this range; overflow is checked and refused.
struct Rational {
  __int128 num = 0, den = 1;
  Rational() = default;
  Rational(long long n) : num(n), den(1) {}
  Rational(__int128 n, __int128 d);
  Rational operator+(const Rational& o) const;
  Rational operator-(const Rational& o) const;
  Rational operator*(const Rational& o) const;
  Rational operator/(const Rational& o) const;
  bool operator==(const Rational& o) const { return num == o.num && den == o.den; }
  bool operator<=(const Rational& o) const;
  // Correctly rounded, as T(::Rational{BigInt}) is (src/operators.jl:597-606).
  double to_double() const;
};

enum Family : int { PointValue = 0, Conservative = 1 };

struct Operators {
  int family = PointValue;
  int prolongation = 0;
  int restriction = 0;
  bool bounded = false;
  int vertex_prolongation = 0;
  Operators() : family(PointValue), prolongation(2), restriction(2) {}
  // src/operators.jl:327-411: the checks of the keyword constructor.
  Operators(int family, int prolongation, int restriction, bool bounded,
            int vertex_prolongation);
};

// src/operators.jl:467-469
int ghost_layers_read(const Operators& ops);
// src/operators.jl:448-449
int vertex_order(const Operators& ops);
// src/operators.jl:541-595, with c[d] = 1 for a vertex-like dimension.
void check_operators(int N, const int* G, const int* c, int D, const Operators& ops);

// Lagrange weights through nodes, evaluated at x (src/operators.jl:627-643).
std::vector<Rational> lagrange_weights(const std::vector<Rational>& nodes, const Rational& x);
// Weights for nodes 0..p-1 at xi (src/operators.jl:675-682), memoised.
const std::vector<Rational>& unit_lagrange_weights(int p, const Rational& xi);
// Weights for the window lo..lo+p-1 at x; refuses extrapolation
// (src/schedule.jl:908-915).
const std::vector<Rational>& interpolation_weights(long lo, int p, const Rational& x,
                                                   const char* what);
// The two subcell weight vectors of conservative prolongation of odd order
// p (src/schedule.jl:1097-1110).
void conservative_prolong_weights(int p, std::vector<Rational>& low, std::vector<Rational>& high);

}  // namespace tah
