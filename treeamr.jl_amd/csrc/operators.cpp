// Operators and exact weights; see operators.hpp.
#include "operators.hpp"

#include <map>
#include <mutex>
#include <string>

namespace tah {

namespace {
__int128 gcd128(__int128 a, __int128 b) {
  if (a < 0) a = -a;
  if (b < 0) b = -b;
  while (b != 0) {
    __int128 t = a % b;
    a = b;
    b = t;
  }
  return a;
}
constexpr __int128 kLimit = (__int128)1 << 100;
void check_range(__int128 v) {
  if (v > kLimit || v < -kLimit) fail("exact weight arithmetic overflowed", kUnsupported);
}
}  // namespace

Rational::Rational(__int128 n, __int128 d) {
  if (d == 0) fail("rational with zero denominator");
  if (d < 0) { n = -n; d = -d; }
  __int128 g = gcd128(n, d);
  if (g == 0) g = 1;
  num = n / g;
  den = d / g;
  check_range(num);
  check_range(den);
}
Rational Rational::operator+(const Rational& o) const {
  __int128 g = gcd128(den, o.den);
  return Rational(num * (o.den / g) + o.num * (den / g), den / g * o.den);
}
Rational Rational::operator-(const Rational& o) const {
  __int128 g = gcd128(den, o.den);
  return Rational(num * (o.den / g) - o.num * (den / g), den / g * o.den);
}
Rational Rational::operator*(const Rational& o) const {
  __int128 g1 = gcd128(num, o.den), g2 = gcd128(o.num, den);
  if (g1 == 0) g1 = 1;
  if (g2 == 0) g2 = 1;
  return Rational((num / g1) * (o.num / g2), (den / g2) * (o.den / g1));
}
Rational Rational::operator/(const Rational& o) const {
  if (o.num == 0) fail("rational division by zero");
  return *this * Rational(o.den, o.num);
}
bool Rational::operator<=(const Rational& o) const {
  return num * o.den <= o.num * den;
}
double Rational::to_double() const {
  // Both parts below 2^53 are exact in double, and IEEE division then
  // rounds the exact quotient once.
  const __int128 lim = (__int128)1 << 53;
  if (num >= lim || num <= -lim || den >= lim)
    fail("weight too large to convert exactly", kUnsupported);
  return double((long long)num) / double((long long)den);
}

Operators::Operators(int fam, int p, int r, bool b, int vp)
    : family(fam), prolongation(p), restriction(r), bounded(b), vertex_prolongation(vp) {
  if (fam != PointValue && fam != Conservative) fail("unknown operator family", kUnsupported);
  if (p <= 0 && fam == PointValue) fail("prolongation order must be at least 2, got " + std::to_string(p));
  if (fam == PointValue) {
    if (p < 2) fail("prolongation order must be at least 2, got " + std::to_string(p));
    if (r < 2) fail("restriction order must be at least 2, got " + std::to_string(r));
    if (p % 2) fail("prolongation order must be even for the point-value family, got " + std::to_string(p));
    if (r % 2) fail("restriction order must be even for the point-value family, got " + std::to_string(r));
  } else {
    if (p < 1) fail("conservative prolongation order must be at least 1, got " + std::to_string(p));
    if (p % 2 == 0) fail("conservative prolongation order must be odd, got " + std::to_string(p));
    if (r != 2)
      fail("conservative restriction is the exact volume average, so restriction must be 2, got " +
           std::to_string(r));
  }
  if (b) fail("bounded operators are not implemented by this library", kUnsupported);
  if (vp != 0) {
    if (fam != Conservative) fail("vertex_prolongation is for the conservative family only");
    if (vp < 2 || vp % 2) fail("vertex_prolongation must be even and at least 2");
  }
}

int ghost_layers_read(const Operators& ops) {
  return ops.family == Conservative ? (ops.prolongation - 1) / 2 : ops.prolongation / 2;
}

int vertex_order(const Operators& ops) {
  return ops.family == Conservative ? ops.vertex_prolongation : ops.prolongation;
}

void check_operators(int N, const int* G, const int* c, int D, const Operators& ops) {
  int pr = ops.restriction;
  int needed = ghost_layers_read(ops);
  for (int d = 0; d < D; ++d) {
    int g = G[d];
    std::string dim = " (dimension " + std::to_string(d + 1) + ")";
    if (c[d] == 1) {
      int pv = vertex_order(ops);
      if (pv == 0)
        fail("a Conservative field set with a vertex-like dimension needs vertex_prolongation" + dim);
      if (g < pv / 2 - 1)
        fail("prolongation of order " + std::to_string(pv) + " into a vertex-like dimension needs G >= " +
             std::to_string(pv / 2 - 1) + dim);
    } else {
      if (g < needed)
        fail("prolongation of order " + std::to_string(ops.prolongation) + " needs G >= " +
             std::to_string(needed) + " ghost layers" + dim);
      if (N < 2 * g + pr / 2 - 1)
        fail("restriction of order " + std::to_string(pr) + " into G=" + std::to_string(g) +
             " ghost layers needs N >= " + std::to_string(2 * g + pr / 2 - 1) + dim);
      if (N < pr)
        fail("restriction of order " + std::to_string(pr) + " needs N >= " + std::to_string(pr) + dim);
    }
  }
}

std::vector<Rational> lagrange_weights(const std::vector<Rational>& nodes, const Rational& x) {
  size_t n = nodes.size();
  std::vector<Rational> w(n);
  for (size_t i = 0; i < n; ++i) {
    Rational num(1), den(1);
    for (size_t j = 0; j < n; ++j) {
      if (j == i) continue;
      num = num * (x - nodes[j]);
      den = den * (nodes[i] - nodes[j]);
    }
    w[i] = num / den;
  }
  return w;
}

namespace {
std::mutex cache_lock;
struct RKey {
  int p;
  long long n, d;
  bool operator<(const RKey& o) const {
    if (p != o.p) return p < o.p;
    if (n != o.n) return n < o.n;
    return d < o.d;
  }
};
std::map<RKey, std::vector<Rational>> lagrange_cache;
}  // namespace

const std::vector<Rational>& unit_lagrange_weights(int p, const Rational& xi) {
  std::lock_guard<std::mutex> g(cache_lock);
  RKey key{p, (long long)xi.num, (long long)xi.den};
  auto it = lagrange_cache.find(key);
  if (it != lagrange_cache.end()) return it->second;
  std::vector<Rational> nodes;
  for (int j = 0; j < p; ++j) nodes.emplace_back(j);
  auto res = lagrange_cache.emplace(key, lagrange_weights(nodes, xi));
  return res.first->second;
}

const std::vector<Rational>& interpolation_weights(long lo, int p, const Rational& x,
                                                   const char* what) {
  Rational l(lo), h(lo + p - 1);
  if (!(l <= x && x <= h))
    fail(std::string(what) + " would extrapolate: the target lies outside the source window; "
         "the block geometry cannot support this order");
  return unit_lagrange_weights(p, x - l);
}

void conservative_prolong_weights(int p, std::vector<Rational>& low, std::vector<Rational>& high) {
  int r = (p - 1) / 2;
  std::vector<Rational> bounds;
  for (int i = 0; i <= p; ++i) bounds.push_back(Rational(-r) - Rational(1, 2) + Rational(i));
  auto L = lagrange_weights(bounds, Rational(0));
  low.assign(p, Rational(0));
  high.assign(p, Rational(0));
  for (int t = 0; t < p; ++t) {
    Rational tail(0);
    for (int i = t + 1; i <= p; ++i) tail = tail + L[i];
    low[t] = Rational(2) * (tail - Rational(t < r ? 1 : 0));
    high[t] = Rational(2) * (Rational(t < r + 1 ? 1 : 0) - tail);
  }
}

}  // namespace tah
