// Host check: the select-only frame of an axis-aligned area vector
// (axis_frame_t) equals the general path (tangent_binormal_t + make_frame_t)
// bit for bit, every member and signed zeros included, in FP64 and FP32, for
// random signs and magnitudes over and beyond the range the kernels use it
// for; one vector in 97 is not axis-aligned and must be refused.
// Usage: axis_frame_check [count]; exits non-zero on any mismatch.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include "physics.hpp"
using namespace ma;
static uint64_t s = 0x243F6A8885A308D3ull;
static uint64_t nxt() { s += 0x9E3779B97F4A7C15ull; uint64_t z = s; z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull; z = (z ^ (z >> 27)) * 0x94D049BB133111EBull; return z ^ (z >> 31); }
template <class T> bool same(const T& x, const T& y) { return std::memcmp(&x, &y, sizeof(T)) == 0; }
template <class T> bool same_frame(const FaceFrameT<T>& a, const FaceFrameT<T>& b) {
  for (int d = 0; d < 3; ++d)
    if (!same(a.n[d], b.n[d]) || !same(a.t[d], b.t[d]) || !same(a.b[d], b.b[d]) || !same(a.nu[d], b.nu[d]) ||
        !same(a.tu[d], b.tu[d]) || !same(a.bu[d], b.bu[d])) return false;
  return same(a.nn, b.nn);
}
template <class T> long run(long n, int emin, int emax) {
  long bad = 0, fast = 0;
  for (long i = 0; i < n; ++i) {
    T a[3];
    const int ax = (int)(nxt() % 3);
    for (int d = 0; d < 3; ++d) a[d] = (nxt() & 1) ? T(-0.0) : T(0.0);
    const double m = std::ldexp(1.0 + (double)(nxt() >> 12) / 4503599627370496.0, emin + (int)(nxt() % (uint64_t)(emax - emin + 1)));
    a[ax] = T((nxt() & 1) ? -m : m);
    const bool skew = i % 97 == 0;
    if (skew) a[(ax + 1) % 3] = T(std::ldexp(1.0, -(int)(nxt() % 40)));  // not axis-aligned
    if (skew && is_axis_area<T>(a)) ++bad;
    FaceFrameT<T> g;
    T t[3], b[3];
    tangent_binormal_t<T>(a, t, b);
    g = make_frame_t<T>(a, t, b);
    if (is_axis_area<T>(a)) {
      const FaceFrameT<T> f = axis_frame_t<T>(a);
      ++fast;
      if (!same_frame(f, g)) { if (bad < 5) printf("mismatch a=(%a %a %a)\n", (double)a[0], (double)a[1], (double)a[2]); ++bad; }
    }
  }
  printf("%s: n=%ld fast=%ld mismatches=%ld\n", sizeof(T) == 8 ? "double" : "float", n, fast, bad);
  return bad;
}
int main(int argc, char** argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 20000000L;
  long bad = run<double>(n, -520, 520) + run<double>(n, -40, 10);
  bad += run<float>(n, -70, 70) + run<float>(n, -30, 5);
  return bad ? 1 : 0;
}
