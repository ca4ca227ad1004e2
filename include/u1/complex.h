// complex.h -- drop-in for the reference's include/complex.h (_complex<Real>,
// complexd, complexs).  Same member names and arithmetic conventions
// (complex.h:205-300): multiplication (a.x b.x - a.y b.y, a.y b.x + a.x b.y),
// abs = sqrt(x^2 + y^2), conj, arg/phase/angle, stream output "re<TAB>im"
// (complex.h:367-372, the format of the reference's data files).
#ifndef U1_DROPIN_COMPLEX_H
#define U1_DROPIN_COMPLEX_H
#include <cmath>
#include <iostream>

namespace U1 {

template <class Real>
class _complex {
 public:
  Real x, y;
  _complex() : x(0), y(0) {}
  _complex(Real re) : x(re), y(0) {}
  _complex(Real re, Real im) : x(re), y(im) {}
  template <class R2>
  explicit _complex(const _complex<R2>& o) : x((Real)o.x), y((Real)o.y) {}

  Real& real() { return x; }
  Real& imag() { return y; }
  Real real() const { return x; }
  Real imag() const { return y; }

  _complex operator+(const _complex& b) const { return _complex(x + b.x, y + b.y); }
  _complex operator-(const _complex& b) const { return _complex(x - b.x, y - b.y); }
  _complex operator-() const { return _complex(-x, -y); }
  _complex operator*(const _complex& b) const { return _complex(x * b.x - y * b.y, y * b.x + x * b.y); }
  _complex operator*(Real b) const { return _complex(x * b, y * b); }
  _complex operator/(Real b) const { return _complex(x / b, y / b); }
  _complex operator/(const _complex& b) const {
    const Real t = (Real)1 / (b.x * b.x + b.y * b.y);
    return _complex((x * b.x + y * b.y) * t, (y * b.x - x * b.y) * t);
  }
  _complex& operator+=(const _complex& b) { x += b.x; y += b.y; return *this; }
  _complex& operator-=(const _complex& b) { x -= b.x; y -= b.y; return *this; }
  _complex& operator*=(const _complex& b) { *this = *this * b; return *this; }
  _complex& operator*=(Real b) { x *= b; y *= b; return *this; }
  _complex& operator/=(Real b) { x /= b; y /= b; return *this; }
  _complex& operator/=(const _complex& b) { *this = *this / b; return *this; }
  bool operator==(const _complex& b) const { return x == b.x && y == b.y; }
  bool operator!=(const _complex& b) const { return !(*this == b); }

  _complex conj() const { return _complex(x, -y); }
  _complex operator~() const { return conj(); }
  Real abs() const { return std::sqrt(x * x + y * y); }
  Real abs2() const { return x * x + y * y; }
  Real phase() const { return std::atan2(y, x); }
  Real angle() const { return std::atan2(y, x); }
  Real arg() const { return std::atan2(y, x); }

  friend _complex operator+(Real a, const _complex& z) { return _complex(a + z.x, z.y); }
  friend _complex operator-(Real a, const _complex& z) { return _complex(a - z.x, -z.y); }
  friend _complex operator*(Real a, const _complex& z) { return _complex(z.x * a, z.y * a); }
  friend _complex conj(const _complex& z) { return z.conj(); }
  friend Real abs(const _complex& z) { return z.abs(); }
  friend Real abs2(const _complex& z) { return z.abs2(); }
  friend Real arg(const _complex& z) { return z.arg(); }
  friend std::ostream& operator<<(std::ostream& out, const _complex& z) {
    return out << z.x << '\t' << z.y;
  }
};

typedef _complex<double> complexd;
typedef _complex<float> complexs;

inline complexd exp_ir(const double a) { return complexd(std::cos(a), std::sin(a)); }

}  // namespace U1
#endif
