// Forward-mode dual number (value + one tangent) for the per-particle
// derivatives of the load-step adjoint.  Branches compare values only, so a
// dual evaluation follows the same branch as the double evaluation and gives
// the derivative of that branch.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define GM_DD __device__ __forceinline__

struct Dual {
  double v, d;
};

GM_DD Dual dmk(double v, double d = 0.0) {
  Dual r;
  r.v = v;
  r.d = d;
  return r;
}
GM_DD double val(double x) { return x; }
GM_DD double val(const Dual& x) { return x.v; }

GM_DD Dual operator+(Dual a, Dual b) { return dmk(a.v + b.v, a.d + b.d); }
GM_DD Dual operator-(Dual a, Dual b) { return dmk(a.v - b.v, a.d - b.d); }
GM_DD Dual operator*(Dual a, Dual b) { return dmk(a.v * b.v, a.d * b.v + a.v * b.d); }
GM_DD Dual operator/(Dual a, Dual b) {
  double q = a.v / b.v;
  return dmk(q, (a.d - q * b.d) / b.v);
}
GM_DD Dual operator-(Dual a) { return dmk(-a.v, -a.d); }
GM_DD Dual operator+(Dual a, double b) { return dmk(a.v + b, a.d); }
GM_DD Dual operator+(double a, Dual b) { return dmk(a + b.v, b.d); }
GM_DD Dual operator-(Dual a, double b) { return dmk(a.v - b, a.d); }
GM_DD Dual operator-(double a, Dual b) { return dmk(a - b.v, -b.d); }
GM_DD Dual operator*(Dual a, double b) { return dmk(a.v * b, a.d * b); }
GM_DD Dual operator*(double a, Dual b) { return dmk(a * b.v, a * b.d); }
GM_DD Dual operator/(Dual a, double b) { return dmk(a.v / b, a.d / b); }
GM_DD Dual operator/(double a, Dual b) {
  double q = a / b.v;
  return dmk(q, -q * b.d / b.v);
}
GM_DD Dual& operator+=(Dual& a, Dual b) {
  a.v += b.v;
  a.d += b.d;
  return a;
}
GM_DD Dual& operator-=(Dual& a, Dual b) {
  a.v -= b.v;
  a.d -= b.d;
  return a;
}
GM_DD Dual& operator*=(Dual& a, Dual b) {
  a = a * b;
  return a;
}
GM_DD Dual& operator+=(Dual& a, double b) {
  a.v += b;
  return a;
}
GM_DD Dual& operator-=(Dual& a, double b) {
  a.v -= b;
  return a;
}

GM_DD Dual sqrt(Dual a) {
  double s = sqrt(a.v);
  return dmk(s, a.d / (2.0 * s));
}
GM_DD Dual log(Dual a) { return dmk(log(a.v), a.d / a.v); }
GM_DD Dual exp(Dual a) {
  double e = exp(a.v);
  return dmk(e, e * a.d);
}
GM_DD Dual sin(Dual a) { return dmk(sin(a.v), cos(a.v) * a.d); }
GM_DD Dual cos(Dual a) { return dmk(cos(a.v), -sin(a.v) * a.d); }
GM_DD Dual fabs(Dual a) { return a.v < 0.0 ? -a : a; }
GM_DD Dual atan2(Dual y, Dual x) {
  double r2 = x.v * x.v + y.v * y.v;
  return dmk(atan2(y.v, x.v), (x.v * y.d - y.v * x.d) / r2);
}
// fmax against a constant floor: the floor branch has zero derivative
GM_DD Dual fmax(Dual a, double b) { return a.v >= b ? a : dmk(b, 0.0); }
