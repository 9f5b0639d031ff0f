/*
 * tlm_dual.h -- forward-mode dual number for the tangent-linear build of the oracle
 * (TEST INFRASTRUCTURE, see tlm.cpp).
 *
 * A dual x = (v, d) carries a value and one directional derivative. Arithmetic follows
 * the differentiation rules; comparisons use the value only, so every branch, mask and
 * clamp of the oracle takes the path the plain build takes and the tangent is the
 * derivative of that path (the tangent-linear model of the discrete step).
 */
#ifndef M70_TLM_DUAL_H_
#define M70_TLM_DUAL_H_

#include <cmath>

struct orc_dual {
    double v, d;
    orc_dual() = default;
    orc_dual(double x) : v(x), d(0.0) {}
    orc_dual(double x, double dx) : v(x), d(dx) {}
    explicit operator float() const { return (float)v; }
    orc_dual &operator+=(const orc_dual &o) { d += o.d; v += o.v; return *this; }
    orc_dual &operator-=(const orc_dual &o) { d -= o.d; v -= o.v; return *this; }
    orc_dual &operator*=(const orc_dual &o) { d = d * o.v + v * o.d; v *= o.v; return *this; }
    orc_dual &operator/=(const orc_dual &o)
    {
        const double q = v / o.v;
        d = (d - q * o.d) / o.v;
        v = q;
        return *this;
    }
};

inline orc_dual operator-(const orc_dual &a) { return orc_dual(-a.v, -a.d); }
inline orc_dual operator+(const orc_dual &a) { return a; }
inline orc_dual operator+(const orc_dual &a, const orc_dual &b) { return orc_dual(a.v + b.v, a.d + b.d); }
inline orc_dual operator-(const orc_dual &a, const orc_dual &b) { return orc_dual(a.v - b.v, a.d - b.d); }
inline orc_dual operator*(const orc_dual &a, const orc_dual &b) { return orc_dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline orc_dual operator/(const orc_dual &a, const orc_dual &b)
{
    const double q = a.v / b.v;
    return orc_dual(q, (a.d - q * b.d) / b.v);
}
inline orc_dual operator+(double a, const orc_dual &b) { return orc_dual(a + b.v, b.d); }
inline orc_dual operator+(const orc_dual &a, double b) { return orc_dual(a.v + b, a.d); }
inline orc_dual operator-(double a, const orc_dual &b) { return orc_dual(a - b.v, -b.d); }
inline orc_dual operator-(const orc_dual &a, double b) { return orc_dual(a.v - b, a.d); }
inline orc_dual operator*(double a, const orc_dual &b) { return orc_dual(a * b.v, a * b.d); }
inline orc_dual operator*(const orc_dual &a, double b) { return orc_dual(a.v * b, a.d * b); }
inline orc_dual operator/(double a, const orc_dual &b)
{
    const double q = a / b.v;
    return orc_dual(q, -q * b.d / b.v);
}
inline orc_dual operator/(const orc_dual &a, double b) { return orc_dual(a.v / b, a.d / b); }

inline bool operator<(const orc_dual &a, const orc_dual &b) { return a.v < b.v; }
inline bool operator>(const orc_dual &a, const orc_dual &b) { return a.v > b.v; }
inline bool operator<=(const orc_dual &a, const orc_dual &b) { return a.v <= b.v; }
inline bool operator>=(const orc_dual &a, const orc_dual &b) { return a.v >= b.v; }
inline bool operator==(const orc_dual &a, const orc_dual &b) { return a.v == b.v; }
inline bool operator!=(const orc_dual &a, const orc_dual &b) { return a.v != b.v; }

// A zero tangent stays zero (also where the derivative is infinite, as cbrt at 0 for a
// dry cell's dissipation: the plain build's value there does not move).
inline orc_dual sqrt(const orc_dual &a)
{
    const double s = std::sqrt(a.v);
    return orc_dual(s, a.d == 0.0 ? 0.0 : a.d / (2.0 * s));
}
inline orc_dual cbrt(const orc_dual &a)
{
    const double c = std::cbrt(a.v);
    return orc_dual(c, a.d == 0.0 ? 0.0 : a.d / (3.0 * c * c));
}
inline orc_dual tanh(const orc_dual &a)
{
    const double t = std::tanh(a.v);
    return orc_dual(t, a.d == 0.0 ? 0.0 : a.d * (1.0 - t * t));
}
inline orc_dual fabs(const orc_dual &a) { return a.v < 0 ? orc_dual(-a.v, -a.d) : a; }
inline orc_dual log(const orc_dual &a) { return orc_dual(std::log(a.v), a.d == 0.0 ? 0.0 : a.d / a.v); }
inline orc_dual exp(const orc_dual &a)
{
    const double e = std::exp(a.v);
    return orc_dual(e, e * a.d);
}

#endif
