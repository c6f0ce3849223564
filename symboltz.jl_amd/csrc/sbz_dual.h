// Forward-mode dual numbers (value + one directional derivative) for host and device code.
//
// The reference obtains the explicit time derivative of the perturbation right-hand side by
// pushing ForwardDiff dual numbers through the generated f! (OrdinaryDiffEqDifferentiation,
// call site src/solve.jl:566).  This header gives the same mechanism in C++: every background
// quantity is evaluated on `Dual`, so d/dτ of every observed expression is exact, including the
// derivative of the cubic Hermite spline.
#pragma once
#ifndef SBZ_HD
#define SBZ_HD __host__ __device__ inline
#endif

struct Dual {
    double v, d;
    SBZ_HD Dual() : v(0.0), d(0.0) {}
    SBZ_HD Dual(double x) : v(x), d(0.0) {}
    SBZ_HD Dual(double x, double dx) : v(x), d(dx) {}
};

SBZ_HD Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
SBZ_HD Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
SBZ_HD Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
SBZ_HD Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
SBZ_HD Dual operator/(Dual a, Dual b) {
    double r = 1.0 / b.v;
    double q = a.v * r;
    return Dual(q, (a.d - q * b.d) * r);
}
SBZ_HD Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
SBZ_HD Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
SBZ_HD Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
SBZ_HD Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
SBZ_HD Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
SBZ_HD Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
SBZ_HD Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
SBZ_HD Dual operator/(double a, Dual b) {
    double q = a / b.v;
    return Dual(q, -q * b.d / b.v);
}

SBZ_HD double dval(double x) { return x; }
SBZ_HD double dval(Dual x) { return x.v; }
SBZ_HD double dder(double) { return 0.0; }
SBZ_HD double dder(Dual x) { return x.d; }

SBZ_HD double sbz_exp(double x) { return exp(x); }
SBZ_HD double sbz_log(double x) { return log(x); }
SBZ_HD double sbz_sqrt(double x) { return sqrt(x); }
SBZ_HD double sbz_tanh(double x) { return tanh(x); }
SBZ_HD double sbz_pow(double x, double p) { return pow(x, p); }
SBZ_HD double sbz_abs(double x) { return fabs(x); }

SBZ_HD Dual sbz_exp(Dual x) { double e = exp(x.v); return Dual(e, e * x.d); }
SBZ_HD Dual sbz_log(Dual x) { return Dual(log(x.v), x.d / x.v); }
SBZ_HD Dual sbz_sqrt(Dual x) { double s = sqrt(x.v); return Dual(s, 0.5 * x.d / s); }
SBZ_HD Dual sbz_tanh(Dual x) { double t = tanh(x.v); return Dual(t, (1.0 - t * t) * x.d); }
SBZ_HD Dual sbz_pow(Dual x, double p) { double y = pow(x.v, p); return Dual(y, p * y / x.v * x.d); }
SBZ_HD Dual sbz_abs(Dual x) { return x.v < 0 ? -x : x; }

// Second-order Taylor numbers (value, d/dτ, d²/dτ²): the CMB source functions need the second
// time derivative of the visibility function v = -κ̇ e^{-κ} (cosmologies.jl:99-105), which the
// reference obtains by expanding D(D(...)) symbolically.
struct Tay2 {
    double v, d, dd;
    SBZ_HD Tay2() : v(0.0), d(0.0), dd(0.0) {}
    SBZ_HD Tay2(double x) : v(x), d(0.0), dd(0.0) {}
    SBZ_HD Tay2(double x, double dx, double ddx) : v(x), d(dx), dd(ddx) {}
};
SBZ_HD Tay2 operator+(Tay2 a, Tay2 b) { return Tay2(a.v + b.v, a.d + b.d, a.dd + b.dd); }
SBZ_HD Tay2 operator-(Tay2 a, Tay2 b) { return Tay2(a.v - b.v, a.d - b.d, a.dd - b.dd); }
SBZ_HD Tay2 operator-(Tay2 a) { return Tay2(-a.v, -a.d, -a.dd); }
SBZ_HD Tay2 operator*(Tay2 a, Tay2 b) {
    return Tay2(a.v * b.v, a.d * b.v + a.v * b.d, a.dd * b.v + 2.0 * a.d * b.d + a.v * b.dd);
}
SBZ_HD Tay2 operator/(Tay2 a, Tay2 b) {
    double q = a.v / b.v;
    double qd = (a.d - q * b.d) / b.v;
    return Tay2(q, qd, (a.dd - 2.0 * qd * b.d - q * b.dd) / b.v);
}
SBZ_HD Tay2 operator+(Tay2 a, double b) { return Tay2(a.v + b, a.d, a.dd); }
SBZ_HD Tay2 operator+(double a, Tay2 b) { return Tay2(a + b.v, b.d, b.dd); }
SBZ_HD Tay2 operator-(Tay2 a, double b) { return Tay2(a.v - b, a.d, a.dd); }
SBZ_HD Tay2 operator-(double a, Tay2 b) { return Tay2(a - b.v, -b.d, -b.dd); }
SBZ_HD Tay2 operator*(Tay2 a, double b) { return Tay2(a.v * b, a.d * b, a.dd * b); }
SBZ_HD Tay2 operator*(double a, Tay2 b) { return Tay2(a * b.v, a * b.d, a * b.dd); }
SBZ_HD Tay2 operator/(Tay2 a, double b) { return Tay2(a.v / b, a.d / b, a.dd / b); }
SBZ_HD Tay2 operator/(double a, Tay2 b) { return Tay2(a) / b; }
SBZ_HD double dval(Tay2 x) { return x.v; }
SBZ_HD Tay2 sbz_exp(Tay2 x) { double e = exp(x.v); return Tay2(e, e * x.d, e * (x.dd + x.d * x.d)); }
SBZ_HD Tay2 sbz_log(Tay2 x) { return Tay2(log(x.v), x.d / x.v, x.dd / x.v - (x.d * x.d) / (x.v * x.v)); }
SBZ_HD Tay2 sbz_sqrt(Tay2 x) {
    double s = sqrt(x.v);
    double sd = 0.5 * x.d / s;
    return Tay2(s, sd, (0.5 * x.dd - sd * sd) / s);
}
SBZ_HD Tay2 sbz_tanh(Tay2 x) {
    double t = tanh(x.v), s = 1.0 - t * t;
    return Tay2(t, s * x.d, s * x.dd - 2.0 * t * s * x.d * x.d);
}
SBZ_HD Tay2 sbz_pow(Tay2 x, double p) {
    double y1 = pow(x.v, p - 1.0), y = y1 * x.v;
    return Tay2(y, p * y1 * x.d, p * y1 * x.dd + p * (p - 1.0) * (y1 / x.v) * x.d * x.d);
}
SBZ_HD Tay2 sbz_abs(Tay2 x) { return x.v < 0 ? -x : x; }
