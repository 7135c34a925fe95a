// Chi-square CDF for the G^2 decision, FP64, usable on host and device.
//
// The reference calls stats::pchisq(g2, df, false) (IndependenceTest.cpp:146,268,355)
// from the third-party StatsLib/gcem headers, which are not part of the source tree.
// pchisq(x, k) is the regularized lower incomplete gamma P(k/2, x/2); it is evaluated
// here by the power series (x < a+1) or the modified-Lentz continued fraction for Q
// (Press et al., Numerical Recipes 3rd ed., section 6.2), written from the mathematics.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace fpgm {

__host__ __device__ inline double gamma_p(double a, double x)
{
    if (!(x > 0.0)) return 0.0;
    const double eps = 1e-16;
    const double lpre = -x + a * log(x) - lgamma(a);
    if (x < a + 1.0) {
        double ap = a, del = 1.0 / a, sum = del;
        for (int it = 0; it < 1000000; ++it) {
            ap += 1.0;
            del *= x / ap;
            sum += del;
            if (fabs(del) < fabs(sum) * eps) break;
        }
        return sum * exp(lpre);
    }
    const double fpmin = 1e-300;
    double b = x + 1.0 - a, c = 1.0 / fpmin, d = 1.0 / b, h = d;
    for (int i = 1; i < 1000000; ++i) {
        const double an = -i * (i - a);
        b += 2.0;
        d = an * d + b;
        if (fabs(d) < fpmin) d = fpmin;
        c = b + an / c;
        if (fabs(c) < fpmin) c = fpmin;
        d = 1.0 / d;
        const double del = d * c;
        h *= del;
        if (fabs(del - 1.0) < eps) break;
    }
    return 1.0 - exp(lpre) * h;
}

// chi-square CDF with df degrees of freedom
__host__ __device__ inline double pchisq(double x, int df)
{
    if (df <= 0) return (x > 0.0) ? 1.0 : 0.0;
    return gamma_p(0.5 * df, 0.5 * x);
}

// density of the chi-square distribution
__host__ inline double dchisq(double x, int df)
{
    if (!(x > 0.0)) return 0.0;
    const double a = 0.5 * df;
    return 0.5 * exp((a - 1.0) * log(0.5 * x) - 0.5 * x - lgamma(a));
}

// z with P(N(0,1) > z) = alpha, by bisection on erfc (host only)
__host__ inline double norm_upper_quantile(double alpha)
{
    double zlo = -40.0, zhi = 40.0;
    for (int i = 0; i < 200; ++i) {
        const double zm = 0.5 * (zlo + zhi);
        if (zm <= zlo || zm >= zhi) break;
        if (0.5 * erfc(zm / sqrt(2.0)) > alpha) zlo = zm; else zhi = zm;
    }
    return 0.5 * (zlo + zhi);
}

// critical value c with 1 - pchisq(c, df) = alpha: Wilson-Hilferty start, then
// Newton steps kept inside a bisection bracket.
__host__ inline double chisq_crit(int df, double alpha)
{
    if (df <= 0 || !(alpha > 0.0) || !(alpha < 1.0)) return NAN;
    const double target = 1.0 - alpha;
    const double z = norm_upper_quantile(alpha);
    const double k = df;
    const double t = 1.0 - 2.0 / (9.0 * k) + z * sqrt(2.0 / (9.0 * k));
    double x = k * t * t * t;
    if (!(x > 0.0)) x = 0.5 * k + 1e-3;
    double lo = 0.0, hi = k + 10.0;
    while (pchisq(hi, df) < target) hi *= 2.0;
    if (x <= lo || x >= hi) x = 0.5 * (lo + hi);
    for (int it = 0; it < 200; ++it) {
        const double f = pchisq(x, df) - target;
        if (f < 0.0) lo = x; else hi = x;
        const double fp = dchisq(x, df);
        double xn = (fp > 0.0) ? x - f / fp : 0.5 * (lo + hi);
        if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
        if (fabs(xn - x) <= 1e-14 * x) {
            x = xn;
            break;
        }
        x = xn;
    }
    return x;
}

}  // namespace fpgm
