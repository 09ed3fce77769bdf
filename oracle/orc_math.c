/*
 * Oracle (test infrastructure): integer/closed-form helpers of the GRPP path.
 */
#include <math.h>
#include <stdlib.h>
#include "orc.h"

/* n! as double; exact up to 20!, recursive product beyond (factorial.c:37-47). */
double orc_factorial(int n)
{
    if (n < 0) return 1.0;
    double f = 1.0;
    uint64_t exact = 1;
    int k = 1;
    for (; k <= n && k <= 20; k++) exact *= (uint64_t) k;
    f = (double) exact;
    for (; k <= n; k++) f *= k;
    return f;
}

/* n!/m! by a running product (factorial.c:55-69). */
double orc_factorial_ratio(int n, int m)
{
    if (n == m) return 1.0;
    if (n < m) return 1.0 / orc_factorial_ratio(m, n);
    double p = 1.0;
    for (int i = m + 1; i <= n; i++) p *= i;
    return p;
}

/* n!! for -1 <= n <= 30 from exact 64-bit products (factorial.c:105-115 asserts n<=30). */
double orc_double_factorial(int n)
{
    if (n > 30) abort();
    if (n <= 0) return 1.0;
    uint64_t p = 1;
    for (int k = n; k > 1; k -= 2) p *= (uint64_t) k;
    return (double) p;
}

/* binomial coefficient, exact in 64-bit for the small arguments of this path (binomial.c:24-59). */
uint64_t orc_binomial(uint64_t n, uint64_t k)
{
    if (k > n) return 0;
    if (k == 0 || k == n) return 1;
    if (k > n - k) k = n - k;
    uint64_t r = 1;
    for (uint64_t d = 1; d <= k; d++) {
        r = r * (n - k + d) / d;   /* r*(n-k+d) is divisible by d at every step */
    }
    return r;
}

/* (2a/pi)^{3/4} (4a)^{(n+l+m)/2}: the per-primitive factor libgrpp applies (norm_gaussian.c:16-23). */
double orc_gaussian_norm_factor(int n, int l, int m, double alpha)
{
    return pow(2 * alpha / M_PI, 0.75) * pow(4 * alpha, 0.5 * (n + l + m));
}

/* Cartesian components of angular momentum L: x-power descending, then y-power
 * descending (parameters.c:94-134; the DIRAC and TURBOMOLE generators give the same list). */
int orc_cartesian_generator(int L, int *cart_list)
{
    int c = 0;
    for (int r = L; r >= 0; r--)
        for (int s = L - r; s >= 0; s--) {
            cart_list[3 * c + 0] = r;
            cart_list[3 * c + 1] = s;
            cart_list[3 * c + 2] = L - r - s;
            c++;
        }
    return c;
}

/* coordinates equal within 1e-12 per component (utils.c:96-107) */
int orc_points_are_equal(const double *a, const double *b)
{
    const double t = 1e-12;
    return fabs(a[0] - b[0]) < t && fabs(a[1] - b[1]) < t && fabs(a[2] - b[2]) < t;
}

/* int_0^inf r^n exp(-a r^2) dr in closed form (screening.c:255-265). */
double orc_gaussian_integral(int n, double a)
{
    if (n % 2 == 0) {
        int k = n / 2;
        return orc_double_factorial(2 * k - 1) / (pow(2.0, k + 1) * pow(a, k)) * sqrt(M_PI / a);
    }
    int k = (n - 1) / 2;
    return orc_factorial(k) / (2.0 * pow(a, k + 1));
}
