// erf_poly.cpp -- host fit of the erf kernel polynomial used by the FP64 pair loops (pair.hpp).
//
// erf(alpha r)/r = (2 alpha/sqrt(pi)) G(alpha^2 r^2),  G(u) = int_0^1 exp(-u t^2) dt,
// and G is entire with |G^(m)| <= 1/(2m + 1), so its Chebyshev coefficients on [0, U] fall off
// like (U/4)^m / (m! (2m + 1)).  The interpolant at 64 Chebyshev points is computed in long
// double, truncated at the degree asked for, converted to monomials in x = s xs - 1 (s = r^2 in
// [0, rmax^2]) and rounded to double once.
#include <cmath>
#include <vector>

#include "internal.hpp"

namespace aleator_gpu {

namespace {

long double g_kernel(long double u) {
    if (u < 0.5L) {
        // sum_n (-u)^n / (n! (2n + 1))
        long double term = 1.0L, sum = 1.0L;
        for (int n = 1; n < 40; ++n) {
            term *= -u / n;
            const long double add = term / (2 * n + 1);
            sum += add;
            if (fabsl(add) < 1e-24L) break;
        }
        return sum;
    }
    const long double su = sqrtl(u);
    return erfl(su) * sqrtl(3.141592653589793238462643383279502884L) / (2.0L * su);
}

}  // namespace

namespace {

constexpr int kN = 64;

// cos(pi j (k + 1/2) / 64) for j, k < 64, computed once (row 1 holds the Chebyshev points)
struct CosTable {
    long double c[kN][kN];
    CosTable() {
        constexpr long double kPiL = 3.141592653589793238462643383279502884L;
        for (int j = 0; j < kN; ++j)
            for (int k = 0; k < kN; ++k) c[j][k] = cosl(kPiL * j * (k + 0.5L) / kN);
    }
};

const CosTable& cos_table() {
    static const CosTable table;  // thread-safe one-time initialisation
    return table;
}

// Chebyshev coefficients of G(alpha^2 s) for s in [smin, smax] (64 points, long double)
void cheb_fit(double alpha, long double smin, long double smax, long double* cheb) {
    const CosTable& t = cos_table();
    const long double a2 = (long double)alpha * alpha;
    long double f[kN];
    for (int k = 0; k < kN; ++k) f[k] = g_kernel(a2 * (smin + (t.c[1][k] + 1.0L) * 0.5L * (smax - smin)));
    for (int j = 0; j < kN; ++j) {
        long double s = 0.0L;
        for (int k = 0; k < kN; ++k) s += f[k] * t.c[j][k];
        cheb[j] = 2.0L * s / kN;
    }
    cheb[0] *= 0.5L;
}

// monomial coefficients (lowest first, times 2 alpha/sqrt(pi)) of the degree-deg truncation
void to_monomials(double alpha, const long double* cheb, int deg, double* coef) {
    constexpr long double kPiL = 3.141592653589793238462643383279502884L;
    std::vector<long double> tprev(deg + 1, 0.0L), tcur(deg + 1, 0.0L), mono(deg + 1, 0.0L);
    tprev[0] = 1.0L;
    mono[0] += cheb[0];
    if (deg >= 1) {
        tcur[1] = 1.0L;
        mono[1] += cheb[1];
    }
    for (int j = 2; j <= deg; ++j) {
        std::vector<long double> next(deg + 1, 0.0L);
        for (int i = 0; i < deg; ++i) next[i + 1] += 2.0L * tcur[i];
        for (int i = 0; i <= deg; ++i) next[i] -= tprev[i];
        for (int i = 0; i <= deg; ++i) mono[i] += cheb[j] * next[i];
        tprev.swap(tcur);
        tcur.swap(next);
    }
    const long double pref = 2.0L * alpha / sqrtl(kPiL);
    for (int i = 0; i <= deg; ++i) coef[i] = (double)(pref * mono[i]);
}

long double tail_after(const long double* cheb, int deg) {
    long double tail = 0.0L;
    for (int j = deg + 1; j < kN; ++j) tail += fabsl(cheb[j]);
    return tail;
}

}  // namespace

bool erf_kernel_poly_range(double alpha, double rmin, double rmax, double tol, int deg, double* coef, double* xs,
                           double* xo) {
    const long double smin = (long double)rmin * rmin, smax = (long double)rmax * rmax;
    long double cheb[kN];
    cheb_fit(alpha, smin, smax, cheb);
    to_monomials(alpha, cheb, deg, coef);
    *xs = (double)(2.0L / (smax - smin));
    *xo = (double)(-(smax + smin) / (smax - smin));
    return tail_after(cheb, deg) <= tol;
}

int erf_kernel_poly(double alpha, double rmax, double tol, int min_deg, double* coef, double* xs) {
    const long double smax = (long double)rmax * rmax;
    long double cheb[kN];
    cheb_fit(alpha, 0.0L, smax, cheb);
    // smallest allowed degree (at least min_deg) whose dropped coefficients sum below tol
    static const int kDegrees[] = {18, 24, kErfPolyMax};
    int deg = kErfPolyMax;
    for (int d : kDegrees) {
        if (d < min_deg) continue;
        if (tail_after(cheb, d) <= tol) {
            deg = d;
            break;
        }
    }
    to_monomials(alpha, cheb, deg, coef);
    *xs = (double)(2.0L / smax);
    return deg;
}

}  // namespace aleator_gpu
