// CPU ORACLE (test infrastructure only) — the CMB angular-spectrum stage of SymBoltz.jl.
//
// Only tests/ and bench.py's cpu_baseline leg load this library, and only as the checker; the
// product (libsbz_hip.so) never calls it.
//
// What it restates:
//   cmbo_bessel_table  SphericalBesselCache (src/observables/angular.jl:9-48): j_l(x) and
//                      j_l'(x) = l/(2l+1) j_{l-1} - (l+1)/(2l+1) j_{l+1} (:60) on x_i = i dx,
//                      i < n_x - 1, plus one duplicate padding point.  The reference takes j_l from
//                      Bessels.jl; here j_l comes from the three-term recurrence
//                      j_{l+1} = (2l+1)/x j_l - j_{l-1}: upward from j_0 = sin x / x and
//                      j_1 = (sin x / x - cos x) / x where every l needed is below x (the
//                      recurrence is stable there), otherwise Miller's downward recurrence from
//                      far above max(l, x), normalised to j_0 and j_1 by least squares.
//   cmbo_kinterp       source_kinterp with a ChebyshevInterpolator (fourier.jl:232-242, 520-532):
//                      barycentric interpolation in y = f(k) with the second-kind Chebyshev weights,
//                      the exact node value when y hits a node.
//   cmbo_los           los_integrate (angular.jl:109-185): trapezoid weights, χ = τ0 - τ, Hermite
//                      reads of the Bessel table, and for l ≥ l_limber the Limber value
//                      √(π/(2l+1)) S(τ0 - (l+½)/k) / k with S read by a cubic Hermite in χ whose
//                      slopes are finite differences of the neighbouring samples.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

namespace {

template <class F>
void parallel_for(long long n, int threads, F f) {
    if (threads <= 1 || n < 2) {
        for (long long i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    std::atomic<long long> next{0};
    for (int t = 0; t < threads; ++t)
        th.emplace_back([&] {
            for (long long i; (i = next.fetch_add(1)) < n;) f(i);
        });
    for (auto& t : th) t.join();
}

// j_0 .. j_{L} at x > 0 into out[0..L]
void bessel_all(double x, int L, std::vector<double>& out) {
    out.assign(L + 1, 0.0);
    const double s = std::sin(x), c = std::cos(x);
    const double j0 = s / x, j1 = (s / x - c) / x;
    if (x > L + 1.0) {
        out[0] = j0;
        if (L >= 1) out[1] = j1;
        for (int l = 1; l < L; ++l) out[l + 1] = (2.0 * l + 1.0) / x * out[l] - out[l - 1];
        return;
    }
    // Miller: start well above the turning point max(L, x); the ratio j/y of the recurrence's
    // two solutions decays like exp(-(4√2/3) ((l - x) / x^(1/3))^(3/2)) past it
    const double top = std::max((double)L, x);
    const int ltop = (int)(top + 40.0 + 12.0 * std::cbrt(top)) + 1;
    std::vector<double> d(ltop + 2, 0.0);
    std::vector<int> ex(ltop + 2, 0);   // d[l] was stored when the running scale was 2^-ex[l]
    int e = 0;
    double dp = 0.0, dc = 1.0;          // d_{l+1}, d_l
    d[ltop] = dc;
    for (int l = ltop; l >= 1; --l) {
        double dm = (2.0 * l + 1.0) / x * dc - dp;
        dp = dc;
        dc = dm;
        if (std::fabs(dc) > 1e200) {
            dc = std::ldexp(dc, -600);
            dp = std::ldexp(dp, -600);
            e += 600;
        }
        d[l - 1] = dc;
        ex[l - 1] = e;
    }
    // the value at l is d[l] 2^(e_final - ex[l]) relative to the final scale
    auto val = [&](int l) { return std::ldexp(d[l], ex[l] - e); };
    const double m = std::max(std::fabs(val(0)), std::fabs(val(1)));
    const double a0 = val(0) / m, a1 = val(1) / m;
    const double C = (j0 * a0 + j1 * a1) / (a0 * a0 + a1 * a1) / m;
    for (int l = 0; l <= L; ++l) out[l] = C * val(l);
}

}  // namespace

extern "C" {

void cmbo_bessel_table(const int* ls, int n_l, int n_x, double dx, double* y, double* dy, int threads) {
    int lmax = 0;
    for (int i = 0; i < n_l; ++i) lmax = std::max(lmax, ls[i]);
    parallel_for(n_x - 1, threads, [&](long long i) {
        double* yr = y + i * n_l;
        double* dr = dy + i * n_l;
        const double x = (double)i * dx;
        if (x == 0.0) {
            for (int il = 0; il < n_l; ++il) {
                yr[il] = ls[il] == 0 ? 1.0 : 0.0;
                dr[il] = ls[il] == 1 ? 1.0 / 3.0 : 0.0;
            }
            return;
        }
        std::vector<double> j;
        bessel_all(x, lmax + 1, j);
        for (int il = 0; il < n_l; ++il) {
            const int l = ls[il];
            yr[il] = j[l];
            dr[il] = l == 0 ? -j[1] : (l * j[l - 1] - (l + 1.0) * j[l + 1]) / (2.0 * l + 1.0);
        }
    });
    for (int il = 0; il < n_l; ++il) {   // padding point = the last grid point
        y[(long long)(n_x - 1) * n_l + il] = y[(long long)(n_x - 2) * n_l + il];
        dy[(long long)(n_x - 1) * n_l + il] = dy[(long long)(n_x - 2) * n_l + il];
    }
}

// S_coarse [n_c][m] (m values per node), nodes y_c and weights w_c; out [n_f][m] at y_f
void cmbo_kinterp(const double* S, int n_c, long long m, const double* yc, const double* wc, const double* yf, int n_f,
                  double* out, int threads) {
    parallel_for(n_f, threads, [&](long long i) {
        double* o = out + i * m;
        const double y = yf[i];
        for (int j = 0; j < n_c; ++j)
            if (y - yc[j] == 0.0) {
                for (long long q = 0; q < m; ++q) o[q] = S[j * m + q];
                return;
            }
        std::vector<double> num(m, 0.0);
        double den = 0.0;
        for (int j = 0; j < n_c; ++j) {
            const double t = wc[j] / (y - yc[j]);
            den += t;
            for (long long q = 0; q < m; ++q) num[q] += t * S[j * m + q];
        }
        for (long long q = 0; q < m; ++q) o[q] = num[q] / den;
    });
}

// S [n_k][n_t][n_s]; out [n_s][n_k][n_l]; l_limber per source (INT_MAX: none)
void cmbo_los(const double* S, int n_s, const double* taus, int n_t, const double* ks, int n_k, const int* ls, int n_l,
              const double* y, const double* dy, int n_x, double dx, const int* l_limber, double* out, int threads) {
    std::vector<double> w(n_t), chi(n_t);
    w[0] = 0.5 * (taus[1] - taus[0]);
    for (int i = 1; i < n_t - 1; ++i) w[i] = 0.5 * (taus[i + 1] - taus[i - 1]);
    w[n_t - 1] = 0.5 * (taus[n_t - 1] - taus[n_t - 2]);
    const double tau0 = taus[n_t - 1];
    for (int i = 0; i < n_t; ++i) chi[i] = tau0 - taus[i];
    const double invdx = 1.0 / dx;
    parallel_for(n_k, threads, [&](long long ik) {
        const double k = ks[ik];
        for (int s = 0; s < n_s; ++s) {
            auto Sv = [&](int it) { return S[(ik * n_t + it) * n_s + s]; };
            double* o = out + ((long long)s * n_k + ik) * n_l;
            int ilim = 0;   // first il with l ≥ l_limber
            while (ilim < n_l && ls[ilim] < l_limber[s]) ++ilim;
            for (int il = 0; il < n_l; ++il) o[il] = 0.0;
            for (int it = 0; it < n_t; ++it) {
                const double x = k * chi[it];
                const double Sw = w[it] * Sv(it);
                double fw = x * invdx;
                const int i = (int)fw;
                fw -= i;
                const double wm1 = fw - 1.0;
                const double* y0 = y + (long long)i * n_l;
                const double* d0 = dy + (long long)i * n_l;
                for (int il = 0; il < ilim; ++il) {
                    const double jl = (1 + 2 * fw) * wm1 * wm1 * y0[il] + fw * fw * (3 - 2 * fw) * y0[il + n_l] +
                                      fw * wm1 * (wm1 * d0[il] + fw * d0[il + n_l]) * dx;
                    o[il] += Sw * jl;
                }
            }
            for (int il = ilim; il < n_l; ++il) {
                const int l = ls[il];
                const double c = (l + 0.5) / k;
                if (!(c <= chi[0])) continue;
                int im = 0;   // first index with τ ≥ τ0 - χ (searchsortedfirst)
                while (im < n_t && taus[im] < tau0 - c) ++im;
                if (im == 0) continue;   // the reference leaves this entry at zero
                const int ip = im - 1;
                const double cm = chi[im], cp = chi[ip], Sm = Sv(im), Sp = Sv(ip);
                const double Dc = cp - cm;
                const double dSm = im <= n_t - 2 ? (Sv(im + 1) - Sp) / (chi[im + 1] - cp) : (Sp - Sm) / Dc;
                const double dSp = ip >= 1 ? (Sm - Sv(im - 2)) / (cm - chi[im - 2]) : (Sp - Sm) / Dc;
                const double t = (c - cm) / Dc, t2 = t * t, t3 = t2 * t;
                const double Sx = (2 * t3 - 3 * t2 + 1) * Sm + (t3 - 2 * t2 + t) * Dc * dSm + (-2 * t3 + 3 * t2) * Sp +
                                  (t3 - t2) * Dc * dSp;
                o[il] = std::sqrt(M_PI / (2.0 * l + 1.0)) * Sx / k;
            }
        }
    });
}

}  // extern "C"
