// Oracle (test infrastructure only): counting, spherical transformation, Boys function,
// Hermite R integrals and Hermite E coefficients, restated from the reference.
#include "oracle.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>

namespace orc
{
    // utils.cpp:12-31
    double doubleFactorial(int n)
    {
        double r = 1;
        if (n == 0 || n == -1)
            return r;
        if (n < -1)
            throw std::runtime_error("doubleFactorial: negative argument");
        for (int i = n; i >= 2; i -= 2)
            r *= i;
        return r;
    }

    // utils.cpp:6-10
    double purePrimitiveNorm(int l, double exp)
    {
        return std::sqrt(std::pow(2 * exp / M_PI, 1.5) * std::pow(4 * exp, l) /
                         doubleFactorial(2 * l - 1));
    }

    // utils.cpp:33-47: x-major descending
    std::vector<std::array<int, 3>> cartExps(int l)
    {
        std::vector<std::array<int, 3>> r;
        for (int x = l; x >= 0; x--)
            for (int y = l - x; y >= 0; y--)
                r.push_back({x, y, l - x - y});
        return r;
    }

    // utils.cpp:106-118: by n = t+u+v, then t descending, u descending
    std::vector<std::array<int, 3>> hermiteIdxs(int l)
    {
        std::vector<std::array<int, 3>> r;
        for (int n = 0; n <= l; n++)
            for (int t = n; t >= 0; t--)
                for (int u = n - t; u >= 0; u--)
                    r.push_back({t, u, n - t - u});
        return r;
    }

    static double fact(int n)
    {
        double r = 1;
        for (int i = 2; i <= n; i++)
            r *= i;
        return r;
    }

    static double binom(int n, int k)
    {
        if (k < 0 || k > n)
            return 0;
        return fact(n) / (fact(k) * fact(n - k));
    }

    // Real solid harmonics in the m order 0, +1, -1, +2, -2, ... (docs/integrals.rst:65-79),
    // from the closed form of Helgaker, Jorgensen, Olsen, "Molecular Electronic-Structure
    // Theory", eqs. (6.4.47)-(6.4.50).  The reference tabulates these coefficients rounded to
    // 14 decimals (spherical_trafo.cpp:9-690); the same rounding is applied here so that the
    // oracle reproduces its table bit for bit (pinned by the sphericalTrafo l=8 checksum).
    std::vector<std::tuple<int, int, double>> sphericalTrafo(int l)
    {
        if (l < 0 || l > 9)
            throw std::runtime_error("sphericalTrafo(): inappropriate angular momentum given");
        std::vector<std::tuple<int, int, double>> out;
        int ncart = nCart(l);
        for (int im = 0; im < 2 * l + 1; im++)
        {
            int m = (im == 0) ? 0 : ((im % 2 == 1) ? (im + 1) / 2 : -(im / 2));
            int am = std::abs(m);
            int vm2 = (m < 0) ? 1 : 0; // 2*v_m
            double norm = std::sqrt(2.0 * fact(l + am) * fact(l - am) / (m == 0 ? 2.0 : 1.0)) /
                          (std::pow(2.0, am) * fact(l));
            std::vector<double> c(ncart, 0.0);
            for (int t = 0; t <= (l - am) / 2; t++)
                for (int u = 0; u <= t; u++)
                    for (int v2 = vm2; v2 <= am; v2 += 2)
                    {
                        // exponent of (-1): t + v - v_m  with v - v_m = (v2 - vm2)/2
                        int sgn_exp = t + (v2 - vm2) / 2;
                        double coef = ((sgn_exp % 2) ? -1.0 : 1.0) * std::pow(0.25, t) *
                                      binom(l, t) * binom(l - t, am + t) * binom(t, u) *
                                      binom(am, v2);
                        int py = 2 * u + v2;
                        int px = 2 * t + am - py;
                        int pz = l - 2 * t - am;
                        if (px < 0)
                            continue;
                        c[idxCart(py, pz)] += norm * coef;
                    }
            for (int ic = 0; ic < ncart; ic++)
            {
                if (c[ic] == 0.0)
                    continue;
                char buf[64];
                std::snprintf(buf, sizeof(buf), "%.14f", c[ic]);
                double rounded = std::strtod(buf, nullptr);
                if (rounded != 0.0)
                    out.emplace_back(im, ic, rounded);
            }
        }
        return out;
    }

    // boys_function.cpp:6-69
    BoysGrid::BoysGrid(int n) : n_(n), max_n_(n + 6)
    {
        n_intervals_ = static_cast<int>(large_x_ / interval_) + 1;
        grid_.assign(static_cast<size_t>(n_intervals_) * (max_n_ + 1), 0.0);
        double x = 0;
        for (int iv = 0; iv < n_intervals_; iv++)
        {
            int k = 1;
            double term = 1.0 / (2.0 * max_n_ + 1.0);
            double sum = term;
            while (true)
            {
                term *= x / (max_n_ + 0.5 + k);
                sum += term;
                if (std::fabs(term) < thresh_)
                    break;
                k++;
            }
            double ex = std::exp(-x);
            double *row = &grid_[static_cast<size_t>(iv) * (max_n_ + 1)];
            row[max_n_] = ex * sum;
            for (int m = max_n_ - 1; m >= 0; m--)
                row[m] = (2.0 * x * row[m + 1] + ex) / (2 * m + 1);
            x += interval_;
        }
    }

    // boys_function.cpp:71-117
    void calcBoysF(int n, double x, const BoysGrid &g, double *fnx)
    {
        if (x == 0)
        {
            fnx[0] = 1;
            for (int k = 1; k <= n; k++)
                fnx[k] = 1.0 / (2 * k + 1);
        }
        else if (x > g.large_x_)
        {
            fnx[0] = 0.5 * std::sqrt(M_PI / x);
            for (int k = 1; k <= n; k++)
                fnx[k] = fnx[k - 1] * (k - 0.5) / x;
        }
        else
        {
            int iv = static_cast<int>(x / g.interval_);
            double dx = x - iv * g.interval_;
            double kf = 1.0, dk = 1.0;
            const double *row = &g.grid_[static_cast<size_t>(iv) * (g.max_n_ + 1)];
            double sum = row[n];
            for (int k = 1; k < 7; k++)
            {
                dk *= -dx;
                kf *= k;
                sum += row[n + k] * dk / kf;
            }
            fnx[n] = sum;
            double ex = std::exp(-x);
            for (int k = n - 1; k >= 0; k--)
                fnx[k] = (2.0 * x * fnx[k + 1] + ex) / (2 * k + 1);
        }
    }

    // rints.cpp:61-114 (HJO eqs. 9.9.18-9.9.20, loop order of the reference)
    std::vector<double> calcRInts3D(int l, double p, const double *xyz, const double *fnx)
    {
        const int d = l + 1;
        auto B = [d](int n, int t, int u, int v) -> size_t
        { return ((static_cast<size_t>(n) * d + t) * d + u) * d + v; };
        std::vector<double> buf(static_cast<size_t>(d) * d * d * d, 0.0);
        buf[B(0, 0, 0, 0)] = fnx[0];
        double xf = -2 * p, y = xf;
        for (int n = 1; n <= l; n++)
        {
            buf[B(n, 0, 0, 0)] = fnx[n] * y;
            y *= xf;
        }
        for (int n = l - 1; n >= 0; n--)
        {
            int nn = l - n;
            for (int t = 0; t <= nn; t++)
                for (int u = 0; u <= nn - t; u++)
                    for (int v = 0; v <= nn - t - u; v++)
                    {
                        double &o = buf[B(n, t, u, v)];
                        if (t > 0)
                        {
                            o = xyz[0] * buf[B(n + 1, t - 1, u, v)];
                            if (t > 1)
                                o += (t - 1) * buf[B(n + 1, t - 2, u, v)];
                        }
                        else if (u > 0)
                        {
                            o = xyz[1] * buf[B(n + 1, t, u - 1, v)];
                            if (u > 1)
                                o += (u - 1) * buf[B(n + 1, t, u - 2, v)];
                        }
                        else if (v > 0)
                        {
                            o = xyz[2] * buf[B(n + 1, t, u, v - 1)];
                            if (v > 1)
                                o += (v - 1) * buf[B(n + 1, t, u, v - 2)];
                        }
                    }
        }
        std::vector<double> r(static_cast<size_t>(d) * d * d, 0.0);
        for (int t = 0; t <= l; t++)
            for (int u = 0; u <= l - t; u++)
                for (int v = 0; v <= l - t - u; v++)
                    r[(static_cast<size_t>(t) * d + u) * d + v] = buf[B(0, t, u, v)];
        return r;
    }

    // ecoeffs.cpp:11-31
    std::vector<double> ecoeffsRecurrence1(double one_o_2a, int l)
    {
        const int d = l + 1;
        std::vector<double> e(static_cast<size_t>(d) * d, 0.0);
        auto E = [&](int i, int t) -> double & { return e[i * d + t]; };
        E(0, 0) = 1;
        for (int i = 1; i <= l; i++)
        {
            if (i % 2 == 0)
                E(i, 0) = E(i - 1, 1);
            for (int t = 1; t < i; t++)
                if ((t + i) % 2 == 0)
                    E(i, t) = one_o_2a * E(i - 1, t - 1) + (t + 1) * E(i - 1, t + 1);
            E(i, i) = one_o_2a * E(i - 1, i - 1);
        }
        return e;
    }

    // ecoeffs.cpp:33-68
    std::vector<double> ecoeffsRecurrence2(double a, double b, int la, int lb, double PA,
                                           double PB, double Kab)
    {
        const double p = a + b;
        const double o2p = 1.0 / (2 * p);
        const int d1 = lb + 1, d2 = la + lb + 1;
        std::vector<double> e(static_cast<size_t>(la + 1) * d1 * d2, 0.0);
        auto E = [&](int i, int j, int t) -> double & { return e[(i * d1 + j) * d2 + t]; };
        E(0, 0, 0) = Kab;
        for (int i = 1; i <= la; i++)
        {
            E(i, 0, 0) = PA * E(i - 1, 0, 0) + E(i - 1, 0, 1);
            for (int t = 1; t < i; t++)
                E(i, 0, t) = o2p * E(i - 1, 0, t - 1) + PA * E(i - 1, 0, t) +
                             (t + 1) * E(i - 1, 0, t + 1);
            E(i, 0, i) = o2p * E(i - 1, 0, i - 1) + PA * E(i - 1, 0, i);
        }
        for (int j = 1; j <= lb; j++)
            for (int i = 0; i <= la; i++)
            {
                E(i, j, 0) = PB * E(i, j - 1, 0) + E(i, j - 1, 1);
                for (int t = 1; t < i + j; t++)
                    E(i, j, t) = o2p * E(i, j - 1, t - 1) + PB * E(i, j - 1, t) +
                                 (t + 1) * E(i, j - 1, t + 1);
                E(i, j, i + j) = o2p * E(i, j - 1, i + j - 1) + PB * E(i, j - 1, i + j);
            }
        return e;
    }

    // ecoeffs.cpp:70-111
    std::vector<double> ecoeffsRecurrence2_n1(double a, double b, int la, int lb, double A,
                                              double B, const std::vector<double> &e0v)
    {
        const int d1 = lb + 1, d2 = la + lb + 1;
        std::vector<double> e1v(static_cast<size_t>(la + 1) * d1 * d2, 0.0);
        auto E0 = [&](int i, int j, int t) -> double { return e0v[(i * d1 + j) * d2 + t]; };
        auto E1 = [&](int i, int j, int t) -> double & { return e1v[(i * d1 + j) * d2 + t]; };
        const double p = a + b;
        const double o2p = 1.0 / (2 * p);
        const double R = A - B;
        E1(0, 0, 0) = -2 * (a * b) * R * E0(0, 0, 0) / p;
        for (int i = 1; i <= la; i++)
        {
            E1(i, 0, 0) = -(b / p) * (R * E1(i - 1, 0, 0) + E0(i - 1, 0, 0)) + E1(i - 1, 0, 1);
            for (int t = 1; t < i; t++)
                E1(i, 0, t) = o2p * E1(i - 1, 0, t - 1) -
                              (b / p) * (R * E1(i - 1, 0, t) + E0(i - 1, 0, t)) +
                              (t + 1) * E1(i - 1, 0, t + 1);
            E1(i, 0, i) = o2p * E1(i - 1, 0, i - 1) -
                          (b / p) * (R * E1(i - 1, 0, i) + E0(i - 1, 0, i));
        }
        for (int j = 1; j <= lb; j++)
            for (int i = 0; i <= la; i++)
            {
                E1(i, j, 0) = (a / p) * (R * E1(i, j - 1, 0) + E0(i, j - 1, 0)) + E1(i, j - 1, 1);
                for (int t = 1; t < i + j; t++)
                    E1(i, j, t) = o2p * E1(i, j - 1, t - 1) +
                                  (a / p) * (R * E1(i, j - 1, t) + E0(i, j - 1, t)) +
                                  (t + 1) * E1(i, j - 1, t + 1);
                E1(i, j, i + j) = o2p * E1(i, j - 1, i + j - 1) +
                                  (a / p) * (R * E1(i, j - 1, i + j) + E0(i, j - 1, i + j));
            }
        return e1v;
    }

    // ecoeffs.cpp:157-185
    std::array<std::vector<double>, 3> ecoeffsPrimitivePair(double a, double b, int la, int lb,
                                                            const double *xa, const double *xb)
    {
        const double p = a + b;
        const double mu = a * b / p;
        std::array<std::vector<double>, 3> r;
        for (int i = 0; i < 3; i++)
        {
            double P = (a * xa[i] + b * xb[i]) / p;
            double K = std::exp(-mu * std::pow(xa[i] - xb[i], 2));
            r[i] = ecoeffsRecurrence2(a, b, la, lb, P - xa[i], P - xb[i], K);
        }
        return r;
    }

    // ecoeffs.cpp:187-197
    std::array<std::vector<double>, 3> ecoeffsPrimitivePair_n1(
            double a, double b, int la, int lb, const double *xa, const double *xb,
            const std::array<std::vector<double>, 3> &e0)
    {
        return {ecoeffsRecurrence2_n1(a, b, la, lb, xa[0], xb[0], e0[0]),
                ecoeffsRecurrence2_n1(a, b, la, lb, xa[1], xb[1], e0[1]),
                ecoeffsRecurrence2_n1(a, b, la, lb, xa[2], xb[2], e0[2])};
    }

    // ecoeffs.cpp:222-251: Cartesian x Cartesian overlap factors E^{ii'}_0 E^{jj'}_0 E^{kk'}_0
    std::vector<std::vector<double>> ecoeffsShell(int l, const std::vector<double> &exps)
    {
        const int dim = nCart(l);
        const size_t K = exps.size();
        const double xyz0[3] = {0, 0, 0};
        auto ce = cartExps(l);
        const int d1 = l + 1, d2 = 2 * l + 1;
        std::vector<std::vector<double>> out(K * K, std::vector<double>(dim * dim, 0.0));
        for (size_t ia = 0, iab = 0; ia < K; ia++)
            for (size_t ib = 0; ib < K; ib++, iab++)
            {
                auto E = ecoeffsPrimitivePair(exps[ia], exps[ib], l, l, xyz0, xyz0);
                for (int mu = 0; mu < dim; mu++)
                    for (int nu = 0; nu < dim; nu++)
                    {
                        auto [i, j, k] = ce[mu];
                        auto [i2, j2, k2] = ce[nu];
                        out[iab][mu * dim + nu] = E[0][(i * d1 + i2) * d2] *
                                                  E[1][(j * d1 + j2) * d2] *
                                                  E[2][(k * d1 + k2) * d2];
                    }
            }
        return out;
    }
}
