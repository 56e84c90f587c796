// Oracle (test infrastructure only): the one-electron integrals of the SCF step around the ERI
// path, restated from the reference: overlap (oneelints.cpp:57-100, 169-197), kinetic energy
// (oneelints.cpp:199-276, 348-376) and the attraction to point charges / nuclei
// (oneelints.cpp:508-585, 1703-1744).  Cartesian per-axis Hermite coefficients, transformed to
// normalised spherical functions; each pair class la >= lb, both triangles filled.
#include "oracle.hpp"

#include <cmath>

namespace orc
{
    namespace
    {
        // ints_sph(mu, nu) = sum_ij T_a(mu, i) T_b(nu, j) ints_cart(i, j)  (trafo2Spherical,
        // spherical_trafo.cpp:9-690 tables)
        std::vector<double> toSpherical(int la, int lb, const std::vector<double> &cart)
        {
            const int ncb = nCart(lb), nsa = nSph(la), nsb = nSph(lb);
            std::vector<double> half(static_cast<size_t>(nsa) * ncb, 0.0), out(static_cast<size_t>(nsa) * nsb, 0.0);
            for (auto [mu, i, c] : sphericalTrafo(la))
                for (int j = 0; j < ncb; j++)
                    half[mu * ncb + j] += c * cart[i * ncb + j];
            for (auto [nu, j, c] : sphericalTrafo(lb))
                for (int mu = 0; mu < nsa; mu++)
                    out[mu * nsb + nu] += c * half[mu * ncb + j];
            return out;
        }

        // E^{ij}_t of one axis from the (la+1)(lb+1)(la+lb+1) array of ecoeffsRecurrence2
        inline double E(const std::vector<double> &e, int lb, int lab, int i, int j, int t)
        {
            return e[(static_cast<size_t>(i) * (lb + 1) + j) * (lab + 1) + t];
        }

        // Per pair of every class: kernel(pair data, pair index) -> (Sa, Sb) spherical block,
        // placed at both (a, b) and (b, a) of the dense nbf x nbf matrix
        template <class F>
        std::vector<double> dense(const Structure &s, F kernel)
        {
            const size_t n = s.nbf;
            std::vector<double> out(n * n, 0.0);
            for (const auto &sp : shellPairData(true, s))
                for (size_t p = 0; p < sp.n_pairs; p++)
                {
                    std::vector<double> blk = kernel(sp, p);
                    const size_t oa = sp.offsets_sph[2 * p], ob = sp.offsets_sph[2 * p + 1];
                    const int sa = nSph(sp.la), sb = nSph(sp.lb);
                    const double *na = &sp.norms[sp.offsets_norms[2 * p]];
                    const double *nb = &sp.norms[sp.offsets_norms[2 * p + 1]];
                    for (int mu = 0; mu < sa; mu++)
                        for (int nu = 0; nu < sb; nu++)
                        {
                            const double v = blk[mu * sb + nu] * na[mu] * nb[nu];
                            out[(oa + mu) * n + ob + nu] = v;
                            out[(ob + nu) * n + oa + mu] = v;
                        }
                }
            return out;
        }
    }

    std::vector<double> overlap(const Structure &s)
    {
        return dense(s, [](const ShellPairData &sp, size_t ip) {
            const int la = sp.la, lb = sp.lb, lab = la + lb;
            const auto ca = cartExps(la), cb = cartExps(lb);
            const double *ex = &sp.exps[sp.offsets_prims[ip]];
            const double *co = &sp.coeffs[sp.offsets_prims[ip]];
            const double *xa = &sp.coords[6 * ip], *xb = &sp.coords[6 * ip + 3];
            std::vector<double> cart(static_cast<size_t>(nCart(la)) * nCart(lb), 0.0);
            for (size_t k = 0; k < sp.nrs_ppairs[ip]; k++)
            {
                const double a = ex[2 * k], b = ex[2 * k + 1], p = a + b;
                const double fac = co[2 * k] * co[2 * k + 1] * std::pow(M_PI / p, 1.5);
                auto [Ex, Ey, Ez] = ecoeffsPrimitivePair(a, b, la, lb, xa, xb);
                for (size_t mu = 0; mu < ca.size(); mu++)
                    for (size_t nu = 0; nu < cb.size(); nu++)
                    {
                        auto [i, j, kk] = ca[mu];
                        auto [i_, j_, k_] = cb[nu];
                        cart[mu * cb.size() + nu] += fac * E(Ex, lb, lab, i, i_, 0) * E(Ey, lb, lab, j, j_, 0) *
                                                     E(Ez, lb, lab, kk, k_, 0);
                    }
            }
            return toSpherical(la, lb, cart);
        });
    }

    std::vector<double> kineticEnergy(const Structure &s)
    {
        return dense(s, [](const ShellPairData &sp, size_t ip) {
            const int la = sp.la, lb = sp.lb, lb2 = lb + 2, lab2 = la + lb2;
            const auto ca = cartExps(la), cb = cartExps(lb);
            const double *ex = &sp.exps[sp.offsets_prims[ip]];
            const double *co = &sp.coeffs[sp.offsets_prims[ip]];
            const double *xa = &sp.coords[6 * ip], *xb = &sp.coords[6 * ip + 3];
            std::vector<double> cart(static_cast<size_t>(nCart(la)) * nCart(lb), 0.0);
            for (size_t k = 0; k < sp.nrs_ppairs[ip]; k++)
            {
                const double a = ex[2 * k], b = ex[2 * k + 1], p = a + b, b2 = b * b;
                const double fac = co[2 * k] * co[2 * k + 1] * std::pow(M_PI / p, 1.5);
                // E^{i,j'}_0 up to j' = lb + 2 (the second derivative raises b's exponent)
                auto [Ex, Ey, Ez] = ecoeffsPrimitivePair(a, b, la, lb2, xa, xb);
                auto e0 = [&](const std::vector<double> &e, int i, int j) { return E(e, lb2, lab2, i, j, 0); };
                // -1/2 d2/dx2 on b's factor x^j exp(-b x^2): b(2j+1) S_j - 2 b^2 S_{j+2} - j(j-1)/2 S_{j-2}
                auto t1 = [&](const std::vector<double> &e, int i, int j) {
                    double t = -2 * b2 * e0(e, i, j + 2) + b * (2 * j + 1) * e0(e, i, j);
                    if (j >= 2)
                        t -= 0.5 * j * (j - 1) * e0(e, i, j - 2);
                    return t;
                };
                for (size_t mu = 0; mu < ca.size(); mu++)
                    for (size_t nu = 0; nu < cb.size(); nu++)
                    {
                        auto [i, j, kk] = ca[mu];
                        auto [i_, j_, k_] = cb[nu];
                        const double sx = e0(Ex, i, i_), sy = e0(Ey, j, j_), sz = e0(Ez, kk, k_);
                        cart[mu * cb.size() + nu] += fac * (t1(Ex, i, i_) * sy * sz + sx * t1(Ey, j, j_) * sz +
                                                            sx * sy * t1(Ez, kk, k_));
                    }
            }
            return toSpherical(la, lb, cart);
        });
    }

    std::vector<double> externalCharges(const Structure &s, const std::vector<std::array<double, 4>> &charges)
    {
        return dense(s, [&](const ShellPairData &sp, size_t ip) {
            const int la = sp.la, lb = sp.lb, lab = la + lb;
            const auto ca = cartExps(la), cb = cartExps(lb);
            const double *ex = &sp.exps[sp.offsets_prims[ip]];
            const double *co = &sp.coeffs[sp.offsets_prims[ip]];
            const double *xa = &sp.coords[6 * ip], *xb = &sp.coords[6 * ip + 3];
            BoysGrid grid(lab);
            std::vector<double> cart(static_cast<size_t>(nCart(la)) * nCart(lb), 0.0), fnx(lab + 1);
            const size_t d = lab + 1;
            for (size_t k = 0; k < sp.nrs_ppairs[ip]; k++)
            {
                const double a = ex[2 * k], b = ex[2 * k + 1], p = a + b;
                const double fac = 2 * (M_PI / p) * co[2 * k] * co[2 * k + 1];
                auto [Ex, Ey, Ez] = ecoeffsPrimitivePair(a, b, la, lb, xa, xb);
                double P[3];
                for (int x = 0; x < 3; x++)
                    P[x] = (a * xa[x] + b * xb[x]) / p;
                std::vector<double> rsum(d * d * d, 0.0);
                for (const auto &c : charges)
                {
                    const double PC[3] = {P[0] - c[0], P[1] - c[1], P[2] - c[2]};
                    calcBoysF(lab, p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]), grid, fnx.data());
                    auto r = calcRInts3D(lab, p, PC, fnx.data());
                    for (size_t e = 0; e < rsum.size(); e++)
                        rsum[e] += c[3] * r[e];
                }
                for (size_t mu = 0; mu < ca.size(); mu++)
                    for (size_t nu = 0; nu < cb.size(); nu++)
                    {
                        auto [i, j, kk] = ca[mu];
                        auto [i_, j_, k_] = cb[nu];
                        double v = 0;
                        for (int t = 0; t <= i + i_; t++)
                            for (int u = 0; u <= j + j_; u++)
                                for (int w = 0; w <= kk + k_; w++)
                                    v += E(Ex, lb, lab, i, i_, t) * E(Ey, lb, lab, j, j_, u) *
                                         E(Ez, lb, lab, kk, k_, w) * rsum[(t * d + u) * d + w];
                        cart[mu * cb.size() + nu] -= fac * v;   // the electron's charge is -1
                    }
            }
            return toSpherical(la, lb, cart);
        });
    }
}
