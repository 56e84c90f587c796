// Oracle (test infrastructure only): shells, normalisation, pair data and SHARK E matrices,
// restated from the reference.
#include "oracle.hpp"

#include <cmath>
#include <map>
#include <stdexcept>

namespace orc
{
    // spherical_trafo.cpp:692-712 (Cartesian -> spherical on both indices)
    static std::vector<double> trafo2Spherical(int la, int lb, const std::vector<double> &cart)
    {
        int nca = nCart(la), ncb = nCart(lb), nsa = nSph(la), nsb = nSph(lb);
        auto ta = sphericalTrafo(la), tb = sphericalTrafo(lb);
        std::vector<double> cs(static_cast<size_t>(nca) * nsb, 0.0);
        for (int ia = 0; ia < nca; ia++)
            for (auto &[is, ic, v] : tb)
                cs[ia * nsb + is] += v * cart[ia * ncb + ic];
        std::vector<double> out(static_cast<size_t>(nsa) * nsb, 0.0);
        for (int ib = 0; ib < nsb; ib++)
            for (auto &[is, ic, v] : ta)
                out[is * nsb + ib] += v * cs[ic * nsb + ib];
        return out;
    }

    // shell.cpp:11-46
    std::vector<double> calcShellNorms(int l, const std::vector<double> &coeffs,
                                       const std::vector<double> &exps,
                                       const std::vector<double> &pn)
    {
        size_t K = exps.size();
        auto ec = ecoeffsShell(l, exps);
        int dc = nCart(l);
        std::vector<double> cart(static_cast<size_t>(dc) * dc, 0.0);
        for (size_t ia = 0, iab = 0; ia < K; ia++)
            for (size_t ib = 0; ib < K; ib++, iab++)
            {
                double a = exps[ia], b = exps[ib];
                double da = coeffs[ia] * pn[ia], db = coeffs[ib] * pn[ib];
                double p = a + b;
                double fac = da * db * std::pow(M_PI / p, 1.5);
                for (int munu = 0; munu < dc * dc; munu++)
                    cart[munu] += fac * ec[iab][munu];
            }
        auto sph = trafo2Spherical(l, l, cart);
        int ds = nSph(l);
        std::vector<double> norms(ds);
        for (int i = 0; i < ds; i++)
            norms[i] = 1.0 / std::sqrt(sph[i * ds + i]);
        return norms;
    }

    // shell.cpp:89-135 (atoms in input order, shells in basis order; global AO offsets)
    static std::vector<Shell> constructShells(const std::vector<std::array<double, 3>> &coords,
                                              const std::vector<std::vector<BasisShell>> &atoms)
    {
        if (atoms.size() != coords.size())
            throw std::runtime_error("constructShells(): atoms/coordinates size mismatch");
        size_t idx = 0, ofs_sph = 0, ofs_cart = 0;
        std::vector<Shell> shells;
        for (size_t ia = 0; ia < atoms.size(); ia++)
            for (const auto &bs : atoms[ia])
            {
                if (bs.exps.size() != bs.coeffs.size())
                    throw std::runtime_error("constructShells(): exps/coeffs size mismatch");
                std::vector<double> pn(bs.exps.size());
                for (size_t i = 0; i < pn.size(); i++)
                    pn[i] = purePrimitiveNorm(bs.l, bs.exps[i]);
                Shell s;
                s.l = bs.l;
                s.z = 0;
                s.dim_cart = nCart(bs.l);
                s.dim_sph = nSph(bs.l);
                s.ofs_cart = ofs_cart;
                s.ofs_sph = ofs_sph;
                s.idx = idx;
                s.idx_atom = ia;
                s.xyz = coords[ia];
                s.exps = bs.exps;
                s.coeffs = bs.coeffs;
                s.norms = calcShellNorms(bs.l, bs.coeffs, bs.exps, pn);
                s.norms_prim = pn;
                shells.push_back(s);
                idx++;
                ofs_sph += s.dim_sph;
                ofs_cart += s.dim_cart;
            }
        return shells;
    }

    Structure makeStructure(const std::vector<std::array<double, 3>> &coords,
                            const std::vector<std::vector<BasisShell>> &atoms,
                            const std::vector<std::vector<BasisShell>> *aux)
    {
        Structure s;
        s.natoms = coords.size();
        s.shells = constructShells(coords, atoms);
        for (auto &sh : s.shells)
        {
            s.nbf += sh.dim_sph;
            s.max_l = std::max(s.max_l, sh.l);
        }
        if (aux)
        {
            s.shells_aux = constructShells(coords, *aux);
            for (auto &sh : s.shells_aux)
            {
                s.nbf_aux += sh.dim_sph;
                s.max_l_aux = std::max(s.max_l_aux, sh.l);
            }
        }
        return s;
    }

    // structure.cpp:323-339 (getShellsL: std::map::at, throws for an absent l)
    std::vector<Shell> Structure::shellsL(int l, bool aux) const
    {
        std::vector<Shell> r;
        for (auto &sh : (aux ? shells_aux : shells))
            if (sh.l == l)
                r.push_back(sh);
        if (r.empty())
            throw std::out_of_range("Structure: no shells with the requested angular momentum");
        return r;
    }

    // shell_pair_data.cpp:73-202
    ShellPairData::ShellPairData(bool use_symm, int la_, int lb_, const std::vector<Shell> &sa,
                                 const std::vector<Shell> &sb, double thrs)
        : uses_symm(use_symm), la(la_), lb(lb_)
    {
        int nsa = nSph(la), nsb = nSph(lb);
        int n_sph_e = nsa * nsb * nHerm(la + lb);
        size_t ofs_prim = 0, ofs_norm = 0, ofs_e = 0, ofs_e1 = 0;
        for (size_t is = 0; is < sa.size(); is++)
        {
            size_t bound = (uses_symm && la == lb) ? is + 1 : sb.size();
            for (size_t js = 0; js < bound; js++)
            {
                const Shell &A = sa[is], &B = sb[js];
                if (A.l != la || B.l != lb)
                    throw std::runtime_error("ShellPairData(): angular momentum mismatch");
                double RAB2 = std::pow(A.xyz[0] - B.xyz[0], 2) + std::pow(A.xyz[1] - B.xyz[1], 2) +
                              std::pow(A.xyz[2] - B.xyz[2], 2);
                size_t nsurv = 0;
                for (size_t ia = 0; ia < A.exps.size(); ia++)
                    for (size_t ib = 0; ib < B.exps.size(); ib++)
                    {
                        double a = A.exps[ia], b = B.exps[ib];
                        double da = A.norms_prim[ia] * A.coeffs[ia];
                        double db = B.norms_prim[ib] * B.coeffs[ib];
                        double mu = a * b / (a + b);
                        double Kab = std::exp(-mu * RAB2);
                        if (std::fabs(da * db * Kab) >= thrs)
                        {
                            exps.push_back(a);
                            exps.push_back(b);
                            coeffs.push_back(da);
                            coeffs.push_back(db);
                            nsurv++;
                        }
                    }
                if (nsurv == 0)
                    continue;
                for (int i = 0; i < 3; i++)
                    coords.push_back(A.xyz[i]);
                for (int i = 0; i < 3; i++)
                    coords.push_back(B.xyz[i]);
                nrs_ppairs.push_back(nsurv);
                offsets_prims.push_back(ofs_prim);
                offsets_sph.push_back(A.ofs_sph);
                offsets_sph.push_back(B.ofs_sph);
                offsets_norms.push_back(ofs_norm);
                for (int m = 0; m < nsa; m++)
                    norms.push_back(A.norms[m]);
                ofs_norm += nsa;
                offsets_norms.push_back(ofs_norm);
                for (int m = 0; m < nsb; m++)
                    norms.push_back(B.norms[m]);
                ofs_norm += nsb;
                offsets_ecoeffs.push_back(ofs_e);
                offsets_ecoeffs_d1.push_back(ofs_e1);
                atomic_idxs.push_back(A.idx_atom);
                atomic_idxs.push_back(B.idx_atom);
                shell_idxs.push_back(A.idx);
                shell_idxs.push_back(B.idx);
                n_pairs++;
                n_ppairs += nsurv;
                ofs_prim += 2 * nsurv;
                ofs_e += nsurv * n_sph_e;
                ofs_e1 += 3 * nsurv * n_sph_e;
            }
        }
    }

    // shell_pair_data.cpp:7-71
    ShellData::ShellData(int l_, const std::vector<Shell> &shells) : l(l_)
    {
        int ns = nSph(l), ne = ns * nHerm(l);
        size_t oc = 0, oe = 0, on = 0;
        for (const Shell &s : shells)
        {
            if (s.l != l)
                throw std::runtime_error("ShellData(): angular momentum mismatch");
            for (size_t i = 0; i < s.coeffs.size(); i++)
            {
                coeffs.push_back(s.coeffs[i] * s.norms_prim[i]);
                exps.push_back(s.exps[i]);
            }
            for (int i = 0; i < 3; i++)
                coords.push_back(s.xyz[i]);
            for (int i = 0; i < ns; i++)
                norms.push_back(s.norms[i]);
            atomic_idxs.push_back(s.idx_atom);
            cdepths.push_back(s.coeffs.size());
            coffsets.push_back(oc);
            offsets_ecoeffs.push_back(oe);
            offsets_norms.push_back(on);
            offsets_sph.push_back(s.ofs_sph);
            shell_idxs.push_back(s.idx);
            oc += s.coeffs.size();
            oe += s.coeffs.size() * ne;
            on += ns;
            n_shells++;
        }
        n_prims = oc;
    }

    // utils.cpp:49-71
    std::vector<std::pair<int, int>> lPairsSymm(int l)
    {
        std::vector<std::pair<int, int>> r;
        for (int la = 0; la <= l; la++)
            for (int lb = 0; lb <= la; lb++)
                r.emplace_back(la, lb);
        return r;
    }

    std::vector<std::pair<int, int>> lPairsNoSymm(int l)
    {
        std::vector<std::pair<int, int>> r;
        for (int la = 0; la <= l; la++)
            for (int lb = 0; lb <= l; lb++)
                r.emplace_back(la, lb);
        return r;
    }

    // shell_pair_data.cpp:294-309
    std::vector<ShellPairData> shellPairData(bool use_symm, const Structure &s)
    {
        auto lp = use_symm ? lPairsSymm(s.max_l) : lPairsNoSymm(s.max_l);
        std::vector<ShellPairData> r;
        for (auto [la, lb] : lp)
            r.emplace_back(use_symm, la, lb, s.shellsL(la, false), s.shellsL(lb, false));
        return r;
    }

    // shell_pair_data.cpp:280-292
    std::vector<ShellData> shellDataAux(const Structure &s)
    {
        if (s.shells_aux.empty())
            throw std::runtime_error("shellDataAux(): RI approximation is not enabled!");
        std::vector<ShellData> r;
        for (int l = 0; l <= s.max_l_aux; l++)
            r.emplace_back(l, s.shellsL(l, true));
        return r;
    }

    static std::vector<int> tuvPositions(int l)
    {
        // getHermiteGaussianPositions (utils.cpp:92-104), dense (l+1)^3
        std::vector<int> pos(static_cast<size_t>(l + 1) * (l + 1) * (l + 1), -1);
        int tuv = 0;
        for (int n = 0; n <= l; n++)
            for (int t = n; t >= 0; t--)
                for (int u = n - t; u >= 0; u--, tuv++)
                    pos[(t * (l + 1) + u) * (l + 1) + (n - t - u)] = tuv;
        return pos;
    }

    // ecoeffs.cpp:253-312
    std::vector<double> ecoeffsSHARK(const ShellData &sh, bool transpose)
    {
        int l = sh.l, ns = nSph(l), nh = nHerm(l);
        auto ijk = cartExps(l);
        auto pos = tuvPositions(l);
        auto st = sphericalTrafo(l);
        std::vector<double> e(sh.n_prims * ns * nh, 0.0);
        for (size_t is = 0; is < sh.n_shells; is++)
        {
            size_t K = sh.cdepths[is], co = sh.coffsets[is], oe = sh.offsets_ecoeffs[is];
            const double *norms = &sh.norms[sh.offsets_norms[is]];
            for (size_t ia = 0; ia < K; ia++)
            {
                double a = sh.exps[co + ia], d = sh.coeffs[co + ia];
                auto E1 = ecoeffsRecurrence1(1.0 / (2 * a), l);
                size_t ofs = oe + ia * ns * nh;
                for (auto &[mu, mu_, val] : st)
                {
                    auto [i, j, k] = ijk[mu_];
                    for (int t = 0; t <= i; t++)
                        for (int u = 0; u <= j; u++)
                            for (int v = 0; v <= k; v++)
                            {
                                double ec = norms[mu] * d * E1[i * (l + 1) + t] *
                                            E1[j * (l + 1) + u] * E1[k * (l + 1) + v] * val;
                                int tuv = pos[(t * (l + 1) + u) * (l + 1) + v];
                                size_t idx = transpose ? ofs + tuv * ns + mu : ofs + mu * nh + tuv;
                                e[idx] += ec;
                            }
                }
            }
        }
        return e;
    }

    // Shared body of ecoeffsSHARK(ShellPairData) (ecoeffs.cpp:314-403) and ecoeffsD1SHARK
    // (ecoeffs.cpp:405-517).  nder = 0: plain E; nder = 1: the three E' blocks (d/dx, d/dy,
    // d/dz of the A-B relative coordinate).
    static std::vector<double> sharkPair(const ShellPairData &sp, bool transpose, bool d1)
    {
        int la = sp.la, lb = sp.lb, lab = la + lb;
        int ncb = nCart(lb), nsa = nSph(la), nsb = nSph(lb), nsab = nsa * nsb;
        int nh = nHerm(lab), ne = nsab * nh;
        auto sta = sphericalTrafo(la), stb = sphericalTrafo(lb);
        auto cea = cartExps(la), ceb = cartExps(lb);
        auto pos = tuvPositions(lab);
        int nblk = d1 ? 3 : 1;
        std::vector<double> e(static_cast<size_t>(nblk) * ne * sp.n_ppairs, 0.0);
        const int d1b = lb + 1, d2 = lab + 1;
        auto EI = [&](const std::vector<double> &v, int i, int j, int t)
        { return v[(i * d1b + j) * d2 + t]; };
        for (size_t ip = 0; ip < sp.n_pairs; ip++)
        {
            const double *xa = &sp.coords[6 * ip], *xb = &sp.coords[6 * ip + 3];
            const double *na = &sp.norms[sp.offsets_norms[2 * ip]];
            const double *nb = &sp.norms[sp.offsets_norms[2 * ip + 1]];
            size_t op = sp.offsets_prims[ip];
            const double *ex = &sp.exps[op], *cf = &sp.coeffs[op];
            size_t oe = d1 ? sp.offsets_ecoeffs_d1[ip] : sp.offsets_ecoeffs[ip];
            for (size_t iab = 0; iab < sp.nrs_ppairs[ip]; iab++)
            {
                double a = ex[2 * iab], b = ex[2 * iab + 1];
                auto E0 = ecoeffsPrimitivePair(a, b, la, lb, xa, xb);
                std::array<std::vector<double>, 3> E1;
                if (d1)
                    E1 = ecoeffsPrimitivePair_n1(a, b, la, lb, xa, xb, E0);
                std::vector<double> sc(static_cast<size_t>(nblk) * nsa * ncb * nh, 0.0);
                auto SC = [&](int blk, int mu, int nu_, int tuv) -> double &
                { return sc[((static_cast<size_t>(blk) * nsa + mu) * ncb + nu_) * nh + tuv]; };
                for (auto &[mu, mu_, val] : sta)
                    for (int nu_ = 0; nu_ < ncb; nu_++)
                    {
                        auto [i, j, k] = cea[mu_];
                        auto [i2, j2, k2] = ceb[nu_];
                        for (int t = 0; t <= i + i2; t++)
                            for (int u = 0; u <= j + j2; u++)
                                for (int v = 0; v <= k + k2; v++)
                                {
                                    int tuv = pos[(t * d2 + u) * d2 + v];
                                    if (!d1)
                                        SC(0, mu, nu_, tuv) += val * EI(E0[0], i, i2, t) *
                                                               EI(E0[1], j, j2, u) *
                                                               EI(E0[2], k, k2, v);
                                    else
                                    {
                                        SC(0, mu, nu_, tuv) += val * EI(E1[0], i, i2, t) *
                                                               EI(E0[1], j, j2, u) *
                                                               EI(E0[2], k, k2, v);
                                        SC(1, mu, nu_, tuv) += val * EI(E0[0], i, i2, t) *
                                                               EI(E1[1], j, j2, u) *
                                                               EI(E0[2], k, k2, v);
                                        SC(2, mu, nu_, tuv) += val * EI(E0[0], i, i2, t) *
                                                               EI(E0[1], j, j2, u) *
                                                               EI(E1[2], k, k2, v);
                                    }
                                }
                    }
                double dadb = cf[2 * iab] * cf[2 * iab + 1];
                for (auto &[nu, nu_, val] : stb)
                    for (int mu = 0; mu < nsa; mu++)
                        for (int tuv = 0; tuv < nh; tuv++)
                        {
                            int munu = mu * nsb + nu;
                            for (int blk = 0; blk < nblk; blk++)
                            {
                                size_t ofs = oe + (static_cast<size_t>(nblk) * iab + blk) * ne;
                                size_t idx = transpose ? ofs + tuv * nsab + munu
                                                       : ofs + munu * nh + tuv;
                                if (!d1)
                                    e[idx] += na[mu] * nb[nu] * dadb * val * SC(0, mu, nu_, tuv);
                                else
                                    e[idx] += (na[mu] * nb[nu]) * dadb * val *
                                              SC(blk, mu, nu_, tuv);
                            }
                        }
            }
        }
        return e;
    }

    std::vector<double> ecoeffsSHARK(const ShellPairData &sp, bool transpose)
    {
        return sharkPair(sp, transpose, false);
    }

    std::vector<double> ecoeffsD1SHARK(const ShellPairData &sp, bool transpose)
    {
        return sharkPair(sp, transpose, true);
    }
}
