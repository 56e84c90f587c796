// Host-side set-up: normalised shells, real-solid-harmonic coefficients, the Boys table and
// the screened shell-pair classes.  Runs once per lbx_create.
#include "lbx_internal.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>

namespace lbx
{
    namespace
    {
        double factorial(int n)
        {
            double r = 1;
            for (int i = 2; i <= n; i++)
                r *= i;
            return r;
        }

        double choose(int n, int k) { return (k < 0 || k > n) ? 0.0 : factorial(n) / (factorial(k) * factorial(n - k)); }

        // (n-1)!! for even n >= 0 (n = 0 -> 1)
        double oddDoubleFactorialBelow(int n)
        {
            double r = 1;
            for (int i = n - 1; i > 1; i -= 2)
                r *= i;
            return r;
        }

        // Integral of x^n exp(-p x^2) over the real line.
        double gauss1d(int n, double p)
        {
            if (n % 2)
                return 0.0;
            return oddDoubleFactorialBelow(n) / std::pow(2 * p, n / 2) * std::sqrt(M_PI / p);
        }

        std::vector<std::array<int, 3>> cartesians(int l)
        {
            std::vector<std::array<int, 3>> r;
            for (int i = l; i >= 0; i--)
                for (int j = l - i; j >= 0; j--)
                    r.push_back({i, j, l - i - j});
            return r;
        }

        std::vector<SphEntry> makeSph(int l)
        {
            // Real solid harmonics S_lm as polynomials in x, y, z; m order 0, +1, -1, +2, -2, ...
            // Closed form of Helgaker, Jorgensen, Olsen, Molecular Electronic-Structure Theory,
            // eqs. (6.4.47)-(6.4.50).  Coefficients are rounded to 14 decimals, the precision of
            // the reference's tabulation (spherical_trafo.cpp:9-690), so both use one table.
            std::vector<SphEntry> out;
            const int nc = nCart(l);
            for (int slot = 0; slot < 2 * l + 1; slot++)
            {
                const int m = slot == 0 ? 0 : (slot % 2 ? (slot + 1) / 2 : -(slot / 2));
                const int am = std::abs(m);
                const double pref = std::sqrt(2.0 * factorial(l + am) * factorial(l - am) /
                                              (m == 0 ? 2.0 : 1.0)) /
                                    (std::ldexp(1.0, am) * factorial(l));
                std::vector<double> c(nc, 0.0);
                // v runs over half-integers for m < 0: use w = 2v, odd for m < 0, even for m >= 0
                const int w0 = m < 0 ? 1 : 0;
                for (int t = 0; 2 * t <= l - am; t++)
                    for (int u = 0; u <= t; u++)
                        for (int w = w0; w <= am; w += 2)
                        {
                            const int sign = ((t + (w - w0) / 2) & 1) ? -1 : 1;
                            const double coef = sign * std::ldexp(1.0, -2 * t) * choose(l, t) *
                                                choose(l - t, am + t) * choose(t, u) * choose(am, w);
                            const int ey = 2 * u + w;
                            const int ez = l - 2 * t - am;
                            c[cartIdx(ey, ez)] += pref * coef;
                        }
                for (int k = 0; k < nc; k++)
                {
                    if (c[k] == 0.0)
                        continue;
                    char buf[64];
                    std::snprintf(buf, sizeof buf, "%.14f", c[k]);
                    const double v = std::strtod(buf, nullptr);
                    if (v != 0.0)
                        out.push_back({slot, k, v});
                }
            }
            return out;
        }
    }

    const std::vector<SphEntry> &sphTable(int l)
    {
        static std::vector<std::vector<SphEntry>> tables = []
        {
            std::vector<std::vector<SphEntry>> t;
            for (int i = 0; i <= kMaxLAux; i++)
                t.push_back(makeSph(i));
            return t;
        }();
        if (l < 0 || l > kMaxLAux)
            throw Error(LBX_ERR_INVALID, "angular momentum out of range");
        return tables[l];
    }

    std::vector<HShell> buildShells(const lbx_basis &b, size_t &nbf, int &max_l, int max_l_allowed)
    {
        if (b.natoms <= 0 || b.nshells < 0 || !b.coords ||
            (b.nshells > 0 && (!b.shell_atom || !b.shell_l || !b.shell_nprim || !b.exps ||
                               !b.coeffs)))
            throw Error(LBX_ERR_INVALID, "lbx_basis: null array or bad size");
        std::vector<HShell> shells;
        size_t off = 0, ao = 0;
        max_l = -1;
        for (int i = 0; i < b.nshells; i++)
        {
            HShell s;
            s.l = b.shell_l[i];
            s.atom = b.shell_atom[i];
            if (s.l < 0 || s.l > max_l_allowed)
                throw Error(LBX_ERR_INVALID, "shell angular momentum " + std::to_string(s.l) +
                                                     " outside the supported range 0.." +
                                                     std::to_string(max_l_allowed));
            if (s.atom < 0 || s.atom >= b.natoms || (i > 0 && s.atom < b.shell_atom[i - 1]))
                throw Error(LBX_ERR_INVALID, "shells must be grouped by atom in atom order");
            int K = b.shell_nprim[i];
            if (K <= 0)
                throw Error(LBX_ERR_INVALID, "shell with no primitives");
            s.idx = i;
            s.ao = ao;
            for (int k = 0; k < 3; k++)
                s.xyz[k] = b.coords[3 * s.atom + k];
            s.exps.assign(b.exps + off, b.exps + off + K);
            s.coeffs.assign(b.coeffs + off, b.coeffs + off + K);
            off += K;
            // primitive norms: sqrt((2a/pi)^{3/2} (4a)^l / (2l-1)!!)   (utils.cpp:6-10)
            for (double a : s.exps)
            {
                if (!(a > 0))
                    throw Error(LBX_ERR_INVALID, "non-positive exponent");
                s.pnorms.push_back(std::sqrt(std::pow(2 * a / M_PI, 1.5) * std::pow(4 * a, s.l) /
                                             oddDoubleFactorialBelow(2 * s.l)));
            }
            // contracted norms: 1/sqrt(S_mm) of the spherical self-overlap (shell.cpp:11-46)
            auto cart = cartesians(s.l);
            const int nc = nCart(s.l), ns = nSph(s.l);
            std::vector<double> Sc(nc * nc, 0.0);
            for (int x = 0; x < K; x++)
                for (int y = 0; y < K; y++)
                {
                    double p = s.exps[x] + s.exps[y];
                    double d = s.coeffs[x] * s.pnorms[x] * s.coeffs[y] * s.pnorms[y];
                    for (int m = 0; m < nc; m++)
                        for (int n = 0; n < nc; n++)
                            Sc[m * nc + n] += d * gauss1d(cart[m][0] + cart[n][0], p) *
                                              gauss1d(cart[m][1] + cart[n][1], p) *
                                              gauss1d(cart[m][2] + cart[n][2], p);
                }
            const auto &T = sphTable(s.l);
            std::vector<double> diag(ns, 0.0);
            for (const auto &e1 : T)
                for (const auto &e2 : T)
                    if (e1.sph == e2.sph)
                        diag[e1.sph] += e1.val * e2.val * Sc[e1.cart * nc + e2.cart];
            for (int m = 0; m < ns; m++)
                s.norms.push_back(1.0 / std::sqrt(diag[m]));
            max_l = std::max(max_l, s.l);
            ao += ns;
            shells.push_back(std::move(s));
        }
        nbf = ao;
        return shells;
    }

    // Boys function grid: F_n(x_i) for n = 0..kBoysCols-1 on the points x_i = 0.01 i.  Top column
    // by its power series (HJO eq. 9.8.11, terms until < 1e-16), lower columns by downward
    // recursion (HJO eq. 9.8.14) -- the construction of boys_function.cpp:37-69.  The grid
    // abscissae are accumulated by repeated addition of 0.01 as there, so that the tabulated
    // values (and hence the integrals) agree with the reference's to rounding level.
    std::vector<double> boysTable()
    {
        std::vector<double> t(static_cast<size_t>(kBoysRows) * kBoysCols);
        const int top = kBoysCols - 1;
        double x = 0.0;
        for (int i = 0; i < kBoysRows; i++, x += kBoysStep)
        {
            double term = 1.0 / (2.0 * top + 1.0), sum = term;
            for (int k = 1;; k++)
            {
                term *= x / (top + 0.5 + k);
                sum += term;
                if (std::fabs(term) < 1e-16)
                    break;
            }
            const double ex = std::exp(-x);
            double *row = &t[static_cast<size_t>(i) * kBoysCols];
            row[top] = ex * sum;
            for (int n = top - 1; n >= 0; n--)
                row[n] = (2.0 * x * row[n + 1] + ex) / (2 * n + 1);
        }
        return t;
    }

    // Shell pairs of every class (la >= lb) with primitive-pair screening
    // |N_a d_a N_b d_b exp(-ab/(a+b) R_AB^2)| >= 1e-15, pair order of
    // shell_pair_data.cpp:73-202 (shells a of class la in basis order; b of class lb, b <= a
    // within the same class).  Single-shell classes hold each auxiliary shell once.
    void buildPairClasses(Context &c)
    {
        std::map<int, std::vector<const HShell *>> byl;
        for (const auto &s : c.shells)
            byl[s.l].push_back(&s);
        // molecule of a shell, and per l the range of byl[l] holding each molecule's shells
        // (shells follow atom order, atoms are grouped by molecule)
        const bool batched = !c.atom_mol.empty();
        auto molOf = [&](const HShell &s) { return batched ? c.atom_mol[s.atom] : 0; };
        std::map<int, std::vector<size_t>> mol_lo, mol_hi;
        for (auto &kv : byl)
        {
            auto &lo = mol_lo[kv.first], &hi = mol_hi[kv.first];
            lo.assign(c.nmol, 0);
            hi.assign(c.nmol, 0);
            for (size_t i = kv.second.size(); i-- > 0;)
                lo[molOf(*kv.second[i])] = i;
            for (size_t i = 0; i < kv.second.size(); i++)
                hi[molOf(*kv.second[i])] = i + 1;
        }
        c.cls.assign(pairClassIdx(kMaxLMain, kMaxLMain) + 1, PairClass{});
        for (int la = 0; la <= c.max_l; la++)
            for (int lb = 0; lb <= la; lb++)
            {
                PairClass &pc = c.cls[pairClassIdx(la, lb)];
                pc.la = la;
                pc.lb = lb;
                const auto &A = byl[la];
                const auto &B = byl[lb];
                int nprim = 0;
                for (size_t i = 0; i < A.size(); i++)
                {
                    size_t jbeg = 0, jend = (la == lb) ? i + 1 : B.size();
                    if (batched)
                    {
                        const int m = molOf(*A[i]);
                        jbeg = mol_lo[lb][m];
                        jend = (la == lb) ? i + 1 : mol_hi[lb][m];
                        if (mol_hi[lb][m] == 0)
                            jend = jbeg;   // no shell of this l in the molecule
                    }
                    for (size_t j = jbeg; j < jend; j++)
                    {
                        const HShell &sa = *A[i], &sb = *B[j];
                        double r2 = 0;
                        for (int k = 0; k < 3; k++)
                            r2 += (sa.xyz[k] - sb.xyz[k]) * (sa.xyz[k] - sb.xyz[k]);
                        int surv = 0;
                        for (size_t x = 0; x < sa.exps.size(); x++)
                            for (size_t y = 0; y < sb.exps.size(); y++)
                            {
                                double a = sa.exps[x], b = sb.exps[y];
                                double d = (sa.pnorms[x] * sa.coeffs[x]) * (sb.pnorms[y] * sb.coeffs[y]);
                                if (std::fabs(d * std::exp(-(a * b / (a + b)) * r2)) >= 1e-15)
                                {
                                    pc.prim_x.push_back(static_cast<int>(x));
                                    pc.prim_y.push_back(static_cast<int>(y));
                                    surv++;
                                }
                            }
                        if (surv == 0)
                            continue;
                        pc.ao_a.push_back(static_cast<int>(sa.ao));
                        pc.ao_b.push_back(static_cast<int>(sb.ao));
                        pc.sh_a.push_back(static_cast<int>(sa.idx));
                        pc.sh_b.push_back(static_cast<int>(sb.idx));
                        pc.atom_a.push_back(sa.atom);
                        pc.atom_b.push_back(sb.atom);
                        pc.mol.push_back(molOf(sa));
                        pc.poff.push_back(nprim);
                        pc.pcnt.push_back(surv);
                        pc.kmax = std::max(pc.kmax, surv);
                        nprim += surv;
                    }
                }
                pc.n_pairs = static_cast<int>(pc.pcnt.size());
                pc.n_ppairs = nprim;
            }
        c.aux.assign(kMaxLAux + 1, PairClass{});
        for (int l = 0; l <= c.max_l_aux; l++)
        {
            PairClass &pc = c.aux[l];
            pc.la = l;
            pc.lb = 0;
            pc.single = true;
            int nprim = 0;
            for (const auto &s : c.shells_aux)
            {
                if (s.l != l)
                    continue;
                pc.ao_a.push_back(static_cast<int>(s.ao));
                pc.ao_b.push_back(0);
                pc.sh_a.push_back(static_cast<int>(s.idx));
                pc.sh_b.push_back(-1);
                pc.atom_a.push_back(s.atom);
                pc.atom_b.push_back(s.atom);
                pc.mol.push_back(molOf(s));
                pc.poff.push_back(nprim);
                int K = static_cast<int>(s.exps.size());
                for (int x = 0; x < K; x++)
                {
                    pc.prim_x.push_back(x);
                    pc.prim_y.push_back(0);
                }
                pc.pcnt.push_back(K);
                pc.kmax = std::max(pc.kmax, K);
                nprim += K;
            }
            pc.n_pairs = static_cast<int>(pc.pcnt.size());
            pc.n_ppairs = nprim;
        }
    }

    void Context::release()
    {
        for (void *p : allocations)
            (void)hipFree(p);
        allocations.clear();
    }
}
