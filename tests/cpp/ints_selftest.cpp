// Known-answer checks of the lible::ints C++ interface, written like the reference's own
// test/tests_ints.cpp: Structures from named basis sets, the integral functions, and the sum of
// |values| against the reference's answers (tests_ints.cpp:143-166, 694-760).
//
//   ints_selftest <basis_dir> cpu   basis reading, Structure bookkeeping, VectorMD (no device)
//   ints_selftest <basis_dir> gpu   integrals on the current HIP device
//
// Prints one "PASS <name>" / "FAIL <name>: ..." line per check; exit status = number of failures.
#include <lible/ints/ints.hpp>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <unistd.h>
#include <functional>
#include <string>
#include <vector>

namespace lints = lible::ints;
using lible::vec2d;
using lible::vec3d;
using lible::vec4d;

namespace
{
    int failures = 0;

    void report(const std::string &name, bool ok, const std::string &detail = "")
    {
        std::printf("%s %s%s%s\n", ok ? "PASS" : "FAIL", name.c_str(), detail.empty() ? "" : ": ",
                    detail.c_str());
        std::fflush(stdout);
        failures += ok ? 0 : 1;
    }

    void guarded(const std::string &name, const std::function<void()> &fn)
    {
        try
        {
            fn();
        }
        catch (const std::exception &e)
        {
            report(name, false, std::string("exception: ") + e.what());
        }
    }

    template <typename T>
    double absSum(const T &v)
    {
        double s = 0;
        for (double x : v)
            s += std::fabs(x);
        return s;
    }

    // The reference compares these sums with an absolute 1e-12 (tests_ints.cpp:15), i.e. to the
    // last bits of a ~1e4 sum; a different summation order moves them by ~1e-13 relative, so the
    // bar here is 1e-12 relative.
    void checkSum(const std::string &name, double got, double want)
    {
        char buf[160];
        std::snprintf(buf, sizeof buf, "sum %.15g, reference %.15g, rel diff %.2e", got, want,
                      std::fabs(got - want) / want);
        report(name, std::fabs(got - want) <= 1e-12 * std::fabs(want), buf);
    }

    const std::vector<int> nrs_h2o{8, 1, 1};
    const std::vector<std::array<double, 3>> xyz_h2o{
        {0.0, 0.0, 0.0}, {0.757, 0.586, 0.0}, {-0.757, 0.586, 0.0}};   // tests_ints.cpp:62-73
    const std::vector<int> nrs_o3{8, 8, 8};
    const std::vector<std::array<double, 3>> xyz_o3{
        {0, 0, 0}, {0, 1, 1}, {0, 0, 2}};                               // tests_ints.cpp:115-120

    const std::vector<int> nrs_c2h4{6, 6, 1, 1, 1, 1};
    const std::vector<std::array<double, 3>> xyz_c2h4{
        {3.402, 0.773, -9.252}, {4.697, 0.791, -8.909}, {2.933, -0.150, -9.521},
        {2.837, 1.682, -9.258}, {5.262, -0.118, -8.904}, {5.167, 1.714, -8.641}};   // :51-58
    const std::vector<int> nrs_co2{6, 8, 8};
    const std::vector<std::array<double, 3>> xyz_co2{
        {0.0, 0.0, 0.32942}, {0.0, 1.15362, -0.12353}, {0.0, -1.15362, -0.12353}};  // :18-23

    void cpuChecks()
    {
        guarded("vectormd_layout", [] {
            vec3d v(2, 3, 4);
            bool ok = v.size() == 24 && v.dim<0>() == 2 && v.dim<1>() == 3 && v.dim<2>() == 4;
            v(1, 2, 3) = 7.0;
            ok = ok && v.memptr()[1 * 12 + 2 * 4 + 3] == 7.0;
            lible::VectorMD<double, 2> f(lible::Fill(2.5), 3, 3);
            ok = ok && absSum(f) == 22.5;
            report("vectormd_layout", ok);
        });
        guarded("shell_pair_data", [] {
            // O3 / def2-SVP: 9 s, 6 p, 3 d shells = 171 unique pairs, of which the core s pair
            // of the atoms 3.78 bohr apart has no primitive pair above 1e-15; the ordered pairs
            // are both orders of the distinct-shell ones plus the 18 same-shell pairs
            lints::Structure s("def2-svp", nrs_o3, xyz_o3);
            size_t symm = 0, all = 0, prim = 0;
            bool ok = true;
            for (const auto &d : lints::shellPairData(true, s))
            {
                symm += d.n_pairs_;
                ok = ok && d.offsets_sph_.size() == 2 * d.n_pairs_ && d.nrs_ppairs_.size() == d.n_pairs_ &&
                     d.exps_.size() == 2 * d.n_ppairs_ && d.la_ >= d.lb_;
            }
            for (const auto &d : lints::shellPairData(false, s))
            {
                all += d.n_pairs_;
                prim += d.n_ppairs_total_;
            }
            ok = ok && symm == 170 && all == 2 * 170 - 18;
            report("shell_pair_data", ok, "pairs " + std::to_string(symm) + " / " + std::to_string(all));
        });
        guarded("shell_pair_data_ctor", [] {
            // the constructor over getShellsL (shell_pair_data.hpp:50-57) as the reference's
            // one-electron code calls it: the same pairs as shellPairData; the shells carry the
            // reference's fields (O: Z = 8, 2l + 1 spherical and (l+1)(l+2)/2 Cartesian AOs)
            lints::Structure s("def2-svp", nrs_o3, xyz_o3);
            bool ok = s.getShells().size() == 18 && s.getShellsL(1).size() == 6;
            for (const auto &sh : s.getShells())
                ok = ok && sh.z_ == 8 && sh.dim_sph_ == size_t(2 * sh.l_ + 1) &&
                     sh.dim_cart_ == size_t((sh.l_ + 1) * (sh.l_ + 2) / 2);
            for (const auto &d : lints::shellPairData(true, s))
            {
                lints::ShellPairData c(true, d.la_, d.lb_, s.getShellsL(d.la_), s.getShellsL(d.lb_));
                ok = ok && c.n_pairs_ == d.n_pairs_ && c.n_ppairs_ == d.n_ppairs_ && c.exps_ == d.exps_ &&
                     c.coeffs_ == d.coeffs_ && c.offsets_sph_ == d.offsets_sph_ && c.shell_idxs_ == d.shell_idxs_;
            }
            // no screening at threshold 0: every primitive pair kept
            lints::ShellPairData z(false, 0, 0, s.getShellsL(0), s.getShellsL(0), 0.0);
            ok = ok && z.n_ppairs_ == z.n_ppairs_total_ && z.n_pairs_ == 81;
            bool threw = false;
            try
            {
                lints::ShellPairData bad(true, 1, 0, s.getShellsL(0), s.getShellsL(0));
            }
            catch (const std::runtime_error &)
            {
                threw = true;
            }
            report("shell_pair_data_ctor", ok && threw);
        });
        guarded("dims_o3_svp", [] {
            lints::Structure s("def2-svp", nrs_o3, xyz_o3);
            // O in def2-SVP: 3s 2p 1d = 14 spherical functions
            report("dims_o3_svp", s.getDimAO() == 42 && s.getMaxL() == 2 && !s.getUseRI() &&
                                      s.getNAtoms() == 3);
        });
        guarded("dims_h2o_ri", [] {
            lints::Structure s("def2-SVP", "def2-universal-jkfit", nrs_h2o, xyz_h2o);
            const auto h = lints::basisForAtomAux(1, "def2-universal-jkfit");
            const auto o = lints::basisForAtomAux(8, "def2-universal-jkfit");
            size_t naux = 0;
            for (const auto *a : {&o, &h, &h})
                for (const auto &sh : a->basis_shells_)
                    naux += 2 * sh.l_ + 1;
            report("dims_h2o_ri", s.getDimAO() == 24 && s.getUseRI() && s.getDimAOAux() == naux);
        });
        guarded("coords_bohr", [] {
            lints::Structure s("def2-svp", nrs_o3, xyz_o3);
            const double z = s.getCoordsAtom(2)[2];
            report("coords_bohr", std::fabs(z - 2.0 / 0.5291772083) < 1e-14);
        });
        guarded("ghost_dims", [] {
            lints::Structure s("def2-svp", "def2-svp", {8, 1}, {1}, {xyz_h2o[0], xyz_h2o[1]},
                               {xyz_h2o[2]});
            report("ghost_dims", s.getDimAO() == 24 && s.getNAtoms() == 2);
        });
        guarded("calc_boys_f", [] {
            // tests_ints.cpp:1296-1313: sum of F_0..F_5 at x = 2.5, absolute 1e-12
            lints::BoysGrid grid(5);
            double sum = 0;
            for (double v : lints::calcBoysF(5, 2.5, grid))
                sum += v;
            const auto f0 = lints::calcBoysF(0, 0.0, grid), f30 = lints::calcBoysF(3, 45.0, grid);
            report("calc_boys_f", std::fabs(sum - 0.728561124180) < 1e-12 && f0[0] == 1.0 &&
                                      std::fabs(f30[0] - 0.5 * std::sqrt(M_PI / 45.0)) < 1e-16,
                   "sum " + std::to_string(sum));
        });
        guarded("available_basis_sets", [] {
            const auto m = lints::availableBasisSets(), a = lints::availableBasisSetsAux();
            report("available_basis_sets", m.count("def2-svp") && m.count("def2-tzvp") &&
                                               a.count("def2-universal-jkfit"));
        });
        guarded("missing_element", [] {
            bool threw = false;
            try
            {
                lints::basisForAtom(92, "def2-svp");
            }
            catch (const std::runtime_error &)
            {
                threw = true;
            }
            report("missing_element", threw);
        });
        guarded("missing_basis", [] {
            bool threw = false;
            try
            {
                lints::Structure s("no-such-basis", nrs_o3, xyz_o3);
            }
            catch (const std::runtime_error &)
            {
                threw = true;
            }
            report("missing_basis", threw);
        });
    }

    void gpuChecks()
    {
        guarded("eri4", [] {
            lints::Structure s("def2-svp", nrs_o3, xyz_o3);
            vec4d e = lints::eri4(s);
            checkSum("eri4", absSum(e), 8246.711763196334);   // tests_ints.cpp:748
            const size_t n = e.dim<0>();
            double asym = 0;
            for (size_t a = 0; a < n; a += 5)
                for (size_t b = 0; b < n; b += 3)
                    for (size_t c = 0; c < n; c += 2)
                        for (size_t d = 0; d < n; d++)
                            asym = std::max({asym, std::fabs(e(a, b, c, d) - e(b, a, c, d)),
                                             std::fabs(e(a, b, c, d) - e(c, d, a, b))});
            report("eri4_symmetry", asym <= 1e-14, "max asym " + std::to_string(asym));

            // coulombExchange against the contraction of the dense tensor
            vec2d D(n, n);
            for (size_t i = 0; i < n; i++)
                for (size_t j = 0; j <= i; j++)
                    D(i, j) = D(j, i) = std::cos(0.37 * i + 0.11 * j) / (1.0 + 0.1 * (i + j));
            auto [J, K] = lints::coulombExchange(s, D, 0.0);
            double err = 0, scale = 0;
            for (size_t m = 0; m < n; m++)
                for (size_t k = 0; k < n; k++)
                {
                    double j = 0, x = 0;
                    for (size_t p = 0; p < n; p++)
                        for (size_t q = 0; q < n; q++)
                        {
                            j += e(m, k, p, q) * D(p, q);
                            x += e(m, p, k, q) * D(p, q);
                        }
                    err = std::max({err, std::fabs(J(m, k) - j), std::fabs(K(m, k) - x)});
                    scale = std::max({scale, std::fabs(j), std::fabs(x)});
                }
            report("coulomb_exchange", err <= 1e-12 * scale,
                   "max err " + std::to_string(err / scale) + " of scale");
        });
        guarded("one_electron", [] {
            // the reference's one-electron known answers (tests_ints.cpp:167-183, 236-252,
            // 303-319) and nuclearAttraction = externalCharges at the nuclei
            lints::Structure s("cc-pvdz", nrs_c2h4, xyz_c2h4);
            checkSum("one_electron_overlap", absSum(lints::overlap(s)), 226.099184393591);
            checkSum("one_electron_kinetic", absSum(lints::kineticEnergy(s)), 252.755644635021);
            lints::Structure s2("6-31+g", nrs_co2, xyz_co2);
            vec2d V = lints::nuclearAttraction(s2);
            checkSum("one_electron_nuclear", absSum(V), 2141.927982294058);
            std::vector<std::array<double, 4>> q;
            for (size_t a = 0; a < s2.getNAtoms(); a++)
            {
                const auto c = s2.getCoordsAtom(a);
                q.push_back({c[0], c[1], c[2], double(s2.getZ(a))});
            }
            vec2d V2 = lints::externalCharges(q, s2);
            // two evaluations: the atomic accumulation order may differ in the last bits
            double d = 0, vmax = 0;
            for (size_t i = 0; i < V.size(); i++)
            {
                d = std::max(d, std::fabs(V.memptr()[i] - V2.memptr()[i]));
                vmax = std::max(vmax, std::fabs(V.memptr()[i]));
            }
            report("one_electron_external_charges", d <= 1e-13 * vmax,
                   "max diff " + std::to_string(d / vmax) + " of the largest element");
        });
        guarded("coulomb_exchange_gradient", [] {
            // the backward entry against central differences (h = 1e-4 bohr) of the energy of
            // coulombExchange, each difference formed as 1/2 tr(D dJ) - 1/4 tr(D dK), and the
            // translational invariance sum_A dE/dR_A = 0
            std::vector<std::array<double, 3>> xyz = xyz_h2o;
            lints::Structure s("def2-svp", nrs_h2o, xyz);
            const size_t n = s.getDimAO();
            vec2d D(n, n);
            for (size_t i = 0; i < n; i++)
                for (size_t j = 0; j <= i; j++)
                    D(i, j) = D(j, i) = std::cos(0.29 * i - 0.17 * j) / (1.0 + 0.2 * (i + j));
            vec2d g = lints::coulombExchangeGradient(s, D, 0.0);
            const double h = 1e-4;
            double err = 0, gmax = 0, tsum[3] = {0, 0, 0};
            for (size_t a = 0; a < g.dim<0>(); a++)
                for (int x = 0; x < 3; x++)
                {
                    tsum[x] += g(a, x);
                    gmax = std::max(gmax, std::fabs(g(a, x)));
                    auto shifted = [&](double sgn) {
                        auto c = xyz;
                        c[a][x] += sgn * h * lints::_bohr_to_ang_;
                        return lints::coulombExchange(lints::Structure("def2-svp", nrs_h2o, c), D, 0.0);
                    };
                    auto [Jp, Kp] = shifted(1.0);
                    auto [Jm, Km] = shifted(-1.0);
                    double dE = 0;
                    for (size_t i = 0; i < n * n; i++)
                        dE += D.memptr()[i] * (0.5 * (Jp.memptr()[i] - Jm.memptr()[i]) -
                                               0.25 * (Kp.memptr()[i] - Km.memptr()[i]));
                    err = std::max(err, std::fabs(g(a, x) - dE / (2 * h)));
                }
            const double tmax = std::max({std::fabs(tsum[0]), std::fabs(tsum[1]), std::fabs(tsum[2])});
            report("coulomb_exchange_gradient", g.dim<0>() == 3 && err <= 1e-7 && tmax <= 1e-10 * gmax,
                   "max |g - fd| " + std::to_string(err) + ", |sum| " + std::to_string(tmax));
        });
        guarded("eri4_diagonal", [] {
            lints::Structure s("def2-tzvp", nrs_o3, xyz_o3);
            checkSum("eri4_diagonal", absSum(lints::eri4Diagonal(s)), 306.806419250122);  // :712
        });
        guarded("eri3", [] {
            lints::Structure s("def2-tzvp", "def2-qzvp-rifit", nrs_o3, xyz_o3);
            checkSum("eri3", absSum(lints::eri3(s)), 25854.105851585416531);             // :730
        });
        guarded("eri2", [] {
            lints::Structure s("def2-SVP", "def2-qzvp-rifit", nrs_h2o, xyz_h2o);
            checkSum("eri2", absSum(lints::eri2(s)), 8801.334703460838);                 // :694
        });
        guarded("eri2_diagonal", [] {
            lints::Structure s("def2-SVP", "def2-universal-jkfit", nrs_h2o, xyz_h2o);
            checkSum("eri2_diagonal", absSum(lints::eri2Diagonal(s)), 830.788976485982); // :149
        });
        guarded("g_shells_def2qzvp", [] {
            // a main basis with g shells (O in def2-QZVP, 7s4p3d2f1g: 57 functions): the
            // per-class-pair functors' blocks add up to the dense tensor, and J/K equal its
            // contraction
            lints::Structure s("def2-qzvp", {8}, {{0.0, 0.0, 0.0}});
            vec4d e = lints::eri4(s);
            const size_t n = e.dim<0>();
            auto sp = lints::shellPairData(false, s);
            double sum = 0;
            for (const auto &ab : sp)
                for (const auto &cd : sp)
                {
                    lints::ERI4Kernel kernel(ab, cd);
                    for (size_t p = 0; p < ab.n_pairs_; p++)
                        for (size_t q = 0; q < cd.n_pairs_; q++)
                            sum += absSum(kernel(p, q, ab, cd));
                }
            const double dense = absSum(e);
            vec2d D(n, n);
            for (size_t i = 0; i < n; i++)
                for (size_t j = 0; j <= i; j++)
                    D(i, j) = D(j, i) = std::cos(0.29 * i + 0.13 * j) / (1.0 + 0.1 * (i + j));
            auto [J, K] = lints::coulombExchange(s, D, 0.0);
            double err = 0, scale = 0;
            for (size_t m = 0; m < n; m++)
                for (size_t k = 0; k < n; k++)
                {
                    double j = 0, x = 0;
                    for (size_t p = 0; p < n; p++)
                        for (size_t q = 0; q < n; q++)
                        {
                            j += e(m, k, p, q) * D(p, q);
                            x += e(m, p, k, q) * D(p, q);
                        }
                    err = std::max({err, std::fabs(J(m, k) - j), std::fabs(K(m, k) - x)});
                    scale = std::max({scale, std::fabs(j), std::fabs(x)});
                }
            report("g_shells_def2qzvp", s.getMaxL() == 4 && n == 57 && std::fabs(sum - dense) <= 1e-11 * dense &&
                                                err <= 1e-12 * scale,
                   "functor sum " + std::to_string(sum) + ", dense " + std::to_string(dense) + ", J/K err " +
                           std::to_string(err / scale) + " of scale");
        });
        // the per-block kernels, driven as tests_ints.cpp:873-1053 drives them
        guarded("deploy_eri4_kernel", [] {
            lints::Structure structure("def2-svp", nrs_o3, xyz_o3);
            auto sp = lints::shellPairData(false, structure);
            double sum = 0;
            for (const auto &ab : sp)
                for (const auto &cd : sp)
                {
                    lints::ERI4Kernel kernel(ab, cd);
                    for (size_t p = 0; p < ab.n_pairs_; p++)
                        for (size_t q = 0; q < cd.n_pairs_; q++)
                            sum += absSum(kernel(p, q, ab, cd));
                }
            checkSum("deploy_eri4_kernel", sum, 8246.711763197583);   // :873-900
        });
        guarded("deploy_eri3_kernel", [] {
            lints::Structure structure("def2-svp", "def2-universal-jkfit", nrs_o3, xyz_o3);
            auto sp = lints::shellPairData(false, structure);
            auto sh = lints::shellDataAux(structure);
            double sum = 0;
            for (const auto &ab : sp)
                for (const auto &c : sh)
                {
                    lints::ERI3Kernel kernel(ab, c);
                    for (size_t p = 0; p < ab.n_pairs_; p++)
                        for (size_t k = 0; k < c.n_shells_; k++)
                            sum += absSum(kernel(p, k, ab, c));
                }
            checkSum("deploy_eri3_kernel", sum, 12350.130354189232);  // :902-930
        });
        guarded("deploy_eri2_kernel", [] {
            lints::Structure structure("def2-svp", "def2-universal-jkfit", nrs_o3, xyz_o3);
            auto sh = lints::shellDataAux(structure);
            double sum = 0;
            for (const auto &a : sh)
                for (const auto &b : sh)
                {
                    lints::ERI2Kernel kernel(a, b);
                    for (size_t i = 0; i < a.n_shells_; i++)
                        for (size_t j = 0; j < b.n_shells_; j++)
                            sum += absSum(kernel(i, j, a, b));
                }
            checkSum("deploy_eri2_kernel", sum, 21014.512841176023);  // :932-959
        });
        guarded("deploy_eri4d1_kernel", [] {
            lints::Structure structure("def2-svp", nrs_o3, xyz_o3);
            auto sp = lints::shellPairData(true, structure);
            double sum = 0;
            for (const auto &ab : sp)
                for (const auto &cd : sp)
                {
                    lints::ERI4D1Kernel kernel(ab, cd);
                    for (size_t p = 0; p < ab.n_pairs_; p++)
                        for (size_t q = 0; q < cd.n_pairs_; q++)
                            for (const vec4d &blk : kernel(p, q, ab, cd))
                                sum += absSum(blk);
                }
            checkSum("deploy_eri4d1_kernel", sum, 30546.780073140635);  // :961-990
        });
        guarded("deploy_eri3d1_kernel", [] {
            lints::Structure structure("def2-svp", "def2-universal-jkfit", nrs_o3, xyz_o3);
            auto sp = lints::shellPairData(false, structure);
            auto sh = lints::shellDataAux(structure);
            double sum = 0;
            for (const auto &ab : sp)
                for (const auto &c : sh)
                {
                    lints::ERI3D1Kernel kernel(ab, c);
                    for (size_t p = 0; p < ab.n_pairs_; p++)
                        for (size_t k = 0; k < c.n_shells_; k++)
                            for (const vec3d &blk : kernel(p, k, ab, c))
                                sum += absSum(blk);
                }
            checkSum("deploy_eri3d1_kernel", sum, 78466.360848033844);  // :992-1022
        });
        guarded("deploy_eri2d1_kernel", [] {
            lints::Structure structure("def2-svp", "def2-universal-jkfit", nrs_o3, xyz_o3);
            auto sh = lints::shellDataAux(structure);
            double sum = 0;
            for (const auto &a : sh)
                for (const auto &b : sh)
                {
                    lints::ERI2D1Kernel kernel(a, b);
                    for (size_t i = 0; i < a.n_shells_; i++)
                        for (size_t j = 0; j < b.n_shells_; j++)
                            for (const vec2d &blk : kernel(i, j, a, b))
                                sum += absSum(blk);
                }
            checkSum("deploy_eri2d1_kernel", sum, 39949.218421684251);  // :1024-1053
        });
        guarded("deploy_eri2d2_kernel", [] {
            lints::Structure structure("def2-svp", "def2-universal-jkfit", nrs_o3, xyz_o3);
            auto sh = lints::shellDataAux(structure);
            double sum = 0;
            for (const auto &a : sh)
                for (const auto &b : sh)
                {
                    lints::ERI2D2Kernel kernel(a, b);
                    for (size_t i = 0; i < a.n_shells_; i++)
                        for (size_t j = 0; j < b.n_shells_; j++)
                            for (const auto &row : kernel(i, j, a, b))
                                for (const vec2d &blk : row)
                                    sum += absSum(blk);
                }
            checkSum("deploy_eri2d2_kernel", sum, 175567.635699242965);  // :1055-1085
        });
        guarded("deploy_eri4soc_kernel", [] {
            lints::Structure structure("def2-svp", nrs_o3, xyz_o3);
            auto sp = lints::shellPairData(false, structure);
            double sum = 0;
            for (const auto &ab : sp)
                for (const auto &cd : sp)
                {
                    lints::ERI4SOCKernel kernel(ab, cd);
                    for (size_t p = 0; p < ab.n_pairs_; p++)
                        for (size_t q = 0; q < cd.n_pairs_; q++)
                            for (const vec4d &blk : kernel(p, q, ab, cd))
                                sum += absSum(blk);
                }
            checkSum("deploy_eri4soc_kernel", sum, 12940.953099658431);  // :1087-1116
        });
        guarded("deploy_eri3soc_kernel", [] {
            lints::Structure structure("def2-svp", "def2-universal-jkfit", nrs_o3, xyz_o3);
            auto sp = lints::shellPairData(false, structure);
            auto sh = lints::shellDataAux(structure);
            double sum = 0;
            for (const auto &ab : sp)
                for (const auto &c : sh)
                {
                    lints::ERI3SOCKernel kernel(ab, c);
                    for (size_t p = 0; p < ab.n_pairs_; p++)
                        for (size_t k = 0; k < c.n_shells_; k++)
                            for (const vec3d &blk : kernel(p, k, ab, c))
                                sum += absSum(blk);
                }
            checkSum("deploy_eri3soc_kernel", sum, 18124.713152990007);  // :1118-1148
        });
        guarded("eri4_benchmark", [] {
            // eri4Benchmark prints per class pair and the sum of |(mn|kl)| over the unique
            // quartets; its output is captured and the sum checked against eri4Diagonal's
            // companion checksum of the same molecule (tests_ints.cpp:748 sums the dense tensor;
            // over unique quartets only, the sum is the one printed by the reference's benchmark)
            lints::Structure s("def2-svp", nrs_o3, xyz_o3);
            std::fflush(stdout);
            const int saved = dup(fileno(stdout));
            std::FILE *tmp = std::tmpfile();
            dup2(fileno(tmp), fileno(stdout));
            lints::eri4Benchmark(s);
            std::fflush(stdout);
            dup2(saved, fileno(stdout));
            close(saved);
            std::rewind(tmp);
            std::string text;
            char buf[512];
            while (std::fgets(buf, sizeof buf, tmp))
                text += buf;
            std::fclose(tmp);
            const size_t at = text.find("sum_eri4 = ");
            const double got = at == std::string::npos ? -1.0 : std::atof(text.c_str() + at + 11);
            // the same sum from the dense tensor restricted to unique quartets
            vec4d e = lints::eri4(s);
            const auto sp = lints::shellPairData(true, s);
            double want = 0;
            for (size_t i = 0; i < sp.size(); i++)
                for (size_t j = 0; j <= i; j++)
                    for (size_t p = 0; p < sp[i].n_pairs_; p++)
                        for (size_t q = 0; q < (i == j ? p + 1 : sp[j].n_pairs_); q++)
                        {
                            const auto &ab = sp[i], &cd = sp[j];
                            for (size_t a = 0; a < 2 * size_t(ab.la_) + 1; a++)
                                for (size_t b = 0; b < 2 * size_t(ab.lb_) + 1; b++)
                                    for (size_t c = 0; c < 2 * size_t(cd.la_) + 1; c++)
                                        for (size_t d = 0; d < 2 * size_t(cd.lb_) + 1; d++)
                                            want += std::fabs(e(ab.offsets_sph_[2 * p] + a, ab.offsets_sph_[2 * p + 1] + b,
                                                                cd.offsets_sph_[2 * q] + c, cd.offsets_sph_[2 * q + 1] + d));
                        }
            report("eri4_benchmark", std::fabs(got - want) <= 1e-11 * want && text.find("done") != std::string::npos,
                   "printed " + std::to_string(got) + ", unique-quartet sum " + std::to_string(want));
        });
        guarded("ghost_atoms", [] {
            // Integrals do not see nuclear charges: with the last H a ghost atom, the AO order is
            // unchanged and the tensor equals the all-real one.
            lints::Structure full("def2-svp", nrs_h2o, xyz_h2o);
            lints::Structure ghost("def2-svp", "def2-svp", {8, 1}, {1}, {xyz_h2o[0], xyz_h2o[1]},
                                   {xyz_h2o[2]});
            vec4d a = lints::eri4(full), b = lints::eri4(ghost);
            double d = 0;
            for (size_t i = 0; i < a.size(); i++)
                d = std::max(d, std::fabs(a[i] - b[i]));
            report("ghost_atoms", a.size() == b.size() && d == 0.0);
        });
    }
}

int main(int argc, char **argv)
{
    if (argc < 3 || (std::strcmp(argv[2], "cpu") != 0 && std::strcmp(argv[2], "gpu") != 0))
    {
        std::fprintf(stderr, "usage: %s <basis_dir> cpu|gpu\n", argv[0]);
        return 2;
    }
    lints::BasisPaths::setMainBasisSetsPath(argv[1]);
    lints::BasisPaths::setAuxBasisSetsPath(argv[1]);
    if (std::strcmp(argv[2], "cpu") == 0)
        cpuChecks();
    else
        gpuChecks();
    return failures;
}
