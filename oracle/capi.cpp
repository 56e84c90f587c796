// Oracle (test infrastructure only): C entry points for ctypes.  Loaded by tests/ and by
// bench.py's cpu_baseline leg; never by the product.
#include "oracle.hpp"
#include <type_traits>
#include <algorithm>
#include <memory>
#include <map>

#include <cmath>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>

using namespace orc;

namespace
{
    template <class T>
    struct Tag
    {
        using type = T;
    };
}

namespace
{
    thread_local std::string g_err;

    template <class F>
    int guard(F &&f)
    {
        try
        {
            f();
            return 0;
        }
        catch (const std::exception &e)
        {
            g_err = e.what();
            return -1;
        }
    }

    std::vector<std::vector<BasisShell>> unpack(int natoms, int nsh, const int *atom, const int *l,
                                                const int *nprim, const double *exps,
                                                const double *coeffs)
    {
        std::vector<std::vector<BasisShell>> atoms(natoms);
        size_t o = 0;
        for (int i = 0; i < nsh; i++)
        {
            if (atom[i] < 0 || atom[i] >= natoms)
                throw std::runtime_error("shell atom index out of range");
            if (i > 0 && atom[i] < atom[i - 1])
                throw std::runtime_error("shells must be ordered by atom");
            BasisShell b;
            b.l = l[i];
            b.exps.assign(exps + o, exps + o + nprim[i]);
            b.coeffs.assign(coeffs + o, coeffs + o + nprim[i]);
            o += nprim[i];
            atoms[atom[i]].push_back(b);
        }
        return atoms;
    }

    double absSum(const std::vector<double> &v)
    {
        double s = 0;
        for (double x : v)
            s += std::fabs(x);
        return s;
    }
}

extern "C"
{
    const char *orc_last_error() { return g_err.c_str(); }

    void orc_boys(int n, double x, double *out)
    {
        BoysGrid g(n);
        calcBoysF(n, x, g, out);
    }

    // Boys values for many x at one n (tests of the GPU kernel).
    void orc_boys_many(int n, long cnt, const double *x, double *out)
    {
        BoysGrid g(n);
#pragma omp parallel for
        for (long i = 0; i < cnt; i++)
            calcBoysF(n, x[i], g, out + i * (n + 1));
    }

    void orc_rints3d(int l, double p, const double *xyz, const double *fnx, double *out)
    {
        auto r = calcRInts3D(l, p, xyz, fnx);
        std::memcpy(out, r.data(), r.size() * sizeof(double));
    }

    int orc_sph_trafo(int l, int cap, int *sph, int *cart, double *val)
    {
        int n = -1;
        guard([&]
              {
                  auto t = sphericalTrafo(l);
                  n = static_cast<int>(t.size());
                  for (int i = 0; i < n && i < cap; i++)
                  {
                      sph[i] = std::get<0>(t[i]);
                      cart[i] = std::get<1>(t[i]);
                      val[i] = std::get<2>(t[i]);
                  }
              });
        return n;
    }

    double orc_pure_primitive_norm(int l, double e) { return purePrimitiveNorm(l, e); }

    void orc_ecoeffs_recurrence1(double one_o_2a, int l, double *out)
    {
        auto e = ecoeffsRecurrence1(one_o_2a, l);
        std::memcpy(out, e.data(), e.size() * sizeof(double));
    }

    void orc_ecoeffs_recurrence2(double a, double b, int la, int lb, double PA, double PB,
                                 double Kab, double *out)
    {
        auto e = ecoeffsRecurrence2(a, b, la, lb, PA, PB, Kab);
        std::memcpy(out, e.data(), e.size() * sizeof(double));
    }

    void orc_ecoeffs_recurrence2_n1(double a, double b, int la, int lb, double PA, double PB,
                                    double Kab, double A, double B, double *out)
    {
        auto e0 = ecoeffsRecurrence2(a, b, la, lb, PA, PB, Kab);
        auto e1 = ecoeffsRecurrence2_n1(a, b, la, lb, A, B, e0);
        std::memcpy(out, e1.data(), e1.size() * sizeof(double));
    }

    // 3 consecutive blocks (x, y, z), each (la+1)(lb+1)(la+lb+1)
    void orc_ecoeffs_primitive_pair(double a, double b, int la, int lb, const double *xa,
                                    const double *xb, const double *A, const double *B,
                                    int deriv, double *out)
    {
        auto e0 = ecoeffsPrimitivePair(a, b, la, lb, xa, xb);
        auto e = deriv ? ecoeffsPrimitivePair_n1(a, b, la, lb, A, B, e0) : e0;
        size_t n = e[0].size();
        for (int k = 0; k < 3; k++)
            std::memcpy(out + k * n, e[k].data(), n * sizeof(double));
    }

    double orc_ecoeffs_shell_abs_sum(int l, int n, const double *exps)
    {
        auto e = ecoeffsShell(l, std::vector<double>(exps, exps + n));
        double s = 0;
        for (auto &v : e)
            s += absSum(v);
        return s;
    }

    void *orc_structure_create(int natoms, const double *coords_bohr, int nsh, const int *atom,
                               const int *l, const int *nprim, const double *exps,
                               const double *coeffs, int nsh_aux, const int *atom_aux,
                               const int *l_aux, const int *nprim_aux, const double *exps_aux,
                               const double *coeffs_aux)
    {
        Structure *s = nullptr;
        guard([&]
              {
                  std::vector<std::array<double, 3>> xyz(natoms);
                  for (int i = 0; i < natoms; i++)
                      xyz[i] = {coords_bohr[3 * i], coords_bohr[3 * i + 1], coords_bohr[3 * i + 2]};
                  auto main = unpack(natoms, nsh, atom, l, nprim, exps, coeffs);
                  if (nsh_aux > 0)
                  {
                      auto aux = unpack(natoms, nsh_aux, atom_aux, l_aux, nprim_aux, exps_aux,
                                        coeffs_aux);
                      s = new Structure(makeStructure(xyz, main, &aux));
                  }
                  else
                      s = new Structure(makeStructure(xyz, main, nullptr));
              });
        return s;
    }

    void orc_structure_destroy(void *h) { delete static_cast<Structure *>(h); }

    void orc_structure_dims(void *h, long *nbf, long *nbf_aux, long *natoms)
    {
        auto *s = static_cast<Structure *>(h);
        *nbf = s->nbf;
        *nbf_aux = s->nbf_aux;
        *natoms = s->natoms;
    }

    // shell normalisation constants, concatenated in shell order (main basis)
    void orc_structure_norms(void *h, double *out)
    {
        auto *s = static_cast<Structure *>(h);
        size_t o = 0;
        for (auto &sh : s->shells)
            for (double v : sh.norms)
                out[o++] = v;
    }

    int orc_eri4(void *h, double *out)
    {
        return guard([&]
                     {
                         auto v = eri4(*static_cast<Structure *>(h));
                         std::memcpy(out, v.data(), v.size() * sizeof(double));
                     });
    }

    int orc_eri4_diag(void *h, double *out)
    {
        return guard([&]
                     {
                         auto v = eri4Diagonal(*static_cast<Structure *>(h));
                         std::memcpy(out, v.data(), v.size() * sizeof(double));
                     });
    }

    int orc_eri3(void *h, double *out)
    {
        return guard([&]
                     {
                         auto v = eri3(*static_cast<Structure *>(h));
                         std::memcpy(out, v.data(), v.size() * sizeof(double));
                     });
    }

    // Sums |eri3|, eri3, |eri2|, eri2 over the dense tensors without forming them.
    int orc_ri_sums(void *h, double *out)
    {
        return guard([&] { riSums(*static_cast<Structure *>(h), out); });
    }

    int orc_eri2(void *h, double *out)
    {
        return guard([&]
                     {
                         auto v = eri2(*static_cast<Structure *>(h));
                         std::memcpy(out, v.data(), v.size() * sizeof(double));
                     });
    }

    int orc_eri2_diag(void *h, double *out)
    {
        return guard([&]
                     {
                         auto v = eri2Diagonal(*static_cast<Structure *>(h));
                         std::memcpy(out, v.data(), v.size() * sizeof(double));
                     });
    }

    // Dense first-derivative tensors from the per-block D1 kernels over every ordered shell pair
    // (shellPairData(false), the pairs of the reference's deploy tests, tests_ints.cpp:961-1053):
    // component c = 3 k + x is the derivative with respect to coordinate x of the centre of the
    // k-th index's shell.
    //   eri4d1 (12, nbf, nbf, nbf, nbf), eri3d1 (9, nbf, nbf, naux), eri2d1 (6, naux, naux)
    int orc_eri4d1(void *h, double *out)
    {
        return guard([&]
                     {
                         const Structure &s = *static_cast<Structure *>(h);
                         const size_t n = s.nbf, n4 = n * n * n * n;
                         std::fill(out, out + 12 * n4, 0.0);
                         auto sp = shellPairData(false, s);
                         for (const auto &A : sp)
                             for (const auto &C : sp)
                             {
                                 const int na = nSph(A.la), nb = nSph(A.lb), nc = nSph(C.la), nd = nSph(C.lb);
                                 const size_t blk = size_t(na) * nb * nc * nd;
                                 ERI4D1Kernel kern(A, C);
#pragma omp parallel for schedule(dynamic)
                                 for (size_t p = 0; p < A.n_pairs; p++)
                                     for (size_t q = 0; q < C.n_pairs; q++)
                                     {
                                         auto b = kern(p, q, A, C);
                                         const size_t oa = A.offsets_sph[2 * p], ob = A.offsets_sph[2 * p + 1];
                                         const size_t oc = C.offsets_sph[2 * q], od = C.offsets_sph[2 * q + 1];
                                         for (int c = 0; c < 12; c++)
                                             for (int ia = 0; ia < na; ia++)
                                                 for (int ib = 0; ib < nb; ib++)
                                                     for (int ic = 0; ic < nc; ic++)
                                                         for (int id = 0; id < nd; id++)
                                                             out[c * n4 + (((oa + ia) * n + ob + ib) * n + oc + ic) * n + od + id] =
                                                                     b[c * blk + ((ia * nb + ib) * nc + ic) * nd + id];
                                     }
                             }
                     });
    }

    int orc_eri3d1(void *h, double *out)
    {
        return guard([&]
                     {
                         const Structure &s = *static_cast<Structure *>(h);
                         const size_t n = s.nbf, nx = s.nbf_aux, n3 = n * n * nx;
                         std::fill(out, out + 9 * n3, 0.0);
                         auto sp = shellPairData(false, s);
                         auto sh = shellDataAux(s);
                         for (const auto &A : sp)
                             for (const auto &C : sh)
                             {
                                 const int na = nSph(A.la), nb = nSph(A.lb), nc = nSph(C.l);
                                 const size_t blk = size_t(na) * nb * nc;
                                 ERI3D1Kernel kern(A, C);
#pragma omp parallel for schedule(dynamic)
                                 for (size_t p = 0; p < A.n_pairs; p++)
                                     for (size_t q = 0; q < C.n_shells; q++)
                                     {
                                         auto b = kern(p, q, A, C);
                                         const size_t oa = A.offsets_sph[2 * p], ob = A.offsets_sph[2 * p + 1];
                                         const size_t oc = C.offsets_sph[q];
                                         for (int c = 0; c < 9; c++)
                                             for (int ia = 0; ia < na; ia++)
                                                 for (int ib = 0; ib < nb; ib++)
                                                     for (int ic = 0; ic < nc; ic++)
                                                         out[c * n3 + ((oa + ia) * n + ob + ib) * nx + oc + ic] =
                                                                 b[c * blk + (ia * nb + ib) * nc + ic];
                                     }
                             }
                     });
    }

    int orc_eri2d1(void *h, double *out)
    {
        return guard([&]
                     {
                         const Structure &s = *static_cast<Structure *>(h);
                         const size_t nx = s.nbf_aux, n2 = nx * nx;
                         std::fill(out, out + 6 * n2, 0.0);
                         auto sh = shellDataAux(s);
                         for (const auto &A : sh)
                             for (const auto &B : sh)
                             {
                                 const int na = nSph(A.l), nb = nSph(B.l);
                                 const size_t blk = size_t(na) * nb;
                                 ERI2D1Kernel kern(A, B);
                                 for (size_t p = 0; p < A.n_shells; p++)
                                     for (size_t q = 0; q < B.n_shells; q++)
                                     {
                                         auto b = kern(p, q, A, B);
                                         const size_t oa = A.offsets_sph[p], ob = B.offsets_sph[q];
                                         for (int c = 0; c < 6; c++)
                                             for (int ia = 0; ia < na; ia++)
                                                 for (int ib = 0; ib < nb; ib++)
                                                     out[c * n2 + (oa + ia) * nx + ob + ib] = b[c * blk + ia * nb + ib];
                                     }
                             }
                     });
    }

    // Sum of |values| over every kernel call of the reference's "deploy" tests
    // (tests_ints.cpp:873-1053).  which: 0 ERI4, 1 ERI3, 2 ERI2, 3 ERI4D1, 4 ERI3D1, 5 ERI2D1.
    int orc_deploy_sum(void *h, int which, int use_symm, double *sum)
    {
        return guard([&]
                     {
                         const Structure &s = *static_cast<Structure *>(h);
                         double tot = 0;
                         if (which == 0 || which == 3)
                         {
                             auto sp = shellPairData(use_symm != 0, s);
                             for (auto &A : sp)
                                 for (auto &C : sp)
                                 {
                                     if (which == 0)
                                     {
                                         ERI4Kernel k(A, C);
                                         for (size_t p = 0; p < A.n_pairs; p++)
                                             for (size_t q = 0; q < C.n_pairs; q++)
                                                 tot += absSum(k(p, q, A, C));
                                     }
                                     else
                                     {
                                         ERI4D1Kernel k(A, C);
                                         for (size_t p = 0; p < A.n_pairs; p++)
                                             for (size_t q = 0; q < C.n_pairs; q++)
                                                 tot += absSum(k(p, q, A, C));
                                     }
                                 }
                         }
                         else if (which == 1 || which == 4)
                         {
                             auto sp = shellPairData(use_symm != 0, s);
                             auto sh = shellDataAux(s);
                             for (auto &A : sp)
                                 for (auto &C : sh)
                                 {
                                     if (which == 1)
                                     {
                                         ERI3Kernel k(A, C);
                                         for (size_t p = 0; p < A.n_pairs; p++)
                                             for (size_t q = 0; q < C.n_shells; q++)
                                                 tot += absSum(k(p, q, A, C));
                                     }
                                     else
                                     {
                                         ERI3D1Kernel k(A, C);
                                         for (size_t p = 0; p < A.n_pairs; p++)
                                             for (size_t q = 0; q < C.n_shells; q++)
                                                 tot += absSum(k(p, q, A, C));
                                     }
                                 }
                         }
                         else
                         {
                             auto sh = shellDataAux(s);
                             for (auto &A : sh)
                                 for (auto &B : sh)
                                 {
                                     if (which == 2)
                                     {
                                         ERI2Kernel k(A, B);
                                         for (size_t p = 0; p < A.n_shells; p++)
                                             for (size_t q = 0; q < B.n_shells; q++)
                                                 tot += absSum(k(p, q, A, B));
                                     }
                                     else
                                     {
                                         ERI2D1Kernel k(A, B);
                                         for (size_t p = 0; p < A.n_shells; p++)
                                             for (size_t q = 0; q < B.n_shells; q++)
                                                 tot += absSum(k(p, q, A, B));
                                     }
                                 }
                         }
                         *sum = tot;
                     });
    }

    // Sums of |x| over the blocks of the second-derivative and spin-orbit kernels, driven as the
    // reference's deployERI2D2Kernel / deployERI4SOCKernel / deployERI3SOCKernel
    // (tests_ints.cpp:1055-1148): which 6 eri2d2, 7 eri4soc, 8 eri3soc; all ordered pairs.
    int orc_deploy_sum2(void *h, int which, double *sum)
    {
        return guard([&]
                     {
                         const Structure &s = *static_cast<Structure *>(h);
                         double tot = 0;
                         if (which == 6)
                         {
                             auto sh = shellDataAux(s);
                             for (auto &A : sh)
                                 for (auto &B : sh)
                                 {
                                     ERI2D2Kernel k(A, B);
                                     for (size_t p = 0; p < A.n_shells; p++)
                                         for (size_t q = 0; q < B.n_shells; q++)
                                             tot += absSum(k(p, q, A, B));
                                 }
                         }
                         else if (which == 7)
                         {
                             auto sp = shellPairData(false, s);
                             for (auto &A : sp)
                                 for (auto &C : sp)
                                 {
                                     ERI4SOCKernel k(A, C);
                                     for (size_t p = 0; p < A.n_pairs; p++)
                                         for (size_t q = 0; q < C.n_pairs; q++)
                                             tot += absSum(k(p, q, A, C));
                                 }
                         }
                         else if (which == 8)
                         {
                             auto sp = shellPairData(false, s);
                             auto sh = shellDataAux(s);
                             for (auto &A : sp)
                                 for (auto &C : sh)
                                 {
                                     ERI3SOCKernel k(A, C);
                                     for (size_t p = 0; p < A.n_pairs; p++)
                                         for (size_t q = 0; q < C.n_shells; q++)
                                             tot += absSum(k(p, q, A, C));
                                 }
                         }
                         else
                             throw std::runtime_error("orc_deploy_sum2: unknown kernel");
                         *sum = tot;
                     });
    }

    // Blocks of listed shell quartets / triplets / doublets in the caller's shell order, as the
    // per-block kernels return them for the ordered pairs of shellPairData(false):
    // which 0 eri4, 1 eri3, 2 eri2, 3 eri4d1, 4 eri3d1, 5 eri2d1, 6 eri2d2, 7 eri4soc,
    // 8 eri3soc; shells: n entries of 4 / 3 / 2 shell indices (main, main, main, main / main,
    // main, aux / aux, aux); out: n consecutive kernel outputs.  Entries whose pair has no
    // primitive pair above the screen are zero.
    int orc_blocks(void *h, int which, long n, const int *shells, double *out)
    {
        return guard([&]
                     {
                         const Structure &s = *static_cast<Structure *>(h);
                         const int kind = (which == 0 || which == 3 || which == 7) ? 4
                                                                                   : (which == 1 || which == 4 || which == 8) ? 3 : 2;
                         auto sp = shellPairData(false, s);
                         auto sh = kind == 4 ? std::vector<ShellData>{} : shellDataAux(s);
                         std::map<std::pair<size_t, size_t>, std::pair<int, size_t>> pairAt;
                         for (size_t c = 0; c < sp.size(); c++)
                             for (size_t p = 0; p < sp[c].n_pairs; p++)
                                 pairAt[{sp[c].shell_idxs[2 * p], sp[c].shell_idxs[2 * p + 1]}] = {int(c), p};
                         std::map<size_t, std::pair<int, size_t>> auxAt;
                         for (size_t c = 0; c < sh.size(); c++)
                             for (size_t p = 0; p < sh[c].n_shells; p++)
                                 auxAt[sh[c].shell_idxs[p]] = {int(c), p};
                         const int per = kind;
                         auto lOf = [&](int pos, int idx) {
                             const bool aux = kind == 2 || (kind == 3 && pos == 2);
                             const auto &v = aux ? s.shells_aux : s.shells;
                             if (idx < 0 || static_cast<size_t>(idx) >= v.size())
                                 throw std::runtime_error("orc_blocks: shell index out of range");
                             return v[idx].l;
                         };
                         size_t bsize = 1;
                         for (int k = 0; k < per; k++)
                             bsize *= static_cast<size_t>(2 * lOf(k, shells[k]) + 1);
                         const size_t ncomp = which < 3 ? 1 : which < 6 ? (which == 3 ? 12 : which == 4 ? 9 : 6)
                                                                        : which == 6 ? 36 : 3;
                         const size_t stride = ncomp * bsize;
                         std::map<std::pair<int, int>, std::shared_ptr<void>> kernels;
                         auto kern = [&](auto tag, int ca, int cb, auto make) {
                             using K = typename decltype(tag)::type;
                             auto &slot = kernels[{ca, cb}];
                             if (!slot)
                                 slot = std::shared_ptr<void>(make(), [](void *q) { delete static_cast<K *>(q); });
                             return static_cast<K *>(slot.get());
                         };
                         for (long t = 0; t < n; t++)
                         {
                             const int *e = shells + per * t;
                             double *dst = out + t * stride;
                             std::fill(dst, dst + stride, 0.0);
                             std::vector<double> blk;
                             if (kind == 2)
                             {
                                 (void)lOf(0, e[0]);
                                 (void)lOf(1, e[1]);
                                 auto a = auxAt.at(e[0]), b = auxAt.at(e[1]);
                                 const ShellData &A = sh[a.first], &B = sh[b.first];
                                 if (which == 2)
                                     blk = (*kern(Tag<ERI2Kernel>{}, a.first, b.first,
                                                  [&] { return new ERI2Kernel(A, B); }))(a.second, b.second, A, B);
                                 else if (which == 5)
                                     blk = (*kern(Tag<ERI2D1Kernel>{}, a.first, b.first,
                                                  [&] { return new ERI2D1Kernel(A, B); }))(a.second, b.second, A, B);
                                 else
                                     blk = (*kern(Tag<ERI2D2Kernel>{}, a.first, b.first,
                                                  [&] { return new ERI2D2Kernel(A, B); }))(a.second, b.second, A, B);
                             }
                             else
                             {
                                 for (int k = 0; k < per; k++)
                                     (void)lOf(k, e[k]);
                                 auto ia = pairAt.find({static_cast<size_t>(e[0]), static_cast<size_t>(e[1])});
                                 if (ia == pairAt.end())
                                     continue;
                                 const ShellPairData &A = sp[ia->second.first];
                                 const size_t pa = ia->second.second;
                                 if (kind == 3)
                                 {
                                     auto c = auxAt.at(e[2]);
                                     const ShellData &C = sh[c.first];
                                     if (which == 1)
                                         blk = (*kern(Tag<ERI3Kernel>{}, ia->second.first, c.first,
                                                      [&] { return new ERI3Kernel(A, C); }))(pa, c.second, A, C);
                                     else if (which == 4)
                                         blk = (*kern(Tag<ERI3D1Kernel>{}, ia->second.first, c.first,
                                                      [&] { return new ERI3D1Kernel(A, C); }))(pa, c.second, A, C);
                                     else
                                         blk = (*kern(Tag<ERI3SOCKernel>{}, ia->second.first, c.first,
                                                      [&] { return new ERI3SOCKernel(A, C); }))(pa, c.second, A, C);
                                 }
                                 else
                                 {
                                     auto ic = pairAt.find({static_cast<size_t>(e[2]), static_cast<size_t>(e[3])});
                                     if (ic == pairAt.end())
                                         continue;
                                     const ShellPairData &C = sp[ic->second.first];
                                     const size_t pc = ic->second.second;
                                     if (which == 0)
                                         blk = (*kern(Tag<ERI4Kernel>{}, ia->second.first, ic->second.first,
                                                      [&] { return new ERI4Kernel(A, C); }))(pa, pc, A, C);
                                     else if (which == 3)
                                         blk = (*kern(Tag<ERI4D1Kernel>{}, ia->second.first, ic->second.first,
                                                      [&] { return new ERI4D1Kernel(A, C); }))(pa, pc, A, C);
                                     else
                                         blk = (*kern(Tag<ERI4SOCKernel>{}, ia->second.first, ic->second.first,
                                                      [&] { return new ERI4SOCKernel(A, C); }))(pa, pc, A, C);
                                 }
                             }
                             if (blk.size() != stride)
                                 throw std::runtime_error("orc_blocks: entries of different angular momenta");
                             std::copy(blk.begin(), blk.end(), dst);
                         }
                     });
    }

    // which 0 overlap, 1 kinetic energy, 2 attraction to the n point charges (x, y, z, q)
    int orc_oneel(void *h, int which, long n, const double *charges, double *out)
    {
        return guard([&]
                     {
                         const Structure &s = *static_cast<Structure *>(h);
                         std::vector<double> r;
                         if (which == 0)
                             r = overlap(s);
                         else if (which == 1)
                             r = kineticEnergy(s);
                         else
                         {
                             std::vector<std::array<double, 4>> c(n);
                             for (long i = 0; i < n; i++)
                                 c[i] = {charges[4 * i], charges[4 * i + 1], charges[4 * i + 2], charges[4 * i + 3]};
                             r = externalCharges(s, c);
                         }
                         std::copy(r.begin(), r.end(), out);
                     });
    }

    // stats[0] = screened unique quartets processed, stats[1] = algorithmic FLOPs
    int orc_jk(void *h, const double *D, double *J, double *K, double thresh, int stride,
               double *stats)
    {
        return guard([&]
                     {
                         auto st = jk(*static_cast<Structure *>(h), D, J, K, thresh, stride);
                         if (stats)
                         {
                             stats[0] = st.n_quartets;
                             stats[1] = st.flops;
                             stats[2] = st.seconds_setup;
                             stats[3] = st.seconds_jk;
                         }
                     });
    }

    int orc_jk_grad(void *h, const double *D, double *grad, double thresh)
    {
        return guard([&] { jkGradient(*static_cast<Structure *>(h), D, grad, thresh); });
    }
}
