// Oracle (test infrastructure only): Coulomb/exchange build, energy and nuclear gradient from
// the oracle's SHARK kernels.  The reference has no J/K code; these loops define the quantity:
//   J_mn = sum_kl (mn|kl) D_kl,  K_mk = sum_nl (mn|kl) D_nl,  E = 1/2 tr(DJ) - 1/4 tr(DK)
// over unique shell quartets (eri4.cpp:44-61 enumeration) with permutational weights, and a
// Schwarz bound Q_ab = sqrt(max |(ab|ab)|) from the diagonal quartets (eri4.cpp:152-174).
#include "oracle.hpp"

#include <algorithm>
#include <cmath>
#include <omp.h>

namespace orc
{
    namespace
    {
        double rFlops(int L)
        {
            double s = 0;
            for (int n = 0; n < L; n++)
                for (int m = 1; m <= L - n; m++)
                    s += (m + 1) * (m + 2) / 2;
            return 2 * s;
        }

        // Algorithmic FLOPs of one quartet, the same formula the product reports.
        double quartetFlops(int la, int lb, int lc, int ld, double kab, double kcd)
        {
            double Hab = nHerm(la + lb), Hcd = nHerm(lc + ld);
            double Sab = nSph(la) * nSph(lb), Scd = nSph(lc) * nSph(ld);
            return 2 * kab * kcd * Hab * Hcd * Scd + 2 * kab * Sab * Hab * Scd +
                   kab * kcd * (rFlops(la + lb + lc + ld) + Hab * Hcd) + 12 * Sab * Scd;
        }

        std::vector<std::vector<double>> schwarz(const std::vector<ShellPairData> &sp)
        {
            std::vector<std::vector<double>> Q(sp.size());
            for (size_t i = 0; i < sp.size(); i++)
            {
                const auto &A = sp[i];
                ERI4Kernel kern(A, A);
                Q[i].assign(A.n_pairs, 0.0);
                int nab = nSph(A.la) * nSph(A.lb);
#pragma omp parallel for schedule(dynamic)
                for (size_t p = 0; p < A.n_pairs; p++)
                {
                    auto b = kern(p, p, A, A);
                    double m = 0;
                    for (int k = 0; k < nab; k++)
                        m = std::max(m, std::fabs(b[k * nab + k]));
                    Q[i][p] = std::sqrt(m);
                }
            }
            return Q;
        }
    }

    JKStats jk(const Structure &s, const double *D, double *J, double *K, double thresh,
               int stride)
    {
        const double t0 = omp_get_wtime();
        auto sp = shellPairData(true, s);
        auto Q = schwarz(sp);
        const double t1 = omp_get_wtime();
        const size_t n = s.nbf;
        JKStats st;
        int nth = omp_get_max_threads();
        std::vector<std::vector<double>> Jt(nth, std::vector<double>(n * n, 0.0)),
                Kt(nth, std::vector<double>(n * n, 0.0));
        std::vector<double> cnt(nth, 0.0), fl(nth, 0.0);
        for (size_t iab = 0; iab < sp.size(); iab++)
            for (size_t icd = 0; icd <= iab; icd++)
            {
                const auto &A = sp[iab], &C = sp[icd];
                ERI4Kernel kern(A, C);
                const int na = nSph(A.la), nb = nSph(A.lb), nc = nSph(C.la), nd = nSph(C.lb);
#pragma omp parallel for schedule(dynamic)
                for (size_t p = 0; p < A.n_pairs; p++)
                {
                    if (stride > 1 && (p % stride) != 0)
                        continue;
                    int tid = omp_get_thread_num();
                    double *Jl = Jt[tid].data(), *Kl = Kt[tid].data();
                    size_t bound = (iab == icd) ? p + 1 : C.n_pairs;
                    for (size_t q = 0; q < bound; q++)
                    {
                        if (Q[iab][p] * Q[icd][q] < thresh)
                            continue;
                        auto I = kern(p, q, A, C);
                        double f = 1.0;
                        if (A.shell_idxs[2 * p] == A.shell_idxs[2 * p + 1])
                            f *= 0.5;
                        if (C.shell_idxs[2 * q] == C.shell_idxs[2 * q + 1])
                            f *= 0.5;
                        if (iab == icd && p == q)
                            f *= 0.5;
                        size_t oa = A.offsets_sph[2 * p], ob = A.offsets_sph[2 * p + 1];
                        size_t oc = C.offsets_sph[2 * q], od = C.offsets_sph[2 * q + 1];
                        for (int ia = 0; ia < na; ia++)
                            for (int ib = 0; ib < nb; ib++)
                                for (int ic = 0; ic < nc; ic++)
                                    for (int id = 0; id < nd; id++)
                                    {
                                        double v = f * I[((ia * nb + ib) * nc + ic) * nd + id];
                                        size_t mu = oa + ia, nu = ob + ib, ka = oc + ic, la = od + id;
                                        Jl[mu * n + nu] += 2 * v * D[ka * n + la];
                                        Jl[ka * n + la] += 2 * v * D[mu * n + nu];
                                        Kl[mu * n + ka] += v * D[nu * n + la];
                                        Kl[nu * n + ka] += v * D[mu * n + la];
                                        Kl[mu * n + la] += v * D[nu * n + ka];
                                        Kl[nu * n + la] += v * D[mu * n + ka];
                                    }
                        cnt[tid] += 1;
                        fl[tid] += quartetFlops(A.la, A.lb, C.la, C.lb, A.nrs_ppairs[p],
                                                C.nrs_ppairs[q]);
                    }
                }
            }
        st.seconds_setup = t1 - t0;
        st.seconds_jk = omp_get_wtime() - t1;
        for (size_t i = 0; i < n * n; i++)
        {
            J[i] = 0;
            K[i] = 0;
        }
        for (int t = 0; t < nth; t++)
        {
            for (size_t i = 0; i < n * n; i++)
            {
                J[i] += Jt[t][i];
                K[i] += Kt[t][i];
            }
            st.n_quartets += cnt[t];
            st.flops += fl[t];
        }
        for (size_t i = 0; i < n; i++)
            for (size_t j = 0; j <= i; j++)
            {
                double jj = J[i * n + j] + J[j * n + i], kk = K[i * n + j] + K[j * n + i];
                J[i * n + j] = J[j * n + i] = jj;
                K[i * n + j] = K[j * n + i] = kk;
            }
        return st;
    }

    JKStats jkGradient(const Structure &s, const double *D, double *grad, double thresh)
    {
        auto sp = shellPairData(true, s);
        auto Q = schwarz(sp);
        const size_t n = s.nbf;
        JKStats st;
        int nth = omp_get_max_threads();
        std::vector<std::vector<double>> gt(nth, std::vector<double>(3 * s.natoms, 0.0));
        for (size_t iab = 0; iab < sp.size(); iab++)
            for (size_t icd = 0; icd <= iab; icd++)
            {
                const auto &A = sp[iab], &C = sp[icd];
                ERI4D1Kernel kern(A, C);
                const int na = nSph(A.la), nb = nSph(A.lb), nc = nSph(C.la), nd = nSph(C.lb);
                const size_t nblk = static_cast<size_t>(na) * nb * nc * nd;
#pragma omp parallel for schedule(dynamic)
                for (size_t p = 0; p < A.n_pairs; p++)
                {
                    int tid = omp_get_thread_num();
                    size_t bound = (iab == icd) ? p + 1 : C.n_pairs;
                    for (size_t q = 0; q < bound; q++)
                    {
                        if (Q[iab][p] * Q[icd][q] < thresh)
                            continue;
                        auto G = kern(p, q, A, C);
                        double f = 1.0;
                        if (A.shell_idxs[2 * p] == A.shell_idxs[2 * p + 1])
                            f *= 0.5;
                        if (C.shell_idxs[2 * q] == C.shell_idxs[2 * q + 1])
                            f *= 0.5;
                        if (iab == icd && p == q)
                            f *= 0.5;
                        size_t oa = A.offsets_sph[2 * p], ob = A.offsets_sph[2 * p + 1];
                        size_t oc = C.offsets_sph[2 * q], od = C.offsets_sph[2 * q + 1];
                        size_t atoms[4] = {A.atomic_idxs[2 * p], A.atomic_idxs[2 * p + 1],
                                           C.atomic_idxs[2 * q], C.atomic_idxs[2 * q + 1]};
                        double acc[12] = {0};
                        size_t e = 0;
                        for (int ia = 0; ia < na; ia++)
                            for (int ib = 0; ib < nb; ib++)
                                for (int ic = 0; ic < nc; ic++)
                                    for (int id = 0; id < nd; id++, e++)
                                    {
                                        size_t mu = oa + ia, nu = ob + ib, ka = oc + ic, la = od + id;
                                        double w = f * (4 * D[mu * n + nu] * D[ka * n + la] -
                                                        D[mu * n + ka] * D[nu * n + la] -
                                                        D[mu * n + la] * D[nu * n + ka]);
                                        for (int k = 0; k < 12; k++)
                                            acc[k] += w * G[k * nblk + e];
                                    }
                        for (int c = 0; c < 4; c++)
                            for (int x = 0; x < 3; x++)
                                gt[tid][3 * atoms[c] + x] += acc[3 * c + x];
                    }
                }
            }
        for (size_t i = 0; i < 3 * s.natoms; i++)
        {
            grad[i] = 0;
            for (int t = 0; t < nth; t++)
                grad[i] += gt[t][i];
        }
        return st;
    }
}
