// Oracle (test infrastructure only): SHARK ERI kernels (4-, 3-, 2-centre, plain and first
// derivative) and the dense drivers, restated from the reference generic-L code paths.
#include "oracle.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace orc
{
    namespace
    {
        // Hermite R matrix R(tuv, t'u'v') = (-1)^{t'+u'+v'} fac R_{t+t',u+u',v+v'} with an optional
        // shift of the summed index (rints.cpp:116-144 for no shift; rints.cpp:187-230 for the
        // +x/+y/+z shifts used by the derivative kernels).
        void rMatrix(const std::vector<double> &r3, int dim, double fac,
                     const std::vector<std::array<int, 3>> &hb,
                     const std::vector<std::array<int, 3>> &hk, int sx, int sy, int sz,
                     double *out)
        {
            size_t nk = hk.size();
            for (size_t j = 0; j < nk; j++)
            {
                auto [t_, u_, v_] = hk[j];
                double sign = ((t_ + u_ + v_) % 2 != 0) ? -1.0 : 1.0;
                for (size_t i = 0; i < hb.size(); i++)
                {
                    auto [t, u, v] = hb[i];
                    out[i * nk + j] = sign * fac *
                                      r3[((static_cast<size_t>(t + t_ + sx)) * dim + (u + u_ + sy)) *
                                                 dim + (v + v_ + sz)];
                }
            }
        }

        // C[m x n] += A[m x k] B[k x n]  (shark_mm_kernels.cpp:11-23, row-major dgemm, beta = 1)
        void gemmAcc(int m, int n, int k, const double *A, const double *B, double *C)
        {
            for (int i = 0; i < m; i++)
                for (int p = 0; p < k; p++)
                {
                    double a = A[i * k + p];
                    for (int j = 0; j < n; j++)
                        C[i * n + j] += a * B[p * n + j];
                }
        }

        void axpy(size_t n, double a, const double *x, double *y)
        {
            for (size_t i = 0; i < n; i++)
                y[i] += a * x[i];
        }

        const double k2pi52 = 2.0 * std::pow(M_PI, 2.5);
    }

    // ---------------------------------------------------------------- ERI4
    // eri_kernels.cpp:1101-1120
    ERI4Kernel::ERI4Kernel(const ShellPairData &ab, const ShellPairData &cd)
        : e_bra(ecoeffsSHARK(ab, false)), e_ket(ecoeffsSHARK(cd, true)),
          grid(ab.la + ab.lb + cd.la + cd.lb)
    {
    }

    // eri_kernel_funs.cpp:7-101
    std::vector<double> ERI4Kernel::operator()(size_t ipab, size_t ipcd, const ShellPairData &ab,
                                               const ShellPairData &cd) const
    {
        int lab = ab.la + ab.lb, lcd = cd.la + cd.lb, L = lab + lcd;
        int nsab = nSph(ab.la) * nSph(ab.lb), nscd = nSph(cd.la) * nSph(cd.lb);
        int nhab = nHerm(lab), nhcd = nHerm(lcd);
        int neab = nsab * nhab, necd = nscd * nhcd;
        const double *eab = &e_bra[ab.offsets_ecoeffs[ipab]];
        const double *ecd = &e_ket[cd.offsets_ecoeffs[ipcd]];
        const double *exab = &ab.exps[ab.offsets_prims[ipab]];
        const double *excd = &cd.exps[cd.offsets_prims[ipcd]];
        const double *xa = &ab.coords[6 * ipab], *xb = &ab.coords[6 * ipab + 3];
        const double *xc = &cd.coords[6 * ipcd], *xd = &cd.coords[6 * ipcd + 3];
        auto hb = hermiteIdxs(lab), hk = hermiteIdxs(lcd);
        std::vector<double> out(static_cast<size_t>(nsab) * nscd, 0.0);
        std::vector<double> fnx(L + 1), R(static_cast<size_t>(nhab) * nhcd);
        for (size_t iab = 0; iab < ab.nrs_ppairs[ipab]; iab++)
        {
            std::vector<double> RxE(static_cast<size_t>(nhab) * nscd, 0.0);
            for (size_t icd = 0; icd < cd.nrs_ppairs[ipcd]; icd++)
            {
                double a = exab[2 * iab], b = exab[2 * iab + 1];
                double c = excd[2 * icd], d = excd[2 * icd + 1];
                double p = a + b, q = c + d, alpha = p * q / (p + q);
                double P[3], Q[3], PQ[3];
                for (int i = 0; i < 3; i++)
                {
                    P[i] = (a * xa[i] + b * xb[i]) / p;
                    Q[i] = (c * xc[i] + d * xd[i]) / q;
                }
                for (int i = 0; i < 3; i++)
                    PQ[i] = P[i] - Q[i];
                double x = alpha * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
                calcBoysF(L, x, grid, fnx.data());
                double fac = k2pi52 / (p * q * std::sqrt(p + q));
                auto r3 = calcRInts3D(L, alpha, PQ, fnx.data());
                rMatrix(r3, L + 1, fac, hb, hk, 0, 0, 0, R.data());
                gemmAcc(nhab, nscd, nhcd, R.data(), &ecd[icd * necd], RxE.data());
            }
            gemmAcc(nsab, nscd, nhab, &eab[iab * neab], RxE.data(), out.data());
        }
        return out;
    }

    // ---------------------------------------------------------------- ERI3
    // eri_kernels.cpp:1122-1141
    ERI3Kernel::ERI3Kernel(const ShellPairData &ab, const ShellData &c)
        : e_bra(ecoeffsSHARK(ab, false)), e_ket(ecoeffsSHARK(c, true)), grid(ab.la + ab.lb + c.l)
    {
    }

    // eri_kernel_funs.cpp:103-184
    std::vector<double> ERI3Kernel::operator()(size_t ipab, size_t ishc, const ShellPairData &ab,
                                               const ShellData &c) const
    {
        int lab = ab.la + ab.lb, L = lab + c.l;
        int nsab = nSph(ab.la) * nSph(ab.lb), nsc = nSph(c.l);
        int nhab = nHerm(lab), nhc = nHerm(c.l);
        int neab = nsab * nhab, nec = nsc * nhc;
        size_t Kc = c.cdepths[ishc], co = c.coffsets[ishc];
        const double *eab = &e_bra[ab.offsets_ecoeffs[ipab]];
        const double *ec = &e_ket[c.offsets_ecoeffs[ishc]];
        const double *ex = &ab.exps[ab.offsets_prims[ipab]];
        const double *exc = &c.exps[co];
        const double *xa = &ab.coords[6 * ipab], *xb = &ab.coords[6 * ipab + 3];
        const double *xc = &c.coords[3 * ishc];
        auto hb = hermiteIdxs(lab), hk = hermiteIdxs(c.l);
        std::vector<double> out(static_cast<size_t>(nsab) * nsc, 0.0);
        std::vector<double> fnx(L + 1), R(static_cast<size_t>(nhab) * nhc);
        for (size_t iab = 0; iab < ab.nrs_ppairs[ipab]; iab++)
        {
            std::vector<double> RxE(static_cast<size_t>(nhab) * nsc, 0.0);
            for (size_t ic = 0; ic < Kc; ic++)
            {
                double a = ex[2 * iab], b = ex[2 * iab + 1], cc = exc[ic];
                double p = a + b, alpha = p * cc / (p + cc);
                double PC[3];
                for (int i = 0; i < 3; i++)
                    PC[i] = (a * xa[i] + b * xb[i]) / p - xc[i];
                double x = alpha * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
                calcBoysF(L, x, grid, fnx.data());
                double fac = k2pi52 / (p * cc * std::sqrt(p + cc));
                auto r3 = calcRInts3D(L, alpha, PC, fnx.data());
                rMatrix(r3, L + 1, fac, hb, hk, 0, 0, 0, R.data());
                gemmAcc(nhab, nsc, nhc, R.data(), &ec[ic * nec], RxE.data());
            }
            gemmAcc(nsab, nsc, nhab, &eab[iab * neab], RxE.data(), out.data());
        }
        return out;
    }

    // ---------------------------------------------------------------- ERI2
    // eri_kernels.cpp:1143-1162
    ERI2Kernel::ERI2Kernel(const ShellData &a, const ShellData &b)
        : e_bra(ecoeffsSHARK(a, false)), e_ket(ecoeffsSHARK(b, true)), grid(a.l + b.l)
    {
    }

    // eri_kernel_funs.cpp:186-255
    std::vector<double> ERI2Kernel::operator()(size_t isa, size_t isb, const ShellData &A,
                                               const ShellData &B) const
    {
        int la = A.l, lb = B.l, L = la + lb;
        int nsa = nSph(la), nsb = nSph(lb), nha = nHerm(la), nhb = nHerm(lb);
        int nea = nsa * nha, neb = nsb * nhb;
        size_t Ka = A.cdepths[isa], Kb = B.cdepths[isb];
        const double *ea = &e_bra[A.offsets_ecoeffs[isa]];
        const double *eb = &e_ket[B.offsets_ecoeffs[isb]];
        const double *exa = &A.exps[A.coffsets[isa]], *exb = &B.exps[B.coffsets[isb]];
        const double *xa = &A.coords[3 * isa], *xb = &B.coords[3 * isb];
        auto hb = hermiteIdxs(la), hk = hermiteIdxs(lb);
        double AB[3] = {xa[0] - xb[0], xa[1] - xb[1], xa[2] - xb[2]};
        double r2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
        std::vector<double> out(static_cast<size_t>(nsa) * nsb, 0.0);
        std::vector<double> fnx(L + 1), R(static_cast<size_t>(nha) * nhb);
        for (size_t ia = 0; ia < Ka; ia++)
        {
            std::vector<double> RxE(static_cast<size_t>(nha) * nsb, 0.0);
            for (size_t ib = 0; ib < Kb; ib++)
            {
                double a = exa[ia], b = exb[ib];
                double alpha = a * b / (a + b);
                calcBoysF(L, alpha * r2, grid, fnx.data());
                double fac = k2pi52 / (a * b * std::sqrt(a + b));
                auto r3 = calcRInts3D(L, alpha, AB, fnx.data());
                rMatrix(r3, L + 1, fac, hb, hk, 0, 0, 0, R.data());
                gemmAcc(nha, nsb, nhb, R.data(), &eb[ib * neb], RxE.data());
            }
            gemmAcc(nsa, nsb, nha, &ea[ia * nea], RxE.data(), out.data());
        }
        return out;
    }

    // ---------------------------------------------------------------- ERI4 D1
    // eri_kernels.cpp:1164-1185
    ERI4D1Kernel::ERI4D1Kernel(const ShellPairData &ab, const ShellPairData &cd)
        : e0_bra(ecoeffsSHARK(ab, false)), e1_bra(ecoeffsD1SHARK(ab, false)),
          e0_ket(ecoeffsSHARK(cd, true)), e1_ket(ecoeffsD1SHARK(cd, true)),
          grid(ab.la + ab.lb + cd.la + cd.lb + 1)
    {
    }

    // eri_kernel_funs.cpp:582-748
    std::vector<double> ERI4D1Kernel::operator()(size_t ipab, size_t ipcd,
                                                 const ShellPairData &ab,
                                                 const ShellPairData &cd) const
    {
        int lab = ab.la + ab.lb, lcd = cd.la + cd.lb, L = lab + lcd;
        int nsab = nSph(ab.la) * nSph(ab.lb), nscd = nSph(cd.la) * nSph(cd.lb);
        size_t nsabcd = static_cast<size_t>(nsab) * nscd;
        int nhab = nHerm(lab), nhcd = nHerm(lcd);
        size_t nr = static_cast<size_t>(nhab) * nhcd;
        int neab = nsab * nhab, necd = nscd * nhcd;
        size_t nrxe = static_cast<size_t>(nhab) * nscd;
        const double *ex_ab = &ab.exps[ab.offsets_prims[ipab]];
        const double *ex_cd = &cd.exps[cd.offsets_prims[ipcd]];
        const double *xa = &ab.coords[6 * ipab], *xb = &ab.coords[6 * ipab + 3];
        const double *xc = &cd.coords[6 * ipcd], *xd = &cd.coords[6 * ipcd + 3];
        const double *e0ab = &e0_bra[ab.offsets_ecoeffs[ipab]];
        const double *e1ab = &e1_bra[ab.offsets_ecoeffs_d1[ipab]];
        const double *e0cd = &e0_ket[cd.offsets_ecoeffs[ipcd]];
        const double *e1cd = &e1_ket[cd.offsets_ecoeffs_d1[ipcd]];
        auto hb = hermiteIdxs(lab), hk = hermiteIdxs(lcd);
        std::vector<double> out(12 * nsabcd, 0.0);
        std::vector<double> fnx(L + 2), R(4 * nr);
        for (size_t iab = 0; iab < ab.nrs_ppairs[ipab]; iab++)
        {
            double a = ex_ab[2 * iab], b = ex_ab[2 * iab + 1], p = a + b;
            std::vector<double> RxE(7 * nrxe, 0.0);
            for (size_t icd = 0; icd < cd.nrs_ppairs[ipcd]; icd++)
            {
                double c = ex_cd[2 * icd], d = ex_cd[2 * icd + 1];
                double q = c + d, alpha = p * q / (p + q);
                double PQ[3];
                for (int i = 0; i < 3; i++)
                    PQ[i] = (a * xa[i] + b * xb[i]) / p - (c * xc[i] + d * xd[i]) / q;
                double x = alpha * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
                calcBoysF(L + 1, x, grid, fnx.data());
                double fac = k2pi52 / (p * q * std::sqrt(p + q));
                auto r3 = calcRInts3D(L + 1, alpha, PQ, fnx.data());
                // calcRInts_ERI3D1 (rints.cpp:187-230): +x, +y, +z shifted, unshifted
                rMatrix(r3, L + 2, fac, hb, hk, 1, 0, 0, &R[0 * nr]);
                rMatrix(r3, L + 2, fac, hb, hk, 0, 1, 0, &R[1 * nr]);
                rMatrix(r3, L + 2, fac, hb, hk, 0, 0, 1, &R[2 * nr]);
                rMatrix(r3, L + 2, fac, hb, hk, 0, 0, 0, &R[3 * nr]);
                std::vector<double> I(3 * nrxe, 0.0);
                const double *E0 = &e0cd[icd * necd];
                for (int k = 0; k < 3; k++)
                    gemmAcc(nhab, nscd, nhcd, &R[k * nr], E0, &I[k * nrxe]);
                gemmAcc(nhab, nscd, nhcd, &R[3 * nr], E0, &RxE[3 * nrxe]);
                for (int k = 0; k < 3; k++)
                    gemmAcc(nhab, nscd, nhcd, &R[3 * nr], &e1cd[(3 * icd + k) * necd],
                            &RxE[(4 + k) * nrxe]);
                axpy(3 * nrxe, 1.0, I.data(), &RxE[0]);
                axpy(3 * nrxe, -(c / q), I.data(), &RxE[4 * nrxe]);
            }
            std::vector<double> P(3 * nsabcd, 0.0), Rr(3 * nsabcd, 0.0);
            const double *E0 = &e0ab[iab * neab];
            for (int k = 0; k < 3; k++)
                gemmAcc(nsab, nscd, nhab, E0, &RxE[k * nrxe], &P[k * nsabcd]);
            for (int k = 0; k < 3; k++)
                gemmAcc(nsab, nscd, nhab, &e1ab[(3 * iab + k) * neab], &RxE[3 * nrxe],
                        &Rr[k * nsabcd]);
            for (int k = 0; k < 3; k++)
            {
                axpy(nsabcd, a / p, &P[k * nsabcd], &out[k * nsabcd]);
                axpy(nsabcd, 1.0, &Rr[k * nsabcd], &out[k * nsabcd]);
                axpy(nsabcd, b / p, &P[k * nsabcd], &out[(3 + k) * nsabcd]);
                axpy(nsabcd, -1.0, &Rr[k * nsabcd], &out[(3 + k) * nsabcd]);
                gemmAcc(nsab, nscd, nhab, E0, &RxE[(4 + k) * nrxe], &out[(6 + k) * nsabcd]);
            }
        }
        for (int k = 0; k < 3; k++)
            for (size_t i = 0; i < nsabcd; i++)
                out[(9 + k) * nsabcd + i] = -1 * (out[k * nsabcd + i] + out[(3 + k) * nsabcd + i] +
                                                   out[(6 + k) * nsabcd + i]);
        return out;
    }

    // ---------------------------------------------------------------- ERI3 D1
    // eri_kernels.cpp:1187-1206
    ERI3D1Kernel::ERI3D1Kernel(const ShellPairData &ab, const ShellData &c)
        : e0_bra(ecoeffsSHARK(ab, false)), e1_bra(ecoeffsD1SHARK(ab, false)),
          e0_ket(ecoeffsSHARK(c, true)), grid(ab.la + ab.lb + c.l + 1)
    {
    }

    // eri_kernel_funs.cpp:480-580
    std::vector<double> ERI3D1Kernel::operator()(size_t ipab, size_t ishc,
                                                 const ShellPairData &ab, const ShellData &C) const
    {
        int lab = ab.la + ab.lb, L = lab + C.l;
        int nsab = nSph(ab.la) * nSph(ab.lb), nsc = nSph(C.l);
        size_t nsabc = static_cast<size_t>(nsab) * nsc;
        int nhab = nHerm(lab), nhc = nHerm(C.l);
        size_t nr = static_cast<size_t>(nhab) * nhc, nrxe = static_cast<size_t>(nhab) * nsc;
        int neab = nsab * nhab, nec = nsc * nhc;
        size_t Kc = C.cdepths[ishc], co = C.coffsets[ishc];
        const double *ex_ab = &ab.exps[ab.offsets_prims[ipab]];
        const double *exc = &C.exps[co];
        const double *xa = &ab.coords[6 * ipab], *xb = &ab.coords[6 * ipab + 3];
        const double *xc = &C.coords[3 * ishc];
        const double *e0ab = &e0_bra[ab.offsets_ecoeffs[ipab]];
        const double *e1ab = &e1_bra[ab.offsets_ecoeffs_d1[ipab]];
        const double *e0c = &e0_ket[C.offsets_ecoeffs[ishc]];
        auto hb = hermiteIdxs(lab), hk = hermiteIdxs(C.l);
        std::vector<double> out(9 * nsabc, 0.0);
        std::vector<double> fnx(L + 2), R(4 * nr);
        for (size_t iab = 0; iab < ab.nrs_ppairs[ipab]; iab++)
        {
            double a = ex_ab[2 * iab], b = ex_ab[2 * iab + 1], p = a + b;
            std::vector<double> RxE(4 * nrxe, 0.0);
            for (size_t ic = 0; ic < Kc; ic++)
            {
                double c = exc[ic], alpha = p * c / (p + c);
                double PC[3];
                for (int i = 0; i < 3; i++)
                    PC[i] = (a * xa[i] + b * xb[i]) / p - xc[i];
                double x = alpha * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
                calcBoysF(L + 1, x, grid, fnx.data());
                double fac = k2pi52 / (p * c * std::sqrt(p + c));
                auto r3 = calcRInts3D(L + 1, alpha, PC, fnx.data());
                rMatrix(r3, L + 2, fac, hb, hk, 1, 0, 0, &R[0 * nr]);
                rMatrix(r3, L + 2, fac, hb, hk, 0, 1, 0, &R[1 * nr]);
                rMatrix(r3, L + 2, fac, hb, hk, 0, 0, 1, &R[2 * nr]);
                rMatrix(r3, L + 2, fac, hb, hk, 0, 0, 0, &R[3 * nr]);
                for (int k = 0; k < 4; k++)
                    gemmAcc(nhab, nsc, nhc, &R[k * nr], &e0c[ic * nec], &RxE[k * nrxe]);
            }
            std::vector<double> P(3 * nsabc, 0.0), Rr(3 * nsabc, 0.0);
            for (int k = 0; k < 3; k++)
                gemmAcc(nsab, nsc, nhab, &e0ab[iab * neab], &RxE[k * nrxe], &P[k * nsabc]);
            for (int k = 0; k < 3; k++)
                gemmAcc(nsab, nsc, nhab, &e1ab[(3 * iab + k) * neab], &RxE[3 * nrxe],
                        &Rr[k * nsabc]);
            for (int k = 0; k < 3; k++)
            {
                axpy(nsabc, a / p, &P[k * nsabc], &out[k * nsabc]);
                axpy(nsabc, 1.0, &Rr[k * nsabc], &out[k * nsabc]);
                axpy(nsabc, -1.0, &P[k * nsabc], &out[(6 + k) * nsabc]);
            }
        }
        for (int k = 0; k < 3; k++)
            for (size_t i = 0; i < nsabc; i++)
                out[(3 + k) * nsabc + i] = -1 * (out[k * nsabc + i] + out[(6 + k) * nsabc + i]);
        return out;
    }

    // ---------------------------------------------------------------- ERI2 D1
    // eri_kernels.cpp:1208-1228
    ERI2D1Kernel::ERI2D1Kernel(const ShellData &a, const ShellData &b)
        : e_bra(ecoeffsSHARK(a, false)), e_ket(ecoeffsSHARK(b, true)), grid(a.l + b.l + 1)
    {
    }

    // eri_kernel_funs.cpp:257-333
    std::vector<double> ERI2D1Kernel::operator()(size_t isa, size_t isb, const ShellData &A,
                                                 const ShellData &B) const
    {
        int la = A.l, lb = B.l, L = la + lb;
        int nsa = nSph(la), nsb = nSph(lb), nha = nHerm(la), nhb = nHerm(lb);
        int nea = nsa * nha, neb = nsb * nhb;
        size_t nr = static_cast<size_t>(nha) * nhb, nrxe = static_cast<size_t>(nha) * nsb;
        size_t nsab = static_cast<size_t>(nsa) * nsb;
        size_t Ka = A.cdepths[isa], Kb = B.cdepths[isb];
        const double *ea = &e_bra[A.offsets_ecoeffs[isa]];
        const double *eb = &e_ket[B.offsets_ecoeffs[isb]];
        const double *exa = &A.exps[A.coffsets[isa]], *exb = &B.exps[B.coffsets[isb]];
        const double *xa = &A.coords[3 * isa], *xb = &B.coords[3 * isb];
        auto hb = hermiteIdxs(la), hk = hermiteIdxs(lb);
        double AB[3] = {xa[0] - xb[0], xa[1] - xb[1], xa[2] - xb[2]};
        double r2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
        std::vector<double> out(6 * nsab, 0.0);
        std::vector<double> fnx(L + 2), R(3 * nr);
        for (size_t ia = 0; ia < Ka; ia++)
        {
            std::vector<double> RxE(3 * nrxe, 0.0);
            for (size_t ib = 0; ib < Kb; ib++)
            {
                double a = exa[ia], b = exb[ib], alpha = a * b / (a + b);
                calcBoysF(L + 1, alpha * r2, grid, fnx.data());
                double fac = k2pi52 / (a * b * std::sqrt(a + b));
                auto r3 = calcRInts3D(L + 1, alpha, AB, fnx.data());
                rMatrix(r3, L + 2, fac, hb, hk, 1, 0, 0, &R[0 * nr]);
                rMatrix(r3, L + 2, fac, hb, hk, 0, 1, 0, &R[1 * nr]);
                rMatrix(r3, L + 2, fac, hb, hk, 0, 0, 1, &R[2 * nr]);
                for (int k = 0; k < 3; k++)
                    gemmAcc(nha, nsb, nhb, &R[k * nr], &eb[ib * neb], &RxE[k * nrxe]);
            }
            for (int k = 0; k < 3; k++)
                gemmAcc(nsa, nsb, nha, &ea[ia * nea], &RxE[k * nrxe], &out[k * nsab]);
        }
        for (int k = 3; k < 6; k++)
            for (size_t i = 0; i < nsab; i++)
                out[k * nsab + i] = -1 * out[(k - 3) * nsab + i];
        return out;
    }

    // ---------------------------------------------------------------- ERI2 D2
    // eri_kernels.cpp:1230-1249
    ERI2D2Kernel::ERI2D2Kernel(const ShellData &a, const ShellData &b)
        : e_bra(ecoeffsSHARK(a, false)), e_ket(ecoeffsSHARK(b, true)), grid(a.l + b.l + 2)
    {
    }

    // eri_kernel_funs.cpp:343-433, R of rints.cpp:228-270: the six second-order shifts
    // 200, 110, 101, 020, 011, 002 of the Hermite sum; AA from them, AB = BA = -AA, BB = AA
    std::vector<double> ERI2D2Kernel::operator()(size_t isa, size_t isb, const ShellData &A,
                                                 const ShellData &B) const
    {
        int la = A.l, lb = B.l, L = la + lb;
        int nsa = nSph(la), nsb = nSph(lb), nha = nHerm(la), nhb = nHerm(lb);
        int nea = nsa * nha, neb = nsb * nhb;
        size_t nr = static_cast<size_t>(nha) * nhb, nrxe = static_cast<size_t>(nha) * nsb;
        size_t nsab = static_cast<size_t>(nsa) * nsb;
        size_t Ka = A.cdepths[isa], Kb = B.cdepths[isb];
        const double *ea = &e_bra[A.offsets_ecoeffs[isa]];
        const double *eb = &e_ket[B.offsets_ecoeffs[isb]];
        const double *exa = &A.exps[A.coffsets[isa]], *exb = &B.exps[B.coffsets[isb]];
        const double *xa = &A.coords[3 * isa], *xb = &B.coords[3 * isb];
        auto hb = hermiteIdxs(la), hk = hermiteIdxs(lb);
        double AB[3] = {xa[0] - xb[0], xa[1] - xb[1], xa[2] - xb[2]};
        double r2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
        static const int shifts[6][3] = {{2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}};
        static const int slot[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
        std::vector<double> aa(9 * nsab, 0.0);   // AA[i][j] for i <= j
        std::vector<double> fnx(L + 3), R(6 * nr);
        for (size_t ia = 0; ia < Ka; ia++)
        {
            std::vector<double> RxE(6 * nrxe, 0.0);
            for (size_t ib = 0; ib < Kb; ib++)
            {
                double a = exa[ia], b = exb[ib], alpha = a * b / (a + b);
                calcBoysF(L + 2, alpha * r2, grid, fnx.data());
                double fac = k2pi52 / (a * b * std::sqrt(a + b));
                auto r3 = calcRInts3D(L + 2, alpha, AB, fnx.data());
                for (int k = 0; k < 6; k++)
                {
                    rMatrix(r3, L + 3, fac, hb, hk, shifts[k][0], shifts[k][1], shifts[k][2], &R[k * nr]);
                    gemmAcc(nha, nsb, nhb, &R[k * nr], &eb[ib * neb], &RxE[k * nrxe]);
                }
            }
            for (int k = 0; k < 6; k++)
                gemmAcc(nsa, nsb, nha, &ea[ia * nea], &RxE[k * nrxe],
                        &aa[(3 * slot[k][0] + slot[k][1]) * nsab]);
        }
        for (int i = 0; i < 3; i++)
            for (int j = i + 1; j < 3; j++)
                std::copy_n(&aa[(3 * i + j) * nsab], nsab, &aa[(3 * j + i) * nsab]);
        std::vector<double> out(36 * nsab, 0.0);
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++)
            {
                const double sg = ((i < 3) == (j < 3)) ? 1.0 : -1.0;
                const double *src = &aa[(3 * (i % 3) + (j % 3)) * nsab];
                for (size_t e = 0; e < nsab; e++)
                    out[(6 * i + j) * nsab + e] = sg * src[e];
            }
        return out;
    }

    // ---------------------------------------------------------------- ERI4 SOC
    // eri_kernels.cpp:1251-1270
    ERI4SOCKernel::ERI4SOCKernel(const ShellPairData &ab, const ShellPairData &cd)
        : e1_bra(ecoeffsD1SHARK(ab, false)), e0_ket(ecoeffsSHARK(cd, true)),
          grid(ab.la + ab.lb + cd.la + cd.lb + 1)
    {
    }

    namespace
    {
        // The bra's E' (d/d(A-B), per axis) against the ket-transformed R shifted by one unit
        // along each axis (calcRInts_ERISOC, rints.cpp:272-308): component x = E'_y R_z - E'_z R_y,
        // y = E'_z R_x - E'_x R_z, z = E'_x R_y - E'_y R_x (eri_kernel_funs.cpp:932-971)
        void socBra(int nsab, int nscd, int nhab, const double *e1, size_t neab, const double *RxE,
                    size_t nrxe, size_t nsabcd, double *out)
        {
            static const int ax[3][2] = {{1, 2}, {2, 0}, {0, 1}};
            for (int k = 0; k < 3; k++)
            {
                std::vector<double> p(nsabcd, 0.0), m(nsabcd, 0.0);
                gemmAcc(nsab, nscd, nhab, &e1[ax[k][0] * neab], &RxE[ax[k][1] * nrxe], p.data());
                gemmAcc(nsab, nscd, nhab, &e1[ax[k][1] * neab], &RxE[ax[k][0] * nrxe], m.data());
                axpy(nsabcd, 1.0, p.data(), &out[k * nsabcd]);
                axpy(nsabcd, -1.0, m.data(), &out[k * nsabcd]);
            }
        }
    }

    // eri_kernel_funs.cpp:846-974
    std::vector<double> ERI4SOCKernel::operator()(size_t ipab, size_t ipcd, const ShellPairData &ab,
                                                  const ShellPairData &cd) const
    {
        int lab = ab.la + ab.lb, lcd = cd.la + cd.lb, L = lab + lcd;
        int nsab = nSph(ab.la) * nSph(ab.lb), nscd = nSph(cd.la) * nSph(cd.lb);
        size_t nsabcd = static_cast<size_t>(nsab) * nscd;
        int nhab = nHerm(lab), nhcd = nHerm(lcd);
        size_t nr = static_cast<size_t>(nhab) * nhcd, nrxe = static_cast<size_t>(nhab) * nscd;
        size_t neab = static_cast<size_t>(nsab) * nhab;
        int necd = nscd * nhcd;
        const double *ex_ab = &ab.exps[ab.offsets_prims[ipab]];
        const double *ex_cd = &cd.exps[cd.offsets_prims[ipcd]];
        const double *xa = &ab.coords[6 * ipab], *xb = &ab.coords[6 * ipab + 3];
        const double *xc = &cd.coords[6 * ipcd], *xd = &cd.coords[6 * ipcd + 3];
        const double *e1ab = &e1_bra[ab.offsets_ecoeffs_d1[ipab]];
        const double *e0cd = &e0_ket[cd.offsets_ecoeffs[ipcd]];
        auto hb = hermiteIdxs(lab), hk = hermiteIdxs(lcd);
        std::vector<double> out(3 * nsabcd, 0.0);
        std::vector<double> fnx(L + 2), R(3 * nr);
        for (size_t iab = 0; iab < ab.nrs_ppairs[ipab]; iab++)
        {
            double a = ex_ab[2 * iab], b = ex_ab[2 * iab + 1], p = a + b;
            std::vector<double> RxE(3 * nrxe, 0.0);
            for (size_t icd = 0; icd < cd.nrs_ppairs[ipcd]; icd++)
            {
                double c = ex_cd[2 * icd], d = ex_cd[2 * icd + 1];
                double q = c + d, alpha = p * q / (p + q);
                double PQ[3];
                for (int i = 0; i < 3; i++)
                    PQ[i] = (a * xa[i] + b * xb[i]) / p - (c * xc[i] + d * xd[i]) / q;
                double x = alpha * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
                calcBoysF(L + 1, x, grid, fnx.data());
                double fac = k2pi52 / (p * q * std::sqrt(p + q));
                auto r3 = calcRInts3D(L + 1, alpha, PQ, fnx.data());
                rMatrix(r3, L + 2, fac, hb, hk, 1, 0, 0, &R[0 * nr]);
                rMatrix(r3, L + 2, fac, hb, hk, 0, 1, 0, &R[1 * nr]);
                rMatrix(r3, L + 2, fac, hb, hk, 0, 0, 1, &R[2 * nr]);
                for (int k = 0; k < 3; k++)
                    gemmAcc(nhab, nscd, nhcd, &R[k * nr], &e0cd[icd * necd], &RxE[k * nrxe]);
            }
            socBra(nsab, nscd, nhab, &e1ab[3 * iab * neab], neab, RxE.data(), nrxe, nsabcd, out.data());
        }
        return out;
    }

    // ---------------------------------------------------------------- ERI3 SOC
    // eri_kernels.cpp:1272-1291
    ERI3SOCKernel::ERI3SOCKernel(const ShellPairData &ab, const ShellData &c)
        : e1_bra(ecoeffsD1SHARK(ab, false)), e0_ket(ecoeffsSHARK(c, true)), grid(ab.la + ab.lb + c.l + 1)
    {
    }

    // eri_kernel_funs.cpp:976-1090
    std::vector<double> ERI3SOCKernel::operator()(size_t ipab, size_t ishc, const ShellPairData &ab,
                                                  const ShellData &C) const
    {
        int lab = ab.la + ab.lb, L = lab + C.l;
        int nsab = nSph(ab.la) * nSph(ab.lb), nsc = nSph(C.l);
        size_t nsabc = static_cast<size_t>(nsab) * nsc;
        int nhab = nHerm(lab), nhc = nHerm(C.l);
        size_t nr = static_cast<size_t>(nhab) * nhc, nrxe = static_cast<size_t>(nhab) * nsc;
        size_t neab = static_cast<size_t>(nsab) * nhab;
        int nec = nsc * nhc;
        size_t Kc = C.cdepths[ishc], co = C.coffsets[ishc];
        const double *ex_ab = &ab.exps[ab.offsets_prims[ipab]];
        const double *exc = &C.exps[co];
        const double *xa = &ab.coords[6 * ipab], *xb = &ab.coords[6 * ipab + 3];
        const double *xc = &C.coords[3 * ishc];
        const double *e1ab = &e1_bra[ab.offsets_ecoeffs_d1[ipab]];
        const double *e0c = &e0_ket[C.offsets_ecoeffs[ishc]];
        auto hb = hermiteIdxs(lab), hk = hermiteIdxs(C.l);
        std::vector<double> out(3 * nsabc, 0.0);
        std::vector<double> fnx(L + 2), R(3 * nr);
        for (size_t iab = 0; iab < ab.nrs_ppairs[ipab]; iab++)
        {
            double a = ex_ab[2 * iab], b = ex_ab[2 * iab + 1], p = a + b;
            std::vector<double> RxE(3 * nrxe, 0.0);
            for (size_t ic = 0; ic < Kc; ic++)
            {
                double c = exc[ic], alpha = p * c / (p + c);
                double PC[3];
                for (int i = 0; i < 3; i++)
                    PC[i] = (a * xa[i] + b * xb[i]) / p - xc[i];
                double x = alpha * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
                calcBoysF(L + 1, x, grid, fnx.data());
                double fac = k2pi52 / (p * c * std::sqrt(p + c));
                auto r3 = calcRInts3D(L + 1, alpha, PC, fnx.data());
                rMatrix(r3, L + 2, fac, hb, hk, 1, 0, 0, &R[0 * nr]);
                rMatrix(r3, L + 2, fac, hb, hk, 0, 1, 0, &R[1 * nr]);
                rMatrix(r3, L + 2, fac, hb, hk, 0, 0, 1, &R[2 * nr]);
                for (int k = 0; k < 3; k++)
                    gemmAcc(nhab, nsc, nhc, &R[k * nr], &e0c[ic * nec], &RxE[k * nrxe]);
            }
            socBra(nsab, nsc, nhab, &e1ab[3 * iab * neab], neab, RxE.data(), nrxe, nsabc, out.data());
        }
        return out;
    }

    // ---------------------------------------------------------------- drivers
    // eri4.cpp:38-95
    std::vector<double> eri4(const Structure &s)
    {
        auto sp = shellPairData(true, s);
        size_t n = s.nbf;
        std::vector<double> out(n * n * n * n, 0.0);
        auto at = [n](size_t i, size_t j, size_t k, size_t l) { return ((i * n + j) * n + k) * n + l; };
        for (size_t iab = 0; iab < sp.size(); iab++)
            for (size_t icd = 0; icd <= iab; icd++)
            {
                const auto &A = sp[iab], &C = sp[icd];
                int na = nSph(A.la), nb = nSph(A.lb), nc = nSph(C.la), nd = nSph(C.lb);
                ERI4Kernel kern(A, C);
#pragma omp parallel for schedule(dynamic)
                for (size_t p = 0; p < A.n_pairs; p++)
                {
                    size_t bound = (iab == icd) ? p + 1 : C.n_pairs;
                    for (size_t q = 0; q < bound; q++)
                    {
                        auto b = kern(p, q, A, C);
                        size_t oa = A.offsets_sph[2 * p], ob = A.offsets_sph[2 * p + 1];
                        size_t oc = C.offsets_sph[2 * q], od = C.offsets_sph[2 * q + 1];
                        for (int ia = 0; ia < na; ia++)
                            for (int ib = 0; ib < nb; ib++)
                                for (int ic = 0; ic < nc; ic++)
                                    for (int id = 0; id < nd; id++)
                                    {
                                        size_t mu = oa + ia, nu = ob + ib, ka = oc + ic, ta = od + id;
                                        double v = b[((ia * nb + ib) * nc + ic) * nd + id];
                                        out[at(mu, nu, ka, ta)] = v;
                                        out[at(mu, nu, ta, ka)] = v;
                                        out[at(nu, mu, ka, ta)] = v;
                                        out[at(nu, mu, ta, ka)] = v;
                                        out[at(ka, ta, mu, nu)] = v;
                                        out[at(ka, ta, nu, mu)] = v;
                                        out[at(ta, ka, mu, nu)] = v;
                                        out[at(ta, ka, nu, mu)] = v;
                                    }
                    }
                }
            }
        return out;
    }

    // eri4.cpp:152-174 and transferIntsERI4Diag (eri4.cpp:15-36)
    std::vector<double> eri4Diagonal(const Structure &s)
    {
        auto sp = shellPairData(true, s);
        size_t n = s.nbf;
        std::vector<double> out(n * n, 0.0);
        for (const auto &A : sp)
        {
            ERI4Kernel kern(A, A);
            int na = nSph(A.la), nb = nSph(A.lb), nab = na * nb;
#pragma omp parallel for schedule(dynamic)
            for (size_t p = 0; p < A.n_pairs; p++)
            {
                auto b = kern(p, p, A, A);
                size_t oa = A.offsets_sph[2 * p], ob = A.offsets_sph[2 * p + 1];
                for (int ia = 0; ia < na; ia++)
                    for (int ib = 0; ib < nb; ib++)
                    {
                        double v = b[(ia * nb + ib) * nab + ia * nb + ib];
                        out[(oa + ia) * n + ob + ib] = v;
                        out[(ob + ib) * n + oa + ia] = v;
                    }
            }
        }
        return out;
    }

    // eri3.cpp:7-46
    std::vector<double> eri3(const Structure &s)
    {
        auto shs = shellDataAux(s);
        auto sp = shellPairData(true, s);
        size_t n = s.nbf, nx = s.nbf_aux;
        std::vector<double> out(n * n * nx, 0.0);
        for (const auto &A : sp)
            for (const auto &C : shs)
            {
                ERI3Kernel kern(A, C);
                int na = nSph(A.la), nb = nSph(A.lb), nc = nSph(C.l);
#pragma omp parallel for schedule(dynamic)
                for (size_t p = 0; p < A.n_pairs; p++)
                    for (size_t q = 0; q < C.n_shells; q++)
                    {
                        auto b = kern(p, q, A, C);
                        size_t oa = A.offsets_sph[2 * p], ob = A.offsets_sph[2 * p + 1];
                        size_t oc = C.offsets_sph[q];
                        for (int ia = 0; ia < na; ia++)
                            for (int ib = 0; ib < nb; ib++)
                                for (int ic = 0; ic < nc; ic++)
                                {
                                    double v = b[(ia * nb + ib) * nc + ic];
                                    out[((oa + ia) * n + ob + ib) * nx + oc + ic] = v;
                                    out[((ob + ib) * n + oa + ia) * nx + oc + ic] = v;
                                }
                    }
            }
        return out;
    }

    // eri2.cpp:27-72
    std::vector<double> eri2(const Structure &s)
    {
        auto shs = shellDataAux(s);
        size_t nx = s.nbf_aux;
        std::vector<double> out(nx * nx, 0.0);
        for (size_t ia = 0; ia < shs.size(); ia++)
            for (size_t ib = 0; ib <= ia; ib++)
            {
                const auto &A = shs[ia], &B = shs[ib];
                ERI2Kernel kern(A, B);
                int na = nSph(A.l), nb = nSph(B.l);
#pragma omp parallel for schedule(dynamic)
                for (size_t p = 0; p < A.n_shells; p++)
                {
                    size_t bound = (A.l == B.l) ? p + 1 : B.n_shells;
                    for (size_t q = 0; q < bound; q++)
                    {
                        auto b = kern(p, q, A, B);
                        size_t oa = A.offsets_sph[p], ob = B.offsets_sph[q];
                        for (int i = 0; i < na; i++)
                            for (int j = 0; j < nb; j++)
                            {
                                out[(oa + i) * nx + ob + j] = b[i * nb + j];
                                out[(ob + j) * nx + oa + i] = b[i * nb + j];
                            }
                    }
                }
            }
        return out;
    }

    // eri2.cpp:74-96
    std::vector<double> eri2Diagonal(const Structure &s)
    {
        auto shs = shellDataAux(s);
        std::vector<double> out(s.nbf_aux, 0.0);
        for (const auto &A : shs)
        {
            ERI2Kernel kern(A, A);
            int na = nSph(A.l);
            for (size_t p = 0; p < A.n_shells; p++)
            {
                auto b = kern(p, p, A, A);
                for (int i = 0; i < na; i++)
                    out[A.offsets_sph[p] + i] = b[i * na + i];
            }
        }
        return out;
    }

    // The checksums of eri3() and eri2() without the dense tensors: a block of shell pair (ab)
    // with a != b appears twice in the dense eri3 tensor ((ab|P) and (ba|P)), a block (P|Q) of
    // distinct shells twice in eri2 ((P|Q) and (Q|P)); diagonal blocks once.
    void riSums(const Structure &s, double out[4])
    {
        auto shs = shellDataAux(s);
        auto sp = shellPairData(true, s);
        double a3 = 0, s3 = 0, a2 = 0, s2 = 0;
        for (const auto &A : sp)
            for (const auto &C : shs)
            {
                ERI3Kernel kern(A, C);
#pragma omp parallel for schedule(dynamic) reduction(+ : a3, s3)
                for (size_t p = 0; p < A.n_pairs; p++)
                {
                    const double w = A.shell_idxs[2 * p] == A.shell_idxs[2 * p + 1] ? 1.0 : 2.0;
                    for (size_t q = 0; q < C.n_shells; q++)
                    {
                        auto b = kern(p, q, A, C);
                        double ba = 0, bs = 0;
                        for (double v : b)
                        {
                            ba += std::fabs(v);
                            bs += v;
                        }
                        a3 += w * ba;
                        s3 += w * bs;
                    }
                }
            }
        for (size_t ia = 0; ia < shs.size(); ia++)
            for (size_t ib = 0; ib <= ia; ib++)
            {
                const auto &A = shs[ia], &B = shs[ib];
                ERI2Kernel kern(A, B);
#pragma omp parallel for schedule(dynamic) reduction(+ : a2, s2)
                for (size_t p = 0; p < A.n_shells; p++)
                {
                    const size_t bound = (A.l == B.l) ? p + 1 : B.n_shells;
                    for (size_t q = 0; q < bound; q++)
                    {
                        const double w = (A.l == B.l && p == q) ? 1.0 : 2.0;
                        auto b = kern(p, q, A, B);
                        double ba = 0, bs = 0;
                        for (double v : b)
                        {
                            ba += std::fabs(v);
                            bs += v;
                        }
                        a2 += w * ba;
                        s2 += w * bs;
                    }
                }
            }
        out[0] = a3;
        out[1] = s3;
        out[2] = a2;
        out[3] = s2;
    }
}
