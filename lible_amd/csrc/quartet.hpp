// SHARK shell-quartet kernels for gfx950.
//
// Per shell quartet (ab|cd) of class (LA LB | LC LD), following the reference's SHARK scheme
// (eri_kernel_funs.hpp:104-193, after doi:10.1002/jcc.26942 eqs. 25-26):
//   for each bra primitive pair i:  X_i = sum_k R_{ik} . Ek_k          (ket transform)
//   I = sum_i Eb_i^T X_i                                               (bra transform)
// with R_{ik}(tuv, t'u'v') = fac R_{t+t',u+u',v+v'} (Hermite integrals, sign of the ket folded
// into Ek) and Eb, Ek the spherical Hermite-expansion matrices of the primitive pairs.
//
// Two kernels:
//   quartetV  one lane per quartet, everything in registers (small classes).  A wave holds 64
//             bra pairs against one ket pair, so every ket operand is a wave-uniform (scalar)
//             load and the transforms are VALU FMAs with an SGPR operand.
//   quartetM  one wave per quartet, both transforms on FP64 MFMA (v_mfma_f64_16x16x4f64).  The
//             bra primitive pairs are stacked along M, (prim, tuv) rows, so one GEMM per ket
//             primitive covers all bra primitives: X[(i,tuv), kl] += R[(i,tuv), t'u'v'] Ek[t'u'v', kl];
//             the bra transform contracts over the stacked rows, I^T = X^T Eb_stack, reading X^T
//             straight out of the accumulator registers: accumulator element c of a 16-row tile
//             holds rows 4c..4c+3 across the lane groups, which is exactly the A-operand layout
//             of K-chunk c.  R is built cooperatively in LDS.
// Epilogues: J/K contraction with atomics, dense tensor scatter, or the (ab|ab) diagonal.
#pragma once

#include "device_common.hpp"
#include "lbx_internal.hpp"
#include "registry.hpp"

namespace lbx
{
    enum Mode
    {
        MODE_JK = 0,
        MODE_DENSE = 1,
        MODE_DIAG = 2
    };

    struct KArgs
    {
        DPairs bra, ket;
        const long long *prefix;   // per ket position q: first task (nket + 1 entries)
        const int *plo, *phi;      // bra sorted-position range [plo, phi) for each q
        int nket;
        long long ntasks;
        int same_class;            // bra class == ket class: quartet p == q is its own image
        const double *boys;
        // MODE_JK
        const double *D;
        double *J, *K;
        long long nbf;
        // MODE_DENSE: kind 4 eri4 (nbf^4), 3 eri3 (nbf, nbf, naux), 2 eri2 (naux, naux)
        double *out;
        int kind;
        long long naux;
        // MODE_DIAG
        double *qout;   // sqrt(max |(ab|ab)|) per pair (pair order)
        double *dout;   // (mn|mn) matrix (pair classes) or vector (single classes), may be null
    };

    template <int LA, int LB, int LC, int LD>
    struct Cls
    {
        static constexpr int L = LA + LB + LC + LD;
        static constexpr int HP = nHerm(LA + LB), HQ = nHerm(LC + LD), HL = nHerm(L);
        static constexpr int SA = nSph(LA), SB = nSph(LB), SC = nSph(LC), SD = nSph(LD);
        static constexpr int SP = SA * SB, SQ = SC * SD;
        // registers of the one-lane-per-quartet form: I + X + R
        static constexpr int kValuRegs = SP * SQ + HP * SQ + HL + L;
        static constexpr bool kValu = kValuRegs <= 160;
        // MFMA tiling
        static constexpr int NQ = ceilDiv(SQ, 16), NP = ceilDiv(SP, 16), KQ = ceilDiv(HQ, 4);
        // row tiles per chunk: accumulators bounded by 64 doubles per lane, and at most 16
        // primitive pairs' worth of rows (bounds the LDS R buffers)
        static constexpr int MT0 = (16 / NQ) < 1 ? 1 : (16 / NQ);
        static constexpr int MT = MT0 < HP ? MT0 : HP;
        static constexpr int RC = 16 * MT;                          // rows per chunk
        static constexpr int CHP = (RC - 1) / HP + 2;               // primitives a chunk touches
    };

    // ------------------------------------------------------------------------------- epilogues
    // Quartet integrals available through an accessor I(mn, kl); f = permutational weight.
    template <class C, class Acc>
    __device__ __forceinline__ void denseScatter(const KArgs &A, Acc I, int a0, int b0, int c0,
                                                 int d0, int e0, int ecount, int estep)
    {
        const long long n = A.nbf;
        for (int e = e0; e < ecount; e += estep)
        {
            const int mn = e / C::SQ, kl = e - mn * C::SQ;
            const int mu = mn / C::SB, nu = mn - mu * C::SB;
            const int ka = kl / C::SD, la = kl - ka * C::SD;
            const double v = I(mn, kl);
            const long long i = a0 + mu, j = b0 + nu, k = c0 + ka, l = d0 + la;
            if (A.kind == 4)
            {
                double *o = A.out;
                o[((i * n + j) * n + k) * n + l] = v;
                o[((i * n + j) * n + l) * n + k] = v;
                o[((j * n + i) * n + k) * n + l] = v;
                o[((j * n + i) * n + l) * n + k] = v;
                o[((k * n + l) * n + i) * n + j] = v;
                o[((k * n + l) * n + j) * n + i] = v;
                o[((l * n + k) * n + i) * n + j] = v;
                o[((l * n + k) * n + j) * n + i] = v;
            }
            else if (A.kind == 3)
            {
                A.out[(i * n + j) * A.naux + k] = v;
                A.out[(j * n + i) * A.naux + k] = v;
            }
            else
            {
                A.out[i * A.naux + k] = v;
                A.out[k * A.naux + i] = v;
            }
        }
    }

    // ------------------------------------------------------------------------------- quartetV
    template <int LA, int LB, int LC, int LD, int MODE>
    __device__ __forceinline__ void quartetVTask(const KArgs &A, long long t, int lane)
    {
        using C = Cls<LA, LB, LC, LD>;
        constexpr int L = C::L, HP = C::HP, HQ = C::HQ, SP = C::SP, SQ = C::SQ, HL = C::HL;

        const int q = uniform(findKet(A.prefix, A.nket, t));
        const long long local = t - A.prefix[q];
        const int plo = uniform(A.plo[q]), phi = uniform(A.phi[q]);
        const int p = plo + static_cast<int>(local) * 64 + lane;
        const bool active = p < phi;
        const int cd = uniform(A.ket.order[q]);
        const int ab = A.bra.order[active ? p : plo];
        const int kpoff = uniform(A.ket.poff[cd]), kcnt = uniform(A.ket.pcnt[cd]);
        const int bpoff = A.bra.poff[ab], bcnt = active ? A.bra.pcnt[ab] : 0;

        double I[SP * SQ];
#pragma unroll
        for (int e = 0; e < SP * SQ; e++)
            I[e] = 0.0;

        for (int i = 0; i < bcnt; i++)
        {
            const double4 P = A.bra.pp[bpoff + i];
            double X[HP * SQ];
#pragma unroll
            for (int e = 0; e < HP * SQ; e++)
                X[e] = 0.0;
            for (int k = 0; k < kcnt; k++)
            {
                const double4 Q = A.ket.pp[kpoff + k];
                const double pq = P.x * Q.x, ps = P.x + Q.x;
                const double alpha = pq / ps;
                const double X_ = P.y - Q.y, Y_ = P.z - Q.z, Z_ = P.w - Q.w;
                const double T = alpha * (X_ * X_ + Y_ * Y_ + Z_ * Z_);
                double R[HL + L];
                boysF<L>(T, A.boys, R);
                const double fac = kTwoPi52 / (pq * sqrt(ps));
                double sc = fac;
#pragma unroll
                for (int n = 0; n <= L; n++)
                {
                    R[n] *= sc;
                    sc *= -2.0 * alpha;
                }
                hermiteR<L>(R, X_, Y_, Z_);
                const double *__restrict__ ek = A.ket.ek + static_cast<size_t>(kpoff + k) * HQ * SQ;
#pragma unroll
                for (int h = 0; h < HP; h++)
#pragma unroll
                    for (int g = 0; g < HQ; g++)
                    {
                        const double r = R[rSum<L>(h, g)];
#pragma unroll
                        for (int c = 0; c < SQ; c++)
                            X[h * SQ + c] = fma(r, ek[g * SQ + c], X[h * SQ + c]);
                    }
            }
            const double *eb = A.bra.eb + static_cast<size_t>(bpoff + i) * HP * SP;
#pragma unroll
            for (int h = 0; h < HP; h++)
#pragma unroll
                for (int mn = 0; mn < SP; mn++)
                {
                    const double e = eb[h * SP + mn];
#pragma unroll
                    for (int c = 0; c < SQ; c++)
                        I[mn * SQ + c] = fma(e, X[h * SQ + c], I[mn * SQ + c]);
                }
        }

        const int a0 = active ? A.bra.ao_a[ab] : 0, b0 = active ? A.bra.ao_b[ab] : 0;
        const int c0 = uniform(A.ket.ao_a[cd]), d0 = uniform(A.ket.ao_b[cd]);
        if constexpr (MODE == MODE_JK)
        {
            double f = 1.0;
            if (A.bra.sh_a[ab] == A.bra.sh_b[ab])
                f *= 0.5;
            if (A.ket.sh_a[cd] == A.ket.sh_b[cd])
                f *= 0.5;
            if (A.same_class && p == q)
                f *= 0.5;
            if (!active)
                f = 0.0;
            const long long n = A.nbf;
            const double *D = A.D;
            constexpr int SA = C::SA, SB = C::SB, SC = C::SC, SD = C::SD;
            // J_ab += 2f I D_cd ; J_cd += 2f I^T D_ab
            double dcd[SQ];
#pragma unroll
            for (int ka = 0; ka < SC; ka++)
#pragma unroll
                for (int la = 0; la < SD; la++)
                    dcd[ka * SD + la] = D[(c0 + ka) * n + d0 + la];
            double jcd[SQ];
#pragma unroll
            for (int kl = 0; kl < SQ; kl++)
                jcd[kl] = 0.0;
#pragma unroll
            for (int mu = 0; mu < SA; mu++)
#pragma unroll
                for (int nu = 0; nu < SB; nu++)
                {
                    const int mn = mu * SB + nu;
                    const double dab = active ? D[(a0 + mu) * n + b0 + nu] : 0.0;
                    double s = 0.0;
#pragma unroll
                    for (int kl = 0; kl < SQ; kl++)
                    {
                        s = fma(I[mn * SQ + kl], dcd[kl], s);
                        jcd[kl] = fma(I[mn * SQ + kl], dab, jcd[kl]);
                    }
                    if (active)
                        unsafeAtomicAdd(&A.J[(a0 + mu) * n + b0 + nu], 2.0 * f * s);
                }
#pragma unroll
            for (int kl = 0; kl < SQ; kl++)
            {
                const double s = waveSum(2.0 * f * jcd[kl]);
                if (lane == 0)
                    unsafeAtomicAdd(&A.J[(c0 + kl / SD) * n + d0 + kl % SD], s);
            }
            if (active)
            {
                // K_ac += f sum_{nu la} I D_bd ; K_bc += f sum_{mu la} I D_ad ;
                // K_ad += f sum_{nu ka} I D_bc ; K_bd += f sum_{mu ka} I D_ac
#pragma unroll
                for (int mu = 0; mu < SA; mu++)
#pragma unroll
                    for (int ka = 0; ka < SC; ka++)
                    {
                        double s1 = 0.0, s3 = 0.0;
#pragma unroll
                        for (int nu = 0; nu < SB; nu++)
#pragma unroll
                            for (int la = 0; la < SD; la++)
                                s1 = fma(I[(mu * SB + nu) * SQ + ka * SD + la],
                                         D[(b0 + nu) * n + d0 + la], s1);
                        unsafeAtomicAdd(&A.K[(a0 + mu) * n + c0 + ka], f * s1);
                        (void)s3;
                    }
#pragma unroll
                for (int nu = 0; nu < SB; nu++)
#pragma unroll
                    for (int ka = 0; ka < SC; ka++)
                    {
                        double s = 0.0;
#pragma unroll
                        for (int mu = 0; mu < SA; mu++)
#pragma unroll
                            for (int la = 0; la < SD; la++)
                                s = fma(I[(mu * SB + nu) * SQ + ka * SD + la],
                                        D[(a0 + mu) * n + d0 + la], s);
                        unsafeAtomicAdd(&A.K[(b0 + nu) * n + c0 + ka], f * s);
                    }
#pragma unroll
                for (int mu = 0; mu < SA; mu++)
#pragma unroll
                    for (int la = 0; la < SD; la++)
                    {
                        double s = 0.0;
#pragma unroll
                        for (int nu = 0; nu < SB; nu++)
#pragma unroll
                            for (int ka = 0; ka < SC; ka++)
                                s = fma(I[(mu * SB + nu) * SQ + ka * SD + la],
                                        D[(b0 + nu) * n + c0 + ka], s);
                        unsafeAtomicAdd(&A.K[(a0 + mu) * n + d0 + la], f * s);
                    }
#pragma unroll
                for (int nu = 0; nu < SB; nu++)
#pragma unroll
                    for (int la = 0; la < SD; la++)
                    {
                        double s = 0.0;
#pragma unroll
                        for (int mu = 0; mu < SA; mu++)
#pragma unroll
                            for (int ka = 0; ka < SC; ka++)
                                s = fma(I[(mu * SB + nu) * SQ + ka * SD + la],
                                        D[(a0 + mu) * n + c0 + ka], s);
                        unsafeAtomicAdd(&A.K[(b0 + nu) * n + d0 + la], f * s);
                    }
            }
        }
        else if constexpr (MODE == MODE_DENSE)
        {
            // unrolled so that I stays in registers
            if (active)
#pragma unroll
                for (int e = 0; e < SP * SQ; e++)
                {
                    const double v = I[e];
                    denseScatter<C>(A, [v](int, int) { return v; }, a0, b0, c0, d0, e, e + 1, 1);
                }
        }
        else
        {
            // (ab|ab): the Schwarz factor and the diagonal (mn|mn)
            static_assert(MODE != MODE_DIAG || (LA == LC && LB == LD), "diag needs ab == cd");
            if (active)
            {
                double m = 0.0;
#pragma unroll
                for (int mn = 0; mn < SP; mn++)
                    m = fmax(m, fabs(I[mn * SQ + (mn < SQ ? mn : 0)]));
                A.qout[ab] = sqrt(m);
                if (A.dout)
                {
                    constexpr int SB = C::SB;
#pragma unroll
                    for (int mn = 0; mn < SP; mn++)
                    {
                        const int mu = mn / SB, nu = mn % SB;
                        const double v = I[mn * SQ + (mn < SQ ? mn : 0)];
                        if (A.kind == 2)
                            A.dout[a0 + mu] = v;
                        else
                        {
                            A.dout[(a0 + mu) * A.nbf + b0 + nu] = v;
                            A.dout[(b0 + nu) * A.nbf + a0 + mu] = v;
                        }
                    }
                }
            }
        }
    }

    template <int LA, int LB, int LC, int LD, int MODE>
    __global__ __launch_bounds__(256) void quartetV(const KArgs A)
    {
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        for (long long t = static_cast<long long>(blockIdx.x) * 4 + wave; t < A.ntasks;
             t += static_cast<long long>(gridDim.x) * 4)
            quartetVTask<LA, LB, LC, LD, MODE>(A, t, lane);
    }

    // ------------------------------------------------------------------------------- quartetM
    template <int LA, int LB, int LC, int LD, int MODE>
    __global__ __launch_bounds__(64) void quartetM(const KArgs A)
    {
        using C = Cls<LA, LB, LC, LD>;
        constexpr int L = C::L, HP = C::HP, HQ = C::HQ, SP = C::SP, SQ = C::SQ, HL = C::HL;
        constexpr int NQ = C::NQ, NP = C::NP, KQ = C::KQ, MT = C::MT, RC = C::RC, CHP = C::CHP;
        constexpr int SA = C::SA, SB = C::SB, SC = C::SC, SD = C::SD;
        static_assert(CHP <= 64, "one lane per primitive pair of a chunk");
        __shared__ double sR[2][CHP][HL];
        __shared__ double sInit[CHP][L + 1];
        __shared__ double sPQ[CHP][3];
        __shared__ unsigned short sTab[HP][4 * KQ];
        __shared__ double sI[SP * SQ];
        __shared__ double sD[6][49];

        const int lane = threadIdx.x;
        const int g = lane >> 4, j = lane & 15;

        // Hermite-sum table: row tuv, column t'u'v' -> index of the summed Hermite function
        for (int e = lane; e < HP * 4 * KQ; e += 64)
        {
            const int h = e / (4 * KQ), k = e % (4 * KQ);
            int v = 0;
            if (k < HQ)
            {
                const int a = hermTuv(h), b = hermTuv(k);
                v = hermIdx((a & 255) + (b & 255), ((a >> 8) & 255) + ((b >> 8) & 255),
                            (a >> 16) + (b >> 16));
            }
            sTab[h][k] = static_cast<unsigned short>(v);
        }
        for (int e = lane; e < 2 * CHP * HL; e += 64)
            (&sR[0][0][0])[e] = 0.0;
        __syncthreads();

        for (long long t = blockIdx.x; t < A.ntasks; t += gridDim.x)
        {
            const int q = uniform(findKet(A.prefix, A.nket, t));
            const int p = uniform(A.plo[q] + static_cast<int>(t - A.prefix[q]));
            const int cd = uniform(A.ket.order[q]);
            const int ab = uniform(A.bra.order[p]);
            const int kpoff = uniform(A.ket.poff[cd]), kcnt = uniform(A.ket.pcnt[cd]);
            const int bpoff = uniform(A.bra.poff[ab]), bcnt = uniform(A.bra.pcnt[ab]);
            const int nrows = bcnt * HP;

            v4d acc2[NQ][NP];
#pragma unroll
            for (int a = 0; a < NQ; a++)
#pragma unroll
                for (int b = 0; b < NP; b++)
                    acc2[a][b] = v4d{0.0, 0.0, 0.0, 0.0};

            for (int rowbase = 0; rowbase < nrows; rowbase += RC)
            {
                const int prim0 = rowbase / HP;
                const int primEnd = min(bcnt, (rowbase + RC - 1) / HP + 1);
                const int nprim = primEnd - prim0;

                v4d acc1[MT][NQ];
#pragma unroll
                for (int a = 0; a < MT; a++)
#pragma unroll
                    for (int b = 0; b < NQ; b++)
                        acc1[a][b] = v4d{0.0, 0.0, 0.0, 0.0};

                for (int k = 0; k < kcnt; k++)
                {
                    const double4 Q = A.ket.pp[kpoff + k];
                    // Boys + scaled R^{(n)}_000 for the chunk's primitive quartets
                    if (lane < nprim)
                    {
                        const double4 P = A.bra.pp[bpoff + prim0 + lane];
                        const double pq = P.x * Q.x, ps = P.x + Q.x;
                        const double alpha = pq / ps;
                        const double X_ = P.y - Q.y, Y_ = P.z - Q.z, Z_ = P.w - Q.w;
                        const double T = alpha * (X_ * X_ + Y_ * Y_ + Z_ * Z_);
                        double F[L + 1];
                        boysF<L>(T, A.boys, F);
                        double sc = kTwoPi52 / (pq * sqrt(ps));
#pragma unroll
                        for (int n = 0; n <= L; n++)
                        {
                            sInit[lane][n] = F[n] * sc;
                            sc *= -2.0 * alpha;
                        }
                        sPQ[lane][0] = X_;
                        sPQ[lane][1] = Y_;
                        sPQ[lane][2] = Z_;
                        sR[L & 1][lane][0] = sInit[lane][L];
                    }
                    __syncthreads();
                    // levels n = L-1 .. 0, ping-pong buffers
#pragma unroll
                    for (int n = L - 1; n >= 0; n--)
                    {
                        const int nh = nHerm(L - n) - 1;
                        const double(*in)[HL] = sR[(n + 1) & 1];
                        double(*out)[HL] = sR[n & 1];
                        const int items = nprim * nh;
                        for (int it = lane; it < items; it += 64)
                        {
                            const int pr = it / nh;
                            const int h = it - pr * nh + 1;
                            const unsigned r = kRecTab.v[h];
                            const int s1 = r & 2047, s2 = (r >> 11) & 2047, ax = (r >> 22) & 3,
                                      c = (r >> 24);
                            out[pr][h] = fma(sPQ[pr][ax], in[pr][s1], c * in[pr][s2]);
                        }
                        if (lane < nprim)
                            out[lane][0] = sInit[lane][n];
                        __syncthreads();
                    }
                    // ket transform on MFMA: acc1[mt][nt] += Rg[rows, k] Ek[k, cols]
                    const double *ek = A.ket.ek + static_cast<size_t>(kpoff + k) * HQ * SQ;
#pragma unroll
                    for (int kt = 0; kt < KQ; kt++)
                    {
                        const int kk = 4 * kt + g;
                        double b1[NQ];
#pragma unroll
                        for (int nt = 0; nt < NQ; nt++)
                        {
                            const int col = 16 * nt + j;
                            b1[nt] = (kk < HQ && col < SQ) ? ek[kk * SQ + col] : 0.0;
                        }
#pragma unroll
                        for (int mt = 0; mt < MT; mt++)
                        {
                            if (rowbase + 16 * mt >= nrows)
                                continue;
                            const int gr = rowbase + 16 * mt + j;
                            const int pr = gr / HP;
                            const int h = gr - pr * HP;
                            const double a = (pr < bcnt) ? sR[0][pr - prim0][sTab[h][kk]] : 0.0;
#pragma unroll
                            for (int nt = 0; nt < NQ; nt++)
                                acc1[mt][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                        a, b1[nt], acc1[mt][nt], 0, 0, 0);
                        }
                    }
                    __syncthreads();
                }
                // bra transform: I^T[kl, mn] += X^T[kl, rows] Eb[rows, mn]
#pragma unroll
                for (int mt = 0; mt < MT; mt++)
                {
                    if (rowbase + 16 * mt >= nrows)
                        continue;
#pragma unroll
                    for (int c = 0; c < 4; c++)
                    {
                        // K-chunk c of X^T = accumulator element c = rows 16 mt + 4 c + (0..3)
                        const int gr = rowbase + 16 * mt + 4 * c + g;
                        const int pr = gr / HP;
                        const int h = gr - pr * HP;
                        const bool rv = pr < bcnt;
                        const double *eb = A.bra.eb + (static_cast<size_t>(bpoff + pr) * HP + h) * SP;
#pragma unroll
                        for (int np = 0; np < NP; np++)
                        {
                            const int col = 16 * np + j;
                            const double b2 = (rv && col < SP) ? eb[col] : 0.0;
#pragma unroll
                            for (int nq = 0; nq < NQ; nq++)
                                acc2[nq][np] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                        acc1[mt][nq][c], b2, acc2[nq][np], 0, 0, 0);
                        }
                    }
                }
            }
            // accumulator layout of v_mfma_f64_16x16x4f64: lane (g, j), element r holds
            // D[4 r + g][j]; here D = I^T[16 nq + 4 r + g][16 np + j]
#pragma unroll
            for (int nq = 0; nq < NQ; nq++)
#pragma unroll
                for (int np = 0; np < NP; np++)
#pragma unroll
                    for (int r = 0; r < 4; r++)
                    {
                        const int kl = 16 * nq + 4 * r + g, mn = 16 * np + j;
                        if (kl < SQ && mn < SP)
                            sI[mn * SQ + kl] = acc2[nq][np][r];
                    }
            __syncthreads();

            const int a0 = A.bra.ao_a[ab], b0 = A.bra.ao_b[ab];
            const int c0 = A.ket.ao_a[cd], d0 = A.ket.ao_b[cd];
            if constexpr (MODE == MODE_JK)
            {
                double f = 1.0;
                if (A.bra.sh_a[ab] == A.bra.sh_b[ab])
                    f *= 0.5;
                if (A.ket.sh_a[cd] == A.ket.sh_b[cd])
                    f *= 0.5;
                if (A.same_class && p == q)
                    f *= 0.5;
                const long long n = A.nbf;
                const double *D = A.D;
                // D blocks: 0 ab, 1 cd, 2 bd, 3 ad, 4 bc, 5 ac
                for (int e = lane; e < SA * SB; e += 64)
                    sD[0][e] = D[(a0 + e / SB) * n + b0 + e % SB];
                for (int e = lane; e < SC * SD; e += 64)
                    sD[1][e] = D[(c0 + e / SD) * n + d0 + e % SD];
                for (int e = lane; e < SB * SD; e += 64)
                    sD[2][e] = D[(b0 + e / SD) * n + d0 + e % SD];
                for (int e = lane; e < SA * SD; e += 64)
                    sD[3][e] = D[(a0 + e / SD) * n + d0 + e % SD];
                for (int e = lane; e < SB * SC; e += 64)
                    sD[4][e] = D[(b0 + e / SC) * n + c0 + e % SC];
                for (int e = lane; e < SA * SC; e += 64)
                    sD[5][e] = D[(a0 + e / SC) * n + c0 + e % SC];
                __syncthreads();
                for (int mn = lane; mn < SP; mn += 64)
                {
                    double s = 0.0;
                    for (int kl = 0; kl < SQ; kl++)
                        s = fma(sI[mn * SQ + kl], sD[1][kl], s);
                    unsafeAtomicAdd(&A.J[(a0 + mn / SB) * n + b0 + mn % SB], 2.0 * f * s);
                }
                for (int kl = lane; kl < SQ; kl += 64)
                {
                    double s = 0.0;
                    for (int mn = 0; mn < SP; mn++)
                        s = fma(sI[mn * SQ + kl], sD[0][mn], s);
                    unsafeAtomicAdd(&A.J[(c0 + kl / SD) * n + d0 + kl % SD], 2.0 * f * s);
                }
                // K_ac[mu,ka] += f sum_{nu,la} I D_bd[nu,la]
                for (int e = lane; e < SA * SC; e += 64)
                {
                    const int mu = e / SC, ka = e % SC;
                    double s = 0.0;
                    for (int nu = 0; nu < SB; nu++)
                        for (int la = 0; la < SD; la++)
                            s = fma(sI[(mu * SB + nu) * SQ + ka * SD + la], sD[2][nu * SD + la], s);
                    unsafeAtomicAdd(&A.K[(a0 + mu) * n + c0 + ka], f * s);
                }
                // K_bc[nu,ka] += f sum_{mu,la} I D_ad[mu,la]
                for (int e = lane; e < SB * SC; e += 64)
                {
                    const int nu = e / SC, ka = e % SC;
                    double s = 0.0;
                    for (int mu = 0; mu < SA; mu++)
                        for (int la = 0; la < SD; la++)
                            s = fma(sI[(mu * SB + nu) * SQ + ka * SD + la], sD[3][mu * SD + la], s);
                    unsafeAtomicAdd(&A.K[(b0 + nu) * n + c0 + ka], f * s);
                }
                // K_ad[mu,la] += f sum_{nu,ka} I D_bc[nu,ka]
                for (int e = lane; e < SA * SD; e += 64)
                {
                    const int mu = e / SD, la = e % SD;
                    double s = 0.0;
                    for (int nu = 0; nu < SB; nu++)
                        for (int ka = 0; ka < SC; ka++)
                            s = fma(sI[(mu * SB + nu) * SQ + ka * SD + la], sD[4][nu * SC + ka], s);
                    unsafeAtomicAdd(&A.K[(a0 + mu) * n + d0 + la], f * s);
                }
                // K_bd[nu,la] += f sum_{mu,ka} I D_ac[mu,ka]
                for (int e = lane; e < SB * SD; e += 64)
                {
                    const int nu = e / SD, la = e % SD;
                    double s = 0.0;
                    for (int mu = 0; mu < SA; mu++)
                        for (int ka = 0; ka < SC; ka++)
                            s = fma(sI[(mu * SB + nu) * SQ + ka * SD + la], sD[5][mu * SC + ka], s);
                    unsafeAtomicAdd(&A.K[(b0 + nu) * n + d0 + la], f * s);
                }
            }
            else if constexpr (MODE == MODE_DENSE)
            {
                denseScatter<C>(A, [&](int mn, int kl) { return sI[mn * SQ + kl]; }, a0, b0, c0,
                                d0, lane, SP * SQ, 64);
            }
            else
            {
                if (lane == 0)
                {
                    double m = 0.0;
                    for (int mn = 0; mn < SP; mn++)
                        m = fmax(m, fabs(sI[mn * SQ + mn]));
                    A.qout[ab] = sqrt(m);
                }
                if (A.dout)
                    for (int mn = lane; mn < SP; mn += 64)
                    {
                        const int mu = mn / SB, nu = mn % SB;
                        const double v = sI[mn * SQ + mn];
                        if (A.kind == 2)
                            A.dout[a0 + mu] = v;
                        else
                        {
                            A.dout[(a0 + mu) * A.nbf + b0 + nu] = v;
                            A.dout[(b0 + nu) * A.nbf + a0 + mu] = v;
                        }
                    }
            }
            __syncthreads();
        }
    }

    template <int LA, int LB, int LC, int LD, int MODE>
    void launchQuartet(const KArgs &A, hipStream_t s)
    {
        using C = Cls<LA, LB, LC, LD>;
        if (A.ntasks <= 0)
            return;
        if constexpr (C::kValu)
        {
            long long wg = (A.ntasks + 3) / 4;
            int grid = static_cast<int>(wg < 65536 ? wg : 65536);
            hipLaunchKernelGGL((quartetV<LA, LB, LC, LD, MODE>), dim3(grid), dim3(256), 0, s, A);
        }
        else
        {
            int grid = static_cast<int>(A.ntasks < 131072 ? A.ntasks : 131072);
            hipLaunchKernelGGL((quartetM<LA, LB, LC, LD, MODE>), dim3(grid), dim3(64), 0, s, A);
        }
    }

    template <int LA, int LB, int LC, int LD, int MODE>
    void registerQuartet()
    {
        KernelSlot &s = kernelSlot(MODE, LA, LB, LC, LD);
        s.fn = &launchQuartet<LA, LB, LC, LD, MODE>;
        s.width = Cls<LA, LB, LC, LD>::kValu ? 64 : 1;
    }
}
