// SHARK shell-quartet kernels for gfx950.
//
// Per shell quartet (ab|cd) of class (LA LB | LC LD), following the reference's SHARK scheme
// (eri_kernel_funs.hpp:104-193, after doi:10.1002/jcc.26942 eqs. 25-26):
//   for each bra primitive pair i:  X_i = sum_k R_{ik} . Ek_k          (ket transform)
//   I = sum_i Eb_i^T X_i                                               (bra transform)
// with R_{ik}(tuv, t'u'v') = fac R_{t+t',u+u',v+v'} (Hermite integrals; the ket's sign
// (-1)^{t'+u'+v'} folded into Ek) and Eb, Ek the spherical Hermite-expansion matrices of the
// primitive pairs.
//
// Two kernels:
//   quartetV  one lane per quartet, everything in registers.  A wave holds 64 pairs of one side
//             (the "lane" side) against one pair of the other (the "uniform" side), so every
//             uniform-side operand is a scalar load and the transforms are VALU FMAs with an
//             SGPR operand.  Registers are bounded by splitting the uniform side's spherical
//             components into column passes (Boys + R recomputed per pass); which side runs on
//             the lanes is chosen per class to minimise work.
//   quartetM  one wave per quartet, both transforms on FP64 MFMA (v_mfma_f64_16x16x4f64).  The
//             bra primitive pairs are stacked along M, (prim, tuv) rows, so one GEMM per ket
//             primitive covers all bra primitives: X[(i,tuv), kl] += R[(i,tuv), t'u'v'] Ek[t'u'v', kl];
//             the bra transform contracts over the stacked rows, I^T = X^T Eb_stack, reading X^T
//             straight out of the accumulator registers: accumulator element c of a 16-row tile
//             holds rows 4c..4c+3 across the lane groups, which is exactly the A-operand layout
//             of K-chunk c.  R is built cooperatively in LDS.  Used where R alone would not fit
//             a lane's registers (high total angular momentum).
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
        MODE_DIAG = 2,
        MODE_SUM = 3,   // per-molecule sums of |value| and value of the dense eri3 / eri2 tensors
        N_MODES = 4
    };

    struct KArgs
    {
        DPairs bra, ket;
        const long long *prefix;   // per uniform-side position q: first task (nket + 1 entries)
        const int *plo, *phi;      // lane-side sorted-position range [plo, phi) for each q
        int nket;
        long long ntasks;
        int same_class;            // bra class == ket class: quartet p == q is its own image
        int bra_major;             // quartetM: q indexes bra pairs, p ket pairs (classes differ)
        // quartetV lane-block-major order: lane block b = lane positions 64b .. 64b+63; group g =
        // blocks g G .. g G + G - 1 (G = 2^gshift).  Group g's tasks pair its uniform positions
        // q = bq0[g], bq0[g] + 1, ... with its G blocks in turn (bprefix[g+1] - bprefix[g] = G x
        // the count of those q).
        int block_major, nblk, gshift;
        const long long *bprefix;
        const int *bq0;
        // lane-block-major J/K: lane position -> Schwarz-order position (the lane side ordered by
        // shell within Schwarz-factor buckets, so lanes sharing a shell are adjacent), and per
        // lane block the range [bqlo, bqhi) of Schwarz-order positions it holds; null otherwise
        const int *lperm, *bqlo, *bqhi;
        const double *boys;
        // MODE_JK
        const double *D;
        double *J, *K;
        long long nbf;
        // MODE_DENSE: kind 4 eri4 (nbf^4), 3 eri3 (nbf, nbf, naux), 2 eri2 (naux, naux)
        // MODE_SUM: out = (nmol, 4) sums |eri3|, eri3, |eri2|, eri2 by the ket's molecule
        double *out;
        int kind;
        long long naux;
        // MODE_DIAG
        double *qout;   // sqrt(max |(ab|ab)|) per pair (pair order)
        double *dout;   // (mn|mn) matrix (pair classes) or vector (single classes), may be null
    };

    constexpr LBX_HD int rFlops(int L)
    {
        int r = 0;
        for (int n = 0; n < L; n++)
            for (int m = 1; m <= L - n; m++)
                r += (m + 1) * (m + 2) / 2;
        return 2 * r;
    }

    template <int LA, int LB, int LC, int LD>
    struct Cls
    {
        static constexpr int L = LA + LB + LC + LD;
        static constexpr int HP = nHerm(LA + LB), HQ = nHerm(LC + LD), HL = nHerm(L);
        static constexpr int SA = nSph(LA), SB = nSph(LB), SC = nSph(LC), SD = nSph(LD);
        static constexpr int SP = SA * SB, SQ = SC * SD;

        // ---- one-lane-per-quartet plan: per-lane doubles of I (S_lane x NC) + X (H_lane x NC)
        // + R stay within kBudget (the region where the compiler keeps them in registers)
        static constexpr int kBudget = 160;
        // Pinned: at 192 the ket-on-lanes (ff|ss) form (one static pass of all 49 columns
        // instead of two of 35) gives a wrong lane-side J on the GPU (a few hundred elements off
        // by up to 5%; its K and uniform-side J are right; cause not found).  Raise only with
        // every changed form checked element-wise against the oracle.
        static_assert(kBudget == 160, "one-lane register budget: see the comment above");
        static constexpr int ncFor(int sl, int hl, int su)
        {
            const int nc = (kBudget - (HL + L)) / (sl + hl);
            return nc < 1 ? 0 : (nc > su ? su : nc);
        }
        static constexpr int NC0 = ncFor(SP, HP, SQ);   // lanes = bra, uniform = ket
        static constexpr int NC1 = ncFor(SQ, HQ, SP);   // lanes = ket, uniform = bra
        // FP32 form: the same register budget holds twice as many values
        static constexpr int ncFor32(int sl, int hl, int su)
        {
            const int nc = (2 * kBudget - (HL + L)) / (sl + hl);
            return nc < 1 ? 0 : (nc > su ? su : nc);
        }
        static constexpr int NC0f = ncFor32(SP, HP, SQ);
        static constexpr int NC1f = ncFor32(SQ, HQ, SP);
        template <typename T, bool SWAP>
        static constexpr int nc() { return sizeof(T) == 4 ? (SWAP ? NC1f : NC0f) : (SWAP ? NC1 : NC0); }

        // ---- MFMA tiling
        static constexpr int NQ = ceilDiv(SQ, 16), NP = ceilDiv(SP, 16), KQ = ceilDiv(HQ, 4);
        // row tiles per chunk: accumulators (acc1 + acc2) within ~96 doubles per lane, rows for
        // about four primitive pairs, at least one tile
        static constexpr int mtBudget = (96 - 4 * NQ * NP) / (4 * NQ);
        static constexpr int mtRows = ceilDiv(4 * HP, 16);
        static constexpr int MT0 = mtBudget < mtRows ? mtBudget : mtRows;
        static constexpr int MT = MT0 < 1 ? 1 : MT0;
        static constexpr int RC = 16 * MT;                          // rows per chunk
        static constexpr int CHP = (RC - 1) / HP + 2;               // primitives a chunk touches
    };

    // ------------------------------------------------------------------------------- dense out
    template <class C>
    __device__ __forceinline__ void denseStore(const KArgs &A, double v, int mn, int kl, int a0,
                                               int b0, int c0, int d0)
    {
        const long long n = A.nbf;
        const int mu = mn / C::SB, nu = mn - mu * C::SB;
        const int ka = kl / C::SD, la = kl - ka * C::SD;
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

    // Runs of adjacent lanes with equal key (a shell of the lane-side pair): per-lane masks of a
    // segmented suffix sum (step i adds the lane 2^i above when it is in the same run), the run
    // head, and whether any lane of the run holds a quartet.
    struct LaneRuns
    {
        unsigned ok;    // bit i: lane + 2^i is in this lane's run
        bool head;      // first lane of its run
        bool any;       // (head lanes) some lane of the run is active
    };

    __device__ __forceinline__ LaneRuns laneRuns(int key, bool active, int lane)
    {
        LaneRuns r;
        const int prev = __shfl_up(key, 1, 64);
        r.head = lane == 0 || prev != key;
        // run id = lane index of the run head (prefix max of head ? lane : 0)
        int id = r.head ? lane : 0;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
        {
            const int x = __shfl_up(id, o, 64);
            if (lane >= o)
                id = max(id, x);
        }
        r.ok = 0;
        int act = active ? 1 : 0;
#pragma unroll
        for (int i = 0; i < 6; i++)
        {
            const int o = 1 << i;
            const int xid = __shfl_down(id, o, 64);
            const int xa = __shfl_down(act, o, 64);
            const bool same = lane + o < 64 && xid == id;
            if (same)
            {
                r.ok |= 1u << i;
                act |= xa;
            }
        }
        r.any = act != 0;
        return r;
    }

    // Adds v of every lane of a run into one address, by the run's head lane (all lanes call).
    __device__ __forceinline__ void runAtomicAdd(const LaneRuns &r, double *addr, double v)
    {
#pragma unroll
        for (int i = 0; i < 6; i++)
        {
            const double x = __shfl_down(v, 1 << i, 64);
            if (r.ok & (1u << i))
                v += x;
        }
        if (r.head && r.any)
            unsafeAtomicAdd(addr, v);
    }

    // ------------------------------------------------------------------------------- quartetV
    // One quartet per lane: lane-side pair l (64 consecutive sorted positions) against the
    // wave-uniform pair u.  SWAP = false: lanes = bra, u = ket; SWAP = true: lanes = ket, u = bra.
    // T: type of R, the E matrices and both transforms (double; float for the reduced-precision
    // J/K quartets, whose Boys values, epilogue and accumulation stay double).
    // jl (J/K only, SL entries per lane): this lane's J contributions to its lane-side pair,
    // scaled by 2f.  CARRY = false: zeroed here and added to J at the end of the task; CARRY =
    // true: accumulated into the caller's array, which adds it to J once for a run of tasks that
    // share the lane-side pairs.
    template <int LA, int LB, int LC, int LD, int MODE, bool SWAP, typename T, bool CARRY = false>
    __device__ __forceinline__ void quartetVTask(const KArgs &A, int q, int local, int lane, double *jl)
    {
        using C = Cls<LA, LB, LC, LD>;
        constexpr int L = C::L, HL = C::HL;
        // lane side (La, Lb), uniform side (Ua, Ub)
        constexpr int La = SWAP ? LC : LA, Lb = SWAP ? LD : LB;
        constexpr int Ua = SWAP ? LA : LC, Ub = SWAP ? LB : LD;
        constexpr int SLa = nSph(La), SLb = nSph(Lb), SUa = nSph(Ua), SUb = nSph(Ub);
        constexpr int SL = SLa * SLb, SU = SUa * SUb;
        constexpr int HLn = nHerm(La + Lb), HU = nHerm(Ua + Ub);
        constexpr int NC = C::template nc<T, SWAP>();
        constexpr int NPASS = ceilDiv(SU, NC);
        constexpr bool F32 = sizeof(T) == 4;
        // Skipping structural zeros scatters the loads of E (no wide contiguous loads); it pays
        // where each loaded value feeds several FMAs: lane side with >= 2 columns per pass,
        // uniform side with >= 4 Hermite rows of X.
        constexpr bool kSkipL = NC >= 2;
        constexpr bool kSkipU = HLn >= 4;
        static_assert(!F32 || MODE == MODE_JK || MODE == MODE_SUM, "FP32 forms: J/K and RI sums");
        const DPairs &LP = SWAP ? A.ket : A.bra;
        const DPairs &UP = SWAP ? A.bra : A.ket;

        // lanes cover lane positions pbase .. pbase + 63 (pbase = `local`); p is the lane's
        // Schwarz-order position (through A.lperm when the lane side is shell-ordered)
        const int plo = uniform(A.plo[q]), phi = uniform(A.phi[q]);
        const bool permuted = A.lperm != nullptr;
        // the last group of a lane-block-major table runs G blocks, some past the lane side
        if (local >= LP.n)
            return;
        if (permuted ? (plo >= A.bqhi[local >> 6] || phi <= A.bqlo[local >> 6])
                     : (plo >= local + 64 || phi <= local))
            return;   // no lane has a quartet
        const int pl = local + lane;
        const bool inside = !permuted || pl < LP.n;
        const int p = permuted ? A.lperm[inside ? pl : local] : pl;
        const bool active = inside && p >= plo && p < phi;
        const int u = uniform(UP.order[q]);
        const int l = LP.order[active || permuted ? p : local];
        const int upoff = uniform(UP.poff[u]), ucnt = uniform(UP.pcnt[u]);
        const int lpoff = LP.poff[l], lcnt = active ? LP.pcnt[l] : 0;
        // the ket's E carries the sign (-1)^{t'+u'+v'}
        const T *eUb, *eLb;
        if constexpr (F32)
        {
            eUb = SWAP ? UP.eb32 : UP.ek32;
            eLb = SWAP ? LP.ek32 : LP.eb32;
        }
        else
        {
            eUb = SWAP ? UP.eb : UP.ek;
            eLb = SWAP ? LP.ek : LP.eb;
        }
        const cptr<T> eU = constView(eUb + static_cast<size_t>(upoff) * HU * SU);
        const cptr<double> ppU = constView(reinterpret_cast<const double *>(UP.pp + upoff));
        const T *__restrict__ eL = eLb + static_cast<size_t>(lpoff) * HLn * SL;

        const int l0 = active ? LP.ao_a[l] : 0, l1 = active ? LP.ao_b[l] : 0;
        const int u0 = uniform(UP.ao_a[u]), u1 = uniform(UP.ao_b[u]);
        double f = 1.0;
        if constexpr (MODE == MODE_JK)
        {
            if (LP.sh_a[l] == LP.sh_b[l])
                f *= 0.5;
            if (UP.sh_a[u] == UP.sh_b[u])
                f *= 0.5;
            if (A.same_class && p == q)
                f *= 0.5;
            if (!active)
                f = 0.0;
        }
        const long long n = A.nbf;
        const double *D = A.D;
        // K updates on the columns of the lane pair's first shell: with a shell-ordered lane side
        // the lanes of one shell add their sum once (runs of adjacent lanes), else each lane
        // adds its own
        LaneRuns runs{};
        int la0 = l0;
        if constexpr (MODE == MODE_JK)
            if (permuted)
            {
                la0 = LP.ao_a[l];
                runs = laneRuns(la0, active, lane);
            }
        auto addA = [&](long long row, int al, double v) {
            if (permuted)
                runAtomicAdd(runs, &A.K[row * n + la0 + al], v);
            else if (active)
                unsafeAtomicAdd(&A.K[row * n + l0 + al], v);
        };
        if constexpr (MODE == MODE_JK && !CARRY)
#pragma unroll
            for (int s = 0; s < SL; s++)
                jl[s] = 0.0;
        double dmax = 0.0;
        double sabs = 0.0, ssum = 0.0;   // MODE_SUM

        // Passes stay a runtime loop (unrolling them lets the compiler overlap the live ranges of
        // successive passes and spill), but each pass runs its own instance of the body with a
        // compile-time pass index, so the column tests fold away and the uniform side's
        // structural zeros are skipped per column.
        auto passBody = [&](auto passC) {
            // a compile-time constant for the instanced passes, a runtime index otherwise
            const int pass = static_cast<int>(passC);
            constexpr bool kStaticPass = !std::is_same_v<decltype(passC), int>;
            T I[SL * NC];
#pragma unroll
            for (int e = 0; e < SL * NC; e++)
                I[e] = T(0);

            for (int i = 0; i < lcnt; i++)
            {
                const double4 Pl = LP.pp[lpoff + i];
                T X[HLn * NC];
#pragma unroll
                for (int e = 0; e < HLn * NC; e++)
                    X[e] = T(0);
                for (int k = 0; k < ucnt; k++)
                {
                    const double4 Pu{ppU[4 * k], ppU[4 * k + 1], ppU[4 * k + 2], ppU[4 * k + 3]};
                    const double4 &Pb = SWAP ? Pu : Pl;
                    const double4 &Pk = SWAP ? Pl : Pu;
                    const double pq = Pb.x * Pk.x, ps = Pb.x + Pk.x;
                    const double alpha = pq / ps;
                    const double X_ = Pb.y - Pk.y, Y_ = Pb.z - Pk.z, Z_ = Pb.w - Pk.w;
                    const double Tb = alpha * (X_ * X_ + Y_ * Y_ + Z_ * Z_);
                    double F[L + 1];
                    boysF<L>(Tb, A.boys, F);
                    T R[HL + L];
                    double sc = kTwoPi52 / (pq * sqrt(ps));
#pragma unroll
                    for (int m = 0; m <= L; m++)
                    {
                        R[m] = T(F[m] * sc);
                        sc *= -2.0 * alpha;
                    }
                    hermiteR<L, T>(R, T(X_), T(Y_), T(Z_));
                    const cptr<T> ek = eU + static_cast<size_t>(k) * HU * SU;
                    constexpr RSumTab<L, HLn, HU> RS{};
                    constexpr EPattern<Ua, Ub> PU{};   // the uniform side's structural zeros
#pragma unroll
                    for (int h = 0; h < HLn; h++)
#pragma unroll
                        for (int g = 0; g < HU; g++)
                        {
                            const T r = R[RS.v[h * HU + g]];
#pragma unroll
                            for (int c = 0; c < NC; c++)
                                if (pass * NC + c < SU &&
                                    (!kSkipU || !kStaticPass || PU.nz[g][pass * NC + c]))
                                    X[h * NC + c] = fma(r, ek[g * SU + pass * NC + c], X[h * NC + c]);
                        }
                }
                const T *el = eL + static_cast<size_t>(i) * HLn * SL;
                // structural zeros of the lane side's E (about half the entries) are skipped
                constexpr EPattern<La, Lb> PL{};
#pragma unroll
                for (int h = 0; h < HLn; h++)
#pragma unroll
                    for (int s = 0; s < SL; s++)
                    {
                        if (kSkipL && !PL.nz[h][s])
                            continue;
                        const T e = el[h * SL + s];
#pragma unroll
                        for (int c = 0; c < NC; c++)
                            I[s * NC + c] = fma(e, X[h * NC + c], I[s * NC + c]);
                    }
            }

            // ---- epilogue for the columns of this pass: I[s][c] = (l s | u pass*NC + c)
            if constexpr (MODE == MODE_JK)
            {
                // J_l += 2f I D_u (accumulated over passes); J_u += 2f I^T D_l (wave sum)
#pragma unroll
                for (int c = 0; c < NC; c++)
                {
                    const int su = pass * NC + c;
                    if (su >= SU)
                        continue;
                    const int ga = su / SUb, de = su % SUb;
                    const double du = 2.0 * f * D[(u0 + ga) * n + u1 + de];
                    double ju = 0.0;
#pragma unroll
                    for (int s = 0; s < SL; s++)
                    {
                        const int al = s / SLb, be = s % SLb;
                        const double iv = I[s * NC + c];
                        jl[s] = fma(iv, du, jl[s]);
                        ju = fma(iv, active ? D[(l0 + al) * n + l1 + be] : 0.0, ju);
                    }
                    ju = waveSum(2.0 * f * ju);
                    if (lane == 0)
                        unsafeAtomicAdd(&A.J[(u0 + ga) * n + u1 + de], ju);
                }
                {
                    // K[u0+ga][l0+al] += f sum_{be, de} I D[u1+de][l1+be]
                    // K[u0+ga][l1+be] += f sum_{al, de} I D[u1+de][l0+al]
                    // K[u1+de][l0+al] += f sum_{be, ga} I D[u0+ga][l1+be]
                    // K[u1+de][l1+be] += f sum_{al, ga} I D[u0+ga][l0+al]
                    // The wave-uniform shell indexes the row of K (and of the symmetric D): the
                    // unique-quartet contributions are symmetrised at the end (K + K^T), so
                    // either triangle may receive them, and with the row fixed across the wave
                    // its atomics and loads stay on a few rows that remain in L2.
#pragma unroll
                    for (int ga = 0; ga < SUa; ga++)
                    {
                        // columns of this pass with first index ga
                        bool any = false;
#pragma unroll
                        for (int de = 0; de < SUb; de++)
                            any |= (ga * SUb + de) / NC == pass;
                        if (!any)
                            continue;
#pragma unroll
                        for (int al = 0; al < SLa; al++)
                        {
                            double s = 0.0;
#pragma unroll
                            for (int be = 0; be < SLb; be++)
#pragma unroll
                                for (int de = 0; de < SUb; de++)
                                {
                                    const int su = ga * SUb + de;
                                    if (su / NC != pass)
                                        continue;
                                    s = fma(double(I[(al * SLb + be) * NC + su % NC]),
                                            D[(u1 + de) * n + l1 + be], s);
                                }
                            addA(u0 + ga, al, f * s);
                        }
#pragma unroll
                        for (int be = 0; be < SLb; be++)
                        {
                            double s = 0.0;
#pragma unroll
                            for (int al = 0; al < SLa; al++)
#pragma unroll
                                for (int de = 0; de < SUb; de++)
                                {
                                    const int su = ga * SUb + de;
                                    if (su / NC != pass)
                                        continue;
                                    s = fma(double(I[(al * SLb + be) * NC + su % NC]),
                                            D[(u1 + de) * n + l0 + al], s);
                                }
                            if (active)
                                unsafeAtomicAdd(&A.K[(u0 + ga) * n + l1 + be], f * s);
                        }
                    }
#pragma unroll
                    for (int de = 0; de < SUb; de++)
                    {
                        bool any = false;
#pragma unroll
                        for (int ga = 0; ga < SUa; ga++)
                            any |= (ga * SUb + de) / NC == pass;
                        if (!any)
                            continue;
#pragma unroll
                        for (int al = 0; al < SLa; al++)
                        {
                            double s = 0.0;
#pragma unroll
                            for (int be = 0; be < SLb; be++)
#pragma unroll
                                for (int ga = 0; ga < SUa; ga++)
                                {
                                    const int su = ga * SUb + de;
                                    if (su / NC != pass)
                                        continue;
                                    s = fma(double(I[(al * SLb + be) * NC + su % NC]),
                                            D[(u0 + ga) * n + l1 + be], s);
                                }
                            addA(u1 + de, al, f * s);
                        }
#pragma unroll
                        for (int be = 0; be < SLb; be++)
                        {
                            double s = 0.0;
#pragma unroll
                            for (int al = 0; al < SLa; al++)
#pragma unroll
                                for (int ga = 0; ga < SUa; ga++)
                                {
                                    const int su = ga * SUb + de;
                                    if (su / NC != pass)
                                        continue;
                                    s = fma(double(I[(al * SLb + be) * NC + su % NC]),
                                            D[(u0 + ga) * n + l0 + al], s);
                                }
                            if (active)
                                unsafeAtomicAdd(&A.K[(u1 + de) * n + l1 + be], f * s);
                        }
                    }
                }
            }
            else if constexpr (MODE == MODE_DENSE)
            {
                static_assert(!SWAP, "dense output keeps the bra on the lanes");
                if (active)
#pragma unroll
                    for (int s = 0; s < SL; s++)
#pragma unroll
                        for (int c = 0; c < NC; c++)
                            if (pass * NC + c < SU)
                                denseStore<C>(A, I[s * NC + c], s, pass * NC + c, l0, l1, u0, u1);
            }
            else if constexpr (MODE == MODE_SUM)
            {
                static_assert(!SWAP, "sums keep the bra on the lanes");
                if (active)
#pragma unroll
                    for (int s = 0; s < SL; s++)
#pragma unroll
                        for (int c = 0; c < NC; c++)
                            if (pass * NC + c < SU)
                            {
                                const double v = I[s * NC + c];
                                sabs += fabs(v);
                                ssum += v;
                            }
            }
            else
            {
                static_assert(MODE != MODE_DIAG || (LA == LC && LB == LD && !SWAP),
                              "diagonal needs ab == cd");
                if (active)
#pragma unroll
                    for (int c = 0; c < NC; c++)
                    {
                        const int s = pass * NC + c;
                        if (s >= SU)
                            continue;
                        const double v = I[s * NC + c];
                        dmax = fmax(dmax, fabs(v));
                        if (A.dout)
                        {
                            const int mu = s / SLb, nu = s % SLb;
                            if (A.kind == 2)
                                A.dout[l0 + mu] = v;
                            else
                            {
                                A.dout[(l0 + mu) * A.nbf + l1 + nu] = v;
                                A.dout[(l1 + nu) * A.nbf + l0 + mu] = v;
                            }
                        }
                    }
            }
        };
        forEachPass<NPASS>(passBody);

        if constexpr (MODE == MODE_JK && !CARRY)
        {
            if (active)
#pragma unroll
                for (int s = 0; s < SL; s++)
                    unsafeAtomicAdd(&A.J[(l0 + s / SLb) * n + l1 + s % SLb], jl[s]);
        }
        else if constexpr (MODE == MODE_DIAG)
        {
            if (active)
                A.qout[l] = sqrt(dmax);
        }
        else if constexpr (MODE == MODE_SUM)
        {
            // images in the dense tensor: (ab|P) and (ba|P) for distinct shells; (P|Q), (Q|P)
            const bool one = A.kind == 3 ? LP.sh_a[l] == LP.sh_b[l] : (A.same_class && p == q);
            const double w = active ? (one ? 1.0 : 2.0) : 0.0;
            sabs = waveSum(w * sabs);
            ssum = waveSum(w * ssum);
            if (lane == 0)
            {
                double *o = A.out + 4 * static_cast<long long>(uniform(UP.mol[u])) + (A.kind == 3 ? 0 : 2);
                unsafeAtomicAdd(o, sabs);
                unsafeAtomicAdd(o + 1, ssum);
            }
        }
    }

    // Uniform-major tables: each wave takes a contiguous range of tasks, which share the uniform
    // pair q and step through its 64-lane blocks (the task -> q search runs once per wave).
    // Lane-block-major tables (A.block_major): a task is one uniform pair q against one lane
    // block of a group of G blocks; a group's tasks are consecutive (q outer, block inner) and
    // wave w takes tasks w, w + nw, ... with nw the resident waves, so the waves in flight share
    // one group of lane-side pairs, whose E matrices (the larger per-task traffic) stay in L2,
    // and each uniform pair's E is read once per group.  With G dividing nw a wave stays on one
    // lane block within a group, so its J contributions to the lane-side pairs are summed in
    // registers and added to J once per block.
    template <int LA, int LB, int LC, int LD, int MODE, bool SWAP, typename T>
    __global__ __launch_bounds__(256) void quartetV(const KArgs A)
    {
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const long long nw = static_cast<long long>(gridDim.x) * 4;
        const long long w = static_cast<long long>(blockIdx.x) * 4 + wave;
        constexpr int SLb = nSph(SWAP ? LD : LB);
        constexpr int SL = nSph(SWAP ? LC : LA) * SLb;
        double jl[SL];
        if constexpr (MODE == MODE_JK)
        {
            if (A.block_major)
            {
                const DPairs &LP = SWAP ? A.ket : A.bra;
                auto flushJ = [&](int b) {
                    const int p = 64 * b + lane;
                    if (p < LP.n)
                    {
                        const int l = LP.order[A.lperm ? A.lperm[p] : p];
                        const long long r0 = LP.ao_a[l], r1 = LP.ao_b[l];
#pragma unroll
                        for (int s = 0; s < SL; s++)
                            unsafeAtomicAdd(&A.J[(r0 + s / SLb) * A.nbf + r1 + s % SLb], jl[s]);
                    }
                };
                const int gs = A.gshift;
                const int ng = (A.nblk + (1 << gs) - 1) >> gs;
                int g = 0, q0 = 0, b = -1;
                long long base = 0, next = 0;
                for (long long t = w; t < A.ntasks; t += nw)
                {
                    if (t >= next)
                    {
                        g = uniform(findKet(A.bprefix, ng, t));
                        base = A.bprefix[g];
                        next = A.bprefix[g + 1];
                        q0 = uniform(A.bq0[g]);
                    }
                    const long long local = t - base;
                    const int bt = (g << gs) + static_cast<int>(local & ((1LL << gs) - 1));
                    if (bt != b)
                    {
                        if (b >= 0)
                            flushJ(b);
#pragma unroll
                        for (int s = 0; s < SL; s++)
                            jl[s] = 0.0;
                        b = bt;
                    }
                    quartetVTask<LA, LB, LC, LD, MODE, SWAP, T, true>(A, q0 + static_cast<int>(local >> gs), 64 * b,
                                                                      lane, jl);
                }
                if (b >= 0)
                    flushJ(b);
                return;
            }
        }
        const long long t0 = A.ntasks * w / nw, t1 = A.ntasks * (w + 1) / nw;
        if (t0 >= t1)
            return;
        int q = uniform(findKet(A.prefix, A.nket, t0));
        long long base = A.prefix[q], next = A.prefix[q + 1];
        for (long long t = t0; t < t1; t++)
        {
            while (t >= next)
            {
                q++;
                base = next;
                next = A.prefix[q + 1];
            }
            quartetVTask<LA, LB, LC, LD, MODE, SWAP, T>(A, q, uniform(A.plo[q]) + 64 * static_cast<int>(t - base),
                                                        lane, jl);
        }
    }

    // ------------------------------------------------------------------------------- quartetM
    typedef float v4f __attribute__((ext_vector_type(4)));

    // 16x16x4 MFMA in FP64 or FP32.  Both take A[j][g] and B[g][j] from lane (g, j) = (lane / 16,
    // lane % 16); accumulator element r of that lane holds D[4 r + g][j] for FP64 and
    // D[4 g + r][j] for FP32.
    __device__ __forceinline__ v4d mfma16(double a, double b, v4d c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ __forceinline__ v4f mfma16(float a, float b, v4f c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }

    // T = double, or float (the small-bound quartets of mixed-precision J/K: R, the E matrices
    // and both transforms in FP32 on the FP32 MFMA; Boys values and the epilogue in FP64).
    template <int LA, int LB, int LC, int LD, int MODE, typename T = double>
    __global__ __launch_bounds__(64) void quartetM(const KArgs A)
    {
        using C = Cls<LA, LB, LC, LD>;
        constexpr int L = C::L, HP = C::HP, HQ = C::HQ, SP = C::SP, SQ = C::SQ, HL = C::HL;
        constexpr int NQ = C::NQ, NP = C::NP, KQ = C::KQ, MT = C::MT, RC = C::RC, CHP = C::CHP;
        constexpr int SA = C::SA, SB = C::SB, SC = C::SC, SD = C::SD;
        constexpr bool F32 = sizeof(T) == 4;
        using Acc = std::conditional_t<F32, v4f, v4d>;
        static_assert(!F32 || MODE == MODE_JK, "FP32 MFMA form: J/K only");
        static_assert(CHP <= 64, "one lane per primitive pair of a chunk");
        __shared__ T sR[2][CHP][HL];
        __shared__ double sInit[CHP][L + 1];
        __shared__ T sPQ[CHP][3];
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
            (&sR[0][0][0])[e] = T(0);
        __syncthreads();

        // contiguous task range per workgroup
        const long long t0 = A.ntasks * blockIdx.x / gridDim.x;
        const long long t1 = A.ntasks * (blockIdx.x + 1) / gridDim.x;
        if (t0 >= t1)
            return;
        int q = uniform(findKet(A.prefix, A.nket, t0));
        long long base = A.prefix[q], next = A.prefix[q + 1];
        for (long long t = t0; t < t1; t++)
        {
            while (t >= next)
            {
                q++;
                base = next;
                next = A.prefix[q + 1];
            }
            const int p = uniform(A.plo[q] + static_cast<int>(t - base));
            // task tables run over the ket pairs of one bra pair when bra_major is set (the
            // larger E matrix is then shared by consecutive tasks and stays in cache)
            const int cd = uniform(A.ket.order[A.bra_major ? p : q]);
            const int ab = uniform(A.bra.order[A.bra_major ? q : p]);
            const int kpoff = uniform(A.ket.poff[cd]), kcnt = uniform(A.ket.pcnt[cd]);
            const int bpoff = uniform(A.bra.poff[ab]), bcnt = uniform(A.bra.pcnt[ab]);
            const int nrows = bcnt * HP;

            Acc acc2[NQ][NP];
#pragma unroll
            for (int a = 0; a < NQ; a++)
#pragma unroll
                for (int b = 0; b < NP; b++)
                    acc2[a][b] = Acc{0, 0, 0, 0};

            for (int rowbase = 0; rowbase < nrows; rowbase += RC)
            {
                const int prim0 = rowbase / HP;
                const int primEnd = min(bcnt, (rowbase + RC - 1) / HP + 1);
                const int nprim = primEnd - prim0;

                Acc acc1[MT][NQ];
#pragma unroll
                for (int a = 0; a < MT; a++)
#pragma unroll
                    for (int b = 0; b < NQ; b++)
                        acc1[a][b] = Acc{0, 0, 0, 0};

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
                        const double Tb = alpha * (X_ * X_ + Y_ * Y_ + Z_ * Z_);
                        double F[L + 1];
                        boysF<L>(Tb, A.boys, F);
                        double sc = kTwoPi52 / (pq * sqrt(ps));
#pragma unroll
                        for (int m = 0; m <= L; m++)
                        {
                            sInit[lane][m] = F[m] * sc;
                            sc *= -2.0 * alpha;
                        }
                        sPQ[lane][0] = T(X_);
                        sPQ[lane][1] = T(Y_);
                        sPQ[lane][2] = T(Z_);
                        sR[L & 1][lane][0] = T(sInit[lane][L]);
                    }
                    __syncthreads();
                    // levels n = L-1 .. 0, ping-pong buffers
#pragma unroll
                    for (int n = L - 1; n >= 0; n--)
                    {
                        const int nh = nHerm(L - n) - 1;
                        const T(*in)[HL] = sR[(n + 1) & 1];
                        T(*out)[HL] = sR[n & 1];
                        const int items = nprim * nh;
                        for (int it = lane; it < items; it += 64)
                        {
                            const int pr = it / nh;
                            const int h = it - pr * nh + 1;
                            const unsigned r = kRecTab.v[h];
                            const int s1 = r & 2047, s2 = (r >> 11) & 2047, ax = (r >> 22) & 3,
                                      c = (r >> 24);
                            out[pr][h] = fma(sPQ[pr][ax], in[pr][s1], T(c) * in[pr][s2]);
                        }
                        if (lane < nprim)
                            out[lane][0] = T(sInit[lane][n]);
                        __syncthreads();
                    }
                    // ket transform on MFMA: acc1[mt][nt] += Rg[rows, k] Ek[k, cols]
                    const T *ek;
                    if constexpr (F32)
                        ek = A.ket.ek32 + static_cast<size_t>(kpoff + k) * HQ * SQ;
                    else
                        ek = A.ket.ek + static_cast<size_t>(kpoff + k) * HQ * SQ;
#pragma unroll
                    for (int kt = 0; kt < KQ; kt++)
                    {
                        const int kk = 4 * kt + g;
                        T b1[NQ];
#pragma unroll
                        for (int nt = 0; nt < NQ; nt++)
                        {
                            const int col = 16 * nt + j;
                            b1[nt] = (kk < HQ && col < SQ) ? ek[kk * SQ + col] : T(0);
                        }
#pragma unroll
                        for (int mt = 0; mt < MT; mt++)
                        {
                            if (rowbase + 16 * mt >= nrows)
                                continue;
                            const int gr = rowbase + 16 * mt + j;
                            const int pr = gr / HP;
                            const int h = gr - pr * HP;
                            const T a = (pr < bcnt) ? sR[0][pr - prim0][sTab[h][kk]] : T(0);
#pragma unroll
                            for (int nt = 0; nt < NQ; nt++)
                                acc1[mt][nt] = mfma16(a, b1[nt], acc1[mt][nt]);
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
                        // K-chunk c of X^T = accumulator element c of every lane: rows
                        // 16 mt + 4 c + (0..3) in FP64 (element c holds row 4 c + g), rows
                        // 16 mt + c + 4 (0..3) in FP32 (element c holds row 4 g + c)
                        const int gr = rowbase + 16 * mt + (F32 ? 4 * g + c : 4 * c + g);
                        const int pr = gr / HP;
                        const int h = gr - pr * HP;
                        const bool rv = pr < bcnt;
                        const T *eb;
                        if constexpr (F32)
                            eb = A.bra.eb32 + (static_cast<size_t>(bpoff + pr) * HP + h) * SP;
                        else
                            eb = A.bra.eb + (static_cast<size_t>(bpoff + pr) * HP + h) * SP;
#pragma unroll
                        for (int np = 0; np < NP; np++)
                        {
                            const int col = 16 * np + j;
                            const T b2 = (rv && col < SP) ? eb[col] : T(0);
#pragma unroll
                            for (int nq = 0; nq < NQ; nq++)
                                acc2[nq][np] = mfma16(acc1[mt][nq][c], b2, acc2[nq][np]);
                        }
                    }
                }
            }
            // accumulator element r of lane (g, j) holds D[4 r + g][j] (FP64) or D[4 g + r][j]
            // (FP32); here D = I^T[16 nq + row][16 np + j]
#pragma unroll
            for (int nq = 0; nq < NQ; nq++)
#pragma unroll
                for (int np = 0; np < NP; np++)
#pragma unroll
                    for (int r = 0; r < 4; r++)
                    {
                        const int kl = 16 * nq + (F32 ? 4 * g + r : 4 * r + g), mn = 16 * np + j;
                        if (kl < SQ && mn < SP)
                            sI[mn * SQ + kl] = double(acc2[nq][np][r]);
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
                    unsafeAtomicAdd(&A.K[(c0 + ka) * n + a0 + mu], f * s);
                }
                // K_bc[nu,ka] += f sum_{mu,la} I D_ad[mu,la]
                for (int e = lane; e < SB * SC; e += 64)
                {
                    const int nu = e / SC, ka = e % SC;
                    double s = 0.0;
                    for (int mu = 0; mu < SA; mu++)
                        for (int la = 0; la < SD; la++)
                            s = fma(sI[(mu * SB + nu) * SQ + ka * SD + la], sD[3][mu * SD + la], s);
                    unsafeAtomicAdd(&A.K[(c0 + ka) * n + b0 + nu], f * s);
                }
                // K_ad[mu,la] += f sum_{nu,ka} I D_bc[nu,ka]
                for (int e = lane; e < SA * SD; e += 64)
                {
                    const int mu = e / SD, la = e % SD;
                    double s = 0.0;
                    for (int nu = 0; nu < SB; nu++)
                        for (int ka = 0; ka < SC; ka++)
                            s = fma(sI[(mu * SB + nu) * SQ + ka * SD + la], sD[4][nu * SC + ka], s);
                    unsafeAtomicAdd(&A.K[(d0 + la) * n + a0 + mu], f * s);
                }
                // K_bd[nu,la] += f sum_{mu,ka} I D_ac[mu,ka]
                for (int e = lane; e < SB * SD; e += 64)
                {
                    const int nu = e / SD, la = e % SD;
                    double s = 0.0;
                    for (int mu = 0; mu < SA; mu++)
                        for (int ka = 0; ka < SC; ka++)
                            s = fma(sI[(mu * SB + nu) * SQ + ka * SD + la], sD[5][mu * SC + ka], s);
                    unsafeAtomicAdd(&A.K[(d0 + la) * n + b0 + nu], f * s);
                }
            }
            else if constexpr (MODE == MODE_DENSE)
            {
                for (int e = lane; e < SP * SQ; e += 64)
                    denseStore<C>(A, sI[e], e / SQ, e % SQ, a0, b0, c0, d0);
            }
            else if constexpr (MODE == MODE_SUM)
            {
                double sa = 0.0, ss = 0.0;
                for (int e = lane; e < SP * SQ; e += 64)
                {
                    sa += fabs(sI[e]);
                    ss += sI[e];
                }
                const bool one = A.kind == 3 ? A.bra.sh_a[ab] == A.bra.sh_b[ab] : (A.same_class && p == q);
                const double w = one ? 1.0 : 2.0;
                sa = waveSum(w * sa);
                ss = waveSum(w * ss);
                if (lane == 0)
                {
                    double *o = A.out + 4 * static_cast<long long>(A.ket.mol[cd]) + (A.kind == 3 ? 0 : 2);
                    unsafeAtomicAdd(o, sa);
                    unsafeAtomicAdd(o + 1, ss);
                }
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

    template <int LA, int LB, int LC, int LD, int MODE, bool SWAP, typename T>
    void launchValu(const KArgs &A, hipStream_t s)
    {
        if (A.ntasks <= 0)
            return;
        // uniform-major: contiguous task ranges per wave, up to 8192 workgroups of 4 waves;
        // lane-block-major: as many workgroups as are resident at once, so that one sweep of the
        // waves over the task order keeps them all on the same few lane blocks
        static const int resident = [] {
            int per_cu = 0, dev = 0, ncu = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                        &per_cu, reinterpret_cast<const void *>(&quartetV<LA, LB, LC, LD, MODE, SWAP, T>), 256, 0) !=
                        hipSuccess ||
                hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
                return 8192;
            return std::max(1, per_cu) * std::max(1, ncu);
        }();
        const long long cap = A.block_major ? std::min(resident, 8192) : 8192;
        long long wg = (A.ntasks + 3) / 4;
        int grid = static_cast<int>(wg < cap ? wg : cap);
        hipLaunchKernelGGL((quartetV<LA, LB, LC, LD, MODE, SWAP, T>), dim3(grid), dim3(256), 0, s, A);
    }

    template <int LA, int LB, int LC, int LD, int MODE>
    void launchMfma(const KArgs &A, hipStream_t s)
    {
        if (A.ntasks <= 0)
            return;
        int grid = static_cast<int>(A.ntasks < 16384 ? A.ntasks : 16384);
        hipLaunchKernelGGL((quartetM<LA, LB, LC, LD, MODE>), dim3(grid), dim3(64), 0, s, A);
    }

    template <int LA, int LB, int LC, int LD>
    void launchMfma32(const KArgs &A, hipStream_t s)
    {
        if (A.ntasks <= 0)
            return;
        int grid = static_cast<int>(A.ntasks < 16384 ? A.ntasks : 16384);
        hipLaunchKernelGGL((quartetM<LA, LB, LC, LD, MODE_JK, float>), dim3(grid), dim3(64), 0, s, A);
    }

    // The FP32 forms: one lane per quartet (J/K and RI sums), one wave per quartet on the FP32
    // MFMA (J/K).
    template <int LA, int LB, int LC, int LD, int MODE>
    void registerQuartet32()
    {
        using C = Cls<LA, LB, LC, LD>;
        constexpr bool same = LA == LC && LB == LD;
        if constexpr ((MODE == MODE_JK || MODE == MODE_SUM) && C::NC0f > 0)
        {
            KernelSlot &s = kernelSlot(MODE, LA, LB, LC, LD, VAR_VALU_BRA32);
            s.fn = &launchValu<LA, LB, LC, LD, MODE, false, float>;
            s.width = 64;
            s.swap = false;
            s.nc = C::NC0f;
        }
        if constexpr (MODE == MODE_JK && C::NC1f > 0 && !same)
        {
            KernelSlot &s = kernelSlot(MODE, LA, LB, LC, LD, VAR_VALU_KET32);
            s.fn = &launchValu<LA, LB, LC, LD, MODE, true, float>;
            s.width = 64;
            s.swap = true;
            s.nc = C::NC1f;
        }
        if constexpr (MODE == MODE_JK)
        {
            KernelSlot &s = kernelSlot(MODE, LA, LB, LC, LD, VAR_MFMA32);
            s.fn = &launchMfma32<LA, LB, LC, LD>;
            s.width = 1;
            s.swap = false;
            s.kind = KIND_WAVE;
        }
    }

    // Registers every variant the class supports in this mode.  The ket-on-lanes VALU form
    // only serves J/K between different classes (dense and diagonal outputs keep the bra on
    // the lanes; for one class both forms coincide).  The MFMA form is always available.
    // FORMS: 1 the FP64 forms, 2 the FP32 forms, 3 both (lets the build compile them apart).
    template <int LA, int LB, int LC, int LD, int MODE, int FORMS = 3>
    void registerQuartet()
    {
        using C = Cls<LA, LB, LC, LD>;
        if constexpr (FORMS & 2)
            registerQuartet32<LA, LB, LC, LD, MODE>();
        if constexpr (!(FORMS & 1))
            return;
        if constexpr (C::NC0 > 0)
        {
            KernelSlot &s = kernelSlot(MODE, LA, LB, LC, LD, VAR_VALU_BRA);
            s.fn = &launchValu<LA, LB, LC, LD, MODE, false, double>;
            s.width = 64;
            s.swap = false;
            s.nc = C::NC0;
        }
        constexpr bool same = LA == LC && LB == LD;
        if constexpr (MODE == MODE_JK && C::NC1 > 0 && !same)
        {
            KernelSlot &s = kernelSlot(MODE, LA, LB, LC, LD, VAR_VALU_KET);
            s.fn = &launchValu<LA, LB, LC, LD, MODE, true, double>;
            s.width = 64;
            s.swap = true;
            s.nc = C::NC1;
        }
        KernelSlot &s = kernelSlot(MODE, LA, LB, LC, LD, VAR_MFMA);
        s.fn = &launchMfma<LA, LB, LC, LD, MODE>;
        s.width = 1;
        s.swap = false;
        s.kind = KIND_WAVE;
    }
}
