// Block form of the J/K quartet kernel for gfx950 (variants VAR_BLOCK / VAR_BLOCK32).
//
// One workgroup of four waves per task: one bra pair (ab) against a block of up to NB ket pairs
// (cd)_j, consecutive positions of the ket class's Schwarz-sorted list.  The work follows the
// reference's SHARK order (eri_kernel_funs.hpp:104-193): per bra primitive pair i
//   X_i[tuv][(j, kl)] = sum_k sum_t'u'v' R_ijk(tuv + t'u'v') Ek_jk[t'u'v'][kl]   ket transform
//   I[mn][(j, kl)]   += sum_tuv Eb_i[tuv][mn] X_i[tuv][(j, kl)]                   bra transform
// so the bra transform runs once per bra primitive for the whole block, with the block's ket
// pairs stacked along N of the FP64 MFMA (v_mfma_f64_16x16x4f64), and the ket transform once
// per primitive quartet.  This is the canonical amount of work for every class pair, which the
// one-lane form misses for unbalanced pairs such as (fd|ps) (it repeats the bra transform per
// ket primitive).
//
// Phases of a task, per bra primitive i:
//   R + X  wave w takes ket pairs w, w + 4, ...; per chunk of up to PB ket primitives: Boys
//          function and R^(n)_000 (one lane per primitive quartet), the Hermite recursion level
//          by level in the wave's LDS buffers (lanes over (primitive, tuv)), then the ket
//          transform with one lane per row tuv of X, the ket's E matrix through scalar loads
//          (wave-uniform), X held in registers and written to the shared LDS buffer.
//   I      the four waves split the 16x16 tiles of I; A operand Eb_i from global memory (L2),
//          B operand X from LDS; accumulators stay in registers across i.
// Epilogue: I to LDS (over X), J of the bra pair summed over the block, J of each ket pair,
// and the four K blocks of every quartet by FP64 atomics (K and J are symmetrised by the
// caller).  The ket side's LDS synchronisation is wave-local (in-order LDS within a wave);
// the workgroup synchronises only around the MFMA phase and the epilogue.
#pragma once

#include "quartet.hpp"

namespace lbx
{
    template <int LA, int LB, int LC, int LD, typename T>
    struct BlockPlan
    {
        static constexpr int L = LA + LB + LC + LD;
        static constexpr int HA = nHerm(LA + LB), HB = nHerm(LC + LD), HL = nHerm(L);
        static constexpr int H1 = L >= 1 ? nHerm(L - 1) : 1;   // entries of recursion level 1
        static constexpr int SAa = nSph(LA), SAb = nSph(LB), SBc = nSph(LC), SBd = nSph(LD);
        static constexpr int SA = SAa * SAb, SB = SBc * SBd;
        static constexpr int NB = SB >= 48 ? 1 : 48 / SB;      // ket pairs per task
        static constexpr int NCOL = NB * SB;                    // columns of X and I
        static constexpr int MT = ceilDiv(SA, 16), NT = ceilDiv(NCOL, 16), KT = ceilDiv(HA, 4);
        static constexpr int NTILE = MT * NT, TPW = ceilDiv(NTILE, 4);   // tiles, per wave
        static constexpr int XS = 16 * NT + 1;                 // row stride of X (padded)
        static constexpr int XR = 4 * KT;                      // rows of X (zero from HA on)
        static constexpr int RP = ceilDiv(HA, 64);             // rows of X per lane
        // ket primitives per recursion chunk: about 8 KiB of level buffers per wave
        static constexpr int PB0 = 8192 / ((HL + H1) * static_cast<int>(sizeof(T)));
        static constexpr int PB = PB0 < 1 ? 1 : (PB0 > 16 ? 16 : PB0);
        static constexpr int CW = SB < 16 ? SB : 16;           // X columns per register pass
        static constexpr bool kRegIdx = HB <= 20;              // gather indices in registers
        static_assert(SA <= XR, "I is written over the X buffer");
        static_assert(NCOL <= 64, "one thread per column of the ket pairs' J");
    };

    // Hermite index of (t + t', u + u', v + v') for packed tuv values
    __device__ __forceinline__ int hermSum(int a, int b)
    {
        const int t = (a & 255) + (b & 255), u = ((a >> 8) & 255) + ((b >> 8) & 255),
                  v = (a >> 16) + (b >> 16);
        return hermIdx(t, u, v);
    }

    // LDS ordering inside one wave: a wave's LDS operations execute in issue order, so only the
    // compiler has to be kept from moving loads and stores across this point.
    __device__ __forceinline__ void waveFence()
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }

    template <int LA, int LB, int LC, int LD, typename T>
    __global__ __launch_bounds__(256) void quartetB(const KArgs A)
    {
        using P = BlockPlan<LA, LB, LC, LD, T>;
        constexpr int L = P::L, HA = P::HA, HB = P::HB, HL = P::HL, H1 = P::H1;
        constexpr int SA = P::SA, SB = P::SB, SAa = P::SAa, SAb = P::SAb, SBc = P::SBc, SBd = P::SBd;
        constexpr int NB = P::NB, NCOL = P::NCOL, NT = P::NT, KT = P::KT, NTILE = P::NTILE;
        constexpr int TPW = P::TPW, XS = P::XS, XR = P::XR, RP = P::RP, PB = P::PB, CW = P::CW;
        constexpr bool F32 = sizeof(T) == 4;
        using Acc = std::conditional_t<F32, v4f, v4d>;

        __shared__ T sX[XR * XS];               // X (rows tuv, columns (j, kl)); I in the epilogue
        __shared__ T sRe[4][PB * HL];           // recursion buffers per wave: even levels
        __shared__ T sRo[4][PB * H1];           // odd levels
        __shared__ double sInit[4][PB][L + 1];  // R^(n)_000 of the chunk's primitive quartets
        __shared__ T sPQ[4][PB][3];
        __shared__ unsigned short sTab[P::kRegIdx ? 1 : HB * HA];   // gather index [t'u'v'][tuv]
        __shared__ int sKet[NB][5];             // ket pair: pair index, first prim, prims, AO c, AO d
        __shared__ double sF[NB];               // permutational weight of each quartet
        __shared__ double sDA[SA], sDB[NCOL];   // D of the bra pair, of each ket pair

        const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
        const int g4 = lane >> 4, j16 = lane & 15;
        const long long n = A.nbf;
        const double *__restrict__ D = A.D;

        for (int e = tid; e < XR * XS; e += 256)
            sX[e] = T(0);
        if constexpr (!P::kRegIdx)
            for (int e = tid; e < HB * HA; e += 256)
                sTab[e] = static_cast<unsigned short>(hermSum(hermTuv(e % HA), hermTuv(e / HA)));
        // this lane's gather indices: rows lane + 64 r of X against every ket Hermite index
        int gi[RP][P::kRegIdx ? HB : 1];
        if constexpr (P::kRegIdx)
#pragma unroll
            for (int r = 0; r < RP; r++)
            {
                const int h = lane + 64 * r;
                const int th = h < HA ? hermTuv(h) : 0;
#pragma unroll
                for (int g = 0; g < HB; g++)
                    gi[r][g] = hermSum(th, hermTuv(g));
            }
        __syncthreads();

        // Workgroups b and b + 8 run on the same XCD: each XCD takes a contiguous range of the
        // tasks in flight, so one bra pair's blocks share that XCD's L2.
        const int G = gridDim.x;   // a multiple of 8
        const long long v = static_cast<long long>(blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3);
        for (long long t = v; t < A.ntasks; t += G)
        {
            const int q = uniform(findKet(A.prefix, A.nket, t));
            const int lo = uniform(A.plo[q] + static_cast<int>(t - A.prefix[q]) * NB);
            const int nb = uniform(min(NB, A.phi[q] - lo));
            const int ab = uniform(A.bra.order[q]);
            const int bpoff = uniform(A.bra.poff[ab]), bcnt = uniform(A.bra.pcnt[ab]);
            const int a0 = uniform(A.bra.ao_a[ab]), b0 = uniform(A.bra.ao_b[ab]);
            if (tid < nb)
            {
                const int pos = lo + tid;
                const int cd = A.ket.order[pos];
                sKet[tid][0] = cd;
                sKet[tid][1] = A.ket.poff[cd];
                sKet[tid][2] = A.ket.pcnt[cd];
                sKet[tid][3] = A.ket.ao_a[cd];
                sKet[tid][4] = A.ket.ao_b[cd];
                double f = A.bra.sh_a[ab] == A.bra.sh_b[ab] ? 0.5 : 1.0;
                if (A.ket.sh_a[cd] == A.ket.sh_b[cd])
                    f *= 0.5;
                if (A.same_class && pos == q)
                    f *= 0.5;
                sF[tid] = f;
            }
            if (tid >= 64 && tid < 64 + SA)
                sDA[tid - 64] = D[(a0 + (tid - 64) / SAb) * n + b0 + (tid - 64) % SAb];
            __syncthreads();
            if (tid < NCOL)
            {
                const int jj = tid / SB, c = tid % SB;
                sDB[tid] = jj < nb ? D[(sKet[jj][3] + c / SBd) * n + sKet[jj][4] + c % SBd] : 0.0;
            }

            Acc acc[TPW];
#pragma unroll
            for (int s = 0; s < TPW; s++)
                acc[s] = Acc{0, 0, 0, 0};

            for (int i = 0; i < bcnt; i++)
            {
                const double4 Pa = A.bra.pp[bpoff + i];
                // ---- Hermite integrals and ket transforms: wave w takes ket pairs w, w + 4, ...
                for (int jj = w; jj < NB; jj += 4)
                {
                    const int kcnt = jj < nb ? sKet[jj][2] : 0;
                    if (kcnt == 0)
                    {
                        for (int e = lane; e < HA * SB; e += 64)
                            sX[(e / SB) * XS + jj * SB + e % SB] = T(0);
                        continue;
                    }
                    const int kpoff = uniform(sKet[jj][1]);
                    const T *ekb;
                    if constexpr (F32)
                        ekb = A.ket.ek32;
                    else
                        ekb = A.ket.ek;
                    T *const bufE = sRe[w];
                    T *const bufO = sRo[w];
                    for (int k0 = 0; k0 < kcnt; k0 += PB)
                    {
                        const int np = min(PB, kcnt - k0);
                        if (lane < np)
                        {
                            const double4 Q = A.ket.pp[kpoff + k0 + lane];
                            const double pq = Pa.x * Q.x, ps = Pa.x + Q.x;
                            const double alpha = pq / ps;
                            const double X_ = Pa.y - Q.y, Y_ = Pa.z - Q.z, Z_ = Pa.w - Q.w;
                            double F[L + 1];
                            boysF<L>(alpha * (X_ * X_ + Y_ * Y_ + Z_ * Z_), A.boys, F);
                            double sc = kTwoPi52 / (pq * sqrt(ps));
#pragma unroll
                            for (int m = 0; m <= L; m++)
                            {
                                sInit[w][lane][m] = F[m] * sc;
                                sc *= -2.0 * alpha;
                            }
                            sPQ[w][lane][0] = T(X_);
                            sPQ[w][lane][1] = T(Y_);
                            sPQ[w][lane][2] = T(Z_);
                            if constexpr (L & 1)
                                bufO[lane * H1] = T(sInit[w][lane][L]);
                            else
                                bufE[lane * HL] = T(sInit[w][lane][L]);
                        }
                        waveFence();
                        // levels L-1 .. 0: level n holds R^(n)_tuv for |tuv| <= L - n
#pragma unroll
                        for (int lev = L - 1; lev >= 0; lev--)
                        {
                            const T *in = ((lev + 1) & 1) ? bufO : bufE;
                            T *out = (lev & 1) ? bufO : bufE;
                            const int si = ((lev + 1) & 1) ? H1 : HL, so = (lev & 1) ? H1 : HL;
                            const int nh = nHerm(L - lev) - 1;
                            const int items = np * nh;
                            for (int it = lane; it < items; it += 64)
                            {
                                const int p = it / nh;
                                const int h = it - p * nh + 1;
                                const unsigned r = kRecTab.v[h];
                                const int s1 = r & 2047, s2 = (r >> 11) & 2047, ax = (r >> 22) & 3,
                                          c = static_cast<int>(r >> 24);
                                out[p * so + h] = fma(sPQ[w][p][ax], in[p * si + s1], T(c) * in[p * si + s2]);
                            }
                            if (lane < np)
                                out[lane * so] = T(sInit[w][lane][lev]);
                            waveFence();
                        }
                        // ket transform: X[h][(jj, kl)] += sum_p sum_g R_p[h + g] Ek_p[g][kl]
#pragma unroll
                        for (int r = 0; r < RP; r++)
                        {
                            const int h = lane + 64 * r;
                            if (h < HA)
                            {
#pragma unroll
                                for (int c0 = 0; c0 < SB; c0 += CW)
                                {
                                    T x[CW];
#pragma unroll
                                    for (int c = 0; c < CW; c++)
                                        x[c] = (k0 > 0 && c0 + c < SB) ? sX[h * XS + jj * SB + c0 + c] : T(0);
                                    for (int p = 0; p < np; p++)
                                    {
                                        const cptr<T> e =
                                                constView(ekb + static_cast<size_t>(kpoff + k0 + p) * HB * SB);
                                        const T *rr = bufE + p * HL;
#pragma unroll
                                        for (int g = 0; g < HB; g++)
                                        {
                                            int idx;
                                            if constexpr (P::kRegIdx)
                                                idx = gi[r][g];
                                            else
                                                idx = sTab[g * HA + h];
                                            const T rv = rr[idx];
#pragma unroll
                                            for (int c = 0; c < CW; c++)
                                                if (c0 + c < SB)
                                                    x[c] = fma(rv, e[g * SB + c0 + c], x[c]);
                                        }
                                    }
#pragma unroll
                                    for (int c = 0; c < CW; c++)
                                        if (c0 + c < SB)
                                            sX[h * XS + jj * SB + c0 + c] = x[c];
                                }
                            }
                        }
                        waveFence();
                    }
                }
                __syncthreads();
                // ---- bra transform on MFMA: I[mn][col] += sum_h Eb_i[h][mn] X[h][col]
                const T *eb;
                if constexpr (F32)
                    eb = A.bra.eb32 + static_cast<size_t>(bpoff + i) * HA * SA;
                else
                    eb = A.bra.eb + static_cast<size_t>(bpoff + i) * HA * SA;
#pragma unroll 2
                for (int kt = 0; kt < KT; kt++)
                {
                    const int kk = 4 * kt + g4;
#pragma unroll
                    for (int s = 0; s < TPW; s++)
                    {
                        const int tile = w + 4 * s;
                        if (tile < NTILE)
                        {
                            const int mt = tile / NT, nt = tile - mt * NT;
                            const int row = 16 * mt + j16;
                            const T a = (kk < HA && row < SA) ? eb[kk * SA + row] : T(0);
                            const T b = sX[kk * XS + 16 * nt + j16];
                            acc[s] = mfma16(a, b, acc[s]);
                        }
                    }
                }
                __syncthreads();
            }

            // ---- epilogue: I (rows mn, columns (j, kl)) over the X buffer
#pragma unroll
            for (int s = 0; s < TPW; s++)
            {
                const int tile = w + 4 * s;
                if (tile < NTILE)
                {
                    const int mt = tile / NT, nt = tile - mt * NT;
#pragma unroll
                    for (int r = 0; r < 4; r++)
                    {
                        // accumulator element r of lane (g, j): row 4 r + g (FP64), 4 g + r (FP32)
                        const int row = 16 * mt + (F32 ? 4 * g4 + r : 4 * r + g4), col = 16 * nt + j16;
                        if (row < SA && col < NCOL)
                            sX[row * XS + col] = acc[s][r];
                    }
                }
            }
            __syncthreads();
            // J of the bra pair, summed over the block: J_ab += sum_j 2 f_j I_j D_cd_j
            if (tid < SA)
            {
                double sum = 0.0;
                for (int jj = 0; jj < nb; jj++)
                {
                    double sj = 0.0;
#pragma unroll
                    for (int c = 0; c < SB; c++)
                        sj = fma(static_cast<double>(sX[tid * XS + jj * SB + c]), sDB[jj * SB + c], sj);
                    sum = fma(2.0 * sF[jj], sj, sum);
                }
                unsafeAtomicAdd(&A.J[(a0 + tid / SAb) * n + b0 + tid % SAb], sum);
            }
            // J of each ket pair: J_cd += 2 f I^T D_ab
            if (tid >= 64 && tid < 64 + nb * SB)
            {
                const int col = tid - 64, jj = col / SB, c = col % SB;
                double sum = 0.0;
#pragma unroll
                for (int s = 0; s < SA; s++)
                    sum = fma(static_cast<double>(sX[s * XS + col]), sDA[s], sum);
                unsafeAtomicAdd(&A.J[(sKet[jj][3] + c / SBd) * n + sKet[jj][4] + c % SBd], 2.0 * sF[jj] * sum);
            }
            // K: per quartet the blocks (a c), (b c), (a d), (b d), each contracted with D of the
            // other two shells:  K[a][c] += f sum_{b,d} I D[b][d],  K[b][c] += f sum_{a,d} I D[a][d],
            //                    K[a][d] += f sum_{b,c} I D[b][c],  K[b][d] += f sum_{a,c} I D[a][c]
            constexpr int K0 = SAa * SBc, K1 = SAb * SBc, K2 = SAa * SBd, K3 = SAb * SBd;
            constexpr int NK = K0 + K1 + K2 + K3;
            for (int e = tid; e < nb * NK; e += 256)
            {
                const int jj = e / NK;
                int k = e - jj * NK;
                const int c0 = sKet[jj][3], d0 = sKet[jj][4];
                const T *I = sX + jj * SB;   // I[mn][kl] = I[mn * XS + kl]
                double sum = 0.0;
                long long dst;
                if (k < K0)
                {
                    const int mu = k / SBc, ka = k % SBc;
#pragma unroll
                    for (int nu = 0; nu < SAb; nu++)
#pragma unroll
                        for (int la = 0; la < SBd; la++)
                            sum = fma(static_cast<double>(I[(mu * SAb + nu) * XS + ka * SBd + la]),
                                      D[(b0 + nu) * n + d0 + la], sum);
                    dst = (a0 + mu) * n + c0 + ka;
                }
                else if ((k -= K0) < K1)
                {
                    const int nu = k / SBc, ka = k % SBc;
#pragma unroll
                    for (int mu = 0; mu < SAa; mu++)
#pragma unroll
                        for (int la = 0; la < SBd; la++)
                            sum = fma(static_cast<double>(I[(mu * SAb + nu) * XS + ka * SBd + la]),
                                      D[(a0 + mu) * n + d0 + la], sum);
                    dst = (b0 + nu) * n + c0 + ka;
                }
                else if ((k -= K1) < K2)
                {
                    const int mu = k / SBd, la = k % SBd;
#pragma unroll
                    for (int nu = 0; nu < SAb; nu++)
#pragma unroll
                        for (int ka = 0; ka < SBc; ka++)
                            sum = fma(static_cast<double>(I[(mu * SAb + nu) * XS + ka * SBd + la]),
                                      D[(b0 + nu) * n + c0 + ka], sum);
                    dst = (a0 + mu) * n + d0 + la;
                }
                else
                {
                    k -= K2;
                    const int nu = k / SBd, la = k % SBd;
#pragma unroll
                    for (int mu = 0; mu < SAa; mu++)
#pragma unroll
                        for (int ka = 0; ka < SBc; ka++)
                            sum = fma(static_cast<double>(I[(mu * SAb + nu) * XS + ka * SBd + la]),
                                      D[(a0 + mu) * n + c0 + ka], sum);
                    dst = (b0 + nu) * n + d0 + la;
                }
                unsafeAtomicAdd(&A.K[dst], sF[jj] * sum);
            }
            __syncthreads();
        }
    }

    template <int LA, int LB, int LC, int LD, typename T>
    void launchBlock(const KArgs &A, hipStream_t s)
    {
        if (A.ntasks <= 0)
            return;
        // persistent grid: the resident workgroups, a multiple of 8 (one share per XCD)
        static const int resident = [] {
            int per_cu = 0, dev = 0, ncu = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                        &per_cu, reinterpret_cast<const void *>(&quartetB<LA, LB, LC, LD, T>), 256, 0) !=
                        hipSuccess ||
                hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
                return 1024;
            return std::max(1, per_cu) * std::max(1, ncu);
        }();
        const long long want = std::min<long long>(A.ntasks, resident);
        const int grid = static_cast<int>((want + 7) / 8 * 8);
        hipLaunchKernelGGL((quartetB<LA, LB, LC, LD, T>), dim3(grid), dim3(256), 0, s, A);
    }

    // FORMS: 1 the FP64 form, 2 the FP32 form, 3 both.
    template <int LA, int LB, int LC, int LD, int FORMS>
    void registerBlock()
    {
        if constexpr (FORMS & 1)
        {
            KernelSlot &s = kernelSlot(MODE_JK, LA, LB, LC, LD, VAR_BLOCK);
            s.fn = &launchBlock<LA, LB, LC, LD, double>;
            s.width = BlockPlan<LA, LB, LC, LD, double>::NB;
            s.swap = false;
            s.kind = KIND_BLOCK;
        }
        if constexpr (FORMS & 2)
        {
            KernelSlot &s = kernelSlot(MODE_JK, LA, LB, LC, LD, VAR_BLOCK32);
            s.fn = &launchBlock<LA, LB, LC, LD, float>;
            s.width = BlockPlan<LA, LB, LC, LD, float>::NB;
            s.swap = false;
            s.kind = KIND_BLOCK;
        }
    }
}
