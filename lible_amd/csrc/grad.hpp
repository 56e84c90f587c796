// Nuclear gradient of the two-electron energy E = 1/2 tr(DJ) - 1/4 tr(DK), contracted on the
// fly (adjoint form): no derivative integral is stored.
//
// For a shell quartet (ab|cd) with weight W[ab][cd] = f (4 D_ab D_cd - D_ac D_bd - D_ad D_bc)
// (f the permutational weight of the unique quartet) the gradient is sum W d(ab|cd)/dX.  With
// the Hermite form (ab|cd) = sum_{ij} Eb_i[ab][h] Ek_j[cd][g] R_{h+g} (prim pairs i, j) and
// E a function of the relative coordinate A - B only (P = (aA + bB)/p moves the Hermite
// Gaussian, d/dP_x raises its x index by one):
//   dA = sum_ij [ (a/p) T_ij + U_ij ],   dB = sum_ij [ (b/p) T_ij - U_ij ],
//   dC = sum_ij [ V_ij - (c/q) T_ij ],   dD = -(dA + dB + dC)
// with, per Cartesian direction x,
//   T_ij = sum W Eb_i Ek_j R_{h+g+e_x},  U_ij = sum W Eb'_i,x Ek_j R_{h+g},
//   V_ij = sum W Eb_i Ek'_j,x R_{h+g}   (E' = dE/d(A-B)_x; the ket's (-1)^{t'+u'+v'} in Ek, Ek').
// This is the quantity of the reference's ERI4D1 kernel (eri_kernel_funs.cpp:582-748) contracted
// with the density weights.
//
// Evaluation, per primitive pair i of the lane-side pair and column pass of cd:
//   Y[h][cd] = sum_ab Eb_i[ab][h] W[ab][cd]
//   per ket primitive j:  T += sum_{h,g} (sum_cd Y Ek_j) R_{h+g+e_x},  V += sum_{h,g} R_{h+g}
//                         (sum_cd Y Ek'_j,x),   X[h][cd] += sum_g R_{h+g} Ek_j[cd][g]
//   then U = sum_{ab,cd} W[ab][cd] sum_h Eb'_i,x[ab][h] X[h][cd].
// Work layout: one wave per task; a task is G lane-side pairs against one uniform-side pair;
// each pair (slot) gets 64/G lanes that split its Hermite index h; R is built cooperatively in
// LDS per slot, W sits in LDS per slot.
#pragma once

#include "device_common.hpp"
#include "lbx_internal.hpp"

namespace lbx
{
    struct GArgs
    {
        DPairs lane, unif;           // lane-side class (the "bra" of the formulas), uniform side
        const long long *prefix;     // per uniform-side position: first task
        const int *plo, *phi;        // lane-side sorted-position range per uniform position
        int nket;
        long long ntasks;
        int same_class;
        const double *boys;
        const double *D;
        long long nbf;
        double *grad;                // 3 natoms
    };

    using GradLauncher = void (*)(const GArgs &, hipStream_t);
    struct GradSlot
    {
        GradLauncher fn = nullptr;     // FP64
        GradLauncher fn32 = nullptr;   // R, E, E' and the transforms in FP32, sums in FP64
        int width = 0;   // lane-side pairs per task
    };
    // (la lb) lane side, (lc ld) uniform side
    GradSlot &gradSlot(int la, int lb, int lc, int ld);

    constexpr LBX_HD int pow2Ceil(int x)
    {
        int p = 1;
        while (p < x)
            p *= 2;
        return p;
    }

    // lane side chosen for an unordered class pair: the larger Hermite dimension (more lanes in
    // use), ties to the larger class index
    constexpr LBX_HD bool gradLaneFirst(int la, int lb, int lc, int ld)
    {
        return nHerm(la + lb) > nHerm(lc + ld) ||
               (nHerm(la + lb) == nHerm(lc + ld) && pairClassIdx(la, lb) >= pairClassIdx(lc, ld));
    }

    template <int LA, int LB, int LC, int LD>
    struct GCls
    {
        static constexpr int L = LA + LB + LC + LD, L1 = L + 1;
        static constexpr int HP = nHerm(LA + LB), HQ = nHerm(LC + LD), HL1 = nHerm(L1);
        static constexpr int SA = nSph(LA), SB = nSph(LB), SC = nSph(LC), SD = nSph(LD);
        static constexpr int SP = SA * SB, SQ = SC * SD;
        static constexpr int kLds = 6144;   // doubles of LDS per wave (48 KiB)
        static constexpr int perSlot = 2 * HL1 + SP * SQ;
        static constexpr int G0 = 64 / (HP >= 64 ? 64 : pow2Ceil(HP));
        static constexpr int gFit(int g) { return (g > 1 && g * perSlot > kLds) ? gFit(g / 2) : g; }
        static constexpr int G = gFit(G0);
        static constexpr int LPS = 64 / G;                 // lanes per slot
        static constexpr int NH = ceilDiv(HP, LPS);        // h values per lane
        static constexpr int NC = SQ < 8 ? SQ : 8;         // cd columns per pass
        static constexpr int NPASS = ceilDiv(SQ, NC);
        static_assert(G * perSlot <= kLds, "gradient LDS budget");
    };

    // sum over the LPS lanes of one slot (contiguous lane groups)
    template <int LPS>
    __device__ __forceinline__ double slotSum(double x)
    {
#pragma unroll
        for (int o = LPS / 2; o >= 1; o >>= 1)
            x += __shfl_xor(x, o, 64);
        return x;
    }

    __device__ __forceinline__ int waveMaxInt(int x)
    {
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1)
            x = max(x, __shfl_xor(x, o, 64));
        return x;
    }

    // T: type of R, the E and E' matrices and the transforms (double, or float for the
    // small-bound quartets of the mixed-precision gradient); Boys values, density weights and the
    // gradient sums stay double.
    template <int LA, int LB, int LC, int LD, typename T>
    __global__ __launch_bounds__(64) void gradKernel(const GArgs A)
    {
        using C = GCls<LA, LB, LC, LD>;
        constexpr int L1 = C::L1, HP = C::HP, HQ = C::HQ, HL1 = C::HL1, SP = C::SP, SQ = C::SQ,
                      SB = C::SB, SD = C::SD, G = C::G, LPS = C::LPS, NH = C::NH, NC = C::NC;
        constexpr bool F32 = sizeof(T) == 4;
        __shared__ T sR[2][G][HL1];
        __shared__ double sW[G][SP * SQ];
        const int lane = threadIdx.x;
        const int slot = lane / LPS, jl = lane % LPS;
        const DPairs &LP = A.lane, &UP = A.unif;
        const long long n = A.nbf;
        const double *D = A.D;

        // (t, u, v) of this lane's Hermite indices
        int th[NH], uh[NH], vh[NH];
        bool hv[NH];
#pragma unroll
        for (int r = 0; r < NH; r++)
        {
            const int h = jl + r * LPS;
            hv[r] = h < HP;
            const int tuv = hermTuv(hv[r] ? h : 0);
            th[r] = tuv & 255;
            uh[r] = (tuv >> 8) & 255;
            vh[r] = tuv >> 16;
        }

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
            const int local = static_cast<int>(t - base);
            const int plo = uniform(A.plo[q]), phi = uniform(A.phi[q]);
            const int p = plo + local * G + slot;
            const bool active = p < phi;
            const int u = uniform(UP.order[q]);
            const int l = LP.order[active ? p : plo];
            const int upoff = uniform(UP.poff[u]), ucnt = uniform(UP.pcnt[u]);
            const int lpoff = LP.poff[l], lcnt = active ? LP.pcnt[l] : 0;
            const int imax = uniform(waveMaxInt(lcnt));
            const int a0 = LP.ao_a[l], b0 = LP.ao_b[l];
            const int c0 = uniform(UP.ao_a[u]), d0 = uniform(UP.ao_b[u]);
            double f = active ? 1.0 : 0.0;
            if (LP.sh_a[l] == LP.sh_b[l])
                f *= 0.5;
            if (UP.sh_a[u] == UP.sh_b[u])
                f *= 0.5;
            if (A.same_class && p == q)
                f *= 0.5;

            // density weights of this slot's quartet
            for (int e = jl; e < SP * SQ; e += LPS)
            {
                const int ab = e / SQ, cd = e - ab * SQ;
                const int al = ab / SB, be = ab - al * SB, ga = cd / SD, de = cd - ga * SD;
                const long long ia = a0 + al, ib = b0 + be, ic = c0 + ga, id = d0 + de;
                sW[slot][e] = active ? f * (4.0 * D[ia * n + ib] * D[ic * n + id] -
                                            D[ia * n + ic] * D[ib * n + id] - D[ia * n + id] * D[ib * n + ic])
                                     : 0.0;
            }
            __syncthreads();

            double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0}, gC[3] = {0, 0, 0};
#pragma unroll 1
            for (int pass = 0; pass < C::NPASS; pass++)
            {
                const int cd0 = pass * NC;
#pragma unroll 1
                for (int i = 0; i < imax; i++)
                {
                    const bool valid = i < lcnt;
                    const int ip = lpoff + (valid ? i : 0);
                    // Y[h][c] = sum_ab Eb_i[h][ab] W[ab][cd0 + c]
                    T Y[NH][NC], X[NH][NC];
                    const T *eb;
                    if constexpr (F32)
                        eb = LP.eb32 + static_cast<size_t>(ip) * HP * SP;
                    else
                        eb = LP.eb + static_cast<size_t>(ip) * HP * SP;
#pragma unroll
                    for (int r = 0; r < NH; r++)
#pragma unroll
                        for (int c = 0; c < NC; c++)
                        {
                            Y[r][c] = T(0);
                            X[r][c] = T(0);
                        }
                    for (int ab = 0; ab < SP; ab++)
                    {
                        T w[NC];
#pragma unroll
                        for (int c = 0; c < NC; c++)
                            w[c] = cd0 + c < SQ ? T(sW[slot][ab * SQ + cd0 + c]) : T(0);
#pragma unroll
                        for (int r = 0; r < NH; r++)
                        {
                            const T e = hv[r] ? eb[(jl + r * LPS) * SP + ab] : T(0);
#pragma unroll
                            for (int c = 0; c < NC; c++)
                                Y[r][c] = fma(e, w[c], Y[r][c]);
                        }
                    }
                    const double4 Pl = LP.pp[ip];
                    double Ts[3] = {0, 0, 0};
                    const cptr<double> ppU = constView(reinterpret_cast<const double *>(UP.pp + upoff));
                    // Boys table row of ket primitive k + 1 fetched while k is computed
                    auto argOf = [&](int k) {
                        const double X_ = Pl.y - ppU[4 * k + 1], Y_ = Pl.z - ppU[4 * k + 2],
                                     Z_ = Pl.w - ppU[4 * k + 3];
                        const double pq = Pl.x * ppU[4 * k], ps = Pl.x + ppU[4 * k];
                        return pq / ps * (X_ * X_ + Y_ * Y_ + Z_ * Z_);
                    };
                    BoysRow next;
                    boysFetch<L1>(argOf(0), A.boys, next);
#pragma unroll 1
                    for (int k = 0; k < ucnt; k++)
                    {
                        const BoysRow cur = next;
                        if (k + 1 < ucnt)
                            boysFetch<L1>(argOf(k + 1), A.boys, next);
                        const double4 Pu{ppU[4 * k], ppU[4 * k + 1], ppU[4 * k + 2], ppU[4 * k + 3]};
                        const double pq = Pl.x * Pu.x, ps = Pl.x + Pu.x, alpha = pq / ps;
                        const double PX = Pl.y - Pu.y, PY = Pl.z - Pu.z, PZ = Pl.w - Pu.w;
                        double F[L1 + 1];
                        boysEval<L1>(cur, F);
                        const double fac = kTwoPi52 / (pq * sqrt(ps));
                        double init[L1 + 1];
                        {
                            double sc = fac;
#pragma unroll
                            for (int m = 0; m <= L1; m++)
                            {
                                init[m] = F[m] * sc;
                                sc *= -2.0 * alpha;
                            }
                        }
                        // levels n = L1 .. 0 of R^{(n)}_{tuv}, |tuv| <= L1 - n, ping-pong in LDS
                        if (jl == 0)
                            sR[L1 & 1][slot][0] = T(init[L1]);
                        __syncthreads();
#pragma unroll
                        for (int lev = L1 - 1; lev >= 0; lev--)
                        {
                            const int nh = nHerm(L1 - lev);
                            const T *in = sR[(lev + 1) & 1][slot];
                            T *out = sR[lev & 1][slot];
                            for (int h = jl; h < nh; h += LPS)
                            {
                                if (h == 0)
                                {
                                    out[0] = T(init[lev]);
                                    continue;
                                }
                                const unsigned rt = kRecTab.v[h];
                                const int s1 = rt & 2047, s2 = (rt >> 11) & 2047, ax = (rt >> 22) & 3,
                                          cm = rt >> 24;
                                const T xa = T(ax == 0 ? PX : (ax == 1 ? PY : PZ));
                                out[h] = fma(xa, in[s1], T(cm) * in[s2]);
                            }
                            __syncthreads();
                        }
                        const T *R = sR[0][slot];
                        if (valid)
                        {
                            const size_t kk = static_cast<size_t>(upoff + k);
                            const T *ekb, *e1b;
                            if constexpr (F32)
                            {
                                ekb = UP.ek32;
                                e1b = UP.e1k32;
                            }
                            else
                            {
                                ekb = UP.ek;
                                e1b = UP.e1k;
                            }
                            const cptr<T> ek = constView(ekb + kk * HQ * SQ + cd0);
                            const cptr<T> e1 = constView(e1b + kk * 3 * HQ * SQ + cd0);
                            double tk[3] = {0, 0, 0};
#pragma unroll
                            for (int g = 0; g < HQ; g++)
                            {
                                const int tuvg = hermTuv(g);
                                const int tg = tuvg & 255, ug = (tuvg >> 8) & 255, vg = tuvg >> 16;
                                T ekv[NC], e1v[3][NC];
#pragma unroll
                                for (int c = 0; c < NC; c++)
                                {
                                    const bool in = cd0 + c < SQ;
                                    ekv[c] = in ? ek[g * SQ + c] : T(0);
#pragma unroll
                                    for (int x = 0; x < 3; x++)
                                        e1v[x][c] = in ? e1[(x * HQ + g) * SQ + c] : T(0);
                                }
#pragma unroll
                                for (int r = 0; r < NH; r++)
                                {
                                    if (!hv[r])
                                        continue;
                                    const int tt = th[r] + tg, uu = uh[r] + ug, vv = vh[r] + vg;
                                    const T r0 = R[hermIdx(tt, uu, vv)];
                                    const T rx = R[hermIdx(tt + 1, uu, vv)];
                                    const T ry = R[hermIdx(tt, uu + 1, vv)];
                                    const T rz = R[hermIdx(tt, uu, vv + 1)];
                                    T z = T(0), zx = T(0), zy = T(0), zz = T(0);
#pragma unroll
                                    for (int c = 0; c < NC; c++)
                                    {
                                        z = fma(Y[r][c], ekv[c], z);
                                        zx = fma(Y[r][c], e1v[0][c], zx);
                                        zy = fma(Y[r][c], e1v[1][c], zy);
                                        zz = fma(Y[r][c], e1v[2][c], zz);
                                        X[r][c] = fma(r0, ekv[c], X[r][c]);
                                    }
                                    tk[0] = fma(double(z), double(rx), tk[0]);
                                    tk[1] = fma(double(z), double(ry), tk[1]);
                                    tk[2] = fma(double(z), double(rz), tk[2]);
                                    gC[0] = fma(double(r0), double(zx), gC[0]);
                                    gC[1] = fma(double(r0), double(zy), gC[1]);
                                    gC[2] = fma(double(r0), double(zz), gC[2]);
                                }
                            }
                            const double rc = constView(UP.ra)[upoff + k];   // c / q
#pragma unroll
                            for (int x = 0; x < 3; x++)
                            {
                                Ts[x] += tk[x];
                                gC[x] = fma(-rc, tk[x], gC[x]);
                            }
                        }
                        __syncthreads();
                    }
                    if (valid)
                    {
                        const double ra = LP.ra[ip];   // a / p
#pragma unroll
                        for (int x = 0; x < 3; x++)
                        {
                            gA[x] = fma(ra, Ts[x], gA[x]);
                            gB[x] = fma(1.0 - ra, Ts[x], gB[x]);
                        }
                        // U_x = sum_ab sum_c W[ab][cd0 + c] sum_h Eb'_x[h][ab] X[h][c]
                        const T *e1;
                        if constexpr (F32)
                            e1 = LP.e1b32 + static_cast<size_t>(ip) * 3 * HP * SP;
                        else
                            e1 = LP.e1b + static_cast<size_t>(ip) * 3 * HP * SP;
                        double ux[3] = {0, 0, 0};
                        for (int ab = 0; ab < SP; ab++)
                        {
                            T w[NC];
#pragma unroll
                            for (int c = 0; c < NC; c++)
                                w[c] = cd0 + c < SQ ? T(sW[slot][ab * SQ + cd0 + c]) : T(0);
#pragma unroll
                            for (int r = 0; r < NH; r++)
                            {
                                if (!hv[r])
                                    continue;
                                T s = T(0);
#pragma unroll
                                for (int c = 0; c < NC; c++)
                                    s = fma(w[c], X[r][c], s);
                                const int h = jl + r * LPS;
#pragma unroll
                                for (int x = 0; x < 3; x++)
                                    ux[x] = fma(double(e1[(x * HP + h) * SP + ab]), double(s), ux[x]);
                            }
                        }
#pragma unroll
                        for (int x = 0; x < 3; x++)
                        {
                            gA[x] += ux[x];
                            gB[x] -= ux[x];
                        }
                    }
                }
            }
#pragma unroll
            for (int x = 0; x < 3; x++)
            {
                gA[x] = slotSum<LPS>(gA[x]);
                gB[x] = slotSum<LPS>(gB[x]);
                gC[x] = slotSum<LPS>(gC[x]);
            }
            if (active && jl == 0)
            {
                const int at[4] = {LP.atom_a[l], LP.atom_b[l], UP.atom_a[u], UP.atom_b[u]};
#pragma unroll
                for (int x = 0; x < 3; x++)
                {
                    unsafeAtomicAdd(&A.grad[3 * at[0] + x], gA[x]);
                    unsafeAtomicAdd(&A.grad[3 * at[1] + x], gB[x]);
                    unsafeAtomicAdd(&A.grad[3 * at[2] + x], gC[x]);
                    unsafeAtomicAdd(&A.grad[3 * at[3] + x], -(gA[x] + gB[x] + gC[x]));
                }
            }
            __syncthreads();   // sW of the next task
        }
    }

    template <int LA, int LB, int LC, int LD, typename T>
    void launchGrad(const GArgs &A, hipStream_t s)
    {
        if (A.ntasks <= 0)
            return;
        const int grid = static_cast<int>(A.ntasks < 16384 ? A.ntasks : 16384);
        hipLaunchKernelGGL((gradKernel<LA, LB, LC, LD, T>), dim3(grid), dim3(64), 0, s, A);
    }

    template <int LA, int LB, int LC, int LD>
    void registerGrad()
    {
        GradSlot &s = gradSlot(LA, LB, LC, LD);
        s.fn = &launchGrad<LA, LB, LC, LD, double>;
        s.fn32 = &launchGrad<LA, LB, LC, LD, float>;
        s.width = GCls<LA, LB, LC, LD>::G;
    }
}
