// Quartet kernels for the class pairs with a g shell in the main basis: one kernel for all of
// them, with the angular momenta at run time (the reference's generic-L kernels,
// eri_kernel_funs.cpp:7-101, serve every class above its rolled-out ones the same way).
//
// One wave per shell quartet (the task tables of the one-wave MFMA form, so every J/K, dense,
// diagonal, per-molecule-sum and gradient driver runs these classes unchanged).  Per quartet,
// the ket columns are taken in chunks of whole c functions sized to about 24 KiB of LDS (at
// most 64 KiB), so that several workgroups stay resident per CU; per chunk and bra primitive
// pair i (the chunk's block I in registers, then in LDS for the epilogue)
//     Y[h][cd] = sum_k sum_g R^{ik}_{h+g} Ek_k[g][cd]         (R of each primitive quartet in LDS)
//     I[ab][cd] += sum_h Eb_i[h][ab] Y[h][cd]
// and the chunk's block I goes through the epilogue of the mode (the contractions of the MFMA
// form's epilogue, restricted to the chunk's columns and added atomically).
//
// The gradient (E = 1/2 tr(DJ) - 1/4 tr(DK)) in the adjoint form of the compiled gradient
// kernels: with W = f (4 D_ab D_cd - D_ac D_bd - D_ad D_bc), per chunk and bra primitive i,
//     Y[h][cd] = sum_ab Eb_i[h][ab] W[ab][cd]
//     per ket primitive k:  T_x = sum_{h,g} (sum_cd Y Ek_k) R_{h+g+e_x},
//                           V_x = sum_{h,g} (sum_cd Y Ek'_k,x) R_{h+g},
//                           X[h][cd] += sum_g R_{h+g} Ek_k[g][cd]
//     U_x = sum_{ab,cd} W[ab][cd] sum_h Eb'_i,x[h][ab] X[h][cd]
//     dA += (a/p) T + U,  dB += (b/p) T - U,  dC += V - (c/q) T,  dD = -(dA + dB + dC).
#include "grad.hpp"
#include "quartet.hpp"
#include "runtime_l.hpp"

#include <utility>

namespace lbx
{
    namespace
    {
        constexpr int kGenThreads = 256;         // four waves share one quartet of a large class
        constexpr int kGenLds = 64 * 1024 / 8;   // doubles of LDS per workgroup, at most
        // below this many doubles the column chunk grows no further, so that four workgroups
        // (16 waves) stay resident per CU (R is rebuilt per chunk, a small cost)
        constexpr int kGenLdsTarget = 5 * 1024;

        // LDS layout of one class pair (offsets in doubles) and its column chunk
        struct GenPlan
        {
            int la, lb, lc, ld;
            int L;                   // highest R level (the class's total L, + 1 for the gradient)
            int nc;                  // c functions per column chunk
            int oR0, oR1, oTuv, oF;  // R ping-pong, packed (t, u, v) table, Boys values
            int oDab, oDbd, oDad;    // density blocks fixed per quartet
            int oDcd, oDbc, oDac;    // density blocks of the chunk
            int oY, oX, oW, oI;      // chunk work arrays
            int oRed;                // one double per wave: cross-wave maximum
            int oSum;                // quartet kernel: R index of (h, g), Hab x Hcd ushorts, or -1
            int oE;                  // quartet kernel: the ket primitive's E columns of the chunk, or -1
            int nt;                  // threads per workgroup: 64 or kGenThreads
            int doubles;
        };

        // jk: J/K density blocks; grad: the gradient's arrays (W, Y, X) instead of (Y, I)
        GenPlan makePlan(int la, int lb, int lc, int ld, bool jk, bool grad)
        {
            GenPlan g{};
            g.la = la;
            g.lb = lb;
            g.lc = lc;
            g.ld = ld;
            g.L = la + lb + lc + ld + (grad ? 1 : 0);
            const int HL = nHerm(g.L), Hab = nHerm(la + lb);
            const int Sa = nSph(la), Sb = nSph(lb), Sc = nSph(lc), Sd = nSph(ld), Sab = Sa * Sb;
            const bool dens = jk || grad;
            int o = 0;
            auto take = [&](int n) {
                const int at = o;
                o += n;
                return at;
            };
            g.oR0 = take(HL);
            g.oR1 = take(HL);
            g.oTuv = take((HL + 1) / 2);
            g.oF = take(g.L + 1);
            g.oDab = take(dens ? Sab : 0);
            g.oDbd = take(dens ? Sb * Sd : 0);
            g.oDad = take(dens ? Sa * Sd : 0);
            const int fixed = o;
            // per c function of the chunk
            const int perC = (dens ? Sd + Sb + Sa : 0) + (grad ? (Sab + 2 * Hab) * Sd : (Hab + Sab) * Sd);
            g.nc = std::min(Sc, std::max(1, (kGenLdsTarget - fixed - 4) / perC));
            if (g.nc < 1 || fixed + 4 + g.nc * perC > kGenLds)
                throw Error(LBX_ERR_INVALID, "class pair too large for the run-time-L kernel");
            const int cw = g.nc * Sd;
            g.oDcd = take(dens ? cw : 0);
            g.oDbc = take(dens ? Sb * g.nc : 0);
            g.oDac = take(dens ? Sa * g.nc : 0);
            g.oY = take(Hab * cw);
            g.oX = take(grad ? Hab * cw : 0);
            g.oW = take(grad ? Sab * cw : 0);
            g.oI = take(grad ? 0 : Sab * cw);
            g.oRed = take(kGenThreads / 64);
            // the R index of every (h, g) pair of the ket transform, when it takes at most 8 KiB
            const int nsum = Hab * nHerm(lc + ld);
            g.oSum = !grad && nsum <= 4096 ? take((nsum + 3) / 4) : -1;
            // the ket primitive's E for the chunk's columns, staged once for all h, when it
            // takes at most 8 KiB
            const int ne = nHerm(lc + ld) * cw;
            g.oE = !grad && ne <= 1024 ? take(ne) : -1;
            g.doubles = o;
            if (o > kGenLds)
                throw Error(LBX_ERR_INVALID, "class pair too large for the run-time-L kernel");
            // one wave when the widest loop per primitive quartet (Y, or the gradient's (h, g)
            // pairs) has little work, four otherwise
            const int widest = grad ? Hab * nHerm(lc + ld) : Hab * cw;
            g.nt = widest < 512 ? 64 : kGenThreads;
            return g;
        }

        __device__ __forceinline__ int hsum(int a, int b)
        {
            return hermIdx((a & 255) + (b & 255), ((a >> 8) & 255) + ((b >> 8) & 255), (a >> 16) + (b >> 16));
        }

        template <int NT>
        __global__ __launch_bounds__(NT) void genericQuartet(const KArgs A, const GenPlan G, const int mode)
        {
            extern __shared__ double smem[];
            double *const sR0 = smem + G.oR0, *const sR1 = smem + G.oR1, *const sF = smem + G.oF;
            int *const sTuv = reinterpret_cast<int *>(smem + G.oTuv);
            double *const sY = smem + G.oY, *const sI = smem + G.oI;
            double *const dab = smem + G.oDab, *const dbd = smem + G.oDbd, *const dad = smem + G.oDad;
            double *const dcd = smem + G.oDcd, *const dbc = smem + G.oDbc, *const dac = smem + G.oDac;
            const int lane = threadIdx.x;
            const int Hab = nHerm(G.la + G.lb), Hcd = nHerm(G.lc + G.ld);
            const int Sa = nSph(G.la), Sb = nSph(G.lb), Sc = nSph(G.lc), Sd = nSph(G.ld);
            const int Sab = Sa * Sb, Scd = Sc * Sd;
            const long long n = A.nbf;
            for (int h = lane; h < nHerm(G.L); h += NT)
                sTuv[h] = hermTuv(h);
            __syncthreads();
            unsigned short *const sSum = G.oSum >= 0 ? reinterpret_cast<unsigned short *>(smem + G.oSum) : nullptr;
            double *const sE = G.oE >= 0 ? smem + G.oE : nullptr;
            if (sSum)
            {
                for (int e = lane; e < Hab * Hcd; e += NT)
                    sSum[e] = static_cast<unsigned short>(hsum(sTuv[e / Hcd], sTuv[e % Hcd]));
                __syncthreads();
            }

            // contiguous task range per workgroup (the one-wave MFMA form's walk)
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
                const int cd = uniform(A.ket.order[A.bra_major ? p : q]);
                const int ab = uniform(A.bra.order[A.bra_major ? q : p]);
                const int kpoff = uniform(A.ket.poff[cd]), kcnt = uniform(A.ket.pcnt[cd]);
                const int bpoff = uniform(A.bra.poff[ab]), bcnt = uniform(A.bra.pcnt[ab]);
                const int a0 = A.bra.ao_a[ab], b0 = A.bra.ao_b[ab];
                const int c0 = A.ket.ao_a[cd], d0 = A.ket.ao_b[cd];
                double f = 1.0;
                if (mode == MODE_JK)
                {
                    if (A.bra.sh_a[ab] == A.bra.sh_b[ab])
                        f *= 0.5;
                    if (A.ket.sh_a[cd] == A.ket.sh_b[cd])
                        f *= 0.5;
                    if (A.same_class && p == q)
                        f *= 0.5;
                    for (int e = lane; e < Sab; e += NT)
                        dab[e] = A.D[(a0 + e / Sb) * n + b0 + e % Sb];
                    for (int e = lane; e < Sb * Sd; e += NT)
                        dbd[e] = A.D[(b0 + e / Sd) * n + d0 + e % Sd];
                    for (int e = lane; e < Sa * Sd; e += NT)
                        dad[e] = A.D[(a0 + e / Sd) * n + d0 + e % Sd];
                }
                double dmax = 0.0, sabs = 0.0, ssum = 0.0;

                for (int kc0 = 0; kc0 < Sc; kc0 += G.nc)
                {
                    const int ncw = min(G.nc, Sc - kc0), cw = ncw * Sd, col0 = kc0 * Sd;
                    for (int e = lane; e < Sab * cw; e += NT)
                        sI[e] = 0.0;
                    if (mode == MODE_JK)
                    {
                        for (int e = lane; e < cw; e += NT)
                            dcd[e] = A.D[(c0 + kc0 + e / Sd) * n + d0 + e % Sd];
                        for (int e = lane; e < Sb * ncw; e += NT)
                            dbc[e] = A.D[(b0 + e / ncw) * n + c0 + kc0 + e % ncw];
                        for (int e = lane; e < Sa * ncw; e += NT)
                            dac[e] = A.D[(a0 + e / ncw) * n + c0 + kc0 + e % ncw];
                    }
                    __syncthreads();
                    for (int i = bpoff; i < bpoff + bcnt; i++)
                    {
                        const double4 P = A.bra.pp[i];
                        // the lane that zeroes an element of Y is the lane that accumulates it
                        for (int e = lane; e < Hab * cw; e += NT)
                            sY[e] = 0.0;
                        for (int k = kpoff; k < kpoff + kcnt; k++)
                        {
                            const double4 Q = A.ket.pp[k];
                            const double pq = P.x * Q.x, ps = P.x + Q.x, alpha = pq / ps;
                            const double X = P.y - Q.y, Y = P.z - Q.z, Z = P.w - Q.w;
                            boysLane(alpha * (X * X + Y * Y + Z * Z), A.boys, G.L, sF, lane);
                            const double *ekg = A.ket.ek + static_cast<size_t>(k) * Hcd * Scd + col0;
                            if (sE)
                                for (int e = lane; e < Hcd * cw; e += NT)
                                    sE[e] = ekg[(e / cw) * Scd + e % cw];
                            __syncthreads();
                            const double *R =
                                    hermiteRLds(sR0, sR1, sTuv, sF, G.L, kTwoPi52 / (pq * sqrt(ps)), alpha, X, Y, Z, lane, NT);
                            // E of column s, Hermite row g at ek[g * es + s]
                            const double *ek = sE ? sE : ekg;
                            const int es = sE ? cw : Scd;
                            if (sSum)
                                for (int e = lane; e < Hab * cw; e += NT)
                                {
                                    const int h = e / cw, s = e - h * cw;
                                    const unsigned short *row = sSum + h * Hcd;
                                    double y = sY[e];
                                    for (int g = 0; g < Hcd; g++)
                                        y = fma(R[row[g]], ek[g * es + s], y);
                                    sY[e] = y;
                                }
                            else
                                for (int e = lane; e < Hab * cw; e += NT)
                                {
                                    const int h = e / cw, s = e - h * cw;
                                    const int th = sTuv[h];
                                    double y = sY[e];
                                    for (int g = 0; g < Hcd; g++)
                                        y = fma(R[hsum(th, sTuv[g])], ek[g * es + s], y);
                                    sY[e] = y;
                                }
                            __syncthreads();
                        }
                        const double *eb = A.bra.eb + static_cast<size_t>(i) * Hab * Sab;
                        for (int e = lane; e < Sab * cw; e += NT)
                        {
                            const int sab = e / cw, s = e - sab * cw;
                            double a = sI[e];
                            for (int h = 0; h < Hab; h++)
                                a = fma(eb[h * Sab + sab], sY[h * cw + s], a);
                            sI[e] = a;
                        }
                        __syncthreads();
                    }

                    // epilogue of the chunk: columns col0 .. col0 + cw - 1 of the block, c functions
                    // kc0 .. kc0 + ncw - 1
                    if (mode == MODE_JK)
                    {
                        double *J = A.J, *K = A.K;
                        for (int mn = lane; mn < Sab; mn += NT)
                        {
                            double s = 0.0;
                            for (int c = 0; c < cw; c++)
                                s = fma(sI[mn * cw + c], dcd[c], s);
                            unsafeAtomicAdd(&J[(a0 + mn / Sb) * n + b0 + mn % Sb], 2.0 * f * s);
                        }
                        for (int c = lane; c < cw; c += NT)
                        {
                            double s = 0.0;
                            for (int mn = 0; mn < Sab; mn++)
                                s = fma(sI[mn * cw + c], dab[mn], s);
                            unsafeAtomicAdd(&J[(c0 + kc0 + c / Sd) * n + d0 + c % Sd], 2.0 * f * s);
                        }
                        // K_ac[mu, ka] += f sum_{nu, la} I D_bd[nu, la]
                        for (int e = lane; e < Sa * ncw; e += NT)
                        {
                            const int mu = e / ncw, kk = e % ncw;
                            double s = 0.0;
                            for (int nu = 0; nu < Sb; nu++)
                                for (int la = 0; la < Sd; la++)
                                    s = fma(sI[(mu * Sb + nu) * cw + kk * Sd + la], dbd[nu * Sd + la], s);
                            unsafeAtomicAdd(&K[(c0 + kc0 + kk) * n + a0 + mu], f * s);
                        }
                        // K_bc[nu, ka] += f sum_{mu, la} I D_ad[mu, la]
                        for (int e = lane; e < Sb * ncw; e += NT)
                        {
                            const int nu = e / ncw, kk = e % ncw;
                            double s = 0.0;
                            for (int mu = 0; mu < Sa; mu++)
                                for (int la = 0; la < Sd; la++)
                                    s = fma(sI[(mu * Sb + nu) * cw + kk * Sd + la], dad[mu * Sd + la], s);
                            unsafeAtomicAdd(&K[(c0 + kc0 + kk) * n + b0 + nu], f * s);
                        }
                        // K_ad[mu, la] += f sum_{nu, ka} I D_bc[nu, ka] (this chunk's ka)
                        for (int e = lane; e < Sa * Sd; e += NT)
                        {
                            const int mu = e / Sd, la = e % Sd;
                            double s = 0.0;
                            for (int nu = 0; nu < Sb; nu++)
                                for (int kk = 0; kk < ncw; kk++)
                                    s = fma(sI[(mu * Sb + nu) * cw + kk * Sd + la], dbc[nu * ncw + kk], s);
                            unsafeAtomicAdd(&K[(d0 + la) * n + a0 + mu], f * s);
                        }
                        // K_bd[nu, la] += f sum_{mu, ka} I D_ac[mu, ka] (this chunk's ka)
                        for (int e = lane; e < Sb * Sd; e += NT)
                        {
                            const int nu = e / Sd, la = e % Sd;
                            double s = 0.0;
                            for (int mu = 0; mu < Sa; mu++)
                                for (int kk = 0; kk < ncw; kk++)
                                    s = fma(sI[(mu * Sb + nu) * cw + kk * Sd + la], dac[mu * ncw + kk], s);
                            unsafeAtomicAdd(&K[(d0 + la) * n + b0 + nu], f * s);
                        }
                    }
                    else if (mode == MODE_DENSE)
                    {
                        for (int e = lane; e < Sab * cw; e += NT)
                        {
                            const int mn = e / cw, kl = col0 + e % cw;
                            const double v = sI[e];
                            const long long i0 = a0 + mn / Sb, j0 = b0 + mn % Sb, k0 = c0 + kl / Sd, l0 = d0 + kl % Sd;
                            double *o = A.out;
                            if (A.kind == 4)
                            {
                                o[((i0 * n + j0) * n + k0) * n + l0] = v;
                                o[((i0 * n + j0) * n + l0) * n + k0] = v;
                                o[((j0 * n + i0) * n + k0) * n + l0] = v;
                                o[((j0 * n + i0) * n + l0) * n + k0] = v;
                                o[((k0 * n + l0) * n + i0) * n + j0] = v;
                                o[((k0 * n + l0) * n + j0) * n + i0] = v;
                                o[((l0 * n + k0) * n + i0) * n + j0] = v;
                                o[((l0 * n + k0) * n + j0) * n + i0] = v;
                            }
                            else if (A.kind == 3)
                            {
                                o[(i0 * n + j0) * A.naux + k0] = v;
                                o[(j0 * n + i0) * A.naux + k0] = v;
                            }
                            else
                            {
                                o[i0 * A.naux + k0] = v;
                                o[k0 * A.naux + i0] = v;
                            }
                        }
                    }
                    else if (mode == MODE_SUM)
                    {
                        for (int e = lane; e < Sab * cw; e += NT)
                        {
                            sabs += fabs(sI[e]);
                            ssum += sI[e];
                        }
                    }
                    else
                    {
                        // (ab|ab) diagonal: element (mn, mn) lies in this chunk when its column does
                        for (int mn = lane; mn < Sab; mn += NT)
                        {
                            const int c = mn - col0;
                            if (c < 0 || c >= cw)
                                continue;
                            const double v = sI[mn * cw + c];
                            dmax = fmax(dmax, fabs(v));
                            if (A.dout)
                            {
                                const int mu = mn / Sb, nu = mn % Sb;
                                if (A.kind == 2)
                                    A.dout[a0 + mu] = v;
                                else
                                {
                                    A.dout[(a0 + mu) * n + b0 + nu] = v;
                                    A.dout[(b0 + nu) * n + a0 + mu] = v;
                                }
                            }
                        }
                    }
                    __syncthreads();
                }

                if (mode == MODE_SUM)
                {
                    // images in the dense tensor: (ab|P) and (ba|P) for distinct shells; (P|Q), (Q|P)
                    const bool one = A.kind == 3 ? A.bra.sh_a[ab] == A.bra.sh_b[ab] : (A.same_class && p == q);
                    const double w = one ? 1.0 : 2.0;
                    sabs = waveSum(w * sabs);
                    ssum = waveSum(w * ssum);
                    if ((lane & 63) == 0)   // each wave adds its part
                    {
                        double *o = A.out + 4 * static_cast<long long>(A.ket.mol[cd]) + (A.kind == 3 ? 0 : 2);
                        unsafeAtomicAdd(o, sabs);
                        unsafeAtomicAdd(o + 1, ssum);
                    }
                }
                else if (mode == MODE_DIAG)
                {
                    dmax = waveMax(dmax);
                    double *const red = smem + G.oRed;
                    if ((lane & 63) == 0)
                        red[lane >> 6] = dmax;
                    __syncthreads();
                    if (lane == 0)
                    {
                        for (int w = 1; w < NT / 64; w++)
                            dmax = fmax(dmax, red[w]);
                        A.qout[ab] = sqrt(dmax);
                    }
                    __syncthreads();   // red of the next task
                }
            }
        }

        template <int NT>
        __global__ __launch_bounds__(NT) void genericGrad(const GArgs A, const GenPlan G)
        {
            extern __shared__ double smem[];
            double *const sR0 = smem + G.oR0, *const sR1 = smem + G.oR1, *const sF = smem + G.oF;
            int *const sTuv = reinterpret_cast<int *>(smem + G.oTuv);
            double *const sY = smem + G.oY, *const sX = smem + G.oX, *const sW = smem + G.oW;
            double *const dab = smem + G.oDab, *const dbd = smem + G.oDbd, *const dad = smem + G.oDad;
            double *const dcd = smem + G.oDcd, *const dbc = smem + G.oDbc, *const dac = smem + G.oDac;
            const int lane = threadIdx.x;
            const int Hab = nHerm(G.la + G.lb), Hcd = nHerm(G.lc + G.ld);
            const int Sa = nSph(G.la), Sb = nSph(G.lb), Sc = nSph(G.lc), Sd = nSph(G.ld);
            const int Sab = Sa * Sb, Scd = Sc * Sd;
            const long long n = A.nbf;
            const double *D = A.D;
            const DPairs &LP = A.lane, &UP = A.unif;
            for (int h = lane; h < nHerm(G.L); h += NT)
                sTuv[h] = hermTuv(h);
            __syncthreads();

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
                const int u = uniform(UP.order[q]), l = uniform(LP.order[p]);
                const int upoff = uniform(UP.poff[u]), ucnt = uniform(UP.pcnt[u]);
                const int lpoff = uniform(LP.poff[l]), lcnt = uniform(LP.pcnt[l]);
                const int a0 = LP.ao_a[l], b0 = LP.ao_b[l], c0 = UP.ao_a[u], d0 = UP.ao_b[u];
                double f = 1.0;
                if (LP.sh_a[l] == LP.sh_b[l])
                    f *= 0.5;
                if (UP.sh_a[u] == UP.sh_b[u])
                    f *= 0.5;
                if (A.same_class && p == q)
                    f *= 0.5;
                for (int e = lane; e < Sab; e += NT)
                    dab[e] = D[(a0 + e / Sb) * n + b0 + e % Sb];
                for (int e = lane; e < Sb * Sd; e += NT)
                    dbd[e] = D[(b0 + e / Sd) * n + d0 + e % Sd];
                for (int e = lane; e < Sa * Sd; e += NT)
                    dad[e] = D[(a0 + e / Sd) * n + d0 + e % Sd];
                double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0}, gC[3] = {0, 0, 0};

                for (int kc0 = 0; kc0 < Sc; kc0 += G.nc)
                {
                    const int ncw = min(G.nc, Sc - kc0), cw = ncw * Sd, col0 = kc0 * Sd;
                    for (int e = lane; e < cw; e += NT)
                        dcd[e] = D[(c0 + kc0 + e / Sd) * n + d0 + e % Sd];
                    for (int e = lane; e < Sb * ncw; e += NT)
                        dbc[e] = D[(b0 + e / ncw) * n + c0 + kc0 + e % ncw];
                    for (int e = lane; e < Sa * ncw; e += NT)
                        dac[e] = D[(a0 + e / ncw) * n + c0 + kc0 + e % ncw];
                    __syncthreads();
                    // W[ab][c] = f (4 D_ab D_cd - D_ac D_bd - D_ad D_bc) over this chunk's columns
                    for (int e = lane; e < Sab * cw; e += NT)
                    {
                        const int mn = e / cw, c = e - mn * cw;
                        const int mu = mn / Sb, nu = mn % Sb, kk = c / Sd, la = c % Sd;
                        sW[e] = f * (4.0 * dab[mn] * dcd[c] - dac[mu * ncw + kk] * dbd[nu * Sd + la] -
                                     dad[mu * Sd + la] * dbc[nu * ncw + kk]);
                    }
                    __syncthreads();
                    for (int i = lpoff; i < lpoff + lcnt; i++)
                    {
                        const double4 P = LP.pp[i];
                        const double rai = LP.ra[i];
                        const double *eb = LP.eb + static_cast<size_t>(i) * Hab * Sab;
                        for (int e = lane; e < Hab * cw; e += NT)
                        {
                            const int h = e / cw, c = e - h * cw;
                            double y = 0.0;
                            for (int mn = 0; mn < Sab; mn++)
                                y = fma(eb[h * Sab + mn], sW[mn * cw + c], y);
                            sY[e] = y;
                            sX[e] = 0.0;
                        }
                        __syncthreads();
                        for (int k = upoff; k < upoff + ucnt; k++)
                        {
                            const double4 Q = UP.pp[k];
                            const double rck = UP.ra[k];
                            const double pq = P.x * Q.x, ps = P.x + Q.x, alpha = pq / ps;
                            const double X = P.y - Q.y, Y = P.z - Q.z, Z = P.w - Q.w;
                            boysLane(alpha * (X * X + Y * Y + Z * Z), A.boys, G.L, sF, lane);
                            __syncthreads();
                            const double *R =
                                    hermiteRLds(sR0, sR1, sTuv, sF, G.L, kTwoPi52 / (pq * sqrt(ps)), alpha, X, Y, Z, lane, NT);
                            const size_t ecount = static_cast<size_t>(Hcd) * Scd;
                            const double *ek = UP.ek + static_cast<size_t>(k) * ecount + col0;
                            const double *e1k = UP.e1k + static_cast<size_t>(k) * 3 * ecount + col0;
                            // T and V over the (h, g) pairs
                            for (int e = lane; e < Hab * Hcd; e += NT)
                            {
                                const int h = e / Hcd, g = e - h * Hcd;
                                double z = 0.0, zx = 0.0, zy = 0.0, zz = 0.0;
                                for (int c = 0; c < cw; c++)
                                {
                                    const double y = sY[h * cw + c];
                                    const size_t o = static_cast<size_t>(g) * Scd + c;
                                    z = fma(y, ek[o], z);
                                    zx = fma(y, e1k[o], zx);
                                    zy = fma(y, e1k[ecount + o], zy);
                                    zz = fma(y, e1k[2 * ecount + o], zz);
                                }
                                const int w = sTuv[h], v = sTuv[g];
                                const int tt = (w & 255) + (v & 255), uu = ((w >> 8) & 255) + ((v >> 8) & 255),
                                          vv = (w >> 16) + (v >> 16);
                                const double T[3] = {z * R[hermIdx(tt + 1, uu, vv)], z * R[hermIdx(tt, uu + 1, vv)],
                                                     z * R[hermIdx(tt, uu, vv + 1)]};
                                const double r0 = R[hermIdx(tt, uu, vv)];
                                const double V[3] = {zx * r0, zy * r0, zz * r0};
#pragma unroll
                                for (int x = 0; x < 3; x++)
                                {
                                    gA[x] = fma(rai, T[x], gA[x]);
                                    gB[x] = fma(1.0 - rai, T[x], gB[x]);
                                    gC[x] += V[x] - rck * T[x];
                                }
                            }
                            for (int e = lane; e < Hab * cw; e += NT)
                            {
                                const int h = e / cw, c = e - h * cw;
                                const int th = sTuv[h];
                                double x = sX[e];
                                for (int g = 0; g < Hcd; g++)
                                    x = fma(R[hsum(th, sTuv[g])], ek[g * Scd + c], x);
                                sX[e] = x;
                            }
                            __syncthreads();
                        }
                        // U_x = sum W[ab][c] sum_h Eb'_i,x[h][ab] X[h][c]
                        const double *e1b = LP.e1b + static_cast<size_t>(i) * 3 * Hab * Sab;
                        for (int e = lane; e < Sab * cw; e += NT)
                        {
                            const int mn = e / cw, c = e - mn * cw;
                            const double w = sW[e];
#pragma unroll
                            for (int x = 0; x < 3; x++)
                            {
                                const double *d = e1b + static_cast<size_t>(x) * Hab * Sab;
                                double s = 0.0;
                                for (int h = 0; h < Hab; h++)
                                    s = fma(d[h * Sab + mn], sX[h * cw + c], s);
                                gA[x] = fma(w, s, gA[x]);
                                gB[x] = fma(-w, s, gB[x]);
                            }
                        }
                        __syncthreads();
                    }
                }
#pragma unroll
                for (int x = 0; x < 3; x++)
                {
                    gA[x] = waveSum(gA[x]);
                    gB[x] = waveSum(gB[x]);
                    gC[x] = waveSum(gC[x]);
                }
                if ((lane & 63) == 0)   // each wave adds its part
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
                __syncthreads();   // density blocks of the next task
            }
        }

        template <int MODE, int LA, int LB, int LC, int LD>
        void launchGenericQuartet(const KArgs &A, hipStream_t s)
        {
            if (A.ntasks <= 0)
                return;
            static const GenPlan G = makePlan(LA, LB, LC, LD, MODE == MODE_JK, false);
            const int grid = static_cast<int>(A.ntasks < 16384 ? A.ntasks : 16384);
            const size_t lds = static_cast<size_t>(G.doubles) * 8;
            if (G.nt == 64)
                hipLaunchKernelGGL(genericQuartet<64>, dim3(grid), dim3(64), lds, s, A, G, MODE);
            else
                hipLaunchKernelGGL(genericQuartet<kGenThreads>, dim3(grid), dim3(kGenThreads), lds, s, A, G, MODE);
        }

        template <int LA, int LB, int LC, int LD>
        void launchGenericGrad(const GArgs &A, hipStream_t s)
        {
            if (A.ntasks <= 0)
                return;
            static const GenPlan G = makePlan(LA, LB, LC, LD, false, true);
            const int grid = static_cast<int>(A.ntasks < 16384 ? A.ntasks : 16384);
            const size_t lds = static_cast<size_t>(G.doubles) * 8;
            if (G.nt == 64)
                hipLaunchKernelGGL(genericGrad<64>, dim3(grid), dim3(64), lds, s, A, G);
            else
                hipLaunchKernelGGL(genericGrad<kGenThreads>, dim3(grid), dim3(kGenThreads), lds, s, A, G);
        }

        // The compiled units register, among the classes with a 4 on the bra's first index, only
        // the single auxiliary g shell's eri2 forms: dense and per-molecule sums (4 0 | l 0) for
        // l <= 4 and the diagonal (4 0 | 4 0).  Those kernels are the same code for a main (g s)
        // pair, so the run-time-L kernel takes every other slot with a main g shell (a disjoint set:
        // the registration order of the units does not matter).
        constexpr bool compiledSlot(int mode, int la, int lb, int lc, int ld)
        {
            if (la != 4 || lb != 0 || ld != 0)
                return false;
            if (mode == MODE_DENSE || mode == MODE_SUM)
                return lc <= 4;
            return mode == MODE_DIAG && lc == 4;
        }

        // slots looked up with a main g pair (4 lb) as the bra: J/K and dense eri4 against every
        // main ket class up to it, dense eri3 and sums against every auxiliary shell, the diagonal
        constexpr bool wantedSlot(int mode, int la, int lb, int lc, int ld)
        {
            if (la != 4 || lb > 4 || compiledSlot(mode, la, lb, lc, ld))
                return false;
            const bool mainKet = lc <= 4 && ld <= lc && pairClassIdx(lc, ld) <= pairClassIdx(la, lb);
            const bool auxKet = ld == 0 && lc <= kMaxLAux;
            switch (mode)
            {
            case MODE_JK:
                return mainKet;
            case MODE_DENSE:
                return mainKet || auxKet;
            case MODE_SUM:
                return auxKet;
            case MODE_DIAG:
                return lc == la && ld == lb;
            default:
                return false;
            }
        }

        template <int MODE, int N>
        void regQuartetSlot()
        {
            constexpr int lb = N / 36, lc = (N / 6) % 6, ld = N % 6;
            if constexpr (wantedSlot(MODE, 4, lb, lc, ld))
            {
                KernelSlot &s = kernelSlot(MODE, 4, lb, lc, ld, VAR_MFMA);
                s.fn = &launchGenericQuartet<MODE, 4, lb, lc, ld>;
                s.width = 1;
                s.swap = false;
                s.kind = KIND_WAVE;
            }
        }

        // gradient: lane side (la lb), uniform side (lc ld), either of them a main g pair
        template <int N>
        void regGradSlot()
        {
            constexpr int la = N / 125, lb = (N / 25) % 5, lc = (N / 5) % 5, ld = N % 5;
            if constexpr (lb <= la && ld <= lc && (la == 4 || lc == 4))
            {
                GradSlot &s = gradSlot(la, lb, lc, ld);
                s.fn = &launchGenericGrad<la, lb, lc, ld>;
                s.fn32 = s.fn;   // no FP32 form: the small-bound tier of these classes stays FP64
                s.width = 1;
            }
        }

        template <int... N>
        void regQuartetSlots(std::integer_sequence<int, N...>)
        {
            (regQuartetSlot<MODE_JK, N>(), ...);
            (regQuartetSlot<MODE_DENSE, N>(), ...);
            (regQuartetSlot<MODE_DIAG, N>(), ...);
            (regQuartetSlot<MODE_SUM, N>(), ...);
        }

        template <int... N>
        void regGradSlots(std::integer_sequence<int, N...>)
        {
            (regGradSlot<N>(), ...);
        }

        struct Registrar
        {
            Registrar()
            {
                regQuartetSlots(std::make_integer_sequence<int, 5 * 6 * 6>{});
                regGradSlots(std::make_integer_sequence<int, 5 * 5 * 5 * 5>{});
            }
        } registrar;
    }
}
