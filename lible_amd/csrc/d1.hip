// Dense first-derivative ERI tensors: the quantities of the reference's ERI4D1 / ERI3D1 /
// ERI2D1 kernels (eri_kernel_funs.hpp:350-732, eri_kernel_funs.cpp:582-748), written as dense
// tensors whose component 3k + x is the derivative with respect to coordinate x of the centre of
// the k-th index's shell.
//
// For a quartet (ab|cd) with bra primitive pairs i (exponents a, b, p = a + b) and ket primitive
// pairs k (c, d, q = c + d), and the Hermite form (ab|cd) = sum_ik Eb_i R_{h+g} Ek_k:
//   d/dA_x = sum_ik (a/p) Eb_i R_{h+g+e_x} Ek_k + Eb'_i,x R_{h+g} Ek_k
//   d/dB_x = sum_ik (b/p) Eb_i R_{h+g+e_x} Ek_k - Eb'_i,x R_{h+g} Ek_k
//   d/dC_x = sum_ik -(c/q) Eb_i R_{h+g+e_x} Ek_k + Eb_i R_{h+g} Ek'_k,x
//   d/dD_x = sum_ik -(d/q) Eb_i R_{h+g+e_x} Ek_k - Eb_i R_{h+g} Ek'_k,x
// (E' = dE/d(A - B)_x of the pair, the ket's (-1)^{t'+u'+v'} in Ek and Ek'; d/dP_x of the Hermite
// Gaussian raises its x index).  Single-shell sides (the auxiliary shells of eri3 / eri2) are an
// s function of exponent 0 at the shell's centre: E' = 0, a/p = 1, and the dummy centre's
// derivative vanishes, so the 3- and 2-centre tensors are the 4-centre formulas restricted to
// the real centres.
//
// This is the compatibility path of the derivative kernels (dense tensors are only feasible for
// small systems; the nuclear gradient of J/K contracts them on the fly in grad.hpp): one kernel
// for every class pair with the angular momenta at run time, one wave per quartet, Hermite R and
// the ket transform staged in LDS, the output block of one component accumulated in registers
// (lanes own output elements) a chunk of ket columns at a time (up to 2432 elements, so (gg|gg)
// takes three chunks), one pass over the primitive quartets per component and chunk.
#include "lbx_internal.hpp"
#include "runtime_l.hpp"

#include <algorithm>

namespace lbx
{
    namespace
    {
        // output elements per lane (a block column chunk holds at most 64 kD1Acc elements) and
        // LDS per workgroup in doubles: R ping-pong, Y of a column chunk, Boys values, (t, u, v)
        constexpr int kD1Acc = ceilDiv(49 * 49, 64);
        constexpr int kD1Lds = 64 * 1024 / 8;

        // image permutations: image position j holds original index perm[j]
        // second derivatives of (a|b): the six distinct (i, j), i <= j, of the A block
        __constant__ int kD2Pair[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};

        __constant__ int kImg4[8][4] = {{0, 1, 2, 3}, {1, 0, 2, 3}, {0, 1, 3, 2}, {1, 0, 3, 2},
                                        {2, 3, 0, 1}, {3, 2, 0, 1}, {2, 3, 1, 0}, {3, 2, 1, 0}};

        // cw: ket columns (spherical cd indices) per chunk of the block
        __global__ __launch_bounds__(64) void d1Kernel(const D1Args A, const int cw)
        {
            extern __shared__ double smem[];
            const int lane = threadIdx.x;
            const int la = A.la, lb = A.lb, lc = A.lc, ld = A.ld;
            // R to order L + 1 (first derivatives, spin-orbit) or L + 2 (second derivatives)
            const int L1 = la + lb + lc + ld + (A.op == D1_OP_D2 ? 2 : 1);
            const int Hab = nHerm(la + lb), Hcd = nHerm(lc + ld), H1 = nHerm(L1);
            const int Sa = nSph(la), Sb = nSph(lb), Sc = nSph(lc), Sd = nSph(ld);
            const int Sab = Sa * Sb, Scd = Sc * Sd, nOut = Sab * Scd;
            double *const sR0 = smem, *const sR1 = smem + H1, *const sY = smem + 2 * H1;
            double *const sF = sY + Hab * cw;
            int *const sTuv = reinterpret_cast<int *>(sF + L1 + 1);
            for (int h = lane; h < H1; h += 64)
                sTuv[h] = hermTuv(h);
            __syncthreads();

            // components: kind 4 centres A B C D, kind 3 A B C, kind 2 A C
            const int ncent = A.kind == 4 ? 4 : (A.kind == 3 ? 3 : 2);
            // components computed per entry, and components per entry in the output
            const int ncomp = A.op == D1_OP_PLAIN ? 1 : A.op == D1_OP_SOC ? 3 : A.op == D1_OP_D2 ? 6 : 3 * ncent;
            const int nstore = A.op == D1_OP_D2 ? 36 : ncomp;
            const bool plain = A.op == D1_OP_PLAIN, extra = A.op == D1_OP_SOC || A.op == D1_OP_D2;
            const long long n = A.nbf;
            const int ts = A.blocks ? 4 : 2;
            for (long long task = blockIdx.x; task < A.ntasks; task += gridDim.x)
            {
                const int p = A.tasks[ts * task], q = A.tasks[ts * task + 1];
                const int poff = A.bra.poff[p], pcnt = A.bra.pcnt[p];
                const int qoff = A.ket.poff[q], qcnt = A.ket.pcnt[q];
                const int ao[4] = {A.bra.ao_a[p], A.bra.ao_b[p], A.ket.ao_a[q], A.ket.ao_b[q]};
                // per component, the block's ket columns in chunks of cw
                for (int comp = 0; comp < ncomp; comp++)
                    for (int c0 = 0; c0 < Scd; c0 += cw)
                {
                    // columns c0 .. c0 + cwc - 1 of the block: nO elements, lane + 64 r in acc[r]
                    const int cwc = min(cw, Scd - c0), nO = Sab * cwc;
                    const int ci = comp / 3, x = comp % 3;
                    // internal centre: 0 A, 1 B, 2 C, 3 D; -1 for the integrals themselves
                    const int cent = (plain || extra) ? -1 : (A.kind == 2 ? (ci == 0 ? 0 : 2) : ci);
                    const int ex = cent >= 0 && x == 0, ey = cent >= 0 && x == 1, ez = cent >= 0 && x == 2;
                    double acc[kD1Acc];
#pragma unroll
                    for (int r = 0; r < kD1Acc; r++)
                        acc[r] = 0.0;

                    for (int i = poff; i < poff + pcnt; i++)
                    {
                        const double4 P = A.bra.pp[i];
                        const double rai = (A.bra_single || plain || extra) ? 1.0 : A.bra.ra[i];
                        for (int k = qoff; k < qoff + qcnt; k++)
                        {
                            const double4 Q = A.ket.pp[k];
                            const double rck = (A.ket_single || plain || extra) ? 1.0 : A.ket.ra[k];
                            const double pq = P.x * Q.x, ps = P.x + Q.x, alpha = pq / ps;
                            const double X = P.y - Q.y, Y = P.z - Q.z, Z = P.w - Q.w;
                            if (lane == 0)
                                boysRuntime(alpha * (X * X + Y * Y + Z * Z), A.boys, L1, sF);
                            __syncthreads();
                            const double *R = hermiteRLds(sR0, sR1, sTuv, sF, L1, kTwoPi52 / (pq * sqrt(ps)),
                                                          alpha, X, Y, Z, lane, 64);

                            if (extra)
                            {
                                // spin-orbit: component x = sum Eb'_{d1} R_{h+g+e_d2} Ek - Eb'_{d2}
                                // R_{h+g+e_d1} Ek with (d1, d2) = (y, z), (z, x), (x, y) for x, y, z;
                                // second derivatives of (a|b): component (i, j) = Eb R_{h+g+e_i+e_j} Ek
                                const double *ek = A.ket.ek + static_cast<size_t>(k) * Hcd * Scd;
                                const int nterm = A.op == D1_OP_SOC ? 2 : 1;
                                for (int term = 0; term < nterm; term++)
                                {
                                    int sh0 = 0, sh1 = 0, sh2 = 0;
                                    const double *bm;
                                    double sg = 1.0;
                                    if (A.op == D1_OP_SOC)
                                    {
                                        const int d1 = (x + 1) % 3, d2 = (x + 2) % 3;
                                        const int dsh = term == 0 ? d2 : d1, dbra = term == 0 ? d1 : d2;
                                        sh0 = dsh == 0;
                                        sh1 = dsh == 1;
                                        sh2 = dsh == 2;
                                        bm = A.bra.e1b + (static_cast<size_t>(i) * 3 + dbra) * Hab * Sab;
                                        sg = term == 0 ? 1.0 : -1.0;
                                    }
                                    else
                                    {
                                        const int ii = kD2Pair[comp][0], jj = kD2Pair[comp][1];
                                        sh0 = (ii == 0) + (jj == 0);
                                        sh1 = (ii == 1) + (jj == 1);
                                        sh2 = (ii == 2) + (jj == 2);
                                        bm = A.bra.eb + static_cast<size_t>(i) * Hab * Sab;
                                    }
                                    for (int e = lane; e < Hab * cwc; e += 64)
                                    {
                                        const int h = e / cwc, s = c0 + e - h * cwc;
                                        const int th = sTuv[h];
                                        const int t = th & 255, u = (th >> 8) & 255, v = th >> 16;
                                        double y = 0.0;
                                        for (int g = 0; g < Hcd; g++)
                                        {
                                            const int tg = sTuv[g];
                                            y = fma(R[hermIdx(t + (tg & 255) + sh0, u + ((tg >> 8) & 255) + sh1,
                                                              v + (tg >> 16) + sh2)],
                                                    ek[g * Scd + s], y);
                                        }
                                        sY[e] = y;
                                    }
                                    __syncthreads();
#pragma unroll
                                    for (int r = 0; r < kD1Acc; r++)
                                    {
                                        const int o = lane + 64 * r;
                                        if (o < nO)
                                        {
                                            const int sab = o / cwc, s = o - sab * cwc;
                                            double a = 0.0;
                                            for (int h = 0; h < Hab; h++)
                                                a = fma(bm[h * Sab + sab], sY[h * cwc + s], a);
                                            acc[r] = fma(sg, a, acc[r]);
                                        }
                                    }
                                    __syncthreads();
                                }
                                continue;
                            }

                            // ket transform: Y = aT sum_g R_{h+g+e_x} Ek_g + a0 sum_g R_{h+g} Ek'_g
                            double aT, a0 = 0.0;
                            if (cent < 0)
                                aT = 1.0;
                            else if (cent == 0)
                                aT = rai;
                            else if (cent == 1)
                                aT = 1.0 - rai;
                            else if (cent == 2)
                            {
                                aT = -rck;
                                a0 = A.ket_single ? 0.0 : 1.0;
                            }
                            else
                            {
                                aT = -(1.0 - rck);
                                a0 = A.ket_single ? 0.0 : -1.0;
                            }
                            const double *ek = A.ket.ek + static_cast<size_t>(k) * Hcd * Scd;
                            const double *e1k =
                                    a0 != 0.0 ? A.ket.e1k + (static_cast<size_t>(k) * 3 + x) * Hcd * Scd : nullptr;
                            for (int e = lane; e < Hab * cwc; e += 64)
                            {
                                const int h = e / cwc, s = c0 + e - h * cwc;
                                const int th = sTuv[h];
                                const int t = th & 255, u = (th >> 8) & 255, v = th >> 16;
                                double y = 0.0;
                                for (int g = 0; g < Hcd; g++)
                                {
                                    const int tg = sTuv[g];
                                    const int t2 = t + (tg & 255), u2 = u + ((tg >> 8) & 255), v2 = v + (tg >> 16);
                                    double w = aT * R[hermIdx(t2 + ex, u2 + ey, v2 + ez)] * ek[g * Scd + s];
                                    if (e1k)
                                        w = fma(a0 * R[hermIdx(t2, u2, v2)], e1k[g * Scd + s], w);
                                    y += w;
                                }
                                sY[e] = y;
                            }
                            __syncthreads();
                            // bra transform into this lane's output elements
                            const double *eb = A.bra.eb + static_cast<size_t>(i) * Hab * Sab;
#pragma unroll
                            for (int r = 0; r < kD1Acc; r++)
                            {
                                const int o = lane + 64 * r;
                                if (o < nO)
                                {
                                    const int sab = o / cwc, s = o - sab * cwc;
                                    double a = acc[r];
                                    for (int h = 0; h < Hab; h++)
                                        a = fma(eb[h * Sab + sab], sY[h * cwc + s], a);
                                    acc[r] = a;
                                }
                            }
                            __syncthreads();
                            // the bra pair's own derivative: +- Eb'_x R_{h+g} Ek
                            if ((cent == 0 || cent == 1) && !A.bra_single)
                            {
                                for (int e = lane; e < Hab * cwc; e += 64)
                                {
                                    const int h = e / cwc, s = c0 + e - h * cwc;
                                    const int th = sTuv[h];
                                    const int t = th & 255, u = (th >> 8) & 255, v = th >> 16;
                                    double y = 0.0;
                                    for (int g = 0; g < Hcd; g++)
                                    {
                                        const int tg = sTuv[g];
                                        y = fma(R[hermIdx(t + (tg & 255), u + ((tg >> 8) & 255), v + (tg >> 16))],
                                                ek[g * Scd + s], y);
                                    }
                                    sY[e] = y;
                                }
                                __syncthreads();
                                const double sg = cent == 0 ? 1.0 : -1.0;
                                const double *e1b = A.bra.e1b + (static_cast<size_t>(i) * 3 + x) * Hab * Sab;
#pragma unroll
                                for (int r = 0; r < kD1Acc; r++)
                                {
                                    const int o = lane + 64 * r;
                                    if (o < nO)
                                    {
                                        const int sab = o / cwc, s = o - sab * cwc;
                                        double a = 0.0;
                                        for (int h = 0; h < Hab; h++)
                                            a = fma(e1b[h * Sab + sab], sY[h * cwc + s], a);
                                        acc[r] = fma(sg, a, acc[r]);
                                    }
                                }
                                __syncthreads();
                            }
                        }
                    }

                    if (A.blocks)
                    {
                        // the block in the caller's shell order: a <-> b (flag 1), c <-> d (2)
                        const int flags = A.tasks[4 * task + 2];
                        const long long slot = A.tasks[4 * task + 3];
                        const bool swab = flags & 1, swcd = flags & 2;
                        const int rb = swab ? Sa : Sb, rc = swcd ? Sd : Sc, rd = swcd ? Sc : Sd;
                        // component of the caller's centre: its shell position k, 3 k + x
                        int rcomp = comp;
                        double sg = 1.0;
                        if (cent >= 0)
                        {
                            int pos = cent;
                            if (A.kind == 2)
                                pos = cent == 0 ? 0 : 1;
                            else if (cent <= 1)
                                pos = swab ? 1 - cent : cent;
                            else if (A.kind == 4)
                                pos = swcd ? 5 - cent : cent;
                            rcomp = 3 * pos + x;
                        }
                        else if (A.op == D1_OP_SOC && swab)
                            sg = -1.0;   // E' is odd in A - B: the (b, a) pair's is minus the (a, b) one's
                        // second derivatives: AA (i, j) -> blocks [i][j], [j][i], [3+i][3+j], [3+j][3+i]
                        // and minus it at the AB / BA positions [i][3+j], [j][3+i], [3+i][j], [3+j][i]
                        const int nwrite = A.op == D1_OP_D2 ? 8 : 1;
                        for (int wi = 0; wi < nwrite; wi++)
                        {
                            double wsg = sg;
                            if (A.op == D1_OP_D2)
                            {
                                const int ii = kD2Pair[comp][0], jj = kD2Pair[comp][1];
                                const int a = (wi & 1) ? jj : ii, b = (wi & 1) ? ii : jj;
                                const int sa = (wi >> 1) & 1, sb = ((wi >> 1) ^ (wi >> 2)) & 1;
                                rcomp = 6 * (a + 3 * sa) + b + 3 * sb;
                                wsg = sa == sb ? 1.0 : -1.0;
                            }
                            double *dst = A.out + (slot * nstore + rcomp) * nOut;
#pragma unroll
                            for (int r = 0; r < kD1Acc; r++)
                            {
                                const int o = lane + 64 * r;
                                if (o >= nO)
                                    continue;
                                const int sab = o / cwc, scd = c0 + o - sab * cwc;
                                const int i0 = sab / Sb, i1 = sab % Sb, i2 = scd / Sd, i3 = scd % Sd;
                                const int m0 = swab ? i1 : i0, m1 = swab ? i0 : i1;
                                const int m2 = swcd ? i3 : i2, m3 = swcd ? i2 : i3;
                                dst[((m0 * rb + m1) * rc + m2) * rd + m3] = wsg * acc[r];
                            }
                        }
                        continue;
                    }
                    // scatter this component's block to every image
#pragma unroll
                    for (int r = 0; r < kD1Acc; r++)
                    {
                        const int o = lane + 64 * r;
                        if (o >= nO)
                            continue;
                        const int sab = o / cwc, scd = c0 + o - sab * cwc;
                        const int idx[4] = {ao[0] + sab / Sb, ao[1] + sab % Sb, ao[2] + scd / Sd,
                                            ao[3] + scd % Sd};
                        const double val = acc[r];
                        if (A.kind == 4)
                        {
                            const long long n4 = n * n * n * n;
                            for (int im = 0; im < 8; im++)
                            {
                                int pos = 0;
                                for (int j = 0; j < 4; j++)
                                    if (kImg4[im][j] == cent)
                                        pos = j;
                                const long long e = ((static_cast<long long>(idx[kImg4[im][0]]) * n +
                                                      idx[kImg4[im][1]]) * n + idx[kImg4[im][2]]) * n +
                                                    idx[kImg4[im][3]];
                                A.out[(3 * pos + x) * n4 + e] = val;
                            }
                        }
                        else if (A.kind == 3)
                        {
                            const long long nx = A.naux, n3 = n * n * nx;
                            const int pos0 = cent, pos1 = cent == 2 ? 2 : 1 - cent;
                            A.out[(3 * pos0 + x) * n3 + (static_cast<long long>(idx[0]) * n + idx[1]) * nx + idx[2]] = val;
                            A.out[(3 * pos1 + x) * n3 + (static_cast<long long>(idx[1]) * n + idx[0]) * nx + idx[2]] = val;
                        }
                        else
                        {
                            const long long nx = A.naux, n2 = nx * nx;
                            const int pos = cent == 0 ? 0 : 1;
                            A.out[(3 * pos + x) * n2 + static_cast<long long>(idx[0]) * nx + idx[2]] = val;
                            A.out[(3 * (1 - pos) + x) * n2 + static_cast<long long>(idx[2]) * nx + idx[0]] = val;
                        }
                    }
                }
            }
        }
    }

    void launchD1(const D1Args &a, hipStream_t s)
    {
        if (a.ntasks <= 0)
            return;
        // the block's columns in chunks: at most 64 kD1Acc elements per chunk, Y of a chunk and
        // the fixed arrays within kD1Lds
        const int L1 = a.la + a.lb + a.lc + a.ld + (a.op == D1_OP_D2 ? 2 : 1), H1 = nHerm(L1);
        const int Hab = nHerm(a.la + a.lb), Sab = nSph(a.la) * nSph(a.lb), Scd = nSph(a.lc) * nSph(a.ld);
        const int fixed = 2 * H1 + L1 + 1 + (H1 + 1) / 2;
        const int cw = std::min({Scd, 64 * kD1Acc / Sab, (kD1Lds - fixed) / Hab});
        if (L1 > kMaxLBoys || cw < 1)
            throw Error(LBX_ERR_INVALID, "derivative tensors: class beyond the kernel's limits");
        const int grid = static_cast<int>(a.ntasks < 16384 ? a.ntasks : 16384);
        hipLaunchKernelGGL(d1Kernel, dim3(grid), dim3(64), static_cast<size_t>(fixed + Hab * cw) * 8, s, a, cw);
        LBX_HIP(hipGetLastError());
    }
}
