// Device helpers of the kernels whose angular momenta are run-time values (the dense derivative
// tensors and the classes above f in the main basis): the Boys function to a run-time order and
// the Hermite R recursion staged in LDS by the threads of a workgroup.
#pragma once

#include "device_common.hpp"

namespace lbx
{
    // F_0..F_L(x), L at run time: the evaluation of boysF (boys_function.hpp:68-109) with the
    // loops bounded by L.
    __device__ __forceinline__ void boysRuntime(double x, const double *__restrict__ tab, int L, double *F)
    {
        if (x == 0.0)
        {
            F[0] = 1.0;
            for (int k = 1; k <= L; k++)
                F[k] = 1.0 / (2 * k + 1);
        }
        else if (x > kBoysLargeX)
        {
            const double rx = 1.0 / x;
            F[0] = 0.5 * sqrt(M_PI * rx);
            for (int k = 1; k <= L; k++)
                F[k] = F[k - 1] * (k - 0.5) * rx;
        }
        else
        {
            const int iv = static_cast<int>(x * 100.0);
            const double dx = x - iv * kBoysStep;
            const double *row = tab + iv * kBoysCols + L;
            const double m = -dx;
            double s = row[6] * (1.0 / 720.0);
            s = fma(s, m, row[5] * (1.0 / 120.0));
            s = fma(s, m, row[4] * (1.0 / 24.0));
            s = fma(s, m, row[3] * (1.0 / 6.0));
            s = fma(s, m, row[2] * 0.5);
            s = fma(s, m, row[1]);
            s = fma(s, m, row[0]);
            F[L] = s;
            const double ex = exp(-x);
            const double x2 = 2.0 * x;
            for (int k = L - 1; k >= 0; k--)
                F[k] = fma(x2, F[k + 1], ex) * (1.0 / (2 * k + 1));
        }
    }

    // F_m(x) of order m = lane into sF[m], lanes 0..L (L < 64): each lane takes its own column
    // of the table (the 7-term Taylor expansion about the tabulated point below x, as boysF does
    // for F_L; the table holds columns up to kMaxLBoys + 6) instead of one lane running the
    // downward recursion; x > 30 the asymptotic form, x = 0 the exact values.
    __device__ __forceinline__ void boysLane(double x, const double *__restrict__ tab, int L, double *sF, int lane)
    {
        if (lane > L)
            return;
        const int m = lane;
        double f;
        if (x == 0.0)
            f = 1.0 / (2 * m + 1);
        else if (x > kBoysLargeX)
        {
            const double rx = 1.0 / x;
            f = 0.5 * sqrt(M_PI * rx);
            for (int k = 1; k <= m; k++)
                f *= (k - 0.5) * rx;
        }
        else
        {
            const int iv = static_cast<int>(x * 100.0);
            const double dx = x - iv * kBoysStep;
            const double *row = tab + iv * kBoysCols + m;
            const double mx = -dx;
            double s = row[6] * (1.0 / 720.0);
            s = fma(s, mx, row[5] * (1.0 / 120.0));
            s = fma(s, mx, row[4] * (1.0 / 24.0));
            s = fma(s, mx, row[3] * (1.0 / 6.0));
            s = fma(s, mx, row[2] * 0.5);
            s = fma(s, mx, row[1]);
            f = fma(s, mx, row[0]);
        }
        sF[m] = f;
    }

    // Hermite integrals R_tuv(alpha, PQ) for t + u + v <= L of one primitive quartet, by the
    // nthr threads of a workgroup: level m computed from level m + 1 (HJO 9.9.18-20),
    // m = L .. 0, ping-ponging between r0 and r1 (nHerm(L) doubles each).  sF holds F_0..F_L
    // (written by the caller before the call, visible after a barrier); tuv[h] is hermTuv(h).
    // fac is 2 pi^{5/2} / (p q sqrt(p + q)).  Returns the buffer holding level 0; ends with a
    // barrier.
    __device__ __forceinline__ const double *hermiteRLds(double *r0, double *r1, const int *__restrict__ tuv,
                                                         const double *sF, int L, double fac, double alpha, double X,
                                                         double Y, double Z, int lane, int nthr)
    {
        double *cur = r0, *nxt = r1;
        const double m2a = -2.0 * alpha;
        // (-2 alpha)^m, the prefactor of level m
        double pw = 1.0;
        for (int k = 0; k < L; k++)
            pw *= m2a;
        if (lane == 0)
            cur[0] = sF[L] * fac * pw;
        __syncthreads();
        for (int m = L - 1; m >= 0; m--)
        {
            pw = 1.0;
            for (int k = 0; k < m; k++)
                pw *= m2a;
            const int top = nHerm(L - m);
            for (int e = lane; e < top; e += nthr)
            {
                const int w = tuv[e];
                const int t = w & 255, u = (w >> 8) & 255, v = w >> 16;
                double r;
                if (t > 0)
                    r = (t > 1 ? (t - 1) * cur[hermIdx(t - 2, u, v)] : 0.0) + X * cur[hermIdx(t - 1, u, v)];
                else if (u > 0)
                    r = (u > 1 ? (u - 1) * cur[hermIdx(t, u - 2, v)] : 0.0) + Y * cur[hermIdx(t, u - 1, v)];
                else if (v > 0)
                    r = (v > 1 ? (v - 1) * cur[hermIdx(t, u, v - 2)] : 0.0) + Z * cur[hermIdx(t, u, v - 1)];
                else
                    r = sF[m] * fac * pw;
                nxt[e] = r;
            }
            __syncthreads();
            double *sw = cur;
            cur = nxt;
            nxt = sw;
        }
        return cur;
    }

    // largest value over the 64 lanes (every lane gets it)
    __device__ __forceinline__ double waveMax(double x)
    {
        for (int o = 32; o > 0; o >>= 1)
            x = fmax(x, __shfl_xor(x, o, 64));
        return x;
    }
}
