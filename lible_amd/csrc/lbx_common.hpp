// Shared compile-time helpers of the lbx library (host and device).
#pragma once

#include <cstdint>

#define LBX_HD __host__ __device__ __forceinline__

namespace lbx
{
    constexpr int kMaxLMain = 4;    // main-basis shells: s to g (compiled kernels to f, g at run-time L)
    constexpr int kMaxLAux = 5;     // auxiliary shells: up to h
    constexpr int kMaxLBoys = 17;   // highest Boys order any kernel asks for: (gg|gg) gradient
    constexpr int kBoysCols = kMaxLBoys + 7;   // table columns F_0..F_{n+6}
    constexpr int kBoysRows = 3001;            // x = 0, 0.01, ..., 30
    constexpr double kBoysStep = 0.01;
    constexpr double kBoysLargeX = 30.0;

    constexpr LBX_HD int nCart(int l) { return (l + 1) * (l + 2) / 2; }
    constexpr LBX_HD int nSph(int l) { return 2 * l + 1; }
    constexpr LBX_HD int nHerm(int l) { return l < 0 ? 0 : (l + 1) * (l + 2) * (l + 3) / 6; }
    // position of (j,k) (y, z exponents) inside one Cartesian/Hermite shell of order j+k+i
    constexpr LBX_HD int cartIdx(int j, int k) { return (j + k) * (j + k + 1) / 2 + k; }
    // position of (t,u,v) in the Hermite list ordered by t+u+v, then t, then u descending
    constexpr LBX_HD int hermIdx(int t, int u, int v) { return nHerm(t + u + v - 1) + cartIdx(u, v); }

    // inverse of hermIdx: packed (t | u<<8 | v<<16)
    constexpr LBX_HD int hermTuv(int idx)
    {
        int n = 0;
        while (nHerm(n) <= idx)
            n++;
        int r = idx - nHerm(n - 1);
        // r = cartIdx(u, v) with u + v = s:  s(s+1)/2 + v = r
        int s = 0;
        while ((s + 1) * (s + 2) / 2 <= r)
            s++;
        int v = r - s * (s + 1) / 2;
        int u = s - v;
        int t = n - s;
        return t | (u << 8) | (v << 16);
    }

    // class index of a pair class la >= lb in the order (0,0),(1,0),(1,1),(2,0),...
    // (getLPairsSymm, utils.cpp:49-59)
    constexpr LBX_HD int pairClassIdx(int la, int lb) { return la * (la + 1) / 2 + lb; }

    constexpr LBX_HD int ceilDiv(int a, int b) { return (a + b - 1) / b; }

    // Algorithmic FLOPs of one shell quartet with kab, kcd primitive pairs (ket transform,
    // bra transform, Hermite recursion + gather, J/K contraction).
    inline double quartetFlops(int la, int lb, int lc, int ld, double kab, double kcd)
    {
        double Hab = nHerm(la + lb), Hcd = nHerm(lc + ld);
        double Sab = nSph(la) * nSph(lb), Scd = nSph(lc) * nSph(ld);
        int L = la + lb + lc + ld;
        double r = 0;
        for (int n = 0; n < L; n++)
            for (int m = 1; m <= L - n; m++)
                r += (m + 1) * (m + 2) / 2;
        return 2 * kab * kcd * Hab * Hcd * Scd + 2 * kab * Sab * Hab * Scd +
               kab * kcd * (2 * r + Hab * Hcd) + 12 * Sab * Scd;
    }
}
