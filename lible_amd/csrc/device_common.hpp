// Device-side building blocks shared by the quartet kernels: Boys function, Hermite R
// recursion, recursion tables, wave helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include "lbx_common.hpp"

namespace lbx
{
    typedef double v4d __attribute__((ext_vector_type(4)));

    constexpr double kTwoPi52 = 34.98683665524972;   // 2 pi^{5/2}

    // ---- Boys function F_0..F_L(x) (boys_function.hpp:68-109): x = 0 exact values; x > 30
    // asymptotic form with upward recursion; otherwise 7-term Taylor expansion of F_L about the
    // tabulated point below x and downward recursion F_k = (2x F_{k+1} + e^-x)/(2k+1).
    template <int L>
    __device__ __forceinline__ void boysF(double x, const double *__restrict__ tab, double *F)
    {
        if (x == 0.0)
        {
            F[0] = 1.0;
#pragma unroll
            for (int k = 1; k <= L; k++)
                F[k] = 1.0 / (2 * k + 1);
        }
        else if (x > kBoysLargeX)
        {
            const double rx = 1.0 / x;
            F[0] = 0.5 * sqrt(M_PI * rx);
#pragma unroll
            for (int k = 1; k <= L; k++)
                F[k] = F[k - 1] * (k - 0.5) * rx;
        }
        else
        {
            const int iv = static_cast<int>(x * 100.0);
            const double dx = x - iv * kBoysStep;
            const double *row = tab + iv * kBoysCols + L;
            // sum_k F_{L+k}(x0) (-dx)^k / k!
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
#pragma unroll
            for (int k = L - 1; k >= 0; k--)
                F[k] = fma(x2, F[k + 1], ex) * (1.0 / (2 * k + 1));
        }
    }

    // The same evaluation split in two so that the table row can be fetched one iteration ahead
    // of its use (the row load is a random gather whose latency otherwise stalls the wave):
    // boysFetch loads the 7 tabulated values (any x: the row index is clamped to the table),
    // boysEval finishes with exactly the operations of boysF.
    struct BoysRow
    {
        double x;
        double r[7];
    };

    template <int L>
    __device__ __forceinline__ void boysFetch(double x, const double *__restrict__ tab, BoysRow &b)
    {
        b.x = x;
        int iv = static_cast<int>(x * 100.0);
        iv = iv < kBoysRows - 1 ? iv : kBoysRows - 1;
        iv = iv > 0 ? iv : 0;
        const double *row = tab + iv * kBoysCols + L;
#pragma unroll
        for (int j = 0; j < 7; j++)
            b.r[j] = row[j];
    }

    template <int L>
    __device__ __forceinline__ void boysEval(const BoysRow &b, double *F)
    {
        const double x = b.x;
        if (x == 0.0)
        {
            F[0] = 1.0;
#pragma unroll
            for (int k = 1; k <= L; k++)
                F[k] = 1.0 / (2 * k + 1);
        }
        else if (x > kBoysLargeX)
        {
            const double rx = 1.0 / x;
            F[0] = 0.5 * sqrt(M_PI * rx);
#pragma unroll
            for (int k = 1; k <= L; k++)
                F[k] = F[k - 1] * (k - 0.5) * rx;
        }
        else
        {
            const int iv = static_cast<int>(x * 100.0);
            const double dx = x - iv * kBoysStep;
            const double m = -dx;
            double s = b.r[6] * (1.0 / 720.0);
            s = fma(s, m, b.r[5] * (1.0 / 120.0));
            s = fma(s, m, b.r[4] * (1.0 / 24.0));
            s = fma(s, m, b.r[3] * (1.0 / 6.0));
            s = fma(s, m, b.r[2] * 0.5);
            s = fma(s, m, b.r[1]);
            s = fma(s, m, b.r[0]);
            F[L] = s;
            const double ex = exp(-x);
            const double x2 = 2.0 * x;
#pragma unroll
            for (int k = L - 1; k >= 0; k--)
                F[k] = fma(x2, F[k + 1], ex) * (1.0 / (2 * k + 1));
        }
    }

    // ---- Hermite R integrals in registers (HJO eqs. 9.9.18-9.9.20).  R holds nHerm(L)+L
    // slots: R[n] = R^{(n)}_{000} for n = 0..L on entry; on exit R_{tuv} (n = 0) sits at
    // slot 0 for (0,0,0) and at slot L + hermIdx(t,u,v) otherwise.  Levels n = L-1..0 update the
    // |tuv| = m slots in place in descending m, so reads of |tuv| = m-1, m-2 still see level n+1.
    template <int L>
    constexpr LBX_HD int rSlot(int t, int u, int v, int n)
    {
        return (t + u + v == 0) ? n : L + hermIdx(t, u, v);
    }

    // The recursion as a straight-line program generated at compile time: one operation
    // R[dst] = axis * R[s1] + coef * R[s2] per (level, tuv), in the in-place order above.
    struct ROp
    {
        short dst, s1, s2;
        signed char ax, coef;
    };

    constexpr LBX_HD int rOpCount(int L)
    {
        int n = 0;
        for (int lev = L - 1; lev >= 0; lev--)
            n += nHerm(L - lev) - 1;
        return n;
    }

    template <int L>
    struct RProgram
    {
        static constexpr int N = rOpCount(L);
        ROp op[N > 0 ? N : 1];
        constexpr RProgram() : op()
        {
            int k = 0;
            for (int n = L - 1; n >= 0; n--)
                for (int m = L - n; m >= 1; m--)
                    for (int t = m; t >= 0; t--)
                        for (int u = m - t; u >= 0; u--)
                        {
                            const int v = m - t - u;
                            ROp o{};
                            o.dst = static_cast<short>(L + hermIdx(t, u, v));
                            const int ax = t > 0 ? 0 : (u > 0 ? 1 : 2);
                            const int c = ax == 0 ? t : (ax == 1 ? u : v);
                            const int dt = ax == 0, du = ax == 1, dv = ax == 2;
                            o.ax = static_cast<signed char>(ax);
                            o.s1 = static_cast<short>(rSlot<L>(t - dt, u - du, v - dv, n + 1));
                            o.coef = static_cast<signed char>(c - 1);
                            o.s2 = static_cast<short>(c > 1 ? rSlot<L>(t - 2 * dt, u - 2 * du, v - 2 * dv, n + 1) : 0);
                            op[k++] = o;
                        }
        }
    };

    template <int L, typename T, size_t... I>
    __device__ __forceinline__ void hermiteRRun(T *R, T X, T Y, T Z, std::index_sequence<I...>)
    {
        constexpr RProgram<L> P{};
        ((R[P.op[I].dst] = (P.op[I].coef > 0)
                                   ? fma(T(P.op[I].coef), R[P.op[I].s2],
                                         (P.op[I].ax == 0 ? X : (P.op[I].ax == 1 ? Y : Z)) * R[P.op[I].s1])
                                   : (P.op[I].ax == 0 ? X : (P.op[I].ax == 1 ? Y : Z)) * R[P.op[I].s1]),
         ...);
    }

    // T = double, or float for the reduced-precision quartets (the R^{(n)}_{000} inputs are
    // formed in double from the double Boys values and rounded once)
    template <int L, typename T>
    __device__ __forceinline__ void hermiteR(T *R, T X, T Y, T Z)
    {
        hermiteRRun<L, T>(R, X, Y, Z, std::make_index_sequence<static_cast<size_t>(rOpCount(L))>{});
    }

    // slot in the hermiteR buffer of the Hermite index sum (tuv of bra index i) + (tuv of ket j)
    template <int L>
    constexpr LBX_HD int rSum(int i, int j)
    {
        const int a = hermTuv(i), b = hermTuv(j);
        const int t = (a & 255) + (b & 255), u = ((a >> 8) & 255) + ((b >> 8) & 255),
                  v = (a >> 16) + (b >> 16);
        return rSlot<L>(t, u, v, 0);
    }

    // rSum over a whole (i < NI, j < NJ) block, evaluated at compile time
    template <int L, int NI, int NJ>
    struct RSumTab
    {
        short v[NI * NJ];
        constexpr RSumTab() : v()
        {
            for (int i = 0; i < NI; i++)
                for (int j = 0; j < NJ; j++)
                    v[i * NJ + j] = static_cast<short>(rSum<L>(i, j));
        }
    };

    // ---- Structural zeros of the spherical E matrices.  The real solid harmonic S_{l,m}
    // (m = 0, +1, -1, +2, -2, ... for spherical slot 0, 1, 2, ...) is a combination of the
    // Cartesian monomials x^{|m|-p+2s} y^{p+2(t-s)} z^{l-|m|-2t} (0 <= s <= t <= (l-|m|)/2, p <= |m|,
    // p even for m >= 0 and odd for m < 0), and the 1-D Hermite coefficient E^{ij}_t vanishes
    // for t > i + j; so E[tuv][mu nu] is identically zero unless some monomial pair of mu and nu
    // covers (t, u, v).  Every other entry is generically nonzero.
    constexpr LBX_HD bool ePatternNz(int la, int lb, int h, int mn)
    {
        const int tuv = hermTuv(h);
        const int t = tuv & 255, u = (tuv >> 8) & 255, v = tuv >> 16;
        const int sb = 2 * lb + 1;
        const int mu = mn / sb, nu = mn % sb;
        const int ma = mu == 0 ? 0 : ((mu & 1) ? (mu + 1) / 2 : -(mu / 2));
        const int mb = nu == 0 ? 0 : ((nu & 1) ? (nu + 1) / 2 : -(nu / 2));
        const int aa = ma < 0 ? -ma : ma, ab = mb < 0 ? -mb : mb;
        for (int ta = 0; 2 * ta <= la - aa; ta++)
            for (int sa = 0; sa <= ta; sa++)
                for (int pa = 0; pa <= aa; pa++)
                {
                    if ((ma >= 0) != (pa % 2 == 0))
                        continue;
                    const int xa = aa - pa + 2 * sa, ya = pa + 2 * (ta - sa), za = la - aa - 2 * ta;
                    for (int tb = 0; 2 * tb <= lb - ab; tb++)
                        for (int sb2 = 0; sb2 <= tb; sb2++)
                            for (int pb = 0; pb <= ab; pb++)
                            {
                                if ((mb >= 0) != (pb % 2 == 0))
                                    continue;
                                const int xb = ab - pb + 2 * sb2, yb = pb + 2 * (tb - sb2),
                                          zb = lb - ab - 2 * tb;
                                if (t <= xa + xb && u <= ya + yb && v <= za + zb)
                                    return true;
                            }
                }
        return false;
    }

    template <int LA, int LB>
    struct EPattern
    {
        static constexpr int H = nHerm(LA + LB), S = (2 * LA + 1) * (2 * LB + 1);
        bool nz[H][S];
        constexpr EPattern() : nz()
        {
            for (int h = 0; h < H; h++)
                for (int s = 0; s < S; s++)
                    nz[h][s] = ePatternNz(LA, LB, h, s);
        }
    };

    constexpr LBX_HD int hermLevel(int i)
    {
        const int a = hermTuv(i);
        return (a & 255) + ((a >> 8) & 255) + (a >> 16);
    }

    // ---- Cooperative (LDS) recursion table: for Hermite index h >= 1, the source indices,
    // axis and multiplicity of R^{(n)}_h = X_axis R^{(n+1)}_{h-e} + (c-1) R^{(n+1)}_{h-2e}.
    struct RecTab
    {
        unsigned int v[nHerm(kMaxLBoys)];
        constexpr RecTab() : v()
        {
            for (int h = 1; h < nHerm(kMaxLBoys); h++)
            {
                const int a = hermTuv(h);
                const int t = a & 255, u = (a >> 8) & 255, w = a >> 16;
                int ax, c;
                if (t > 0)
                {
                    ax = 0;
                    c = t;
                }
                else if (u > 0)
                {
                    ax = 1;
                    c = u;
                }
                else
                {
                    ax = 2;
                    c = w;
                }
                const int dt = ax == 0, du = ax == 1, dw = ax == 2;
                const int s1 = hermIdx(t - dt, u - du, w - dw);
                const int s2 = c > 1 ? hermIdx(t - 2 * dt, u - 2 * du, w - 2 * dw) : 0;
                v[h] = static_cast<unsigned>(s1) | (static_cast<unsigned>(s2) << 11) |
                       (static_cast<unsigned>(ax) << 22) | (static_cast<unsigned>(c - 1) << 24);
            }
        }
    };

    static __constant__ RecTab kRecTab = RecTab();

    __device__ __forceinline__ double waveSum(double x)
    {
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1)
            x += __shfl_xor(x, o, 64);
        return x;
    }

    __device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

    // for (p = 0; p < N; p++) f(p) as a runtime loop.  Up to kStaticPasses passes each index
    // gets its own instance of f (called with integral_constant<int, p>) and one instance runs
    // per iteration; beyond that (code size and compile time) f is called with a plain int.
    constexpr int kStaticPasses = 3;

    template <int N, class F, int... P>
    __device__ __forceinline__ void forEachPassImpl(F &f, std::integer_sequence<int, P...>)
    {
#pragma unroll 1
        for (int p = 0; p < N; p++)
            ((p == P ? (f(std::integral_constant<int, P>{}), 0) : 0), ...);
    }

    template <int N, class F>
    __device__ __forceinline__ void forEachPass(F &f)
    {
        if constexpr (N <= kStaticPasses)
            forEachPassImpl<N>(f, std::make_integer_sequence<int, N>{});
        else
        {
#pragma unroll 1
            for (int p = 0; p < N; p++)
                f(p);
        }
    }

    // Read-only data at a wave-uniform address (the uniform side's E matrices and primitive
    // data), viewed through the constant address space: the compiler then emits scalar loads
    // into SGPRs, which the VALU FMAs take as operands directly, instead of vector loads that
    // broadcast one address into 64 VGPR lanes.  Only for memory no kernel writes meanwhile.
    template <typename T>
    using cptr = const T __attribute__((address_space(4))) *;

    template <typename T>
    __device__ __forceinline__ cptr<T> constView(const T *p)
    {
        return (cptr<T>)p;
    }

    // task t -> ket position q with prefix[q] <= t < prefix[q+1]
    __device__ __forceinline__ int findKet(const long long *__restrict__ prefix, int nket,
                                           long long t)
    {
        int lo = 0, hi = nket - 1;
        while (lo < hi)
        {
            int mid = (lo + hi + 1) >> 1;
            if (prefix[mid] <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
        return lo;
    }
}
