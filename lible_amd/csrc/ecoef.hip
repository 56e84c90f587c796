// SHARK E matrices on the device: for each primitive pair of a pair class and each spherical
// AO pair (mu, nu), the Hermite expansion
//   E[tuv][mu nu] = N_a[mu] N_b[nu] d_a d_b sum_{cart a, cart b} c_a c_b E^x_t E^y_u E^z_v
// of the product of two normalised spherical Gaussians (the quantity of ecoeffsSHARK,
// ecoeffs.cpp:314-403; single-shell form ecoeffs.cpp:253-312).  One thread per
// (primitive pair, mu nu).  Written twice: Eb unsigned (bra use) and Ek with (-1)^{t+u+v}
// folded in (ket use).  Both in [ppair][tuv][mu nu] order.
#include "device_common.hpp"
#include "lbx_internal.hpp"

#include <algorithm>

namespace lbx
{
    namespace
    {
        constexpr int kSphMax = 256;
        struct SphConst
        {
            int start[kMaxLAux + 1][2 * kMaxLAux + 2];   // entry range of each (l, m slot)
            int cart[kSphMax];
            double val[kSphMax];
        };
        __constant__ SphConst cSph;

        __device__ __forceinline__ void cartOf(int l, int idx, int &i, int &j, int &k)
        {
            // inverse of cartIdx within shell l: idx = s(s+1)/2 + k with s = j + k
            int s = 0;
            while ((s + 1) * (s + 2) / 2 <= idx)
                s++;
            k = idx - s * (s + 1) / 2;
            j = s - k;
            i = l - s;
        }

        // 1-D Hermite coefficients E^{ij}_t, i <= LA, j <= LB (HJO eqs. 9.5.6-9.5.7).
        template <int LA, int LB>
        __device__ __forceinline__ void hermite1d(double (&E)[LA + 1][LB + 1][LA + LB + 1], double K,
                                                  double PA, double PB, double o2p)
        {
            for (int i = 0; i <= LA; i++)
                for (int j = 0; j <= LB; j++)
                    for (int t = 0; t <= LA + LB; t++)
                        E[i][j][t] = 0.0;
            E[0][0][0] = K;
            for (int i = 1; i <= LA; i++)
                for (int t = 0; t <= i; t++)
                {
                    double v = PA * E[i - 1][0][t];
                    if (t > 0)
                        v += o2p * E[i - 1][0][t - 1];
                    if (t + 1 <= i - 1)
                        v += (t + 1) * E[i - 1][0][t + 1];
                    E[i][0][t] = v;
                }
            for (int j = 1; j <= LB; j++)
                for (int i = 0; i <= LA; i++)
                    for (int t = 0; t <= i + j; t++)
                    {
                        double v = PB * E[i][j - 1][t];
                        if (t > 0)
                            v += o2p * E[i][j - 1][t - 1];
                        if (t + 1 <= i + j - 1)
                            v += (t + 1) * E[i][j - 1][t + 1];
                        E[i][j][t] = v;
                    }
        }

        // The same coefficients and their derivatives with respect to X = A_x - B_x.  E depends
        // on X through K = exp(-mu X^2), PA = -(b/p) X and PB = (a/p) X, so differentiating the
        // recurrences term by term gives
        //   dE^{i+1,0}_t = -(b/p) E^{i0}_t + PA dE^{i0}_t + dE^{i0}_{t-1}/(2p) + (t+1) dE^{i0}_{t+1}
        //   dE^{i,j+1}_t =  (a/p) E^{ij}_t + PB dE^{ij}_t + dE^{ij}_{t-1}/(2p) + (t+1) dE^{ij}_{t+1}
        // from dE^{00}_0 = -2 mu X K.
        template <int LA, int LB>
        __device__ __forceinline__ void hermite1dD1(double (&E)[LA + 1][LB + 1][LA + LB + 1],
                                                    double (&dE)[LA + 1][LB + 1][LA + LB + 1],
                                                    double X, double mur, double a, double b)
        {
            const double p = a + b, o2p = 1.0 / (2.0 * p), K = exp(-mur * X * X);
            const double PA = -(b / p) * X, PB = (a / p) * X, dPA = -b / p, dPB = a / p;
            for (int i = 0; i <= LA; i++)
                for (int j = 0; j <= LB; j++)
                    for (int t = 0; t <= LA + LB; t++)
                    {
                        E[i][j][t] = 0.0;
                        dE[i][j][t] = 0.0;
                    }
            E[0][0][0] = K;
            dE[0][0][0] = -2.0 * mur * X * K;
            for (int i = 1; i <= LA; i++)
                for (int t = 0; t <= i; t++)
                {
                    double v = PA * E[i - 1][0][t], dv = dPA * E[i - 1][0][t] + PA * dE[i - 1][0][t];
                    if (t > 0)
                    {
                        v += o2p * E[i - 1][0][t - 1];
                        dv += o2p * dE[i - 1][0][t - 1];
                    }
                    if (t + 1 <= i - 1)
                    {
                        v += (t + 1) * E[i - 1][0][t + 1];
                        dv += (t + 1) * dE[i - 1][0][t + 1];
                    }
                    E[i][0][t] = v;
                    dE[i][0][t] = dv;
                }
            for (int j = 1; j <= LB; j++)
                for (int i = 0; i <= LA; i++)
                    for (int t = 0; t <= i + j; t++)
                    {
                        double v = PB * E[i][j - 1][t], dv = dPB * E[i][j - 1][t] + PB * dE[i][j - 1][t];
                        if (t > 0)
                        {
                            v += o2p * E[i][j - 1][t - 1];
                            dv += o2p * dE[i][j - 1][t - 1];
                        }
                        if (t + 1 <= i + j - 1)
                        {
                            v += (t + 1) * E[i][j - 1][t + 1];
                            dv += (t + 1) * dE[i][j - 1][t + 1];
                        }
                        E[i][j][t] = v;
                        dE[i][j][t] = dv;
                    }
        }

        // d E[tuv][mu nu] / d(A - B)_x for x = 0, 1, 2, one thread per (primitive pair, mu nu);
        // written [ppair][x][tuv][mu nu], unsigned (e1b) and with (-1)^{t+u+v} (e1k).
        template <int LA, int LB>
        __global__ __launch_bounds__(256) void ecoefD1Kernel(int nppairs, const PrimIn *__restrict__ prim,
                                                             const double *__restrict__ centers,
                                                             const double *__restrict__ norms,
                                                             double *__restrict__ e1b,
                                                             double *__restrict__ e1k)
        {
            constexpr int SA = nSph(LA), SB = nSph(LB), SP = SA * SB, HP = nHerm(LA + LB);
            const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
            if (gid >= static_cast<long long>(nppairs) * SP)
                return;
            const int pp = static_cast<int>(gid / SP), mn = static_cast<int>(gid % SP);
            const int mu = mn / SB, nu = mn % SB;
            const PrimIn pr = prim[pp];
            const double a = pr.a, b = pr.b, mur = a * b / (a + b);
            const double *C = centers + 6 * pr.pair;
            double E[3][LA + 1][LB + 1][LA + LB + 1], dE[3][LA + 1][LB + 1][LA + LB + 1];
            for (int x = 0; x < 3; x++)
                hermite1dD1<LA, LB>(E[x], dE[x], C[x] - C[3 + x], mur, a, b);
            const double *nrm = norms + static_cast<size_t>(pr.pair) * (SA + SB);
            const double sc = nrm[mu] * nrm[SA + nu] * pr.dadb;
            for (int x = 0; x < 3; x++)
            {
                double acc[HP];
                for (int h = 0; h < HP; h++)
                    acc[h] = 0.0;
                for (int ea = cSph.start[LA][mu]; ea < cSph.start[LA][mu + 1]; ea++)
                {
                    int ia[3];
                    cartOf(LA, cSph.cart[ea], ia[0], ia[1], ia[2]);
                    for (int eb_ = cSph.start[LB][nu]; eb_ < cSph.start[LB][nu + 1]; eb_++)
                    {
                        int jb[3];
                        cartOf(LB, cSph.cart[eb_], jb[0], jb[1], jb[2]);
                        const double v = cSph.val[ea] * cSph.val[eb_];
                        for (int t = 0; t <= ia[0] + jb[0]; t++)
                            for (int u = 0; u <= ia[1] + jb[1]; u++)
                                for (int w = 0; w <= ia[2] + jb[2]; w++)
                                {
                                    const double fx = x == 0 ? dE[0][ia[0]][jb[0]][t] : E[0][ia[0]][jb[0]][t];
                                    const double fy = x == 1 ? dE[1][ia[1]][jb[1]][u] : E[1][ia[1]][jb[1]][u];
                                    const double fz = x == 2 ? dE[2][ia[2]][jb[2]][w] : E[2][ia[2]][jb[2]][w];
                                    acc[hermIdx(t, u, w)] += v * fx * fy * fz;
                                }
                    }
                }
                for (int h = 0; h < HP; h++)
                {
                    const double val = sc * acc[h];
                    const size_t o = ((static_cast<size_t>(pp) * 3 + x) * HP + h) * SP + mn;
                    e1b[o] = val;
                    e1k[o] = (hermLevel(h) & 1) ? -val : val;
                }
            }
        }

        template <int LA, int LB>
        __global__ __launch_bounds__(256) void ecoefKernel(int nppairs, const PrimIn *__restrict__ prim,
                                                           const double *__restrict__ centers,
                                                           const double *__restrict__ norms,
                                                           double *__restrict__ eb,
                                                           double *__restrict__ ek)
        {
            constexpr int SA = nSph(LA), SB = nSph(LB), SP = SA * SB, HP = nHerm(LA + LB);
            const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
            if (gid >= static_cast<long long>(nppairs) * SP)
                return;
            const int pp = static_cast<int>(gid / SP), mn = static_cast<int>(gid % SP);
            const int mu = mn / SB, nu = mn % SB;
            const PrimIn pr = prim[pp];
            const double a = pr.a, b = pr.b, p = a + b, o2p = 1.0 / (2.0 * p), mur = a * b / p;
            const double *C = centers + 6 * pr.pair;
            double Ex[LA + 1][LB + 1][LA + LB + 1], Ey[LA + 1][LB + 1][LA + LB + 1],
                    Ez[LA + 1][LB + 1][LA + LB + 1];
            {
                double P[3], K[3];
                for (int x = 0; x < 3; x++)
                {
                    P[x] = (a * C[x] + b * C[3 + x]) / p;
                    const double d = C[x] - C[3 + x];
                    K[x] = exp(-mur * d * d);
                }
                hermite1d<LA, LB>(Ex, K[0], P[0] - C[0], P[0] - C[3], o2p);
                hermite1d<LA, LB>(Ey, K[1], P[1] - C[1], P[1] - C[4], o2p);
                hermite1d<LA, LB>(Ez, K[2], P[2] - C[2], P[2] - C[5], o2p);
            }
            double acc[HP];
            for (int h = 0; h < HP; h++)
                acc[h] = 0.0;
            for (int ea = cSph.start[LA][mu]; ea < cSph.start[LA][mu + 1]; ea++)
            {
                int ix, iy, iz;
                cartOf(LA, cSph.cart[ea], ix, iy, iz);
                for (int eb_ = cSph.start[LB][nu]; eb_ < cSph.start[LB][nu + 1]; eb_++)
                {
                    int jx, jy, jz;
                    cartOf(LB, cSph.cart[eb_], jx, jy, jz);
                    const double v = cSph.val[ea] * cSph.val[eb_];
                    for (int t = 0; t <= ix + jx; t++)
                        for (int u = 0; u <= iy + jy; u++)
                            for (int w = 0; w <= iz + jz; w++)
                                acc[hermIdx(t, u, w)] += v * Ex[ix][jx][t] * Ey[iy][jy][u] * Ez[iz][jz][w];
                }
            }
            const double *nrm = norms + static_cast<size_t>(pr.pair) * (SA + SB);
            const double sc = nrm[mu] * nrm[SA + nu] * pr.dadb;
            for (int h = 0; h < HP; h++)
            {
                const double val = sc * acc[h];
                const size_t o = (static_cast<size_t>(pp) * HP + h) * SP + mn;
                eb[o] = val;
                ek[o] = (hermLevel(h) & 1) ? -val : val;
            }
        }

        using EcoefFn = void (*)(int, const PrimIn *, const double *, const double *, double *,
                                 double *, hipStream_t);

        template <int LA, int LB>
        void launchEcoef(int n, const PrimIn *prim, const double *centers, const double *norms,
                         double *eb, double *ek, hipStream_t s)
        {
            constexpr int SP = nSph(LA) * nSph(LB);
            const long long total = static_cast<long long>(n) * SP;
            if (total == 0)
                return;
            const int grid = static_cast<int>((total + 255) / 256);
            hipLaunchKernelGGL((ecoefKernel<LA, LB>), dim3(grid), dim3(256), 0, s, n, prim, centers,
                               norms, eb, ek);
        }

        template <int LA, int LB>
        void launchEcoefD1(int n, const PrimIn *prim, const double *centers, const double *norms,
                           double *e1b, double *e1k, hipStream_t s)
        {
            constexpr int SP = nSph(LA) * nSph(LB);
            const long long total = static_cast<long long>(n) * SP;
            if (total == 0)
                return;
            const int grid = static_cast<int>((total + 255) / 256);
            hipLaunchKernelGGL((ecoefD1Kernel<LA, LB>), dim3(grid), dim3(256), 0, s, n, prim, centers,
                               norms, e1b, e1k);
        }

        EcoefFn ecoefD1Fn(int la, int lb)
        {
#define LBX_E(A, B)                                                                                \
    if (la == A && lb == B)                                                                        \
        return &launchEcoefD1<A, B>;
            LBX_E(0, 0) LBX_E(1, 0) LBX_E(1, 1) LBX_E(2, 0) LBX_E(2, 1) LBX_E(2, 2) LBX_E(3, 0)
            LBX_E(3, 1) LBX_E(3, 2) LBX_E(3, 3) LBX_E(4, 0) LBX_E(4, 1) LBX_E(4, 2) LBX_E(4, 3) LBX_E(4, 4)
#undef LBX_E
            throw Error(LBX_ERR_INVALID, "no derivative E-matrix kernel for this angular momentum pair");
        }

        EcoefFn ecoefFn(int la, int lb)
        {
#define LBX_E(A, B)                                                                                \
    if (la == A && lb == B)                                                                        \
        return &launchEcoef<A, B>;
            LBX_E(0, 0) LBX_E(1, 0) LBX_E(1, 1) LBX_E(2, 0) LBX_E(2, 1) LBX_E(2, 2) LBX_E(3, 0)
            LBX_E(3, 1) LBX_E(3, 2) LBX_E(3, 3) LBX_E(4, 0) LBX_E(4, 1) LBX_E(4, 2) LBX_E(4, 3) LBX_E(4, 4)
            LBX_E(5, 0)
#undef LBX_E
            throw Error(LBX_ERR_INVALID, "no E-matrix kernel for this angular momentum pair");
        }

        // ---- one-electron integrals (overlap, kinetic energy, attraction to point charges) ----
        // E^{ij}_0 alone: with E^{ij}_1 = (i E^{i-1,j}_0 + j E^{i,j-1}_0) / (2p) the t = 0 row of
        // the Hermite recurrences closes on itself (the Obara-Saika overlap recurrence, HJO
        // eq. 9.3.8), so S holds (LA+1)(LB+1) values instead of the full E^{ij}_t table.
        template <int LA, int LB>
        __device__ __forceinline__ void overlap1d(double (&S)[LA + 1][LB + 1], double K, double PA, double PB,
                                                  double o2p)
        {
            S[0][0] = K;
            for (int i = 0; i < LA; i++)
                S[i + 1][0] = PA * S[i][0] + (i > 0 ? i * o2p * S[i - 1][0] : 0.0);
            for (int j = 0; j < LB; j++)
                for (int i = 0; i <= LA; i++)
                {
                    double v = PB * S[i][j];
                    if (i > 0)
                        v += i * o2p * S[i - 1][j];
                    if (j > 0)
                        v += j * o2p * S[i][j - 1];
                    S[i][j + 1] = v;
                }
        }

        // Overlap and kinetic energy of one (primitive pair, mu nu) from the Cartesian 1-D
        // Hermite coefficients E^{ij}_0 (oneelints.cpp:57-100, 199-276): per axis
        //   S_ij = E^{ij}_0,   T_ij = b (2j + 1) E^{ij}_0 - 2 b^2 E^{i,j+2}_0 - j (j - 1)/2 E^{i,j-2}_0
        // (-1/2 d^2/dx^2 acting on b's factor), times (pi/p)^{3/2}, the contraction coefficients
        // and the spherical norms; added into the dense S and T at (a, b) and (b, a).
        template <int LA, int LB>
        __global__ __launch_bounds__(256) void stKernel(int nppairs, const PrimIn *__restrict__ prim,
                                                        const double *__restrict__ centers,
                                                        const double *__restrict__ norms, const int *__restrict__ ao_a,
                                                        const int *__restrict__ ao_b, const int *__restrict__ sh_a,
                                                        const int *__restrict__ sh_b, long long n, double *S,
                                                        double *T)
        {
            constexpr int SA = nSph(LA), SB = nSph(LB), SP = SA * SB, LB2 = LB + 2;
            const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
            if (gid >= static_cast<long long>(nppairs) * SP)
                return;
            const int pp = static_cast<int>(gid / SP), mn = static_cast<int>(gid % SP);
            const int mu = mn / SB, nu = mn % SB;
            const PrimIn pr = prim[pp];
            const double a = pr.a, b = pr.b, p = a + b, o2p = 1.0 / (2.0 * p), mur = a * b / p;
            const double *C = centers + 6 * pr.pair;
            double E[3][LA + 1][LB2 + 1];
            for (int x = 0; x < 3; x++)
            {
                const double P = (a * C[x] + b * C[3 + x]) / p, d = C[x] - C[3 + x];
                overlap1d<LA, LB2>(E[x], exp(-mur * d * d), P - C[x], P - C[3 + x], o2p);
            }
            auto t1 = [&](int x, int i, int j) {
                double t = b * (2 * j + 1) * E[x][i][j] - 2.0 * b * b * E[x][i][j + 2];
                if (j >= 2)
                    t -= 0.5 * j * (j - 1) * E[x][i][j - 2];
                return t;
            };
            double s = 0.0, t = 0.0;
            for (int ea = cSph.start[LA][mu]; ea < cSph.start[LA][mu + 1]; ea++)
            {
                int ix, iy, iz;
                cartOf(LA, cSph.cart[ea], ix, iy, iz);
                for (int eb_ = cSph.start[LB][nu]; eb_ < cSph.start[LB][nu + 1]; eb_++)
                {
                    int jx, jy, jz;
                    cartOf(LB, cSph.cart[eb_], jx, jy, jz);
                    const double v = cSph.val[ea] * cSph.val[eb_];
                    const double sx = E[0][ix][jx], sy = E[1][iy][jy], sz = E[2][iz][jz];
                    s += v * sx * sy * sz;
                    t += v * (t1(0, ix, jx) * sy * sz + sx * t1(1, iy, jy) * sz + sx * sy * t1(2, iz, jz));
                }
            }
            const double *nrm = norms + static_cast<size_t>(pr.pair) * (SA + SB);
            const double sc = nrm[mu] * nrm[SA + nu] * pr.dadb * pow(M_PI / p, 1.5);
            const long long r = ao_a[pr.pair] + mu, c = ao_b[pr.pair] + nu;
            const bool both = sh_a[pr.pair] != sh_b[pr.pair];
            if (S)
            {
                unsafeAtomicAdd(&S[r * n + c], sc * s);
                if (both)
                    unsafeAtomicAdd(&S[c * n + r], sc * s);
            }
            if (T)
            {
                unsafeAtomicAdd(&T[r * n + c], sc * t);
                if (both)
                    unsafeAtomicAdd(&T[c * n + r], sc * t);
            }
        }

        // Attraction of an electron to point charges q_C at C (oneelints.cpp:508-585):
        //   V(mu nu) = -(2 pi / p) sum_tuv E[tuv][mu nu] sum_C q_C R_tuv(p, P - C)
        // with the pair's spherical E matrix (norms and coefficients included) and R from
        // R^(n)_000 = (-2p)^n F_n(p |P - C|^2).
        template <int LA, int LB>
        __global__ __launch_bounds__(256) void vKernel(int nppairs, const PrimIn *__restrict__ prim,
                                                       const double4 *__restrict__ ppd,
                                                       const double *__restrict__ eb, const double4 *__restrict__ charges,
                                                       int ncharges, const double *__restrict__ boys,
                                                       const int *__restrict__ ao_a, const int *__restrict__ ao_b,
                                                       const int *__restrict__ sh_a, const int *__restrict__ sh_b,
                                                       long long n, double *V)
        {
            constexpr int SA = nSph(LA), SB = nSph(LB), SP = SA * SB, L = LA + LB, HP = nHerm(L);
            const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
            if (gid >= static_cast<long long>(nppairs) * SP)
                return;
            const int pp = static_cast<int>(gid / SP), mn = static_cast<int>(gid % SP);
            const int mu = mn / SB, nu = mn % SB;
            const PrimIn pr = prim[pp];
            const double4 P = ppd[pp];
            // this thread's column of E; the SP threads of one primitive pair read consecutive
            // addresses, and re-reads per charge hit the cache
            const double *e = eb + static_cast<size_t>(pp) * HP * SP + mn;
            double v = 0.0;
            for (int k = 0; k < ncharges; k++)
            {
                const double4 q = charges[k];
                const double X = P.y - q.x, Y = P.z - q.y, Z = P.w - q.z;
                double R[HP + L];
                boysF<L>(P.x * (X * X + Y * Y + Z * Z), boys, R);
                double sc = -2.0 * P.x;
#pragma unroll
                for (int m = 1; m <= L; m++)
                {
                    R[m] *= sc;
                    sc *= -2.0 * P.x;
                }
                hermiteR<L, double>(R, X, Y, Z);
                // R_tuv sits at slot 0 for tuv = 000 and at L + hermIdx(t, u, v) otherwise
                double acc = e[0] * R[0];
#pragma unroll
                for (int h = 1; h < HP; h++)
                    acc = fma(e[static_cast<size_t>(h) * SP], R[L + h], acc);
                v = fma(q.w, acc, v);
            }
            v *= -2.0 * M_PI / P.x;
            const long long r = ao_a[pr.pair] + mu, c = ao_b[pr.pair] + nu;
            unsafeAtomicAdd(&V[r * n + c], v);
            if (sh_a[pr.pair] != sh_b[pr.pair])
                unsafeAtomicAdd(&V[c * n + r], v);
        }

        // ---- nuclear gradient of the one-electron energy tr(D (T + V)) - tr(W S) ----
        // Overlap and kinetic parts (oneelints.cpp overlapD1Kernel / kineticEnergyD1Kernel): the
        // derivative of a Cartesian Gaussian with respect to its centre, d/dA_x x^i e^{-a x^2} ->
        // 2a (i+1 term) - i (i-1 term), applied per axis to the 1-D overlap and kinetic factors;
        // d/dB = -d/dA (the integrals depend on A - B only), so pairs within one shell give zero.
        // One thread per primitive pair, contracted with D and W over the pair's AO block.
        template <int LA, int LB>
        __global__ __launch_bounds__(256) void stGradKernel(int nppairs, const PrimIn *__restrict__ prim,
                                                            const double *__restrict__ centers,
                                                            const double *__restrict__ norms,
                                                            const int *__restrict__ ao_a, const int *__restrict__ ao_b,
                                                            const int *__restrict__ sh_a, const int *__restrict__ sh_b,
                                                            const int *__restrict__ atom_a,
                                                            const int *__restrict__ atom_b, long long n,
                                                            const double *__restrict__ D,
                                                            const double *__restrict__ W, double *grad)
        {
            constexpr int SA = nSph(LA), SB = nSph(LB), LA1 = LA + 1, LB2 = LB + 2;
            const int pp = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
            if (pp >= nppairs)
                return;
            const PrimIn pr = prim[pp];
            if (sh_a[pr.pair] == sh_b[pr.pair])
                return;
            const double a = pr.a, b = pr.b, p = a + b, o2p = 1.0 / (2.0 * p), mur = a * b / p;
            const double *C = centers + 6 * pr.pair;
            double S[3][LA1 + 1][LB2 + 1];
            for (int x = 0; x < 3; x++)
            {
                const double P = (a * C[x] + b * C[3 + x]) / p, d = C[x] - C[3 + x];
                overlap1d<LA1, LB2>(S[x], exp(-mur * d * d), P - C[x], P - C[3 + x], o2p);
            }
            auto t1 = [&](int x, int i, int j) {
                double t = b * (2 * j + 1) * S[x][i][j] - 2.0 * b * b * S[x][i][j + 2];
                if (j >= 2)
                    t -= 0.5 * j * (j - 1) * S[x][i][j - 2];
                return t;
            };
            auto ds = [&](int x, int i, int j) { return 2.0 * a * S[x][i + 1][j] - (i > 0 ? i * S[x][i - 1][j] : 0.0); };
            auto dt = [&](int x, int i, int j) { return 2.0 * a * t1(x, i + 1, j) - (i > 0 ? i * t1(x, i - 1, j) : 0.0); };
            const double *nrm = norms + static_cast<size_t>(pr.pair) * (SA + SB);
            const double pre = pr.dadb * pow(M_PI / p, 1.5);
            const long long r0 = ao_a[pr.pair], c0 = ao_b[pr.pair];
            double g[3] = {0.0, 0.0, 0.0};
            for (int mu = 0; mu < SA; mu++)
                for (int nu = 0; nu < SB; nu++)
                {
                    const long long r = r0 + mu, c = c0 + nu;
                    const double wd = D[r * n + c] + D[c * n + r];
                    const double ww = W ? W[r * n + c] + W[c * n + r] : 0.0;
                    double dS[3] = {0.0, 0.0, 0.0}, dT[3] = {0.0, 0.0, 0.0};
                    for (int ea = cSph.start[LA][mu]; ea < cSph.start[LA][mu + 1]; ea++)
                    {
                        int i[3];
                        cartOf(LA, cSph.cart[ea], i[0], i[1], i[2]);
                        for (int eb_ = cSph.start[LB][nu]; eb_ < cSph.start[LB][nu + 1]; eb_++)
                        {
                            int j[3];
                            cartOf(LB, cSph.cart[eb_], j[0], j[1], j[2]);
                            const double v = cSph.val[ea] * cSph.val[eb_];
                            double s1[3], k1[3];
                            for (int x = 0; x < 3; x++)
                            {
                                s1[x] = S[x][i[x]][j[x]];
                                k1[x] = t1(x, i[x], j[x]);
                            }
                            for (int x = 0; x < 3; x++)
                            {
                                const int y = (x + 1) % 3, z = (x + 2) % 3;
                                const double dsx = ds(x, i[x], j[x]);
                                dS[x] += v * dsx * s1[y] * s1[z];
                                dT[x] += v * (dt(x, i[x], j[x]) * s1[y] * s1[z] +
                                              dsx * (k1[y] * s1[z] + s1[y] * k1[z]));
                            }
                        }
                    }
                    const double sc = nrm[mu] * nrm[SA + nu] * pre;
                    for (int x = 0; x < 3; x++)
                        g[x] += sc * (wd * dT[x] - ww * dS[x]);
                }
            const int ia = atom_a[pr.pair], ib = atom_b[pr.pair];
            for (int x = 0; x < 3; x++)
            {
                unsafeAtomicAdd(&grad[3 * ia + x], g[x]);
                unsafeAtomicAdd(&grad[3 * ib + x], -g[x]);
            }
        }

        // Point-charge attraction, step 1: the pair's E matrix and its A - B derivatives
        // contracted with D over the AO block, ed[q][h][ppair] (q = 0: E, 1..3: dE/dX_x), so that
        // step 2 touches one number per Hermite index and charge.
        template <int LA, int LB>
        __global__ __launch_bounds__(256) void vContractKernel(int nppairs, const PrimIn *__restrict__ prim,
                                                               const double *__restrict__ eb,
                                                               const double *__restrict__ e1b,
                                                               const int *__restrict__ ao_a,
                                                               const int *__restrict__ ao_b,
                                                               const int *__restrict__ sh_a,
                                                               const int *__restrict__ sh_b, long long n,
                                                               const double *__restrict__ D, double *ed)
        {
            constexpr int SA = nSph(LA), SB = nSph(LB), SP = SA * SB, HP = nHerm(LA + LB);
            const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
            if (gid >= static_cast<long long>(nppairs) * HP)
                return;
            const int pp = static_cast<int>(gid / HP), h = static_cast<int>(gid % HP);
            const int pair = prim[pp].pair;
            const bool same = sh_a[pair] == sh_b[pair];
            const long long r0 = ao_a[pair], c0 = ao_b[pair];
            const double *e = eb + (static_cast<size_t>(pp) * HP + h) * SP;
            const double *e1 = e1b + static_cast<size_t>(pp) * 3 * HP * SP + static_cast<size_t>(h) * SP;
            double s[4] = {0.0, 0.0, 0.0, 0.0};
            for (int mu = 0; mu < SA; mu++)
                for (int nu = 0; nu < SB; nu++)
                {
                    const long long r = r0 + mu, c = c0 + nu;
                    const double w = same ? D[r * n + c] : D[r * n + c] + D[c * n + r];
                    const int mn = mu * SB + nu;
                    s[0] = fma(w, e[mn], s[0]);
                    for (int x = 0; x < 3; x++)
                        s[1 + x] = fma(w, e1[static_cast<size_t>(x) * HP * SP + mn], s[1 + x]);
                }
            for (int q = 0; q < 4; q++)
                ed[(static_cast<size_t>(q) * HP + h) * nppairs + pp] = s[q];
        }

        // Step 2, one thread per primitive pair, charges in a uniform loop.  With
        // V = -(2 pi / p) q sum_h E_h R_h(p, P - C), X = A - B and P = (a A + b B) / p:
        //   dV/dX_x = -(2 pi / p) q sum_h (dE_h/dX_x) R_h,   dV/dP_x = -(2 pi / p) q sum_h E_h R_{h+1_x}
        //   d/dA = d/dX + (a/p) d/dP,   d/dB = -d/dX + (b/p) d/dP,   d/dC = -d/dP
        // (the orbital part, externalChargesD1Kernel, and the operator part,
        // externalChargesOperatorD1Kernel, oneelints.cpp).  The operator part of each charge is
        // summed over the wave before one atomic add.
        template <int LA, int LB>
        __global__ __launch_bounds__(256) void vGradKernel(int nppairs, const PrimIn *__restrict__ prim,
                                                           const double4 *__restrict__ ppd,
                                                           const double *__restrict__ ed,
                                                           const double4 *__restrict__ charges, int ncharges,
                                                           const double *__restrict__ boys,
                                                           const int *__restrict__ atom_a,
                                                           const int *__restrict__ atom_b, double *grad,
                                                           double *charge_grad)
        {
            constexpr int L = LA + LB, L1 = L + 1, HP = nHerm(L);
            const int pp0 = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
            const bool valid = pp0 < nppairs;
            const int pp = valid ? pp0 : 0;
            const PrimIn pr = prim[pp];
            const double4 P = ppd[pp];
            const double ra = pr.a / P.x;
            double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};
            for (int k = 0; k < ncharges; k++)
            {
                const double4 q = charges[k];
                const double X = P.y - q.x, Y = P.z - q.y, Z = P.w - q.z;
                double R[nHerm(L1) + L1];
                boysF<L1>(P.x * (X * X + Y * Y + Z * Z), boys, R);
                double sc = -2.0 * P.x;
#pragma unroll
                for (int m = 1; m <= L1; m++)
                {
                    R[m] *= sc;
                    sc *= -2.0 * P.x;
                }
                hermiteR<L1, double>(R, X, Y, Z);
                double vP[3] = {0.0, 0.0, 0.0}, vX[3] = {0.0, 0.0, 0.0};
#pragma unroll
                for (int h = 0; h < HP; h++)
                {
                    const int tuv = hermTuv(h);
                    const int t = tuv & 255, u = (tuv >> 8) & 255, v = tuv >> 16;
                    const double e0 = ed[static_cast<size_t>(h) * nppairs + pp];
                    const double rh = R[h == 0 ? 0 : L1 + h];
                    vP[0] = fma(e0, R[L1 + hermIdx(t + 1, u, v)], vP[0]);
                    vP[1] = fma(e0, R[L1 + hermIdx(t, u + 1, v)], vP[1]);
                    vP[2] = fma(e0, R[L1 + hermIdx(t, u, v + 1)], vP[2]);
                    for (int x = 0; x < 3; x++)
                        vX[x] = fma(ed[(static_cast<size_t>(1 + x) * HP + h) * nppairs + pp], rh, vX[x]);
                }
                const double pre = valid ? -2.0 * M_PI / P.x * q.w : 0.0;
                for (int x = 0; x < 3; x++)
                {
                    const double dP = pre * vP[x], dX = pre * vX[x];
                    gA[x] += dX + ra * dP;
                    gB[x] += -dX + (1.0 - ra) * dP;
                    if (charge_grad)
                    {
                        const double cg = waveSum(-dP);
                        if ((threadIdx.x & 63) == 0)
                            unsafeAtomicAdd(&charge_grad[3 * k + x], cg);
                    }
                }
            }
            if (!valid)
                return;
            const int ia = atom_a[pr.pair], ib = atom_b[pr.pair];
            for (int x = 0; x < 3; x++)
            {
                unsafeAtomicAdd(&grad[3 * ia + x], gA[x]);
                unsafeAtomicAdd(&grad[3 * ib + x], gB[x]);
            }
        }

        using StFn = void (*)(const PairClass &, long long, double *, double *, hipStream_t);
        using VFn = void (*)(const PairClass &, const double4 *, int, const double *, long long, double *,
                             hipStream_t);

        template <int LA, int LB>
        void launchST(const PairClass &pc, long long n, double *S, double *T, hipStream_t s)
        {
            const long long total = static_cast<long long>(pc.n_ppairs) * nSph(LA) * nSph(LB);
            if (total == 0)
                return;
            hipLaunchKernelGGL((stKernel<LA, LB>), dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, s,
                               pc.n_ppairs, pc.d_prim, pc.d_centers, pc.d_norms, pc.d_ao_a, pc.d_ao_b,
                               pc.d_sh_a, pc.d_sh_b, n, S, T);
        }

        template <int LA, int LB>
        void launchV(const PairClass &pc, const double4 *charges, int ncharges, const double *boys, long long n,
                     double *V, hipStream_t s)
        {
            const long long total = static_cast<long long>(pc.n_ppairs) * nSph(LA) * nSph(LB);
            if (total == 0)
                return;
            hipLaunchKernelGGL((vKernel<LA, LB>), dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, s,
                               pc.n_ppairs, pc.d_prim, pc.d_pp, pc.d_eb, charges, ncharges, boys, pc.d_ao_a,
                               pc.d_ao_b, pc.d_sh_a, pc.d_sh_b, n, V);
        }

        using GradFn = void (*)(const PairClass &, long long, const double *, const double *, const double4 *, int,
                                const double *, double *, double *, double *, hipStream_t);

        template <int LA, int LB>
        void launchOneGrad(const PairClass &pc, long long n, const double *D, const double *W, const double4 *charges,
                           int ncharges, const double *boys, double *ed, double *grad, double *charge_grad,
                           hipStream_t s)
        {
            const int npp = pc.n_ppairs;
            if (npp == 0)
                return;
            const unsigned g1 = static_cast<unsigned>((npp + 255) / 256);
            hipLaunchKernelGGL((stGradKernel<LA, LB>), dim3(g1), dim3(256), 0, s, npp, pc.d_prim, pc.d_centers,
                               pc.d_norms, pc.d_ao_a, pc.d_ao_b, pc.d_sh_a, pc.d_sh_b, pc.d_atom_a, pc.d_atom_b, n, D,
                               W, grad);
            if (ncharges == 0)
                return;
            const long long tot = static_cast<long long>(npp) * nHerm(LA + LB);
            hipLaunchKernelGGL((vContractKernel<LA, LB>), dim3(static_cast<unsigned>((tot + 255) / 256)), dim3(256), 0,
                               s, npp, pc.d_prim, pc.d_eb, pc.d_e1b, pc.d_ao_a, pc.d_ao_b, pc.d_sh_a, pc.d_sh_b, n, D,
                               ed);
            hipLaunchKernelGGL((vGradKernel<LA, LB>), dim3(g1), dim3(256), 0, s, npp, pc.d_prim, pc.d_pp, ed, charges,
                               ncharges, boys, pc.d_atom_a, pc.d_atom_b, grad, charge_grad);
        }

        GradFn oneGradFn(int la, int lb)
        {
#define LBX_E(A, B)                                                                                \
    if (la == A && lb == B)                                                                        \
        return &launchOneGrad<A, B>;
            LBX_E(0, 0) LBX_E(1, 0) LBX_E(1, 1) LBX_E(2, 0) LBX_E(2, 1) LBX_E(2, 2) LBX_E(3, 0)
            LBX_E(3, 1) LBX_E(3, 2) LBX_E(3, 3) LBX_E(4, 0) LBX_E(4, 1) LBX_E(4, 2) LBX_E(4, 3) LBX_E(4, 4)
#undef LBX_E
            throw Error(LBX_ERR_INVALID, "no one-electron gradient kernel for this angular momentum pair");
        }

        StFn stFn(int la, int lb)
        {
#define LBX_E(A, B)                                                                                \
    if (la == A && lb == B)                                                                        \
        return &launchST<A, B>;
            LBX_E(0, 0) LBX_E(1, 0) LBX_E(1, 1) LBX_E(2, 0) LBX_E(2, 1) LBX_E(2, 2) LBX_E(3, 0)
            LBX_E(3, 1) LBX_E(3, 2) LBX_E(3, 3) LBX_E(4, 0) LBX_E(4, 1) LBX_E(4, 2) LBX_E(4, 3) LBX_E(4, 4)
#undef LBX_E
            throw Error(LBX_ERR_INVALID, "no one-electron kernel for this angular momentum pair");
        }

        VFn vFn(int la, int lb)
        {
#define LBX_E(A, B)                                                                                \
    if (la == A && lb == B)                                                                        \
        return &launchV<A, B>;
            LBX_E(0, 0) LBX_E(1, 0) LBX_E(1, 1) LBX_E(2, 0) LBX_E(2, 1) LBX_E(2, 2) LBX_E(3, 0)
            LBX_E(3, 1) LBX_E(3, 2) LBX_E(3, 3) LBX_E(4, 0) LBX_E(4, 1) LBX_E(4, 2) LBX_E(4, 3) LBX_E(4, 4)
#undef LBX_E
            throw Error(LBX_ERR_INVALID, "no one-electron kernel for this angular momentum pair");
        }

        void uploadSphTables()
        {
            static bool done = false;   // per process; the table is identical on every device
            static int done_dev = -1;
            int dev = 0;
            LBX_HIP(hipGetDevice(&dev));
            if (done && done_dev == dev)
                return;
            SphConst h{};
            int e = 0;
            for (int l = 0; l <= kMaxLAux; l++)
            {
                const auto &T = sphTable(l);
                int k = 0;
                for (int m = 0; m <= 2 * l + 1; m++)
                {
                    h.start[l][m] = e;
                    while (m < 2 * l + 1 && k < static_cast<int>(T.size()) && T[k].sph == m)
                    {
                        if (e >= kSphMax)
                            throw Error(LBX_ERR_INVALID, "spherical table overflow");
                        h.cart[e] = T[k].cart;
                        h.val[e] = T[k].val;
                        e++;
                        k++;
                    }
                }
            }
            LBX_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cSph), &h, sizeof h));
            done = true;
            done_dev = dev;
        }
    }

    // Uploads the pair class to the device and computes its E matrices.
    void computeEcoeffs(Context &c, PairClass &pc, const std::vector<HShell> &shells)
    {
        uploadSphTables();
        const int SA = nSph(pc.la), SB = nSph(pc.lb), SP = SA * SB, HP = nHerm(pc.la + pc.lb);
        std::vector<PrimIn> prim(pc.n_ppairs);
        std::vector<double> centers(6 * static_cast<size_t>(pc.n_pairs));
        std::vector<double> norms(static_cast<size_t>(pc.n_pairs) * (SA + SB));
        std::vector<double4> pp(pc.n_ppairs);
        for (int ip = 0; ip < pc.n_pairs; ip++)
        {
            const HShell &A = shells[pc.sh_a[ip]];
            const HShell *B = pc.single ? nullptr : &shells[pc.sh_b[ip]];
            for (int x = 0; x < 3; x++)
            {
                centers[6 * ip + x] = A.xyz[x];
                centers[6 * ip + 3 + x] = B ? B->xyz[x] : A.xyz[x];
            }
            for (int m = 0; m < SA; m++)
                norms[static_cast<size_t>(ip) * (SA + SB) + m] = A.norms[m];
            for (int m = 0; m < SB; m++)
                norms[static_cast<size_t>(ip) * (SA + SB) + SA + m] = B ? B->norms[m] : 1.0;
            for (int o = pc.poff[ip]; o < pc.poff[ip] + pc.pcnt[ip]; o++)
            {
                const int x = pc.prim_x[o], y = pc.prim_y[o];
                const double a = A.exps[x], b = B ? B->exps[y] : 0.0;
                const double da = A.pnorms[x] * A.coeffs[x];
                const double db = B ? B->pnorms[y] * B->coeffs[y] : 1.0;
                prim[o] = PrimIn{a, b, da * db, ip};
                const double p = a + b;
                double P[3];
                for (int k = 0; k < 3; k++)
                    P[k] = (a * A.xyz[k] + b * (B ? B->xyz[k] : A.xyz[k])) / p;
                pp[o] = double4{p, P[0], P[1], P[2]};
            }
        }
        pc.d_pp = c.upload(pp);
        pc.d_ao_a = c.upload(pc.ao_a);
        pc.d_ao_b = c.upload(pc.ao_b);
        pc.d_sh_a = c.upload(pc.sh_a);
        pc.d_sh_b = c.upload(pc.sh_b);
        pc.d_poff = c.upload(pc.poff);
        pc.d_pcnt = c.upload(pc.pcnt);
        std::vector<int> ident(pc.n_pairs);
        for (int i = 0; i < pc.n_pairs; i++)
            ident[i] = i;
        pc.order = ident;
        pc.d_order = c.upload(ident);
        pc.d_q = c.alloc<double>(pc.n_pairs);
        const size_t ne = static_cast<size_t>(pc.n_ppairs) * HP * SP;
        pc.d_eb = c.alloc<double>(ne);
        pc.d_ek = c.alloc<double>(ne);
        if (pc.n_ppairs == 0)
            return;
        pc.d_prim = c.upload(prim);
        pc.d_centers = c.upload(centers);
        pc.d_norms = c.upload(norms);
        ecoefFn(pc.la, pc.lb)(pc.n_ppairs, pc.d_prim, pc.d_centers, pc.d_norms, pc.d_eb, pc.d_ek,
                              c.stream);
        LBX_HIP(hipGetLastError());
    }

    void computeEcoeffsD1(Context &c, PairClass &pc)
    {
        if (pc.d_e1b || pc.n_pairs == 0)
            return;
        pc.d_atom_a = c.upload(pc.atom_a);
        pc.d_atom_b = c.upload(pc.atom_b);
        std::vector<PrimIn> prim(pc.n_ppairs);
        LBX_HIP(hipMemcpy(prim.data(), pc.d_prim, prim.size() * sizeof(PrimIn), hipMemcpyDeviceToHost));
        std::vector<double> ra(pc.n_ppairs);
        for (int o = 0; o < pc.n_ppairs; o++)
            ra[o] = prim[o].a / (prim[o].a + prim[o].b);
        pc.d_ra = c.upload(ra);
        const size_t ne = 3 * pc.eCount();
        pc.d_e1b = c.alloc<double>(ne);
        pc.d_e1k = c.alloc<double>(ne);
        ecoefD1Fn(pc.la, pc.lb)(pc.n_ppairs, pc.d_prim, pc.d_centers, pc.d_norms, pc.d_e1b, pc.d_e1k,
                                c.stream);
        LBX_HIP(hipGetLastError());
    }

    // Dense overlap, kinetic-energy and point-charge attraction matrices (any may be null),
    // accumulated into zeroed n x n device buffers on the context's stream.  charges: device,
    // (x, y, z, q) per charge.
    void computeOneElectron(Context &c, double *S, double *T, const double4 *charges, int ncharges, double *V)
    {
        const long long n = static_cast<long long>(c.nbf);
        for (const PairClass &pc : c.cls)
        {
            if (pc.n_ppairs == 0)
                continue;
            if (S || T)
                stFn(pc.la, pc.lb)(pc, n, S, T, c.stream);
            if (V && ncharges > 0)
                vFn(pc.la, pc.lb)(pc, charges, ncharges, c.d_boys, n, V, c.stream);
        }
        LBX_HIP(hipGetLastError());
    }

    // Nuclear gradient of tr(D (T + V)) - tr(W S) (W may be null), natoms x 3, and the
    // derivative with respect to each charge position (charge_grad, ncharges x 3, may be null);
    // grad and charge_grad are added into (zeroed by the caller).
    void computeOneElectronGrad(Context &c, const double *D, const double *W, const double4 *charges, int ncharges,
                                double *grad, double *charge_grad)
    {
        size_t ed_doubles = 0;
        for (PairClass &pc : c.cls)
            if (pc.n_ppairs > 0)
            {
                computeEcoeffsD1(c, pc);
                ed_doubles = std::max(ed_doubles, static_cast<size_t>(4) * nHerm(pc.la + pc.lb) * pc.n_ppairs);
            }
        double *ed = nullptr;
        if (ncharges > 0 && ed_doubles > 0)
            LBX_HIP(hipMalloc(&ed, ed_doubles * sizeof(double)));
        try
        {
            const long long n = static_cast<long long>(c.nbf);
            for (const PairClass &pc : c.cls)
                if (pc.n_ppairs > 0)
                    oneGradFn(pc.la, pc.lb)(pc, n, D, W, charges, ncharges, c.d_boys, ed, grad, charge_grad,
                                            c.stream);
            LBX_HIP(hipGetLastError());
            LBX_HIP(hipStreamSynchronize(c.stream));
        }
        catch (...)
        {
            (void)hipStreamSynchronize(c.stream);
            (void)hipFree(ed);
            throw;
        }
        if (ed)
            LBX_HIP(hipFree(ed));
    }
}
