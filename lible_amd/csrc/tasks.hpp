// Host-side task-table arithmetic of the quartet kernels: Schwarz cuts over sorted pair lists,
// the lane-block-major enumeration of the one-lane kernels and the shell-ordered lane side.
// Plain C++ on std::vector (no HIP), so the same code is built into the library and into the
// CPU tests, which check it against brute-force enumerations under the address and
// undefined-behaviour sanitizers.
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

namespace lbx
{
    namespace tasks
    {
        constexpr double kNoUpper = std::numeric_limits<double>::infinity();

        // first position p with qs[p] * qk < thr (qs sorted descending)
        inline int firstBelow(const std::vector<double> &qs, double qk, double thr)
        {
            int a = 0, b = static_cast<int>(qs.size());
            while (a < b)
            {
                const int m = a + (b - a) / 2;
                if (qs[m] * qk >= thr)
                    a = m + 1;
                else
                    b = m;
            }
            return a;
        }

        // Lane-side ranges [plo[q], phi[q]) of every uniform position q for the quartets with
        // thresh <= Q_p Q_q < thresh_hi, both lists sorted by Q descending (thresh <= 0: no lower
        // cut); `same`: only p >= q (one class on both sides, unique pairs).  flo[q] and fhi[q]
        // are the cuts themselves (first p below thresh, first p below thresh_hi), both
        // non-increasing in q, so that a lane position s is in q's range exactly when
        // fhi[q] <= s < flo[q] (and q <= s when `same`).
        struct Cuts
        {
            std::vector<int> plo, phi, flo, fhi;
            bool same = false;
        };

        inline Cuts cut(const std::vector<double> &qlane, const std::vector<double> &qunif, bool same,
                        double thresh, double thresh_hi)
        {
            const int nl = static_cast<int>(qlane.size()), nu = static_cast<int>(qunif.size());
            Cuts c;
            c.same = same;
            c.plo.resize(nu);
            c.phi.resize(nu);
            c.flo.resize(nu);
            c.fhi.resize(nu);
            for (int q = 0; q < nu; q++)
            {
                const double qk = qunif[q];
                c.flo[q] = thresh > 0 ? firstBelow(qlane, qk, thresh) : nl;
                c.fhi[q] = thresh_hi < kNoUpper ? firstBelow(qlane, qk, thresh_hi) : 0;
                const int lo = std::max(same ? q : 0, c.fhi[q]);
                c.plo[q] = lo;
                c.phi[q] = std::max(c.flo[q], lo);
            }
            return c;
        }

        // Lane-block-major enumeration (one-lane kernels): lane block b = lane positions
        // 64 b .. 64 b + 63, group g = blocks g G .. g G + G - 1 (G = 2^gshift).  Group g takes the
        // uniform positions bq0[g] .. bq0[g] + n_g - 1, the smallest range holding every q whose
        // range meets one of the group's lanes, each q with the group's G blocks in turn: G n_g
        // tasks from bprefix[g] on.  Only groups with g % stride == 0 get tasks (the autotuner's
        // samples).  perm maps lane position -> Schwarz-order position (identity when empty).
        struct BlockMajor
        {
            std::vector<long long> bprefix;
            std::vector<int> bq0;
            long long ntasks = 0;
            double quartets = 0;   // quartets the enumerated groups hold
            int nblk = 0;
        };

        inline BlockMajor blockMajor(const Cuts &c, int nlane, int stride, int gshift,
                                     const std::vector<int> &perm)
        {
            const int nu = static_cast<int>(c.plo.size());
            const int gsz = 64 << gshift;
            const int ng = (nlane + gsz - 1) / gsz;
            if (!perm.empty() && static_cast<int>(perm.size()) != nlane)
                throw std::runtime_error("blockMajor: permutation size differs from the lane count");
            std::vector<int> qmin(ng, std::numeric_limits<int>::max()), qmax(ng, -1);
            std::vector<double> quart(ng, 0.0);
            for (int p = 0; p < nlane; p++)
            {
                const int s = perm.empty() ? p : perm[p];
                // q with fhi[q] <= s: a suffix [a, nu); q with s < flo[q]: a prefix [0, b)
                int a = 0, b = nu;
                {
                    int x = 0, y = nu;
                    while (x < y)
                    {
                        const int m = x + (y - x) / 2;
                        if (c.fhi[m] <= s)
                            y = m;
                        else
                            x = m + 1;
                    }
                    a = x;
                }
                {
                    int x = 0, y = nu;
                    while (x < y)
                    {
                        const int m = x + (y - x) / 2;
                        if (c.flo[m] <= s)
                            y = m;
                        else
                            x = m + 1;
                    }
                    b = x;
                }
                if (c.same)
                    b = std::min(b, s + 1);
                if (a >= b)
                    continue;
                const int g = p / gsz;
                qmin[g] = std::min(qmin[g], a);
                qmax[g] = std::max(qmax[g], b - 1);
                quart[g] += b - a;
            }
            BlockMajor r;
            r.nblk = (nlane + 63) / 64;
            r.bprefix.assign(ng + 1, 0);
            r.bq0.assign(ng, 0);
            long long acc = 0;
            for (int g = 0; g < ng; g++)
            {
                r.bprefix[g] = acc;
                if (qmax[g] >= 0 && g % stride == 0)
                {
                    r.bq0[g] = qmin[g];
                    acc += static_cast<long long>(qmax[g] - qmin[g] + 1) << gshift;
                    r.quartets += quart[g];
                }
            }
            r.bprefix[ng] = acc;
            r.ntasks = acc;
            return r;
        }

        // Checks what the one-lane kernels read through a lane-block-major table: every group's
        // uniform range inside [0, nu), task counts consistent with the prefix, and (with a
        // permutation) perm a permutation of 0..nlane-1 and each block's Schwarz-position range
        // [bqlo, bqhi) inside [0, nlane].  Throws on the first violation.
        inline void checkBlockMajor(const BlockMajor &b, int nu, int nlane, int gshift,
                                    const std::vector<int> &perm, const std::vector<int> &bqlo,
                                    const std::vector<int> &bqhi)
        {
            const int gsz = 64 << gshift;
            const int ng = (nlane + gsz - 1) / gsz;
            auto fail = [](const std::string &m) { throw std::runtime_error("task table check: " + m); };
            if (static_cast<int>(b.bprefix.size()) != ng + 1 || static_cast<int>(b.bq0.size()) != ng)
                fail("group array sizes");
            if (b.nblk != (nlane + 63) / 64)
                fail("block count");
            for (int g = 0; g < ng; g++)
            {
                const long long cnt = b.bprefix[g + 1] - b.bprefix[g];
                if (cnt < 0 || (cnt & ((1LL << gshift) - 1)) != 0)
                    fail("group task count");
                if (cnt > 0 && (b.bq0[g] < 0 || b.bq0[g] + (cnt >> gshift) > nu))
                    fail("group uniform range outside the uniform side");
            }
            if (b.bprefix[ng] != b.ntasks)
                fail("task total");
            if (!perm.empty())
            {
                if (static_cast<int>(perm.size()) != nlane)
                    fail("permutation size");
                std::vector<char> seen(nlane, 0);
                for (int p : perm)
                {
                    if (p < 0 || p >= nlane || seen[p])
                        fail("not a permutation");
                    seen[p] = 1;
                }
                const int nblk = (nlane + 63) / 64;
                if (static_cast<int>(bqlo.size()) != nblk || static_cast<int>(bqhi.size()) != nblk)
                    fail("block range sizes");
                for (int k = 0; k < nblk; k++)
                    if (bqlo[k] < 0 || bqhi[k] > nlane || bqlo[k] >= bqhi[k])
                        fail("block Schwarz-position range");
            }
        }

        // Shell-ordered lane side: Schwarz-order positions regrouped by the pair's first shell
        // (then second) within half-decade buckets of the Schwarz factor, so lanes sharing a
        // first shell are adjacent; qsorted[p] is the factor at Schwarz position p, sha/shb the
        // shells of the pair at that position.  Returns perm (lane position -> Schwarz
        // position) and per 64-lane block the range [bqlo, bqhi) of positions it holds.
        struct LaneOrder
        {
            std::vector<int> perm, bqlo, bqhi;
        };

        inline LaneOrder laneOrder(const std::vector<double> &qsorted, const std::vector<int> &sha,
                                   const std::vector<int> &shb)
        {
            const int n = static_cast<int>(qsorted.size());
            std::vector<int> bk(n);
            LaneOrder r;
            r.perm.resize(n);
            for (int p = 0; p < n; p++)
            {
                const double q = qsorted[p];
                bk[p] = q > 0 ? static_cast<int>(std::floor(2.0 * std::log10(q))) : -1000000;
                r.perm[p] = p;
            }
            std::stable_sort(r.perm.begin(), r.perm.end(), [&](int x, int y) {
                if (bk[x] != bk[y])
                    return bk[x] > bk[y];
                if (sha[x] != sha[y])
                    return sha[x] < sha[y];
                return shb[x] < shb[y];
            });
            const int nblk = (n + 63) / 64;
            r.bqlo.assign(nblk, std::numeric_limits<int>::max());
            r.bqhi.assign(nblk, 0);
            for (int p = 0; p < n; p++)
            {
                r.bqlo[p >> 6] = std::min(r.bqlo[p >> 6], r.perm[p]);
                r.bqhi[p >> 6] = std::max(r.bqhi[p >> 6], r.perm[p] + 1);
            }
            return r;
        }
    }
}
