// Test infrastructure: brute-force checks of the library's task-table arithmetic
// (lible_amd/csrc/tasks.hpp), built by tests/test_tasks_host.py with g++ and the address and
// undefined-behaviour sanitizers and run on the CPU.  Exit status 0 when every check passes.
//
// Checked on random sorted Schwarz-factor lists (with ties, zeros and one-sided cuts):
//   cut          every [plo, phi) equals the brute-force set {s : thresh <= Q_s Q_q < hi,
//                s >= q when same}, and flo / fhi are non-increasing;
//   blockMajor   per lane group, the first uniform position, the task count and the quartet
//                count equal a brute-force scan over all (q, s), for random permutations,
//                group sizes and sampling strides; checkBlockMajor accepts the table and
//                rejects planted corruptions (a group's range past the uniform side, a
//                repeated permutation entry, a block range past the lane side);
//   laneOrder    a permutation, Schwarz buckets in descending order, lanes of one (bucket,
//                first shell) adjacent, and bqlo / bqhi the exact per-block ranges.
#include "../../lible_amd/csrc/tasks.hpp"

#include <cstdio>
#include <random>

using namespace lbx::tasks;

namespace
{
    int g_fail = 0;

#define CHECK(cond, ...)                                                                           \
    do                                                                                             \
    {                                                                                              \
        if (!(cond))                                                                               \
        {                                                                                          \
            if (g_fail < 20)                                                                       \
            {                                                                                      \
                std::printf("FAIL %s:%d: ", __FILE__, __LINE__);                                   \
                std::printf(__VA_ARGS__);                                                          \
                std::printf("\n");                                                                 \
            }                                                                                      \
            g_fail++;                                                                              \
        }                                                                                          \
    } while (0)

    std::vector<double> sortedQ(std::mt19937_64 &rng, int n)
    {
        std::uniform_real_distribution<double> u(-16.0, 1.0);
        std::uniform_int_distribution<int> k(0, 9);
        std::vector<double> q(n);
        for (auto &x : q)
        {
            const int r = k(rng);
            x = r == 0 ? 0.0 : std::pow(10.0, u(rng));
        }
        // ties: copy some neighbours
        for (int i = 1; i < n; i++)
            if (k(rng) == 1)
                q[i] = q[i - 1];
        std::sort(q.begin(), q.end(), [](double a, double b) { return a > b; });
        return q;
    }

    bool inRange(double ql, double qu, double thr, double hi) { return ql * qu >= thr && ql * qu < hi; }

    void checkCase(std::mt19937_64 &rng, int nl, int nu, bool same, double thr, double hi)
    {
        std::vector<double> ql = sortedQ(rng, nl), qu = same ? ql : sortedQ(rng, nu);
        if (same)
            nu = nl;
        const Cuts c = cut(ql, qu, same, thr, hi);
        for (int q = 0; q < nu; q++)
        {
            if (q > 0)
                CHECK(c.flo[q] <= c.flo[q - 1] && c.fhi[q] <= c.fhi[q - 1], "cuts not monotone at q=%d", q);
            for (int s = 0; s < nl; s++)
            {
                const bool want = (thr <= 0 || ql[s] * qu[q] >= thr) && ql[s] * qu[q] < hi && (!same || s >= q);
                const bool got = s >= c.plo[q] && s < c.phi[q];
                CHECK(want == got, "cut membership q=%d s=%d want %d got %d (nl %d nu %d same %d)", q, s, want, got,
                      nl, nu, same);
            }
        }
        (void)inRange;
        // lane-block-major enumeration against a scan over all (q, s)
        std::uniform_int_distribution<int> gs(0, 3), st(1, 3), pm(0, 2);
        for (int rep = 0; rep < 3; rep++)
        {
            const int gshift = gs(rng), stride = rep == 0 ? 1 : st(rng);
            std::vector<int> perm;
            const int mode = pm(rng);
            if (mode >= 1)
            {
                // a lane order from random shells, or a random permutation
                if (mode == 1)
                {
                    std::uniform_int_distribution<int> sh(0, std::max(1, nl / 5));
                    std::vector<int> sa(nl), sb(nl);
                    for (int i = 0; i < nl; i++)
                    {
                        sa[i] = sh(rng);
                        sb[i] = sh(rng);
                    }
                    perm = laneOrder(ql, sa, sb).perm;
                }
                else
                {
                    perm.resize(nl);
                    for (int i = 0; i < nl; i++)
                        perm[i] = i;
                    std::shuffle(perm.begin(), perm.end(), rng);
                }
            }
            const BlockMajor b = blockMajor(c, nl, stride, gshift, perm);
            const int gsz = 64 << gshift, ng = (nl + gsz - 1) / gsz;
            CHECK(static_cast<int>(b.bq0.size()) == ng, "group count");
            double quartets = 0;
            long long tasks = 0;
            for (int g = 0; g < ng; g++)
            {
                int qmin = -1, qmax = -1;
                double qq = 0;
                for (int q = 0; q < nu; q++)
                {
                    int cnt = 0;
                    for (int p = g * gsz; p < std::min(nl, (g + 1) * gsz); p++)
                    {
                        const int s = perm.empty() ? p : perm[p];
                        cnt += s >= c.plo[q] && s < c.phi[q];
                    }
                    if (cnt > 0)
                    {
                        if (qmin < 0)
                            qmin = q;
                        qmax = q;
                        qq += cnt;
                    }
                }
                const long long n_g = b.bprefix[g + 1] - b.bprefix[g];
                CHECK(b.bprefix[g] == tasks, "prefix of group %d", g);
                if (qmin >= 0 && g % stride == 0)
                {
                    CHECK(b.bq0[g] == qmin, "group %d first q %d want %d", g, b.bq0[g], qmin);
                    CHECK(n_g == (static_cast<long long>(qmax - qmin + 1) << gshift), "group %d tasks", g);
                    quartets += qq;
                    tasks += static_cast<long long>(qmax - qmin + 1) << gshift;
                }
                else
                    CHECK(n_g == 0, "group %d should have no tasks", g);
            }
            CHECK(b.ntasks == tasks, "task total");
            CHECK(b.quartets == quartets, "quartet total %g want %g", b.quartets, quartets);
            // the bounds check accepts the table and rejects corrupted ones
            std::vector<int> lo, hi2;
            if (!perm.empty())
            {
                const int nblk = (nl + 63) / 64;
                lo.assign(nblk, 1 << 30);
                hi2.assign(nblk, 0);
                for (int k = 0; k < nl; k++)
                {
                    lo[k >> 6] = std::min(lo[k >> 6], perm[k]);
                    hi2[k >> 6] = std::max(hi2[k >> 6], perm[k] + 1);
                }
            }
            bool ok = true;
            try
            {
                checkBlockMajor(b, nu, nl, gshift, perm, lo, hi2);
            }
            catch (const std::exception &)
            {
                ok = false;
            }
            CHECK(ok, "valid table rejected");
            for (int g = 0; g < ng; g++)
                if (b.bprefix[g + 1] > b.bprefix[g])
                {
                    BlockMajor bad = b;
                    bad.bq0[g] = nu;   // the group's uniform range now runs past the side
                    bool threw = false;
                    try
                    {
                        checkBlockMajor(bad, nu, nl, gshift, perm, lo, hi2);
                    }
                    catch (const std::exception &)
                    {
                        threw = true;
                    }
                    CHECK(threw, "corrupted group range accepted");
                    break;
                }
            if (perm.size() >= 2)
            {
                std::vector<int> badp = perm;
                badp[1] = badp[0];
                bool threw = false;
                try
                {
                    checkBlockMajor(b, nu, nl, gshift, badp, lo, hi2);
                }
                catch (const std::exception &)
                {
                    threw = true;
                }
                CHECK(threw, "repeated permutation entry accepted");
                std::vector<int> badhi = hi2;
                badhi.back() = nl + 1;
                threw = false;
                try
                {
                    checkBlockMajor(b, nu, nl, gshift, perm, lo, badhi);
                }
                catch (const std::exception &)
                {
                    threw = true;
                }
                CHECK(threw, "block range past the lane side accepted");
            }
        }
    }

    void checkLaneOrder(std::mt19937_64 &rng, int n)
    {
        std::vector<double> q = sortedQ(rng, n);
        std::uniform_int_distribution<int> sh(0, std::max(1, n / 7));
        std::vector<int> sa(n), sb(n);
        for (int i = 0; i < n; i++)
        {
            sa[i] = sh(rng);
            sb[i] = sh(rng);
        }
        const LaneOrder r = laneOrder(q, sa, sb);
        std::vector<char> seen(n, 0);
        for (int p : r.perm)
        {
            CHECK(p >= 0 && p < n && !seen[p], "lane order is not a permutation");
            if (p >= 0 && p < n)
                seen[p] = 1;
        }
        auto bucket = [&](int s) { return q[s] > 0 ? static_cast<int>(std::floor(2.0 * std::log10(q[s]))) : -1000000; };
        for (int p = 1; p < n; p++)
        {
            const int x = r.perm[p - 1], y = r.perm[p];
            CHECK(bucket(x) >= bucket(y), "buckets out of order at %d", p);
            if (bucket(x) == bucket(y))
                CHECK(sa[x] <= sa[y], "first shells out of order inside a bucket at %d", p);
        }
        const int nblk = (n + 63) / 64;
        for (int k = 0; k < nblk; k++)
        {
            int lo = 1 << 30, hi = 0;
            for (int p = 64 * k; p < std::min(n, 64 * k + 64); p++)
            {
                lo = std::min(lo, r.perm[p]);
                hi = std::max(hi, r.perm[p] + 1);
            }
            CHECK(r.bqlo[k] == lo && r.bqhi[k] == hi, "block %d range", k);
        }
    }
}

int main()
{
    std::mt19937_64 rng(20261203);
    const double inf = kNoUpper;
    const int sizes[][2] = {{1, 1}, {3, 7}, {64, 64}, {65, 130}, {200, 37}, {517, 300}, {1100, 90}};
    for (auto &sz : sizes)
        for (int same = 0; same < 2; same++)
        {
            checkCase(rng, sz[0], sz[1], same, 0.0, inf);
            checkCase(rng, sz[0], sz[1], same, 1e-12, inf);
            checkCase(rng, sz[0], sz[1], same, 1e-12, 1e-5);
            checkCase(rng, sz[0], sz[1], same, 1e-5, inf);
            checkCase(rng, sz[0], sz[1], same, 1e-8, 1e-3);
        }
    for (int n : {1, 5, 63, 64, 65, 700, 2000})
        checkLaneOrder(rng, n);
    if (g_fail)
    {
        std::printf("%d checks failed\n", g_fail);
        return 1;
    }
    std::printf("task tables: all checks passed\n");
    return 0;
}
