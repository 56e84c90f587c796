// The C-ABI of lbx (include/lbx.h): context set-up, Schwarz screening, task tables and the
// drivers that launch the quartet kernels class pair by class pair.
#include "grad.hpp"
#include "quartet.hpp"
#include "registry.hpp"
#include "tasks.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <tuple>
#include <unordered_map>

// Staging area of the allocation-free block entry (lbx_eri4_batch_static): one call's task list,
// four ints per entry, at most kStaticEntries entries (16 MiB).  Each device holds its own copy
// for the life of the process.  Host code never passes this name to a kernel or a copy: it uses
// the address hipGetSymbolAddress returned for the device (scratchAddress below).
constexpr long long kStaticEntries = 1LL << 20;
static __device__ __align__(256) int scratch[4 * kStaticEntries];

namespace lbx
{
    namespace
    {
        thread_local std::string g_err;

        // scratch device buffer owned by one call
        struct DevBuf
        {
            double *p = nullptr;
            explicit DevBuf(size_t n) { LBX_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(double))); }
            ~DevBuf()
            {
                if (p)
                    (void)hipFree(p);
            }
            DevBuf(const DevBuf &) = delete;
            DevBuf &operator=(const DevBuf &) = delete;
        };

        template <class F>
        int guarded(F &&f)
        {
            try
            {
                f();
                return LBX_OK;
            }
            catch (const Error &e)
            {
                g_err = e.what();
                return e.code;
            }
            catch (const std::exception &e)
            {
                g_err = e.what();
                return LBX_ERR_INVALID;
            }
        }

        __global__ void fillKernel(double *p, long long n, double v)
        {
            for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
                 i += static_cast<long long>(gridDim.x) * blockDim.x)
                p[i] = v;
        }

        // M <- M + M^T for an n x n row-major matrix
        __global__ void symmetrizeKernel(double *M, int n)
        {
            const long long tot = static_cast<long long>(n) * n;
            for (long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; e < tot;
                 e += static_cast<long long>(gridDim.x) * blockDim.x)
            {
                const int i = static_cast<int>(e / n), j = static_cast<int>(e % n);
                if (j < i)
                    continue;
                const double v = M[static_cast<long long>(i) * n + j] + M[static_cast<long long>(j) * n + i];
                M[static_cast<long long>(i) * n + j] = v;
                M[static_cast<long long>(j) * n + i] = v;
            }
        }

        void fill(double *p, long long n, double v, hipStream_t s)
        {
            if (n <= 0)
                return;
            long long b = (n + 255) / 256;
            hipLaunchKernelGGL(fillKernel, dim3(static_cast<int>(std::min(b, 65536LL))), dim3(256), 0, s,
                               p, n, v);
            LBX_HIP(hipGetLastError());
        }

        // Device address of the scratch array on device dev, looked up once per device with that
        // device current (context set-up) and kept for every later context on it
        int *scratchAddress(int dev)
        {
            static std::mutex m;
            static std::map<int, int *> addr;
            std::lock_guard<std::mutex> lk(m);
            auto it = addr.find(dev);
            if (it != addr.end())
                return it->second;
            void *p = nullptr;
            LBX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            if (!p)
                throw Error(LBX_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
            addr[dev] = static_cast<int *>(p);
            return static_cast<int *>(p);
        }

        // Boys table, one copy per device for the life of the process
        const double *boysTableDevice()
        {
            static std::mutex m;
            static std::map<int, double *> tabs;
            int dev = 0;
            LBX_HIP(hipGetDevice(&dev));
            std::lock_guard<std::mutex> lk(m);
            auto it = tabs.find(dev);
            if (it != tabs.end())
                return it->second;
            auto h = boysTable();
            double *d = nullptr;
            LBX_HIP(hipMalloc(&d, h.size() * sizeof(double)));
            LBX_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
            tabs[dev] = d;
            return d;
        }

        template <int N>
        __global__ void boysKernel(long long cnt, const double *x, double *out, const double *tab)
        {
            for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < cnt;
                 i += static_cast<long long>(gridDim.x) * blockDim.x)
            {
                double F[N + 1];
                boysF<N>(x[i], tab, F);
                for (int k = 0; k <= N; k++)
                    out[i * (N + 1) + k] = F[k];
            }
        }

        template <int N>
        void launchBoys(long long cnt, const double *x, double *out, const double *tab, hipStream_t s)
        {
            long long b = (cnt + 255) / 256;
            hipLaunchKernelGGL((boysKernel<N>), dim3(static_cast<int>(std::min(b, 65536LL))), dim3(256),
                               0, s, cnt, x, out, tab);
        }

        using BoysFn = void (*)(long long, const double *, double *, const double *, hipStream_t);
        template <int... Ns>
        constexpr std::array<BoysFn, sizeof...(Ns)> boysFns(std::integer_sequence<int, Ns...>)
        {
            return {&launchBoys<Ns>...};
        }

        // ------------------------------------------------------------------ task tables
        struct TaskTable
        {
            long long ntasks = 0;
            int nket = 0;
            std::vector<long long> prefix;
            std::vector<int> plo, phi;
            tasks::Cuts cuts;   // the Schwarz cuts behind plo/phi (buildTasks)
            long long *d_prefix = nullptr;
            int *d_plo = nullptr, *d_phi = nullptr;
            double quartets = 0, flops = 0;
            // over the quartets: sum of K_bra K_ket, of K_bra, of K_ket (canonical roles)
            double sum_kk = 0, sum_kbra = 0, sum_kket = 0;
            // lane-block-major enumeration of a 64-wide table (empty unless built)
            int nblk = 0, gshift = 0;
            long long btasks = 0;
            std::vector<long long> bprefix;
            std::vector<int> bq0;
            long long *d_bprefix = nullptr;
            int *d_bq0 = nullptr;
        };

        enum class Pairing
        {
            Unique,     // p >= q when bra and ket are the same class, all p otherwise
            All,        // every (p, q)
            Diagonal    // p == q only
        };

        using tasks::firstBelow;

        constexpr double kNoUpper = std::numeric_limits<double>::infinity();
        // quartets per class pair and variant that the autotuner times
        constexpr long long kTuneSample = 2LL << 20;

        // Lane-side positions [plo, phi) per uniform-side position q, with the Schwarz cut
        // thresh <= Q_p Q_q < thresh_hi on descending-sorted Q (thresh <= 0: no lower cut).
        // `bra` is the lane side and `ket` the uniform side; lane_is_bra says whether that
        // matches the canonical (bra class >= ket class) roles, which the FLOP count follows.
        TaskTable buildTasks(Context &c, const PairClass &bra, const PairClass &ket, Pairing mode,
                             bool same, double thresh, int width, bool lane_is_bra = true,
                             double thresh_hi = kNoUpper)
        {
            TaskTable t;
            t.nket = ket.n_pairs;
            t.plo.resize(t.nket);
            t.phi.resize(t.nket);
            t.prefix.resize(t.nket + 1);
            std::vector<double> qb(bra.n_pairs), kab_prefix(bra.n_pairs + 1, 0.0);
            for (int i = 0; i < bra.n_pairs; i++)
            {
                qb[i] = bra.Q.empty() ? 1.0 : bra.Q[bra.order[i]];
                kab_prefix[i + 1] = kab_prefix[i] + bra.pcnt[bra.order[i]];
            }
            const PairClass &cb = lane_is_bra ? bra : ket, &ck = lane_is_bra ? ket : bra;
            const int la = cb.la, lb = cb.lb, lc = ck.la, ld = ck.lb;
            const double Hab = nHerm(la + lb), Hcd = nHerm(lc + ld);
            const double Sab = nSph(la) * nSph(lb), Scd = nSph(lc) * nSph(ld);
            const double perKK = quartetFlops(la, lb, lc, ld, 1, 1) - 2 * Sab * Hab * Scd - 12 * Sab * Scd;
            if (mode != Pairing::Diagonal)
            {
                std::vector<double> qk(t.nket);
                for (int q = 0; q < t.nket; q++)
                    qk[q] = ket.Q.empty() ? 1.0 : ket.Q[ket.order[q]];
                t.cuts = tasks::cut(qb, qk, mode == Pairing::Unique && same, thresh, thresh_hi);
            }
            long long acc = 0;
            for (int q = 0; q < t.nket; q++)
            {
                int lo, hi;
                if (mode == Pairing::Diagonal)
                {
                    lo = q;
                    hi = q + 1;
                }
                else
                {
                    lo = t.cuts.plo[q];
                    hi = t.cuts.phi[q];
                }
                t.plo[q] = lo;
                t.phi[q] = hi;
                t.prefix[q] = acc;
                const long long cnt = hi - lo;
                acc += (cnt + width - 1) / width;
                if (cnt > 0)
                {
                    const double kq = ket.pcnt[ket.order[q]];
                    const double skl = kab_prefix[hi] - kab_prefix[lo];
                    t.quartets += cnt;
                    t.sum_kk += kq * skl;
                    t.sum_kbra += lane_is_bra ? skl : kq * cnt;
                    t.sum_kket += lane_is_bra ? kq * cnt : skl;
                    if (lane_is_bra)
                        t.flops += perKK * kq * skl + 2 * Sab * Hab * Scd * skl + 12 * Sab * Scd * cnt;
                    else
                        t.flops += perKK * kq * skl + 2 * Sab * Hab * Scd * kq * cnt + 12 * Sab * Scd * cnt;
                }
            }
            t.prefix[t.nket] = acc;
            t.ntasks = acc;
            t.d_prefix = c.upload(t.prefix);
            t.d_plo = c.upload(t.plo);
            t.d_phi = c.upload(t.phi);
            return t;
        }

        // Lane-block-major enumeration of a 64-wide table.  Lane block b = positions 64b .. 64b+63,
        // group g = blocks g G .. g G + G - 1 (G = 2^gshift).  Group g takes the uniform positions
        // q = bq0[g] .. bq0[g] + n_g - 1, the smallest range holding every q whose [plo, phi)
        // meets the group, each q with the group's G blocks in turn: G n_g tasks (a task whose
        // block misses q's range has every lane idle and returns at once).  Only groups with
        // g % stride == 0 get tasks (the autotuner's samples); returns the quartets they hold.
        // With a shell-ordered lane side (perm: lane position -> Schwarz-order position) a group
        // holds a subset of the Schwarz-order positions [lowest, highest]: its q range is then
        // conservative (the kernel tests each lane) and its quartets are counted exactly.
        double buildBlockMajor(Context &c, TaskTable &t, int nlane, int stride, int gshift,
                               const std::vector<int> &perm, const std::vector<int> &bqlo,
                               const std::vector<int> &bqhi)
        {
            if (t.cuts.plo.size() != static_cast<size_t>(t.nket))
                throw Error(LBX_ERR_STATE, "lane-block-major table without Schwarz cuts");
            tasks::BlockMajor b = tasks::blockMajor(t.cuts, nlane, stride, gshift, perm);
            try
            {
                tasks::checkBlockMajor(b, t.nket, nlane, gshift, perm, bqlo, bqhi);
            }
            catch (const std::runtime_error &e)
            {
                throw Error(LBX_ERR_STATE, e.what());
            }
            t.nblk = b.nblk;
            t.gshift = gshift;
            t.bprefix = std::move(b.bprefix);
            t.bq0 = std::move(b.bq0);
            t.btasks = b.ntasks;
            t.d_bprefix = c.upload(t.bprefix);
            t.d_bq0 = c.upload(t.bq0);
            return b.quartets;
        }

        // Tasks of the batched RI sums: for every uniform-side position q the lane-side pairs of
        // q's molecule (each molecule's positions are contiguous), with `same` only lane
        // positions p >= q, and with bounds (both sides sorted by Q within each molecule) only
        // those with lo <= Q_p Q_q < hi.  No screening below: every triplet of the tensors.
        TaskTable buildTasksMol(Context &c, const PairClass &lane, const PairClass &unif, bool same,
                                int width, double lo_bound = 0.0, double hi_bound = kNoUpper)
        {
            const bool bounded = lo_bound > 0.0 || hi_bound < kNoUpper;
            auto qOf = [](const PairClass &pc, int pos) { return pc.Q.empty() ? 1.0 : pc.Q[pc.order[pos]]; };
            // first position in [a, b) whose bound with qk falls below thr (Q descending there)
            auto firstBelowIn = [&](int a, int b, double qk, double thr) {
                while (a < b)
                {
                    const int m = (a + b) / 2;
                    if (qOf(lane, m) * qk >= thr)
                        a = m + 1;
                    else
                        b = m;
                }
                return a;
            };
            TaskTable t;
            t.nket = unif.n_pairs;
            t.plo.resize(t.nket);
            t.phi.resize(t.nket);
            t.prefix.resize(t.nket + 1);
            std::vector<int> lo(c.nmol, 0), hi(c.nmol, 0);
            std::vector<double> kl_prefix(lane.n_pairs + 1, 0.0);
            for (int p = lane.n_pairs; p-- > 0;)
                lo[lane.mol[lane.order[p]]] = p;
            for (int p = 0; p < lane.n_pairs; p++)
            {
                hi[lane.mol[lane.order[p]]] = p + 1;
                kl_prefix[p + 1] = kl_prefix[p] + lane.pcnt[lane.order[p]];
            }
            const int la = lane.la, lb = lane.lb, lc = unif.la, ld = unif.lb;
            const double Hab = nHerm(la + lb), Sab = nSph(la) * nSph(lb), Scd = nSph(lc) * nSph(ld);
            const double perKK = quartetFlops(la, lb, lc, ld, 1, 1) - 2 * Sab * Hab * Scd - 12 * Sab * Scd;
            long long acc = 0;
            for (int q = 0; q < t.nket; q++)
            {
                const int m = unif.mol[unif.order[q]];
                int a = hi[m] > 0 ? lo[m] : 0, b = hi[m];
                if (bounded && b > a)
                {
                    const double qk = qOf(unif, q);
                    const int a0 = a, b0 = b;
                    if (hi_bound < kNoUpper)
                        a = firstBelowIn(a0, b0, qk, hi_bound);
                    if (lo_bound > 0.0)
                        b = firstBelowIn(a0, b0, qk, lo_bound);
                }
                if (same)
                    a = std::max(a, q);
                b = std::max(a, b);
                t.plo[q] = a;
                t.phi[q] = b;
                t.prefix[q] = acc;
                const long long cnt = b - a;
                acc += (cnt + width - 1) / width;
                if (cnt > 0)
                {
                    const double kq = unif.pcnt[unif.order[q]];
                    const double skl = kl_prefix[b] - kl_prefix[a];
                    t.quartets += cnt;
                    t.sum_kk += kq * skl;
                    t.sum_kbra += skl;
                    t.sum_kket += kq * cnt;
                    // transforms, R and Boys as for J/K; the epilogue sums 2 values per element
                    t.flops += perKK * kq * skl + 2 * Sab * Hab * Scd * skl + 2 * Sab * Scd * cnt;
                }
            }
            t.prefix[t.nket] = acc;
            t.ntasks = acc;
            t.d_prefix = c.upload(t.prefix);
            t.d_plo = c.upload(t.plo);
            t.d_phi = c.upload(t.phi);
            return t;
        }

        void requireSingle(const Context &c, const char *who)
        {
            if (!c.atom_mol.empty())
                throw Error(LBX_ERR_STATE, std::string(who) + ": not available on a batched context "
                                                              "(pairs only exist within each molecule)");
        }

        KArgs baseArgs(Context &c, const PairClass &bra, const PairClass &ket, const TaskTable &t,
                       bool same)
        {
            KArgs a{};
            a.bra = bra.view();
            a.ket = ket.view();
            a.prefix = t.d_prefix;
            a.plo = t.d_plo;
            a.phi = t.d_phi;
            a.nket = t.nket;
            a.ntasks = t.ntasks;
            a.same_class = same ? 1 : 0;
            a.boys = c.d_boys;
            a.nbf = static_cast<long long>(c.nbf);
            a.naux = static_cast<long long>(c.nbf_aux);
            return a;
        }

        int rFlopsHost(int L)
        {
            int r = 0;
            for (int n = 0; n < L; n++)
                for (int m = 1; m <= L - n; m++)
                    r += (m + 1) * (m + 2) / 2;
            return 2 * r;
        }

        // Work estimate of a one-lane-per-quartet variant per quartet, from the primitive-pair
        // counts (the lane side's as its waves see them, `divergent`; else its mean): ket
        // transform and Boys + R per primitive quartet (R once per column pass), bra transform per
        // lane-side primitive pair.
        double valuCost(const PairClass &lane, const PairClass &unif, int nc, bool divergent = false)
        {
            const double kmean = lane.n_pairs ? double(lane.n_ppairs) / lane.n_pairs : 1;
            const double kl = divergent && lane.klane > 0 ? lane.klane : kmean;
            const double ku = unif.n_pairs ? double(unif.n_ppairs) / unif.n_pairs : 1;
            const double hl = nHerm(lane.la + lane.lb), hu = nHerm(unif.la + unif.lb);
            const double sl = nSph(lane.la) * nSph(lane.lb), su = nSph(unif.la) * nSph(unif.lb);
            const int L = lane.la + lane.lb + unif.la + unif.lb;
            const double passes = std::ceil(su / nc);
            return kl * ku * (hl * hu * su + passes * (60 + rFlopsHost(L))) + kl * sl * hl * su;
        }

        // A one-lane form whose waves wait on their longest lane for more than twice the work the
        // mean contraction would give (long, uneven contractions on the lane side, as in the
        // quadruple-zeta bases): for J/K the block form takes the class pair by default, whose
        // workgroup runs one bra pair's primitives for a block of ket pairs (on def2-QZVP water
        // clusters 7-25x faster there than the one-lane forms; on the evenly contracted bases
        // the one-lane forms stay the default).
        bool laneDivergent(const PairClass &lane, const PairClass &unif, int nc)
        {
            return valuCost(lane, unif, nc, true) > 2.0 * valuCost(lane, unif, nc);
        }

        // Default variant before tuning: the cheaper one-lane-per-quartet form by the work
        // estimate (with the lane side's longest-lane contraction); MFMA when no such form fits the
        // registers; for J/K the block form when the cheaper one-lane form diverges.
        int defaultVariant(int mode, const PairClass &bra, const PairClass &ket)
        {
            const KernelSlot &vb = kernelSlot(mode, bra.la, bra.lb, ket.la, ket.lb, VAR_VALU_BRA);
            const KernelSlot &vk = kernelSlot(mode, bra.la, bra.lb, ket.la, ket.lb, VAR_VALU_KET);
            double cb = vb.fn ? valuCost(bra, ket, vb.nc, true) : 1e300;
            double ck = vk.fn ? valuCost(ket, bra, vk.nc, true) : 1e300;
            if (!vb.fn && !vk.fn)
                return VAR_MFMA;
            const bool bra_lanes = cb <= ck;
            if (mode == MODE_JK && kernelSlot(mode, bra.la, bra.lb, ket.la, ket.lb, VAR_BLOCK).fn &&
                (bra_lanes ? laneDivergent(bra, ket, vb.nc) : laneDivergent(ket, bra, vk.nc)))
                return VAR_BLOCK;
            return bra_lanes ? VAR_VALU_BRA : VAR_VALU_KET;
        }

        // Default for the FP32 tier of a mixed-precision J/K: the cheaper FP32 one-lane form by
        // the same estimate (the FP32 block form when it diverges), else the FP32 MFMA form, or
        // -1 when the class has neither (its quartets then stay FP64).
        int defaultVariant32(const PairClass &bra, const PairClass &ket)
        {
            const KernelSlot &vb = kernelSlot(MODE_JK, bra.la, bra.lb, ket.la, ket.lb, VAR_VALU_BRA32);
            const KernelSlot &vk = kernelSlot(MODE_JK, bra.la, bra.lb, ket.la, ket.lb, VAR_VALU_KET32);
            if (!vb.fn && !vk.fn)
                return kernelSlot(MODE_JK, bra.la, bra.lb, ket.la, ket.lb, VAR_MFMA32).fn ? VAR_MFMA32 : -1;
            double cb = vb.fn ? valuCost(bra, ket, vb.nc, true) : 1e300;
            double ck = vk.fn ? valuCost(ket, bra, vk.nc, true) : 1e300;
            const bool bra_lanes = cb <= ck;
            if (kernelSlot(MODE_JK, bra.la, bra.lb, ket.la, ket.lb, VAR_BLOCK32).fn &&
                (bra_lanes ? laneDivergent(bra, ket, vb.nc) : laneDivergent(ket, bra, vk.nc)))
                return VAR_BLOCK32;
            return bra_lanes ? VAR_VALU_BRA32 : VAR_VALU_KET32;
        }

        __global__ void toFloatKernel(const double *__restrict__ in, float *__restrict__ out, long long n)
        {
            for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
                 i += static_cast<long long>(gridDim.x) * blockDim.x)
                out[i] = static_cast<float>(in[i]);
        }

        // FP32 copies of the E matrices, made once per context on first mixed-precision use
        void ensureF32(Context &c)
        {
            for (auto &pc : c.cls)
            {
                if (pc.d_eb32 || pc.n_ppairs == 0)
                    continue;
                const long long n = static_cast<long long>(pc.eCount());
                pc.d_eb32 = c.alloc<float>(n);
                pc.d_ek32 = c.alloc<float>(n);
                const int grid = static_cast<int>(std::min<long long>((n + 255) / 256, 65536));
                hipLaunchKernelGGL(toFloatKernel, dim3(grid), dim3(256), 0, c.stream, pc.d_eb, pc.d_eb32, n);
                hipLaunchKernelGGL(toFloatKernel, dim3(grid), dim3(256), 0, c.stream, pc.d_ek, pc.d_ek32, n);
                LBX_HIP(hipGetLastError());
            }
        }

        // FP32 copies of the auxiliary shells' E matrices (mixed-precision RI sums)
        void ensureF32Aux(Context &c)
        {
            for (auto &pc : c.aux)
            {
                if (pc.d_eb32 || pc.n_ppairs == 0)
                    continue;
                const long long n = static_cast<long long>(pc.eCount());
                pc.d_eb32 = c.alloc<float>(n);
                pc.d_ek32 = c.alloc<float>(n);
                const int grid = static_cast<int>(std::min<long long>((n + 255) / 256, 65536));
                hipLaunchKernelGGL(toFloatKernel, dim3(grid), dim3(256), 0, c.stream, pc.d_eb, pc.d_eb32, n);
                hipLaunchKernelGGL(toFloatKernel, dim3(grid), dim3(256), 0, c.stream, pc.d_ek, pc.d_ek32, n);
                LBX_HIP(hipGetLastError());
            }
        }

        // FP32 copies of the derivative E matrices (mixed-precision gradient)
        void ensureF32D1(Context &c)
        {
            for (auto &pc : c.cls)
            {
                if (pc.d_e1b32 || pc.n_ppairs == 0 || !pc.d_e1b)
                    continue;
                const long long n = 3 * static_cast<long long>(pc.eCount());
                pc.d_e1b32 = c.alloc<float>(n);
                pc.d_e1k32 = c.alloc<float>(n);
                const int grid = static_cast<int>(std::min<long long>((n + 255) / 256, 65536));
                hipLaunchKernelGGL(toFloatKernel, dim3(grid), dim3(256), 0, c.stream, pc.d_e1b, pc.d_e1b32, n);
                hipLaunchKernelGGL(toFloatKernel, dim3(grid), dim3(256), 0, c.stream, pc.d_e1k, pc.d_e1k32, n);
                LBX_HIP(hipGetLastError());
            }
        }

        const KernelSlot &slotFor(int mode, const PairClass &bra, const PairClass &ket, int variant)
        {
            const KernelSlot &s = kernelSlot(mode, bra.la, bra.lb, ket.la, ket.lb, variant);
            if (!s.fn)
                throw Error(LBX_ERR_INVALID, "no kernel for class (" + std::to_string(bra.la) +
                                                     std::to_string(bra.lb) + "|" +
                                                     std::to_string(ket.la) + std::to_string(ket.lb) +
                                                     ") mode " + std::to_string(mode) + " variant " +
                                                     std::to_string(variant));
            return s;
        }

        void run(Context &c, int mode, const PairClass &bra, const PairClass &ket, Pairing pairing,
                 bool same, double thresh, KArgs extra)
        {
            if (bra.n_pairs == 0 || ket.n_pairs == 0)
                return;
            const KernelSlot &s = slotFor(mode, bra, ket, defaultVariant(mode, bra, ket));
            TaskTable t = buildTasks(c, bra, ket, pairing, same, thresh, s.width);
            KArgs a = baseArgs(c, bra, ket, t, same);
            a.D = extra.D;
            a.J = extra.J;
            a.K = extra.K;
            a.out = extra.out;
            a.kind = extra.kind;
            a.qout = extra.qout;
            a.dout = extra.dout;
            s.fn(a, c.stream);
            LBX_HIP(hipGetLastError());
            LBX_HIP(hipStreamSynchronize(c.stream));
            // the task table's device arrays were allocated from the context; free them now
            for (void *p : {static_cast<void *>(t.d_prefix), static_cast<void *>(t.d_plo),
                            static_cast<void *>(t.d_phi)})
            {
                auto it = std::find(c.allocations.begin(), c.allocations.end(), p);
                if (it != c.allocations.end())
                {
                    (void)hipFree(p);
                    c.allocations.erase(it);
                }
            }
        }
    }

    KernelSlot &kernelSlot(int mode, int la, int lb, int lc, int ld, int variant)
    {
        static KernelSlot tab[N_MODES][kMaxLAux + 1][kMaxLMain + 1][kMaxLAux + 1][kMaxLMain + 1][N_VARIANTS];
        if (mode < 0 || mode >= N_MODES || la < 0 || la > kMaxLAux || lc < 0 || lc > kMaxLAux || lb < 0 ||
            lb > kMaxLMain || ld < 0 || ld > kMaxLMain || variant < 0 || variant >= N_VARIANTS)
            throw Error(LBX_ERR_INVALID, "kernel slot out of range");
        return tab[mode][la][lb][lc][ld][variant];
    }

    GradSlot &gradSlot(int la, int lb, int lc, int ld)
    {
        static GradSlot tab[kMaxLMain + 1][kMaxLMain + 1][kMaxLMain + 1][kMaxLMain + 1];
        if (la < 0 || la > kMaxLMain || lb < 0 || lb > kMaxLMain || lc < 0 || lc > kMaxLMain || ld < 0 ||
            ld > kMaxLMain)
            throw Error(LBX_ERR_INVALID, "gradient kernel slot out of range");
        return tab[la][lb][lc][ld];
    }

    // Algorithmic FLOPs of the gradient per quartet (7 ket-sized and 9 bra-sized transforms, R to
    // order L + 1, the density contraction of 12 derivative blocks), from a task table's sums.
    double gradFlops(const TaskTable &t, int la, int lb, int lc, int ld)
    {
        const double Hab = nHerm(la + lb), Hcd = nHerm(lc + ld);
        const double Sab = nSph(la) * nSph(lb), Scd = nSph(lc) * nSph(ld);
        return t.sum_kk * (7 * 2 * Hab * Hcd * Scd + rFlopsHost(la + lb + lc + ld + 1) + Hab * Hcd) +
               t.sum_kbra * 9 * 2 * Sab * Hab * Scd + t.quartets * 72 * Sab * Scd;
    }

    // Schwarz factors Q_ab = sqrt(max |(ab|ab)|) per pair of every main class; pairs sorted by
    // descending Q so that each ket's surviving bra pairs form a prefix.
    // Schwarz factors Q = sqrt(max |(ab|ab)|) of one class (kind 4: pairs, kind 2: single
    // auxiliary shells) and the sorted order: by molecule, then Q descending (one molecule in an
    // ordinary context), so each molecule's pairs keep their contiguous range of positions.
    void sortByBound(Context &c, PairClass &pc, int kind)
    {
        if (pc.n_pairs == 0)
            return;
        KArgs extra{};
        extra.qout = pc.d_q;
        extra.kind = kind;
        run(c, MODE_DIAG, pc, pc, Pairing::Diagonal, true, 0.0, extra);
        pc.Q.resize(pc.n_pairs);
        LBX_HIP(hipMemcpy(pc.Q.data(), pc.d_q, pc.n_pairs * sizeof(double), hipMemcpyDeviceToHost));
        std::vector<int> order(pc.n_pairs);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            return pc.mol[x] != pc.mol[y] ? pc.mol[x] < pc.mol[y] : pc.Q[x] > pc.Q[y];
        });
        std::vector<double> qs(pc.n_pairs);
        for (int i = 0; i < pc.n_pairs; i++)
            qs[i] = pc.Q[order[i]];
        pc.order = order;
        double kl = 0.0;
        for (int b = 0; b < pc.n_pairs; b += 64)
        {
            int m = 0;
            for (int i = b; i < std::min(b + 64, pc.n_pairs); i++)
                m = std::max(m, pc.pcnt[order[i]]);
            kl += m;
        }
        pc.klane = kl / ((pc.n_pairs + 63) / 64);
        LBX_HIP(hipMemcpy(pc.d_order, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice));
        LBX_HIP(hipMemcpy(pc.d_q, qs.data(), qs.size() * sizeof(double), hipMemcpyHostToDevice));
    }

    void ensureSchwarz(Context &c)
    {
        if (c.have_schwarz)
            return;
        for (auto &pc : c.cls)
            sortByBound(c, pc, 4);
        c.have_schwarz = true;
    }

    // the same for the auxiliary shells (bounds of the RI sums' precision split)
    void ensureSchwarzAux(Context &c)
    {
        if (c.have_schwarz_aux)
            return;
        for (auto &pc : c.aux)
            sortByBound(c, pc, 2);
        c.have_schwarz_aux = true;
    }
}

using namespace lbx;

struct lbx_context : Context
{
    // J/K task tables per (bra class, ket class, bound range [lo, hi), variant); Q does not
    // depend on D
    std::map<std::tuple<int, int, double, double, int>, TaskTable> jk_tasks;
    // kernel variant per class pair chosen by lbx_tune (absent: work-estimate default), for
    // the FP64 build and for the FP32 tier of the mixed build
    std::map<std::pair<int, int>, int> tuned, tuned32;
    // lbx_set_class_filter: the one class pair (bra class, ket class) lbx_jk evaluates; -1: all
    int filter_i = -1, filter_j = -1;
    // lbx_set_graphs: capture the J/K launch sequences and the RI sums into HIP graphs and
    // replay them (off by default: a capture in progress makes synchronous HIP calls of other
    // threads on the legacy stream fail)
    bool graphs = false;
    // lbx_set_bound_window: lbx_jk evaluates only quartets with window_lo <= Q_ab Q_cd < window_hi
    double window_lo = 0.0, window_hi = kNoUpper;
    // lbx_eri*_batch: (shell a, shell b) -> pair index per main pair class
    std::vector<std::unordered_map<long long, int>> pair_maps;
    // this device's scratch array (lbx_eri4_batch_static)
    int *d_scratch = nullptr;

    int variantFor(int i, int j)
    {
        auto it = tuned.find({i, j});
        return it != tuned.end() ? it->second : defaultVariant(MODE_JK, cls[i], cls[j]);
    }

    // variant for the small-bound quartets of a mixed build: an FP32 form, or an FP64 one
    // where the class has no FP32 form or tuning found the FP64 form faster
    int variant32For(int i, int j)
    {
        auto it = tuned32.find({i, j});
        if (it != tuned32.end())
            return it->second;
        const int v = defaultVariant32(cls[i], cls[j]);
        return v >= 0 ? v : variantFor(i, j);
    }

    // Lane side of the one-lane J/K kernels, per pair class: Schwarz-order positions regrouped
    // by shell within half-decade buckets of the Schwarz factor (lane position -> Schwarz-order
    // position), so that lanes sharing the first shell are adjacent and combine their K updates,
    // and the range of Schwarz-order positions each 64-lane block holds.
    struct LanePerm
    {
        std::vector<int> perm, bqlo, bqhi;
        int *d_perm = nullptr, *d_bqlo = nullptr, *d_bqhi = nullptr;
    };
    std::map<int, LanePerm> lane_perm;

    const LanePerm &lanePerm(int idx)
    {
        auto it = lane_perm.find(idx);
        if (it != lane_perm.end())
            return it->second;
        const PairClass &pc = cls[idx];
        const int n = pc.n_pairs;
        std::vector<double> qs(n);
        std::vector<int> sa(n), sb(n);
        for (int p = 0; p < n; p++)
        {
            qs[p] = pc.Q.empty() ? 1.0 : pc.Q[pc.order[p]];
            sa[p] = pc.sh_a[pc.order[p]];
            sb[p] = pc.sh_b[pc.order[p]];
        }
        tasks::LaneOrder lo = tasks::laneOrder(qs, sa, sb);
        LanePerm lp;
        std::vector<int> perm = std::move(lo.perm);
        lp.bqlo = std::move(lo.bqlo);
        lp.bqhi = std::move(lo.bqhi);
        lp.d_perm = upload(perm);
        lp.d_bqlo = upload(lp.bqlo);
        lp.d_bqhi = upload(lp.bqhi);
        lp.perm = std::move(perm);
        return lane_perm.emplace(idx, std::move(lp)).first->second;
    }

    // Queues the J/K kernel of class pair (i >= j) with the given variant on the context stream,
    // over the quartets with lo <= Q_ab Q_cd < hi.
    const TaskTable &launchJK(int i, int j, int var, const double *D, double *J, double *K,
                              double thresh, double thresh_hi = kNoUpper)
    {
        const PairClass &bra = cls[i], &ket = cls[j];
        const KernelSlot &s = slotFor(MODE_JK, bra, ket, var);
        // the one-wave-per-quartet kernel walks consecutive tasks of one pair of the side whose
        // E matrices are larger, so those stay in cache across the tasks
        auto eBytes = [](const PairClass &c)
        { return c.n_pairs ? double(c.n_ppairs) / c.n_pairs * nHerm(c.la + c.lb) * nSph(c.la) * nSph(c.lb) : 0.0; };
        const bool braMajor = !s.swap && s.kind == KIND_WAVE && i != j && eBytes(bra) > eBytes(ket);
        // the block kernel's tasks: per bra position, blocks of s.width consecutive ket positions
        const bool blockKind = s.kind == KIND_BLOCK;
        const auto key = std::make_tuple(i, j, thresh, thresh_hi, var);
        auto it = jk_tasks.find(key);
        if (it == jk_tasks.end())
            it = jk_tasks
                         .emplace(key, (s.swap || braMajor || blockKind)
                                               ? buildTasks(*this, ket, bra, Pairing::Unique, blockKind && i == j,
                                                            thresh, s.width, false, thresh_hi)
                                               : buildTasks(*this, bra, ket, Pairing::Unique, i == j, thresh,
                                                            s.width, true, thresh_hi))
                         .first;
        // The one-lane-per-quartet kernels run their tasks by groups of lane blocks: as many
        // blocks (a power of two, up to 64) as keep the lane side's E matrices of a group within
        // about 1 MiB of L2, so each uniform pair's E is read once per group.
        const bool blockMajor = s.kind == KIND_LANE;
        const PairClass &lside = s.swap ? ket : bra;
        const int nlane = lside.n_pairs;
        int gshift = 0;
        if (blockMajor)
        {
            const double groupBytes = 64.0 * eBytes(lside) * (isF32Variant(var) ? 4 : 8);
            while (gshift < 6 && groupBytes * (2 << gshift) <= double(1 << 20))
                gshift++;
        }
        const LanePerm *lp = blockMajor && nlane > 0 ? &lanePerm(s.swap ? j : i) : nullptr;
        if (blockMajor && it->second.nblk == 0 && nlane > 0)
            buildBlockMajor(*this, it->second, nlane, 1, gshift, lp->perm, lp->bqlo, lp->bqhi);
        const TaskTable &t = it->second;
        KArgs a = baseArgs(*this, bra, ket, t, i == j);
        a.bra_major = braMajor ? 1 : 0;
        a.D = D;
        a.J = J;
        a.K = K;
        if (blockMajor && t.nblk > 0)
        {
            a.block_major = 1;
            a.nblk = t.nblk;
            a.gshift = t.gshift;
            a.lperm = lp->d_perm;
            a.bqlo = lp->d_bqlo;
            a.bqhi = lp->d_bqhi;
            a.bprefix = t.d_bprefix;
            a.bq0 = t.d_bq0;
            a.ntasks = t.btasks;
        }
        if (sample_quartets > 0 && a.block_major)
        {
            // every stride-th lane block with all its uniform-side partners
            auto st = tune_tasks.find(key);
            if (st == tune_tasks.end())
            {
                const int stride = static_cast<int>(
                        std::max(1.0, std::ceil(t.quartets / static_cast<double>(sample_quartets))));
                TaskTable ts;
                ts.nket = t.nket;
                ts.plo = t.plo;
                ts.phi = t.phi;
                ts.cuts = t.cuts;
                ts.quartets = buildBlockMajor(*this, ts, nlane, stride, t.gshift, lp->perm, lp->bqlo, lp->bqhi);
                st = tune_tasks.emplace(key, std::move(ts)).first;
            }
            const TaskTable &ts = st->second;
            a.bprefix = ts.d_bprefix;
            a.bq0 = ts.d_bq0;
            a.ntasks = ts.btasks;
            sampled = ts.quartets;
        }
        else if (sample_quartets > 0)
        {
            // every stride-th uniform-side position, about sample_quartets quartets spread over
            // the whole range of Schwarz bounds
            auto st = tune_tasks.find(key);
            if (st == tune_tasks.end())
            {
                const int stride = static_cast<int>(
                        std::max(1.0, std::ceil(t.quartets / static_cast<double>(sample_quartets))));
                TaskTable ts;
                ts.nket = t.nket;
                ts.plo = t.plo;
                ts.phi = t.phi;
                ts.prefix.resize(t.nket + 1);
                long long acc = 0;
                for (int q = 0; q < t.nket; q++)
                {
                    if (q % stride != 0)
                        ts.phi[q] = ts.plo[q];
                    ts.prefix[q] = acc;
                    const long long cnt = ts.phi[q] - ts.plo[q];
                    acc += (cnt + s.width - 1) / s.width;
                    ts.quartets += static_cast<double>(cnt);
                }
                ts.prefix[t.nket] = acc;
                ts.ntasks = acc;
                ts.d_prefix = upload(ts.prefix);
                ts.d_plo = upload(ts.plo);
                ts.d_phi = upload(ts.phi);
                st = tune_tasks.emplace(key, std::move(ts)).first;
            }
            const TaskTable &ts = st->second;
            a.prefix = ts.d_prefix;
            a.plo = ts.d_plo;
            a.phi = ts.d_phi;
            a.ntasks = ts.ntasks;
            sampled = ts.quartets;
        }
        s.fn(a, stream);
        LBX_HIP(hipGetLastError());
        return t;
    }

    // when > 0, launchJK runs only a sample of ~sample_quartets quartets (autotuner timing) and
    // leaves their count in `sampled`
    long long sample_quartets = 0;
    double sampled = 0;
    std::map<std::tuple<int, int, double, double, int>, TaskTable> tune_tasks;   // the samples

    // gradient task tables per (bra class, ket class, Schwarz threshold)
    std::map<std::tuple<int, int, double, double>, TaskTable> grad_tasks;

    // batched RI sums: the whole launch sequence captured once into a graph (per output buffer)
    hipGraph_t ri_graph = nullptr;
    hipGraphExec_t ri_exec = nullptr;
    double *ri_out = nullptr;
    double ri_triplets = 0, ri_doublets = 0, ri_flops = 0, ri_triplets32 = 0, ri_doublets32 = 0;
    double ri_split = 0.0;   // precision split the prepared launches were built for (0: FP64)

    void dropRiGraph()
    {
        if (ri_exec)
            (void)hipGraphExecDestroy(ri_exec);
        if (ri_graph)
            (void)hipGraphDestroy(ri_graph);
        ri_exec = nullptr;
        ri_graph = nullptr;
        ri_out = nullptr;
    }

    // Kernel launches of the RI sums: every eri3 (pair class | aux l) and eri2
    // (aux l | aux l' <= l) class pair, with task tables built once per precision split (device
    // memory cannot be allocated while a stream is being captured).  split > 0: the triplets and
    // doublets with bound Q_ab Q_P below it in the FP32 form where the class has one.
    std::vector<std::pair<Launcher, KArgs>> ri_launches;
    std::map<double, std::vector<std::pair<Launcher, KArgs>>> ri_prepared;
    std::map<double, std::array<double, 5>> ri_counts;   // triplets, doublets, flops, FP32 ones
    // FP64 variant of each RI class pair chosen by lbx_tune on a batched context, keyed by
    // (kind, lane class, uniform class): kind 3 (main pair class | aux l), 2 (aux l | aux l')
    std::map<std::tuple<int, int, int>, int> tuned_sum;

    int sumVariant(int kind, int il, int iu, const PairClass &lane, const PairClass &unif) const
    {
        auto it = tuned_sum.find({kind, il, iu});
        if (it != tuned_sum.end())
            return it->second;
        return kernelSlot(MODE_SUM, lane.la, lane.lb, unif.la, unif.lb, VAR_VALU_BRA).fn ? VAR_VALU_BRA : VAR_MFMA;
    }

    // Times the FP64 sum kernels of every RI class pair (the one-lane form and the MFMA form,
    // each over the class pair's whole task table into a scratch output) and keeps the faster.
    void tuneRiSums()
    {
        DevBuf scratch(4LL * nmol);
        hipEvent_t e0, e1;
        LBX_HIP(hipEventCreate(&e0));
        LBX_HIP(hipEventCreate(&e1));
        auto one = [&](int kind, int il, int iu, const PairClass &lane, const PairClass &unif, bool same)
        {
            if (lane.n_pairs == 0 || unif.n_pairs == 0)
                return;
            int best = -1;
            float best_ms = 0.0f;
            for (int v : {static_cast<int>(VAR_VALU_BRA), static_cast<int>(VAR_MFMA)})
            {
                const KernelSlot &k = kernelSlot(MODE_SUM, lane.la, lane.lb, unif.la, unif.lb, v);
                if (!k.fn)
                    continue;
                TaskTable t = buildTasksMol(*this, lane, unif, same, k.width);
                KArgs a = baseArgs(*this, lane, unif, t, same);
                a.kind = kind;
                a.out = scratch.p;
                k.fn(a, stream);   // loads the code
                LBX_HIP(hipEventRecord(e0, stream));
                k.fn(a, stream);
                LBX_HIP(hipEventRecord(e1, stream));
                LBX_HIP(hipGetLastError());
                LBX_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                LBX_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (best < 0 || ms < best_ms)
                {
                    best = v;
                    best_ms = ms;
                }
                // the table's device arrays were allocated from the context; free them now
                for (void *p : {static_cast<void *>(t.d_prefix), static_cast<void *>(t.d_plo),
                                static_cast<void *>(t.d_phi), static_cast<void *>(t.d_bprefix),
                                static_cast<void *>(t.d_bq0)})
                {
                    auto it = std::find(allocations.begin(), allocations.end(), p);
                    if (p && it != allocations.end())
                    {
                        (void)hipFree(p);
                        allocations.erase(it);
                    }
                }
            }
            if (best >= 0)
                tuned_sum[{kind, il, iu}] = best;
        };
        try
        {
            for (int i = 0; i < static_cast<int>(cls.size()); i++)
                for (int l = 0; l < static_cast<int>(aux.size()); l++)
                    one(3, i, l, cls[i], aux[l], false);
            for (int la = 0; la < static_cast<int>(aux.size()); la++)
                for (int lc = 0; lc <= la; lc++)
                    one(2, la, lc, aux[la], aux[lc], la == lc);
        }
        catch (...)
        {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        dropRiGraph();
        ri_prepared.clear();
        ri_counts.clear();
    }

    void prepareRiSums(double split)
    {
        auto it = ri_prepared.find(split);
        if (it == ri_prepared.end())
        {
            std::vector<std::pair<Launcher, KArgs>> launches;
            std::array<double, 5> n{0, 0, 0, 0, 0};
            auto add = [&](const PairClass &lane, const PairClass &unif, bool same, int kind, int il, int iu)
            {
                if (lane.n_pairs == 0 || unif.n_pairs == 0)
                    return;
                const int var = sumVariant(kind, il, iu, lane, unif);
                const KernelSlot &ks = slotFor(MODE_SUM, lane, unif, var);
                const KernelSlot &k32 = kernelSlot(MODE_SUM, lane.la, lane.lb, unif.la, unif.lb, VAR_VALU_BRA32);
                const bool tiers = split > 0.0 && k32.fn;
                auto one = [&](const KernelSlot &k, double lo, double hi, bool fp32)
                {
                    TaskTable t = buildTasksMol(*this, lane, unif, same, k.width, lo, hi);
                    KArgs a = baseArgs(*this, lane, unif, t, same);
                    a.kind = kind;
                    launches.emplace_back(k.fn, a);
                    n[kind == 3 ? 0 : 1] += t.quartets;
                    n[2] += t.flops;
                    if (fp32)
                        n[kind == 3 ? 3 : 4] += t.quartets;
                };
                if (tiers)
                {
                    one(ks, split, kNoUpper, false);
                    one(k32, 0.0, split, true);
                }
                else
                    one(ks, 0.0, kNoUpper, false);
            };
            for (int i = 0; i < static_cast<int>(cls.size()); i++)
                for (int l = 0; l < static_cast<int>(aux.size()); l++)
                    add(cls[i], aux[l], false, 3, i, l);
            for (int la = 0; la < static_cast<int>(aux.size()); la++)
                for (int lc = 0; lc <= la; lc++)
                    add(aux[la], aux[lc], la == lc, 2, la, lc);
            it = ri_prepared.emplace(split, std::move(launches)).first;
            ri_counts[split] = n;
        }
        ri_launches = it->second;
        const auto &n = ri_counts[split];
        ri_triplets = n[0];
        ri_doublets = n[1];
        ri_flops = n[2];
        ri_triplets32 = n[3];
        ri_doublets32 = n[4];
    }

    // fill + launches on stream s (the capture stream)
    void queueRiSums(double *out, hipStream_t s)
    {
        fill(out, 4LL * nmol, 0.0, s);
        for (auto &l : ri_launches)
        {
            KArgs a = l.second;
            a.out = out;
            l.first(a, s);
            LBX_HIP(hipGetLastError());
        }
    }

    // Queues the gradient kernel of class pair (i >= j) over the quartets with
    // thresh <= Q_ab Q_cd < thresh_hi, in FP32 arithmetic when fp32; returns its task table.
    const TaskTable &launchGrad(int i, int j, const double *D, double *grad, double thresh,
                                double thresh_hi = kNoUpper, bool fp32 = false)
    {
        const PairClass &ci = cls[i], &cj = cls[j];
        const bool lane_i = gradLaneFirst(ci.la, ci.lb, cj.la, cj.lb);
        const PairClass &lane = lane_i ? ci : cj, &unif = lane_i ? cj : ci;
        const GradSlot &s = gradSlot(lane.la, lane.lb, unif.la, unif.lb);
        if (!s.fn || !s.fn32)
            throw Error(LBX_ERR_INVALID, "no gradient kernel for this class pair");
        const auto key = std::make_tuple(i, j, thresh, thresh_hi);
        auto it = grad_tasks.find(key);
        if (it == grad_tasks.end())
            it = grad_tasks
                         .emplace(key, buildTasks(*this, lane, unif, Pairing::Unique, i == j, thresh, s.width,
                                                  lane_i, thresh_hi))
                         .first;
        const TaskTable &t = it->second;
        GArgs a{};
        a.lane = lane.view();
        a.unif = unif.view();
        a.prefix = t.d_prefix;
        a.plo = t.d_plo;
        a.phi = t.d_phi;
        a.nket = t.nket;
        a.ntasks = t.ntasks;
        a.same_class = i == j ? 1 : 0;
        a.boys = d_boys;
        a.D = D;
        a.nbf = static_cast<long long>(nbf);
        a.grad = grad;
        (fp32 ? s.fn32 : s.fn)(a, stream);
        LBX_HIP(hipGetLastError());
        return t;
    }

    // Queues one J/K build on the context stream: J and K cleared, every class pair's kernel
    // (the FP32 tier of a mixed build below `split`), J and K symmetrised.
    lbx_jk_stats queueJK(const double *D, double *J, double *K, double schwarz_thresh, double split,
                         bool mixed, bool prof_on);

    // J/K launch sequences captured into graphs, per (D, J, K, Schwarz threshold, split, mixed);
    // dropped when the variant choice changes (lbx_tune, lbx_set_variant)
    using JKKey = std::tuple<const double *, double *, double *, double, double, int>;
    struct JKGraph
    {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        lbx_jk_stats stats{};
    };
    std::map<JKKey, JKGraph> jk_graphs;
    static constexpr size_t kMaxJKGraphs = 8;

    void dropJKGraphs()
    {
        for (auto &kv : jk_graphs)
        {
            if (kv.second.exec)
                (void)hipGraphExecDestroy(kv.second.exec);
            if (kv.second.graph)
                (void)hipGraphDestroy(kv.second.graph);
        }
        jk_graphs.clear();
    }

    // Captures the launch sequence of a build just run eagerly (its task tables exist, so the
    // capture allocates nothing) on a private stream of default priority.
    void captureJK(const JKKey &key, const lbx_jk_stats &st)
    {
        if (jk_graphs.size() >= kMaxJKGraphs)
            dropJKGraphs();
        hipStream_t cs = nullptr;
        LBX_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        hipStream_t saved = stream;
        JKGraph g;
        g.stats = st;
        try
        {
            stream = cs;
            LBX_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
            try
            {
                queueJK(std::get<0>(key), std::get<1>(key), std::get<2>(key), std::get<3>(key), std::get<4>(key),
                        std::get<5>(key) != 0, false);
            }
            catch (...)
            {
                hipGraph_t tmp = nullptr;
                (void)hipStreamEndCapture(cs, &tmp);
                if (tmp)
                    (void)hipGraphDestroy(tmp);
                throw;
            }
            LBX_HIP(hipStreamEndCapture(cs, &g.graph));
            stream = saved;
            LBX_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
        }
        catch (...)
        {
            stream = saved;
            if (g.graph)
                (void)hipGraphDestroy(g.graph);
            (void)hipStreamDestroy(cs);
            throw;
        }
        (void)hipStreamDestroy(cs);
        jk_graphs.emplace(key, g);
    }

    // per-class timing (lbx_profile_enable)
    struct Prof
    {
        lbx_class_profile rec{};
        hipEvent_t start = nullptr, stop = nullptr;
        bool pending = false;
    };
    bool profiling = false;
    std::map<std::tuple<int, int, int>, Prof> prof;   // (bra class, ket class, tier 0 FP64 / 1 FP32)

    ~lbx_context()
    {
        dropRiGraph();
        dropJKGraphs();
        for (auto &kv : prof)
        {
            if (kv.second.start)
                (void)hipEventDestroy(kv.second.start);
            if (kv.second.stop)
                (void)hipEventDestroy(kv.second.stop);
        }
    }

    void collectProfile()
    {
        for (auto &kv : prof)
        {
            Prof &p = kv.second;
            if (!p.pending)
                continue;
            float ms = 0;
            LBX_HIP(hipEventElapsedTime(&ms, p.start, p.stop));
            p.rec.ms += ms;
            p.pending = false;
        }
    }
};

namespace
{
    // Shells, pair classes, E matrices and per-pair molecule ids of a new context.
    void setupContext(lbx_context &ctx, const lbx_basis &main_basis, const lbx_basis *aux_basis)
    {
        Context &c = ctx;
        LBX_HIP(hipGetDevice(&c.device));
        ctx.d_scratch = scratchAddress(c.device);
        c.natoms = main_basis.natoms;
        c.shells = buildShells(main_basis, c.nbf, c.max_l, kMaxLMain);
        if (c.shells.empty())
            throw Error(LBX_ERR_INVALID, "lbx_create: empty main basis");
        if (aux_basis)
        {
            if (aux_basis->natoms != c.natoms)
                throw Error(LBX_ERR_INVALID, "lbx_create: aux basis natoms differs");
            c.shells_aux = buildShells(*aux_basis, c.nbf_aux, c.max_l_aux, kMaxLAux);
        }
        c.d_boys = const_cast<double *>(boysTableDevice());
        buildPairClasses(c);
        for (auto &pc : c.cls)
        {
            computeEcoeffs(c, pc, c.shells);
            pc.d_mol = c.upload(pc.mol);
        }
        for (auto &pc : c.aux)
        {
            computeEcoeffs(c, pc, c.shells_aux);
            pc.d_mol = c.upload(pc.mol);
        }
        LBX_HIP(hipStreamSynchronize(c.stream));
    }

    // Concatenation of several molecules' lbx_basis arrays (atom indices offset).
    struct FlatBatch
    {
        std::vector<double> coords, exps, coeffs;
        std::vector<int> atom, l, nprim;
        int natoms = 0;
        lbx_basis b{};

        void add(const lbx_basis &x)
        {
            if (x.natoms < 0 || x.nshells < 0 || (x.natoms > 0 && !x.coords) ||
                (x.nshells > 0 && (!x.shell_atom || !x.shell_l || !x.shell_nprim || !x.exps || !x.coeffs)))
                throw Error(LBX_ERR_INVALID, "lbx_create_batch: bad basis");
            coords.insert(coords.end(), x.coords, x.coords + 3 * x.natoms);
            long long np = 0;
            for (int s = 0; s < x.nshells; s++)
            {
                if (x.shell_atom[s] < 0 || x.shell_atom[s] >= x.natoms || x.shell_nprim[s] <= 0)
                    throw Error(LBX_ERR_INVALID, "lbx_create_batch: bad shell");
                atom.push_back(x.shell_atom[s] + natoms);
                l.push_back(x.shell_l[s]);
                nprim.push_back(x.shell_nprim[s]);
                np += x.shell_nprim[s];
            }
            exps.insert(exps.end(), x.exps, x.exps + np);
            coeffs.insert(coeffs.end(), x.coeffs, x.coeffs + np);
            natoms += x.natoms;
        }

        const lbx_basis &view()
        {
            b = lbx_basis{natoms, coords.data(), static_cast<int>(l.size()), atom.data(), l.data(),
                          nprim.data(), exps.data(), coeffs.data()};
            return b;
        }
    };
}

extern "C"
{
    const char *lbx_last_error(void) { return g_err.c_str(); }

    int lbx_create(lbx_context **out, const lbx_basis *main_basis, const lbx_basis *aux_basis)
    {
        if (!out || !main_basis)
        {
            g_err = "lbx_create: null argument";
            return LBX_ERR_INVALID;
        }
        *out = nullptr;
        auto ctx = std::make_unique<lbx_context>();
        int rc = guarded([&] { setupContext(*ctx, *main_basis, aux_basis); });
        if (rc != LBX_OK)
        {
            ctx->release();
            return rc;
        }
        *out = ctx.release();
        return LBX_OK;
    }

    int lbx_create_batch(lbx_context **out, int nmol, const lbx_basis *main_bases,
                         const lbx_basis *aux_bases)
    {
        if (!out || !main_bases || !aux_bases || nmol <= 0)
        {
            g_err = "lbx_create_batch: null argument or nmol <= 0";
            return LBX_ERR_INVALID;
        }
        *out = nullptr;
        auto ctx = std::make_unique<lbx_context>();
        int rc = guarded([&]
                         {
                             // one basis over all molecules' atoms, atoms grouped by molecule
                             FlatBatch m, a;
                             for (int k = 0; k < nmol; k++)
                             {
                                 if (main_bases[k].natoms != aux_bases[k].natoms)
                                     throw Error(LBX_ERR_INVALID, "lbx_create_batch: aux basis natoms differs");
                                 for (int i = 0; i < main_bases[k].natoms; i++)
                                     ctx->atom_mol.push_back(k);
                                 m.add(main_bases[k]);
                                 a.add(aux_bases[k]);
                             }
                             ctx->nmol = nmol;
                             setupContext(*ctx, m.view(), &a.view());
                         });
        if (rc != LBX_OK)
        {
            ctx->release();
            return rc;
        }
        *out = ctx.release();
        return LBX_OK;
    }

    int lbx_destroy(lbx_context *ctx)
    {
        if (!ctx)
            return LBX_OK;
        int rc = guarded([&]
                         {
                             (void)hipStreamSynchronize(ctx->stream);
                             ctx->release();
                         });
        delete ctx;
        return rc;
    }

    int lbx_dims(const lbx_context *ctx, long *nbf, long *nbf_aux, long *natoms)
    {
        if (!ctx)
        {
            g_err = "lbx_dims: null context";
            return LBX_ERR_INVALID;
        }
        if (nbf)
            *nbf = static_cast<long>(ctx->nbf);
        if (nbf_aux)
            *nbf_aux = static_cast<long>(ctx->nbf_aux);
        if (natoms)
            *natoms = ctx->natoms;
        return LBX_OK;
    }

    int lbx_set_stream(lbx_context *ctx, void *stream)
    {
        if (!ctx)
        {
            g_err = "lbx_set_stream: null context";
            return LBX_ERR_INVALID;
        }
        ctx->stream = static_cast<hipStream_t>(stream);
        return LBX_OK;
    }

    int lbx_synchronize(lbx_context *ctx)
    {
        if (!ctx)
        {
            g_err = "lbx_synchronize: null context";
            return LBX_ERR_INVALID;
        }
        return guarded([&] { LBX_HIP(hipStreamSynchronize(ctx->stream)); });
    }

    int lbx_boys(int n, long count, const double *x, double *out, void *stream)
    {
        return guarded([&]
                       {
                           if (n < 0 || n > kMaxLBoys || count < 0 || (count > 0 && (!x || !out)))
                               throw Error(LBX_ERR_INVALID, "lbx_boys: bad argument");
                           if (count == 0)
                               return;
                           static const auto fns = boysFns(std::make_integer_sequence<int, kMaxLBoys + 1>{});
                           fns[n](count, x, out, boysTableDevice(), static_cast<hipStream_t>(stream));
                           LBX_HIP(hipGetLastError());
                       });
    }

    int lbx_eri4(lbx_context *ctx, double *out)
    {
        return guarded([&]
                       {
                           if (!ctx || !out)
                               throw Error(LBX_ERR_INVALID, "lbx_eri4: null argument");
                           std::lock_guard<std::mutex> lk(ctx->mu);
                           Context &c = *ctx;
                           requireSingle(c, "lbx_eri4");
                           const long long n = static_cast<long long>(c.nbf);
                           fill(out, n * n * n * n, 0.0, c.stream);
                           KArgs extra{};
                           extra.out = out;
                           extra.kind = 4;
                           for (size_t i = 0; i < c.cls.size(); i++)
                               for (size_t j = 0; j <= i; j++)
                                   run(c, MODE_DENSE, c.cls[i], c.cls[j], Pairing::Unique, i == j, 0.0, extra);
                       });
    }

    int lbx_eri4_diag(lbx_context *ctx, double *out)
    {
        return guarded([&]
                       {
                           if (!ctx || !out)
                               throw Error(LBX_ERR_INVALID, "lbx_eri4_diag: null argument");
                           std::lock_guard<std::mutex> lk(ctx->mu);
                           Context &c = *ctx;
                           requireSingle(c, "lbx_eri4_diag");
                           const long long n = static_cast<long long>(c.nbf);
                           fill(out, n * n, 0.0, c.stream);
                           for (auto &pc : c.cls)
                           {
                               if (pc.n_pairs == 0)
                                   continue;
                               DevBuf q(pc.n_pairs);
                               KArgs extra{};
                               extra.qout = q.p;
                               extra.dout = out;
                               extra.kind = 4;
                               run(c, MODE_DIAG, pc, pc, Pairing::Diagonal, true, 0.0, extra);
                           }
                       });
    }

    int lbx_eri3(lbx_context *ctx, double *out)
    {
        return guarded([&]
                       {
                           if (!ctx || !out)
                               throw Error(LBX_ERR_INVALID, "lbx_eri3: null argument");
                           std::lock_guard<std::mutex> lk(ctx->mu);
                           Context &c = *ctx;
                           requireSingle(c, "lbx_eri3");
                           if (c.shells_aux.empty())
                               throw Error(LBX_ERR_STATE, "eri3(): RI approximation is not enabled");
                           const long long n = static_cast<long long>(c.nbf);
                           fill(out, n * n * static_cast<long long>(c.nbf_aux), 0.0, c.stream);
                           KArgs extra{};
                           extra.out = out;
                           extra.kind = 3;
                           for (auto &bra : c.cls)
                               for (auto &ket : c.aux)
                                   run(c, MODE_DENSE, bra, ket, Pairing::All, false, 0.0, extra);
                       });
    }

    int lbx_eri2(lbx_context *ctx, double *out)
    {
        return guarded([&]
                       {
                           if (!ctx || !out)
                               throw Error(LBX_ERR_INVALID, "lbx_eri2: null argument");
                           std::lock_guard<std::mutex> lk(ctx->mu);
                           Context &c = *ctx;
                           requireSingle(c, "lbx_eri2");
                           if (c.shells_aux.empty())
                               throw Error(LBX_ERR_STATE, "eri2(): RI approximation is not enabled");
                           const long long n = static_cast<long long>(c.nbf_aux);
                           fill(out, n * n, 0.0, c.stream);
                           KArgs extra{};
                           extra.out = out;
                           extra.kind = 2;
                           for (size_t i = 0; i < c.aux.size(); i++)
                               for (size_t j = 0; j <= i; j++)
                                   run(c, MODE_DENSE, c.aux[i], c.aux[j], Pairing::Unique, i == j, 0.0, extra);
                       });
    }

    int lbx_eri2_diag(lbx_context *ctx, double *out)
    {
        return guarded([&]
                       {
                           if (!ctx || !out)
                               throw Error(LBX_ERR_INVALID, "lbx_eri2_diag: null argument");
                           std::lock_guard<std::mutex> lk(ctx->mu);
                           Context &c = *ctx;
                           requireSingle(c, "lbx_eri2_diag");
                           if (c.shells_aux.empty())
                               throw Error(LBX_ERR_STATE, "eri2Diagonal(): RI approximation is not enabled");
                           fill(out, static_cast<long long>(c.nbf_aux), 0.0, c.stream);
                           for (auto &pc : c.aux)
                           {
                               if (pc.n_pairs == 0)
                                   continue;
                               DevBuf q(pc.n_pairs);
                               KArgs extra{};
                               extra.qout = q.p;
                               extra.dout = out;
                               extra.kind = 2;
                               run(c, MODE_DIAG, pc, pc, Pairing::Diagonal, true, 0.0, extra);
                           }
                       });
    }

    // Dense derivative tensors (kind 4, 3, 2): every unique class pair's unique quartets, one
    // launch per class pair, each writing its blocks' images.
    namespace
    {
        void runD1(Context &c, int kind, double *out)
        {
            const long long n = static_cast<long long>(c.nbf), nx = static_cast<long long>(c.nbf_aux);
            const long long total = kind == 4 ? 12 * n * n * n * n : (kind == 3 ? 9 * n * n * nx : 6 * nx * nx);
            fill(out, total, 0.0, c.stream);
            if (kind != 2)
                for (auto &pc : c.cls)
                    computeEcoeffsD1(c, pc);
            std::vector<void *> bufs;
            auto one = [&](const PairClass &bra, const PairClass &ket, bool same)
            {
                if (bra.n_pairs == 0 || ket.n_pairs == 0)
                    return;
                std::vector<int> tasks;
                for (int p = 0; p < bra.n_pairs; p++)
                    for (int q = 0; q < (same ? p + 1 : ket.n_pairs); q++)
                    {
                        tasks.push_back(p);
                        tasks.push_back(q);
                    }
                void *d = nullptr;
                LBX_HIP(hipMalloc(&d, tasks.size() * sizeof(int)));
                bufs.push_back(d);
                LBX_HIP(hipMemcpyAsync(d, tasks.data(), tasks.size() * sizeof(int), hipMemcpyHostToDevice,
                                       c.stream));
                D1Args a{};
                a.bra = bra.view();
                a.ket = ket.view();
                a.la = bra.la;
                a.lb = bra.lb;
                a.lc = ket.la;
                a.ld = ket.lb;
                a.bra_single = bra.single ? 1 : 0;
                a.ket_single = ket.single ? 1 : 0;
                a.tasks = static_cast<const int *>(d);
                a.ntasks = static_cast<long long>(tasks.size() / 2);
                a.boys = c.d_boys;
                a.out = out;
                a.kind = kind;
                a.nbf = n;
                a.naux = nx;
                launchD1(a, c.stream);
                // the host task list must outlive the asynchronous copy
                LBX_HIP(hipStreamSynchronize(c.stream));
            };
            try
            {
                if (kind == 4)
                {
                    for (size_t i = 0; i < c.cls.size(); i++)
                        for (size_t j = 0; j <= i; j++)
                            one(c.cls[i], c.cls[j], i == j);
                }
                else if (kind == 3)
                {
                    for (auto &bra : c.cls)
                        for (auto &ket : c.aux)
                            one(bra, ket, false);
                }
                else
                {
                    for (size_t i = 0; i < c.aux.size(); i++)
                        for (size_t j = 0; j <= i; j++)
                            one(c.aux[i], c.aux[j], i == j);
                }
            }
            catch (...)
            {
                (void)hipStreamSynchronize(c.stream);
                for (void *d : bufs)
                    (void)hipFree(d);
                throw;
            }
            LBX_HIP(hipStreamSynchronize(c.stream));
            for (void *d : bufs)
                (void)hipFree(d);
        }
    }

    // Blocks of listed shell quartets / triplets / doublets (kind 4, 3, 2) in the caller's
    // shell order: the integrals themselves, their first derivatives, the spin-orbit-type
    // integrals (kinds 4, 3) or the second derivatives (kind 2) (op: D1_OP_*); every entry of
    // one call has the same angular momenta.  Entries whose shell pair has no primitive pair
    // above the 1e-15 screen (the reference's ShellPairData drops such pairs) are zero.
    namespace
    {
        // (shell a, shell b) of a pair class -> pair index, made on first use
        const std::unordered_map<long long, int> &pairIndex(Context &c, std::vector<std::unordered_map<long long, int>> &maps,
                                                           int cls_idx)
        {
            if (maps.size() < c.cls.size())
                maps.resize(c.cls.size());
            auto &m = maps[cls_idx];
            const PairClass &pc = c.cls[cls_idx];
            if (m.empty() && pc.n_pairs > 0)
            {
                const long long nsh = static_cast<long long>(c.shells.size());
                for (int p = 0; p < pc.n_pairs; p++)
                    m.emplace(static_cast<long long>(pc.sh_a[p]) * nsh + pc.sh_b[p], p);
            }
            return m;
        }

        void runBlocks(lbx_context &c, std::vector<std::unordered_map<long long, int>> &maps, int kind,
                       int op, long long nq, const int *sh, double *out, const char *who,
                       int *staging = nullptr)
        {
            const int per = kind == 4 ? 4 : (kind == 3 ? 3 : 2);
            const int nsh = static_cast<int>(c.shells.size()), nax = static_cast<int>(c.shells_aux.size());
            auto isAux = [&](int pos) { return kind == 2 || (kind == 3 && pos == 2); };
            auto lOf = [&](int pos, int idx) {
                const int n = isAux(pos) ? nax : nsh;
                if (idx < 0 || idx >= n)
                    throw Error(LBX_ERR_INVALID, std::string(who) + ": shell index " + std::to_string(idx) +
                                                         " out of range");
                return isAux(pos) ? c.shells_aux[idx].l : c.shells[idx].l;
            };
            if (nq == 0)
                return;
            int l[4] = {0, 0, 0, 0};
            for (int k = 0; k < per; k++)
                l[k] = lOf(k, sh[k]);
            for (long long t = 0; t < nq; t++)
                for (int k = 0; k < per; k++)
                    if (lOf(k, sh[per * t + k]) != l[k])
                        throw Error(LBX_ERR_INVALID, std::string(who) + ": every entry must have the "
                                                                        "angular momenta of the first");
            const int ncomp = op == D1_OP_PLAIN ? 1 : op == D1_OP_SOC ? 3 : op == D1_OP_D2 ? 36
                                                                                       : 3 * (kind == 4 ? 4 : (kind == 3 ? 3 : 2));
            long long bsize = 1;
            for (int k = 0; k < per; k++)
                bsize *= nSph(l[k]);
            fill(out, nq * ncomp * bsize, 0.0, c.stream);
            // library pair classes of the two sides
            PairClass *bra, *ket;
            int ibra = -1, iket = -1;
            if (kind == 2)
            {
                bra = &c.aux[l[0]];
                ket = &c.aux[l[1]];
            }
            else
            {
                ibra = pairClassIdx(std::max(l[0], l[1]), std::min(l[0], l[1]));
                bra = &c.cls[ibra];
                if (kind == 4)
                {
                    iket = pairClassIdx(std::max(l[2], l[3]), std::min(l[2], l[3]));
                    ket = &c.cls[iket];
                }
                else
                    ket = &c.aux[l[2]];
            }
            // pair of two main shells in the library's order (l_a > l_b, or equal l and a >= b)
            auto mainPair = [&](int cls_idx, int a, int b, bool &swapped) {
                swapped = c.shells[a].l < c.shells[b].l || (c.shells[a].l == c.shells[b].l && a < b);
                const long long key = static_cast<long long>(swapped ? b : a) * nsh + (swapped ? a : b);
                const auto &m = pairIndex(c, maps, cls_idx);
                auto it = m.find(key);
                return it == m.end() ? -1 : it->second;
            };
            // single auxiliary shell -> its position in the class of its l
            auto auxPair = [&](const PairClass &pc, int s) {
                auto it = std::lower_bound(pc.sh_a.begin(), pc.sh_a.end(), s);
                return it != pc.sh_a.end() && *it == s ? static_cast<int>(it - pc.sh_a.begin()) : -1;
            };
            std::vector<int> tasks;
            tasks.reserve(4 * nq);
            for (long long t = 0; t < nq; t++)
            {
                const int *e = sh + per * t;
                bool sab = false, scd = false;
                int p, q;
                if (kind == 2)
                {
                    p = auxPair(*bra, e[0]);
                    q = auxPair(*ket, e[1]);
                }
                else
                {
                    p = mainPair(ibra, e[0], e[1], sab);
                    q = kind == 4 ? mainPair(iket, e[2], e[3], scd) : auxPair(*ket, e[2]);
                }
                if (p < 0 || q < 0)
                    continue;
                tasks.insert(tasks.end(), {p, q, (sab ? 1 : 0) | (scd ? 2 : 0), static_cast<int>(t)});
            }
            if (tasks.empty())
                return;
            if (op == D1_OP_DERIV || op == D1_OP_SOC)
            {
                if (kind != 2)
                    computeEcoeffsD1(c, *bra);
                if (kind == 4 && op == D1_OP_DERIV)
                    computeEcoeffsD1(c, *ket);
            }
            // the task list: in a buffer of this call, or in the device's scratch array
            // (staging != nullptr: the allocation-free entry, at most kStaticEntries entries)
            void *d = staging;
            if (!staging)
                LBX_HIP(hipMalloc(&d, tasks.size() * sizeof(int)));
            try
            {
                LBX_HIP(hipMemcpyAsync(d, tasks.data(), tasks.size() * sizeof(int), hipMemcpyHostToDevice,
                                       c.stream));
                D1Args a{};
                a.bra = bra->view();
                a.ket = ket->view();
                a.la = bra->la;
                a.lb = bra->lb;
                a.lc = ket->la;
                a.ld = ket->lb;
                a.bra_single = bra->single ? 1 : 0;
                a.ket_single = ket->single ? 1 : 0;
                a.tasks = static_cast<const int *>(d);
                a.ntasks = static_cast<long long>(tasks.size() / 4);
                a.boys = c.d_boys;
                a.out = out;
                a.kind = kind;
                a.nbf = static_cast<long long>(c.nbf);
                a.naux = static_cast<long long>(c.nbf_aux);
                a.blocks = 1;
                a.op = op;
                launchD1(a, c.stream);
                LBX_HIP(hipStreamSynchronize(c.stream));
            }
            catch (...)
            {
                (void)hipStreamSynchronize(c.stream);
                if (!staging)
                    (void)hipFree(d);
                throw;
            }
            if (!staging)
                (void)hipFree(d);
        }

        int blocksEntry(lbx_context *ctx, int kind, int op, long nq, const int *shells, double *out,
                        const char *who, bool no_alloc = false)
        {
            return guarded([&]
                           {
                               if (!ctx || nq < 0 || (nq > 0 && (!shells || !out)))
                                   throw Error(LBX_ERR_INVALID, std::string(who) + ": null argument or negative count");
                               if (no_alloc && nq > kStaticEntries)
                                   throw Error(LBX_ERR_INVALID, std::string(who) + ": " + std::to_string(nq) +
                                                                        " entries, at most " +
                                                                        std::to_string(kStaticEntries) + " per call");
                               std::lock_guard<std::mutex> lk(ctx->mu);
                               requireSingle(*ctx, who);
                               if (kind != 4 && ctx->shells_aux.empty())
                                   throw Error(LBX_ERR_STATE, std::string(who) + ": RI approximation is not enabled");
                               runBlocks(*ctx, ctx->pair_maps, kind, op, nq, shells, out, who,
                                         no_alloc ? ctx->d_scratch : nullptr);
                           });
        }
    }

    int lbx_eri4_batch_static(lbx_context *ctx, long nq, const int *shells, double *out)
    {
        return blocksEntry(ctx, 4, D1_OP_PLAIN, nq, shells, out, "lbx_eri4_batch_static", true);
    }

    int lbx_eri4_batch(lbx_context *ctx, long nq, const int *shells, double *out)
    {
        return blocksEntry(ctx, 4, D1_OP_PLAIN, nq, shells, out, "lbx_eri4_batch");
    }

    int lbx_eri3_batch(lbx_context *ctx, long nq, const int *shells, double *out)
    {
        return blocksEntry(ctx, 3, D1_OP_PLAIN, nq, shells, out, "lbx_eri3_batch");
    }

    int lbx_eri2_batch(lbx_context *ctx, long nq, const int *shells, double *out)
    {
        return blocksEntry(ctx, 2, D1_OP_PLAIN, nq, shells, out, "lbx_eri2_batch");
    }

    int lbx_eri4d1_batch(lbx_context *ctx, long nq, const int *shells, double *out)
    {
        return blocksEntry(ctx, 4, D1_OP_DERIV, nq, shells, out, "lbx_eri4d1_batch");
    }

    int lbx_eri3d1_batch(lbx_context *ctx, long nq, const int *shells, double *out)
    {
        return blocksEntry(ctx, 3, D1_OP_DERIV, nq, shells, out, "lbx_eri3d1_batch");
    }

    int lbx_eri2d1_batch(lbx_context *ctx, long nq, const int *shells, double *out)
    {
        return blocksEntry(ctx, 2, D1_OP_DERIV, nq, shells, out, "lbx_eri2d1_batch");
    }

    int lbx_eri4soc_batch(lbx_context *ctx, long nq, const int *shells, double *out)
    {
        return blocksEntry(ctx, 4, D1_OP_SOC, nq, shells, out, "lbx_eri4soc_batch");
    }

    int lbx_eri3soc_batch(lbx_context *ctx, long nq, const int *shells, double *out)
    {
        return blocksEntry(ctx, 3, D1_OP_SOC, nq, shells, out, "lbx_eri3soc_batch");
    }

    int lbx_eri2d2_batch(lbx_context *ctx, long nq, const int *shells, double *out)
    {
        return blocksEntry(ctx, 2, D1_OP_D2, nq, shells, out, "lbx_eri2d2_batch");
    }

    int lbx_eri4d1(lbx_context *ctx, double *out)
    {
        return guarded([&]
                       {
                           if (!ctx || !out)
                               throw Error(LBX_ERR_INVALID, "lbx_eri4d1: null argument");
                           std::lock_guard<std::mutex> lk(ctx->mu);
                           requireSingle(*ctx, "lbx_eri4d1");
                           runD1(*ctx, 4, out);
                       });
    }

    int lbx_eri3d1(lbx_context *ctx, double *out)
    {
        return guarded([&]
                       {
                           if (!ctx || !out)
                               throw Error(LBX_ERR_INVALID, "lbx_eri3d1: null argument");
                           std::lock_guard<std::mutex> lk(ctx->mu);
                           requireSingle(*ctx, "lbx_eri3d1");
                           if (ctx->shells_aux.empty())
                               throw Error(LBX_ERR_STATE, "eri3d1: RI approximation is not enabled");
                           runD1(*ctx, 3, out);
                       });
    }

    int lbx_eri2d1(lbx_context *ctx, double *out)
    {
        return guarded([&]
                       {
                           if (!ctx || !out)
                               throw Error(LBX_ERR_INVALID, "lbx_eri2d1: null argument");
                           std::lock_guard<std::mutex> lk(ctx->mu);
                           requireSingle(*ctx, "lbx_eri2d1");
                           if (ctx->shells_aux.empty())
                               throw Error(LBX_ERR_STATE, "eri2d1: RI approximation is not enabled");
                           runD1(*ctx, 2, out);
                       });
    }

    lbx_jk_stats lbx_context::queueJK(const double *D, double *J, double *K, double schwarz_thresh,
                                      double split, bool mixed, bool prof_on)
    {
        Context &c = *this;
        const long long n = static_cast<long long>(c.nbf);
        fill(J, n * n, 0.0, c.stream);
        fill(K, n * n, 0.0, c.stream);
        lbx_jk_stats st{};
        auto launch = [&](int i, int j, int tier, int var, double lo, double hi)
        {
            const PairClass &bra = c.cls[i], &ket = c.cls[j];
            Prof *pr = nullptr;
            if (prof_on)
            {
                pr = &prof[{i, j, tier}];
                if (!pr->start)
                {
                    LBX_HIP(hipEventCreate(&pr->start));
                    LBX_HIP(hipEventCreate(&pr->stop));
                    pr->rec.la = bra.la;
                    pr->rec.lb = bra.lb;
                    pr->rec.lc = ket.la;
                    pr->rec.ld = ket.lb;
                }
                pr->rec.kernel = var;
                LBX_HIP(hipEventRecord(pr->start, c.stream));
            }
            const TaskTable &t = launchJK(i, j, var, D, J, K, lo, hi);
            st.quartets += t.quartets;
            st.flops += t.flops;
            if (isF32Variant(var))
                st.quartets_fp32 += t.quartets;
            if (pr)
            {
                LBX_HIP(hipEventRecord(pr->stop, c.stream));
                pr->pending = true;
                pr->rec.calls += 1;
                pr->rec.quartets += t.quartets;
                pr->rec.flops += t.flops;
            }
        };
        for (int i = 0; i < static_cast<int>(c.cls.size()); i++)
            for (int j = 0; j <= i; j++)
            {
                if (c.cls[i].n_pairs == 0 || c.cls[j].n_pairs == 0)
                    continue;
                if (filter_i >= 0 && (i != filter_i || j != filter_j))
                    continue;
                // the tiers [split, inf) and [schwarz, split), cut to the bound window
                const double lo0 = std::max(split, window_lo), hi0 = window_hi;
                const double lo1 = std::max(schwarz_thresh, window_lo), hi1 = std::min(split, window_hi);
                if (lo0 < hi0)
                    launch(i, j, 0, variantFor(i, j), lo0, hi0);
                if (mixed && lo1 < hi1)
                    launch(i, j, 1, variant32For(i, j), lo1, hi1);
            }
        const int gridn = static_cast<int>(std::min<long long>((n * n + 255) / 256, 65536));
        hipLaunchKernelGGL(symmetrizeKernel, dim3(gridn), dim3(256), 0, c.stream, J, static_cast<int>(n));
        hipLaunchKernelGGL(symmetrizeKernel, dim3(gridn), dim3(256), 0, c.stream, K, static_cast<int>(n));
        LBX_HIP(hipGetLastError());
        return st;
    }

    int lbx_jk(lbx_context *ctx, const double *D, double *J, double *K, double schwarz_thresh,
               int precision, double mixed_thresh, lbx_jk_stats *stats)
    {
        return guarded([&]
                       {
                           if (!ctx || !D || !J || !K)
                               throw Error(LBX_ERR_INVALID, "lbx_jk: null argument");
                           if (precision != LBX_FP64 && precision != LBX_MIXED)
                               throw Error(LBX_ERR_INVALID, "lbx_jk: precision must be LBX_FP64 or LBX_MIXED");
                           if (!(schwarz_thresh >= 0))
                               throw Error(LBX_ERR_INVALID, "lbx_jk: negative Schwarz threshold");
                           if (precision == LBX_MIXED && !(mixed_thresh >= 0))
                               throw Error(LBX_ERR_INVALID, "lbx_jk: negative mixed-precision threshold");
                           std::lock_guard<std::mutex> lk(ctx->mu);
                           Context &c = *ctx;
                           requireSingle(c, "lbx_jk");
                           ensureSchwarz(c);
                           // quartets with schwarz <= Q_ab Q_cd < split run in FP32
                           const bool mixed = precision == LBX_MIXED && mixed_thresh > schwarz_thresh;
                           const double split = mixed ? mixed_thresh : schwarz_thresh;
                           if (mixed)
                               ensureF32(c);
                           // Replays: the launch sequence of a (D, J, K, thresholds) set is
                           // captured into a HIP graph after its first (eager) build and replayed
                           // on later calls; per-class timing and the class filter run eagerly.
                           const bool graphs = ctx->graphs && !ctx->profiling && ctx->filter_i < 0 &&
                                               ctx->window_lo <= 0.0 && ctx->window_hi == kNoUpper;
                           const lbx_context::JKKey gkey{D, J, K, schwarz_thresh, split, mixed ? 1 : 0};
                           if (graphs)
                           {
                               auto it = ctx->jk_graphs.find(gkey);
                               if (it != ctx->jk_graphs.end())
                               {
                                   LBX_HIP(hipGraphLaunch(it->second.exec, c.stream));
                                   LBX_HIP(hipStreamSynchronize(c.stream));
                                   if (stats)
                                       *stats = it->second.stats;
                                   return;
                               }
                           }
                           lbx_jk_stats st = ctx->queueJK(D, J, K, schwarz_thresh, split, mixed, ctx->profiling);
                           LBX_HIP(hipStreamSynchronize(c.stream));
                           if (ctx->profiling)
                               ctx->collectProfile();
                           if (graphs)
                               ctx->captureJK(gkey, st);
                           if (stats)
                               *stats = st;
                       });
    }

    int lbx_ri_sums(lbx_context *ctx, double *out, int precision, double mixed_thresh,
                    lbx_ri_stats *stats)
    {
        return guarded([&]
                       {
                           if (!ctx || !out)
                               throw Error(LBX_ERR_INVALID, "lbx_ri_sums: null argument");
                           if (precision != LBX_FP64 && precision != LBX_MIXED)
                               throw Error(LBX_ERR_INVALID, "lbx_ri_sums: precision must be LBX_FP64 or LBX_MIXED");
                           if (precision == LBX_MIXED && !(mixed_thresh >= 0))
                               throw Error(LBX_ERR_INVALID, "lbx_ri_sums: negative mixed-precision threshold");
                           std::lock_guard<std::mutex> lk(ctx->mu);
                           Context &c = *ctx;
                           if (c.shells_aux.empty())
                               throw Error(LBX_ERR_STATE, "lbx_ri_sums: the context has no auxiliary basis");
                           const double split = precision == LBX_MIXED ? mixed_thresh : 0.0;
                           auto report = [&]
                           {
                               if (stats)
                               {
                                   stats->triplets = ctx->ri_triplets;
                                   stats->doublets = ctx->ri_doublets;
                                   stats->flops = ctx->ri_flops;
                                   stats->triplets_fp32 = ctx->ri_triplets32;
                                   stats->doublets_fp32 = ctx->ri_doublets32;
                               }
                           };
                           if (!ctx->graphs)
                           {
                               // eager launches on the context stream
                               if (split > 0.0)
                               {
                                   ensureSchwarz(c);
                                   ensureSchwarzAux(c);
                                   ensureF32(c);
                                   ensureF32Aux(c);
                               }
                               ctx->prepareRiSums(split);
                               ctx->queueRiSums(out, c.stream);
                               LBX_HIP(hipStreamSynchronize(c.stream));
                               report();
                               return;
                           }
                           if (!ctx->ri_exec || ctx->ri_out != out || ctx->ri_split != split)
                           {
                               ctx->dropRiGraph();
                               if (split > 0.0)
                               {
                                   ensureSchwarz(c);
                                   ensureSchwarzAux(c);
                                   ensureF32(c);
                                   ensureF32Aux(c);
                               }
                               ctx->prepareRiSums(split);
                               ctx->ri_split = split;
                               // capture on a private stream (default priority), replay on the
                               // context stream
                               hipStream_t cs = nullptr;
                               LBX_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
                               try
                               {
                                   LBX_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
                                   try
                                   {
                                       ctx->queueRiSums(out, cs);
                                   }
                                   catch (...)
                                   {
                                       hipGraph_t g = nullptr;
                                       (void)hipStreamEndCapture(cs, &g);
                                       if (g)
                                           (void)hipGraphDestroy(g);
                                       throw;
                                   }
                                   LBX_HIP(hipStreamEndCapture(cs, &ctx->ri_graph));
                                   LBX_HIP(hipGraphInstantiate(&ctx->ri_exec, ctx->ri_graph, nullptr, nullptr, 0));
                               }
                               catch (...)
                               {
                                   (void)hipStreamDestroy(cs);
                                   ctx->dropRiGraph();
                                   throw;
                               }
                               (void)hipStreamDestroy(cs);
                               ctx->ri_out = out;
                           }
                           LBX_HIP(hipGraphLaunch(ctx->ri_exec, c.stream));
                           LBX_HIP(hipStreamSynchronize(c.stream));
                           report();
                       });
    }

    int lbx_set_graphs(lbx_context *ctx, int on)
    {
        return guarded([&]
                       {
                           if (!ctx)
                               throw Error(LBX_ERR_INVALID, "lbx_set_graphs: null context");
                           std::lock_guard<std::mutex> lk(ctx->mu);
                           ctx->graphs = on != 0;
                           if (!ctx->graphs)
                           {
                               ctx->dropJKGraphs();
                               ctx->dropRiGraph();
                           }
                       });
    }

    int lbx_jk_grad(lbx_context *ctx, const double *D, double *grad, double schwarz_thresh,
                    int precision, double mixed_thresh, lbx_jk_stats *stats)
    {
        return guarded([&]
                       {
                           if (!ctx || !D || !grad)
                               throw Error(LBX_ERR_INVALID, "lbx_jk_grad: null argument");
                           if (!(schwarz_thresh >= 0))
                               throw Error(LBX_ERR_INVALID, "lbx_jk_grad: negative Schwarz threshold");
                           if (precision != LBX_FP64 && precision != LBX_MIXED)
                               throw Error(LBX_ERR_INVALID, "lbx_jk_grad: precision must be LBX_FP64 or LBX_MIXED");
                           if (precision == LBX_MIXED && !(mixed_thresh >= 0))
                               throw Error(LBX_ERR_INVALID, "lbx_jk_grad: negative mixed-precision threshold");
                           std::lock_guard<std::mutex> lk(ctx->mu);
                           Context &c = *ctx;
                           requireSingle(c, "lbx_jk_grad");
                           ensureSchwarz(c);
                           for (auto &pc : c.cls)
                               computeEcoeffsD1(c, pc);
                           const bool mixed = precision == LBX_MIXED && mixed_thresh > schwarz_thresh;
                           if (mixed)
                           {
                               ensureF32(c);
                               ensureF32D1(c);
                           }
                           fill(grad, 3LL * c.natoms, 0.0, c.stream);
                           double nq = 0, fl = 0, nq32 = 0;
                           for (int i = 0; i < static_cast<int>(c.cls.size()); i++)
                               for (int j = 0; j <= i; j++)
                               {
                                   const PairClass &bra = c.cls[i], &ket = c.cls[j];
                                   if (bra.n_pairs == 0 || ket.n_pairs == 0)
                                       continue;
                                   const TaskTable &t = ctx->launchGrad(i, j, D, grad,
                                                                        mixed ? mixed_thresh : schwarz_thresh);
                                   nq += t.quartets;
                                   fl += gradFlops(t, bra.la, bra.lb, ket.la, ket.lb);
                                   if (mixed)
                                   {
                                       const TaskTable &t32 = ctx->launchGrad(i, j, D, grad, schwarz_thresh,
                                                                              mixed_thresh, true);
                                       nq += t32.quartets;
                                       nq32 += t32.quartets;
                                       fl += gradFlops(t32, bra.la, bra.lb, ket.la, ket.lb);
                                   }
                               }
                           LBX_HIP(hipStreamSynchronize(c.stream));
                           if (stats)
                           {
                               stats->quartets = nq;
                               stats->flops = fl;
                               stats->quartets_fp32 = nq32;
                           }
                       });
    }

    int lbx_tune(lbx_context *ctx, const double *D, double schwarz_thresh)
    {
        return guarded([&]
                       {
                           if (!ctx)
                               throw Error(LBX_ERR_INVALID, "lbx_tune: null argument");
                           if (!ctx->atom_mol.empty())
                           {
                               // a batched context: the RI sums' kernels (D is not used)
                               std::lock_guard<std::mutex> lk(ctx->mu);
                               if (ctx->shells_aux.empty())
                                   throw Error(LBX_ERR_STATE, "lbx_tune: the batched context has no auxiliary basis");
                               ctx->tuneRiSums();
                               return;
                           }
                           if (!D)
                               throw Error(LBX_ERR_INVALID, "lbx_tune: null argument");
                           if (!(schwarz_thresh >= 0))
                               throw Error(LBX_ERR_INVALID, "lbx_tune: negative Schwarz threshold");
                           std::lock_guard<std::mutex> lk(ctx->mu);
                           Context &c = *ctx;
                           ensureSchwarz(c);
                           const long long n = static_cast<long long>(c.nbf);
                           DevBuf J(n * n), K(n * n);
                           hipEvent_t e0, e1;
                           LBX_HIP(hipEventCreate(&e0));
                           LBX_HIP(hipEventCreate(&e1));
                           try
                           {
                               for (size_t i = 0; i < c.cls.size(); i++)
                                   for (size_t j = 0; j <= i; j++)
                                   {
                                       const PairClass &bra = c.cls[i], &ket = c.cls[j];
                                       if (bra.n_pairs == 0 || ket.n_pairs == 0)
                                           continue;
                                       int best = -1, best32 = -1;
                                       double best_ms = 0, best32_ms = 0;
                                       for (int v = 0; v < N_VARIANTS; v++)
                                       {
                                           if (!kernelSlot(MODE_JK, bra.la, bra.lb, ket.la, ket.lb, v).fn)
                                               continue;
                                           if (isF32Variant(v))
                                               ensureF32(c);
                                           // first launch builds the task table and loads the code;
                                           // the timed one runs a sample of leading quartets
                                           ctx->sample_quartets = kTuneSample;
                                           ctx->launchJK(static_cast<int>(i), static_cast<int>(j), v, D, J.p, K.p,
                                                         schwarz_thresh);
                                           LBX_HIP(hipEventRecord(e0, c.stream));
                                           ctx->launchJK(static_cast<int>(i), static_cast<int>(j), v, D, J.p, K.p,
                                                         schwarz_thresh);
                                           LBX_HIP(hipEventRecord(e1, c.stream));
                                           ctx->sample_quartets = 0;
                                           LBX_HIP(hipEventSynchronize(e1));
                                           float ev = 0;
                                           LBX_HIP(hipEventElapsedTime(&ev, e0, e1));
                                           // time per quartet: the variants' samples differ
                                           const double ms = ev / std::max(ctx->sampled, 1.0);
                                           if (!isF32Variant(v) && (best < 0 || ms < best_ms))
                                           {
                                               best = v;
                                               best_ms = ms;
                                           }
                                           if (best32 < 0 || ms < best32_ms)
                                           {
                                               best32 = v;
                                               best32_ms = ms;
                                           }
                                       }
                                       ctx->dropJKGraphs();
                                       if (best >= 0)
                                           ctx->tuned[{static_cast<int>(i), static_cast<int>(j)}] = best;
                                       if (best32 >= 0)
                                           ctx->tuned32[{static_cast<int>(i), static_cast<int>(j)}] = best32;
                                   }
                           }
                           catch (...)
                           {
                               ctx->sample_quartets = 0;
                               (void)hipEventDestroy(e0);
                               (void)hipEventDestroy(e1);
                               throw;
                           }
                           (void)hipEventDestroy(e0);
                           (void)hipEventDestroy(e1);
                       });
    }

    int lbx_set_variant(lbx_context *ctx, int la, int lb, int lc, int ld, int variant)
    {
        return guarded([&]
                       {
                           if (!ctx)
                               throw Error(LBX_ERR_INVALID, "lbx_set_variant: null context");
                           if (la < lb || lc < ld || la > kMaxLMain || lc > kMaxLMain || lb < 0 || ld < 0 ||
                               pairClassIdx(la, lb) < pairClassIdx(lc, ld))
                               throw Error(LBX_ERR_INVALID, "lbx_set_variant: not a canonical class pair");
                           std::lock_guard<std::mutex> lk(ctx->mu);
                           const std::pair<int, int> key{pairClassIdx(la, lb), pairClassIdx(lc, ld)};
                           ctx->dropJKGraphs();
                           if (variant < 0)
                           {
                               ctx->tuned.erase(key);
                               ctx->tuned32.erase(key);
                               return;
                           }
                           if (variant >= N_VARIANTS || !kernelSlot(MODE_JK, la, lb, lc, ld, variant).fn)
                               throw Error(LBX_ERR_INVALID, "lbx_set_variant: class pair has no such variant");
                           if (isF32Variant(variant))
                               ctx->tuned32[key] = variant;
                           else
                               ctx->tuned[key] = variant;
                       });
    }

    int lbx_set_class_filter(lbx_context *ctx, int la, int lb, int lc, int ld)
    {
        return guarded([&]
                       {
                           if (!ctx)
                               throw Error(LBX_ERR_INVALID, "lbx_set_class_filter: null context");
                           std::lock_guard<std::mutex> lk(ctx->mu);
                           if (la == -1 && lb == -1 && lc == -1 && ld == -1)
                           {
                               ctx->filter_i = ctx->filter_j = -1;
                               return;
                           }
                           if (la < lb || lc < ld || la > kMaxLMain || lc > kMaxLMain || lb < 0 || ld < 0 ||
                               pairClassIdx(la, lb) < pairClassIdx(lc, ld))
                               throw Error(LBX_ERR_INVALID, "lbx_set_class_filter: not a canonical class pair");
                           ctx->filter_i = pairClassIdx(la, lb);
                           ctx->filter_j = pairClassIdx(lc, ld);
                       });
    }

    int lbx_ri_jk(lbx_context *ctx, const double *D, double *J, double *K)
    {
        return guarded([&]
                       {
                           if (!ctx || !D || !J)
                               throw Error(LBX_ERR_INVALID, "lbx_ri_jk: null argument");
                           long long n = 0, naux = 0;
                           hipStream_t s = nullptr;
                           {
                               std::lock_guard<std::mutex> lk(ctx->mu);
                               requireSingle(*ctx, "lbx_ri_jk");
                               if (ctx->shells_aux.empty())
                                   throw Error(LBX_ERR_STATE, "lbx_ri_jk: the context has no auxiliary basis");
                               n = static_cast<long long>(ctx->nbf);
                               naux = static_cast<long long>(ctx->nbf_aux);
                               s = ctx->stream;
                           }
                           riJK(ctx, *ctx, s, n, naux, D, J, K);
                       });
    }

    int lbx_one_electron(lbx_context *ctx, double *S, double *T, int ncharges, const double *charges, double *V)
    {
        return guarded([&]
                       {
                           if (!ctx || (!S && !T && !V))
                               throw Error(LBX_ERR_INVALID, "lbx_one_electron: null context or no output");
                           if (V && (ncharges < 0 || (ncharges > 0 && !charges)))
                               throw Error(LBX_ERR_INVALID, "lbx_one_electron: bad point charges");
                           std::lock_guard<std::mutex> lk(ctx->mu);
                           Context &c = *ctx;
                           requireSingle(c, "lbx_one_electron");
                           const size_t bytes = c.nbf * c.nbf * sizeof(double);
                           for (double *m : {S, T, V})
                               if (m)
                                   LBX_HIP(hipMemsetAsync(m, 0, bytes, c.stream));
                           double4 *d_q = nullptr;
                           if (V && ncharges > 0)
                           {
                               std::vector<double4> q(ncharges);
                               for (int k = 0; k < ncharges; k++)
                                   q[k] = double4{charges[4 * k], charges[4 * k + 1], charges[4 * k + 2],
                                                  charges[4 * k + 3]};
                               LBX_HIP(hipMalloc(&d_q, q.size() * sizeof(double4)));
                               const hipError_t e = hipMemcpyAsync(d_q, q.data(), q.size() * sizeof(double4),
                                                                   hipMemcpyHostToDevice, c.stream);
                               if (e != hipSuccess)
                               {
                                   (void)hipFree(d_q);
                                   LBX_HIP(e);
                               }
                           }
                           try
                           {
                               computeOneElectron(c, S, T, d_q, d_q ? ncharges : 0, V);
                               LBX_HIP(hipStreamSynchronize(c.stream));
                           }
                           catch (...)
                           {
                               (void)hipStreamSynchronize(c.stream);
                               (void)hipFree(d_q);
                               throw;
                           }
                           if (d_q)
                               LBX_HIP(hipFree(d_q));
                       });
    }

    int lbx_one_electron_grad(lbx_context *ctx, const double *D, const double *W, int ncharges,
                              const double *charges, double *grad, double *charge_grad)
    {
        return guarded([&]
                       {
                           if (!ctx || !D || !grad)
                               throw Error(LBX_ERR_INVALID, "lbx_one_electron_grad: null argument");
                           if (ncharges < 0 || (ncharges > 0 && !charges))
                               throw Error(LBX_ERR_INVALID, "lbx_one_electron_grad: bad point charges");
                           std::lock_guard<std::mutex> lk(ctx->mu);
                           Context &c = *ctx;
                           requireSingle(c, "lbx_one_electron_grad");
                           LBX_HIP(hipMemsetAsync(grad, 0, static_cast<size_t>(c.natoms) * 3 * sizeof(double), c.stream));
                           if (charge_grad && ncharges > 0)
                               LBX_HIP(hipMemsetAsync(charge_grad, 0, static_cast<size_t>(ncharges) * 3 * sizeof(double),
                                                      c.stream));
                           double4 *d_q = nullptr;
                           if (ncharges > 0)
                           {
                               std::vector<double4> q(ncharges);
                               for (int k = 0; k < ncharges; k++)
                                   q[k] = double4{charges[4 * k], charges[4 * k + 1], charges[4 * k + 2],
                                                  charges[4 * k + 3]};
                               LBX_HIP(hipMalloc(&d_q, q.size() * sizeof(double4)));
                               const hipError_t e = hipMemcpyAsync(d_q, q.data(), q.size() * sizeof(double4),
                                                                   hipMemcpyHostToDevice, c.stream);
                               if (e != hipSuccess)
                               {
                                   (void)hipFree(d_q);
                                   LBX_HIP(e);
                               }
                           }
                           try
                           {
                               computeOneElectronGrad(c, D, W, d_q, ncharges, grad, charge_grad);
                           }
                           catch (...)
                           {
                               (void)hipStreamSynchronize(c.stream);
                               (void)hipFree(d_q);
                               throw;
                           }
                           if (d_q)
                               LBX_HIP(hipFree(d_q));
                       });
    }

    int lbx_set_bound_window(lbx_context *ctx, double lo, double hi)
    {
        return guarded([&]
                       {
                           if (!ctx)
                               throw Error(LBX_ERR_INVALID, "lbx_set_bound_window: null context");
                           if (!(lo >= 0.0) || !(hi > lo))
                               throw Error(LBX_ERR_INVALID, "lbx_set_bound_window: need 0 <= lo < hi");
                           std::lock_guard<std::mutex> lk(ctx->mu);
                           ctx->window_lo = lo;
                           ctx->window_hi = hi;
                       });
    }

    int lbx_profile_enable(lbx_context *ctx, int on)
    {
        return guarded([&]
                       {
                           if (!ctx)
                               throw Error(LBX_ERR_INVALID, "lbx_profile_enable: null context");
                           std::lock_guard<std::mutex> lk(ctx->mu);
                           if (on)
                               for (auto &kv : ctx->prof)
                               {
                                   lbx_class_profile keep = kv.second.rec;
                                   kv.second.rec = lbx_class_profile{};
                                   kv.second.rec.la = keep.la;
                                   kv.second.rec.lb = keep.lb;
                                   kv.second.rec.lc = keep.lc;
                                   kv.second.rec.ld = keep.ld;
                                   kv.second.rec.kernel = keep.kernel;
                                   kv.second.pending = false;
                               }
                           ctx->profiling = on != 0;
                       });
    }

    int lbx_profile_read(lbx_context *ctx, lbx_class_profile *out, int capacity, int *count)
    {
        return guarded([&]
                       {
                           if (!ctx || !count || (capacity > 0 && !out))
                               throw Error(LBX_ERR_INVALID, "lbx_profile_read: null argument");
                           std::lock_guard<std::mutex> lk(ctx->mu);
                           int n = 0;
                           for (auto &kv : ctx->prof)
                           {
                               if (kv.second.rec.calls == 0)
                                   continue;
                               if (n < capacity)
                                   out[n] = kv.second.rec;
                               n++;
                           }
                           *count = n;
                       });
    }
}
