// The C-ABI of lbx (include/lbx.h): context set-up, Schwarz screening, task tables and the
// drivers that launch the quartet kernels class pair by class pair.
#include "quartet.hpp"
#include "registry.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <tuple>

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
            long long *d_prefix = nullptr;
            int *d_plo = nullptr, *d_phi = nullptr;
            double quartets = 0, flops = 0;
        };

        enum class Pairing
        {
            Unique,     // p >= q when bra and ket are the same class, all p otherwise
            All,        // every (p, q)
            Diagonal    // p == q only
        };

        // first sorted position p with qb[p] * qk < thr (qb descending)
        int firstBelow(const std::vector<double> &qb, double qk, double thr)
        {
            int a = 0, b = static_cast<int>(qb.size());
            while (a < b)
            {
                const int m = (a + b) / 2;
                if (qb[m] * qk >= thr)
                    a = m + 1;
                else
                    b = m;
            }
            return a;
        }

        constexpr double kNoUpper = std::numeric_limits<double>::infinity();

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
                    lo = (mode == Pairing::Unique && same) ? q : 0;
                    hi = bra.n_pairs;
                    const double qk = ket.Q.empty() ? 1.0 : ket.Q[ket.order[q]];
                    if (thresh > 0)
                        hi = firstBelow(qb, qk, thresh);
                    if (thresh_hi < kNoUpper)
                        lo = std::max(lo, firstBelow(qb, qk, thresh_hi));
                    hi = std::max(hi, lo);
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

        // Work estimate of a one-lane-per-quartet variant per quartet, from the classes' mean
        // primitive-pair counts: ket transform and Boys + R per primitive quartet (R once per
        // column pass), bra transform per lane-side primitive pair.
        double valuCost(const PairClass &lane, const PairClass &unif, int nc)
        {
            const double kl = lane.n_pairs ? double(lane.n_ppairs) / lane.n_pairs : 1;
            const double ku = unif.n_pairs ? double(unif.n_ppairs) / unif.n_pairs : 1;
            const double hl = nHerm(lane.la + lane.lb), hu = nHerm(unif.la + unif.lb);
            const double sl = nSph(lane.la) * nSph(lane.lb), su = nSph(unif.la) * nSph(unif.lb);
            const int L = lane.la + lane.lb + unif.la + unif.lb;
            const double passes = std::ceil(su / nc);
            return kl * ku * (hl * hu * su + passes * (60 + rFlopsHost(L))) + kl * sl * hl * su;
        }

        // Default variant before tuning: the cheaper one-lane-per-quartet form by the work
        // estimate; MFMA when no such form fits the registers.
        int defaultVariant(int mode, const PairClass &bra, const PairClass &ket)
        {
            const KernelSlot &vb = kernelSlot(mode, bra.la, bra.lb, ket.la, ket.lb, VAR_VALU_BRA);
            const KernelSlot &vk = kernelSlot(mode, bra.la, bra.lb, ket.la, ket.lb, VAR_VALU_KET);
            double cb = vb.fn ? valuCost(bra, ket, vb.nc) : 1e300;
            double ck = vk.fn ? valuCost(ket, bra, vk.nc) : 1e300;
            if (!vb.fn && !vk.fn)
                return VAR_MFMA;
            return cb <= ck ? VAR_VALU_BRA : VAR_VALU_KET;
        }

        // Default for the FP32 tier of a mixed-precision J/K: the cheaper FP32 one-lane form by
        // the same estimate, or -1 when the class has none (its quartets then stay FP64).
        int defaultVariant32(const PairClass &bra, const PairClass &ket)
        {
            const KernelSlot &vb = kernelSlot(MODE_JK, bra.la, bra.lb, ket.la, ket.lb, VAR_VALU_BRA32);
            const KernelSlot &vk = kernelSlot(MODE_JK, bra.la, bra.lb, ket.la, ket.lb, VAR_VALU_KET32);
            if (!vb.fn && !vk.fn)
                return -1;
            double cb = vb.fn ? valuCost(bra, ket, vb.nc) : 1e300;
            double ck = vk.fn ? valuCost(ket, bra, vk.nc) : 1e300;
            return cb <= ck ? VAR_VALU_BRA32 : VAR_VALU_KET32;
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
        static KernelSlot tab[3][kMaxLAux + 1][kMaxLMain + 1][kMaxLAux + 1][kMaxLMain + 1][N_VARIANTS];
        if (mode < 0 || mode > 2 || la < 0 || la > kMaxLAux || lc < 0 || lc > kMaxLAux || lb < 0 ||
            lb > kMaxLMain || ld < 0 || ld > kMaxLMain || variant < 0 || variant >= N_VARIANTS)
            throw Error(LBX_ERR_INVALID, "kernel slot out of range");
        return tab[mode][la][lb][lc][ld][variant];
    }

    // Schwarz factors Q_ab = sqrt(max |(ab|ab)|) per pair of every main class; pairs sorted by
    // descending Q so that each ket's surviving bra pairs form a prefix.
    void ensureSchwarz(Context &c)
    {
        if (c.have_schwarz)
            return;
        for (auto &pc : c.cls)
        {
            if (pc.n_pairs == 0)
                continue;
            KArgs extra{};
            extra.qout = pc.d_q;
            extra.kind = 4;
            run(c, MODE_DIAG, pc, pc, Pairing::Diagonal, true, 0.0, extra);
            pc.Q.resize(pc.n_pairs);
            LBX_HIP(hipMemcpy(pc.Q.data(), pc.d_q, pc.n_pairs * sizeof(double), hipMemcpyDeviceToHost));
            std::vector<int> order(pc.n_pairs);
            std::iota(order.begin(), order.end(), 0);
            std::stable_sort(order.begin(), order.end(),
                             [&](int x, int y) { return pc.Q[x] > pc.Q[y]; });
            std::vector<double> qs(pc.n_pairs);
            for (int i = 0; i < pc.n_pairs; i++)
                qs[i] = pc.Q[order[i]];
            pc.order = order;
            LBX_HIP(hipMemcpy(pc.d_order, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice));
            LBX_HIP(hipMemcpy(pc.d_q, qs.data(), qs.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        c.have_schwarz = true;
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

    // Queues the J/K kernel of class pair (i >= j) with the given variant on the context stream,
    // over the quartets with lo <= Q_ab Q_cd < hi.
    const TaskTable &launchJK(int i, int j, int var, const double *D, double *J, double *K,
                              double thresh, double thresh_hi = kNoUpper)
    {
        const PairClass &bra = cls[i], &ket = cls[j];
        const KernelSlot &s = slotFor(MODE_JK, bra, ket, var);
        const auto key = std::make_tuple(i, j, thresh, thresh_hi, var);
        auto it = jk_tasks.find(key);
        if (it == jk_tasks.end())
            it = jk_tasks
                         .emplace(key, s.swap ? buildTasks(*this, ket, bra, Pairing::Unique, false, thresh,
                                                           s.width, false, thresh_hi)
                                              : buildTasks(*this, bra, ket, Pairing::Unique, i == j, thresh,
                                                           s.width, true, thresh_hi))
                         .first;
        const TaskTable &t = it->second;
        KArgs a = baseArgs(*this, bra, ket, t, i == j);
        a.D = D;
        a.J = J;
        a.K = K;
        s.fn(a, stream);
        LBX_HIP(hipGetLastError());
        return t;
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
        int rc = guarded([&]
                         {
                             Context &c = *ctx;
                             LBX_HIP(hipGetDevice(&c.device));
                             c.natoms = main_basis->natoms;
                             c.shells = buildShells(*main_basis, c.nbf, c.max_l, kMaxLMain);
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
                                 computeEcoeffs(c, pc, c.shells);
                             for (auto &pc : c.aux)
                                 computeEcoeffs(c, pc, c.shells_aux);
                             LBX_HIP(hipStreamSynchronize(c.stream));
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
                           ensureSchwarz(c);
                           // quartets with schwarz <= Q_ab Q_cd < split run in FP32
                           const bool mixed = precision == LBX_MIXED && mixed_thresh > schwarz_thresh;
                           const double split = mixed ? mixed_thresh : schwarz_thresh;
                           if (mixed)
                               ensureF32(c);
                           const long long n = static_cast<long long>(c.nbf);
                           fill(J, n * n, 0.0, c.stream);
                           fill(K, n * n, 0.0, c.stream);
                           double nq = 0, fl = 0, nq32 = 0;
                           auto launch = [&](int i, int j, int tier, int var, double lo, double hi)
                           {
                               const PairClass &bra = c.cls[i], &ket = c.cls[j];
                               lbx_context::Prof *pr = nullptr;
                               if (ctx->profiling)
                               {
                                   pr = &ctx->prof[{i, j, tier}];
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
                               const TaskTable &t = ctx->launchJK(i, j, var, D, J, K, lo, hi);
                               nq += t.quartets;
                               fl += t.flops;
                               if (var >= N_FP64_VARIANTS)
                                   nq32 += t.quartets;
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
                                   launch(i, j, 0, ctx->variantFor(i, j), split, kNoUpper);
                                   if (mixed)
                                       launch(i, j, 1, ctx->variant32For(i, j), schwarz_thresh, split);
                               }
                           const int gridn = static_cast<int>(std::min<long long>((n * n + 255) / 256, 65536));
                           hipLaunchKernelGGL(symmetrizeKernel, dim3(gridn), dim3(256), 0, c.stream, J, static_cast<int>(n));
                           hipLaunchKernelGGL(symmetrizeKernel, dim3(gridn), dim3(256), 0, c.stream, K, static_cast<int>(n));
                           LBX_HIP(hipGetLastError());
                           LBX_HIP(hipStreamSynchronize(c.stream));
                           if (ctx->profiling)
                               ctx->collectProfile();
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
                           if (!ctx || !D)
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
                                       float best_ms = 0, best32_ms = 0;
                                       for (int v = 0; v < N_VARIANTS; v++)
                                       {
                                           if (!kernelSlot(MODE_JK, bra.la, bra.lb, ket.la, ket.lb, v).fn)
                                               continue;
                                           if (v >= N_FP64_VARIANTS)
                                               ensureF32(c);
                                           // first launch builds the task table and loads the code
                                           ctx->launchJK(static_cast<int>(i), static_cast<int>(j), v, D, J.p, K.p,
                                                         schwarz_thresh);
                                           LBX_HIP(hipEventRecord(e0, c.stream));
                                           ctx->launchJK(static_cast<int>(i), static_cast<int>(j), v, D, J.p, K.p,
                                                         schwarz_thresh);
                                           LBX_HIP(hipEventRecord(e1, c.stream));
                                           LBX_HIP(hipEventSynchronize(e1));
                                           float ms = 0;
                                           LBX_HIP(hipEventElapsedTime(&ms, e0, e1));
                                           if (v < N_FP64_VARIANTS && (best < 0 || ms < best_ms))
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
                                       if (best >= 0)
                                           ctx->tuned[{static_cast<int>(i), static_cast<int>(j)}] = best;
                                       if (best32 >= 0)
                                           ctx->tuned32[{static_cast<int>(i), static_cast<int>(j)}] = best32;
                                   }
                           }
                           catch (...)
                           {
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
                           if (variant < 0)
                           {
                               ctx->tuned.erase(key);
                               ctx->tuned32.erase(key);
                               return;
                           }
                           if (variant >= N_VARIANTS || !kernelSlot(MODE_JK, la, lb, lc, ld, variant).fn)
                               throw Error(LBX_ERR_INVALID, "lbx_set_variant: class pair has no such variant");
                           if (variant >= N_FP64_VARIANTS)
                               ctx->tuned32[key] = variant;
                           else
                               ctx->tuned[key] = variant;
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
