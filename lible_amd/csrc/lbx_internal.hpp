// Host-side data model of the lbx library: shells, pair classes and the device arrays behind
// an lbx_context.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lbx.h"
#include "lbx_common.hpp"

namespace lbx
{
    struct Error : std::runtime_error
    {
        int code;
        Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
    };

#define LBX_HIP(call)                                                                              \
    do                                                                                             \
    {                                                                                              \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            throw ::lbx::Error(e_ == hipErrorOutOfMemory ? LBX_ERR_NOMEM : LBX_ERR_HIP,             \
                               std::string(#call) + ": " + hipGetErrorString(e_));                 \
    } while (0)

    // One normalised contracted shell (the reference's Shell, shell.hpp:36-67).
    struct HShell
    {
        int l = 0, atom = 0;
        size_t ao = 0, idx = 0;
        std::array<double, 3> xyz{};
        std::vector<double> exps, coeffs;   // raw coefficients
        std::vector<double> pnorms;         // primitive norms
        std::vector<double> norms;          // contracted spherical norms (2l+1)
    };

    // Device view of one pair class, passed by value to kernels.
    struct DPairs
    {
        const int *ao_a, *ao_b;     // AO offsets of the two shells, per pair
        const int *sh_a, *sh_b;     // shell indices, per pair (a == b detects diagonal pairs)
        const int *poff, *pcnt;     // first primitive pair and count, per pair
        const double4 *pp;          // per primitive pair: (p, Px, Py, Pz)
        const double *eb;           // E^T [ppair][tuv][mu nu]           (bra use)
        const double *ek;           // (-1)^{t+u+v} E^T [ppair][tuv][mu nu] (ket use)
        const int *order;           // sorted position -> pair index
        const double *q;            // Schwarz factor per sorted position
        int n;
        const float *eb32, *ek32;   // eb, ek rounded to FP32 (mixed precision; may be null)
        // nuclear gradient (may be null until the first gradient call)
        const int *atom_a, *atom_b;   // atom of each shell, per pair
        const double *ra;             // a / (a + b) per primitive pair
        const double *e1b, *e1k;      // d eb / d(A - B)_x, d ek / d(A - B)_x: [ppair][x][tuv][mu nu]
        const int *mol;               // molecule of each pair (batched contexts; may be null)
        const float *e1b32, *e1k32;   // e1b, e1k rounded to FP32 (mixed-precision gradient)
    };

    // Per primitive pair input of the E-matrix kernels.
    struct PrimIn
    {
        double a, b, dadb;   // exponents, product of the normalised contraction coefficients
        int pair;
    };

    // Pairs of shells (la >= lb), or single shells (lb = 0, single = true: an auxiliary shell
    // paired with a unit s function at the same centre, so 3- and 2-centre integrals run
    // through the 4-centre kernels).
    struct PairClass
    {
        int la = 0, lb = 0;
        bool single = false;
        int n_pairs = 0, n_ppairs = 0, kmax = 0;
        // host
        std::vector<int> ao_a, ao_b, sh_a, sh_b, atom_a, atom_b, poff, pcnt;
        std::vector<int> mol;              // molecule per pair (batched contexts)
        std::vector<int> prim_x, prim_y;   // surviving primitive pair -> primitive of a, of b
        std::vector<double> Q;          // Schwarz factor per pair (pair order)
        std::vector<int> order;         // sorted by Q descending
        // mean over blocks of 64 consecutive sorted pairs of the block's largest primitive-pair
        // count: the primitive loop a wave of a one-lane form runs when its lanes take this class
        // (a wave waits for its longest lane); 0 until the pairs are sorted
        double klane = 0.0;
        // device
        int *d_ao_a = nullptr, *d_ao_b = nullptr, *d_sh_a = nullptr, *d_sh_b = nullptr;
        int *d_poff = nullptr, *d_pcnt = nullptr, *d_order = nullptr;
        double4 *d_pp = nullptr;
        double *d_eb = nullptr, *d_ek = nullptr, *d_q = nullptr;
        float *d_eb32 = nullptr, *d_ek32 = nullptr;   // made on first mixed-precision use
        // E-matrix inputs kept for the derivative matrices
        PrimIn *d_prim = nullptr;
        double *d_centers = nullptr, *d_norms = nullptr;
        // made on first gradient use
        int *d_atom_a = nullptr, *d_atom_b = nullptr;
        double *d_ra = nullptr, *d_e1b = nullptr, *d_e1k = nullptr;
        int *d_mol = nullptr;
        float *d_e1b32 = nullptr, *d_e1k32 = nullptr;   // made on first mixed gradient

        // doubles in each of d_eb, d_ek
        size_t eCount() const
        {
            return static_cast<size_t>(n_ppairs) * nHerm(la + lb) * nSph(la) * nSph(lb);
        }

        DPairs view() const
        {
            return DPairs{d_ao_a, d_ao_b, d_sh_a, d_sh_b, d_poff, d_pcnt, d_pp, d_eb, d_ek,
                          d_order, d_q, n_pairs, d_eb32, d_ek32, d_atom_a, d_atom_b, d_ra,
                          d_e1b, d_e1k, d_mol, d_e1b32, d_e1k32};
        }
    };

    struct Context
    {
        int device = 0;
        hipStream_t stream = nullptr;
        std::mutex mu;
        int natoms = 0;
        size_t nbf = 0, nbf_aux = 0;
        int max_l = 0, max_l_aux = -1;
        std::vector<HShell> shells, shells_aux;
        std::vector<PairClass> cls;      // index pairClassIdx(la, lb)
        std::vector<PairClass> aux;      // index l
        double *d_boys = nullptr;        // Boys table kBoysRows x kBoysCols
        bool have_schwarz = false, have_schwarz_aux = false;
        // batched contexts: molecule of each atom (atoms grouped by molecule); pairs only form
        // within a molecule.  Empty for an ordinary (one-molecule) context.
        std::vector<int> atom_mol;
        int nmol = 1;
        std::vector<void *> allocations;
        // density fitting (ri.hip): the fitted 3-index tensor W (nbf^2 naux doubles, P-major),
        // built on the first lbx_ri_jk call and kept, and the exchange work buffer (same size,
        // made on the first call with K); both in allocations
        double *ri_w = nullptr, *ri_t = nullptr;
        bool ri_ready = false;
        std::mutex ri_mu;

        template <class T>
        T *alloc(size_t n)
        {
            void *p = nullptr;
            if (n == 0)
                n = 1;
            LBX_HIP(hipMalloc(&p, n * sizeof(T)));
            allocations.push_back(p);
            return static_cast<T *>(p);
        }

        template <class T>
        T *upload(const std::vector<T> &v)
        {
            T *d = alloc<T>(v.size());
            if (!v.empty())
                LBX_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
            return d;
        }

        void release();
    };

    // host_setup.cpp
    std::vector<HShell> buildShells(const lbx_basis &b, size_t &nbf, int &max_l, int max_l_allowed);
    std::vector<double> boysTable();
    struct SphEntry
    {
        int sph, cart;
        double val;
    };
    const std::vector<SphEntry> &sphTable(int l);
    void buildPairClasses(Context &c);
    void uploadPairClass(Context &c, PairClass &pc, const std::vector<HShell> &shells,
                         bool single);

    // ecoef.hip
    void computeEcoeffs(Context &c, PairClass &pc, const std::vector<HShell> &shells);
    // d E / d (A - B)_x,y,z and the per-pair atoms and a/p ratios for the nuclear gradient
    // (once per pair class)
    void computeEcoeffsD1(Context &c, PairClass &pc);
    // dense overlap S, kinetic energy T and attraction V to ncharges point charges (device,
    // (x, y, z, q) each) over the main basis, added into zeroed nbf x nbf device buffers (any
    // may be null)
    void computeOneElectron(Context &c, double *S, double *T, const double4 *charges, int ncharges, double *V);
    // nuclear gradient of tr(D (T + V)) - tr(W S) into grad (natoms x 3; W may be null) and the
    // derivatives with respect to the charge positions into charge_grad (ncharges x 3, may be
    // null); both added into
    void computeOneElectronGrad(Context &c, const double *D, const double *W, const double4 *charges, int ncharges,
                                double *grad, double *charge_grad);

    // schwarz / task tables (capi.cpp)
    void ensureSchwarz(Context &c);

    // d1.hip: dense first-derivative tensors of one class pair
    struct D1Args
    {
        DPairs bra, ket;
        int la, lb, lc, ld;          // classes; a single-shell side has lb (ld) = 0
        int bra_single, ket_single;  // single-shell sides (no E', a/p = 1)
        const int *tasks;            // (bra pair, ket pair) per task, pair indices
        long long ntasks;
        const double *boys;
        double *out;
        int kind;                    // 4: (12,n,n,n,n); 3: (9,n,n,naux); 2: (6,naux,naux)
        long long nbf, naux;
        // Block output (lbx_eri*_batch): tasks are (bra pair, ket pair, flags, slot) and task t
        // writes its block(s) to out + slot * ncomp * Sa Sb Sc Sd in the caller's shell order
        // (flags bit 0: the caller's bra shells are the pair's (b, a); bit 1: the ket's (d, c)).
        int blocks;
        int op;                      // D1_OP_*: what each entry holds
    };
    enum D1Op
    {
        D1_OP_DERIV = 0,   // first derivatives (12 / 9 / 6 components)
        D1_OP_PLAIN = 1,   // the integrals themselves (one component)
        D1_OP_SOC = 2,     // the reference's spin-orbit-type integrals (3 components; eri4, eri3)
        D1_OP_D2 = 3       // second derivatives of (a|b) (36 components; eri2)
    };
    void launchD1(const D1Args &a, hipStream_t s);

    // ri.hip: density-fitted J (and K when K != nullptr) from the fitted 3-index tensor that the
    // first call builds from the dense eri3 / eri2 of the context and keeps (called without the
    // context's lock, since it calls lbx_eri3 / lbx_eri2; serialised by c.ri_mu)
    void riJK(lbx_context *ctx, Context &c, hipStream_t stream, long long nbf, long long naux, const double *D,
              double *J, double *K);
}
