// CPU FP64 oracle: a restatement of the reference's SHARK electron-repulsion-integral path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (lible_amd/, include/) may include, link or
// call this code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
// and only as the reference results are compared against (or, in bench.py, as the timed CPU
// baseline).  Each function names the reference file:line it follows.
//
// Parity pins (what fixes each part independently of retyping the formula):
//   Boys ............ KAT calcBoysF (tests_ints.cpp:1296-1313), F_m(0)=1/(2m+1), the erf closed
//                     form of F_0, the downward-recursion identity, brute-force quadrature.
//   R integrals ..... KAT calcRInts3D (tests_ints.cpp:1315-1335), finite differences of R_000.
//   E coefficients .. KATs ecoeffsRecurrence1/2/2_n1, ecoeffsPrimitive*, ecoeffsShell
//                     (tests_ints.cpp:1150-1294); overlap sum rule; E' by finite difference.
//   Spherical trafo . KAT sphericalTrafo l=8 (tests_ints.cpp:856-871); normalised self-overlap=1.
//   ERI4/3/2 ........ KATs eri4, deployERI4Kernel, eri4Diagonal, eri3, eri2, eri2Diagonal,
//                     deployERI3/2Kernel (tests_ints.cpp:149-959); (ss|ss) closed form;
//                     8-fold symmetry; Schwarz inequality.
//   D1 integrals .... KATs deployERI4D1/ERI3D1/ERI2D1Kernel (tests_ints.cpp:961-1053); central
//                     finite differences; translational invariance.
//   D2, SOC ......... KATs deployERI2D2/ERI4SOC/ERI3SOCKernel (tests_ints.cpp:1055-1148); ERI2D2
//                     = difference quotient of ERI2D1; ERI4SOC = d_Ay d_Bz - d_Az d_By of the
//                     integrals (one-primitive shells, finite differences).
//   One-electron .... KATs overlap, kineticEnergy (C2H4/cc-pVDZ), nuclearAttraction (CO2/6-31+G)
//                     (tests_ints.cpp:167-319); unit diagonal of the normalised overlap.
//   J/K/E/gradient .. contraction of the dense oracle eri4 tensor; J, K symmetric; E invariant
//                     under rigid translation; gradient = central finite difference of E.
#pragma once

#include <array>
#include <cstddef>
#include <string>
#include <tuple>
#include <vector>

namespace orc
{
    // ---- counting / indexing (utils.hpp:20-60, utils.cpp:33-118) ----
    inline int nCart(int l) { return (l + 1) * (l + 2) / 2; }
    inline int nSph(int l) { return 2 * l + 1; }
    inline int nHerm(int l) { return l < 0 ? 0 : (l + 1) * (l + 2) * (l + 3) / 6; }
    inline int idxCart(int j, int k) { int jk = j + k; return jk * (jk + 1) / 2 + k; }
    // Hermite (t,u,v) -> position in {(0,0,0),(1,0,0),(0,1,0),(0,0,1),(2,0,0),...}
    inline int idxHerm(int t, int u, int v) { return nHerm(t + u + v - 1) + idxCart(u, v); }

    double doubleFactorial(int n);
    double purePrimitiveNorm(int l, double exp);
    std::vector<std::array<int, 3>> cartExps(int l);
    std::vector<std::array<int, 3>> hermiteIdxs(int l);

    // {sph index, cart index, coefficient}, reference ordering (spherical_trafo.cpp:9-690)
    std::vector<std::tuple<int, int, double>> sphericalTrafo(int l);

    // ---- Boys function (boys_function.cpp:6-117) ----
    struct BoysGrid
    {
        explicit BoysGrid(int n);
        int n_, max_n_, n_intervals_;
        double interval_ = 0.01, large_x_ = 30.0, thresh_ = 1e-16;
        std::vector<double> grid_;
    };
    void calcBoysF(int n, double x, const BoysGrid &grid, double *fnx);

    // ---- Hermite R integrals (rints.cpp:61-114) ----
    // R(t,u,v) for t+u+v <= l, dense (l+1)^3 row-major
    std::vector<double> calcRInts3D(int l, double p, const double *xyz, const double *fnx);

    // ---- Hermite E coefficients (ecoeffs.cpp:11-220) ----
    std::vector<double> ecoeffsRecurrence1(double one_o_2a, int l);              // (l+1)^2
    std::vector<double> ecoeffsRecurrence2(double a, double b, int la, int lb, double PA,
                                           double PB, double Kab);              // (la+1)(lb+1)(la+lb+1)
    std::vector<double> ecoeffsRecurrence2_n1(double a, double b, int la, int lb, double A,
                                              double B, const std::vector<double> &e0);
    std::array<std::vector<double>, 3> ecoeffsPrimitivePair(double a, double b, int la, int lb,
                                                            const double *xyz_a,
                                                            const double *xyz_b);
    std::array<std::vector<double>, 3> ecoeffsPrimitivePair_n1(
            double a, double b, int la, int lb, const double *xyz_a, const double *xyz_b,
            const std::array<std::vector<double>, 3> &e0);
    std::vector<std::vector<double>> ecoeffsShell(int l, const std::vector<double> &exps);

    // ---- shells and pair data (shell.hpp, shell.cpp, shell_pair_data.cpp) ----
    struct Shell
    {
        int l, z;
        size_t dim_cart, dim_sph, ofs_cart, ofs_sph, idx, idx_atom;
        std::array<double, 3> xyz;
        std::vector<double> exps, coeffs, norms, norms_prim;
    };

    struct BasisShell
    {
        int l;
        std::vector<double> exps, coeffs;
    };

    std::vector<double> calcShellNorms(int l, const std::vector<double> &coeffs,
                                       const std::vector<double> &exps,
                                       const std::vector<double> &prim_norms);

    // One structure = main shells (+ optional auxiliary shells), coordinates in bohr.
    struct Structure
    {
        std::vector<Shell> shells, shells_aux;
        size_t nbf = 0, nbf_aux = 0, natoms = 0;
        int max_l = 0, max_l_aux = 0;
        std::vector<Shell> shellsL(int l, bool aux) const;
    };

    Structure makeStructure(const std::vector<std::array<double, 3>> &coords_bohr,
                            const std::vector<std::vector<BasisShell>> &basis_atoms,
                            const std::vector<std::vector<BasisShell>> *aux_atoms);

    struct ShellPairData
    {
        ShellPairData(bool use_symm, int la, int lb, const std::vector<Shell> &sa,
                      const std::vector<Shell> &sb, double thrs = 1e-15);
        bool uses_symm;
        int la, lb;
        size_t n_pairs = 0, n_ppairs = 0;
        std::vector<double> exps, coeffs, coords, norms;
        std::vector<size_t> nrs_ppairs, offsets_prims, offsets_sph, offsets_norms,
                offsets_ecoeffs, offsets_ecoeffs_d1, atomic_idxs, shell_idxs;
    };

    struct ShellData
    {
        ShellData(int l, const std::vector<Shell> &shells);
        int l;
        size_t n_shells = 0, n_prims = 0;
        std::vector<double> coeffs, coords, exps, norms;
        std::vector<size_t> atomic_idxs, cdepths, coffsets, offsets_ecoeffs, offsets_norms,
                offsets_sph, shell_idxs;
    };

    std::vector<std::pair<int, int>> lPairsSymm(int l);
    std::vector<std::pair<int, int>> lPairsNoSymm(int l);
    std::vector<ShellPairData> shellPairData(bool use_symm, const Structure &s);
    std::vector<ShellData> shellDataAux(const Structure &s);

    // SHARK E matrices (ecoeffs.cpp:253-517)
    std::vector<double> ecoeffsSHARK(const ShellPairData &sp, bool transpose);
    std::vector<double> ecoeffsSHARK(const ShellData &sh, bool transpose);
    std::vector<double> ecoeffsD1SHARK(const ShellPairData &sp, bool transpose);

    // ---- ERI kernels (eri_kernel_funs.cpp) ----
    struct ERI4Kernel
    {
        ERI4Kernel(const ShellPairData &ab, const ShellPairData &cd);
        std::vector<double> operator()(size_t ipab, size_t ipcd, const ShellPairData &ab,
                                       const ShellPairData &cd) const;
        std::vector<double> e_bra, e_ket;
        BoysGrid grid;
    };

    struct ERI3Kernel
    {
        ERI3Kernel(const ShellPairData &ab, const ShellData &c);
        std::vector<double> operator()(size_t ipab, size_t ishc, const ShellPairData &ab,
                                       const ShellData &c) const;
        std::vector<double> e_bra, e_ket;
        BoysGrid grid;
    };

    struct ERI2Kernel
    {
        ERI2Kernel(const ShellData &a, const ShellData &b);
        std::vector<double> operator()(size_t isha, size_t ishb, const ShellData &a,
                                       const ShellData &b) const;
        std::vector<double> e_bra, e_ket;
        BoysGrid grid;
    };

    // 12 blocks [Ax,Ay,Az,Bx,..,Dz], each (Sa,Sb,Sc,Sd) row-major, concatenated
    struct ERI4D1Kernel
    {
        ERI4D1Kernel(const ShellPairData &ab, const ShellPairData &cd);
        std::vector<double> operator()(size_t ipab, size_t ipcd, const ShellPairData &ab,
                                       const ShellPairData &cd) const;
        std::vector<double> e0_bra, e1_bra, e0_ket, e1_ket;
        BoysGrid grid;
    };

    // 9 blocks [Ax..Cz]
    struct ERI3D1Kernel
    {
        ERI3D1Kernel(const ShellPairData &ab, const ShellData &c);
        std::vector<double> operator()(size_t ipab, size_t ishc, const ShellPairData &ab,
                                       const ShellData &c) const;
        std::vector<double> e0_bra, e1_bra, e0_ket;
        BoysGrid grid;
    };

    // 6 blocks [Ax..Bz]
    struct ERI2D1Kernel
    {
        ERI2D1Kernel(const ShellData &a, const ShellData &b);
        std::vector<double> operator()(size_t isha, size_t ishb, const ShellData &a,
                                       const ShellData &b) const;
        std::vector<double> e_bra, e_ket;
        BoysGrid grid;
    };

    // Second derivatives of (a|b): 36 blocks [i][j], i, j over [Ax,Ay,Az,Bx,By,Bz], each
    // (Sa,Sb), concatenated row-major in (i, j)  (eri_kernel_funs.cpp:343-433)
    struct ERI2D2Kernel
    {
        ERI2D2Kernel(const ShellData &a, const ShellData &b);
        std::vector<double> operator()(size_t isha, size_t ishb, const ShellData &a,
                                       const ShellData &b) const;
        std::vector<double> e_bra, e_ket;
        BoysGrid grid;
    };

    // Spin-orbit-type integrals: 3 blocks (x, y, z), each (Sa,Sb,Sc,Sd)
    // (eri_kernel_funs.cpp:846-974)
    struct ERI4SOCKernel
    {
        ERI4SOCKernel(const ShellPairData &ab, const ShellPairData &cd);
        std::vector<double> operator()(size_t ipab, size_t ipcd, const ShellPairData &ab,
                                       const ShellPairData &cd) const;
        std::vector<double> e1_bra, e0_ket;
        BoysGrid grid;
    };

    // 3 blocks (x, y, z), each (Sa,Sb,Sc)  (eri_kernel_funs.cpp:976-1090)
    struct ERI3SOCKernel
    {
        ERI3SOCKernel(const ShellPairData &ab, const ShellData &c);
        std::vector<double> operator()(size_t ipab, size_t ishc, const ShellPairData &ab,
                                       const ShellData &c) const;
        std::vector<double> e1_bra, e0_ket;
        BoysGrid grid;
    };

    // ---- drivers (eri4.cpp, eri3.cpp, eri2.cpp) ----
    std::vector<double> eri4(const Structure &s);          // nbf^4
    std::vector<double> eri4Diagonal(const Structure &s);  // nbf^2
    std::vector<double> eri3(const Structure &s);          // nbf*nbf*naux
    std::vector<double> eri2(const Structure &s);          // naux^2
    std::vector<double> eri2Diagonal(const Structure &s);  // naux
    // Sums of |x| and x over the dense eri3 tensor and over the dense eri2 tensor (out[0..3]),
    // accumulated block by block over the same enumeration as eri3() / eri2() with each block
    // weighted by the number of its images in the dense tensor; no tensor is formed.
    void riSums(const Structure &s, double out[4]);

    // ---- one-electron integrals (oneelints.cpp), dense nbf x nbf ----
    std::vector<double> overlap(const Structure &s);
    std::vector<double> kineticEnergy(const Structure &s);
    // attraction of the electrons to point charges (x, y, z in bohr, charge): with the nuclei
    // as charges, nuclearAttraction (oneelints.cpp:1703-1744)
    std::vector<double> externalCharges(const Structure &s, const std::vector<std::array<double, 4>> &charges);

    // ---- J/K, energy and gradient (new functionality; no reference counterpart) ----
    struct JKStats
    {
        double n_quartets = 0, flops = 0;
        double seconds_setup = 0;   // pair data + Schwarz factors
        double seconds_jk = 0;      // quartet loop
    };
    // J, K (nbf^2) from symmetric D over unique quartets with Schwarz screening; when
    // sample_stride > 1 only every sample_stride-th bra pair of each class is processed
    // (timing sample; J/K then incomplete).
    JKStats jk(const Structure &s, const double *D, double *J, double *K, double schwarz_thresh,
               int sample_stride);
    // dE/dR for E = 1/2 tr(DJ) - 1/4 tr(DK), D held fixed, grad: natoms*3
    JKStats jkGradient(const Structure &s, const double *D, double *grad, double schwarz_thresh);
}
