// lible::ints — the C++ interface of the reference's two-electron-integral path
// (src/lible/ints/ints.hpp:165-179, shell.hpp:10-34, structure.hpp:12-164), implemented over the
// lbx C-ABI (include/lbx.h) on the current HIP device.  A caller of the reference's eri4 / eri3 /
// eri2 / eri4Diagonal / eri2Diagonal keeps its code: same types, same dense row-major output
// layout (AO order of the shells, spherical m order 0, +1, -1, +2, -2, ...).  Added: the
// Schwarz-screened Coulomb/exchange build coulombExchange().
#pragma once

#include <array>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../types.hpp"
#include "../vectormd.hpp"

struct lbx_context;

namespace lible::ints
{
    namespace detail
    {
        struct Tensors;      // ties shell data and kernel functors to their structure
        struct BlockCache;   // a kernel functor's blocks, computed on the GPU in chunks
    }

    /// Angular momentum, exponents and raw contraction coefficients of one shell
    /// (shell.hpp:10-19).
    struct BasisShell
    {
        int l_{};
        std::vector<double> exps_;
        std::vector<double> coeffs_;
    };

    using basis_shells_t = std::vector<BasisShell>;

    /// Atomic number and the shells of one atom (shell.hpp:24-34).
    struct BasisAtom
    {
        int atomic_nr_{};
        basis_shells_t basis_shells_;
    };

    using basis_atoms_t = std::vector<BasisAtom>;

    /// A normalised shell of atomic orbitals of a structure (shell.hpp:36-67).
    struct Shell
    {
        int l_{};
        int z_{};                 // atomic number of the shell's atom (0 for a ghost atom)
        size_t dim_cart_{};
        size_t dim_sph_{};
        size_t ofs_cart_{};
        size_t ofs_sph_{};
        size_t idx_{};            // position in the list of all shells (of its basis)
        size_t idx_atom_{};
        std::array<double, 3> xyz_coords_{};   // bohr
        std::vector<double> exps_;
        std::vector<double> coeffs_;            // raw contraction coefficients
        std::vector<double> norms_;             // spherical AO norms
        std::vector<double> norms_prim_;        // primitive norms
        /// (not in the reference) the structure the shell belongs to, so that shell-pair data
        /// built from shells can feed the kernel functors
        std::shared_ptr<detail::Tensors> tensors_;
        bool aux_ = false;                      // (not in the reference) auxiliary basis shell
    };

    /// Bohr <-> angstrom (defs.hpp:9-12).
    constexpr double _bohr_to_ang_ = 0.5291772083;
    constexpr double _ang_to_bohr_ = 1.0 / _bohr_to_ang_;

    /// Directories holding the basis-set text files (ints.hpp:333-362).  Files are found as
    /// <dir>/**/<name>.txt with the name lower-cased.
    class BasisPaths
    {
    public:
        static std::string getMainBasisSetsPath();
        static std::string getAuxBasisSetsPath();
        static void setMainBasisSetsPath(const std::string &path);
        static void setAuxBasisSetsPath(const std::string &path);
    };

    /// Shells of one element from a named basis (basis_sets.cpp:214-304).
    BasisAtom basisForAtom(int atomic_nr, const std::string &basis_set);
    BasisAtom basisForAtomAux(int atomic_nr, const std::string &aux_basis_set);
    basis_atoms_t basisForAtoms(const std::vector<int> &atomic_nrs, const std::string &basis_set);
    basis_atoms_t basisForAtomsAux(const std::vector<int> &atomic_nrs,
                                   const std::string &aux_basis_set);

    /// Atoms, coordinates and basis (structure.hpp:12-164).  Coordinates are given in angstrom
    /// and stored in bohr.  The device-side data (normalised shells, pair lists, E matrices,
    /// Schwarz factors) is built on first use and shared by copies of the Structure.
    class Structure
    {
    public:
        Structure() = default;
        Structure(const std::string &basis_set, const std::vector<int> &atomic_nrs,
                  const std::vector<std::array<double, 3>> &coords_angstrom);
        Structure(const std::string &basis_set, const std::string &basis_set_aux,
                  const std::vector<int> &atomic_nrs,
                  const std::vector<std::array<double, 3>> &coords_angstrom);
        Structure(const basis_atoms_t &basis_set, const std::vector<int> &atomic_nrs,
                  const std::vector<std::array<double, 3>> &coords_angstrom);
        Structure(const basis_atoms_t &basis_set, const basis_atoms_t &basis_set_aux,
                  const std::vector<int> &atomic_nrs,
                  const std::vector<std::array<double, 3>> &coords_angstrom);

        /// Ghost atoms carry basis functions and no nucleus; their shells follow those of the
        /// real atoms in AO order (structure.cpp:117-243, shell.cpp:137-194).
        Structure(const std::string &basis_set, const std::string &basis_set_ghost,
                  const std::vector<int> &atomic_nrs, const std::vector<int> &atomic_nrs_ghost,
                  const std::vector<std::array<double, 3>> &coords_angstrom,
                  const std::vector<std::array<double, 3>> &coords_angstrom_ghost);
        Structure(const std::string &basis_set, const std::string &basis_set_ghost,
                  const std::string &basis_set_aux, const std::string &basis_set_aux_ghost,
                  const std::vector<int> &atomic_nrs, const std::vector<int> &atomic_nrs_ghost,
                  const std::vector<std::array<double, 3>> &coords_angstrom,
                  const std::vector<std::array<double, 3>> &coords_angstrom_ghost);
        Structure(const basis_atoms_t &basis_set, const basis_atoms_t &basis_set_ghost,
                  const std::vector<int> &atomic_nrs, const std::vector<int> &atomic_nrs_ghost,
                  const std::vector<std::array<double, 3>> &coords_angstrom,
                  const std::vector<std::array<double, 3>> &coords_angstrom_ghost);
        Structure(const basis_atoms_t &basis_set, const basis_atoms_t &basis_set_ghost,
                  const basis_atoms_t &basis_set_aux, const basis_atoms_t &basis_set_ghost_aux,
                  const std::vector<int> &atomic_nrs, const std::vector<int> &atomic_nrs_ghost,
                  const std::vector<std::array<double, 3>> &coords_angstrom,
                  const std::vector<std::array<double, 3>> &coords_angstrom_ghost);

        bool getUseRI() const { return use_ri_; }
        int getMaxL() const { return max_l_; }
        int getMaxLAux() const { return max_l_aux_; }
        size_t getDimAO() const { return dim_ao_; }
        size_t getDimAOAux() const { return dim_ao_aux_; }
        size_t getNAtoms() const { return coords_.size(); }
        int getZ(size_t iatom) const { return atomic_nrs_.at(iatom); }
        std::array<double, 3> getCoordsAtom(size_t iatom) const { return coords_.at(iatom); }

        /// Normalised shells in AO order, all or of one angular momentum, of the main and the
        /// auxiliary basis (structure.hpp:97-106).
        std::vector<Shell> getShells() const;
        std::vector<Shell> getShellsAux() const;
        std::vector<Shell> getShellsL(int l) const;
        std::vector<Shell> getShellsLAux(int l) const;

        /// The lbx context of this structure (created on the current device on first use).
        lbx_context *context() const;
        /// The link from shell data and kernel functors back to this structure (shared by copies
        /// of the Structure).
        std::shared_ptr<detail::Tensors> tensors() const;
        /// Normalised shells in AO order: l, atom, AO offset, centre (bohr), exponents, raw
        /// coefficients, primitive norms, spherical norms (constructShells, shell.cpp:89-135).
        struct ShellInfo
        {
            int l;
            size_t atom, ofs_sph, ofs_cart;
            std::array<double, 3> xyz;
            std::vector<double> exps, coeffs, norms_prim, norms;
        };
        const std::vector<ShellInfo> &shells(bool aux) const;

    private:
        void init(basis_atoms_t main, const basis_atoms_t *aux);

        bool use_ri_ = false;
        int max_l_ = 0, max_l_aux_ = 0;
        size_t dim_ao_ = 0, dim_ao_aux_ = 0;
        std::vector<int> atomic_nrs_, atomic_nrs_ghost_;
        std::vector<std::array<double, 3>> coords_, coords_ghost_;   // bohr
        basis_atoms_t basis_, basis_aux_;                  // real atoms, then ghost atoms
        mutable std::shared_ptr<lbx_context> ctx_;
        mutable std::shared_ptr<detail::Tensors> tensors_;
        mutable std::shared_ptr<std::vector<ShellInfo>> shells_, shells_aux_;
    };

    /// Contiguous data of the shells of one angular momentum (shell_pair_data.hpp:7-43): the
    /// reference's fields, in its order and meaning.
    struct ShellData
    {
        int l_{};
        size_t n_shells_{};
        size_t n_primitives_{};
        std::vector<double> coeffs_;   // contraction coefficients times primitive norms
        std::vector<double> coords_;
        std::vector<double> exps_;
        std::vector<double> norms_;
        std::vector<size_t> atomic_idxs_;
        std::vector<size_t> cdepths_;
        std::vector<size_t> coffsets_;
        std::vector<size_t> offsets_ecoeffs_;
        std::vector<size_t> offsets_norms_;
        std::vector<size_t> offsets_sph_;
        std::vector<size_t> shell_idxs_;
        /// (not in the reference) the tensors the kernel functors read
        std::shared_ptr<detail::Tensors> tensors_;
    };

    /// Contiguous data of the screened shell pairs of one class (shell_pair_data.hpp:45-119),
    /// primitive pairs kept where |N_a d_a N_b d_b exp(-ab/(a+b) R_AB^2)| >= primitives_thrs_.
    struct ShellPairData
    {
        ShellPairData() = default;
        /// The pairs of shells_a (angular momentum la) with shells_b (lb) (shell_pair_data.hpp:
        /// 50-57): with use_symm and la == lb only b <= a; the shells must come from one
        /// structure's getShells / getShellsL.
        ShellPairData(bool use_symm, int la, int lb, const std::vector<Shell> &shells_a,
                      const std::vector<Shell> &shells_b, double primitives_thrs = 1e-15);

        bool uses_symm_{};
        double primitives_thrs_{};
        int la_{};
        int lb_{};
        size_t n_pairs_{};
        size_t n_pairs_total_{};
        size_t n_ppairs_{};
        size_t n_ppairs_total_{};
        std::vector<double> exps_;     // (a, b) per primitive pair
        std::vector<double> coeffs_;   // (N_a d_a, N_b d_b) per primitive pair
        std::vector<double> coords_;   // (A, B) per pair
        std::vector<double> norms_;    // spherical norms of a, then b, per pair
        std::vector<size_t> nrs_ppairs_;
        std::vector<size_t> offsets_primitives_;
        std::vector<size_t> offsets_sph_;   // (a, b) per pair
        std::vector<size_t> offsets_cart_;  // (a, b) per pair
        std::vector<size_t> offsets_norms_;
        std::vector<size_t> offsets_ecoeffs_;
        std::vector<size_t> offsets_ecoeffs_deriv1_;
        std::vector<size_t> atomic_idxs_;   // (a, b) per pair
        std::vector<size_t> shell_idxs_;    // (a, b) per pair
        /// (not in the reference) the tensors the kernel functors read
        std::shared_ptr<detail::Tensors> tensors_;

        std::pair<int, int> getLPair() const { return {la_, lb_}; }
    };

    /// Shell-pair data of every class of the structure (shell_pair_data.cpp:294-309): with
    /// use_symm the classes la >= lb and, within la == lb, pairs b <= a; otherwise every
    /// (la, lb) and every ordered pair.
    std::vector<ShellPairData> shellPairData(bool use_symm, const Structure &structure);
    /// Shell data of the auxiliary basis, one entry per l = 0 .. max l (shell_pair_data.cpp).
    std::vector<ShellData> shellDataAux(const Structure &structure);

    /// Per-block kernels of the reference (eri_kernels.hpp:60-170), same constructors and call
    /// operators.  Each computes only its own class pair's blocks: a call evaluates, on the GPU
    /// (lbx_eri*_batch), the chunk of consecutive first-side indices (bra pairs, or auxiliary
    /// shells) that holds it against every second-side index, about 32 MiB of blocks, and
    /// keeps the last 8 chunks on the host, so callers that walk the first side in order make
    /// one GPU call per chunk.  Construction is not thread-safe, the call operators are.  The
    /// data must come from shellPairData() / shellDataAux().
    class ERI4Kernel
    {
    public:
        ERI4Kernel(const ShellPairData &sp_data_ab, const ShellPairData &sp_data_cd);
        vec4d operator()(size_t ipair_ab, size_t ipair_cd, const ShellPairData &sp_data_ab,
                         const ShellPairData &sp_data_cd) const;
    private:
        std::shared_ptr<detail::BlockCache> c_;
    };

    class ERI3Kernel
    {
    public:
        ERI3Kernel(const ShellPairData &sp_data_ab, const ShellData &sh_data_c);
        vec3d operator()(size_t ipair_ab, size_t ishell_c, const ShellPairData &sp_data_ab,
                         const ShellData &sh_data_c) const;
    private:
        std::shared_ptr<detail::BlockCache> c_;
    };

    class ERI2Kernel
    {
    public:
        ERI2Kernel(const ShellData &sh_data_a, const ShellData &sh_data_b);
        vec2d operator()(size_t ishell_a, size_t ishell_b, const ShellData &sh_data_a,
                         const ShellData &sh_data_b) const;
    private:
        std::shared_ptr<detail::BlockCache> c_;
    };

    /// 12 blocks: derivatives with respect to A, B, C, D (x, y, z each).
    class ERI4D1Kernel
    {
    public:
        ERI4D1Kernel(const ShellPairData &sp_data_ab, const ShellPairData &sp_data_cd);
        std::array<vec4d, 12> operator()(size_t ipair_ab, size_t ipair_cd,
                                         const ShellPairData &sp_data_ab,
                                         const ShellPairData &sp_data_cd) const;
    private:
        std::shared_ptr<detail::BlockCache> c_;
    };

    /// 9 blocks: A, B, C.
    class ERI3D1Kernel
    {
    public:
        ERI3D1Kernel(const ShellPairData &sp_data_ab, const ShellData &sh_data_c);
        std::array<vec3d, 9> operator()(size_t ipair_ab, size_t ishell_c,
                                        const ShellPairData &sp_data_ab,
                                        const ShellData &sh_data_c) const;
    private:
        std::shared_ptr<detail::BlockCache> c_;
    };

    /// 6 blocks: A, B.
    class ERI2D1Kernel
    {
    public:
        ERI2D1Kernel(const ShellData &sh_data_a, const ShellData &sh_data_b);
        std::array<vec2d, 6> operator()(size_t ishell_a, size_t ishell_b, const ShellData &sh_data_a,
                                        const ShellData &sh_data_b) const;
    private:
        std::shared_ptr<detail::BlockCache> c_;
    };

    /// 6 x 6 blocks: second derivatives with respect to [Ax, Ay, Az, Bx, By, Bz] twice
    /// (eri_kernels.hpp:171-186).
    class ERI2D2Kernel
    {
    public:
        ERI2D2Kernel(const ShellData &sh_data_a, const ShellData &sh_data_b);
        arr2d<vec2d, 6, 6> operator()(size_t ishell_a, size_t ishell_b, const ShellData &sh_data_a,
                                      const ShellData &sh_data_b) const;
    private:
        std::shared_ptr<detail::BlockCache> c_;
    };

    /// 3 blocks x, y, z of the spin-orbit-type integrals (eri_kernels.hpp:188-204).
    class ERI4SOCKernel
    {
    public:
        ERI4SOCKernel(const ShellPairData &sp_data_ab, const ShellPairData &sp_data_cd);
        std::array<vec4d, 3> operator()(size_t ipair_ab, size_t ipair_cd,
                                        const ShellPairData &sp_data_ab,
                                        const ShellPairData &sp_data_cd) const;
    private:
        std::shared_ptr<detail::BlockCache> c_;
    };

    /// 3 blocks x, y, z (eri_kernels.hpp:206-222).
    class ERI3SOCKernel
    {
    public:
        ERI3SOCKernel(const ShellPairData &sp_data_ab, const ShellData &sh_data_c);
        std::array<vec3d, 3> operator()(size_t ipair_ab, size_t ishell_c,
                                        const ShellPairData &sp_data_ab,
                                        const ShellData &sh_data_c) const;
    private:
        std::shared_ptr<detail::BlockCache> c_;
    };

    /// Tabulated Boys function (boys_function.hpp:10-59): F_0 .. F_{n+6} on x = 0, 0.01, .., 30,
    /// the table behind the GPU kernels' Taylor expansions; n <= 17.
    class BoysGrid
    {
    public:
        BoysGrid() = default;
        explicit BoysGrid(int n);
        double getLargeX() const { return 30.0; }
        double getIntervalSize() const { return 0.01; }
        int getN() const { return n_; }
        int getMaxN() const { return max_n_; }
        /// rows x = 0, 0.01, .. (getMaxN() + 1 values each)
        const std::vector<double> &getFnxGrid() const { return fnx_grid_; }

    private:
        int n_ = 0, max_n_ = 0;
        std::vector<double> fnx_grid_;
    };

    /// F_0(x) .. F_n(x) (boys_function.cpp:71-117): 6-term Taylor expansion around the grid
    /// point below x and downward recursion; closed forms at x = 0 and beyond x = 30.  A host
    /// function for scalar callers; lbx_boys evaluates arrays of points on the GPU.
    std::vector<double> calcBoysF(int n, double x, const BoysGrid &boys_grid);

    /// Names of the basis sets found in the main / auxiliary basis directories
    /// (basis_sets.cpp:367-382: the reference's catalogue of the files it ships).
    std::set<std::string> availableBasisSets();
    std::set<std::string> availableBasisSetsAux();

    /// Per class pair (la lb | lc ld) of the unique quartets: count and GPU time of their
    /// integrals, then the sum of |(mn|kl)| over them, printed to stdout (eri4.cpp:97-150).
    void eri4Benchmark(const Structure &structure);

    /// Overlap integrals, nbf x nbf (oneelints.cpp:169-197).
    vec2d overlap(const Structure &structure);
    /// Kinetic-energy integrals, nbf x nbf (oneelints.cpp:348-376).
    vec2d kineticEnergy(const Structure &structure);
    /// Attraction of an electron to the nuclei (charges Z at the real atoms), nbf x nbf
    /// (oneelints.cpp:1703-1744).
    vec2d nuclearAttraction(const Structure &structure);
    /// Attraction of an electron to point charges {x, y, z, q} (bohr), nbf x nbf
    /// (oneelints.cpp:1087-1120).
    vec2d externalCharges(const std::vector<std::array<double, 4>> &point_charges,
                          const Structure &structure);

    /// (mn|kl), dense nbf^4 (eri4.cpp:38-95).
    vec4d eri4(const Structure &structure);
    /// (mn|mn), nbf x nbf (eri4.cpp:152-174).
    vec2d eri4Diagonal(const Structure &structure);
    /// (mn|P), nbf x nbf x naux (eri3.cpp:7-46).
    vec3d eri3(const Structure &structure);
    /// (P|Q), naux x naux (eri2.cpp:27-72).
    vec2d eri2(const Structure &structure);
    /// (P|P), naux (eri2.cpp:74-96).
    std::vector<double> eri2Diagonal(const Structure &structure);

    /// J_mn = sum_kl (mn|kl) D_kl and K_mk = sum_nl (mn|kl) D_nl over Schwarz-screened unique
    /// quartets (no reference counterpart).  D symmetric nbf x nbf.
    std::pair<vec2d, vec2d> coulombExchange(const Structure &structure, const vec2d &D,
                                            double schwarz_threshold = 1e-12);

    /// The backward of coulombExchange's energy: the nuclear gradient dE/dR (atoms x 3, Eh/bohr,
    /// real atoms then ghost atoms) of E = 1/2 tr(DJ) - 1/4 tr(DK) at fixed D, over the same
    /// Schwarz-screened unique quartets (no reference counterpart: the reference supplies the
    /// derivative integrals, ERI4D1Kernel, eri_kernels.hpp:113-131; here they are contracted
    /// with the density inside the kernel and never stored).  D symmetric nbf x nbf.
    vec2d coulombExchangeGradient(const Structure &structure, const vec2d &D,
                                  double schwarz_threshold = 1e-12);
}
