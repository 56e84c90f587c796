// lible::ints C++ interface (include/lible/ints/ints.hpp) over the lbx C-ABI.  Host code only:
// basis-file reading, Structure bookkeeping, and the copies of the device results into the
// row-major VectorMD tensors the reference's functions return.
#include <lible/ints/ints.hpp>

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <set>
#include <cstdio>
#include <chrono>
#include <cctype>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <list>
#include <mutex>
#include <sstream>

#include "lbx.h"
#include "lbx_internal.hpp"

namespace fs = std::filesystem;

namespace lible::ints
{
    namespace
    {
        std::mutex g_path_mutex;

        std::string envOr(const char *name, const char *fallback)
        {
            const char *v = std::getenv(name);
            return v ? std::string(v) : std::string(fallback);
        }

        std::string &mainPath()
        {
            static std::string p = envOr("LIBLE_BASIS_PATH", "");
            return p;
        }

        std::string &auxPath()
        {
            static std::string p = envOr("LIBLE_AUX_BASIS_PATH", "");
            return p;
        }

        std::string lower(std::string s)
        {
            std::transform(s.begin(), s.end(), s.begin(),
                           [](unsigned char c) { return static_cast<char>(std::tolower(c)); });
            return s;
        }

        /// <root>/<name>.txt, else the first <root>/**/<name>.txt (the reference keeps one
        /// sub-directory per basis family, basis_sets.cpp:318-364).
        std::string findBasisFile(const std::string &root, const std::string &basis_set)
        {
            const std::string want = lower(basis_set);
            if (root.empty())
                throw std::runtime_error("basis set directory not set: call "
                                         "BasisPaths::setMainBasisSetsPath / "
                                         "setAuxBasisSetsPath or set LIBLE_BASIS_PATH / "
                                         "LIBLE_AUX_BASIS_PATH");
            const fs::path direct = fs::path(root) / (want + ".txt");
            if (fs::exists(direct))
                return direct.string();
            if (fs::is_directory(root))
                for (const auto &e : fs::recursive_directory_iterator(root))
                {
                    if (!e.is_regular_file())
                        continue;
                    std::string name = e.path().filename().string();
                    if (name.substr(0, name.find('.')) == want)
                        return e.path().string();
                }
            throw std::runtime_error("The requested basis set " + basis_set +
                                     " could not be found under " + root);
        }

        /// Element block "element=Z" ... "end=Z"; inside, "l=<l> ..." starts a shell and every
        /// other line is "<exponent> <coefficient>" (basis_sets.cpp:214-278).
        BasisAtom readAtom(int atomic_nr, const std::string &basis_set, const std::string &path)
        {
            std::ifstream in(path);
            if (!in)
                throw std::runtime_error("cannot open basis file " + path);
            const std::string begin = "element=" + std::to_string(atomic_nr);
            const std::string end = "end=" + std::to_string(atomic_nr);
            BasisAtom atom;
            atom.atomic_nr_ = atomic_nr;
            bool inside = false, found = false;
            std::string line;
            while (std::getline(in, line))
            {
                if (!line.empty() && line.back() == '\r')
                    line.pop_back();
                if (inside && line == end)
                    break;
                if (inside)
                {
                    if (line.rfind("l=", 0) == 0)
                    {
                        BasisShell sh;
                        sh.l_ = std::stoi(line.substr(2));
                        atom.basis_shells_.push_back(sh);
                    }
                    else if (!line.empty())
                    {
                        if (atom.basis_shells_.empty())
                            throw std::runtime_error("malformed basis block in " + path);
                        std::istringstream ss(line);
                        double e = 0, c = 0;
                        ss >> e >> c;
                        atom.basis_shells_.back().exps_.push_back(e);
                        atom.basis_shells_.back().coeffs_.push_back(c);
                    }
                }
                if (line == begin)
                    inside = found = true;
            }
            if (!found)
                throw std::runtime_error("Basis set " + basis_set + " not found for element " +
                                         std::to_string(atomic_nr) + "!");
            return atom;
        }

        void check(int rc)
        {
            if (rc != LBX_OK)
                throw std::runtime_error(std::string("lbx: ") + lbx_last_error());
        }

        void hipCheck(hipError_t e, const char *what)
        {
            if (e != hipSuccess)
                throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
        }

        /// Device buffer of n doubles, released on scope exit.
        struct DeviceArray
        {
            explicit DeviceArray(size_t n) : n_(n)
            {
                if (n_)
                    hipCheck(hipMalloc(&p_, n_ * sizeof(double)), "hipMalloc");
            }
            ~DeviceArray()
            {
                if (p_)
                    (void)hipFree(p_);
            }
            DeviceArray(const DeviceArray &) = delete;
            DeviceArray &operator=(const DeviceArray &) = delete;

            void toHost(double *dst) const
            {
                if (n_)
                    hipCheck(hipMemcpy(dst, p_, n_ * sizeof(double), hipMemcpyDeviceToHost),
                             "hipMemcpy");
            }
            void fromHost(const double *src)
            {
                if (n_)
                    hipCheck(hipMemcpy(p_, src, n_ * sizeof(double), hipMemcpyHostToDevice),
                             "hipMemcpy");
            }
            double *get() { return static_cast<double *>(p_); }

            void *p_ = nullptr;
            size_t n_;
        };

        /// Flattened basis arrays for lbx_basis (kept alive for the lbx_create call).
        struct FlatBasis
        {
            FlatBasis(const basis_atoms_t &atoms, const std::vector<std::array<double, 3>> &xyz)
            {
                for (const auto &c : xyz)
                    coords.insert(coords.end(), c.begin(), c.end());
                for (size_t ia = 0; ia < atoms.size(); ia++)
                    for (const BasisShell &sh : atoms[ia].basis_shells_)
                    {
                        if (sh.exps_.size() != sh.coeffs_.size() || sh.exps_.empty())
                            throw std::runtime_error("shell with mismatched or empty "
                                                     "exponents/coefficients");
                        atom.push_back(static_cast<int>(ia));
                        l.push_back(sh.l_);
                        nprim.push_back(static_cast<int>(sh.exps_.size()));
                        exps.insert(exps.end(), sh.exps_.begin(), sh.exps_.end());
                        coeffs.insert(coeffs.end(), sh.coeffs_.begin(), sh.coeffs_.end());
                    }
                b.natoms = static_cast<int>(xyz.size());
                b.coords = coords.data();
                b.nshells = static_cast<int>(l.size());
                b.shell_atom = atom.data();
                b.shell_l = l.data();
                b.shell_nprim = nprim.data();
                b.exps = exps.data();
                b.coeffs = coeffs.data();
            }
            std::vector<double> coords, exps, coeffs;
            std::vector<int> atom, l, nprim;
            lbx_basis b{};
        };
    }

    std::string BasisPaths::getMainBasisSetsPath()
    {
        std::lock_guard<std::mutex> g(g_path_mutex);
        return mainPath();
    }

    std::string BasisPaths::getAuxBasisSetsPath()
    {
        std::lock_guard<std::mutex> g(g_path_mutex);
        return auxPath();
    }

    void BasisPaths::setMainBasisSetsPath(const std::string &path)
    {
        std::lock_guard<std::mutex> g(g_path_mutex);
        mainPath() = path;
    }

    void BasisPaths::setAuxBasisSetsPath(const std::string &path)
    {
        std::lock_guard<std::mutex> g(g_path_mutex);
        auxPath() = path;
    }

    BasisAtom basisForAtom(int atomic_nr, const std::string &basis_set)
    {
        return readAtom(atomic_nr, basis_set,
                        findBasisFile(BasisPaths::getMainBasisSetsPath(), basis_set));
    }

    BasisAtom basisForAtomAux(int atomic_nr, const std::string &aux_basis_set)
    {
        return readAtom(atomic_nr, aux_basis_set,
                        findBasisFile(BasisPaths::getAuxBasisSetsPath(), aux_basis_set));
    }

    basis_atoms_t basisForAtoms(const std::vector<int> &atomic_nrs, const std::string &basis_set)
    {
        const std::string path = findBasisFile(BasisPaths::getMainBasisSetsPath(), basis_set);
        basis_atoms_t out;
        for (int z : atomic_nrs)
            out.push_back(readAtom(z, basis_set, path));
        return out;
    }

    basis_atoms_t basisForAtomsAux(const std::vector<int> &atomic_nrs,
                                   const std::string &aux_basis_set)
    {
        const std::string path = findBasisFile(BasisPaths::getAuxBasisSetsPath(), aux_basis_set);
        basis_atoms_t out;
        for (int z : atomic_nrs)
            out.push_back(readAtom(z, aux_basis_set, path));
        return out;
    }

    Structure::Structure(const std::string &basis_set, const std::vector<int> &atomic_nrs,
                         const std::vector<std::array<double, 3>> &coords_angstrom)
        : atomic_nrs_(atomic_nrs), coords_(coords_angstrom)
    {
        init(basisForAtoms(atomic_nrs, basis_set), nullptr);
    }

    Structure::Structure(const std::string &basis_set, const std::string &basis_set_aux,
                         const std::vector<int> &atomic_nrs,
                         const std::vector<std::array<double, 3>> &coords_angstrom)
        : atomic_nrs_(atomic_nrs), coords_(coords_angstrom)
    {
        const basis_atoms_t aux = basisForAtomsAux(atomic_nrs, basis_set_aux);
        init(basisForAtoms(atomic_nrs, basis_set), &aux);
    }

    Structure::Structure(const basis_atoms_t &basis_set, const std::vector<int> &atomic_nrs,
                         const std::vector<std::array<double, 3>> &coords_angstrom)
        : atomic_nrs_(atomic_nrs), coords_(coords_angstrom)
    {
        init(basis_set, nullptr);
    }

    Structure::Structure(const basis_atoms_t &basis_set, const basis_atoms_t &basis_set_aux,
                         const std::vector<int> &atomic_nrs,
                         const std::vector<std::array<double, 3>> &coords_angstrom)
        : atomic_nrs_(atomic_nrs), coords_(coords_angstrom)
    {
        init(basis_set, &basis_set_aux);
    }

    namespace
    {
        basis_atoms_t joined(basis_atoms_t a, const basis_atoms_t &b)
        {
            a.insert(a.end(), b.begin(), b.end());
            return a;
        }
    }

    Structure::Structure(const std::string &basis_set, const std::string &basis_set_ghost,
                         const std::vector<int> &atomic_nrs,
                         const std::vector<int> &atomic_nrs_ghost,
                         const std::vector<std::array<double, 3>> &coords_angstrom,
                         const std::vector<std::array<double, 3>> &coords_angstrom_ghost)
        : atomic_nrs_(atomic_nrs), atomic_nrs_ghost_(atomic_nrs_ghost), coords_(coords_angstrom),
          coords_ghost_(coords_angstrom_ghost)
    {
        init(joined(basisForAtoms(atomic_nrs, basis_set),
                    basisForAtoms(atomic_nrs_ghost, basis_set_ghost)),
             nullptr);
    }

    Structure::Structure(const std::string &basis_set, const std::string &basis_set_ghost,
                         const std::string &basis_set_aux, const std::string &basis_set_aux_ghost,
                         const std::vector<int> &atomic_nrs,
                         const std::vector<int> &atomic_nrs_ghost,
                         const std::vector<std::array<double, 3>> &coords_angstrom,
                         const std::vector<std::array<double, 3>> &coords_angstrom_ghost)
        : atomic_nrs_(atomic_nrs), atomic_nrs_ghost_(atomic_nrs_ghost), coords_(coords_angstrom),
          coords_ghost_(coords_angstrom_ghost)
    {
        const basis_atoms_t aux = joined(basisForAtomsAux(atomic_nrs, basis_set_aux),
                                         basisForAtomsAux(atomic_nrs_ghost, basis_set_aux_ghost));
        init(joined(basisForAtoms(atomic_nrs, basis_set),
                    basisForAtoms(atomic_nrs_ghost, basis_set_ghost)),
             &aux);
    }

    Structure::Structure(const basis_atoms_t &basis_set, const basis_atoms_t &basis_set_ghost,
                         const std::vector<int> &atomic_nrs,
                         const std::vector<int> &atomic_nrs_ghost,
                         const std::vector<std::array<double, 3>> &coords_angstrom,
                         const std::vector<std::array<double, 3>> &coords_angstrom_ghost)
        : atomic_nrs_(atomic_nrs), atomic_nrs_ghost_(atomic_nrs_ghost), coords_(coords_angstrom),
          coords_ghost_(coords_angstrom_ghost)
    {
        init(joined(basis_set, basis_set_ghost), nullptr);
    }

    Structure::Structure(const basis_atoms_t &basis_set, const basis_atoms_t &basis_set_ghost,
                         const basis_atoms_t &basis_set_aux,
                         const basis_atoms_t &basis_set_ghost_aux,
                         const std::vector<int> &atomic_nrs,
                         const std::vector<int> &atomic_nrs_ghost,
                         const std::vector<std::array<double, 3>> &coords_angstrom,
                         const std::vector<std::array<double, 3>> &coords_angstrom_ghost)
        : atomic_nrs_(atomic_nrs), atomic_nrs_ghost_(atomic_nrs_ghost), coords_(coords_angstrom),
          coords_ghost_(coords_angstrom_ghost)
    {
        const basis_atoms_t aux = joined(basis_set_aux, basis_set_ghost_aux);
        init(joined(basis_set, basis_set_ghost), &aux);
    }

    void Structure::init(basis_atoms_t main, const basis_atoms_t *aux)
    {
        // structure.cpp:12-38, 117-150: same count checks, angstrom -> bohr, dimensions from
        // the shells; ghost atoms after the real ones
        if (atomic_nrs_.size() != coords_.size() ||
            atomic_nrs_ghost_.size() != coords_ghost_.size())
            throw std::runtime_error("Structure::Structure(): number of atomic numbers does not "
                                     "match the number of (x, y, z)-coordinates");
        const size_t nall = coords_.size() + coords_ghost_.size();
        if (main.size() != nall || (aux && aux->size() != nall))
            throw std::runtime_error("Structure::Structure(): number of basis atoms does not "
                                     "match the number of (x, y, z)-coordinates");
        for (auto *v : {&coords_, &coords_ghost_})
            for (auto &c : *v)
                for (double &x : c)
                    x *= _ang_to_bohr_;
        basis_ = std::move(main);
        for (const BasisAtom &a : basis_)
            for (const BasisShell &s : a.basis_shells_)
            {
                max_l_ = std::max(max_l_, s.l_);
                dim_ao_ += 2 * s.l_ + 1;
            }
        if (aux)
        {
            use_ri_ = true;
            basis_aux_ = *aux;
            for (const BasisAtom &a : basis_aux_)
                for (const BasisShell &s : a.basis_shells_)
                {
                    max_l_aux_ = std::max(max_l_aux_, s.l_);
                    dim_ao_aux_ += 2 * s.l_ + 1;
                }
        }
    }

    lbx_context *Structure::context() const
    {
        if (!ctx_)
        {
            std::vector<std::array<double, 3>> xyz = coords_;
            xyz.insert(xyz.end(), coords_ghost_.begin(), coords_ghost_.end());
            FlatBasis m(basis_, xyz);
            lbx_context *raw = nullptr;
            if (use_ri_)
            {
                FlatBasis a(basis_aux_, xyz);
                check(lbx_create(&raw, &m.b, &a.b));
            }
            else
                check(lbx_create(&raw, &m.b, nullptr));
            ctx_ = std::shared_ptr<lbx_context>(raw, [](lbx_context *c) { lbx_destroy(c); });
        }
        return ctx_.get();
    }

    namespace
    {
        template <typename Out, typename Fn>
        Out fetch(const Structure &s, Out out, Fn &&fn)
        {
            lbx_context *ctx = s.context();
            DeviceArray d(out.size());
            check(fn(ctx, d.get()));
            check(lbx_synchronize(ctx));
            d.toHost(out.memptr());
            return out;
        }

        void requireRI(const Structure &s, const char *who)
        {
            if (!s.getUseRI())
                throw std::runtime_error(std::string(who) + ": the structure has no auxiliary "
                                         "basis");
        }
    }

    vec4d eri4(const Structure &s)
    {
        const size_t n = s.getDimAO();
        return fetch(s, vec4d(n, n, n, n), lbx_eri4);
    }

    vec2d eri4Diagonal(const Structure &s)
    {
        const size_t n = s.getDimAO();
        return fetch(s, vec2d(n, n), lbx_eri4_diag);
    }

    vec3d eri3(const Structure &s)
    {
        requireRI(s, "eri3()");
        const size_t n = s.getDimAO(), na = s.getDimAOAux();
        return fetch(s, vec3d(n, n, na), lbx_eri3);
    }

    vec2d eri2(const Structure &s)
    {
        requireRI(s, "eri2()");
        const size_t na = s.getDimAOAux();
        return fetch(s, vec2d(na, na), lbx_eri2);
    }

    std::vector<double> eri2Diagonal(const Structure &s)
    {
        requireRI(s, "eri2Diagonal()");
        const size_t na = s.getDimAOAux();
        VectorMD<double, 1> v = fetch(s, VectorMD<double, 1>(na), lbx_eri2_diag);
        return std::vector<double>(v.begin(), v.end());
    }

    std::pair<vec2d, vec2d> coulombExchange(const Structure &s, const vec2d &D,
                                            double schwarz_threshold)
    {
        const size_t n = s.getDimAO();
        if (D.dim<0>() != n || D.dim<1>() != n)
            throw std::runtime_error("coulombExchange(): density must be nbf x nbf");
        lbx_context *ctx = s.context();
        DeviceArray dD(n * n), dJ(n * n), dK(n * n);
        dD.fromHost(D.memptr());
        check(lbx_jk(ctx, dD.get(), dJ.get(), dK.get(), schwarz_threshold, LBX_FP64, 0.0,
                     nullptr));
        check(lbx_synchronize(ctx));
        vec2d J(n, n), K(n, n);
        dJ.toHost(J.memptr());
        dK.toHost(K.memptr());
        return {std::move(J), std::move(K)};
    }

    vec2d coulombExchangeGradient(const Structure &s, const vec2d &D, double schwarz_threshold)
    {
        const size_t n = s.getDimAO();
        if (D.dim<0>() != n || D.dim<1>() != n)
            throw std::runtime_error("coulombExchangeGradient(): density must be nbf x nbf");
        lbx_context *ctx = s.context();
        long nbf = 0, naux = 0, natoms = 0;
        check(lbx_dims(ctx, &nbf, &naux, &natoms));
        DeviceArray dD(n * n), dG(3 * static_cast<size_t>(natoms));
        dD.fromHost(D.memptr());
        check(lbx_jk_grad(ctx, dD.get(), dG.get(), schwarz_threshold, LBX_FP64, 0.0, nullptr));
        check(lbx_synchronize(ctx));
        vec2d g(static_cast<size_t>(natoms), 3);
        dG.toHost(g.memptr());
        return g;
    }

    // ------------------------------------------------------------------ one-electron integrals
    vec2d overlap(const Structure &s)
    {
        const size_t n = s.getDimAO();
        return fetch(s, vec2d(n, n), [](lbx_context *c, double *o)
                     { return lbx_one_electron(c, o, nullptr, 0, nullptr, nullptr); });
    }

    vec2d kineticEnergy(const Structure &s)
    {
        const size_t n = s.getDimAO();
        return fetch(s, vec2d(n, n), [](lbx_context *c, double *o)
                     { return lbx_one_electron(c, nullptr, o, 0, nullptr, nullptr); });
    }

    vec2d externalCharges(const std::vector<std::array<double, 4>> &point_charges, const Structure &s)
    {
        const size_t n = s.getDimAO();
        std::vector<double> q;
        for (const auto &c : point_charges)
            q.insert(q.end(), c.begin(), c.end());
        const int nq = static_cast<int>(point_charges.size());
        return fetch(s, vec2d(n, n), [&](lbx_context *c, double *o)
                     { return lbx_one_electron(c, nullptr, nullptr, nq, nq ? q.data() : nullptr, o); });
    }

    vec2d nuclearAttraction(const Structure &s)
    {
        std::vector<std::array<double, 4>> nuclei(s.getNAtoms());
        for (size_t a = 0; a < s.getNAtoms(); a++)
        {
            const auto xyz = s.getCoordsAtom(a);
            nuclei[a] = {xyz[0], xyz[1], xyz[2], static_cast<double>(s.getZ(a))};
        }
        return externalCharges(nuclei, s);
    }

    // ------------------------------------------------------------------ Boys, catalogue, benchmark
    BoysGrid::BoysGrid(int n) : n_(n), max_n_(n + 6)
    {
        if (n < 0 || n > lbx::kMaxLBoys)
            throw std::runtime_error("BoysGrid: n must be in 0.." + std::to_string(lbx::kMaxLBoys));
        // the library's table (rows x = 0, 0.01, .., 30; F_0 .. F_{kBoysCols-1}), cut to max_n
        const std::vector<double> tab = lbx::boysTable();
        fnx_grid_.resize(static_cast<size_t>(lbx::kBoysRows) * (max_n_ + 1));
        for (int r = 0; r < lbx::kBoysRows; r++)
            for (int m = 0; m <= max_n_; m++)
                fnx_grid_[static_cast<size_t>(r) * (max_n_ + 1) + m] =
                        tab[static_cast<size_t>(r) * lbx::kBoysCols + m];
    }

    std::vector<double> calcBoysF(int n, double x, const BoysGrid &g)
    {
        if (n < 0 || n > g.getN())
            throw std::runtime_error("calcBoysF: n exceeds the grid's order");
        std::vector<double> f(n + 1);
        if (x == 0.0)
        {
            for (int k = 0; k <= n; k++)
                f[k] = 1.0 / (2 * k + 1);
            return f;
        }
        if (x > g.getLargeX())
        {
            f[0] = 0.5 * std::sqrt(M_PI / x);
            for (int k = 1; k <= n; k++)
                f[k] = f[k - 1] * (k - 0.5) / x;
            return f;
        }
        const int w = g.getMaxN() + 1;
        const int iv = static_cast<int>(x / g.getIntervalSize());
        const double dx = x - iv * g.getIntervalSize();
        const double *row = g.getFnxGrid().data() + static_cast<size_t>(iv) * w + n;
        // F_n(x) = sum_k F_{n+k}(x_i) (-dx)^k / k!, k = 0..6 (d/dx F_m = -F_{m+1})
        double s = row[6] / 720.0;
        s = s * -dx + row[5] / 120.0;
        s = s * -dx + row[4] / 24.0;
        s = s * -dx + row[3] / 6.0;
        s = s * -dx + row[2] / 2.0;
        s = s * -dx + row[1];
        s = s * -dx + row[0];
        f[n] = s;
        const double ex = std::exp(-x);
        for (int k = n - 1; k >= 0; k--)
            f[k] = (2.0 * x * f[k + 1] + ex) / (2 * k + 1);
        return f;
    }

    namespace
    {
        std::set<std::string> basisNames(const std::string &root)
        {
            std::set<std::string> names;
            if (root.empty() || !fs::is_directory(root))
                return names;
            for (const auto &e : fs::recursive_directory_iterator(root))
                if (e.is_regular_file() && e.path().extension() == ".txt")
                    names.insert(lower(e.path().stem().string()));
            return names;
        }
    }

    std::set<std::string> availableBasisSets() { return basisNames(BasisPaths::getMainBasisSetsPath()); }

    std::set<std::string> availableBasisSetsAux() { return basisNames(BasisPaths::getAuxBasisSetsPath()); }

    void eri4Benchmark(const Structure &structure)
    {
        std::printf("Lible::%-40s\n", "ERI4 benchmark...");
        std::fflush(stdout);
        const auto t0 = std::chrono::steady_clock::now();
        lbx_context *ctx = structure.context();
        const auto sp = shellPairData(true, structure);
        double total = 0;
        for (size_t i = 0; i < sp.size(); i++)
            for (size_t j = 0; j <= i; j++)
            {
                const auto &ab = sp[i], &cd = sp[j];
                const auto t1 = std::chrono::steady_clock::now();
                const size_t blk = (2 * ab.la_ + 1) * (2 * ab.lb_ + 1) * (2 * cd.la_ + 1) * (2 * cd.lb_ + 1);
                // unique quartets of the class pair, in chunks of ~32 MiB of blocks
                const size_t chunk = std::max<size_t>(1, (size_t(4) << 20) / blk);
                std::vector<int> q;
                size_t count = 0;
                auto flush = [&]() {
                    const size_t n = q.size() / 4;
                    if (n == 0)
                        return;
                    DeviceArray d(n * blk);
                    check(lbx_eri4_batch(ctx, static_cast<long>(n), q.data(), d.get()));
                    check(lbx_synchronize(ctx));
                    std::vector<double> h(n * blk);
                    d.toHost(h.data());
                    for (double x : h)
                        total += std::fabs(x);
                    q.clear();
                };
                for (size_t p = 0; p < ab.n_pairs_; p++)
                    for (size_t r = 0; r < (i == j ? p + 1 : cd.n_pairs_); r++)
                    {
                        q.insert(q.end(), {int(ab.shell_idxs_[2 * p]), int(ab.shell_idxs_[2 * p + 1]),
                                           int(cd.shell_idxs_[2 * r]), int(cd.shell_idxs_[2 * r + 1])});
                        count++;
                        if (q.size() / 4 >= chunk)
                            flush();
                    }
                flush();
                const std::chrono::duration<double> dt = std::chrono::steady_clock::now() - t1;
                std::printf("   %d %d %d %d ; %10zu ; %.2e s\n", ab.la_, ab.lb_, cd.la_, cd.lb_, count, dt.count());
            }
        std::printf("   sum_eri4 = %16.12f\n", total);
        const std::chrono::duration<double> dt = std::chrono::steady_clock::now() - t0;
        std::printf("done %.2e s\n", dt.count());
        std::fflush(stdout);
    }

    // ------------------------------------------------------------------ shell data, kernels
    namespace detail
    {
        /// Ties shell data and kernel functors to their structure: a copy of the structure made
        /// before this object existed (no ownership cycle), which shares the device context when
        /// one was already made.
        struct Tensors
        {
            std::shared_ptr<const Structure> source;
            size_t n = 0, naux = 0;
        };

        /// The blocks of one kernel functor's class pair, computed on the GPU
        /// (lbx_eri*_batch) in chunks: a run of consecutive first-side indices against every
        /// second-side index, the most recently used chunks kept on the host.
        struct BlockCache
        {
            std::shared_ptr<Tensors> t;
            int (*fn)(lbx_context *, long, const int *, double *) = nullptr;
            int per1 = 0, per2 = 0;          // shells per first-side / second-side index
            std::vector<int> first, second;  // their shell indices
            size_t n1 = 0, n2 = 0, block = 0, chunk = 1;
            std::mutex mu;
            std::list<std::pair<size_t, std::shared_ptr<const std::vector<double>>>> lru;

            static constexpr size_t kChunkDoubles = size_t(4) << 20;   // 32 MiB per chunk
            static constexpr size_t kChunks = 8;

            void init(size_t block_doubles)
            {
                block = block_doubles;
                const size_t row = std::max<size_t>(1, n2 * block);
                chunk = std::max<size_t>(1, kChunkDoubles / row);
            }

            std::shared_ptr<const std::vector<double>> get(size_t c)
            {
                std::lock_guard<std::mutex> lk(mu);
                for (auto it = lru.begin(); it != lru.end(); ++it)
                    if (it->first == c)
                    {
                        lru.splice(lru.begin(), lru, it);
                        return lru.front().second;
                    }
                const size_t lo = c * chunk, hi = std::min(n1, lo + chunk);
                const size_t ne = (hi - lo) * n2, per = per1 + per2;
                std::vector<int> sh(ne * per);
                size_t e = 0;
                for (size_t i = lo; i < hi; i++)
                    for (size_t j = 0; j < n2; j++, e++)
                    {
                        for (int k = 0; k < per1; k++)
                            sh[e * per + k] = first[i * per1 + k];
                        for (int k = 0; k < per2; k++)
                            sh[e * per + per1 + k] = second[j * per2 + k];
                    }
                auto v = std::make_shared<std::vector<double>>(ne * block);
                if (ne > 0)
                {
                    lbx_context *ctx = t->source->context();
                    DeviceArray d(ne * block);
                    check(fn(ctx, static_cast<long>(ne), sh.data(), d.get()));
                    check(lbx_synchronize(ctx));
                    d.toHost(v->data());
                }
                lru.emplace_front(c, v);
                if (lru.size() > kChunks)
                    lru.pop_back();
                return v;
            }

            // the block of (i, j): `block` doubles
            std::pair<std::shared_ptr<const std::vector<double>>, size_t> at(size_t i, size_t j)
            {
                auto v = get(i / chunk);
                return {v, ((i % chunk) * n2 + j) * block};
            }
        };
    }

    std::shared_ptr<detail::Tensors> Structure::tensors() const
    {
        if (!tensors_)
        {
            auto t = std::make_shared<detail::Tensors>();
            t->source = std::make_shared<const Structure>(*this);
            t->n = dim_ao_;
            t->naux = dim_ao_aux_;
            tensors_ = t;
        }
        return tensors_;
    }

    const std::vector<Structure::ShellInfo> &Structure::shells(bool aux) const
    {
        auto &slot = aux ? shells_aux_ : shells_;
        if (!slot)
        {
            if (aux && !use_ri_)
                throw std::runtime_error("shellDataAux(): the structure has no auxiliary basis");
            std::vector<std::array<double, 3>> xyz = coords_;
            xyz.insert(xyz.end(), coords_ghost_.begin(), coords_ghost_.end());
            FlatBasis fb(aux ? basis_aux_ : basis_, xyz);
            size_t nbf = 0;
            int maxl = 0;
            auto hs = lbx::buildShells(fb.b, nbf, maxl, aux ? lbx::kMaxLAux : lbx::kMaxLMain);
            auto v = std::make_shared<std::vector<ShellInfo>>();
            size_t cart = 0;
            for (const auto &h : hs)
            {
                v->push_back(ShellInfo{h.l, static_cast<size_t>(h.atom), h.ao, cart, h.xyz, h.exps,
                                       h.coeffs, h.pnorms, h.norms});
                cart += static_cast<size_t>((h.l + 1) * (h.l + 2) / 2);
            }
            slot = v;
        }
        return *slot;
    }

    namespace
    {
        std::vector<Shell> shellsOf(const Structure &s, bool aux, int l)
        {
            const auto &sh = s.shells(aux);
            const size_t nreal = s.getNAtoms();
            std::vector<Shell> out;
            for (size_t i = 0; i < sh.size(); i++)
            {
                const auto &h = sh[i];
                if (l >= 0 && h.l != l)
                    continue;
                Shell x;
                x.l_ = h.l;
                x.z_ = h.atom < nreal ? s.getZ(h.atom) : 0;
                x.dim_cart_ = static_cast<size_t>((h.l + 1) * (h.l + 2) / 2);
                x.dim_sph_ = static_cast<size_t>(2 * h.l + 1);
                x.ofs_cart_ = h.ofs_cart;
                x.ofs_sph_ = h.ofs_sph;
                x.idx_ = i;
                x.idx_atom_ = h.atom;
                x.xyz_coords_ = h.xyz;
                x.exps_ = h.exps;
                x.coeffs_ = h.coeffs;
                x.norms_ = h.norms;
                x.norms_prim_ = h.norms_prim;
                x.tensors_ = s.tensors();
                x.aux_ = aux;
                out.push_back(std::move(x));
            }
            return out;
        }
    }

    std::vector<Shell> Structure::getShells() const { return shellsOf(*this, false, -1); }
    std::vector<Shell> Structure::getShellsAux() const { return shellsOf(*this, true, -1); }
    std::vector<Shell> Structure::getShellsL(int l) const { return shellsOf(*this, false, l); }
    std::vector<Shell> Structure::getShellsLAux(int l) const { return shellsOf(*this, true, l); }

    // shell_pair_data.cpp:7-202: primitive pairs kept where
    // |N_a d_a N_b d_b exp(-ab/(a+b) R_AB^2)| >= primitives_thrs; pairs without any are dropped
    ShellPairData::ShellPairData(bool use_symm, int la, int lb, const std::vector<Shell> &A,
                                 const std::vector<Shell> &B, double primitives_thrs)
        : uses_symm_(use_symm), primitives_thrs_(primitives_thrs), la_(la), lb_(lb)
    {
        for (const auto *v : {&A, &B})
            for (const Shell &x : *v)
            {
                if (x.l_ != (v == &A ? la : lb))
                    throw std::runtime_error("ShellPairData: a shell's angular momentum differs from la / lb");
                if (x.aux_)
                    throw std::runtime_error("ShellPairData: auxiliary-basis shell");
                if (!tensors_)
                    tensors_ = x.tensors_;
                else if (x.tensors_ != tensors_)
                    throw std::runtime_error("ShellPairData: shells of different structures");
            }
        const size_t sa = 2 * la + 1, sb = 2 * lb + 1;
        const size_t ne = sa * sb * static_cast<size_t>(lbx::nHerm(la + lb));
        size_t nprim = 0, nnorm = 0;
        for (size_t ia = 0; ia < A.size(); ia++)
            for (size_t ib = 0; ib < ((use_symm && la == lb) ? std::min(ia + 1, B.size()) : B.size()); ib++)
            {
                const Shell &a = A[ia], &b = B[ib];
                double r2 = 0;
                for (int k = 0; k < 3; k++)
                    r2 += (a.xyz_coords_[k] - b.xyz_coords_[k]) * (a.xyz_coords_[k] - b.xyz_coords_[k]);
                n_pairs_total_++;
                size_t surv = 0;
                for (size_t x = 0; x < a.exps_.size(); x++)
                    for (size_t y = 0; y < b.exps_.size(); y++)
                    {
                        n_ppairs_total_++;
                        const double ea = a.exps_[x], eb = b.exps_[y];
                        const double da = a.norms_prim_[x] * a.coeffs_[x], db = b.norms_prim_[y] * b.coeffs_[y];
                        if (std::fabs(da * db * std::exp(-(ea * eb / (ea + eb)) * r2)) >= primitives_thrs)
                        {
                            exps_.insert(exps_.end(), {ea, eb});
                            coeffs_.insert(coeffs_.end(), {da, db});
                            surv++;
                        }
                    }
                if (surv == 0)
                    continue;
                coords_.insert(coords_.end(), a.xyz_coords_.begin(), a.xyz_coords_.end());
                coords_.insert(coords_.end(), b.xyz_coords_.begin(), b.xyz_coords_.end());
                norms_.insert(norms_.end(), a.norms_.begin(), a.norms_.end());
                norms_.insert(norms_.end(), b.norms_.begin(), b.norms_.end());
                nrs_ppairs_.push_back(surv);
                offsets_primitives_.push_back(nprim);
                offsets_ecoeffs_.push_back(nprim * ne);
                offsets_ecoeffs_deriv1_.push_back(3 * nprim * ne);
                offsets_norms_.push_back(nnorm);
                offsets_sph_.insert(offsets_sph_.end(), {a.ofs_sph_, b.ofs_sph_});
                offsets_cart_.insert(offsets_cart_.end(), {a.ofs_cart_, b.ofs_cart_});
                atomic_idxs_.insert(atomic_idxs_.end(), {a.idx_atom_, b.idx_atom_});
                shell_idxs_.insert(shell_idxs_.end(), {a.idx_, b.idx_});
                nprim += surv;
                nnorm += sa + sb;
                n_pairs_++;
            }
        n_ppairs_ = nprim;
    }

    std::vector<ShellPairData> shellPairData(bool use_symm, const Structure &s)
    {
        std::vector<ShellPairData> out;
        for (int la = 0; la <= s.getMaxL(); la++)
            for (int lb = 0; lb <= (use_symm ? la : s.getMaxL()); lb++)
            {
                out.emplace_back(use_symm, la, lb, s.getShellsL(la), s.getShellsL(lb));
                out.back().tensors_ = s.tensors();   // also for a class without shells
            }
        return out;
    }

    std::vector<ShellData> shellDataAux(const Structure &s)
    {
        const auto &sh = s.shells(true);
        std::vector<ShellData> out;
        for (int l = 0; l <= s.getMaxLAux(); l++)
        {
            ShellData d;
            d.l_ = l;
            d.tensors_ = s.tensors();
            const size_t ne = static_cast<size_t>((2 * l + 1) * lbx::nHerm(l));
            size_t nprim = 0, nnorm = 0;
            for (size_t i = 0; i < sh.size(); i++)
            {
                const auto &a = sh[i];
                if (a.l != l)
                    continue;
                for (size_t x = 0; x < a.exps.size(); x++)
                {
                    d.coeffs_.push_back(a.coeffs[x] * a.norms_prim[x]);
                    d.exps_.push_back(a.exps[x]);
                }
                d.coords_.insert(d.coords_.end(), a.xyz.begin(), a.xyz.end());
                d.norms_.insert(d.norms_.end(), a.norms.begin(), a.norms.end());
                d.atomic_idxs_.push_back(a.atom);
                d.cdepths_.push_back(a.exps.size());
                d.coffsets_.push_back(nprim);
                d.offsets_ecoeffs_.push_back(nprim * ne);
                d.offsets_norms_.push_back(nnorm);
                d.offsets_sph_.push_back(a.ofs_sph);
                d.shell_idxs_.push_back(i);
                nprim += a.exps.size();
                nnorm += 2 * l + 1;
                d.n_shells_++;
            }
            d.n_primitives_ = nprim;
            out.push_back(std::move(d));
        }
        return out;
    }

    namespace
    {
        template <class D>
        std::shared_ptr<detail::Tensors> tensorsOf(const D &a, const char *who)
        {
            if (!a.tensors_)
                throw std::runtime_error(std::string(who) + ": shell data must come from "
                                         "shellPairData() / shellDataAux()");
            return a.tensors_;
        }

        template <class A, class B>
        std::shared_ptr<detail::Tensors> sameTensors(const A &a, const B &b, const char *who)
        {
            auto ta = tensorsOf(a, who), tb = tensorsOf(b, who);
            if (ta != tb)
                throw std::runtime_error(std::string(who) + ": shell data of different structures");
            return ta;
        }

        void checkIndex(size_t i, size_t n, const char *who)
        {
            if (i >= n)
                throw std::out_of_range(std::string(who) + ": index out of range");
        }

        std::vector<int> shellList(const std::vector<size_t> &idx)
        {
            return std::vector<int>(idx.begin(), idx.end());
        }

        size_t nsph(int l) { return static_cast<size_t>(2 * l + 1); }

        // The functor's cache: first side = bra pairs (or auxiliary shells), second side = ket
        // pairs (or auxiliary shells); ncomp components per entry.
        std::shared_ptr<detail::BlockCache> makeCache(std::shared_ptr<detail::Tensors> t,
                                                      int (*fn)(lbx_context *, long, const int *, double *),
                                                      std::vector<int> first, int per1, size_t n1,
                                                      std::vector<int> second, int per2, size_t n2,
                                                      size_t block)
        {
            auto c = std::make_shared<detail::BlockCache>();
            c->t = std::move(t);
            c->fn = fn;
            c->first = std::move(first);
            c->second = std::move(second);
            c->per1 = per1;
            c->per2 = per2;
            c->n1 = n1;
            c->n2 = n2;
            c->init(block);
            return c;
        }

        template <class V>
        V copyBlock(V r, const std::pair<std::shared_ptr<const std::vector<double>>, size_t> &b, size_t comp)
        {
            std::copy_n(b.first->data() + b.second + comp * r.size(), r.size(), r.memptr());
            return r;
        }
    }

    ERI4Kernel::ERI4Kernel(const ShellPairData &ab, const ShellPairData &cd)
        : c_(makeCache(sameTensors(ab, cd, "ERI4Kernel"), lbx_eri4_batch, shellList(ab.shell_idxs_), 2,
                       ab.n_pairs_, shellList(cd.shell_idxs_), 2, cd.n_pairs_,
                       nsph(ab.la_) * nsph(ab.lb_) * nsph(cd.la_) * nsph(cd.lb_)))
    {
    }

    vec4d ERI4Kernel::operator()(size_t p, size_t q, const ShellPairData &ab, const ShellPairData &cd) const
    {
        checkIndex(p, ab.n_pairs_, "ERI4Kernel");
        checkIndex(q, cd.n_pairs_, "ERI4Kernel");
        return copyBlock(vec4d(nsph(ab.la_), nsph(ab.lb_), nsph(cd.la_), nsph(cd.lb_)), c_->at(p, q), 0);
    }

    ERI3Kernel::ERI3Kernel(const ShellPairData &ab, const ShellData &c)
        : c_(makeCache(sameTensors(ab, c, "ERI3Kernel"), lbx_eri3_batch, shellList(ab.shell_idxs_), 2,
                       ab.n_pairs_, shellList(c.shell_idxs_), 1, c.n_shells_,
                       nsph(ab.la_) * nsph(ab.lb_) * nsph(c.l_)))
    {
    }

    vec3d ERI3Kernel::operator()(size_t p, size_t k, const ShellPairData &ab, const ShellData &c) const
    {
        checkIndex(p, ab.n_pairs_, "ERI3Kernel");
        checkIndex(k, c.n_shells_, "ERI3Kernel");
        return copyBlock(vec3d(nsph(ab.la_), nsph(ab.lb_), nsph(c.l_)), c_->at(p, k), 0);
    }

    ERI2Kernel::ERI2Kernel(const ShellData &a, const ShellData &b)
        : c_(makeCache(sameTensors(a, b, "ERI2Kernel"), lbx_eri2_batch, shellList(a.shell_idxs_), 1,
                       a.n_shells_, shellList(b.shell_idxs_), 1, b.n_shells_, nsph(a.l_) * nsph(b.l_)))
    {
    }

    vec2d ERI2Kernel::operator()(size_t i, size_t j, const ShellData &a, const ShellData &b) const
    {
        checkIndex(i, a.n_shells_, "ERI2Kernel");
        checkIndex(j, b.n_shells_, "ERI2Kernel");
        return copyBlock(vec2d(nsph(a.l_), nsph(b.l_)), c_->at(i, j), 0);
    }

    ERI4D1Kernel::ERI4D1Kernel(const ShellPairData &ab, const ShellPairData &cd)
        : c_(makeCache(sameTensors(ab, cd, "ERI4D1Kernel"), lbx_eri4d1_batch, shellList(ab.shell_idxs_), 2,
                       ab.n_pairs_, shellList(cd.shell_idxs_), 2, cd.n_pairs_,
                       12 * nsph(ab.la_) * nsph(ab.lb_) * nsph(cd.la_) * nsph(cd.lb_)))
    {
    }

    std::array<vec4d, 12> ERI4D1Kernel::operator()(size_t p, size_t q, const ShellPairData &ab,
                                                   const ShellPairData &cd) const
    {
        checkIndex(p, ab.n_pairs_, "ERI4D1Kernel");
        checkIndex(q, cd.n_pairs_, "ERI4D1Kernel");
        const auto b = c_->at(p, q);
        std::array<vec4d, 12> r;
        for (size_t x = 0; x < 12; x++)
            r[x] = copyBlock(vec4d(nsph(ab.la_), nsph(ab.lb_), nsph(cd.la_), nsph(cd.lb_)), b, x);
        return r;
    }

    ERI3D1Kernel::ERI3D1Kernel(const ShellPairData &ab, const ShellData &c)
        : c_(makeCache(sameTensors(ab, c, "ERI3D1Kernel"), lbx_eri3d1_batch, shellList(ab.shell_idxs_), 2,
                       ab.n_pairs_, shellList(c.shell_idxs_), 1, c.n_shells_,
                       9 * nsph(ab.la_) * nsph(ab.lb_) * nsph(c.l_)))
    {
    }

    std::array<vec3d, 9> ERI3D1Kernel::operator()(size_t p, size_t k, const ShellPairData &ab,
                                                  const ShellData &c) const
    {
        checkIndex(p, ab.n_pairs_, "ERI3D1Kernel");
        checkIndex(k, c.n_shells_, "ERI3D1Kernel");
        const auto b = c_->at(p, k);
        std::array<vec3d, 9> r;
        for (size_t x = 0; x < 9; x++)
            r[x] = copyBlock(vec3d(nsph(ab.la_), nsph(ab.lb_), nsph(c.l_)), b, x);
        return r;
    }

    ERI2D2Kernel::ERI2D2Kernel(const ShellData &a, const ShellData &b)
        : c_(makeCache(sameTensors(a, b, "ERI2D2Kernel"), lbx_eri2d2_batch, shellList(a.shell_idxs_), 1,
                       a.n_shells_, shellList(b.shell_idxs_), 1, b.n_shells_, 36 * nsph(a.l_) * nsph(b.l_)))
    {
    }

    arr2d<vec2d, 6, 6> ERI2D2Kernel::operator()(size_t i, size_t j, const ShellData &a,
                                                const ShellData &b) const
    {
        checkIndex(i, a.n_shells_, "ERI2D2Kernel");
        checkIndex(j, b.n_shells_, "ERI2D2Kernel");
        const auto blk = c_->at(i, j);
        arr2d<vec2d, 6, 6> r;
        for (size_t x = 0; x < 6; x++)
            for (size_t y = 0; y < 6; y++)
                r[x][y] = copyBlock(vec2d(nsph(a.l_), nsph(b.l_)), blk, 6 * x + y);
        return r;
    }

    ERI4SOCKernel::ERI4SOCKernel(const ShellPairData &ab, const ShellPairData &cd)
        : c_(makeCache(sameTensors(ab, cd, "ERI4SOCKernel"), lbx_eri4soc_batch, shellList(ab.shell_idxs_), 2,
                       ab.n_pairs_, shellList(cd.shell_idxs_), 2, cd.n_pairs_,
                       3 * nsph(ab.la_) * nsph(ab.lb_) * nsph(cd.la_) * nsph(cd.lb_)))
    {
    }

    std::array<vec4d, 3> ERI4SOCKernel::operator()(size_t p, size_t q, const ShellPairData &ab,
                                                   const ShellPairData &cd) const
    {
        checkIndex(p, ab.n_pairs_, "ERI4SOCKernel");
        checkIndex(q, cd.n_pairs_, "ERI4SOCKernel");
        const auto b = c_->at(p, q);
        std::array<vec4d, 3> r;
        for (size_t x = 0; x < 3; x++)
            r[x] = copyBlock(vec4d(nsph(ab.la_), nsph(ab.lb_), nsph(cd.la_), nsph(cd.lb_)), b, x);
        return r;
    }

    ERI3SOCKernel::ERI3SOCKernel(const ShellPairData &ab, const ShellData &c)
        : c_(makeCache(sameTensors(ab, c, "ERI3SOCKernel"), lbx_eri3soc_batch, shellList(ab.shell_idxs_), 2,
                       ab.n_pairs_, shellList(c.shell_idxs_), 1, c.n_shells_,
                       3 * nsph(ab.la_) * nsph(ab.lb_) * nsph(c.l_)))
    {
    }

    std::array<vec3d, 3> ERI3SOCKernel::operator()(size_t p, size_t k, const ShellPairData &ab,
                                                   const ShellData &c) const
    {
        checkIndex(p, ab.n_pairs_, "ERI3SOCKernel");
        checkIndex(k, c.n_shells_, "ERI3SOCKernel");
        const auto b = c_->at(p, k);
        std::array<vec3d, 3> r;
        for (size_t x = 0; x < 3; x++)
            r[x] = copyBlock(vec3d(nsph(ab.la_), nsph(ab.lb_), nsph(c.l_)), b, x);
        return r;
    }

    ERI2D1Kernel::ERI2D1Kernel(const ShellData &a, const ShellData &b)
        : c_(makeCache(sameTensors(a, b, "ERI2D1Kernel"), lbx_eri2d1_batch, shellList(a.shell_idxs_), 1,
                       a.n_shells_, shellList(b.shell_idxs_), 1, b.n_shells_, 6 * nsph(a.l_) * nsph(b.l_)))
    {
    }

    std::array<vec2d, 6> ERI2D1Kernel::operator()(size_t i, size_t j, const ShellData &a,
                                                  const ShellData &b) const
    {
        checkIndex(i, a.n_shells_, "ERI2D1Kernel");
        checkIndex(j, b.n_shells_, "ERI2D1Kernel");
        const auto blk = c_->at(i, j);
        std::array<vec2d, 6> r;
        for (size_t x = 0; x < 6; x++)
            r[x] = copyBlock(vec2d(nsph(a.l_), nsph(b.l_)), blk, x);
        return r;
    }
}
