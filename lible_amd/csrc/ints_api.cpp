// lible::ints C++ interface (include/lible/ints/ints.hpp) over the lbx C-ABI.  Host code only:
// basis-file reading, Structure bookkeeping, and the copies of the device results into the
// row-major VectorMD tensors the reference's functions return.
#include <lible/ints/ints.hpp>

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <mutex>
#include <sstream>

#include "lbx.h"

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
}
