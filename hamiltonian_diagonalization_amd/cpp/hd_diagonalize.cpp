// hd_diagonalize.cpp -- command-line driver over the drop-in classes.
//
// It runs the reference model callers' stage sequence on the device:
//   SpherePseudopotentialModel::BuildFockBasis / BuildHamiltonian / Diagonalize
//   (src/fqhe_sphere/pseudopotential_model/pseudopotential_model.cpp:160-268,
//    BaseDiagonalize pseudopotential_model_base.hpp:235-285), and the same for
//   term tables read from raw binary files (InteractingOflModel::TermsFromFile /
//   BuildFockBasis / BuildHamiltonian / Diagonalize, interacting_ofl_model.cpp:656-1147).
//
//   hd_diagonalize --nbr-particles N --nbr-orbitals NORB --lz LZ
//                  --pseudopotentials v0,v1,... [options]
//   hd_diagonalize --nbr-particles N --kmax K --layout sphere|ofl --complex 0|1
//                  --ktable F --vtable F [--eps F] --sector lz:a,b|k2d:kx,ky,dx,dy|k1d:ky,dy|none
// options: --nbr-eigenvalues NEV (5)  --tol T (1e-12)  --seed S (0)
//          --precision f64|coded|f32  --dense-below D (60; the flux-lattice caller uses 1000)
//          --dump-crs PREFIX (writes PREFIX.rowstarts/.cols/.values of the upper-triangle CRS)
//          --dump-basis FILE  --initial-vector FILE  --final-vector FILE
// Prints one JSON line with the dimension, the CRS size, the eigenvalues and
// the solver statistics.
#include "hd_dropin.hpp"

#include <map>
#include <sstream>

using namespace diagonalization;

namespace
{
    std::map<std::string, std::string> ParseArgs(int argc, char** argv)
    {
        std::map<std::string, std::string> a;
        for (int i = 1; i < argc; ++i) {
            std::string k = argv[i];
            if (k.rfind("--", 0) != 0) continue;
            k = k.substr(2);
            a[k] = (i + 1 < argc && std::string(argv[i + 1]).rfind("--", 0) != 0) ? argv[++i] : "1";
        }
        return a;
    }

    std::vector<double> ParseList(const std::string& s)
    {
        std::vector<double> v;
        std::stringstream ss(s);
        std::string item;
        while (std::getline(ss, item, ',')) if (!item.empty()) v.push_back(std::stod(item));
        return v;
    }

    template <typename V>
    bool ReadRaw(const std::string& path, std::vector<V>& out)
    {
        std::ifstream f(path, std::ios::binary | std::ios::ate);
        if (!f.is_open()) return false;
        const std::streamsize n = f.tellg();
        f.seekg(0);
        out.resize((size_t)n / sizeof(V));
        f.read(reinterpret_cast<char*>(out.data()), n);
        return (bool)f;
    }

    template <typename V>
    void WriteRaw(const std::string& path, const V* p, size_t n)
    {
        std::ofstream f(path, std::ios::binary);
        f.write(reinterpret_cast<const char*>(p), (std::streamsize)(n * sizeof(V)));
    }

    // Term tables given as raw arrays (the layout is inferred from the size by
    // SpinlessFermionHamiltonian::Add_CdCdCC_Terms).
    template <typename T>
    struct RawQuarticTables : public TermTables<T>
    {
        void Set(std::vector<kState_t> k, std::vector<T> v)
        {
            this->m_kTable = std::move(k);
            this->m_vTable = std::move(v);
        }
    };

    template <typename T, class Quad, class Quart>
    int Run(SpinlessFermionHamiltonian<T>& ham, Quad* quad, Quart* quart, bool haveQuad,
            std::map<std::string, std::string>& args, utilities::MpiWrapper& mpi, const FermionFockBasis& basis)
    {
        const iSize_t N = (iSize_t)std::stoul(args["nbr-particles"]);
        ham.SetFockBasis(basis);
        ham.Initialize(N, (iSize_t)std::stoul(args.count("nbr-orbitals") ? args["nbr-orbitals"] : args["kmax"]),
                       utilities::_SPARSE_MAPPED_, mpi);
        if (args.count("precision")) {
            const std::string p = args["precision"];
            const int prec = p == "coded" ? HD_CODED : p == "f32" ? HD_F32 : HD_F64;
            ham.m_arpackWrapper.SetPrecision(prec);
            ham.SetValueFormats(prec == HD_CODED ? 1 | 4 : prec == HD_F32 ? 1 | 2 : 1);
        } else {
            ham.SetValueFormats(1);
        }
        // pseudopotential_model.cpp:246-250: quadratic terms first, then quartic
        if (haveQuad) ham.Add_CdC_Terms(quad, 0, ham.m_data.m_fockSpaceDim, mpi);
        ham.Add_CdCdCC_Terms(quart, 0, ham.m_data.m_fockSpaceDim, mpi);
        mpi.ExitFlagTest();

        long int nnzUpper = -1;
        if (args.count("dump-crs")) {
            const auto& crs = ham.m_crsSparseMatrix;
            nnzUpper = crs.GetNbrNonZeros();
            const std::string pre = args["dump-crs"];
            WriteRaw(pre + ".rowstarts", crs.GetRowStarts(), (size_t)crs.GetNbrRows() + 1);
            WriteRaw(pre + ".cols", crs.GetCols(), (size_t)nnzUpper);
            WriteRaw(pre + ".values", crs.GetData(), (size_t)nnzUpper);
        }

        iSize_t nev = (iSize_t)std::stoul(args.count("nbr-eigenvalues") ? args["nbr-eigenvalues"] : "5");
        ham.m_arpackWrapper.SetTolerance(args.count("tol") ? std::stod(args["tol"]) : 1e-12);
        ham.m_arpackWrapper.SetSeed(args.count("seed") ? std::stoull(args["seed"]) : 0);
        if (args.count("initial-vector")) {
            ham.m_arpackWrapper.UseInitialVector();
            ham.m_arpackWrapper.SetInitialVectorFile(args["initial-vector"]);
        }
        if (args.count("final-vector")) {
            ham.m_arpackWrapper.StoreFinalVector();
            ham.m_arpackWrapper.SetFinalVectorFile(args["final-vector"]);
        }
        // BaseDiagonalize (pseudopotential_model_base.hpp:243-252): LAPACK below the cutoff
        const fock_t denseBelow = args.count("dense-below") ? std::stoull(args["dense-below"]) : 60;
        std::string method;
        bool ok;
        if (ham.m_data.m_fockSpaceDim < denseBelow) {
            method = "full";
            ok = ham.FullDiagonalize(nev, mpi);
        } else {
            method = "lanczos";
            MatrixVectorFunction<T, Quad, Quart> mv(ham, mpi);
            ok = ham.LanczosDiagonalize(quad, quart, mv, nev, mpi);
        }
        const hd_lanczos_stats& st = ham.m_arpackWrapper.m_lastStats;
        std::printf("{\"dim\": %llu, \"nnz_upper\": %ld, \"method\": \"%s\", \"ok\": %s, \"eigenvalues\": [",
                    (unsigned long long)ham.m_data.m_fockSpaceDim, nnzUpper, method.c_str(), ok ? "true" : "false");
        for (size_t i = 0; i < ham.m_eigenvalues.size(); ++i)
            std::printf("%s%.17g", i ? ", " : "", ham.m_eigenvalues[i]);
        std::printf("], \"matvecs\": %d, \"restarts\": %d, \"residuals\": [", st.matvecs, st.restarts);
        for (iSize_t i = 0; method == "lanczos" && i < nev && i < 64; ++i)
            std::printf("%s%.3e", i ? ", " : "", st.residual[i]);
        std::printf("]}\n");
        return ok ? 0 : 1;
    }

    bool MakeBasis(FermionFockBasis& basis, std::map<std::string, std::string>& args, utilities::MpiWrapper& mpi,
                   iSize_t N, iSize_t norb)
    {
        std::string sec = args.count("sector") ? args["sector"] : "";
        if (sec.empty() && args.count("lz")) sec = "lz:" + args["lz"] + "," + std::to_string(norb - 1);
        bool empty;
        if (sec.rfind("lz:", 0) == 0) {
            auto v = ParseList(sec.substr(3));
            empty = basis.GenerateFockSpace(N, norb, AngularMomentumSector((int)v.at(0), (int)v.at(1)), mpi);
        } else if (sec.rfind("k2d:", 0) == 0) {
            auto v = ParseList(sec.substr(4));
            empty = basis.GenerateFockSpace(
                N, norb, LinearMomentumSector2D((int)v.at(0), (int)v.at(1), (int)v.at(2), (int)v.at(3)), mpi);
        } else if (sec.rfind("k1d:", 0) == 0) {
            auto v = ParseList(sec.substr(4));
            empty = basis.GenerateFockSpace(N, norb, LinearMomentumSector1D((int)v.at(0), (int)v.at(1)), mpi);
        } else {
            empty = basis.GenerateFockSpace(N, norb, NoConstraint(), mpi);
        }
        if (args.count("dump-basis"))
            WriteRaw(args["dump-basis"], basis.m_basis.data(), basis.m_basis.size());
        return !empty;
    }
}

int main(int argc, char** argv)
{
    auto args = ParseArgs(argc, argv);
    utilities::MpiWrapper mpi;
    if (!args.count("nbr-particles")) {
        std::fprintf(stderr, "usage: see the header of hd_diagonalize.cpp\n");
        return 2;
    }
    const iSize_t N = (iSize_t)std::stoul(args["nbr-particles"]);

    if (args.count("pseudopotentials")) {
        // sphere model: tables from V_m (pseudopotential_model_base.hpp:97-229)
        const iSize_t norb = (iSize_t)std::stoul(args["nbr-orbitals"]);
        SphereQuarticTermTables quart;
        if (!quart.FromPseudopotentials((int)norb - 1, ParseList(args["pseudopotentials"]), mpi)) return 1;
        FermionFockBasis basis;
        if (!MakeBasis(basis, args, mpi, N, norb)) {
            std::printf("{\"dim\": 0, \"empty\": true}\n");
            return 0;
        }
        SpinlessFermionHamiltonian<double> ham;
        DiagonalQuadraticTermTables<double> quad;
        return Run(ham, &quad, &quart, false, args, mpi, basis);
    }

    const iSize_t kmax = (iSize_t)std::stoul(args["kmax"]);
    const bool cplx = args.count("complex") && args["complex"] == "1";
    std::vector<kState_t> kt;
    if (!ReadRaw(args["ktable"], kt)) { std::fprintf(stderr, "cannot read --ktable\n"); return 2; }
    FermionFockBasis basis;
    if (!MakeBasis(basis, args, mpi, N, kmax)) {
        std::printf("{\"dim\": 0, \"empty\": true}\n");
        return 0;
    }
    if (cplx) {
        std::vector<dcmplx> vt, eps;
        if (!ReadRaw(args["vtable"], vt)) { std::fprintf(stderr, "cannot read --vtable\n"); return 2; }
        RawQuarticTables<dcmplx> quart;
        quart.Set(std::move(kt), std::move(vt));
        DiagonalQuadraticTermTables<dcmplx> quad;
        quad.Initialize((kState_t)kmax, mpi);
        const bool haveQuad = args.count("eps") && ReadRaw(args["eps"], eps);
        for (iSize_t k = 0; haveQuad && k < kmax; ++k) quad.SetEkk(eps.at(k), (kState_t)k);
        SpinlessFermionHamiltonian<dcmplx> ham;
        return Run(ham, &quad, &quart, haveQuad, args, mpi, basis);
    }
    std::vector<double> vt, eps;
    if (!ReadRaw(args["vtable"], vt)) { std::fprintf(stderr, "cannot read --vtable\n"); return 2; }
    RawQuarticTables<double> quart;
    quart.Set(std::move(kt), std::move(vt));
    DiagonalQuadraticTermTables<double> quad;
    quad.Initialize((kState_t)kmax, mpi);
    const bool haveQuad = args.count("eps") && ReadRaw(args["eps"], eps);
    for (iSize_t k = 0; haveQuad && k < kmax; ++k) quad.SetEkk(eps.at(k), (kState_t)k);
    SpinlessFermionHamiltonian<double> ham;
    return Run(ham, &quad, &quart, haveQuad, args, mpi, basis);
}
