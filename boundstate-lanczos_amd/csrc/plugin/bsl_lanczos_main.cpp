// bsl_lanczos_main.cpp -- standalone Lanczos driver over the HIP calculators,
// with the reference driver's command line, decks and output files
// (lanczosEigenValuesVectors.cpp:17-506):
//
//   bsl_lanczos LanczosInputFile HvInputFile [PotentialFile]
//
// The Lanczos deck holds 7 positional values (numberOfIterations, the
// eigenvalue window (emin, emax], all/partial/none, the number of
// eigenvectors, the HvCalculator string and the output directory).  Loop 1
// runs with the Krylov vectors resident on the GPU (bsl_lanczos); with
// BSL_HV_PATH=host it runs on the host through the Hv function pointer
// instead, exactly as the reference driver does.  The good eigenvalues,
// the Tm eigenvectors and the Ritz vectors are written to boundstates.txt,
// states_zpe.txt, lv and eigVecFile.txt in a time-stamped directory.
#include "bsl_reference_types.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "bsl_hip.h"
#include "bsl_plugin.h"

using std::cerr;
using std::cout;
using std::endl;
using std::string;

namespace {

void check(int rc, const char *what)
{
    if (rc != BSL_OK) {
        cerr << "ERROR: " << what << ": " << bsl_last_error() << endl;
        std::exit(1);
    }
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 3) {
        cerr << "usage: " << argv[0] << " LanczosInputFile HvInputFile [PotentialFile]" << endl;
        return 1;
    }
    // ---- Lanczos deck (:30-80) ----
    const string lanczosInputFilename = argv[1], HvInputFilename = argv[2];
    int niter = 0, numEigVec = 0;
    double emin = 0, emax = 0;
    string numEigVecWord, HvCalculatorSwitch, outputDirPath, junk, line;
    {
        std::ifstream in(lanczosInputFilename.c_str());
        if (!in.is_open()) {
            cerr << "ERROR: Lanczos input file '" << lanczosInputFilename << "' could not be opened." << endl;
            return 1;
        }
        std::getline(in, line);
        std::getline(in, line);
        in >> junk >> niter >> junk >> emin >> junk >> emax >> junk >> numEigVecWord >> junk >> numEigVec;
        in >> junk >> HvCalculatorSwitch >> junk >> outputDirPath;
    }
    if (numEigVecWord != "all" && numEigVecWord != "partial" && numEigVecWord != "none") {
        cerr << "ERROR: Invalid value of " << numEigVecWord << " for NumberOfEigenvectors(all/partial/none) " << endl;
        return 1;
    }
    if (niter <= 0) {
        cerr << "ERROR: numberOfIterations must be positive." << endl;
        return 1;
    }
    cout << std::scientific << std::setprecision(15) << endl;
    cout << "Eigenvalue and Eigenvector Lanczos calculator begun (MI355X HIP calculators)." << endl;
    cout << "The number of Lanczos iterations will be: " << niter << endl;

    // ---- plugin setup and HvPrep (:93-110) ----
    lanczosStor *lanczos_data = new lanczosStor();
    generalStor *general_data = nullptr;
    void (*HvPrepPtr)(int, char **, generalStor *, lanczosStor *) = nullptr;
    void (*HvPtr)(int, char **, generalStor *, lanczosStor *, double *, double *) = nullptr;
    if (!bsl_HvInterfaceSetup(HvCalculatorSwitch, &general_data, &HvPrepPtr, &HvPtr)) {
        cerr << "HvCalculator '" << HvCalculatorSwitch << "' not recognized." << endl;
        return 1;
    }
    (*HvPrepPtr)(argc, argv, general_data, lanczos_data);
    const int ntotbs = lanczos_data->total_basis_size;
    const string sim_descr = lanczos_data->sim_descr;
    cout << "The total basis size is: " << ntotbs << endl;
    cout << "The current simulation has been described as: " << sim_descr << endl;
    cout << std::fixed << std::setprecision(6) << "Eigenvalues will be searched for in the interval (" << emin
         << ", " << emax << "]; all in kJ/mol." << endl;
    cout << std::scientific << std::setprecision(15) << endl;

    // ---- output directory (:123-152) ----
    char stamp[80];
    std::time_t t = std::time(nullptr);
    std::strftime(stamp, sizeof stamp, "%F_%H_%M_%S", std::localtime(&t));
    const string dir = outputDirPath + lanczos_data->sim_descr_short + "_" + stamp;
    if (std::system(("mkdir -p " + dir).c_str()) != 0) {
        cerr << "ERROR: Storage directory could not be created with the command: mkdir -p " << dir << endl;
        return 1;
    }
    cout << "The output directory is: " << dir << endl;
    if (std::system(("cp " + lanczosInputFilename + " " + dir).c_str()) != 0) cerr << "warning: deck not copied" << endl;
    if (std::system(("cp " + HvInputFilename + " " + dir).c_str()) != 0) cerr << "warning: deck not copied" << endl;

    // ---- Loop 1 (:158-242) ----
    std::vector<double> alpha(niter, 0.0), beta(niter + 1, 0.0), beta2(niter + 1, 0.0);
    const char *path = std::getenv("BSL_HV_PATH");
    const bool host_path = path && string(path) == "host";
    bsl_ctx *ctx = bsl_plugin_context(general_data);
    cout << "'Lanczos Loop 1 - Get Tm' beginning." << endl;
    if (host_path) {
        std::vector<double> rvec(ntotbs, 0.0), vec(ntotbs, 1. / std::sqrt((double)ntotbs)), uec(ntotbs, 0.0);
        for (int j = 1; j <= niter; j++) {
            std::fill(uec.begin(), uec.end(), 0.0);
            (*HvPtr)(argc, argv, general_data, lanczos_data, vec.data(), uec.data());
            for (int i = 0; i < ntotbs; i++) rvec[i] += uec[i];
            double a = 0.;
            for (int i = 0; i < ntotbs; i++) a += vec[i] * rvec[i];
            alpha[j - 1] = a;
            for (int i = 0; i < ntotbs; i++) rvec[i] -= a * vec[i];
            double b2 = 0.;
            for (int i = 0; i < ntotbs; i++) b2 += rvec[i] * rvec[i];
            beta2[j] = b2;
            beta[j] = std::sqrt(b2);
            for (int i = 0; i < ntotbs; i++) {
                const double r = (1. / beta[j]) * rvec[i];
                rvec[i] = (-beta[j]) * vec[i];
                vec[i] = r;
            }
        }
    } else {
        check(bsl_lanczos(ctx, niter, nullptr, BSL_START_UNIFORM, 0, alpha.data(), beta.data(), beta2.data()),
              "Lanczos loop 1");
    }
    cout << "Lanczos Coefficients" << endl << "Iteration j alpha_j beta_j" << endl;
    for (int j = 1; j <= niter; j++) cout << j << " " << alpha[j - 1] << " " << beta[j] << endl;
    cout << "'Lanczos Loop 1 - Get Tm' FINISHED." << endl;

    // ---- good eigenvalues (:251-295) ----
    std::vector<double> eval(niter), evalerr(niter);
    int32_t ngood = 0;
    check(bsl_lancbis_tol(niter, alpha.data(), beta.data(), beta2.data(), emin, emax,
                          bsl_plugin_mult_tol(general_data), eval.data(), evalerr.data(), &ngood),
          "lancbis");
    cout << endl << "Number of Good Eigenvalues is: " << ngood << endl;
    if (ngood > 0) {
        cout << "Groundstate Eigenvalue is: " << eval[0] << " kJ/mol" << endl;
        cout << "Groundstate Eigenvalue Error is: " << evalerr[0] << " kJ/mol?" << endl;
    }
    {
        std::ofstream lancout((dir + "/boundstates.txt").c_str());
        std::ofstream zpe((dir + "/states_zpe.txt").c_str());
        lancout << std::scientific << std::setprecision(15) << endl;
        zpe << std::scientific << std::setprecision(15) << endl;
        lancout << "The following are the eigenvalues and errors for the simulation " << sim_descr << endl;
        lancout << "Number of Eigenvalues: " << ngood << endl;
        lancout << "State | Eigenvalue (kJ/mol) | Error (kJ/mol?)" << endl;
        zpe << "The following are the relative energies for the simulation " << sim_descr << endl;
        zpe << "Number of Eigenvalues: " << ngood << endl;
        zpe << "Zero Point Energy = " << (ngood > 0 ? eval[0] : 0.0) << " kJ/mol" << endl;
        zpe << "State | Relative Energy (kJ/mol)" << endl;
        for (int i = 0; i < ngood; i++) {
            lancout << i << " " << eval[i] << " " << evalerr[i] << endl;
            zpe << i << " " << (eval[i] - eval[0]) << endl;
        }
    }
    if (numEigVecWord == "none") {
        cout << "LANCZOS ALGORITHM COMPLETED - no eigenvectors calculated as based on user input." << endl;
        bsl_destroy(ctx);
        return 0;
    }
    if (numEigVecWord == "partial") ngood = std::min<int32_t>(numEigVec, ngood);

    // ---- Tm eigenvectors and lv (:329-364) ----
    std::vector<double> evtr((size_t)niter * std::max(ngood, 1));
    std::vector<int32_t> mamin(std::max(ngood, 1));
    check(bsl_trivec_tol(niter, alpha.data(), beta.data(), beta2.data(), eval.data(), ngood,
                         bsl_plugin_mult_tol(general_data), evtr.data(), mamin.data()),
          "trivec");
    {
        std::ofstream lv((dir + "/lv").c_str());
        lv << std::scientific << std::setprecision(15) << endl;
        lv << "The following are the Coefficients squared of the Eigenvectors of Tm for the simulation " << sim_descr
           << endl;
        lv << "Number of Eigenvectors: " << ngood << endl;
        lv << "Iteration | Coefficient Squared | Cumulative Norm " << endl;
        std::vector<double> cum(ngood, 0.0);
        for (int j = 1; j <= niter; j++) {
            lv << j << " ";
            for (int n = 0; n < ngood; n++) {
                const double c2 = std::pow(evtr[(j - 1) + (size_t)niter * n], 2.);
                cum[n] += c2;
                lv << c2 << " " << cum[n] << " ";
            }
            lv << endl;
        }
    }

    // ---- Loop 2: Ritz vectors (:371-495) ----
    std::vector<double> ritz((size_t)ntotbs * std::max(ngood, 1));
    const double mult_tol = bsl_plugin_mult_tol(general_data);
    if (mult_tol > 0.0) {
        // a mixed-precision operator: its levels are refined on the device by
        // the FP64 Rayleigh quotient of their Ritz vectors (FP64 Hv), and the
        // refined values written beside the reference's files
        std::vector<double> ref_e(std::max(ngood, 1)), ref_r(std::max(ngood, 1));
        check(bsl_ritz_refine(ctx, niter, nullptr, BSL_START_UNIFORM, 0, evtr.data(), ngood, ritz.data(), ref_e.data(),
                              ref_r.data()),
              "Lanczos loop 2 and FP64 refinement");
        std::ofstream rf((dir + "/refined_levels.txt").c_str());
        rf << std::scientific << std::setprecision(15) << endl;
        rf << "FP64 Rayleigh-quotient refinement of the levels of the mixed-precision simulation " << sim_descr << endl;
        rf << "Number of Eigenvalues: " << ngood << endl;
        rf << "State | Lanczos Eigenvalue (kJ/mol) | Refined Eigenvalue (kJ/mol) | Residual |Hy - Ey|/|y| (kJ/mol)"
           << endl;
        for (int i = 0; i < ngood; i++) rf << i << " " << eval[i] << " " << ref_e[i] << " " << ref_r[i] << endl;
    } else {
        check(bsl_ritz(ctx, niter, nullptr, BSL_START_UNIFORM, 0, evtr.data(), ngood, ritz.data()), "Lanczos loop 2");
    }
    {
        std::ofstream ev((dir + "/eigVecFile.txt").c_str());
        ev << std::scientific << std::setprecision(15) << endl;
        ev << "The following are the Ritz vectors of H for the simulation " << sim_descr << endl;
        ev << "Number of Ritz vectors: " << ngood << endl;
        ev << "Basis Index | Vector 1 | Vector 2 ... " << endl;
        for (int row = 0; row < ntotbs; row++) {
            ev << row << " ";
            for (int n = 0; n < ngood; n++) ev << ritz[(size_t)n * ntotbs + row] << " ";
            ev << endl;
        }
    }
    cout << "Boundstate Lanczos Calculator has finished executing." << endl;
    bsl_destroy(ctx);
    return 0;
}
