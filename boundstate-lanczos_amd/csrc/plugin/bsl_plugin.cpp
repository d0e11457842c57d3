// bsl_plugin.cpp -- HvPrep/Hv adapters of the HIP calculators for the
// reference's plugin boundary (see bsl_plugin.h).
#ifndef BSL_REFERENCE_TYPES_HEADER
#define BSL_REFERENCE_TYPES_HEADER "bsl_reference_types.h"
#endif
#include BSL_REFERENCE_TYPES_HEADER

#include "bsl_plugin.h"

#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

namespace {

[[noreturn]] void die(const std::string &msg)
{
    std::cerr << msg << std::endl;
    std::exit(1);
}

void check(int rc, const char *what)
{
    if (rc != BSL_OK) die(std::string("ERROR: ") + what + ": " + bsl_last_error());
}

const char *kPrefix = "linRotCartSph_HIP_";

// potential part of a calculator string: ContGrid_Alavi_* (pair sums at the
// exact H positions), Alavi_* (grid-mapped universe), Coulomb, IsoHarmOsc
bool parse_potential(const std::string &rest, bslPluginStor *st)
{
    if (rest == "ContGrid_Alavi_SPCE" || rest == "ContGrid_Alavi_TIP4P" || rest == "Alavi_SPCE" ||
        rest == "Alavi_TIP4P") {
        st->potential = rest.compare(0, 9, "ContGrid_") == 0 ? BSL_POT_CAGE : BSL_POT_CAGE_GRID;
        st->water = rest.size() > 5 && rest.compare(rest.size() - 5, 5, "TIP4P") == 0 ? 1 : 0;
        return true;
    }
    if (rest == "Coulomb") {
        st->potential = BSL_POT_COULOMB;
        return true;
    }
    if (rest == "IsoHarmOsc") {
        st->potential = BSL_POT_IHO;
        return true;
    }
    return false;
}

bool parse_calculator(const std::string &s, bslPluginStor *st)
{
    // the reference's NoQuad strings (interface.cpp:90-211) are accepted as
    // aliases of the FP64 NoQuad GPU calculators; its on-the-fly strings are
    // not (the reference's on-the-fly Hv computes another operator)
    const std::string ref = "linRotCartSph_NoQuad_";
    if (s.compare(0, ref.size(), ref) == 0) {
        if (!parse_potential(s.substr(ref.size()), st)) return false;
        st->calculator = s;
        st->calc = BSL_CALC_NOQUAD;
        st->precision = BSL_PREC_FP64;
        return true;
    }
    if (s.compare(0, std::string(kPrefix).size(), kPrefix) != 0) return false;
    std::string rest = s.substr(std::string(kPrefix).size());
    int precision = BSL_PREC_FP64;
    const std::string fp32 = "_FP32", bf = "_BF16X3";
    if (rest.size() > fp32.size() && rest.compare(rest.size() - fp32.size(), fp32.size(), fp32) == 0) {
        precision = BSL_PREC_FP32;
        rest = rest.substr(0, rest.size() - fp32.size());
    } else if (rest.size() > bf.size() && rest.compare(rest.size() - bf.size(), bf.size(), bf) == 0) {
        precision = BSL_PREC_BF16X3;
        rest = rest.substr(0, rest.size() - bf.size());
    }
    int calc;
    if (rest.compare(0, 7, "NoQuad_") == 0) {
        calc = BSL_CALC_NOQUAD;
        rest = rest.substr(7);
    } else if (rest.compare(0, 5, "Quad_") == 0) {
        calc = BSL_CALC_QUAD;
        rest = rest.substr(5);
    } else {
        return false;
    }
    bslPluginStor probe = *st;
    if (!parse_potential(rest, &probe)) return false;
    *st = probe;
    st->calculator = s;
    st->calc = calc;
    st->precision = precision;
    return true;
}

// Hv deck: two comment lines, then 22 "name value" pairs whose names are
// ignored (HvPrep_Internal_NoQuad linRotCartSphHvFullVPrepRoutines.cpp:248-321)
struct HvDeck {
    int nx, ny, nz, lmax, thetaPoints, phiPoints, pnx, pny, pnz;
    double x_max, y_max, z_max, px_max, py_max, pz_max, totalMass, rotationalConstant, ceilingPotential;
    std::string geometryFilename, simulationFilename, quadConvergeStudy, symFlag;
};

HvDeck read_hv_deck(const std::string &path)
{
    std::ifstream in(path.c_str());
    if (!in.is_open()) die("Input file '" + path + "' could not be opened.");
    std::string line, junk;
    std::getline(in, line);
    std::getline(in, line);
    HvDeck d;
    in >> junk >> d.nx >> junk >> d.x_max >> junk >> d.ny >> junk >> d.y_max >> junk >> d.nz >> junk >> d.z_max;
    in >> junk >> d.lmax >> junk >> d.thetaPoints >> junk >> d.phiPoints;
    in >> junk >> d.pnx >> junk >> d.px_max >> junk >> d.pny >> junk >> d.py_max >> junk >> d.pnz >> junk >> d.pz_max;
    in >> junk >> d.totalMass >> junk >> d.rotationalConstant >> junk >> d.geometryFilename;
    in >> junk >> d.simulationFilename >> junk >> d.ceilingPotential >> junk >> d.quadConvergeStudy;
    in >> junk >> d.symFlag;
    if (!in) die("Input file '" + path + "' is incomplete: 22 name/value pairs are expected.");
    return d;
}

// COM + ZY'Z'' Euler rows of a geometry file (getSPCEatoms SPCE_AH2_EngRoutines.cpp:15-39)
std::vector<double> read_geometry(const std::string &path)
{
    std::ifstream in(path.c_str());
    if (!in.is_open()) die("File " + path + " could not be opened.");
    int nmol = 0;
    std::string line, name;
    in >> nmol;
    std::getline(in, line);
    std::getline(in, line);
    std::vector<double> rows(6 * (size_t)nmol);
    for (int i = 0; i < nmol; i++) {
        in >> name;
        for (int c = 0; c < 6; c++) in >> rows[6 * i + c];
    }
    if (!in) die("Geometry file " + path + " is incomplete.");
    return rows;
}

}  // namespace

bool bsl_HvInterfaceSetup(const std::string &HvCalculatorSwitch, generalStor **general_data,
                          void (**HvPrepPtr)(int, char **, generalStor *, lanczosStor *),
                          void (**HvPtr)(int, char **, generalStor *, lanczosStor *, double *, double *))
{
    bslPluginStor probe;
    if (!parse_calculator(HvCalculatorSwitch, &probe)) return false;
    bslPluginStor *st = new bslPluginStor(probe);
    *general_data = reinterpret_cast<generalStor *>(st);
    *HvPrepPtr = &Hv_Prep_linRotCartSph_HIP;
    *HvPtr = &Hv_linRotCartSph_HIP;
    return true;
}

bsl_ctx *bsl_plugin_context(generalStor *general_data)
{
    return reinterpret_cast<bslPluginStor *>(general_data)->ctx;
}

double bsl_plugin_mult_tol(generalStor *general_data)
{
    const bslPluginStor *st = reinterpret_cast<bslPluginStor *>(general_data);
    if (st->precision == BSL_PREC_FP64) return 0.0;
    if (st->precision == BSL_PREC_FP32) return BSL_MIXED_MULT_TOL;
    return st->calc == BSL_CALC_QUAD ? BSL_BF16X3_MULT_TOL : BSL_BF16X3_KIN_MULT_TOL;
}

void Hv_Prep_linRotCartSph_HIP(int argc, char **argv, generalStor *general_data, lanczosStor *lanczos)
{
    bslPluginStor *st = reinterpret_cast<bslPluginStor *>(general_data);
    if (argc < 3) die("HvPrep: the Hv input file (argv[2]) is missing.");
    const HvDeck d = read_hv_deck(argv[2]);
    // checks of HvPrep_Internal_NoQuad :330-348
    {
        std::ifstream g(d.geometryFilename.c_str());
        if (!g.is_open() && d.geometryFilename != "NULL")
            die("Atom geometry file '" + d.geometryFilename + "' could not be opened.");
    }
    if (d.x_max > d.px_max || d.y_max > d.py_max || d.z_max > d.pz_max)
        die("Error, the system's physical dimensions must be the same size or smaller than the potential universe "
            "dimensions.");
    if (d.x_max * d.px_max * d.y_max * d.py_max * d.z_max * d.pz_max < DBL_EPSILON ||
        (long long)d.nx * d.ny * d.nz * d.pnx * d.pny * d.pnz == 0) {
        std::ostringstream os;
        os << "Error, none of the system dimensions nor the number of points can be zero and none should be smaller "
              "than the machine epsilon: "
           << DBL_EPSILON;
        die(os.str());
    }
    int sym;
    if (d.symFlag == "All")
        sym = BSL_SYM_ALL;
    else if (d.symFlag == "Even")
        sym = BSL_SYM_EVEN;
    else if (d.symFlag == "Odd")
        sym = BSL_SYM_ODD;
    else
        die("ERROR: Symmeterizer Flag not recognized.");

    bsl_params p{};
    p.nx = d.nx;
    p.ny = d.ny;
    p.nz = d.nz;
    p.x_max = d.x_max;
    p.y_max = d.y_max;
    p.z_max = d.z_max;
    p.lmax = d.lmax;
    p.n_theta = d.thetaPoints;
    p.n_phi = d.phiPoints;
    p.mass = d.totalMass;
    p.rot_const = d.rotationalConstant;
    p.ceiling = d.ceilingPotential;
    p.symmetry = sym;
    p.basis = BSL_BASIS_FULL;  // the driver's vectors carry every lm
    p.calculator = st->calc;
    p.precision = st->precision;
    p.potential = st->potential;
    std::vector<double> pos, q;
    std::vector<int32_t> lj;
    if (st->potential == BSL_POT_CAGE || st->potential == BSL_POT_CAGE_GRID) {
        // Alavi H2 (Alavi_Parameters.h:5-14) on SPC/E (SPCE_AH2_EngRoutines.h:17-38)
        // or TIP4P (TIP4P_AH2_EngRoutines.h:16-41) waters; the grid-mapped
        // potentials are tabulated on the deck's potential universe
        // (preCalcPotential_Alavi Alavi_H2_Routines.cpp:289-317)
        p.pnx = d.pnx;
        p.pny = d.pny;
        p.pnz = d.pnz;
        p.px_max = d.px_max;
        p.py_max = d.py_max;
        p.pz_max = d.pz_max;
        const std::vector<double> rows = read_geometry(d.geometryFilename);
        const int nmol = (int)(rows.size() / 6);
        const int ns = st->water == 0 ? 3 : 4;
        pos.resize((size_t)3 * ns * nmol);
        q.resize((size_t)ns * nmol);
        lj.resize((size_t)ns * nmol);
        check(bsl_water_sites(st->water, nmol, rows.data(), pos.data(), q.data(), lj.data()), "water sites");
        p.n_sites = ns * nmol;
        p.site_pos = pos.data();
        p.site_q = q.data();
        p.site_lj = lj.data();
        p.q_h2_h = 0.4932;
        p.q_h2_cm = -0.9864;
        p.h2_half_bond = 0.03707;
        p.lj_a = st->water == 0 ? 1.364e-06 : 1.362e-06;
        p.lj_b = st->water == 0 ? 1.533e-3 : 1.532e-3;
    } else if (st->potential == BSL_POT_COULOMB) {
        // preCalcPotential_Coulomb Coulomb_Routines.cpp:36-86: argv[3] holds
        // the centre charge, the molecule charge and the |V| cap
        if (argc < 4) die("Coulomb Potential input file (argv[3]) is missing.");
        std::ifstream in(argv[3]);
        if (!in.is_open()) die(std::string("Coulomb Potential input file '") + argv[3] + "' could not be opened.");
        std::string line, junk;
        std::getline(in, line);
        std::getline(in, line);
        in >> junk >> p.coul_q_centre >> junk >> p.coul_q_mol >> junk >> p.coul_ceil;
    } else {
        // preCalcPotential_IsoHarmOsc IsoHarmOsc_Routines.cpp:41-78
        if (argc < 4) die("Isotropic Harmonic Oscillator Potential input file (argv[3]) is missing.");
        std::ifstream in(argv[3]);
        if (!in.is_open())
            die(std::string("Isotropic Harmonic Oscillator Potential input file '") + argv[3] +
                "' could not be opened.");
        std::string line, junk;
        double mass, frequency, ceil_v;
        std::getline(in, line);
        std::getline(in, line);
        in >> junk >> mass >> junk >> frequency >> junk >> ceil_v;
        p.iho_factor = 0.5 * mass * frequency * frequency;
        p.iho_ceil = ceil_v;
    }
    check(bsl_create(&p, &st->ctx), "HvPrep");
    lanczos->total_basis_size = (int)bsl_basis_size(st->ctx);
    lanczos->sim_descr = d.simulationFilename;
    lanczos->sim_descr_short = d.simulationFilename;
}

void Hv_linRotCartSph_HIP(int, char **, generalStor *general_data, lanczosStor *, double *vec, double *uec)
{
    bslPluginStor *st = reinterpret_cast<bslPluginStor *>(general_data);
    check(bsl_hv(st->ctx, vec, uec), "Hv");
}
