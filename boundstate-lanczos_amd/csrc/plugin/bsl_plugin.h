// bsl_plugin.h -- the MI355X Hv calculators behind the reference's
// function-pointer plugin contract (interface.h:40, interface.cpp:3-250).
//
// The reference driver only sees an opaque storage struct and two function
// pointers, HvPrep and Hv, chosen by the HvCalculator string of its Lanczos
// input deck.  bsl_HvInterfaceSetup recognises the HIP calculator strings
// below and returns a storage struct plus HvPrep/Hv adapters with the exact
// reference signatures; the adapters call libbsl_hip.so through its C ABI.
//
//   linRotCartSph_HIP_{NoQuad,Quad}_{ContGrid_Alavi_SPCE, ContGrid_Alavi_TIP4P,
//                                    Alavi_SPCE, Alavi_TIP4P, Coulomb, IsoHarmOsc}[_FP32]
//   linRotCartSph_HIP_Quad_<potential>_BF16X3
//
// ContGrid_Alavi_*: pair sums at the exact H positions; Alavi_*: the
// grid-mapped potential universe of the reference's linRotCartSph_NoQuad_Alavi_*.
// The reference's own NoQuad strings (linRotCartSph_NoQuad_{ContGrid_Alavi_SPCE,
// Alavi_SPCE, Alavi_TIP4P, Coulomb, IsoHarmOsc}) are accepted too, as the FP64
// NoQuad GPU calculators, so that a driver linked against this library reads
// the reference's decks unchanged.
//
// The decks are the reference's: argv[2] is the Hv deck (22 positional
// values, HvPrep_Internal_NoQuad linRotCartSphHvFullVPrepRoutines.cpp:248-321)
// and argv[3] the isotropic harmonic oscillator deck
// (preCalcPotential_IsoHarmOsc IsoHarmOsc_Routines.cpp:41-58) or the Coulomb
// deck (preCalcPotential_Coulomb Coulomb_Routines.cpp:48-71).
//
// The types generalStor and lanczosStor must be the reference's
// (interfaceContainers.h:6-14); a translation unit that includes this header
// has them declared before (the reference's own header, or
// bsl_reference_types.h for the standalone driver).
#pragma once

#include <string>

#include "bsl_hip.h"

// storage struct handed to the driver as a generalStor*
struct bslPluginStor {
    std::string calculator;
    int calc = BSL_CALC_NOQUAD;
    int precision = BSL_PREC_FP64;
    int potential = BSL_POT_CAGE;
    int water = 0;  // 0 SPC/E, 1 TIP4P
    bsl_ctx *ctx = nullptr;
};

// Returns true and fills the three outputs when HvCalculatorSwitch names a
// HIP calculator; returns false (and touches nothing) otherwise, so that the
// reference's own else-if chain can continue.
bool bsl_HvInterfaceSetup(const std::string &HvCalculatorSwitch, generalStor **general_data,
                          void (**HvPrepPtr)(int, char **, generalStor *, lanczosStor *),
                          void (**HvPtr)(int, char **, generalStor *, lanczosStor *, double *, double *));

// HvPrep: parses the decks, prepares the device operator, sets
// total_basis_size, sim_descr and sim_descr_short.  Errors: message on cerr
// and exit(1), as the reference's HvPrep.
void Hv_Prep_linRotCartSph_HIP(int argc, char **argv, generalStor *general_data, lanczosStor *lanczos_data);

// Hv: uec += H vec on host vectors of length total_basis_size.
void Hv_linRotCartSph_HIP(int argc, char **argv, generalStor *general_data, lanczosStor *lanczos_data, double *vec,
                          double *uec);

// the prepared context behind a storage struct from bsl_HvInterfaceSetup
bsl_ctx *bsl_plugin_context(generalStor *general_data);
// multiplicity tolerance of the eigensolve for the calculator's operator:
// 0 (the reference's BISEC) for FP64, BSL_MIXED_MULT_TOL for FP32 and
// BSL_BF16X3_MULT_TOL for BF16x3
double bsl_plugin_mult_tol(generalStor *general_data);
