// ref_harness.cpp -- drives the reference's own Hv calculators (compiled from
// the reference sources by oracle/Makefile.ref into oracle/_ref/).  Test
// infrastructure only: it calls HvInterfaceSetup / HvPrep / Hv exactly as the
// reference driver does (lanczosEigenValuesVectors.cpp:93-103, :195) and
// exposes the results to the oracle tests through a small C API.
#include <cstring>
#include <string>

#include "interface.h"
#include "lanczos.h"

extern "C" void genran_(int *, double *, int *);

namespace {
generalStor *g_general = nullptr;
lanczosStor *g_lanczos = nullptr;
void (*g_prep)(int, char **, generalStor *, lanczosStor *) = nullptr;
void (*g_hv)(int, char **, generalStor *, lanczosStor *, double *, double *) = nullptr;
char g_arg0[8] = "ref", g_arg1[8] = "unused";
std::string g_hvdeck, g_potdeck;
char *g_argv[4];
}  // namespace

extern "C" __attribute__((visibility("default"))) long long ref_setup(const char *calc, const char *hvdeck,
                                                                      const char *potdeck)
{
    g_hvdeck = hvdeck;
    g_potdeck = potdeck ? potdeck : "";
    g_argv[0] = g_arg0;
    g_argv[1] = g_arg1;
    g_argv[2] = const_cast<char *>(g_hvdeck.c_str());
    g_argv[3] = const_cast<char *>(g_potdeck.c_str());
    g_lanczos = new lanczosStor();
    HvInterfaceSetup(std::string(calc), &g_general, &g_prep, &g_hv);
    (*g_prep)(potdeck ? 4 : 3, g_argv, g_general, g_lanczos);
    return g_lanczos->total_basis_size;
}

extern "C" __attribute__((visibility("default"))) void ref_hv(const double *vec, double *uec)
{
    (*g_hv)(4, g_argv, g_general, g_lanczos, const_cast<double *>(vec), uec);
}

// V blocks of the NoQuad calculators, reference order [i][j][k][n][n']
extern "C" __attribute__((visibility("default"))) int ref_vblocks(double *out, long long count)
{
    interfaceStor *s = reinterpret_cast<interfaceStor *>(g_general);
    if (!s->potential || !s->potential->fullPotential) return -1;
    std::memcpy(out, s->potential->fullPotential, sizeof(double) * (size_t)count);
    return 0;
}

// replaces the V blocks of a prepared NoQuad calculator (same order): the
// reference's Hv then applies the given potential with its own kinetic,
// rotor and mask terms
extern "C" __attribute__((visibility("default"))) int ref_set_vblocks(const double *in, long long count)
{
    interfaceStor *s = reinterpret_cast<interfaceStor *>(g_general);
    if (!s->potential || !s->potential->fullPotential) return -1;
    std::memcpy(s->potential->fullPotential, in, sizeof(double) * (size_t)count);
    return 0;
}

// kinetic matrix of axis 0/1/2 and the grid
extern "C" __attribute__((visibility("default"))) void ref_kin(int axis, double *kin, double *grid)
{
    interfaceStor *s = reinterpret_cast<interfaceStor *>(g_general);
    int n = axis == 0 ? s->grids->nx : axis == 1 ? s->grids->ny : s->grids->nz;
    double *k = axis == 0 ? s->grids->xKinMat : axis == 1 ? s->grids->yKinMat : s->grids->zKinMat;
    double *g = axis == 0 ? s->grids->x_Grid : axis == 1 ? s->grids->y_Grid : s->grids->z_Grid;
    std::memcpy(kin, k, sizeof(double) * n * n);
    std::memcpy(grid, g, sizeof(double) * n);
}

// quadrature nodes / weights and harmonic tables L_lm[n][a], S_m[m][b] (both phi halves)
extern "C" __attribute__((visibility("default"))) void ref_quad(double *ct, double *wt, double *cp, double *wp,
                                                                double *L, double *S1, double *S2)
{
    interfaceStor *s = reinterpret_cast<interfaceStor *>(g_general);
    int na = s->quadrature->GLnum, nb = s->quadrature->GCnum;
    int lmax = s->lmBasis->lmax, nn = s->lmBasis->length;
    std::memcpy(ct, s->quadrature->GLabscissae, sizeof(double) * na);
    std::memcpy(wt, s->quadrature->GLweights, sizeof(double) * na);
    std::memcpy(cp, s->quadrature->GCabscissae, sizeof(double) * nb);
    std::memcpy(wp, s->quadrature->GCweights, sizeof(double) * nb);
    for (int n = 0; n < nn; n++)
        for (int a = 0; a < na; a++) L[n * na + a] = s->tesseral->L_lm[n][a];
    for (int m = 0; m < 2 * lmax + 1; m++)
        for (int b = 0; b < nb; b++) {
            S1[m * nb + b] = s->tesseral->S_m[m][b];
            S2[m * nb + b] = s->tesseral2PI->S_m[m][b];
        }
}

// water sites from a .xyz geometry (SPC/E model, getSPCEatoms)
extern "C" __attribute__((visibility("default"))) int ref_spce_sites(const char *xyz, double *pos, char *type)
{
    char *atomType;
    VECT *atomPos;
    int nAtoms;
    getSPCEatoms(&atomType, &atomPos, &nAtoms, std::string(xyz));
    for (int i = 0; i < nAtoms; i++) {
        for (int c = 0; c < 3; c++) pos[3 * i + c] = atomPos[i].COOR(c);
        type[i] = atomType[i];
    }
    return nAtoms;
}

// point energies: Q_SPCE_Eng + LJ_SPCE_Eng_Fast at a position
extern "C" __attribute__((visibility("default"))) double ref_spce_energy(const char *xyz, double x, double y,
                                                                        double z, double q, int with_lj)
{
    char *atomType;
    VECT *atomPos;
    int nAtoms;
    getSPCEatoms(&atomType, &atomPos, &nAtoms, std::string(xyz));
    VECT pos(3);
    pos.COOR(0) = x;
    pos.COOR(1) = y;
    pos.COOR(2) = z;
    double e = Q_SPCE_Eng(pos, q, atomType, atomPos, nAtoms);
    if (with_lj) e += LJ_SPCE_Eng_Fast(pos, atomType, atomPos, nAtoms);
    return e;
}

// Coulomb energy of a charge q at (x,y,z) with the TIP4P sites of a .xyz
// geometry (getTIP4Patoms + Q_TIP4P_Eng), as in the reference's commented test main
extern "C" __attribute__((visibility("default"))) double ref_tip4p_coulomb(const char *xyz, double x, double y,
                                                                          double z, double q)
{
    char *atomType;
    VECT *atomPos;
    int nAtoms;
    getTIP4Patoms(&atomType, &atomPos, &nAtoms, std::string(xyz));
    VECT pos(3);
    pos.COOR(0) = x;
    pos.COOR(1) = y;
    pos.COOR(2) = z;
    return Q_TIP4P_Eng(pos, q, atomType, atomPos, nAtoms);
}

// The reference's Cullum-Willoughby driver lancbis (lancbis.cpp:5-105: BISEC,
// ISOEV, INVERR from its Fortran sources) on plain arrays: alpha[niter],
// beta[niter + 1], beta2[niter + 1] (entry 0 unused); the ngood good
// eigenvalues and their error estimates go to eval / err (cap entries at
// most).  Returns ngood.
extern "C" __attribute__((visibility("default"))) int ref_lancbis(int niter, const double *alpha,
                                                                  const double *beta, const double *beta2,
                                                                  double emin, double emax, double *eval,
                                                                  double *err, int cap)
{
    VECT a(niter), b(niter + 1), b2(niter + 1), ev(niter), ee(niter);
    for (int i = 0; i < niter; i++) a(i) = alpha[i];
    for (int i = 0; i <= niter; i++) {
        b(i) = beta[i];
        b2(i) = beta2[i];
    }
    int ngood = 0;
    lancbis(niter, ev, ee, emin, emax, ngood, a, b, b2);
    for (int i = 0; i < ngood && i < cap; i++) {
        eval[i] = ev(i);
        err[i] = ee(i);
    }
    return ngood;
}

// The reference's T_m eigenvectors (TRIVEC through lanczosvectors,
// lanczosEigenValuesVectors.cpp:509-561): evtr[j * niter + i] = component i
// of the vector of eval[j]; mamin[j] the size of the T_m used for it.
extern "C" __attribute__((visibility("default"))) void ref_trivec(int niter, const double *alpha,
                                                                  const double *beta, const double *beta2,
                                                                  int ngood, const double *eval, double *evtr,
                                                                  double *mamin)
{
    double *la = new double[niter], *lb = new double[niter + 1], *lb2 = new double[niter + 1];
    double *le = new double[ngood], *w1 = new double[niter], *w2 = new double[niter];
    for (int i = 0; i < niter; i++) {
        la[i] = alpha[i];
        w1[i] = w2[i] = 0.0;
    }
    lb[0] = lb2[0] = 0.0;
    for (int j = 1; j <= niter; j++) {
        lb[j] = beta[j];
        lb2[j] = beta2[j];
    }
    for (int j = 0; j < ngood; j++) le[j] = eval[j];
    for (int i = 0; i < niter * ngood; i++) evtr[i] = 0.0;
    for (int i = 0; i < ngood; i++) mamin[i] = 0.0;
    FORTRAN(trivec)(la, lb, lb2, w1, w2, &niter, le, &ngood, evtr, mamin);
    delete[] la;
    delete[] lb;
    delete[] lb2;
    delete[] le;
    delete[] w1;
    delete[] w2;
}

// GENRAN (genran.f:4-32) with the GNU runtime's SRAND / RAND
extern "C" __attribute__((visibility("default"))) void ref_genran(int seed, double *r, int n)
{
    genran_(&seed, r, &n);
}
