"""Physical constants and model parameters (reference units: nm, ps, kJ/mol,
amu, e; lanczosUnits.h:5-14)."""
import math

H_BAR = 0.0635077857002           # kJ/mol ps (lanczosUnits.h:9)
K_COULOMB = 138.9354846111096     # kJ/mol nm / e^2 (lanczosUnits.h:8)
KJMOL_PER_CM1 = 0.0119626566      # 1 cm^-1 in kJ/mol
C_CM_PER_PS = 2.99792458e-2       # speed of light, cm/ps

# Alavi H2 (Alavi_Parameters.h:5-14)
Q_H2_H = 0.4932
Q_H2_CM = -0.9864
H2_HALF_BOND = 0.03707            # nm (0.7414 Angstrom bond)

# SPC/E water with the reference's rounded O-H2 LJ (SPCE_AH2_EngRoutines.h:17-38)
Q_O_SPCE = -0.8476
Q_H_SPCE = 0.4238
LJ_A_REF = 1.364e-06              # kJ nm^12/mol
LJ_B_REF = 1.533e-3               # kJ nm^6/mol

# reference production values (Compilation/HvInputFile_*_Base.txt)
MASS_H2_REF = 2.01588
ROT_CONST_REF = 0.709648713019339  # kJ/mol = 59.322 cm^-1


def cm1_to_kjmol(x):
    return x * KJMOL_PER_CM1


def omega_from_cm1(nu_cm1):
    """Angular frequency (1/ps) of a vibration of nu_cm1 wavenumbers."""
    return 2.0 * math.pi * C_CM_PER_PS * nu_cm1


def lj_ab(eps_a, sigma_a, eps_b, sigma_b, mixing="geometric"):
    """A = 4 eps sigma^12, B = 4 eps sigma^6 for the O - H2 pair.

    mixing "geometric" (the reference's rule, SPCE_AH2_EngRoutines.h:35-38:
    eps = sqrt(eps_a eps_b), sigma = sqrt(sigma_a sigma_b)) or
    "lorentz-berthelot" (sigma arithmetic mean).  sigma in nm, eps in kJ/mol."""
    eps = math.sqrt(eps_a * eps_b)
    if mixing == "geometric":
        sig = math.sqrt(sigma_a * sigma_b)
    else:
        sig = 0.5 * (sigma_a + sigma_b)
    return 4.0 * eps * sig ** 12, 4.0 * eps * sig ** 6
