"""Unit constants of the SlaKoNet hot path.

Values follow slakonet/units.py:62-68 (DFTB+ constants) for input
conversion, and slakonet/main.py:34 for the eV factor used on outputs.
"""

# Bohr -> Angstrom (slakonet/units.py:62)
BOHR_AA = 0.529177249
# Angstrom -> Bohr (slakonet/units.py:64)
AA_BOHR = 1.0 / BOHR_AA
# Hartree -> eV for table/input conversion (slakonet/units.py:66)
HARTREE_EV = 27.2113845
# Hartree -> eV used by every reported property (slakonet/main.py:34)
H2E = 27.211

# Default physics knobs of the reference path.
HS_CUTOFF_BOHR = 10.0       # hs_matrix(cutoff=10.0), slakonet/slaterkoster.py:111
FERMI_KT_HA = 0.01          # fermi_search(kT=0.01), slakonet/fermi.py:143
FERMI_TOL = 1e-6            # fermi_search(tol=1e-6), slakonet/fermi.py:144
FERMI_MAX_ITER = 100        # fermi_search(max_iter=100), slakonet/fermi.py:145
METAL_THRESHOLD_EV = 1e-3   # calculate_band_gap, slakonet/main.py:666

# Species symbols Z=0..99 (index == atomic number), slakonet/elements.py:22.
CHEMICAL_SYMBOLS = (
    "X H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe "
    "Co Ni Cu Zn Ga Ge As Se Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn "
    "Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W "
    "Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf "
    "Es"
).split()
SYMBOL_TO_Z = {s: z for z, s in enumerate(CHEMICAL_SYMBOLS)}


def generate_shell_dict_upto_Z65():
    """Shell lists per Z, following slakonet/main.py:37-59."""
    shell_dict = {}
    for z in range(1, 100):
        if z <= 2:
            shell_dict[z] = [0]
        elif z <= 20:
            shell_dict[z] = [0, 1]
        elif z <= 30:
            shell_dict[z] = [0, 1, 2]
        elif z <= 36:
            shell_dict[z] = [0, 1]
        elif z <= 48:
            shell_dict[z] = [0, 1, 2]
        elif z <= 54:
            shell_dict[z] = [0, 1]
        elif z <= 57:
            shell_dict[z] = [0, 1, 2]
        else:
            shell_dict[z] = [0, 1, 2, 3]
    return shell_dict
