"""kpc - Msun - Gyr constants for the benchmark configurations.

The reference works in code units (hbar/m = 1, Poisson source 4 pi rho,
src/solver/gravity.jl:23); its physical scalings come from
src/simulation_main.jl:94-113.  The configurations here fix hbar/m and G
directly in kpc^2/Gyr and kpc^3 Msun^-1 Gyr^-2 from CODATA values.
"""
import math

HBAR_SI = 1.054571817e-34          # J s
EV_C2_KG = 1.78266192e-36          # kg per eV/c^2
KPC_M = 3.0856775814913673e19      # m
GYR_S = 3.15576e16                 # Julian Gyr in s
GM_SUN_SI = 1.32712440018e20       # m^3 s^-2
KMS_KPC_GYR = 1e3 * GYR_S / KPC_M  # 1 km/s in kpc/Gyr


def hbar_over_m(m22: float = 1.0) -> float:
    """hbar/m in kpc^2/Gyr for a boson of mass m22 * 1e-22 eV."""
    m = m22 * 1e-22 * EV_C2_KG
    return HBAR_SI / m * GYR_S / KPC_M ** 2


G_KPC = GM_SUN_SI * GYR_S ** 2 / KPC_M ** 3  # kpc^3 Msun^-1 Gyr^-2
FOUR_PI_G = 4.0 * math.pi * G_KPC
