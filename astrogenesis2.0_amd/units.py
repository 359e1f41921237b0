"""SI constants and unit conversions of the reference.

Values are those of simulation/src/Math/Constants.h:7-18 and
simulation/src/Math/Units.h:6-19; the engine works in SI like the reference.
"""

G = 6.67430e-11            # Constants.h:7
PI = 3.14159265358979323846
GAMMA = 5.0 / 3.0          # Constants.h:15
K_B = 1.38064852e-23       # Constants.h:16
M_PROTON = 1.6726219e-27   # Constants.h:18 (prtn)
MU_DEFAULT = 0.58          # Particle.h:41

PC = 3.08567758149137e16   # Units.h:7-9
KPC = 3.08567758149137e19
MPC = 3.08567758149137e22
YR = 31557600.0            # Units.h:12-14
KYR = 3.15576e10
MYR = 3.15576e13
MSUN = 1.98847e30          # Units.h:16
KMS = 1.0e3                # Units.h:19

# G in kpc (km/s)^2 / Msun implied by the constants above.
G_KPC_KMS_MSUN = G * MSUN / (KPC * KMS * KMS)

TYPE_STAR, TYPE_GAS, TYPE_DM = 1, 2, 3   # Particle.h:32


def u_from_temperature(T, mu=MU_DEFAULT):
    """Internal energy per mass for temperature T; inverse of the
    T = (gamma-1) U m_p mu / k_B relation of Node.cpp:791."""
    return T * K_B / ((GAMMA - 1.0) * M_PROTON * mu)
