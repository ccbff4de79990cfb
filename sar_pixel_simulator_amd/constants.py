"""Physical constants, pinned to the exact mix the reference uses.

The reference reads some constants from ``scipy.constants`` at run time
(parameters.py:272,293; cpu/main_sim.py:109-110; cpu/incident_sim.py:106-107)
and hard-codes others as CODATA-2018 literals (absorbers.py:50-51,
materials/materials.py:45, near2far.py:55-57).  With scipy >= 1.15 the two
sets differ (CODATA-2022 vs CODATA-2018 vacuum permittivity), so every call
site below names which of the two it follows.
"""
import scipy.constants as _sc

#: Speed of light [m/s] (exact; both sets agree).
C0 = float(_sc.c)
#: Vacuum permittivity from scipy.constants (CODATA-2022 on scipy 1.15).
EPS0_SCIPY = float(_sc.epsilon_0)
#: Vacuum permeability from scipy.constants.
MU0_SCIPY = float(_sc.mu_0)
#: CODATA-2018 vacuum permittivity literal (absorbers.py:51, materials.py:45).
EPS0_2018 = 8.8541878128e-12
#: Free-space impedance literal (absorbers.py:50, near2far.py:55).
ETA0 = 376.730313668
#: Speed-of-light literal used by the far-field transform (near2far.py:57).
C0_N2F = 299792458
