"""Physical constants under the reference's names (py/orbit/utils/consts.py)."""
speed_of_light = 2.99792458e8  # m/s
pi = 3.14159265
mass_proton = 0.938272029  # GeV/c^2
mass_electron = 0.000511  # GeV/c^2
mass_electron_mks = 9.10938291e-31  # kg
charge_electron = 1.602176e-19  # C
