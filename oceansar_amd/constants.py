"""Physical constants used on the raw-data generation path.

Values follow oceansar/constants/constants.py:27-50 (which star-imports
scipy.constants and then overrides a few): g and G come from scipy.constants,
c / r_earth / m_earth / epsilon_sw are the module's own literals.
"""
import scipy.constants as _sc

g = _sc.g                      # 9.80665 m/s^2 (scipy.constants, constants.py:27)
G = _sc.G                      # gravitational constant (scipy.constants)
c = 2.99792458e8               # constants.py:31
r_earth = 6378.137e3           # constants.py:29
m_earth = 5.9742e24            # constants.py:30
epsilon_sw = 73 + 18j          # sea-water relative permittivity, constants.py:50
