"""Physical constants and standard atomic weights used by the kinetics.

R and SIGMA enter the step (Arrhenius rate and surface radiation); MOLEMASS
(g/mol, upper-case element symbols) turns a species composition into a molar
mass, as the configuration interpreter does for ``composition: {H: 2, O: 1}``.
The values are the same physical data the reference uses
(src/LBM/util/constant.py:17,19,21-115).
"""

c = 299792458  # speed of light in vacuum, m/s
R = 8.314441  # universal gas constant, J/(mol K)
SIGMA = 5.670374419e-8  # Stefan-Boltzmann constant, W/(m^2 K^4)

# standard atomic weights (g/mol), abridged, indexed by upper-case symbol
MOLEMASS = {
    "H": 1.008, "HE": 4.0026, "LI": 6.94, "BE": 9.0122, "B": 10.81, "C": 12.011, "N": 14.007,
    "O": 15.999, "F": 18.998, "NE": 20.18, "NA": 22.99, "MG": 24.305, "AL": 26.982, "SI": 28.085,
    "P": 30.974, "S": 32.06, "CL": 35.45, "AR": 39.95, "K": 39.098, "CA": 40.078, "SC": 44.956,
    "TI": 47.867, "V": 50.941, "CR": 51.996, "MN": 54.938, "FE": 55.845, "CO": 58.933, "NI": 58.693,
    "CU": 63.546, "ZN": 65.38, "GA": 69.723, "GE": 72.63, "AS": 74.922, "SE": 78.971, "BR": 79.904,
    "KR": 83.798, "RB": 85.467, "SR": 87.62, "Y": 88.906, "ZR": 91.224, "NB": 92.906, "MO": 95.95,
    "TC": 98, "RU": 101.07, "RH": 102.91, "PD": 106.42, "AG": 107.87, "CD": 112.41, "IN": 114.82,
    "SN": 118.71, "SB": 121.76, "TE": 127.6, "I": 126.9, "XE": 131.29, "CS": 132.91, "BA": 137.33,
    "LA": 138.91, "CE": 140.12, "PR": 140.91, "ND": 144.24, "PM": 145, "SM": 150.36, "EU": 151.96,
    "GD": 157.25, "TB": 158.93, "DY": 162.5, "HO": 164.93, "ER": 167.26, "TM": 168.93, "YB": 173.04,
    "LU": 174.97, "HF": 178.49, "TA": 180.95, "W": 183.84, "RE": 186.21, "OS": 190.23, "IR": 192.22,
    "PT": 195.08, "AU": 196.97, "HG": 200.59, "TL": 204.38, "PB": 207.2, "BI": 208.98, "PO": 209,
    "AT": 210, "RN": 222, "FR": 223, "RA": 226, "AC": 227, "TH": 232.04, "PA": 231.04, "U": 238.03,
}
