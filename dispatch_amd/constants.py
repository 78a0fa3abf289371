"""The physical and compound constants of the reference (HOP/const_mod.f90), as the
real(8) parameters it declares.  The product's host code forms every factor that
the reference forms from these (save_intensity's conversion, h nu/k of the Hermite
solvers, 2h nu^3/c^2 of the rates) from these values, not from newer CODATA ones:
h differs by 1.6e-5 relative, k by 9e-6.
"""
from __future__ import annotations

EE = 1.602189e-12          # 1 eV in erg                          const_mod.f90:38
HH = 6.626176e-27          # Planck's constant [erg s]            :39
CC = 2.99792458e10         # speed of light [cm/s]                 :40
EM = 9.109534e-28          # electron mass [g]                     :41
UU = 1.6605655e-24         # atomic mass unit [g]                  :42
BK = 1.380662e-16          # Boltzmann's constant [erg/K]          :43
EC = 4.80325e-10           # electron charge [statcoulomb]         :44
PI = 3.14159265359         # the reference's pi                    :14
CM_TO_AA = 1e8             # cm to Angstrom                        :56
E_K = EE / BK              # e/k [K/eV]                            :70
H_K = HH / BK              # h/k [K s]                             :71
TWOH_C2 = 2.0 * HH / CC ** 2   # 2h/c^2 [erg s^3 cm^-2]            :66
H_4PI = HH * (1.0 / (4.0 * PI))  # h/(4 pi) [erg s]                :72-73
