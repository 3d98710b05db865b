"""Molecule and basis container in libcint layout."""
from .mole import (M, Mole, BOHR, cart2sph_block, cart_components, ncart, nsph,
                   ATOM_OF, ANG_OF, NPRIM_OF, NCTR_OF, PTR_EXP, PTR_COEFF, PTR_COORD,
                   BAS_SLOTS, ATM_SLOTS)
