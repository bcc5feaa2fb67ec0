"""Element property table for the Fe-Mn-Ni-Co-Cu system.

The values are the ones tabulated in phasefield5d/materials/database.py:34-90
(Fe, Mn, Ni, Co, Cu entries): four atomic-radius conventions in nm, shear
modulus in GPa, Poisson ratio and molar volume in m^3/mol.  Only these five
elements are needed by the explicit step (Vegard misfit strains); the other
entries and the optional ``mendeleev`` lookup of the reference are not on the
step and are not carried.
"""

RADIUS_SOURCES = ("senkov", "kittel", "riedlich", "miracle")

#            senkov  kittel  riedlich miracle  G[GPa]  nu    Vm[m3/mol]
_ROWS = {
    "Fe": (1.241, 1.270, 1.274, 1.260, 82.0, 0.29, 7.09e-6),
    "Mn": (1.350, 1.260, 1.264, 1.320, 76.4, 0.25, 7.35e-6),
    "Ni": (1.246, 1.250, 1.246, 1.240, 76.0, 0.31, 6.59e-6),
    "Co": (1.251, 1.250, 1.252, 1.260, 75.0, 0.31, 6.67e-6),
    "Cu": (1.278, 1.280, 1.278, 1.260, 48.0, 0.34, 7.11e-6),
}

ELEMENT_DATA = {
    sym: {
        **{f"atomic_radius_{src}": row[i] for i, src in enumerate(RADIUS_SOURCES)},
        "shear_modulus_gpa": row[4],
        "poissons_ratio": row[5],
        "molar_volume_m3mol": row[6],
    }
    for sym, row in _ROWS.items()
}


def get_element_data(symbol):
    """Property dict for ``symbol``; ValueError for elements outside the table."""
    key = symbol.strip().capitalize()
    if key not in ELEMENT_DATA:
        raise ValueError(f"Element '{symbol}' is not in the built-in table "
                         f"(known: {sorted(ELEMENT_DATA)})")
    return ELEMENT_DATA[key]
