"""Read a ThermoCalc-style tab-separated CALPHAD table.

Follows phasefield5d/thermodynamics/io.py:9-83 and the column specification
of data/FeMnNiCoCu_fcc/DATA_FORMAT.md:19-49: the first file in ``path`` whose
name contains the temperature tag (case-insensitive) is read; the columns kept
are x(E) and Gm.x(E) for the independent elements and Mob(E) for all five, in
that order, and renamed to X_E / Gx_E / Mob_E.  Column order defines the
9-property layout [mu_Mn, mu_Ni, mu_Co, mu_Cu, Mob_Fe, ..., Mob_Cu].
"""
import os

import pandas as pd

DEFAULT_ELEMENTS = ["Fe", "Mn", "Ni", "Co", "Cu"]


def calphad_columns(elements):
    """(columns to read, rename map) for an element list whose first entry is dependent."""
    indep = elements[1:]
    cols = [f"x({e})" for e in indep] + [f"Gm.x({e})" for e in indep] + [f"Mob({e})" for e in elements]
    rename = {f"x({e})": f"X_{e}" for e in indep}
    rename.update({f"Gm.x({e})": f"Gx_{e}" for e in indep})
    rename.update({f"Mob({e})": f"Mob_{e}" for e in elements})
    return cols, rename


def find_table_file(T, path):
    tag = str(T).lower()
    names = os.listdir(path)
    hits = [f for f in names if tag in f.lower()]
    if not hits:
        raise FileNotFoundError(f"No file containing '{T}' (case-insensitive) found in {path}\n"
                                f"Files present: {names}")
    return os.path.join(path, hits[0])


def load_calphad_dataframe(T, path="data/FeMnNiCoCu_fcc", elements=None):
    """Load the table at temperature tag ``T``; ValueError on missing columns."""
    elements = elements or DEFAULT_ELEMENTS
    cols, rename = calphad_columns(elements)
    full = find_table_file(T, path)
    print(f"Reading CALPHAD file: {os.path.basename(full)}")
    df = pd.read_table(full, sep="\t", lineterminator="\n", index_col=False)
    missing = [c for c in cols if c not in df.columns]
    if missing:
        raise ValueError(f"Missing columns in {os.path.basename(full)}: {missing}")
    df = df[cols].copy()
    df.rename(columns=rename, inplace=True)
    print("Loaded CALPHAD data with shape:", df.shape)
    return df
