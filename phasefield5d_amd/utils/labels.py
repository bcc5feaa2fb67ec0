"""Composition labels for run directories (phasefield5d/utils/labels.py:4-18)."""
import numpy as np


def get_composition_label(composition, unicode=True, elements=None):
    """'Fe20Mn10Ni20Co20Cu30'-style label: the dependent component first, atomic percent rounded."""
    if elements is None:
        elements = ["Fe", "Mn", "Ni", "Co", "Cu"]
    x = np.asarray(composition, dtype=float)
    pct = np.round(np.concatenate([[1.0 - x.sum()], x]) * 100).astype(int)
    if unicode:
        sub = str.maketrans("0123456789", "₀₁₂₃₄₅₆₇₈₉")
        return "".join(f"{el}{str(v).translate(sub)}" for el, v in zip(elements, pct))
    return "".join(f"{el}{v}" for el, v in zip(elements, pct))
