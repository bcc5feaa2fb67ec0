"""Composition labels for run directories (phasefield5d/utils/labels.py:4-18).

The label names every element with its atomic percent, the dependent component
first (its fraction is one minus the others), rounded half to even: for example
``Fe20Mn10Ni20Co20Cu30``, with subscript digits when ``unicode``."""
import numpy as np

DEFAULT_ELEMENTS = ("Fe", "Mn", "Ni", "Co", "Cu")
_SUBSCRIPT = {str(d): chr(0x2080 + d) for d in range(10)}


def atomic_percent(composition):
    """Integer atomic percent of every element, the dependent one first."""
    x = np.asarray(composition, dtype=float)
    return np.rint(np.append(1.0 - x.sum(), x) * 100).astype(int)


def get_composition_label(composition, unicode=True, elements=None):
    names = DEFAULT_ELEMENTS if elements is None else elements
    parts = []
    for name, pct in zip(names, atomic_percent(composition)):
        digits = str(pct)
        if unicode:
            digits = "".join(_SUBSCRIPT.get(c, c) for c in digits)
        parts.append(name + digits)
    return "".join(parts)
