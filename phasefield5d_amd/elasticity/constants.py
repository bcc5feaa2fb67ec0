"""Material constant loaders (phasefield5d/elasticity/constants.py:58-95).

``load_elastic_constants`` keeps the reference's hard-coded FeMnNiCoCu values
(189.8/123.8/139.2 GPa) as its default; callers that need other constants
(e.g. 240/150/120 GPa) pass them through the ``--c11/--c12/--c44`` flags of the
simulate CLI instead.
"""
from phasefield5d_amd.materials.alloy import build_alloy_constants
from phasefield5d_amd.materials.database import RADIUS_SOURCES

DEFAULT_ELEMENTS = ["Fe", "Mn", "Ni", "Co", "Cu"]
DEFAULT_CIJ_GPA = (189.8, 123.8, 139.2)


def load_material_constants(atomic_radius_tag, alloy=None):
    """Return ``(alloy, ri, mu, nu, Vi, qi)`` for a radius convention."""
    if atomic_radius_tag not in RADIUS_SOURCES:
        raise ValueError(f"Unknown atomic_radius_tag '{atomic_radius_tag}'. "
                         f"Choose from {list(RADIUS_SOURCES)}.")
    return build_alloy_constants(alloy or DEFAULT_ELEMENTS, radius_source=atomic_radius_tag)


def load_elastic_constants(composition=None, temperature=None, override=None):
    """Cubic single-crystal constants (c11, c12, c44) in GPa."""
    if override is not None:
        c11, c12, c44 = (float(v) for v in override)
        return c11, c12, c44
    return DEFAULT_CIJ_GPA
