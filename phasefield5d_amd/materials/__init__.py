"""Per-element constants needed by the elastic coupling (radii, molar volumes)."""
from .database import ELEMENT_DATA, RADIUS_SOURCES, get_element_data  # noqa: F401
from .alloy import build_alloy_constants  # noqa: F401
