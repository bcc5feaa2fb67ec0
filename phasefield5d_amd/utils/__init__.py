"""Small numeric helpers shared by host-side setup code."""
from .statistics import round_to_first_nonzero, round_down_to_first_nonzero  # noqa: F401
