"""CALPHAD table ingestion (TSV, DATA_FORMAT column layout) and a synthetic generator."""
from .io import load_calphad_dataframe, calphad_columns  # noqa: F401
from .synthetic import SyntheticTable, make_synthetic_table, write_synthetic_table  # noqa: F401
