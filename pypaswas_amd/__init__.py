"""pypaswas_amd — pyPaSWAS's Smith-Waterman scoring and traceback on AMD MI355X (gfx950).

The product is the HIP library ``libpypaswas_hip.so`` (C ABI in
include/pypaswas_hip.h) and the Python host that mirrors pyPaSWAS's
SmithWaterman engine interface and Aligner/Trimmer plugin contract.
"""
from pypaswas_amd.cli import normalize_file_path, parse_cli, set_logger  # noqa: F401
