"""MI355X-native engine for SCONE's multigroup tracking loop.

``scone_amd.tables`` compiles SCONE input dictionaries into the flat tables of the C ABI
(include/scone_gpu.h); ``scone_amd.engine`` loads the HIP library (libscone_amd.so) and drives it;
``scone_amd.package`` runs whole SCONE input files (eigenvalue and fixed-source packages) on the
engine and writes SCONE's result files; ``scone_amd.serve`` serves fixed-source solves with
continuous batching.
"""
from . import dictio, tables, configs  # noqa: F401

__all__ = ["dictio", "tables", "configs"]
