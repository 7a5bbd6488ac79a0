"""MI355X-native PruneJuice pruning pipeline (LCC + cycle / path / template
walks + enumeration) over a TOTEM CSR graph held in HBM.

Native parts: lib/libprunejuice_hip.so (gfx950 kernels, include/prunejuice.h)
and lib/libpj_host.so (formats, generators, constraint generation, step
machine, include/prunejuice_host.h).
"""
from .api import (STEP_CC, STEP_ENUMERATION, STEP_LCC, STEP_LCC0, STEP_PC, STEP_TDS, Batch, Constraints, Graph,
                  LabelStat, Pruner, State, run_cli)
from .native import STEP_NAMES, AlgoResults, Row

__all__ = ["Batch", "Graph", "Constraints", "LabelStat", "Pruner", "State", "run_cli", "AlgoResults", "Row", "STEP_NAMES", "STEP_LCC0",
           "STEP_LCC", "STEP_CC", "STEP_PC", "STEP_TDS", "STEP_ENUMERATION"]
