"""eclcc -- MI355X (gfx950) race-free ECL-CC connected components.

Python mirror of the reference's host interface for the CC path
(src/racefree/egr-input/CC/ECL-CC-atomic.cu): ``compute_cc`` is
``computeCC``, ``Graph.read``/``Graph.write`` are ``readECLgraph`` /
``writeECLgraph`` (library/ECLgraph.h).  All computation runs in the in-tree
HIP library libeclcc.so; importing works without a GPU, calling needs one.
"""
from ._native import EclccError, EclccParams, EclgraphError, NativeLibraryMissing
from .graph import Graph
from .cc import (DeviceGraph, compute, compute_cc, compute_timed, default_params,
                 params_for_graph, version)
from . import roofline, workloads

__all__ = [
    "Graph", "DeviceGraph", "compute", "compute_cc", "compute_timed", "default_params",
    "params_for_graph",
    "version", "workloads", "EclccError", "EclgraphError", "EclccParams", "NativeLibraryMissing",
]
