"""The research copy with plain (racy) label loads and stores,
lib/libeclcc_plain.so, built only to measure what race freedom costs (as the
reference's racy baseline does, src/baseline/egr-input/CC/ECL-CC_11.cu:81-93,
240-253).  Its labels are checked, not trusted: on the three single-graph
workloads at full size, under both filter modes, they must equal the
oracle's, bit for bit."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import announce
from eclcc import DeviceGraph, default_params, workloads
from eclcc import _native as N
from oracle import oracle

pytestmark = pytest.mark.gpu

WORKLOADS = {
    "er65536": lambda: workloads.er65536(1),
    "grid256": workloads.grid256,
    "rmat24": workloads.rmat24,
}


@pytest.mark.timeout(600)
@pytest.mark.parametrize("name", list(WORKLOADS))
def test_plain_label_build_matches_the_oracle(cuda, name):
    import torch

    g = WORKLOADS[name]()
    expect = oracle.cc_labels(g.nindex, g.nlist)
    P = N.plain_lib()
    assert b"plain labels" in P.eclcc_version()
    stream = torch.cuda.current_stream(cuda)
    for filt in (2, 1):
        params = default_params(giant_filter=filt)
        dg = DeviceGraph(g, "cuda:0", params)
        announce(f"libeclcc_plain.so on {g.name} (n={g.nodes}, m={g.edges}), giant_filter {filt}")
        for rep in range(2):
            dg.nstat.fill_(-1)
            N.check_cc(P.eclcc_run_device(dg.nodes, dg.edges, dg.nindex.data_ptr(), dg.nlist.data_ptr(),
                                          dg.nstat.data_ptr(), dg.work.data_ptr(), dg.work_bytes,
                                          C.byref(params), stream.cuda_stream))
            got = dg.labels()
            bad = np.nonzero(got != expect)[0]
            assert bad.size == 0, f"{g.name} filter {filt} run {rep}: {bad.size} labels differ, first at {bad[:5]}"
        del dg
        torch.cuda.empty_cache()
