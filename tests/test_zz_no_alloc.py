"""eclcc_run_device_noalloc: the device-resident run whose workspace is the
library's fixed device array (no allocation in the call).  Its labels must
equal eclcc_run_device's (workspace passed in) and the oracle's, bit for
bit, on two calls in a row, the second on a smaller graph, whatever the
first call left in the shared array."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import announce
from eclcc import _native as N
from eclcc import default_params, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20  # ECLCC_NOALLOC_SCRATCH_BYTES


def _upload(g, dev):
    import torch

    ni = torch.from_numpy(np.ascontiguousarray(g.nindex)).to(dev)
    nl = (torch.from_numpy(np.ascontiguousarray(g.nlist)).to(dev) if g.edges
          else torch.zeros(1, dtype=torch.int32, device=dev))
    return ni, nl


def _run_with_workspace(L, g, ni, nl, params, stream, dev):
    import torch

    wb = int(L.eclcc_workspace_bytes(g.nodes, g.edges, C.byref(params)))
    work = torch.empty(wb, dtype=torch.uint8, device=dev)
    lab = torch.full((g.nodes,), -1, dtype=torch.int32, device=dev)
    N.check_cc(L.eclcc_run_device(g.nodes, g.edges, ni.data_ptr(), nl.data_ptr(), lab.data_ptr(),
                                  work.data_ptr(), wb, C.byref(params), stream.cuda_stream))
    stream.synchronize()
    return lab.cpu().numpy()


@pytest.mark.parametrize("kw", [dict(), dict(giant_filter=0, thread_max_deg=2, wave_max_deg=40, hub_chunk=64),
                                dict(giant_filter=1, sample_shift=2, path_mode=0)],
                         ids=["default", "f0-hub-tasks", "f1-shift2-mode0"])
def test_noalloc_equals_run_device_on_two_calls(cuda, kw):
    import torch

    L = N.cc_lib()
    stream = torch.cuda.current_stream(cuda)
    params = default_params(**kw)
    first = workloads.rmat(18, 16, 7)          # 262,144 vertices
    second = workloads.grid3d(40, 0.5, 9)      # 64,000 vertices: smaller
    graphs = [first, second]
    for g in graphs:
        assert int(L.eclcc_workspace_bytes(g.nodes, g.edges, C.byref(params))) <= SCRATCH_BYTES
    dev_graphs = [_upload(g, cuda) for g in graphs]
    torch.cuda.synchronize()
    # the existing entry point first, each graph with its own workspace
    expect = []
    for g, (ni, nl) in zip(graphs, dev_graphs):
        announce(f"eclcc_run_device {g.name} params {kw}")
        lab = _run_with_workspace(L, g, ni, nl, params, stream, cuda)
        assert np.array_equal(lab, oracle.cc_labels(g.nindex, g.nlist)), g.name
        expect.append(lab)
    # then the no-allocation entry point, twice in a row: large, then small
    for g, (ni, nl), exp in zip(graphs, dev_graphs, expect):
        announce(f"eclcc_run_device_noalloc {g.name} params {kw}")
        lab = torch.full((g.nodes,), -1, dtype=torch.int32, device=cuda)
        N.check_cc(L.eclcc_run_device_noalloc(g.nodes, g.edges, ni.data_ptr(), nl.data_ptr(),
                                              lab.data_ptr(), C.byref(params), stream.cuda_stream))
        stream.synchronize()
        got = lab.cpu().numpy()
        bad = np.nonzero(got != exp)[0]
        assert bad.size == 0, f"{g.name}: {bad.size} labels differ from eclcc_run_device, first at {bad[:5]}"


def test_noalloc_refuses_graphs_beyond_its_array(cuda):
    import torch

    L = N.cc_lib()
    params = default_params()
    n = 1 << 24  # needs about 82 MB of workspace
    assert int(L.eclcc_workspace_bytes(n, 0, C.byref(params))) > SCRATCH_BYTES
    announce("eclcc_run_device_noalloc size refusal")
    ni = torch.zeros(n + 1, dtype=torch.int32, device=cuda)
    lab = torch.empty(n, dtype=torch.int32, device=cuda)
    rc = L.eclcc_run_device_noalloc(n, 0, ni.data_ptr(), None, lab.data_ptr(), C.byref(params), None)
    assert rc == -6
    assert b"eclcc_run_device_noalloc holds" in L.eclcc_last_error()
    # a graph that fits still runs after the refusal
    g = workloads.er65536(1)
    ni2, nl2 = _upload(g, cuda)
    lab2 = torch.full((g.nodes,), -1, dtype=torch.int32, device=cuda)
    N.check_cc(L.eclcc_run_device_noalloc(g.nodes, g.edges, ni2.data_ptr(), nl2.data_ptr(),
                                          lab2.data_ptr(), None, None))
    torch.cuda.synchronize()
    assert np.array_equal(lab2.cpu().numpy(), oracle.cc_labels(g.nindex, g.nlist))
