"""The batched and served workloads at their stated sizes, checked graph by
graph against the oracle (bit-exact: integer labels, tolerance zero).

  batched: the 4,096-graph mix (n log-uniform on [1e3, 1e6], 50/50
           Erdos-Renyi with average degree U[4, 32] / R-MAT with edge factor
           8, seed 4096), about 1.0e10 CSR entries in total (so the batch's
           64-bit offsets are used past 2^31), with the committed autotuned
           parameters, run as a replayed hipGraph.
  served:  1,024 requests drawn from the same mix (seed 4097), resident in
           HBM, submitted on a seeded Poisson schedule to the continuous-
           batching server; each result equals the oracle's and the same
           graph's labels when it runs alone.

The reference checks each run with its host verifier
(src/racefree/egr-input/CC/ECL-CC-atomic.cu:365-395); the oracle's labels
are the minimum vertex ID per component, which that verifier accepts and
which every interleaving of the kernels produces.
"""
from __future__ import annotations

import ctypes as C
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from conftest import announce
from eclcc import _native as N
from eclcc import autotune, default_params, workloads
from eclcc.batch import Batch
from eclcc.server import Server
from oracle import oracle

pytestmark = pytest.mark.gpu

WORKERS = min(16, len(os.sched_getaffinity(0)))


def oracle_labels(graphs):
    with ThreadPoolExecutor(max_workers=WORKERS) as ex:
        return list(ex.map(lambda g: oracle.cc_labels(g.nindex, g.nlist), graphs))


def first_mismatch(graphs, got, expect):
    for g, a, e in zip(graphs, got, expect):
        bad = np.nonzero(a != e)[0]
        if bad.size:
            return f"{g.name}: {bad.size} labels differ, first at {bad[:5]}"
    return None


def tuned_params(key="batch_mix"):
    p = autotune.load(key)
    assert p is not None, f"tuning/mi355x.json has no {key} entry"
    return p


@pytest.mark.timeout(900)
def test_batched_mix_4096_graphs_against_the_oracle(cuda):
    import torch

    t0 = time.perf_counter()
    graphs = workloads.batch_mix(4096, progress=lambda d, c: announce(f"generated {d}/{c} graphs"))
    n_tot = sum(g.nodes for g in graphs)
    m_tot = sum(g.edges for g in graphs)
    announce(f"generated 4096 mix graphs in {time.perf_counter() - t0:.0f}s: n={n_tot} m={m_tot}")
    assert m_tot > 2 ** 31, "the stated mix must exceed 2^31 entries in total"
    params = tuned_params()
    announce(f"tuned params {params.as_dict()}")
    t0 = time.perf_counter()
    expect = oracle_labels(graphs)
    announce(f"oracle labels in {time.perf_counter() - t0:.0f}s")
    with Batch(graphs, params) as b:
        announce(f"batch uploaded: {b.device_bytes() / 2 ** 30:.1f} GiB")
        stream = torch.cuda.current_stream(cuda)
        # capture once, then replays with the program's own host work between
        for rep in range(3):
            announce(f"hipGraph replay {rep}")
            b.run(stream, use_graph=True)
            torch.cuda.synchronize()
            if rep in (0, 2):
                got = b.all_labels()
                err = first_mismatch(graphs, got, expect)
                assert err is None, f"replay {rep}: {err}"
                del got
        # the graphs far apart in the packed layout (first, last, largest)
        # also run alone, one graph per launch, on the same resident data
        big = max(range(len(graphs)), key=lambda i: graphs[i].edges)
        for i in (0, len(graphs) - 1, big):
            b.run_single(i, stream)
            torch.cuda.synchronize()
            assert np.array_equal(b.labels(i), expect[i]), ("alone", graphs[i].name)


@pytest.mark.timeout(600)
def test_served_poisson_1024_requests_alone_and_in_the_mix(cuda):
    import torch

    R = 1024
    t0 = time.perf_counter()
    graphs = workloads.batch_mix(R, base_seed=workloads.BATCH_SEED + 1,
                                 progress=lambda d, c: announce(f"generated {d}/{c} graphs"))
    announce(f"generated {R} request graphs in {time.perf_counter() - t0:.0f}s, "
             f"m={sum(g.edges for g in graphs)}")
    expect = oracle_labels(graphs)
    params = tuned_params("served")
    # the graph store: every request's CSR resident in HBM
    store = []
    for g in graphs:
        ni = torch.from_numpy(np.ascontiguousarray(g.nindex)).to(cuda)
        nl = (torch.from_numpy(np.ascontiguousarray(g.nlist)).to(cuda) if g.edges
              else torch.zeros(1, dtype=torch.int32, device=cuda))
        store.append((g.nodes, g.edges, ni, nl, torch.full((g.nodes,), -1, dtype=torch.int32, device=cuda)))
    torch.cuda.synchronize()
    L = N.cc_lib()
    # alone: one graph per launch, waiting for each
    wb = int(L.eclcc_workspace_bytes(max(g.nodes for g in graphs), max(g.edges for g in graphs),
                                     C.byref(params)))
    work = torch.empty(wb, dtype=torch.uint8, device=cuda)
    stream = torch.cuda.current_stream(cuda)
    alone = []
    announce("each request alone")
    for n, m, ni, nl, lab in store:
        N.check_cc(L.eclcc_run_device(n, m, ni.data_ptr(), nl.data_ptr(), lab.data_ptr(),
                                      work.data_ptr(), wb, C.byref(params), stream.cuda_stream))
        stream.synchronize()
        alone.append(lab.cpu().numpy())
    err = first_mismatch(graphs, alone, expect)
    assert err is None, f"alone: {err}"
    for *_, lab in store:
        lab.fill_(-1)
    torch.cuda.synchronize()
    # in the mix: a seeded Poisson stream at a stated rate into the server
    qps = 40000.0
    arrivals = np.cumsum(np.random.default_rng(2024).exponential(1.0 / qps, R))
    announce(f"Poisson stream of {R} requests at {qps:.0f}/s")
    with Server(hbm_budget_bytes=16 << 30, streams=3, max_batch=128, params=params,
                batch_wait_us=100) as s:
        sub = L.eclcc_server_submit_device
        ids = []
        t0 = time.perf_counter()
        for i, (n, m, ni, nl, lab) in enumerate(store):
            while time.perf_counter() < t0 + arrivals[i]:
                pass
            ids.append(N.check_cc(int(sub(s._h, n, m, ni.data_ptr(), nl.data_ptr() if m else None,
                                          lab.data_ptr()))))
        lat = C.c_double(0.0)
        lats = []
        for rid in ids:
            N.check_cc(L.eclcc_server_wait(s._h, rid, 120000, C.byref(lat)))
            lats.append(lat.value)
        st = s.stats()
    got = [lab.cpu().numpy() for *_, lab in store]
    err = first_mismatch(graphs, got, expect)
    assert err is None, f"served: {err}"
    assert all(np.array_equal(a, b) for a, b in zip(got, alone)), "served labels differ from alone"
    assert st["completed"] == R and st["max_batch"] > 1, st
    announce(f"served p50 {np.percentile(lats, 50):.3f} ms, p99 {np.percentile(lats, 99):.3f} ms, "
             f"{st['batches']} batches")
