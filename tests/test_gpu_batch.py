"""GPU parity of the batched and served forms: every graph's labels equal
the oracle's and the plain single-graph run's, alone and in the mix."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import announce
from eclcc import compute, default_params, workloads
from eclcc.batch import Batch
from eclcc.server import Server
from oracle import oracle

pytestmark = pytest.mark.gpu


def small_mix(count, seed=77, n_hi=20000):
    return workloads.batch_mix(count, base_seed=seed, n_lo=1000, n_hi=n_hi)


def test_batch_matches_oracle_and_single(cuda):
    import torch

    gs = small_mix(24)
    announce(f"batch of {len(gs)} graphs, total m={sum(g.edges for g in gs)}")
    expect = [oracle.cc_labels(g.nindex, g.nlist) for g in gs]
    with Batch(gs) as b:
        b.run(use_graph=False)
        torch.cuda.synchronize()
        for g, e, lab in zip(gs, expect, b.all_labels()):
            assert np.array_equal(lab, e), g.name
        # hipGraph: capture once, replay several times with host work between
        for rep in range(4):
            b.run(use_graph=True)
            torch.cuda.synchronize()
            got = b.all_labels()
            for g, e, lab in zip(gs, expect, got):
                assert np.array_equal(lab, e), f"replay {rep}: {g.name}"
        # one graph per launch on the same resident data
        for i in range(len(gs)):
            b.run_single(i)
        torch.cuda.synchronize()
        for g, e, lab in zip(gs, expect, b.all_labels()):
            assert np.array_equal(lab, e), g.name


@pytest.mark.parametrize("filt", [0, 1, 2])
def test_batch_with_tuned_params(cuda, filt):
    import torch

    gs = small_mix(10, seed=5, n_hi=60000) + [workloads.grid3d(20, 0.2, 3), workloads.rmat(12, 2, 4)]
    announce(f"batch with non-default params, giant_filter={filt}")
    p = default_params(thread_max_deg=2, wave_max_deg=32, hub_chunk=256, block_threads=512,
                       giant_filter=filt, sample_shift=filt)
    expect = [oracle.cc_labels(g.nindex, g.nlist) for g in gs]
    with Batch(gs, p) as b:
        for rep in range(2):
            b.run(use_graph=True)
            torch.cuda.synchronize()
            for g, e, lab in zip(gs, expect, b.all_labels()):
                assert np.array_equal(lab, e), (rep, g.name)
        b.set_params(default_params(giant_filter=filt))
        b.run(use_graph=True)
        torch.cuda.synchronize()
        for g, e, lab in zip(gs, expect, b.all_labels()):
            assert np.array_equal(lab, e), g.name


def test_server_alone_and_in_mix(cuda):
    gs = small_mix(40, seed=91, n_hi=50000)
    announce(f"server with {len(gs)} requests")
    alone = [compute(g) for g in gs]
    with Server(hbm_budget_bytes=256 << 20, streams=3, max_batch=16) as s:
        # one at a time
        for g, a in zip(gs[:5], alone[:5]):
            lab, lat = s.wait(s.submit(g))
            assert np.array_equal(lab, a) and lat > 0
        # all at once: batched and overlapped
        ids = [s.submit(g) for g in gs]
        for g, a, rid in zip(gs, alone, ids):
            lab, lat = s.wait(rid, timeout_ms=60000)
            assert np.array_equal(lab, a), g.name
            assert np.array_equal(lab, oracle.cc_labels(g.nindex, g.nlist)), g.name
        st = s.stats()
        assert st["completed"] == len(gs) + 5
        assert st["max_batch"] > 1


def test_server_rejects_request_over_budget(cuda):
    g = workloads.rmat(14, 16, 3)
    announce("server budget rejection")
    from eclcc import EclccError

    with Server(hbm_budget_bytes=1 << 20, streams=1) as s:
        with pytest.raises(EclccError) as ei:
            s.submit(g)
        assert ei.value.status == -6


def test_server_rejects_at_submit_what_staging_cannot_hold(cuda):
    # 8 MiB budget -> 1 MiB of pinned staging.  A batch of one host request
    # stages its labels (align256(4n)) plus the graph table (256 B) and the
    # chunk map (4 B per 256 slots).  n = 262,080: the labels alone fit
    # (1,048,320 B) but not with the tables, so the worker could never admit
    # it; it must be refused at submit, not left at the head of the queue.
    # n = 261,000 fills the staging exactly (1,044,224 + 256 + 4,096 B) and
    # is served, as is a request queued behind the refused one.
    import graphs as G
    from eclcc import EclccError

    announce("server staging boundary at an 8 MiB budget")
    big, fits, small = G.path(262080), G.path(261000), G.path(1000)
    with Server(hbm_budget_bytes=8 << 20, streams=1) as s:
        with pytest.raises(EclccError) as ei:
            s.submit(big)
        assert ei.value.status == -6
        r1, r2 = s.submit(fits), s.submit(small)
        lab, _ = s.wait(r1, timeout_ms=60000)
        assert np.array_equal(lab, np.zeros(fits.nodes, dtype=np.int32))
        lab, _ = s.wait(r2, timeout_ms=60000)
        assert np.array_equal(lab, np.zeros(small.nodes, dtype=np.int32))


def test_server_device_resident_requests(cuda):
    import torch

    gs = small_mix(30, seed=93, n_hi=60000)
    announce(f"server with {len(gs)} device-resident requests mixed with host requests")
    dev = torch.device("cuda:0")
    res = []
    for g in gs:
        ni = torch.from_numpy(np.ascontiguousarray(g.nindex)).to(dev)
        nl = torch.from_numpy(np.ascontiguousarray(g.nlist)).to(dev) if g.edges else \
            torch.zeros(1, dtype=torch.int32, device=dev)
        lab = torch.full((g.nodes,), -1, dtype=torch.int32, device=dev)
        res.append((ni, nl, lab))
    torch.cuda.synchronize()
    with Server(hbm_budget_bytes=128 << 20, streams=2, max_batch=32) as s:
        ids = []
        for i, (g, (ni, nl, lab)) in enumerate(zip(gs, res)):
            ids.append(("dev", s.submit_device(g.nodes, g.edges, ni, nl, lab)))
            if i % 3 == 0:
                ids.append(("host", s.submit(g)))
        k = 0
        for kind, rid in ids:
            out, lat = s.wait(rid, timeout_ms=60000)
            got = out.cpu().numpy() if kind == "dev" else out
            g = gs[k] if kind == "dev" else gs[k - 1]
            if kind == "dev":
                k += 1
            assert lat > 0
            assert np.array_equal(got, oracle.cc_labels(g.nindex, g.nlist)), (kind, g.name)
        assert s.stats()["max_batch"] > 1
    with pytest.raises(ValueError):
        with Server(hbm_budget_bytes=64 << 20, streams=1) as s:
            s.submit_device(gs[0].nodes, gs[0].edges, res[0][0].cpu(), res[0][1], res[0][2])


def test_server_bulk_device_requests(cuda):
    import torch

    gs = small_mix(40, seed=95, n_hi=40000)
    announce(f"server bulk submission of {len(gs)} device-resident requests")
    dev = torch.device("cuda:0")
    store = []
    for g in gs:
        ni = torch.from_numpy(np.ascontiguousarray(g.nindex)).to(dev)
        nl = torch.from_numpy(np.ascontiguousarray(g.nlist)).to(dev) if g.edges else \
            torch.zeros(1, dtype=torch.int32, device=dev)
        store.append((g.nodes, g.edges, ni, nl, torch.full((g.nodes,), -1, dtype=torch.int32, device=dev)))
    torch.cuda.synchronize()
    with Server(hbm_budget_bytes=128 << 20, streams=2, max_batch=16) as s:
        for rep in range(2):
            ids = s.submit_device_many(store)
            assert len(set(ids)) == len(gs)
            for g, rid in zip(gs, ids):
                lab, lat = s.wait(rid, timeout_ms=60000)
                assert np.array_equal(lab.cpu().numpy(), oracle.cc_labels(g.nindex, g.nlist)), (rep, g.name)
        assert s.stats()["max_batch"] > 1
        # a bad request anywhere rejects the whole call before anything is queued
        done = s.stats()["completed"]
        bad = list(store[:3]) + [(0, 0, store[0][2], store[0][3], store[0][4])]
        with pytest.raises(Exception):
            s.submit_device_many(bad)
        assert s.stats()["completed"] == done


@pytest.mark.parametrize("kw", [dict(), dict(giant_filter=0, thread_max_deg=2, wave_max_deg=40, hub_chunk=64),
                                dict(giant_filter=1, sample_shift=3),
                                dict(giant_filter=2, thread_max_deg=0, wave_max_deg=1, path_mode=0),
                                dict(path_mode=2, sample_shift=2), dict(path_mode=2, giant_filter=0),
                                dict(path_mode=2, sample_min_deg=16)],
                         ids=["default", "f0-small-bins", "f1-shift3", "f2-all-hubs", "halve-shift2",
                              "halve-f0", "halve-long-lists"])
def test_batch_of_irregular_graphs(cuda, kw):
    # unsorted lists, duplicates, self-loops, stars with the hub anywhere,
    # permuted paths, shuffled cliques, sizes from 1 to 3000 (most not a
    # multiple of 64): each graph's labels equal the oracle's, in the batch
    # and in the one-graph-per-launch form
    import graphs as G
    import torch

    gs = [G.random_structure(s) for s in range(48)]
    expect = [oracle.cc_labels(g.nindex, g.nlist) for g in gs]
    announce(f"batch of {len(gs)} irregular graphs, params {kw}")
    with Batch(gs, default_params(**kw)) as b:
        b.run(use_graph=True)
        torch.cuda.synchronize()
        for g, e, lab in zip(gs, expect, b.all_labels()):
            assert np.array_equal(lab, e), g.name
        for i in range(len(gs)):
            b.run_single(i)
        torch.cuda.synchronize()
        for g, e, lab in zip(gs, expect, b.all_labels()):
            assert np.array_equal(lab, e), ("single", g.name)


@pytest.mark.parametrize("kw", [dict(), dict(wave_max_deg=64, hub_chunk=64), dict(thread_max_deg=4, wave_max_deg=8, hub_chunk=64, path_mode=2)],
                         ids=["default", "wave64-chunk64", "wave8-chunk64-halve"])
def test_batch_with_stars_outside_each_largest_tree(cuda, kw):
    # graphs whose star centres lie outside their largest tree, between
    # ordinary mix graphs: with filter 2 mark lists the centres of all graphs
    # in one hub list, and a task's graph comes from its hub's slot
    import graphs as G
    import torch

    gs = []
    for i in range(6):
        gs.append(G.er_with_stars(20_000 + 3000 * i, 80_000, 1 + i % 3, 4000 + 1000 * i, seed=40 + i))
        gs += small_mix(2, seed=300 + i, n_hi=30000)
    expect = [oracle.cc_labels(g.nindex, g.nlist) for g in gs]
    announce(f"batch of {len(gs)} graphs with stars outside their largest trees, params {kw}")
    with Batch(gs, default_params(**kw)) as b:
        for rep in range(3):
            b.run(use_graph=rep > 0)
            torch.cuda.synchronize()
            for g, e, lab in zip(gs, expect, b.all_labels()):
                assert np.array_equal(lab, e), (rep, g.name)
        for i in range(len(gs)):
            b.run_single(i)
        torch.cuda.synchronize()
        for g, e, lab in zip(gs, expect, b.all_labels()):
            assert np.array_equal(lab, e), ("single", g.name)


def test_server_on_irregular_graphs_under_a_tight_budget(cuda):
    # requests of every irregular shape (down to one vertex) submitted at
    # once to a server whose small HBM budget forces many small batches:
    # each result equals the oracle's
    import graphs as G

    gs = [G.random_structure(s) for s in range(100, 164)]
    announce(f"server with {len(gs)} irregular requests, 8 MiB budget")
    with Server(hbm_budget_bytes=8 << 20, streams=4, max_batch=8) as s:
        ids = [s.submit(g) for g in gs]
        for g, rid in zip(gs, ids):
            lab, lat = s.wait(rid, timeout_ms=60000)
            assert np.array_equal(lab, oracle.cc_labels(g.nindex, g.nlist)), g.name
        st = s.stats()
        assert st["completed"] == len(gs) and st["batches"] >= len(gs) // 8
