"""The drop-in calls' host copies on both sides of the staging thresholds
(ecl-suite_amd/csrc/eclcc_stage.hip): a total upload of 64 MiB or more and a
download of 8 MiB or more go through the pinned staging pool, smaller ones
are plain copies.  Each case mixes the two paths in one call and checks the
result against the oracle: labels for eclcc_compute (the reference's
computeCC, ECL-CC-atomic.cu:275-325), the forest for eclmst_compute (gpuMST,
ECL-MST_10-atomic.cu:285-382)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import announce
from eclcc import Graph, compute_cc, mst, workloads
from oracle import oracle

MIB = 1 << 20


def cc_host_call(g):
    lab = np.full(g.nodes, -7, dtype=np.int32)
    compute_cc(g.nodes, g.edges, np.ascontiguousarray(g.nindex, dtype=np.int32),
               np.ascontiguousarray(g.nlist, dtype=np.int32), lab)
    return lab


@pytest.mark.gpu
@pytest.mark.parametrize("n,pairs,up_staged,down_staged", [
    (4_000_000, 5_000_000, False, True),    # 53 MiB of CSR in plainly, 15 MiB of labels out staged
    (1_000_000, 10_000_000, True, False),   # 80 MiB in staged (chunks of both arrays), 4 MiB out plainly
    (2_097_153, 8_388_609, True, True),     # sizes that are no multiple of the 8 MiB chunk
])
def test_cc_host_call_across_the_staging_thresholds(cuda, n, pairs, up_staged, down_staged):
    g = Graph.er(n, pairs, seed=n % 97)
    up = 4 * (g.nodes + 1) + 4 * g.edges
    assert (up >= 64 * MIB) == up_staged and (4 * g.nodes >= 8 * MIB) == down_staged
    announce(f"staging: ER n={g.nodes} m={g.edges}, upload {up / MIB:.1f} MiB, labels {4 * g.nodes / MIB:.1f} MiB")
    lab = cc_host_call(g)
    assert np.array_equal(lab, oracle.cc_labels(g.nindex, g.nlist))
    # a second call through the same pool (its slots and streams reused)
    assert np.array_equal(cc_host_call(g), lab)


@pytest.mark.gpu
def test_mst_host_call_staged_both_ways(cuda):
    g = workloads.rmat(20)  # 120 MiB of entries and as many of weights in, 30 MiB of in_mst out
    w = mst.mst_weights(g)
    assert 4 * (g.nodes + 1) + 8 * g.edges >= 64 * MIB and g.edges >= 8 * MIB
    announce(f"staging: ECL-MST on {g.name}, m={g.edges}")
    got = mst.compute_mst(g, w)
    assert np.array_equal(got, oracle.mst(g.nindex, g.nlist, w))
    assert np.array_equal(mst.compute_mst(g, w), got)
