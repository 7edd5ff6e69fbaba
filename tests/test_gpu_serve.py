"""Solve server: resident gauge fields, requests of mixed volume / kappa / RHS count
admitted as slots and memory free up into the batch that runs (continuous batching:
every SAP half-sweep one launch over all admitted requests' lattices).  Each result
is the same alone and in the mix (bit for bit: equal digests of the FP64 solutions,
equal iteration counts) and meets the FP64 true residual."""
import numpy as np
import pytest

import qws_amd as Q

pytestmark = pytest.mark.gpu

FIELDS = [([8, 8, 8, 8], [4, 4, 4, 4], 0.3, 8), ([8, 8, 8, 16], [4, 4, 4, 4], 0.3, 9),
          ([16, 8, 8, 8], [4, 4, 4, 4], 0.3, 10)]
PRM = dict(nsap=4, nm=2, tol=1e-11, maxiter=40, tol_s=1e-3, maxiter_s=40, true_tol=1e-10, max_restarts=4)


def _trace(n, seed=7, rate=200.0):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    return [dict(field=int(rng.integers(len(FIELDS))), kappa=float(rng.uniform(0.110, 0.125)),
                 nrhs=int(rng.integers(1, 5)), seed=100 + i, arrival=float(t[i])) for i in range(n)]


@pytest.fixture(scope="module")
def server():
    s = Q.Server(FIELDS, links32=12, slots=4, **PRM)
    yield s
    s.close()


def test_same_alone_and_in_the_mix(server):
    tr = _trace(10)
    mix = server.run(tr)
    for i, r in enumerate(tr):
        alone = server.run([dict(r, arrival=0.0)])[0]
        assert alone["digest"] == mix[i]["digest"], i
        assert alone["inner_iters"] == mix[i]["inner_iters"]
        assert mix[i]["true_residual_max"] <= 1e-10
    # admitted in arrival order, never before arrival
    adm = [m["admit"] for m in mix]
    arr = [m["arrival"] for m in mix]
    order = np.argsort(arr, kind="stable")
    assert all(adm[order[k]] <= adm[order[k + 1]] + 1e-9 for k in range(len(order) - 1))
    assert all(m["admit"] >= m["arrival"] - 1e-6 for m in mix)
    # with 4 slots some requests ran at the same time
    iv = sorted((m["admit"], m["finish"]) for m in mix)
    assert any(iv[k + 1][0] < iv[k][1] for k in range(len(iv) - 1))


def test_memory_admission():
    """A budget that holds one request at a time serialises the requests."""
    one = max(Q.Server.request_bytes(int(np.prod(f[0])), 4) for f in FIELDS)
    s = Q.Server(FIELDS, links32=12, slots=4, mem_budget=one + 1.0, **PRM)
    try:
        tr = _trace(6, seed=3, rate=1000.0)
        for r in tr:
            r["nrhs"] = 4
        res = s.run(tr)
    finally:
        s.close()
    iv = sorted((m["admit"], m["finish"]) for m in res)
    assert all(iv[k + 1][0] >= iv[k][1] - 1e-6 for k in range(len(iv) - 1))
    assert all(m["true_residual_max"] <= 1e-10 for m in res)


def test_request_in_a_larger_pooled_slot():
    """A request served from a free pooled slot of larger RHS capacity (the smallest
    free one on its field with room) gives the same result as from a slot of its own
    size, and takes no new memory: with a budget for one 8-RHS slot, a 2-RHS request
    after an 8-RHS one on the same field reuses that slot."""
    f = FIELDS[2]
    V = int(np.prod(f[0]))
    small = dict(field=2, kappa=0.12, nrhs=2, seed=501, arrival=0.0)
    big = dict(field=2, kappa=0.118, nrhs=8, seed=502, arrival=0.0)
    s1 = Q.Server(FIELDS, links32=12, slots=4, **PRM)
    try:
        own = s1.run([small])[0]  # a fresh pool: a 2-RHS slot
    finally:
        s1.close()
    budget = Q.Server.request_bytes(V, 8) + 1.0
    s2 = Q.Server(FIELDS, links32=12, slots=4, mem_budget=budget, **PRM)
    try:
        first = s2.run([big])[0]
        again = s2.run([small])[0]
    finally:
        s2.close()
    assert first["true_residual_max"] <= 1e-10
    assert again["bytes"] == first["bytes"] > own["bytes"]  # served by the 8-RHS slot
    assert again["digest"] == own["digest"]
    assert again["inner_iters"] == own["inner_iters"] and again["outer_iters"] == own["outer_iters"]
    assert again["true_residual_max"] <= 1e-10
