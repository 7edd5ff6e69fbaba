"""CPU checks of the benchmark's algorithmic flop model of the fused SAP launch:
hop counts by enumerating the sites of a block, the launch sequence by replaying
the sweep loop of prec_ddd_s_ (qws.cc:809-925)."""
import itertools

import numpy as np
import pytest

import bench

HOP = {0: 168, 1: 168, 2: 168, 3: 144}
CLV, UPD = 552, 48


def enumerate_hops(B):
    """Average in-block and face hop flops per site, by visiting every site."""
    h_in = h_face = 0.0
    for c in itertools.product(*[range(b) for b in B]):
        for mu in range(4):
            for d in (+1, -1):
                if 0 <= c[mu] + d < B[mu]:
                    h_in += HOP[mu]
                else:
                    h_face += HOP[mu]
    vb = int(np.prod(B))
    return h_in / vb, h_face / vb


def replay(B, nsap, nm):
    h_in, h_face = enumerate_hops(B)
    launches = []
    for isap in range(nsap):
        for c in range(2):
            pro = isap > 0 or c == 1
            launches.append((pro, "supdate"))
    launches += [(nsap > 0, "out"), (True, "out_face")]
    total = 0.0
    for pro, epi in launches:
        f = nm * (h_in + CLV + UPD)
        if pro:
            f += h_face + CLV + UPD
        f += h_in + CLV + {"supdate": 96, "out": 48, "out_face": 96}[epi]
        total += f
    return total / len(launches)


@pytest.mark.parametrize("B", [[4, 4, 4, 4], [8, 4, 4, 2], [2, 2, 2, 2]])
@pytest.mark.parametrize("nsap,nm", [(2, 6), (1, 1), (4, 3)])
def test_sap_flop_model(B, nsap, nm):
    assert bench.sap_flops_per_site(B, nsap, nm) == pytest.approx(replay(B, nsap, nm), rel=1e-12)


def test_sap_flop_model_4444_value():
    # 4^4 blocks: 4.5 spatial + 1.5 temporal in-block hops, 1.5 + 0.5 face hops per site
    h_in = 4.5 * 168 + 1.5 * 144
    h_face = 1.5 * 168 + 0.5 * 144
    assert enumerate_hops([4, 4, 4, 4]) == pytest.approx((h_in, h_face))
    assert bench.sap_flops_per_site([4, 4, 4, 4], 2, 6) == pytest.approx(11814.0)


def test_bench_host_layout_is_the_reference_layout():
    """bench.py's conversion to QWS's host AoSoA (for the drop-in measurement) equals the
    layout the reference-library parity tests use (oracle/ref_worker.py)."""
    import numpy as np

    from oracle import ref_worker as RW
    L = [32, 2, 2, 4]
    V = int(np.prod(L))
    rng = np.random.default_rng(9)
    U, C, s = rng.standard_normal((4, V, 3, 3, 2)), rng.standard_normal((V, 2, 36)), rng.standard_normal((V, 3, 4, 2))
    for vlen in (8, 16):
        assert np.array_equal(bench.qws_host_layout(U, "links", L, vlen), RW.links_to_aosoa(U, L, vlen))
        assert np.array_equal(bench.qws_host_layout(C, "clover", L, vlen), RW.clover_to_aosoa(C, L, vlen))
        assert np.array_equal(bench.qws_host_layout(s, "spinor", L, vlen), RW.spinor_to_aosoa(s, L, vlen))
