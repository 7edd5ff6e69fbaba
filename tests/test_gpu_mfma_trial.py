"""The multi-RHS clover product on the matrix cores (v_mfma_f32_16x16x16f16, the 16
right-hand sides of a site as N) gives the vector-ALU product's result: same inputs,
FP16 storage, FP32 accumulation, so the two differ only by the order of the FP32
accumulation and the FP16 roundings that follow (a wrong operand or output lane layout
would differ at O(1))."""
import pytest

import qws_amd as Q

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("V,nrep", [(4096, 1), (4099, 6)])
def test_mfma_clover_product_matches_valu(V, nrep):
    ms_valu, ms_mfma, diff = Q.time_clover_rhs(V, nrep, 2)
    print(f"[mfma] V={V} nrep={nrep} valu {ms_valu:.4f} ms mfma {ms_mfma:.4f} ms maxdiff {diff:.2e}", flush=True)
    assert diff < 4e-3 * nrep, diff
    assert ms_valu > 0 and ms_mfma > 0
