"""The allocation-free dEFSApply entry point (dEFSApplyHexBatchedNoAlloc) against dEFSApplyHexBatched: every
mode and insert mode, two calls in a row with the second on a smaller input (fewer elements and a smaller
rule), which must not see the first call's staged tables."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402


@pytest.fixture(scope="module")
def C():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import dohp_amd.capi as capi
    return capi


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def test_no_alloc_apply_matches_allocating_entry_point(C):
    rng = np.random.default_rng(11)
    for amode in (C.APPLY_INTERP, C.APPLY_GRAD, C.APPLY_INTERP_TRANSPOSE, C.APPLY_GRAD_TRANSPOSE):
        for imode in (C.INSERT_VALUES, C.ADD_VALUES):
            for (P, Q, D, ne) in ((7, 7, 1, 9), (4, 5, 3, 4)):  # second call: smaller rule and fewer elements
                print(f"no-alloc amode={amode} imode={imode} P={P} Q={Q} D={D} ne={ne}", flush=True)
                rc, _ = O.gauss_legendre(Q)
                nin = {C.APPLY_INTERP: P ** 3 * D, C.APPLY_GRAD: P ** 3 * D, C.APPLY_INTERP_TRANSPOSE: Q ** 3 * D,
                       C.APPLY_GRAD_TRANSPOSE: Q ** 3 * D * 3}[amode]
                nout = {C.APPLY_INTERP: Q ** 3 * D, C.APPLY_GRAD: Q ** 3 * D * 3,
                        C.APPLY_INTERP_TRANSPOSE: P ** 3 * D, C.APPLY_GRAD_TRANSPOSE: P ** 3 * D}[amode]
                jinv = dev(rng.uniform(-1, 1, (ne, Q ** 3, 9)))
                x = dev(rng.uniform(-1, 1, (ne, nin)))
                y0 = rng.uniform(-1, 1, (ne, nout))
                ya, yb = dev(y0), dev(y0)
                C.efs_apply_hex_batched(P, rc, ne, jinv, D, x, ya, amode, imode)
                C.efs_apply_hex_batched_noalloc(P, rc, ne, jinv, D, x, yb, amode, imode)
                torch.cuda.synchronize()
                assert torch.equal(ya, yb), (amode, imode, P, Q)


def test_no_alloc_rejects_oversized_rule(C):
    rc, _ = O.gauss_legendre(13)
    x = torch.zeros(13 ** 3, dtype=torch.float64, device="cuda")
    with pytest.raises(C.DohpError) as ei:
        C.efs_apply_hex_batched_noalloc(13, rc, 1, None, 1, x, x, C.APPLY_INTERP)
    assert ei.value.code == C.ERR_ARG
