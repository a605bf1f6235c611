"""Boundary entry points added beside the operator: the batched dRuleComputePhysical against the oracle,
and the device-detected error codes (negative and non-finite Jacobians, non-finite inputs of the residual,
the MATSHELL product and the solver)."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from dohp_amd.mesh import block_vertices, splitmix64_uniform  # noqa: E402
from oracle import oracle as O  # noqa: E402


@pytest.fixture(scope="module")
def C():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import dohp_amd.capi as capi
    return capi


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


@pytest.mark.parametrize("Q,nelem", [(1, 3), (2, 5), (5, 7), (7, 11), (11, 2)])
def test_rule_compute_physical_matches_oracle(C, Q, nelem):
    print(f"rule physical Q={Q} nelem={nelem}", flush=True)
    rng = np.random.default_rng(Q * 13 + nelem)
    _, w1 = O.gauss_legendre(Q)
    jac = rng.uniform(-1, 1, (nelem, Q ** 3, 3, 3)) + 2.5 * np.eye(3)
    jinv, jw = torch.zeros(nelem, Q ** 3, 9, dtype=torch.float64, device="cuda"), torch.zeros(
        nelem, Q ** 3, dtype=torch.float64, device="cuda")
    C.rule_compute_physical_batched(w1, nelem, dev(jac), jinv, jw)
    wt = np.einsum("i,j,k->ijk", w1, w1, w1).reshape(-1)
    dp = ctypes.POINTER(ctypes.c_double)
    for e in range(nelem):
        ri, rw = np.zeros((Q ** 3, 9)), np.zeros(Q ** 3)
        je = np.ascontiguousarray(jac[e])
        assert O.lib().orc_rule_compute_physical(Q ** 3, wt.ctypes.data_as(dp), je.ctypes.data_as(dp),
                                                 ri.ctypes.data_as(dp), rw.ctypes.data_as(dp)) == 0
        np.testing.assert_allclose(jinv[e].cpu().numpy(), ri, rtol=1e-14, atol=1e-15)
        np.testing.assert_allclose(jw[e].cpu().numpy(), rw, rtol=1e-14)


def test_rule_compute_physical_errors(C):
    Q, nelem = 3, 4
    _, w1 = O.gauss_legendre(Q)
    jac = np.tile(np.eye(3), (nelem, Q ** 3, 1, 1))
    jinv = torch.zeros(nelem, Q ** 3, 9, dtype=torch.float64, device="cuda")
    jw = torch.zeros(nelem, Q ** 3, dtype=torch.float64, device="cuda")
    C.rule_compute_physical_batched(w1, nelem, dev(jac), jinv, jw)
    np.testing.assert_allclose(jinv.cpu().numpy().reshape(-1, 3, 3), np.tile(np.eye(3), (nelem * Q ** 3, 1, 1)))
    neg = jac.copy()
    neg[2, 5] = np.diag([1.0, -1.0, 1.0])
    with pytest.raises(C.DohpError) as ei:
        C.rule_compute_physical_batched(w1, nelem, dev(neg), jinv, jw)
    assert ei.value.code == C.ERR_NEGATIVE_JACOBIAN
    nan = jac.copy()
    nan[1, 0, 1, 2] = np.nan
    with pytest.raises(C.DohpError) as ei:
        C.rule_compute_physical_batched(w1, nelem, dev(nan), jinv, jw)
    assert ei.value.code == C.ERR_NAN


def _op(C, M=(3, 3, 3), p=4, nonlinear=True):
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=2)
    fs = C.FS(M, p, V)
    op = C.Ellip(fs, p=1.5 if nonlinear else 2.0, eps=float(np.sqrt(0.005)))
    return fs, op


@pytest.mark.parametrize("storage", [0, 1])
def test_nonfinite_inputs_return_nan_code(C, storage):
    print(f"nan codes storage={storage}", flush=True)
    fs, op = _op(C)
    op.set_storage(storage)
    X = fs.nodal_coordinates()
    uc = 0.1 * np.sin(2 * np.pi * X[:, 0]) + 0.05 * X[:, 2]
    fs.dirichlet_commit(uc[fs.nowned:])
    u = dev(uc[:fs.nowned])
    f = torch.zeros_like(u)
    op.function(u, f)
    v = dev(2 * splitmix64_uniform(3, fs.nowned) - 1)
    y = torch.zeros_like(u)
    op.mat_mult(v, y)
    y32 = torch.zeros(fs.nowned, dtype=torch.float32, device="cuda")
    op.mat_mult_f32(v.float(), y32)
    for bad in (np.nan, np.inf):
        vb = v.clone()
        vb[fs.nowned // 2] = bad
        with pytest.raises(C.DohpError) as ei:
            op.mat_mult(vb, y)
        assert ei.value.code == C.ERR_NAN
        with pytest.raises(C.DohpError) as ei:
            op.mat_mult_f32(vb.float(), y32)
        assert ei.value.code == C.ERR_NAN
    # a finite application afterwards is clean again (the flag is cleared per call)
    op.mat_mult(v, y)
    assert torch.isfinite(y).all()
    ub = u.clone()
    ub[7] = np.nan
    with pytest.raises(C.DohpError) as ei:
        op.function(ub, f)
    assert ei.value.code == C.ERR_NAN
    # the Jacobian of that state is NaN too
    with pytest.raises(C.DohpError) as ei:
        op.mat_mult(v, y)
    assert ei.value.code == C.ERR_NAN
    # unchecked: returns at once, the NaN is in the output
    op.set_checks(0)
    op.mat_mult(v, y)
    torch.cuda.synchronize()
    assert not torch.isfinite(y).all()
    op.set_checks(1)
    op.function(u, f)
    op.mat_mult(v, y)
    assert torch.isfinite(y).all()


def test_solver_reports_nonfinite_rhs(C):
    fs, op = _op(C, M=(4, 4, 4), p=3, nonlinear=False)
    b = dev(2 * splitmix64_uniform(5, fs.nowned) - 1)
    x = torch.zeros_like(b)
    q1 = op.assemble_q1(torch.zeros_like(b))
    b[3] = np.nan
    with pytest.raises(C.DohpError) as ei:
        op.solve(q1, b, x, rtol=1e-8)
    assert ei.value.code == C.ERR_NAN
