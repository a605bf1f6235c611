"""C-ABI library checks that need no GPU: it loads, exports every declared symbol, and its
host-side setup tables (rules, bases, numbering) agree with the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    return set(re.findall(r"^\s*(?:dErr|void|int)\s+(\w+)\s*\(", txt, flags=re.M))


def test_library_exports_every_declared_symbol():
    C = pytest.importorskip("dohp_amd.capi")
    syms = _declared_symbols("dohp_amd.h")
    assert len(syms) >= 15
    lib = C.lib()
    for s in sorted(syms):
        assert hasattr(lib, s), s


def test_dhm_library_exports_every_declared_symbol():
    import ctypes
    path = os.path.join(ROOT, "dohp_amd", "libdohp_dhm.so")
    assert os.path.exists(path), "DHM library not built"
    lib = ctypes.CDLL(path)
    syms = _declared_symbols("dohp_dhm.h")
    assert len(syms) >= 10
    for s in sorted(syms):
        assert hasattr(lib, s), s


def test_every_header_is_covered():
    assert sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")) == ["dohp_amd.h", "dohp_dhm.h"]


@pytest.mark.parametrize("method,family,order", [(0, 0, k) for k in range(0, 21)] + [(0, 1, k) for k in range(0, 13)]
                         + [(1, 0, k) for k in (2, 4, 6, 8, 12)] + [(2, 1, k) for k in (2, 4, 8, 12)])
def test_tensor_rule_matches_oracle(method, family, order):
    C = pytest.importorskip("dohp_amd.capi")
    z, w = C.tensor_rule(method, family, order)
    zo, wo = O.rule_1d(method, family, order)
    np.testing.assert_allclose(z, zo, atol=1e-15)
    np.testing.assert_allclose(w, wo, atol=1e-15)


@pytest.mark.parametrize("P", range(2, 13))
def test_tensor_basis_matches_oracle(P):
    C = pytest.importorskip("dohp_amd.capi")
    for Q in (P, P + 1, 2 * (P - 1)):
        rc, _ = O.rule_1d(O.QUAD_FAST if Q != 2 * (P - 1) else O.QUAD_SPARSE, O.GAUSS_LEGENDRE,
                          2 * (Q - 1) if Q != 2 * (P - 1) else Q)
        B, D, node, weight = C.tensor_basis(P, rc)
        Bo, Do, nodeo, weighto = O.tensor_basis(P, rc)
        np.testing.assert_allclose(B, Bo, atol=1e-14)
        np.testing.assert_allclose(D, Do, atol=5e-13 * P * P)
        np.testing.assert_allclose(node, nodeo, atol=1e-15)
        np.testing.assert_allclose(weight, weighto, atol=1e-15)


@pytest.mark.gpu
def test_fs_numbering_matches_oracle_space():
    """The library's element->closure map (E) equals the oracle's, on a perturbed block."""
    C = pytest.importorskip("dohp_amd.capi")
    from dohp_amd.mesh import block_vertices
    for M, p, mask in [((3, 2, 4), 3, 7), ((2, 2, 2), 1, 1), ((4, 3, 2), 4, 6)]:
        V = block_vertices(M, perturb=0.1, seed=3)
        fs = C.FS(M, p, V, dirichlet_mask=mask)
        S = O.BlockSpace(M, p, V, dirichlet_mask=mask)
        assert (fs.nowned, fs.nclosure, fs.nelem) == (S.nowned, S.nclosure, S.nelem)
        np.testing.assert_array_equal(fs.element_nodes(), S.elem_nodes)
        np.testing.assert_allclose(fs.nodal_coordinates(), S.nodal_coordinates(), atol=1e-14)


def test_error_reporting_without_gpu():
    C = pytest.importorskip("dohp_amd.capi")
    with pytest.raises(C.DohpError) as ei:
        C.tensor_basis(1, np.zeros(3))
    assert "basis size" in str(ei.value)
    with pytest.raises(C.DohpError):
        C.tensor_rule(7, 0, 4)


def test_failed_device_allocation_is_an_error_not_a_crash():
    """Without a device every allocation fails: dFSCreateBlock returns dohpERR_OOM naming the size it asked for, and a
    second call asks again (a buffer whose allocation failed keeps no size, so a retry never skips the allocation and
    writes through a null pointer)."""
    C = pytest.importorskip("dohp_amd.capi")
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("exercises the allocation failure of a host without a GPU")
    from dohp_amd.mesh import block_vertices
    V = block_vertices((2, 2, 2))
    msgs = []
    for _ in range(2):
        with pytest.raises(C.DohpError) as ei:
            C.FS((2, 2, 2), 3, V)
        assert ei.value.code == C.ERR_OOM
        msgs.append(str(ei.value))
    assert "hipMalloc failed for" in msgs[0] and msgs[0] == msgs[1]
