"""Parity of the batched mixed-order Poisson operator (one persistent launch over block spaces and
independent elements of degree 1..10) with the CPU oracle on the same seeded inputs.

FP64 bar: relative 2-norm difference <= 1e-12 per segment.  FP32 states its error against the FP64
oracle: bound 2e-5 relative per segment (FP32 values and contractions, FP64 geometry).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from dohp_amd.mesh import (block_vertices, element_vertex_index, element_vertices,  # noqa: E402
                           splitmix64_uniform, uniform_integers)
from oracle import oracle as O  # noqa: E402

TOL64 = 1e-12
TOL32 = 2e-5


@pytest.fixture(scope="module")
def C():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import dohp_amd.capi as capi
    return capi


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def mixed_mesh(M, seed, lo=1, hi=10):
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=seed)
    nel = M[0] * M[1] * M[2]
    return uniform_integers(seed + 1, nel, lo, hi), element_vertices(M, V)


def build(C, spaces, elems):
    """Batch and oracle references: spaces [(M, p, V, kv|None)], elems [(deg, vx, kv|None)]."""
    B = C.Batch()
    keep, segs = [], []
    for (M, p, V, kv) in spaces:
        fs = C.FS(M, p, V)
        keep.append(fs)
        segs.append(("space", B.add_space(fs, kv), (M, p, V, kv)))
    for (deg, vx, kv) in elems:
        segs.append(("elems", B.add_elements(deg, vx, kv), (deg, vx, kv)))
    B.setup()
    return B, keep, segs


def reference(kind, data, x):
    if kind == "elems":
        deg, vx, kv = data
        return O.expanded_poisson_apply(deg, vx, x, kvert=kv)
    M, p, V, kv = data
    S = O.BlockSpace(M, p, V)
    E = O.Ellip(S, p=2.0, kvert=None if kv is None else kv[element_vertex_index(M)])
    xc = np.concatenate([x, np.zeros(S.nclosure - S.nowned)])
    return E.jacobian_apply(np.zeros(S.nclosure), xc)[:S.nowned]


def case(seed):
    spaces = []
    for i, (M, p) in enumerate([((2, 3, 2), 4), ((3, 2, 2), 2), ((2, 2, 2), 7), ((1, 2, 1), 10)]):
        V = block_vertices(M, perturb=0.1, seed=seed + 10 * i)
        kv = None if i % 2 else np.exp(splitmix64_uniform(seed + 100 + i, V.shape[0]) - 0.5)
        spaces.append((M, p, V, kv))
    deg, vx = mixed_mesh((3, 3, 3), seed)
    kv = 1 + splitmix64_uniform(seed + 7, deg.size * 8).reshape(-1, 8)
    deg2, vx2 = mixed_mesh((2, 2, 1), seed + 3, lo=5, hi=10)
    return spaces, [(deg, vx, kv), (deg2, vx2, None)]


def test_batch_fp64_and_fp32_match_oracle(C):
    print("batch mixed 3^3 + 2x2x1 + four spaces", flush=True)
    spaces, elems = case(21)
    B, keep, segs = build(C, spaces, elems)
    n = B.size
    x = 2 * splitmix64_uniform(5, n) - 1
    y = torch.zeros(n, dtype=torch.float64, device="cuda")
    B.mult(dev(x), y)
    y32 = torch.zeros(n, dtype=torch.float32, device="cuda")
    B.mult_f32(dev(x, torch.float32), y32)
    yh, y32h = y.cpu().numpy(), y32.cpu().numpy()
    errs = []
    for kind, s, data in segs:
        off, ln = B.segment(s)
        ref = reference(kind, data, x[off:off + ln])
        assert rel(yh[off:off + ln], ref) < TOL64, (kind, s)
        errs.append(rel(y32h[off:off + ln], ref))
        assert errs[-1] < TOL32, (kind, s, errs[-1])
    print("fp32 relative errors", ["%.2e" % e for e in errs])
    # every degree present was tuned, in both precisions
    present = {int(d) for (deg, _, _) in elems for d in deg} | {int(sp[1]) for sp in spaces}
    for d in present:
        for fp32 in (False, True):
            tile, ms = B.tuning(d, fp32)
            assert tile in (0, 1) and ms[tile] > 0 and ms[2] < 0


def test_batch_workload_generator_matches_oracle(C):
    """The bench's mixed-order workload (mesh.mixed_order_batch) at 8^3 mixed elements (degrees 2..10) and 8
    Poisson problems (n^3 meshes, n in 4..16, degrees 2..8): every segment against the oracle."""
    from dohp_amd.mesh import mixed_order_batch
    print("batch workload generator 8^3 + 8 problems", flush=True)
    deg, vx, probs = mixed_order_batch(8, 8)
    spaces = [(M, p, V, None) for (M, p, V) in probs]
    B, keep, segs = build(C, spaces, [(deg, vx, None)])
    n = B.size
    x = 2 * splitmix64_uniform(12, n) - 1
    y = torch.zeros(n, dtype=torch.float64, device="cuda")
    B.mult(dev(x), y)
    y32 = torch.zeros(n, dtype=torch.float32, device="cuda")
    B.mult_f32(dev(x, torch.float32), y32)
    yh, y32h = y.cpu().numpy(), y32.cpu().numpy()
    worst64 = worst32 = 0.0
    for kind, s, data in segs:
        off, ln = B.segment(s)
        ref = reference(kind, data, x[off:off + ln])
        e64, e32 = rel(yh[off:off + ln], ref), rel(y32h[off:off + ln], ref)
        worst64, worst32 = max(worst64, e64), max(worst32, e32)
        assert e64 < TOL64 and e32 < TOL32, (kind, s, e64, e32)
    print(f"{len(segs)} segments, {n} unknowns: fp64 {worst64:.1e}, fp32 {worst32:.1e}")


def test_batch_repeat_and_layout(C):
    print("batch repeat", flush=True)
    spaces, elems = case(33)
    B, keep, segs = build(C, spaces, elems)
    offs = sorted(B.segment(s) for _, s, _ in segs)
    assert offs[0][0] == 0 and all(a[0] + a[1] == b[0] for a, b in zip(offs, offs[1:]))
    assert offs[-1][0] + offs[-1][1] == B.size
    # block spaces come first in the packed vector
    assert max(B.segment(s)[0] for k, s, _ in segs if k == "space") < min(B.segment(s)[0] for k, s, _ in segs if k == "elems")
    x = dev(2 * splitmix64_uniform(6, B.size) - 1)
    y1 = torch.full((B.size,), 7.0, dtype=torch.float64, device="cuda")  # output is overwritten
    y2 = torch.zeros_like(y1)
    B.mult(x, y1)
    B.mult(x, y2)
    assert rel(y1.cpu().numpy(), y2.cpu().numpy()) < 1e-14


def test_batch_space_equals_operator(C):
    print("batch space vs operator", flush=True)
    M, p = (3, 2, 3), 6
    V = block_vertices(M, perturb=0.1, seed=4)
    fs = C.FS(M, p, V)
    op = C.Ellip(fs, p=2.0)
    B = C.Batch()
    s = B.add_space(fs)
    x = dev(2 * splitmix64_uniform(9, fs.nowned) - 1)
    y = torch.zeros_like(x)
    op.mat_mult(x, y)
    yb = torch.zeros(B.size, dtype=torch.float64, device="cuda")
    B.mult(x, yb)
    assert B.segment(s) == (0, fs.nowned)
    assert rel(yb.cpu().numpy(), y.cpu().numpy()) < TOL64


def test_batch_rejects_bad_input(C):
    B = C.Batch()
    vx = element_vertices((1, 1, 1), block_vertices((1, 1, 1)))
    with pytest.raises(C.DohpError):
        B.add_elements([11], vx)
    mirrored = vx.copy()
    mirrored[..., 0] *= -1
    with pytest.raises(C.DohpError) as ei:
        B.add_elements([3], mirrored)
    assert ei.value.code == C.ERR_NEGATIVE_JACOBIAN


@pytest.mark.parametrize("passes,tol", [(3, 5e-4), (1, 5e-2)])
def test_batch_mfma_tiles(C, passes, tol):
    """FP32 with the MFMA contraction tile (BF16 operands, FP32 accumulation) fixed for every degree >= 4,
    against the FP64 oracle per segment; then the tuner with the MFMA tile offered measures all three."""
    print(f"batch mfma passes={passes}", flush=True)
    spaces, elems = case(45)
    B, keep, segs = build(C, spaces, elems)
    B.set_mfma(passes)
    for d in range(4, 11):
        B.set_tile(d, True, 2)
    n = B.size
    x = 2 * splitmix64_uniform(8, n) - 1
    y32 = torch.zeros(n, dtype=torch.float32, device="cuda")
    B.mult_f32(dev(x, torch.float32), y32)
    y32h = y32.cpu().numpy()
    errs = []
    for kind, s, data in segs:
        off, ln = B.segment(s)
        ref = reference(kind, data, x[off:off + ln])
        errs.append(rel(y32h[off:off + ln], ref))
        assert errs[-1] < tol, (kind, s, errs[-1])
    print("mfma tile errors", ["%.2e" % e for e in errs])
    for d in range(4, 11):
        B.set_tile(d, True, -1)
    B.mult_f32(dev(x, torch.float32), y32)
    for d in range(4, 11):
        tile, ms = B.tuning(d, True)
        if tile >= 0:  # workgroup and MFMA tiles measured (the wavefront tile only for degree <= 7)
            assert ms[0] > 0 and ms[2] > 0 and (ms[1] > 0) == (d <= 7), (d, ms)
    print("tiles", {d: B.tuning(d, True) for d in range(1, 11)})
