"""The Newton-Krylov layer (dohp_amd/snes.py) on small CPU systems, pinned by what the mathematics fixes:
GMRES's k-th residual norm is the least-squares minimum over the k-th Krylov space and PCG's k-th iterate the
energy minimum over it (both computed here independently from an explicit Krylov basis), restarted GMRES
reaches the solution, Newton converges
quadratically near a root, and the backtracking line search takes a step that satisfies its sufficient-
decrease condition where the full Newton step would increase ||F||."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from dohp_amd import snes  # noqa: E402


def matvec(Am):
    T = torch.as_tensor(Am)
    return lambda v, out: out.copy_(T @ v)


def test_gmres_history_is_the_krylov_minimum():
    rng = np.random.default_rng(3)
    n = 40
    A = np.eye(n) * 4 + rng.normal(size=(n, n)) * 0.5  # nonsymmetric
    Minv = np.diag(1 / np.diag(A)) + 0.01 * rng.normal(size=(n, n))
    b = rng.normal(size=n)
    x, hist, reason = snes.gmres(matvec(A), matvec(Minv), torch.as_tensor(b), rtol=1e-10, restart=50)
    assert reason == 1
    B, r0 = Minv @ A, Minv @ b
    assert hist[0] == pytest.approx(np.linalg.norm(r0), rel=1e-14)
    K = [r0]
    for k in range(1, 12):
        Q, _ = np.linalg.qr(np.array(K).T)
        y, *_ = np.linalg.lstsq(B @ Q, r0, rcond=None)
        assert hist[k] == pytest.approx(np.linalg.norm(r0 - B @ Q @ y), rel=1e-8, abs=1e-12), k
        K.append(B @ K[-1])
    assert np.linalg.norm(Minv @ (b - A @ x.numpy())) <= 1e-10 * hist[0]


def test_restarted_gmres_reaches_the_solution():
    rng = np.random.default_rng(5)
    n = 60
    A = np.eye(n) * 3 + rng.normal(size=(n, n)) * 0.3
    b = rng.normal(size=n)
    x, hist, reason = snes.gmres(matvec(A), None, torch.as_tensor(b), rtol=1e-12, restart=7)
    assert reason == 1 and len(hist) > 8
    np.testing.assert_allclose(x.numpy(), np.linalg.solve(A, b), rtol=1e-9, atol=1e-10)


class Cubic:
    """F(x) = A x + c x^3 - b with the Jacobian action of the last residual's state, like the library's operator."""

    def __init__(self, A, c, b):
        self.A, self.c, self.b = torch.as_tensor(A), c, torch.as_tensor(b)
        self.u = None
        self.nfun = 0

    def function(self, x, f):
        self.u = x.clone()
        self.nfun += 1
        f.copy_(self.A @ x + self.c * x ** 3 - self.b)

    def mat_mult(self, v, y):
        y.copy_(self.A @ v + 3 * self.c * self.u ** 2 * v)


def test_newton_converges_quadratically():
    rng = np.random.default_rng(7)
    n = 20
    A = np.eye(n) * 2 + 0.1 * rng.normal(size=(n, n))
    A = A @ A.T
    op = Cubic(A, 0.5, rng.normal(size=n))
    x = torch.zeros(n, dtype=torch.float64)
    x, hist, ksp, reason = snes.newton(op, x, ksp_rtol=1e-12)
    assert reason == 3 and hist[-1] <= 1e-10 * hist[0]
    r = [hist[k + 1] / hist[k] ** 2 for k in range(len(hist) - 1) if hist[k + 1] > 1e-13]
    assert r[-1] < 10 * max(1.0, r[0])  # quadratic: e_{k+1} ~ C e_k^2
    f = torch.empty_like(x)
    op.function(x, f)
    assert float(torch.linalg.vector_norm(f)) <= 1e-10 * hist[0]


class Atan:
    """F(x) = atan(x) componentwise: Newton's full step from |x| > 1.39 overshoots and ||F|| grows."""

    def function(self, x, f):
        self.u = x.clone()
        f.copy_(torch.atan(x))

    def mat_mult(self, v, y):
        y.copy_(v / (1 + self.u ** 2))


def test_backtracking_line_search_decreases_where_the_full_step_fails():
    op = Atan()
    x0 = torch.full((3,), 2.0, dtype=torch.float64)
    f = torch.empty_like(x0)
    op.function(x0, f)
    fn = float(torch.linalg.vector_norm(f))
    Y = f * (1 + x0 ** 2)  # the Newton step J^-1 F
    full = torch.atan(x0 - Y)
    assert float(torch.linalg.vector_norm(full)) > fn  # the basic line search would diverge here
    xn, fnew, gn, lam = snes._line_search_bt(op, x0.clone(), f, fn, Y, torch.empty_like(x0))
    slope = -float(torch.dot(f, f))  # (F, J Y) = ||F||^2
    assert 0 < lam < 1
    assert 0.5 * gn ** 2 <= 0.5 * fn ** 2 + 1e-4 * lam * slope
    assert torch.allclose(xn, x0 - lam * Y)
    x, hist, _, reason = snes.newton(Atan(), x0.clone(), linesearch="bt", ksp_rtol=1e-12)
    assert reason > 0 and hist[-1] < 1e-10 * hist[0]
    x, hist, _, reason = snes.newton(Atan(), x0.clone(), linesearch="basic", ksp_rtol=1e-12, max_it=8)
    assert hist[-1] > hist[0]  # the full step diverges from this start


def test_cg_matches_the_krylov_energy_minimum():
    """PCG's iterate k minimizes the A-norm of the error over x0 + K_k(M^-1 A, M^-1 b); the iterates agree with
    that minimum computed from an explicit Krylov basis, and the monitored norm is ||M^-1 r||."""
    rng = np.random.default_rng(11)
    n = 30
    R = rng.normal(size=(n, n))
    A = R @ R.T + n * np.eye(n)
    Minv = np.diag(1 / np.diag(A))
    b = rng.normal(size=n)
    xstar = np.linalg.solve(A, b)
    for k in (1, 3, 6):
        x, hist, reason = snes.cg(matvec(A), matvec(Minv), torch.as_tensor(b), rtol=0.0, max_it=k)
        K = [Minv @ b]
        for _ in range(k - 1):
            K.append(Minv @ A @ K[-1])
        Q, _ = np.linalg.qr(np.array(K).T)
        y = np.linalg.solve(Q.T @ A @ Q, Q.T @ b)  # Galerkin = energy minimum
        np.testing.assert_allclose(x.numpy(), Q @ y, rtol=1e-8, atol=1e-10)
        assert hist[-1] == pytest.approx(np.linalg.norm(Minv @ (b - A @ x.numpy())), rel=1e-8)
    x, hist, reason = snes.cg(matvec(A), matvec(Minv), torch.as_tensor(b), rtol=1e-12)
    assert reason == 1
    np.testing.assert_allclose(x.numpy(), xstar, rtol=1e-9)
