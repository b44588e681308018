"""CPU tests of the host-side AD mirror: the finite-difference checker
(tenrso-ad/src/gradcheck.rs:43-268, including its own tests at ~375-437) and the
operation tape.  No GPU calls."""
import numpy as np
import pytest

from tenrso_amd.ad import AdContext, GradCheckConfig, check_gradient


def test_gradcheck_identity():
    # gradcheck.rs test_gradcheck_identity: f(x) = x, df = grad_y
    cfg = GradCheckConfig(epsilon=1e-5, rtol=1e-3, atol=1e-5, use_central_diff=True)
    r = check_gradient(lambda x: x.copy(), lambda x, gy: gy.copy(), np.ones((2, 3)), np.ones((2, 3)), cfg)
    assert r.passed and r.max_abs_diff < 1e-6


def test_gradcheck_square():
    # gradcheck.rs test_gradcheck_square: f(x) = x^2, df = 2 x grad_y
    x = np.array([1.0, 2.0, 3.0, 4.0]).reshape(2, 2)
    r = check_gradient(lambda x: x * x, lambda x, gy: 2.0 * x * gy, x, np.ones((2, 2)))
    assert r.passed and r.max_rel_diff < 1e-4


def test_gradcheck_config_default():
    c = GradCheckConfig()
    assert (c.epsilon, c.rtol, c.atol, c.use_central_diff, c.verbose) == (1e-5, 1e-3, 1e-5, True, False)


def test_gradcheck_detects_wrong_gradient_and_shape():
    x = np.array([1.0, -2.0, 0.5])
    r = check_gradient(lambda x: x ** 3, lambda x, gy: 3.0 * x * gy, x, np.ones(3))   # should be 3x^2
    assert not r.passed and r.num_failures == 2   # 3x == 3x^2 at x = 1
    fwd = check_gradient(lambda x: x ** 3, lambda x, gy: 3.0 * x * x * gy, x, np.ones(3),
                         GradCheckConfig(epsilon=1e-7, use_central_diff=False))
    assert fwd.passed
    with pytest.raises(ValueError, match="doesn't match input shape"):
        check_gradient(lambda x: x, lambda x, gy: np.ones(2), x, np.ones(3))


def test_gradcheck_on_mttkrp_vjp_formula():
    # f(X) = mttkrp(X, A, 0) on a 3x4x5 tensor (brute-force einsum), VJP dX = Mbar o A1 o A2
    rng = np.random.default_rng(0)
    A = [rng.standard_normal((n, 2)) for n in (3, 4, 5)]
    X = rng.standard_normal((3, 4, 5))
    Mbar = rng.standard_normal((3, 2))
    r = check_gradient(lambda x: np.einsum("ijk,jr,kr->ir", x, A[1], A[2]),
                       lambda x, gy: np.einsum("ir,jr,kr->ijk", gy, A[1], A[2]), X, Mbar)
    assert r.passed


class _Square:
    def __init__(self, x):
        self.x = x

    def forward(self):
        return [self.x * self.x]

    def backward(self, grads):
        return [2.0 * self.x * grads[0]]

    def name(self):
        return "square"

    def num_inputs(self):
        return 1

    def num_outputs(self):
        return 1


def test_ad_context_tape():
    ctx = AdContext()
    op = ctx.register_operation(_Square(np.array([1.0, 3.0])))
    ctx.backward(op, np.array([1.0, 0.5]))
    assert np.allclose(ctx.get_gradient(op, 0), [2.0, 3.0])
    with pytest.raises(KeyError):
        ctx.backward(op + 1, np.ones(2))
    ctx.clear()
    assert ctx.get_gradient(op, 0) is None
