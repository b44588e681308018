"""GPU parity of the reverse-mode rules (tenrso_amd.ad) through the C ABI.

Pins, each independent of the kernels under test:
* FP64 central finite differences (step 1e-6) of the functional along random
  unit directions: <Mbar, mttkrp(X, A, mode)> for the MTTKRP VJP, and
  0.5 ||X - [[A]]||^2 (evaluated in FP64 by torch) for the CP-loss gradient;
* the FP64 oracle of CpReconstructionGrad (grad.rs:81-120);
* the identity dL/dA_n = A_n Gamma_n - M_n = mttkrp(Xhat - X, A, n).
Bars: directional derivatives within 1e-4 of ||g|| ||d||; FP32 gradients within
relative Frobenius 1e-5 of FP64 (1e-4 where the result is a difference of two
large terms); loss within 1e-6 relative of FP64.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import oracle  # noqa: E402
from tenrso_amd import synth  # noqa: E402
from tenrso_amd.ad import AdContext, CpLoss, CpReconstructionGrad, MttkrpVjp, cp_gd, cp_loss_and_grad  # noqa: E402
from tenrso_amd.kernels import cp_reconstruct  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def recon64(F):
    """FP64 CP reconstruction with torch matmuls (A_0 @ KR(A_1..)^T)."""
    R = F[0].shape[1]
    kr = F[-1]
    for f in reversed(F[1:-1]):
        kr = (f[:, None, :] * kr[None, :, :]).reshape(-1, R)
    return (F[0] @ kr.T).reshape([f.shape[0] for f in F])


def loss64(X64, F):
    return 0.5 * float(torch.sum((X64 - recon64(F)) ** 2))


def unit(rng, shape):
    d = rng.standard_normal(shape)
    return d / np.linalg.norm(d)


@pytest.mark.parametrize("shape,R", [((6, 7, 8), 5), ((5, 6, 4, 3), 8), ((20, 16, 24), 33)])
def test_mttkrp_vjp_fd(shape, R):
    print(f"[case] mttkrp vjp FD {shape} R={R}", flush=True)
    rng = np.random.default_rng(len(shape) * 100 + R)
    X = rng.standard_normal(shape).astype(np.float32)
    F = [rng.standard_normal((n, R)).astype(np.float32) for n in shape]
    X64, F64 = X.astype(np.float64), [f.astype(np.float64) for f in F]
    h = 1e-6
    for mode in range(len(shape)):
        Mbar = rng.standard_normal((shape[mode], R)).astype(np.float32)
        op = MttkrpVjp(cuda(X), [cuda(f) for f in F], mode)
        grads = [g.cpu().numpy().astype(np.float64) for g in op.vjp(cuda(Mbar))]
        others = [m for m in range(len(shape)) if m != mode]

        def f(Xv, Fv):
            return float(np.sum(Mbar.astype(np.float64) * oracle.mttkrp_fast(Xv, Fv, mode)))

        # direction in X
        for _ in range(4):
            d = unit(rng, shape)
            fd = (f(X64 + h * d, F64) - f(X64 - h * d, F64)) / (2 * h)
            assert abs(fd - np.sum(grads[0] * d)) <= 1e-4 * np.linalg.norm(grads[0])
        # directions in each factor that enters the MTTKRP
        for k, m in enumerate(others):
            g = grads[1 + k]
            for _ in range(3):
                d = unit(rng, F[m].shape)
                Fp = [x.copy() for x in F64]
                Fm = [x.copy() for x in F64]
                Fp[m] += h * d
                Fm[m] -= h * d
                fd = (f(X64, Fp) - f(X64, Fm)) / (2 * h)
                assert abs(fd - np.sum(g * d)) <= 1e-4 * np.linalg.norm(g)
            # closed form: dA_m = mttkrp(X, A with A_mode := Mbar, m)
            Fsub = [x.copy() for x in F64]
            Fsub[mode] = Mbar.astype(np.float64)
            assert rel(g, oracle.mttkrp_fast(X64, Fsub, m)) <= 1e-5
        Fsub = [x.copy() for x in F64]
        Fsub[mode] = Mbar.astype(np.float64)
        assert rel(grads[0], oracle.cp_reconstruct(Fsub)) <= 1e-5


@pytest.mark.parametrize("weighted", [False, True])
def test_cp_reconstruction_grad_vs_oracle(weighted):
    rng = np.random.default_rng(3 + weighted)
    shape, R = (9, 10, 11, 4), 6
    F = [rng.standard_normal((n, R)).astype(np.float32) for n in shape]
    G = rng.standard_normal(shape).astype(np.float32)
    w = rng.uniform(0.5, 2.0, R).astype(np.float32) if weighted else None
    op = CpReconstructionGrad([cuda(f) for f in F], None if w is None else cuda(w))
    got = op.compute_factor_gradients(cuda(G))
    ref = oracle.cp_factor_grads(G, [f.astype(np.float64) for f in F], w)
    for n in range(len(shape)):
        assert rel(got[n].cpu().numpy(), ref[n]) <= 1e-5
    with pytest.raises(ValueError, match="doesn't match number of factors"):
        op.compute_factor_gradients(cuda(G[0]))


def test_cp_loss_grad_small_fd_and_identity():
    rng = np.random.default_rng(11)
    shape, R = (7, 8, 9), 4
    X = rng.standard_normal(shape).astype(np.float32)
    F = [rng.standard_normal((n, R)).astype(np.float32) for n in shape]
    Fd = [cuda(f) for f in F]
    loss, grads = cp_loss_and_grad(cuda(X), Fd)
    F64 = [torch.from_numpy(f.astype(np.float64)).cuda() for f in F]
    X64 = torch.from_numpy(X.astype(np.float64)).cuda()
    assert abs(loss - loss64(X64, F64)) <= 1e-6 * loss
    Xh = cp_reconstruct(Fd)
    ident = CpReconstructionGrad(Fd).compute_factor_gradients(Xh - cuda(X))
    h = 1e-6
    for n in range(3):
        g = grads[n].cpu().numpy().astype(np.float64)
        assert rel(g, ident[n].cpu().numpy()) <= 1e-4
        ref = oracle.cp_factor_grads(oracle.cp_reconstruct([f.astype(np.float64) for f in F]) - X,
                                     [f.astype(np.float64) for f in F])
        assert rel(g, ref[n]) <= 1e-5
        for _ in range(4):
            d = torch.from_numpy(unit(rng, F[n].shape)).cuda()
            Fp = list(F64)
            Fm = list(F64)
            Fp[n] = F64[n] + h * d
            Fm[n] = F64[n] - h * d
            fd = (loss64(X64, Fp) - loss64(X64, Fm)) / (2 * h)
            assert abs(fd - float(np.sum(g * d.cpu().numpy()))) <= 1e-4 * np.linalg.norm(g)
    # AdOperation shape and the tape
    ctx = AdContext()
    op_id = ctx.register_operation(CpLoss(cuda(X), Fd))
    ctx.backward(op_id, 2.0)
    assert rel(ctx.get_gradient(op_id, 2).cpu().numpy(), 2.0 * grads[1].cpu().numpy()) <= 1e-6


def test_c4_96x4_vjp_and_loss_grad_fd():
    print("[case] 96^4 R=32 VJP / loss gradient FD, 64 directions", flush=True)
    X, Ftrue, w, init = synth.dense4_96_r32()
    shape, R = X.shape, 32
    rng = np.random.default_rng(42)
    # a point between the truth and the random init, so the residual is not tiny
    F = [(ft * w ** 0.25 + 0.3 * fi).astype(np.float32) for ft, fi in zip(Ftrue, init)]
    Xd = cuda(X)
    Fd = [cuda(f) for f in F]
    loss, grads = cp_loss_and_grad(Xd, Fd)
    X64 = Xd.double()
    F64 = [f.double() for f in Fd]
    L64 = loss64(X64, F64)
    print(f"  loss gpu {loss:.10e} fp64 {L64:.10e}", flush=True)
    assert abs(loss - L64) <= 1e-6 * L64
    # identity with the reconstruction-gradient rule, every mode
    Xh = cp_reconstruct(Fd)
    ident = CpReconstructionGrad(Fd).compute_factor_gradients(Xh - Xd)
    for n in range(4):
        e = rel(grads[n].cpu().numpy(), ident[n].cpu().numpy())
        print(f"  mode {n}: |grad - mttkrp(Xh - X)| rel {e:.3e}", flush=True)
        assert e <= 1e-4
    # FP64 central FD of the loss along 64 random unit directions (16 per mode)
    h = 1e-6
    worst = 0.0
    for n in range(4):
        g = grads[n].double()
        for _ in range(16):
            d = torch.from_numpy(unit(rng, (shape[n], R))).cuda()
            Fp = list(F64)
            Fm = list(F64)
            Fp[n] = F64[n] + h * d
            Fm[n] = F64[n] - h * d
            fd = (loss64(X64, Fp) - loss64(X64, Fm)) / (2 * h)
            err = abs(fd - float(torch.sum(g * d))) / float(torch.linalg.norm(g))
            worst = max(worst, err)
    print(f"  worst directional error / ||g||: {worst:.3e}", flush=True)
    assert worst <= 1e-4
    # MTTKRP VJP, all four modes, along 16 directions of X each (functional linear in X)
    for mode in range(4):
        Mbar = torch.randn((shape[mode], R), generator=torch.Generator().manual_seed(mode)).cuda()
        dX = MttkrpVjp(Xd, Fd, mode).vjp(Mbar)[0].double()
        for _ in range(4):
            d = torch.from_numpy(unit(rng, shape)).cuda()
            # <Mbar, mttkrp(d, A, mode)> exactly (linear): FP64 torch contraction
            Fm64 = [f for k, f in enumerate(F64) if k != mode]
            kr = Fm64[-1]
            for f in reversed(Fm64[:-1]):
                kr = (f[:, None, :] * kr[None, :, :]).reshape(-1, R)
            dm = torch.movedim(d, mode, 0).reshape(shape[mode], -1) @ kr
            exact = float(torch.sum(Mbar.double() * dm))
            assert abs(exact - float(torch.sum(dX * d))) <= 1e-4 * float(torch.linalg.norm(dX))


def test_c4_gradient_descent_200_steps():
    print("[case] 96^4 R=32 gradient descent 200 steps", flush=True)
    X, Ftrue, w, init = synth.dense4_96_r32()
    F = [(ft * w ** 0.25 + 0.3 * fi).astype(np.float32) for ft, fi in zip(Ftrue, init)]
    Xd = cuda(X)
    Fd = [cuda(f) for f in F]
    # first steps against an FP64 torch restatement of the same update
    F64 = [f.double() for f in Fd]
    X64 = Xd.double()
    hist = cp_gd(Xd, [f.clone() for f in Fd], 3, lr=0.0)
    for s in range(3):
        L = loss64(X64, F64)
        assert abs(hist[s] - L) <= 1e-5 * L
        grams = [f.T @ f for f in F64]
        gam = []
        for n in range(4):
            g = torch.ones_like(grams[0])
            for k in range(4):
                if k != n:
                    g = g * grams[k]
            gam.append(g)
        step = 1.0 / max(float(torch.linalg.norm(g)) for g in gam)
        Xh = recon64(F64)
        grads = []
        for n in range(4):
            Fm = [f for k, f in enumerate(F64) if k != n]
            kr = Fm[-1]
            for f in reversed(Fm[:-1]):
                kr = (f[:, None, :] * kr[None, :, :]).reshape(-1, 32)
            grads.append(torch.movedim(Xh - X64, n, 0).reshape(X.shape[n], -1) @ kr)
        F64 = [f - step * g for f, g in zip(F64, grads)]
    hist = cp_gd(Xd, Fd, 200, lr=0.0)
    print(f"  loss {hist[0]:.6e} -> {hist[-1]:.6e}", flush=True)
    assert np.all(np.isfinite(hist))
    assert hist[-1] < 0.5 * hist[0]
    assert np.all(np.diff(hist) <= 1e-6 * hist[0])


@pytest.mark.parametrize("shape,R", [((6, 9, 5, 11), 4), ((12, 7, 10, 8), 16), ((40, 24, 36, 20), 32),
                                     ((3, 64, 2, 70), 8)])
def test_cp_loss_grad_4way_tree_vs_oracle(shape, R):
    """Four modes take the dimension tree: two partial MTTKRPs over views of X
    ([I0 I1, I2, I3] and [I0, I1, I2 I3]), each mode's M contracted from one of
    them with a neighbouring factor.  Distinct mode sizes pin which index each
    contraction sums over.  Loss and all four gradients against the FP64 oracle,
    then five descent steps (step 1 / max_n ||Gamma_n||_F) against the same steps
    taken in FP64."""
    print(f"[case] 4-way loss / gradient / descent {shape} rank {R}", flush=True)
    rng = np.random.default_rng(sum(shape) + R)
    X = rng.standard_normal(shape).astype(np.float32)
    F = [(0.5 * rng.standard_normal((n, R))).astype(np.float32) for n in shape]
    X64 = X.astype(np.float64)
    F64 = [f.astype(np.float64) for f in F]
    loss, grads = cp_loss_and_grad(cuda(X), [cuda(f) for f in F])

    def ref_loss_grads(Fs):
        D = oracle.cp_reconstruct(Fs) - X64
        return 0.5 * float(np.sum(D * D)), oracle.cp_factor_grads(D, Fs)

    L64, G64 = ref_loss_grads(F64)
    assert abs(loss - L64) <= 1e-5 * L64
    for n in range(4):
        e = rel(grads[n].cpu().numpy(), G64[n])
        print(f"  mode {n}: gradient rel {e:.2e}", flush=True)
        assert e <= 1e-5
    Fd = [cuda(f) for f in F]
    hist = cp_gd(cuda(X), Fd, 5)
    Fs = list(F64)
    for s in range(5):
        L, G = ref_loss_grads(Fs)
        assert abs(hist[s] - L) <= 1e-4 * L, (s, hist[s], L)
        grams = [f.T @ f for f in Fs]
        gam = []
        for n in range(4):
            g = np.ones((R, R))
            for k in range(4):
                if k != n:
                    g = g * grams[k]
            gam.append(np.linalg.norm(g))
        step = 1.0 / max(gam)
        Fs = [f - step * g for f, g in zip(Fs, G)]
    for n in range(4):
        assert rel(Fd[n].cpu().numpy(), Fs[n]) <= 1e-4
