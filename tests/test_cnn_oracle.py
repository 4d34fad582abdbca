"""Pins of the learned start's CPU oracle (oracle/cnn.py), on the CPU.

* the periodic 3 x 3 convolution against a circular cross-correlation by FFT;
* the stride-2 convolution against the stride-1 one sampled at even positions;
* the projector loss: closed form for one vector (1 - |u^H y|^2 / |u|^2 |y|^2),
  0 for any basis of the target span, 1 for a block orthogonal to it,
  invariance under Y -> Y M for invertible M, and its gradient against
  central finite differences;
* the whole backward pass against central finite differences of the forward
  pass along random parameter directions (FP64, no BF16 rounding);
* BF16 rounding (ties to even) on hand-picked values.
"""
import numpy as np
import pytest

from oracle import cnn as C


def rng(seed):
    return np.random.default_rng(seed)


def cplx(r, shape):
    return r.standard_normal(shape) + 1j * r.standard_normal(shape)


def test_conv_matches_fft_cross_correlation():
    r = rng(1)
    x = r.standard_normal((2, 3, 8, 12))
    w = r.standard_normal((5, 3, 3, 3))
    got = C.conv3x3(x, w, 1)
    # circular cross-correlation (x * k)[y, x] = sum_{u,v} k[u, v] x[y+u, x+v]
    # = IFFT(FFT(x) conj(FFT(k))) for real k; with k[(dy-1) mod H, (dx-1) mod W]
    # = w[o, c, dy, dx] it is the convolution's sum over (dy, dx) for channel c
    H, W = x.shape[2:]
    want = np.zeros_like(got)
    for o in range(w.shape[0]):
        for c in range(x.shape[1]):
            k = np.zeros((H, W))
            for dy in range(3):
                for dx in range(3):
                    k[(dy - 1) % H, (dx - 1) % W] = w[o, c, dy, dx]
            want[:, o] += np.real(np.fft.ifft2(np.fft.fft2(x[:, c]) * np.conj(np.fft.fft2(k))))
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_stride_two_is_even_sampling():
    r = rng(2)
    x = r.standard_normal((2, 4, 16, 8))
    w = r.standard_normal((6, 4, 3, 3))
    assert np.allclose(C.conv3x3(x, w, 2), C.conv3x3(x, w, 1)[:, :, ::2, ::2], rtol=0, atol=1e-12)


def test_loss_closed_form_one_vector():
    r = rng(3)
    y, u = cplx(r, (1, 9)), cplx(r, (1, 9))
    loss, _ = C.projector_loss(y, u)
    want = 1 - abs(np.vdot(u[0], y[0])) ** 2 / (np.vdot(u[0], u[0]).real * np.vdot(y[0], y[0]).real)
    assert abs(loss - want) <= 1e-9


def test_loss_zero_on_target_span_one_on_orthogonal_block():
    r = rng(4)
    nb, ng = 4, 25
    U = cplx(r, (nb, ng))
    M = cplx(r, (nb, nb))
    loss, g = C.projector_loss(M.T @ U, U)  # rows of M^T U span the rows of U
    assert abs(loss) <= 1e-8 and np.abs(g).max() <= 1e-6 * np.abs(U).max()
    # rows supported on disjoint plane waves are orthogonal
    Ua = np.zeros((nb, ng), complex)
    Yb = np.zeros((nb, ng), complex)
    Ua[:, :10] = cplx(r, (nb, 10))
    Yb[:, 10:] = cplx(r, (nb, 15))
    assert abs(C.projector_loss(Yb, Ua)[0] - 1.0) <= 1e-12


def test_loss_invariant_under_change_of_basis():
    r = rng(5)
    nb, ng = 5, 49
    Y, U, M = cplx(r, (nb, ng)), cplx(r, (nb, ng)), cplx(r, (nb, nb))
    l1, _ = C.projector_loss(Y, U)
    l2, _ = C.projector_loss(M.T @ Y, U)
    l3, _ = C.projector_loss(Y, M.T @ U)
    assert 0 < l1 < 1
    assert abs(l1 - l2) <= 1e-8 and abs(l1 - l3) <= 1e-8


def test_loss_gradient_finite_differences():
    r = rng(6)
    nb, ng = 3, 9
    Y, U = cplx(r, (nb, ng)), cplx(r, (nb, ng))
    _, g = C.projector_loss(Y, U)
    h = 1e-6
    for _ in range(6):
        D = cplx(r, (nb, ng))
        lp, _ = C.projector_loss(Y + h * D, U)
        lm, _ = C.projector_loss(Y - h * D, U)
        fd = (lp - lm) / (2 * h)
        an = np.sum(g.real * D.real + g.imag * D.imag)
        assert abs(fd - an) <= 1e-7 * max(1.0, abs(an)), (fd, an)


@pytest.mark.parametrize("seed", [7, 8])
def test_backward_matches_finite_differences(seed):
    r = rng(seed)
    H, W, nv, h, B = 16, 8, 3, 1, 3
    lay = C.layout(H, W, nv, h)
    p = {}
    for name, shape in lay:
        fan = int(np.prod(shape[1:])) if len(shape) > 1 else 1
        p[name] = r.standard_normal(shape) * (np.sqrt(2.0 / fan) if len(shape) > 1 else 0.1)
    grid = r.uniform(0.05, 1.0, (B, 2, H, W))
    embed = np.column_stack([r.uniform(0, 0.5, B), r.uniform(0, 0.5, B), r.integers(0, 2, B), r.integers(0, 2, B)])
    ng = (2 * h + 1) ** 2
    targets = cplx(r, (B, nv, ng))
    loss, g, _ = C.loss_and_grads(p, grid, embed, targets, nv, h)
    assert 0 < loss < 1
    flat = C.pack(p, lay)
    gflat = C.pack(g, lay)

    def masks(f):
        _, c = C.forward(C.unpack(f, lay), grid, embed, nv, h)
        return np.concatenate([c["eh"].ravel() > 0] + [a.ravel() > 0 for a in c["acts"]])

    def fd_check(d, label):
        # central differences where no ReLU changes state between the two
        # points (a crossed kink makes the difference quotient meaningless)
        for step in (1e-6, 1e-7, 1e-8):
            if np.array_equal(masks(flat + step * d), masks(flat - step * d)):
                lp = C.loss_and_grads(C.unpack(flat + step * d, lay), grid, embed, targets, nv, h)[0]
                lm = C.loss_and_grads(C.unpack(flat - step * d, lay), grid, embed, targets, nv, h)[0]
                fd = (lp - lm) / (2 * step)
                an = float(gflat @ d)
                assert abs(fd - an) <= 2e-6 * max(1e-3, abs(an)), (label, step, fd, an)
                return True
        return False

    checked = sum(fd_check(r.standard_normal(flat.size), "all") for _ in range(6))
    assert checked >= 4
    # each tensor separately (a dropped term in one layer must show)
    o = 0
    for name, shape in lay:
        n = int(np.prod(shape))
        ok = False
        for _ in range(4):
            d = np.zeros(flat.size)
            d[o:o + n] = r.standard_normal(n)
            if fd_check(d, name):
                ok = True
                break
        assert ok, name
        o += n


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7, -2.5, 3.0e-3], dtype=np.float32)
    got = C.bf16(vals)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1;
    # 1 + 3 2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: ties to even -> 1 + 2^-6
    want = [1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, -2.5]
    assert list(got[:5]) == want
    assert abs(got[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_adam_first_step_is_sign_step():
    r = rng(9)
    g = r.standard_normal(50)
    p = r.standard_normal(50)
    new, m, v = C.adam_step(p, g, np.zeros(50), np.zeros(50), 1, 1e-3)
    # m_hat = g, v_hat = g^2: the first step is lr g / (|g| + eps)
    assert np.allclose(new, p - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)
