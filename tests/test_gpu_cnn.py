"""The learned start's network on the GPU (bh_cnn_*) against the CPU oracle.

* forward pass: the predicted coefficients against oracle/cnn.py's FP64
  forward on the same weights rounded as the GPU uses them (matrices BF16,
  biases FP32), 1e-4 of the largest output;
* training step: the batch loss (1e-6 absolute) and the gradient of every
  parameter tensor (1e-3 of its norm) against the oracle's backward pass; the
  GPU accumulates in FP32 where the oracle uses FP64;
* Adam: two steps against the oracle's update applied to the GPU's gradients
  (1e-6 absolute);
* training fits a small fixed set (the loss falls by more than half);
* training steps replayed from HIP graphs equal the same steps launched
  kernel by kernel, bit for bit (losses, gradients, weights), across batch
  sizes, learning rates and evaluation-only steps.
"""
import numpy as np
import pytest

from blaze2d_amd import HipBackend
from blaze2d_amd.learned import CnnNet
from oracle import cnn as C

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def be():
    b = HipBackend(0)
    yield b
    b.close()


def random_case(seed, B, H, W, nv, h):
    r = np.random.default_rng(seed)
    grid = r.uniform(0.05, 1.0, (B, 2, H, W)).astype(np.float32)
    embed = np.column_stack([r.uniform(0, 0.5, B), r.uniform(0, 0.5, B), r.integers(0, 2, B),
                             r.integers(0, 2, B)]).astype(np.float32)
    ng = (2 * h + 1) ** 2
    targets = r.standard_normal((B, nv, ng)) + 1j * r.standard_normal((B, nv, ng))
    return grid, embed, targets


def random_params(net, seed):
    # the network's own initialisation plus non-zero biases
    r = np.random.default_rng(seed)
    lay = C.layout(net.height, net.width, net.n_vectors, net.pw_half)
    p = C.unpack(net.params().astype(np.float64), lay)
    for name, shape in lay:
        if name.startswith("b"):
            p[name] = 0.1 * r.standard_normal(shape)
    flat = C.pack(p, lay).astype(np.float32)
    net.set_params(flat)
    return lay, C.unpack(flat.astype(np.float64), lay)


@pytest.mark.parametrize("H,W,nv,h,B", [(16, 16, 4, 2, 5), (64, 64, 12, 4, 3), (24, 40, 2, 1, 2)])
def test_forward_and_gradients_match_oracle(be, H, W, nv, h, B):
    net = CnnNet(be, H, W, nv, h, max_batch=8, seed=H + W)
    try:
        lay, p = random_params(net, 11)
        grid, embed, targets = random_case(5, B, H, W, nv, h)
        got = net.predict(grid, embed)
        pw = C.gpu_weights(p)
        want, _ = C.forward(pw, grid, embed, nv, h)
        err = np.abs(got - want).max() / np.abs(want).max()
        loss = net.train_step(grid, embed, targets, 0.0)
        lref, gref, _ = C.loss_and_grads(pw, grid, embed, targets, nv, h)
        g = C.unpack(net.grads().astype(np.float64), lay)
        worst = max(np.linalg.norm(g[n] - gref[n]) / np.linalg.norm(gref[n]) for n, _ in lay)
        print(f"{H}x{W} nv={nv} h={h} B={B}: forward rel err {err:.2e}, loss {loss:.6f} vs {lref:.6f}, "
              f"worst gradient rel err {worst:.2e}")
        assert err <= 1e-4
        assert abs(loss - lref) <= 1e-6
        for n, _ in lay:
            assert np.linalg.norm(g[n] - gref[n]) <= 1e-3 * np.linalg.norm(gref[n]), n
    finally:
        net.close()


def test_adam_steps_match_oracle(be):
    net = CnnNet(be, 16, 16, 3, 1, max_batch=4, seed=3)
    try:
        lay, p = random_params(net, 12)
        grid, embed, targets = random_case(6, 4, 16, 16, 3, 1)
        flat = net.params().astype(np.float64)
        m = np.zeros_like(flat)
        v = np.zeros_like(flat)
        for step, lr in ((1, 1e-3), (2, 5e-4)):
            net.train_step(grid, embed, targets, lr)
            g = net.grads().astype(np.float64)
            want, m, v = C.adam_step(flat, g, m, v, step, lr)
            got = net.params().astype(np.float64)
            assert np.abs(got - want).max() <= 1e-6, step
            flat = got
    finally:
        net.close()


def test_training_fits_a_small_set(be):
    net = CnnNet(be, 16, 16, 4, 2, max_batch=16, seed=4)
    try:
        grid, embed, targets = random_case(7, 16, 16, 16, 4, 2)
        first = net.train_step(grid, embed, targets, 0.0)
        for _ in range(150):
            last = net.train_step(grid, embed, targets, 2e-3)
        last = net.train_step(grid, embed, targets, 0.0)
        print(f"fixed set of 16: loss {first:.4f} -> {last:.4f}")
        assert last < 0.5 * first
    finally:
        net.close()


def test_bad_arguments_raise(be):
    with pytest.raises(RuntimeError):
        CnnNet(be, 12, 16, 4, 2)  # not a multiple of 8
    with pytest.raises(RuntimeError):
        CnnNet(be, 16, 16, 40, 2)  # more vectors than plane waves
    net = CnnNet(be, 16, 16, 2, 1, max_batch=2)
    try:
        grid, embed, targets = random_case(8, 3, 16, 16, 2, 1)
        with pytest.raises(RuntimeError):
            net.train_step(grid, embed, targets, 1e-3)  # batch 3 > max_batch 2
    finally:
        net.close()


def test_graph_replays_equal_eager_steps(be):
    # two networks from the same weights: one replays each step from a HIP
    # graph (captured on the first step of each batch size and update kind),
    # the other launches the same kernels one by one; every replay is followed
    # by the step's own synchronisation (the loss readback) and compared
    nets = [CnnNet(be, 16, 16, 4, 2, max_batch=8, seed=3) for _ in range(2)]
    try:
        nets[0].use_graphs(True)
        nets[1].use_graphs(False)
        p0 = nets[0].params()
        nets[1].set_params(p0)
        nets[0].set_params(p0)
        plan = [(5, 1e-3), (5, 8e-4), (3, 1e-3), (5, 0.0), (5, 5e-4), (3, 2e-4), (8, 1e-3), (5, 1e-4)]
        for step, (B, lr) in enumerate(plan):
            grid, embed, targets = random_case(100 + step, B, 16, 16, 4, 2)
            la = nets[0].train_step(grid, embed, targets, lr)
            lb = nets[1].train_step(grid, embed, targets, lr)
            assert la == lb, (step, la, lb)
            assert np.array_equal(nets[0].grads(), nets[1].grads()), step
            assert np.array_equal(nets[0].params(), nets[1].params()), step
    finally:
        for n in nets:
            n.close()
