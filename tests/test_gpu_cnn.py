"""GPU parity of the learned warm-start network (imp_cnn_*) against the
FP64 oracle with bf16 rounding at the points where the GPU stores bf16."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _rand(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_cnn_forward_matches_oracle(ctx):
    from impatient import ops
    from oracle import cnn
    H, W = 8, 64
    net = ops.CNN(ctx, H, W, seed=5)
    p = cnn.random_params(11)
    net.set_params(p)
    z = _rand((H, W), 12).astype(np.complex64)
    got = net.forward(z)
    ref, _ = cnn.forward(p.astype(np.float32).astype(np.float64), z)
    err = rel_l2(got, ref)
    # the network's own contribution, to see the error relative to it
    rel_net = rel_l2(got - z, ref - z)
    print(f"CNN forward: rel_l2 {err:.2e}, on the network term {rel_net:.2e}")
    assert err < 2e-4 and rel_net < 2e-3


def test_cnn_gradient_matches_oracle(ctx):
    from impatient import ops
    from oracle import cnn
    H, W = 8, 64
    net = ops.CNN(ctx, H, W, seed=5)
    p = cnn.random_params(13).astype(np.float32)
    net.set_params(p)
    z = _rand((H, W), 14).astype(np.complex64)
    t = (z + 0.5 * _rand((H, W), 15)).astype(np.complex64)
    loss, g = net.gradient(z, t)
    lref, gref = cnn.loss_and_grad(p.astype(np.float64), z, t)
    print(f"CNN loss {loss:.6e} vs oracle {lref:.6e}")
    assert abs(loss - lref) <= 1e-4 * abs(lref)
    gl = g.reshape(cnn.NLAYER, cnn.PSZ)
    rl = gref.reshape(cnn.NLAYER, cnn.PSZ)
    for layer in range(cnn.NLAYER):
        ew = rel_l2(gl[layer, :cnn.WSZ], rl[layer, :cnn.WSZ])
        eb = rel_l2(gl[layer, cnn.WSZ:], rl[layer, cnn.WSZ:])
        print(f"layer {layer}: dW rel_l2 {ew:.2e}, db rel_l2 {eb:.2e}")
        assert ew < 5e-3 and eb < 5e-3
    assert np.all(g[~cnn.trainable_mask()] == 0)


def test_cnn_training_reduces_loss(ctx):
    """40 Adam steps on one pair (target = z smoothed by a fixed 3x3 blur)."""
    from impatient import ops
    H, W = 32, 64
    net = ops.CNN(ctx, H, W, seed=7)
    z = _rand((H, W), 16).astype(np.complex64)
    zp = np.pad(z, 1, mode="constant")
    t = sum(zp[1 + dy:1 + dy + H, 1 + dx:1 + dx + W] for dy in (-1, 0, 1)
            for dx in (-1, 0, 1)) / 9.0
    t = t.astype(np.complex64)
    dz = [ctx.to_device(np.ascontiguousarray(v, np.float32)) for v in (z.real, z.imag)]
    dt = [ctx.to_device(np.ascontiguousarray(v, np.float32)) for v in (t.real, t.imag)]
    losses = [net.train_step_device(dz[0], dz[1], dt[0], dt[1], 3e-3) for _ in range(40)]
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.5 * losses[0]


def test_cnn_params_roundtrip_and_errors(ctx):
    from impatient import ops, ImpError
    from oracle import cnn
    net = ops.CNN(ctx, 4, 64, seed=3)
    p0 = net.get_params()
    assert p0.shape == (cnn.NPARAM,)
    assert np.all(p0[~cnn.trainable_mask()] == 0)
    # the untrained network is the identity (output layer starts at zero)
    z = _rand((4, 64), 17).astype(np.complex64)
    assert rel_l2(net.forward(z), z) < 1e-6
    p = cnn.random_params(4).astype(np.float32)
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)
    bad = p.copy()
    bad[5] = np.inf
    with pytest.raises(ImpError):
        net.set_params(bad)
    with pytest.raises(ImpError):
        ops.CNN(ctx, 4, 48)


def test_cnn_outliving_its_context_is_detached():
    """A network whose context is destroyed first loses its device memory
    with the context; calls on it then fail with IMP_EINVAL and destroying
    it afterwards frees only the host object (no use of the freed context)."""
    from impatient import ops, Context, ImpError
    c2 = Context(0)
    net = ops.CNN(c2, 8, 64, seed=2)
    c2.close()
    with pytest.raises(ImpError):
        net.get_params()
    net.close()
