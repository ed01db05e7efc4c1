"""The surrogate network's CPU oracle and host-side packing, pinned without a GPU.

* the oracle's analytic gradient against central finite differences (FP64);
* the oracle's BF16 rounding against known bit patterns and ties-to-even;
* the product's inference blob decoded independently back to BF16 weights.
"""
import numpy as np
import pytest

from oracle import mlp as M


def _rand(seed, n=40):
    rng = np.random.default_rng(seed)
    theta = rng.normal(0.0, 0.3, M.COUNT)
    x = rng.normal(size=(n, 8))
    y = rng.normal(size=(n, 4))
    return theta, x, y


def test_gradient_matches_central_differences():
    theta, x, y = _rand(1)
    g = M.gradient(theta, x, y)
    rng = np.random.default_rng(2)
    # every layer's weights and biases, plus random coordinates
    p = M.unpack(np.arange(M.COUNT))
    picks = [int(p[name].reshape(-1)[i]) for name, _ in M.SHAPES for i in (0, 3)] + list(rng.integers(0, M.COUNT, 40))
    picks = [i for i in picks if not (i < 512 and i % 8 == 7)]  # W1's unused input column has zero gradient
    eps = 1e-6
    for i in picks:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += eps
        tm[i] -= eps
        fd = (M.loss(tp, x, y) - M.loss(tm, x, y)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_gradient_of_unused_input_is_zero_and_loss_is_zero_at_target():
    theta, x, y = _rand(3)
    out, _ = M.forward(theta, x)
    assert M.loss(theta, x, out) == 0.0
    g = M.gradient(theta, x, y)
    assert np.all(M.unpack(g)["W1"][:, 7] == 0.0)


def test_bf16_rounding_known_values():
    cases = {1.0: 0x3F80, -2.0: 0xC000, 1.0 + 2 ** -8: 0x3F80,          # tie -> even (down)
             1.0 + 3 * 2 ** -8: 0x3F82,                                  # tie -> even (up)
             1.0 + 2 ** -8 + 2 ** -20: 0x3F81, 3.0e38: 0x7F62, 0.0: 0x0000}
    for value, bits in cases.items():
        r = M.round_bf16(np.array([value], np.float32))
        assert int(r.view(np.uint32)[0] >> 16) == bits, (value, hex(int(r.view(np.uint32)[0] >> 16)), hex(bits))
    x = np.random.default_rng(4).normal(size=10000).astype(np.float32)
    r = M.round_bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8)
    assert np.all(r.view(np.uint32) & 0xFFFF == 0)


def test_inference_blob_layout_decodes_to_the_weights():
    """Decode the product's blob tile by tile (lane l, element i of tile (mu, kt) is W[16mu + l%16][16kt + 4(l//16) + i])."""
    surrogate = pytest.importorskip("baking_amd.surrogate")
    rng = np.random.default_rng(5)
    theta = rng.normal(size=M.COUNT).astype(np.float32)
    shift, scale = rng.normal(size=8).astype(np.float32), rng.uniform(1, 2, 8).astype(np.float32)
    oshift, oscale = rng.normal(size=4).astype(np.float32), rng.uniform(1, 2, 4).astype(np.float32)
    blob = surrogate.build_blob(theta, shift, scale, oshift, oscale)
    assert blob.size == 29856
    head = blob[:96].view(np.float32)
    np.testing.assert_array_equal(head[0:7], shift[:7])
    np.testing.assert_array_equal(head[8:15], scale[:7])
    np.testing.assert_array_equal(head[16:20], oshift)
    np.testing.assert_array_equal(head[20:24], oscale)
    p = M.unpack(theta)
    bias = blob[96:96 + 272 * 4].view(np.float32)
    np.testing.assert_array_equal(bias[:256], np.concatenate([p["b1"], p["b2"], p["b3"], p["b4"]]))
    np.testing.assert_array_equal(bias[256:260], p["b5"])

    def decode(offset, mus, kts, rows, cols):
        t = blob[offset:offset + mus * kts * 512].view(np.uint16).reshape(mus, kts, 64, 4)
        W = np.zeros((16 * mus, 16 * kts), np.float32)
        for mu in range(mus):
            for kt in range(kts):
                for lane in range(64):
                    for i in range(4):
                        W[16 * mu + lane % 16, 16 * kt + 4 * (lane // 16) + i] = \
                            (t[mu, kt, lane, i].astype(np.uint32) << 16).view(np.float32)
        assert np.all(W[rows:, :] == 0) and np.all(W[:, cols:] == 0)
        return W[:rows, :cols]

    off = 96 + 272 * 4
    np.testing.assert_array_equal(decode(off, 4, 1, 64, 8), M.round_bf16(p["W1"]))
    off += 4 * 512
    for name in ("W2", "W3", "W4"):
        np.testing.assert_array_equal(decode(off, 4, 4, 64, 64), M.round_bf16(p[name]))
        off += 16 * 512
    np.testing.assert_array_equal(decode(off, 1, 4, 4, 64), M.round_bf16(p["W5"]))
    assert off + 4 * 512 == blob.size


def test_product_bf16_rounding_matches_the_oracle():
    surrogate = pytest.importorskip("baking_amd.surrogate")
    x = np.random.default_rng(6).normal(size=5000).astype(np.float32) * np.float32(1e3)
    x[:4] = [1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -0.0, 65504.0]
    got = surrogate.bf16_bits(x)
    want = (M.round_bf16(x).view(np.uint32) >> 16).astype(np.uint16)
    np.testing.assert_array_equal(got, want)
