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
    """Decode the product's blob tile by tile: W1's 16x16 tiles (lane l, element i of tile mu is W1[16mu + l%16][4(l//16) + i]),
    the other layers' 16x32 tiles (lane l, element j of tile (mu, kc) is W[16mu + l%16][unit32(kc, l//16, j)])."""
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

    def bits(t):
        return (t.astype(np.uint32) << 16).view(np.float32)

    def decode16(offset, mus, rows, cols):
        t = blob[offset:offset + mus * 512].view(np.uint16).reshape(mus, 64, 4)
        W = np.zeros((16 * mus, 16), np.float32)
        for mu in range(mus):
            for lane in range(64):
                for i in range(4):
                    W[16 * mu + lane % 16, 4 * (lane // 16) + i] = bits(t[mu, lane, i])
        assert np.all(W[rows:, :] == 0) and np.all(W[:, cols:] == 0)
        return W[:rows, :cols]

    def decode32(offset, mus, rows, cols):
        t = blob[offset:offset + mus * 2 * 1024].view(np.uint16).reshape(mus, 2, 64, 8)
        W = np.zeros((16 * mus, 64), np.float32)
        for mu in range(mus):
            for kc in range(2):
                for lane in range(64):
                    for j in range(8):
                        W[16 * mu + lane % 16, unit32(kc, lane // 16, j)] = bits(t[mu, kc, lane, j])
        assert np.all(W[rows:, :] == 0) and np.all(W[:, cols:] == 0)
        return W[:rows, :cols]

    off = 96 + 272 * 4
    np.testing.assert_array_equal(decode16(off, 4, 64, 8), M.round_bf16(p["W1"]))
    off += 4 * 512
    for name in ("W2", "W3", "W4"):
        np.testing.assert_array_equal(decode32(off, 4, 64, 64), M.round_bf16(p[name]))
        off += 8 * 1024
    np.testing.assert_array_equal(decode32(off, 1, 4, 64), M.round_bf16(p["W5"]))
    assert off + 2 * 1024 == blob.size


def unit32(kc, g, j):
    """The unit that K slot (lane group g, element j) of chunk kc stands for in the 16x16x32 layers."""
    return 32 * kc + 4 * g + (j if j < 4 else 16 + j - 4)


def test_register_flow_of_the_16x16x32_layers():
    """The kernel's data flow, emulated lane by lane: output tile mu of a layer leaves lane l units
    16 mu + 4 (l // 16) + i of particle l % 16; chunk kc's B operand is tiles 2 kc and 2 kc + 1 side by side;
    v_mfma_f32_16x16x32_bf16 sums A's element j of lane group g against B's element j of the same group
    (the lane layout the instruction was checked to have on gfx950).  With the blob's tiles as A operands
    the emulated layer equals W @ a for every layer the 16x16x32 form serves."""
    surrogate = pytest.importorskip("baking_amd.surrogate")
    rng = np.random.default_rng(8)
    W = M.round_bf16(rng.normal(size=(64, 64)).astype(np.float32))
    a = M.round_bf16(rng.normal(size=(64, 16)).astype(np.float32))  # units x 16 particles
    tiles = surrogate._tiles32(W, 4, 2).reshape(4, 2, 64, 8)
    A = (tiles.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    # previous layer's output tiles as lanes hold them
    h = np.zeros((4, 64, 4))
    for mu in range(4):
        for lane in range(64):
            h[mu, lane] = a[16 * mu + 4 * (lane // 16) + np.arange(4), lane % 16]
    D = np.zeros((64, 16))
    for mu in range(4):
        for kc in range(2):
            B = np.concatenate([h[2 * kc], h[2 * kc + 1]], axis=1)  # lane x 8
            for m in range(16):
                for n in range(16):
                    D[16 * mu + m, n] += sum(A[mu, kc, m + 16 * g, j] * B[n + 16 * g, j] for g in range(4) for j in range(8))
    np.testing.assert_allclose(D, W.astype(np.float64) @ a.astype(np.float64), rtol=1e-12, atol=1e-12)


def test_product_bf16_rounding_matches_the_oracle():
    surrogate = pytest.importorskip("baking_amd.surrogate")
    x = np.random.default_rng(6).normal(size=5000).astype(np.float32) * np.float32(1e3)
    x[:4] = [1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -0.0, 65504.0]
    got = surrogate.bf16_bits(x)
    want = (M.round_bf16(x).view(np.uint32) >> 16).astype(np.uint16)
    np.testing.assert_array_equal(got, want)
