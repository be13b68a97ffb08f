"""The surrogate's BF16-MFMA kernels against CPU references (psim_amd/csrc/surrogate.hip).

- GEMM: every transpose combination on ragged shapes (not multiples of the 64 x 64 tile or of
  K = 32) against the float64 product of the same bf16-rounded operands.
- Gradients: the GPU backward against torch autograd in float64 on the CPU, run on the same
  bf16-rounded weights and inputs, with the forward's roundings reproduced (activations rounded
  to bf16 between layers; rounding passes gradients straight through).
- Training: Adam on the GPU fits a smooth function of the inputs.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _bf16(x):
    return torch.as_tensor(np.asarray(x, np.float32)).to(torch.bfloat16).to(torch.float64).numpy()


@pytest.fixture(scope="module")
def surrogate():
    from psim_amd import surrogate as s
    return s


@pytest.mark.parametrize("shape", [(70, 45, 100), (72, 40, 104), (4100, 4104, 64), (70, 45, 1100), (512, 512, 8192)],
                         ids=["ragged", "vector-loads", "large-tiles", "split-k", "split-k-16"])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_gemm_matches_host_product(surrogate, ta, tb, shape):
    """Ragged shapes load element by element, shapes whose rows are multiples of 8 load 16-byte
    vectors, products of at least 1024 128 x 128 blocks use the large tile, and a product of few
    blocks with K >= 1024 is split over K and summed in order (512 x 512 x 8192, K-major: 16
    splits)."""
    rng = np.random.default_rng(3 + 2 * ta + tb)
    M, N, K = shape
    a = rng.normal(size=(K, M) if ta else (M, K)).astype(np.float32)
    b = rng.normal(size=(N, K) if tb else (K, N)).astype(np.float32)
    got = surrogate.gemm_check(a, b, ta, tb)
    A = _bf16(a).T if ta else _bf16(a)
    B = _bf16(b).T if tb else _bf16(b)
    want = A @ B
    # FP32 accumulation of 100 exact bf16 x bf16 products
    assert np.max(np.abs(got - want)) < 1e-4 * np.sqrt(K) * np.max(np.abs(want))
    # a transposed or shifted operand would be far off
    assert np.corrcoef(got.ravel(), want.ravel())[0, 1] > 0.9999


def test_gemm_single_tile_identity(surrogate):
    """An identity operand returns the other operand (bf16-rounded) exactly: each MFMA output
    element lands at its own row and column."""
    rng = np.random.default_rng(9)
    a = rng.normal(size=(64, 64)).astype(np.float32)
    eye = np.eye(64, dtype=np.float32)
    np.testing.assert_array_equal(surrogate.gemm_check(a, eye, False, False), _bf16(a).astype(np.float32))
    np.testing.assert_array_equal(surrogate.gemm_check(eye, a, False, False), _bf16(a).astype(np.float32))


def test_device_bf16_rounding_is_nearest_even(surrogate):
    """The device's float -> bf16 rounding (one hardware convert, in every kernel that writes bf16)
    is round-to-nearest-even: exact ties of both parities and signs, values near the largest finite
    bf16 and tiny normal values.  A network whose output is relu(x0) - relu(-x0) (weights +-1, no
    bias) returns the device-rounded first feature unchanged."""
    rng = np.random.default_rng(12)
    base = rng.uniform(1, 2, size=512).astype(np.float32)
    u = base.view(np.uint32) & np.uint32(0xFFFF0000)
    ties = (u | np.uint32(0x8000)).view(np.float32)  # exactly half-way between two bf16 values
    above = (u | np.uint32(0x8001)).view(np.float32)
    below = (u | np.uint32(0x7FFF)).view(np.float32)
    big = rng.uniform(1e37, 3.3e38, size=512).astype(np.float32)
    tiny = rng.uniform(1.2e-38, 1e-36, size=512).astype(np.float32)
    vals = np.concatenate([ties, -ties, above, -below, big, -big, tiny, -tiny])
    net = surrogate.Surrogate(hidden=64, layers=1, seed=2)
    p = np.zeros(net.n_params, np.float32)
    (w1, b1), (w2, b2) = net.layer_views(p)
    w1[0, 0], w1[0, 1] = 1.0, -1.0
    w2[0, 0], w2[1, 0] = 1.0, -1.0
    net.set_params(p)
    x = np.zeros((vals.size, surrogate.N_FEATURES), np.float32)
    x[:, 0] = vals
    got = net.predict(x)
    net.close()
    w = vals.view(np.uint32).astype(np.uint64)
    want = (((w + 0x7FFF + ((w >> 16) & 1)) >> 16) << 16).astype(np.uint32)
    np.testing.assert_array_equal(got.view(np.uint32), want)


def _reference_gradients(net, params, x, t):
    """Loss and flat gradient in float64 on the CPU, with the GPU's bf16 roundings in the forward."""
    views = net.layer_views(params.astype(np.float32))
    h = torch.tensor(_bf16(x), requires_grad=False)
    ws, bs = [], []
    for i, (w, b) in enumerate(views):
        W = torch.tensor(_bf16(w), requires_grad=True)
        B = torch.tensor(b.astype(np.float64), requires_grad=True)
        ws.append(W)
        bs.append(B)
        z = h @ W + B
        if i + 1 < len(views):
            a = torch.relu(z)
            # the next layer reads bf16 activations; the rounding is transparent to the gradient
            h = a + (torch.tensor(_bf16(a.detach().numpy())) - a).detach()
        else:
            h = z
    y = h[:, 0]
    loss = torch.mean((y - torch.tensor(t, dtype=torch.float64)) ** 2)
    loss.backward()
    flat = []
    for W, B in zip(ws, bs):
        flat.append(W.grad.numpy().ravel())
        flat.append(B.grad.numpy().ravel())
    return float(loss.detach()), np.concatenate(flat)


@pytest.mark.parametrize("rows", [150, 2100, 8192])
@pytest.mark.parametrize("hidden,layers", [(96, 3), (512, 4)])
def test_gradients_match_cpu_reference(surrogate, rows, hidden, layers):
    """At 2100 rows the weight gradients (K = rows) are split over K 4 ways, at 8192 rows 16 ways.
    512 x 4 is the warm-start network's own shape (32 -> 512 x 4 -> 1)."""
    net = surrogate.Surrogate(n_in=32, hidden=hidden, layers=layers, seed=1)
    rng = np.random.default_rng(5)
    x = rng.normal(size=(rows, 32)).astype(np.float32)
    t = rng.normal(size=rows).astype(np.float32)
    params = net.params()
    loss, g = net.gradients(x, t)
    ref_loss, ref_g = _reference_gradients(net, params, x, t)
    assert abs(loss - ref_loss) <= 1e-4 * ref_loss
    # per layer: the GPU rounds each back-propagated gradient to bf16 (relative 2^-9)
    off = 0
    for w, b in net.layer_views(params):
        for n in (w.size, b.size):
            got, want = g[off:off + n], ref_g[off:off + n]
            off += n
            scale = np.linalg.norm(want)
            if scale == 0.0:
                np.testing.assert_array_equal(got, 0.0)
                continue
            assert np.linalg.norm(got - want) <= 2e-2 * scale
    assert off == g.size
    # the padded output columns get no gradient
    w_last, b_last = net.layer_views(g)[-1]
    assert np.all(w_last[:, 1:] == 0.0) and np.all(b_last[1:] == 0.0)


def test_gradient_check_rejects_a_wrong_reference(surrogate):
    """The tolerance above is tight enough to see a gradient computed without the ReLU mask."""
    net = surrogate.Surrogate(n_in=32, hidden=96, layers=3, seed=1)
    rng = np.random.default_rng(5)
    x = rng.normal(size=(150, 32)).astype(np.float32)
    t = rng.normal(size=150).astype(np.float32)
    _, g = net.gradients(x, t)
    views = net.layer_views(net.params())
    # rebuild the first layer's weight gradient with the masks dropped (a linear network)
    h = [_bf16(x)]
    for w, b in views[:-1]:
        h.append(_bf16(np.maximum(h[-1] @ _bf16(w) + b, 0.0)))
    y = h[-1] @ _bf16(views[-1][0]) + views[-1][1]
    d = 2.0 * (y[:, 0] - t) / len(t)
    dz = np.outer(d, np.eye(16)[0])
    for w, _ in reversed(views[1:]):
        dz = dz @ _bf16(w).T
    wrong = h[0].T @ dz
    got = g[:views[0][0].size].reshape(views[0][0].shape)
    assert np.linalg.norm(got - wrong) > 2e-2 * np.linalg.norm(wrong)


def test_training_fits_a_smooth_function(surrogate):
    rng = np.random.default_rng(11)
    x = rng.uniform(-1.0, 1.0, size=(8192, 32)).astype(np.float32)
    t = (np.sin(2.0 * x[:, 0]) * x[:, 1] + 0.5 * x[:, 2] ** 2).astype(np.float32)
    net = surrogate.Surrogate(seed=2)
    base = float(np.mean((t - t.mean()) ** 2))
    hist = net.fit(x, t, epochs=160, batch=512, lr=1e-3, seed=0)
    assert hist[-1] < 0.02 * base
    y = net.predict(x[:512])
    assert float(np.mean((y - t[:512]) ** 2)) < 0.03 * base
    # parameters round-trip; prediction depends only on them
    p = net.params()
    other = surrogate.Surrogate(seed=99)
    other.set_params(p)
    np.testing.assert_array_equal(other.predict(x[:512]), y)


def test_too_many_rows_are_rejected(surrogate):
    """Row counts beyond the GEMM grid's reach (65535 blocks of 64 rows) fail with a message
    before any buffer is read."""
    net = surrogate.Surrogate(n_in=32, hidden=64, layers=1, seed=0)
    x = np.zeros((4, 32), np.float32)
    y = np.zeros(4, np.float32)
    L = net._L
    rc = L.psim_mlp_forward(net._h, surrogate._fp(x), 65535 * 64 + 1, surrogate._fp(y))
    assert rc != 0
    assert "rows" in L.psim_mlp_last_error(net._h).decode()
    np.testing.assert_array_equal(net.predict(x), net.predict(x))  # the handle still works


def test_every_tile_and_split_sum_gives_the_same_bits(surrogate, monkeypatch):
    """The autotuner may pick any tile shape of the K-major kernel, and for a product split over K
    either the in-kernel or the separate ordered sum of the partials (PSIM_GEMM_CHOICE forces each
    in turn): a forward product, a split weight-gradient-shaped product and the network's forward
    and gradients must come out bit-identical under every choice."""
    rng = np.random.default_rng(21)
    a = rng.normal(size=(1024, 512)).astype(np.float32)
    b = rng.normal(size=(512, 512)).astype(np.float32)
    aw = rng.normal(size=(512, 4096)).astype(np.float32)
    bw = rng.normal(size=(512, 4096)).astype(np.float32)
    aw16 = rng.normal(size=(512, 8192)).astype(np.float32)
    bw16 = rng.normal(size=(512, 8192)).astype(np.float32)
    # 2048 rows: the hidden layers' weight gradients (K = rows) are split over K
    x = rng.uniform(-1, 1, size=(2048, surrogate.N_FEATURES)).astype(np.float32)
    t = rng.uniform(-1, 1, size=2048).astype(np.float32)
    net = surrogate.Surrogate(seed=4)
    results = []
    for choice in range(12):  # 5 tiles, 5 tiles with the separate split sum, resident weights (4, 8 waves)
        monkeypatch.setenv("PSIM_GEMM_CHOICE", str(choice))
        fwd = surrogate.gemm_check(a, b, False, True)
        split = surrogate.gemm_check(aw, bw, False, True)  # 512 x 512 x 4096: 8 splits over K
        split16 = surrogate.gemm_check(aw16, bw16, False, True)  # 512 x 512 x 8192: 16 splits
        y = net.predict(x)
        loss, grad = net.gradients(x, t)
        results.append((fwd, split, split16, y, grad))
    net.close()
    for r in results[1:]:
        for got, want in zip(r, results[0]):
            np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("shape", [(1, 128, 256), (257, 640, 256), (1000, 512, 512), (4100, 384, 512), (300, 256, 192)],
                         ids=["one-row", "ragged-band", "rows-1000", "three-slices", "k-192"])
@pytest.mark.parametrize("choice", ["10", "11"], ids=["4-waves", "8-waves"])
def test_resident_weight_kernel_matches_the_tiles(surrogate, monkeypatch, shape, choice):
    """The kernel that keeps the weight slice in LDS (PSIM_GEMM_CHOICE 10 with 4 waves, 11 with 8; it
    takes N a multiple of 128 and K a multiple of 256 with 4 waves, of 64 with 8, up to 512: K = 192
    runs on the 8-wave kernel alone, the 4-wave choice falling back to a tile) gives the 64 x 64
    tile's bits on K-major products whose rows end inside a band, matches the host product, and puts
    every element in its place (identity operand)."""
    rng = np.random.default_rng(31)
    M, N, K = shape
    a = rng.normal(size=(M, K)).astype(np.float32)
    b = rng.normal(size=(N, K)).astype(np.float32)
    monkeypatch.setenv("PSIM_GEMM_CHOICE", "4")
    tiled = surrogate.gemm_check(a, b, False, True)
    monkeypatch.setenv("PSIM_GEMM_CHOICE", choice)
    resident = surrogate.gemm_check(a, b, False, True)
    np.testing.assert_array_equal(resident, tiled)
    want = _bf16(a) @ _bf16(b).T
    assert np.max(np.abs(resident - want)) < 1e-4 * np.sqrt(K) * np.max(np.abs(want))
    eye = np.eye(N, K, dtype=np.float32)  # B = [I; 0]: C = A's first N columns
    got = surrogate.gemm_check(a, eye, False, True)
    np.testing.assert_array_equal(got[:, :min(N, K)], _bf16(a)[:, :min(N, K)].astype(np.float32))
    if N > K:
        assert not got[:, K:].any()


def test_resident_weight_kernel_in_the_network(surrogate, monkeypatch):
    """The network's forward and gradients at 300 rows (the last band ragged) are the same bits with
    the resident-weight kernel, at 4 and 8 waves, as with the 64 x 64 tile."""
    rng = np.random.default_rng(33)
    x = rng.uniform(-1, 1, size=(300, surrogate.N_FEATURES)).astype(np.float32)
    t = rng.uniform(-1, 1, size=300).astype(np.float32)
    net = surrogate.Surrogate(seed=5)
    out = []
    for choice in ("4", "10", "11"):
        monkeypatch.setenv("PSIM_GEMM_CHOICE", choice)
        out.append((net.predict(x), net.gradients(x, t)))
    net.close()
    for o in out[1:]:
        np.testing.assert_array_equal(out[0][0], o[0])
        assert out[0][1][0] == o[1][0]
        np.testing.assert_array_equal(out[0][1][1], o[1][1])


def test_graph_replayed_training_steps_equal_eager_steps(surrogate):
    """Training steps replayed from the captured graph (the default: eager at a new batch size,
    captured on the next step, replayed after) give the same losses and parameters bit for bit as
    all-eager steps, with eager work (a forward pass) between the steps and a change of batch
    size that makes the step be captured again."""
    rng = np.random.default_rng(17)
    graphed = surrogate.Surrogate(seed=6)
    eager = surrogate.Surrogate(seed=6)
    eager.set_graph(False)
    probe = rng.uniform(-1, 1, size=(300, surrogate.N_FEATURES)).astype(np.float32)
    for k, rows in enumerate([1024, 1024, 1024, 1024, 2048, 2048, 2048, 1024]):
        print(f"training step {k}, {rows} rows", flush=True)
        x = rng.uniform(-1, 1, size=(rows, surrogate.N_FEATURES)).astype(np.float32)
        t = rng.uniform(-1, 1, size=rows).astype(np.float32)
        lg = graphed.train_step(x, t, 1e-3)
        le = eager.train_step(x, t, 1e-3)
        assert lg == le
        np.testing.assert_array_equal(graphed.params(), eager.params())
        np.testing.assert_array_equal(graphed.predict(probe), eager.predict(probe))
    graphed.close()
    eager.close()
