"""GPU: surrogate MLP kernels against torch FP32 products and the FP64 MLP oracle.

Tolerances: GEMMs take BF16 operands and accumulate in FP32, so they are
compared with an FP32 product of the same BF16-rounded operands (rtol 1e-3,
summation order only).  Whole-network values, gradients and JVPs round every
activation to BF16 and are compared norm-wise against the FP64 oracle
(relative error <= 3e-2).
"""

import numpy as np
import pytest

from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _lib():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pyftle_amd import surrogate

    surrogate._lib()


def _bf(t):
    return t.to(torch.bfloat16).contiguous()


def _rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


@pytest.mark.parametrize("a_kc,b_kc", [(1, 1), (1, 0), (0, 1), (0, 0)])
def test_gemm_layouts(a_kc, b_kc):
    from pyftle_amd.surrogate import EPI_F32, _p, gemm

    g = torch.Generator(device="cuda").manual_seed(1)
    M, N, K = 256, 384, 192
    A = torch.randn(M, K, device="cuda", generator=g)
    B = torch.randn(K, N, device="cuda", generator=g)
    Ab = _bf(A if a_kc else A.t())
    Bb = _bf(B.t() if b_kc else B)
    C = torch.empty(M, N, device="cuda")
    gemm(_p(Ab), Ab.shape[1], a_kc, _p(Bb), Bb.shape[1], b_kc, M, N, K, _p(C), N, epi=EPI_F32)
    Ar = Ab.float() if a_kc else Ab.float().t()
    Br = Bb.float().t() if b_kc else Bb.float()
    ref = Ar @ Br
    torch.testing.assert_close(C, ref, rtol=1e-3, atol=1e-3)


def test_gemm_split_k_and_epilogues():
    from pyftle_amd.surrogate import EPI_BIAS_SILU, EPI_DSILU, EPI_F32, _p, gemm

    g = torch.Generator(device="cuda").manual_seed(2)
    M, N, K = 256, 128, 512
    A = _bf(torch.randn(M, K, device="cuda", generator=g))
    W = _bf(torch.randn(N, K, device="cuda", generator=g) * 0.05)
    bias = torch.randn(N, device="cuda", generator=g)
    ref = A.float() @ W.float().t()
    part = torch.empty(4, M, N, device="cuda")
    gemm(_p(A), K, 1, _p(W), K, 1, M, N, K, _p(part), N, epi=EPI_F32, splits=4)
    torch.testing.assert_close(part.sum(0), ref, rtol=1e-3, atol=1e-3)
    H = torch.empty(M, N, dtype=torch.bfloat16, device="cuda")
    Z = torch.empty(M, N, dtype=torch.bfloat16, device="cuda")
    gemm(_p(A), K, 1, _p(W), K, 1, M, N, K, _p(H), N, epi=EPI_BIAS_SILU, bias=_p(bias), Z=_p(Z))
    z = ref + bias
    torch.testing.assert_close(Z.float(), z, rtol=1e-2, atol=1e-2)
    torch.testing.assert_close(H.float(), torch.nn.functional.silu(z), rtol=1e-2, atol=1e-2)
    # SiLU' epilogue with the row period of stacked tangent blocks (2 blocks of 128 rows)
    D = torch.empty(M, N, dtype=torch.bfloat16, device="cuda")
    gemm(_p(A), K, 1, _p(W), K, 1, M, N, K, _p(D), N, epi=EPI_DSILU, Z=_p(Z), z_rows=128)
    zz = Z.float()[:128].repeat(2, 1)
    s = torch.sigmoid(zz)
    torch.testing.assert_close(D.float(), ref * s * (1 + zz * (1 - s)), rtol=2e-2, atol=2e-2)


def test_gemm_rejects_bad_shapes():
    from pyftle_amd import _native
    from pyftle_amd.surrogate import _p, gemm

    A = torch.zeros(100, 64, dtype=torch.bfloat16, device="cuda")
    with pytest.raises(_native.PyftleError, match="multiples of 128"):
        gemm(_p(A), 64, 1, _p(A), 64, 1, 100, 128, 64, _p(A), 128)


def _small_model(seed=0, width=128, depth=3):
    from pyftle_amd.surrogate import FlowMapMLP, MLPSpec

    spec = MLPSpec(d_in=3, d_out=2, n_freq=64, width=width, depth=depth, sigma=1.5, scale=(0.5, 1.0, 0.1))
    return FlowMapMLP(spec, seed=seed)


def test_forward_and_jvp_match_oracle():
    m = _small_model()
    p = m.to_numpy()
    rng = np.random.default_rng(0)
    x = rng.uniform([0, 0, 0], [2, 1, 10], (300, 3))
    xt = torch.as_tensor(x, dtype=torch.float32, device="cuda")
    y, tans = m.forward(xt, n_tangents=2)
    assert _rel(y.cpu().numpy(), MO.forward(p, x)) < 3e-2
    for a in range(2):
        assert _rel(tans[a].cpu().numpy(), MO.jvp(p, x, a)) < 3e-2


def test_gradients_match_oracle():
    m = _small_model(seed=1)
    p = m.to_numpy()
    rng = np.random.default_rng(1)
    x = rng.uniform([0, 0, 0], [2, 1, 10], (500, 3))
    t = rng.standard_normal((500, 2)) * 0.3
    loss = m.loss_and_backward(torch.as_tensor(x, dtype=torch.float32, device="cuda"),
                               torch.as_tensor(t, dtype=torch.float32, device="cuda"))
    l_ref, g_ref = MO.loss_and_grads(p, x, t)
    assert abs(loss.item() - l_ref) / l_ref < 3e-2
    g = m.grads
    for l in range(m.spec.depth):
        o, shp, ob = m.offsets[l]
        gw = g[o:o + shp[0] * shp[1]].view(shp).cpu().numpy()
        gb = g[ob:ob + shp[0]].cpu().numpy()
        assert _rel(gw, g_ref["W"][l]) < 3e-2, l
        assert _rel(gb, g_ref["b"][l]) < 3e-2, l
    o, shp, ob = m.out_off
    assert _rel(g[o:o + shp[0] * shp[1]].view(shp).cpu().numpy(), g_ref["Wout"]) < 3e-2
    assert _rel(g[ob:ob + 2].cpu().numpy(), g_ref["bout"]) < 3e-2


def test_adam_matches_formula():
    m = _small_model(seed=2)
    w0 = m.params.clone()
    m.grads.copy_(torch.randn_like(m.grads))
    gr = m.grads.clone()
    m.adam_step(1e-2)
    mm = 0.1 * gr
    vv = 0.001 * gr * gr
    exp = w0 - 1e-2 * (mm / 0.1) / (torch.sqrt(vv / 0.001) + 1e-8)
    torch.testing.assert_close(m.params, exp, rtol=1e-5, atol=1e-6)
    torch.testing.assert_close(m.params_bf16.float(), m.params, rtol=1e-2, atol=1e-2)


def test_training_reduces_loss():
    from pyftle_amd.surrogate import train

    m = _small_model(seed=3)
    g = torch.Generator(device="cuda").manual_seed(0)
    X = torch.rand(8192, 3, device="cuda", generator=g) * torch.tensor([2.0, 1.0, 10.0], device="cuda")
    T = torch.stack([torch.sin(np.pi * X[:, 0]) * 0.3, torch.cos(np.pi * X[:, 1]) * 0.2 * torch.sin(X[:, 2])], 1)
    losses = train(m, X, T, epochs=15, batch=2048, lr=3e-3)
    print(f"[gpu] loss {losses[0]:.3e} -> {losses[-1]:.3e}", flush=True)
    assert losses[-1] < 0.1 * losses[0]


def _fused_model(seed=4, depth=3, d_out=2):
    from pyftle_amd.surrogate import FlowMapMLP, MLPSpec

    spec = MLPSpec(d_in=3, d_out=d_out, n_freq=64, width=256, depth=depth, sigma=1.5, scale=(0.5, 1.0, 0.1),
                   out_scale=0.7)
    return FlowMapMLP(spec, seed=seed)


@pytest.mark.parametrize("M", [1, 37, 300, 4096])
def test_fused_forward_and_jvp_match_oracle(M):
    """One-kernel network at width 256 (activations in LDS, 2 tangents) vs FP64."""
    m = _fused_model()
    p = m.to_numpy()
    rng = np.random.default_rng(M)
    x = rng.uniform([0, 0, 0], [2, 1, 10], (M, 3))
    Y = m.forward_fused(torch.as_tensor(x, dtype=torch.float32, device="cuda"), n_tangents=2).cpu().numpy()
    assert Y.shape == (3, M, 2)
    assert _rel(Y[0], MO.forward(p, x)) < 3e-2
    for a in range(2):
        assert _rel(Y[1 + a], MO.jvp(p, x, a)) < 3e-2
    Y0 = m.forward_fused(torch.as_tensor(x, dtype=torch.float32, device="cuda")).cpu().numpy()
    assert np.array_equal(Y0[0], Y[0])  # the primal rows do not depend on the tangents


def test_fused_matches_layered_path():
    m = _fused_model(seed=5, depth=5)
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.rand(1000, 3, device="cuda", generator=g) * torch.tensor([2.0, 1.0, 10.0], device="cuda")
    y, tans = m.forward(x, n_tangents=2)
    Y = m.forward_fused(x, n_tangents=2)
    assert _rel(Y[0].cpu().numpy(), y.cpu().numpy()) < 2e-2
    for a in range(2):
        assert _rel(Y[1 + a].cpu().numpy(), tans[a].cpu().numpy()) < 2e-2


def test_fused_three_outputs_and_repack_after_update():
    m = _fused_model(seed=6, depth=4, d_out=3)
    p = m.to_numpy()
    rng = np.random.default_rng(9)
    x = rng.uniform([0, 0, 0], [2, 1, 10], (200, 3))
    xt = torch.as_tensor(x, dtype=torch.float32, device="cuda")
    assert _rel(m.forward_fused(xt)[0].cpu().numpy(), MO.forward(p, x)) < 3e-2
    m.grads.copy_(torch.randn_like(m.grads))
    m.adam_step(1e-2)  # weights change: the packed copy must follow
    p2 = m.to_numpy()
    assert _rel(m.forward_fused(xt)[0].cpu().numpy(), MO.forward(p2, x)) < 3e-2


def _wide_model(seed=11, depth=3, d_out=3):
    from pyftle_amd.surrogate import FlowMapMLP, MLPSpec

    spec = MLPSpec(d_in=4, d_out=d_out, n_freq=64, width=512, depth=depth, sigma=0.5,
                   scale=(0.5, 0.5, 0.5, 0.1), out_scale=2.0)
    return FlowMapMLP(spec, seed=seed)


@pytest.mark.parametrize("M", [1, 15, 17, 300, 5000])
def test_wide_forward_and_jvp_match_oracle(M):
    """The one-kernel network at width 512 vs FP64: primal and 3 tangents;
    M covers partial tiles of 16 and 64 seeds."""
    m = _wide_model()
    p = m.to_numpy()
    rng = np.random.default_rng(M)
    x = rng.uniform([0, 0, 0, -1], [6.3, 6.3, 6.3, 11], (M, 4))
    xt = torch.as_tensor(x, dtype=torch.float32, device="cuda")
    Y = m.forward_fused(xt, n_tangents=3).cpu().numpy()
    assert Y.shape == (4, M, 3)
    assert _rel(Y[0], MO.forward(p, x)) < 3e-2
    for a in range(3):
        assert _rel(Y[1 + a], MO.jvp(p, x, a)) < 3e-2
    Y0 = m.forward_fused(xt).cpu().numpy()
    assert np.array_equal(Y0[0], Y[0])  # the primal rows do not depend on the tangents


def test_wide_matches_layered_path_and_two_outputs():
    m = _wide_model(seed=12, depth=6)
    g = torch.Generator(device="cuda").manual_seed(4)
    x = torch.rand(3000, 4, device="cuda", generator=g) * torch.tensor([6.3, 6.3, 6.3, 10.0], device="cuda")
    y, tans = m.forward(x, n_tangents=3)
    Y = m.forward_fused(x, n_tangents=3)
    assert _rel(Y[0].cpu().numpy(), y.cpu().numpy()) < 2e-2
    for a in range(3):
        assert _rel(Y[1 + a].cpu().numpy(), tans[a].cpu().numpy()) < 2e-2
    m2 = _wide_model(seed=13, depth=2, d_out=2)
    p2 = m2.to_numpy()
    x2 = np.random.default_rng(5).uniform([0, 0, 0, 0], [2, 1, 1, 10], (700, 4))
    Y2 = m2.forward_fused(torch.as_tensor(x2, dtype=torch.float32, device="cuda"))
    assert Y2.shape == (1, 700, 2)
    assert _rel(Y2[0].cpu().numpy(), MO.forward(p2, x2)) < 3e-2


@pytest.mark.parametrize("width,d_in,nt", [(512, 4, 3), (512, 4, 0), (256, 3, 2), (256, 3, 0)])
def test_fused_persistent_tiles_match_layered_path(width, d_in, nt):
    """Enough seeds that every work-group loops over several tiles (the weight prefetch
    carries over from one tile's last layer to the next tile's first)."""
    from pyftle_amd.surrogate import FlowMapMLP, MLPSpec

    d_out = 3 if width == 512 else 2
    spec = MLPSpec(d_in=d_in, d_out=d_out, n_freq=64, width=width, depth=3, sigma=0.5,
                   scale=(0.5,) * (d_in - 1) + (0.1,), out_scale=1.5)
    m = FlowMapMLP(spec, seed=21)
    M = 70001
    g = torch.Generator(device="cuda").manual_seed(6)
    x = torch.rand(M, d_in, device="cuda", generator=g) * 6.0
    Y = m.forward_fused(x, n_tangents=nt)
    if nt:
        y, tans = m.forward(x, n_tangents=nt)
        refs = [y] + list(tans)
    else:
        refs = [m.forward(x)]
    for k, ref in enumerate(refs):
        assert _rel(Y[k].cpu().numpy(), ref.cpu().numpy()) < 2e-2
    # row by row: the last tiles agree as well as the first
    tail = slice(M - 100, M)
    assert _rel(Y[0][tail].cpu().numpy(), refs[0][tail].cpu().numpy()) < 2e-2


@pytest.mark.parametrize("width,depth", [(256, 2), (256, 8), (512, 2), (512, 8)])
def test_fused_depth_extremes_match_oracle(width, depth):
    """The stored activations carry a factor -log2(e) and the tangent rows a further 1/ln 2
    per hidden layer, undone at the output: the shallowest and deepest supported networks
    still match the FP64 oracle (primal and every tangent)."""
    from pyftle_amd.surrogate import FlowMapMLP, MLPSpec

    d_in, d_out, nt = (3, 2, 2) if width == 256 else (4, 3, 3)
    spec = MLPSpec(d_in=d_in, d_out=d_out, n_freq=64, width=width, depth=depth, sigma=0.7,
                   scale=(0.5,) * (d_in - 1) + (0.1,), out_scale=1.3)
    m = FlowMapMLP(spec, seed=40 + depth)
    p = m.to_numpy()
    x = np.random.default_rng(depth).uniform(0.0, 3.0, (333, d_in))
    Y = m.forward_fused(torch.as_tensor(x, dtype=torch.float32, device="cuda"), n_tangents=nt).cpu().numpy()
    assert _rel(Y[0], MO.forward(p, x)) < 3e-2
    for a in range(nt):
        assert _rel(Y[1 + a], MO.jvp(p, x, a)) < 3e-2


def test_fused_rejects_unsupported_shapes():
    m = _small_model()  # width 128
    with pytest.raises(ValueError, match="fused inference"):
        m.forward_fused(torch.zeros(4, 3, device="cuda"))


@pytest.mark.parametrize("fused", [True, False])
def test_concurrent_ensemble_training_equals_sequential(fused):
    """train_ensemble (the members' runs interleaved, a stream each) leaves every member's
    parameters, Adam moments and bf16 weights, and (fused) its per-epoch losses, exactly as its
    own train() call does."""
    from pyftle_amd.surrogate import train, train_ensemble

    g = torch.Generator(device="cuda").manual_seed(8)
    X = torch.rand(8192, 3, device="cuda", generator=g) * torch.tensor([2.0, 1.0, 10.0], device="cuda")
    T = torch.stack([torch.sin(np.pi * X[:, 0]) * 0.3, torch.cos(np.pi * X[:, 1]) * 0.2], 1)
    seq = [_train_model(256, 3, 3, 2, seed=20 + k) for k in range(3)]
    con = [_train_model(256, 3, 3, 2, seed=20 + k) for k in range(3)]
    l_seq = [train(m, X, T, epochs=3, batch=1024, lr=3e-3, seed=k, fused=fused) for k, m in enumerate(seq)]
    l_con = train_ensemble(con, X, T, epochs=3, batch=1024, lr=3e-3, seeds=range(3), fused=fused)
    torch.cuda.synchronize()
    for k in range(3):
        assert torch.equal(seq[k].params, con[k].params), k
        assert torch.equal(seq[k].m, con[k].m) and torch.equal(seq[k].v, con[k].v), k
        assert torch.equal(seq[k].params_bf16, con[k].params_bf16), k
        if fused:
            assert l_seq[k] == l_con[k], k
        else:  # the layered path adds its loss partials with atomics: the last bit follows their order
            np.testing.assert_allclose(l_con[k], l_seq[k], rtol=1e-6)
        assert seq[k].step == con[k].step == 24
    assert not torch.equal(con[0].params, con[1].params)


def test_graph_captured_training_equals_eager():
    """The training step replayed from a HIP graph (with eager shuffles and host syncs
    between replays) updates the weights exactly as the same step run eagerly."""
    from pyftle_amd.surrogate import train

    g = torch.Generator(device="cuda").manual_seed(5)
    X = torch.rand(8192, 3, device="cuda", generator=g) * torch.tensor([2.0, 1.0, 10.0], device="cuda")
    T = torch.stack([torch.sin(np.pi * X[:, 0]) * 0.3, torch.cos(np.pi * X[:, 1]) * 0.2], 1)
    m_eager, m_graph = _small_model(seed=7), _small_model(seed=7)
    l_eager = train(m_eager, X, T, epochs=3, batch=1024, lr=3e-3, graph=False)
    l_graph = train(m_graph, X, T, epochs=3, batch=1024, lr=3e-3, graph=True)
    assert torch.equal(m_eager.params, m_graph.params)
    assert torch.equal(m_eager.params_bf16, m_graph.params_bf16)
    np.testing.assert_allclose(l_graph, l_eager, rtol=1e-5)
    assert m_graph.step == m_eager.step == 24


def _train_model(width, depth, d_in, d_out, seed=7):
    from pyftle_amd.surrogate import FlowMapMLP, MLPSpec

    spec = MLPSpec(d_in=d_in, d_out=d_out, n_freq=64, width=width, depth=depth, sigma=1.0,
                   scale=(0.5, 1.0, 0.1, 0.3)[:d_in], out_scale=0.7)
    return FlowMapMLP(spec, seed=seed)


@pytest.mark.parametrize("width,depth,d_in,d_out", [(256, 5, 3, 2), (256, 3, 3, 3), (512, 3, 4, 3)])
def test_fused_training_step_gradients_match_oracle(width, depth, d_in, d_out):
    """pyftle_mlp_train_grads: every parameter's MSE gradient of one minibatch against the
    FP64 oracle (BF16 activations and weights: 3e-2 norm-wise, as the layered kernels) and
    against the layered backward pass; the loss is added and the row offset advanced."""
    from pyftle_amd.surrogate import FusedTrainStep

    m = _train_model(width, depth, d_in, d_out)
    p = m.to_numpy()
    rng = np.random.default_rng(11)
    M, extra = 1024, 128
    lo, hi = [0, 0, 0, 0][:d_in], [2, 1, 10, 3][:d_in]
    x = rng.uniform(lo, hi, (M + extra, d_in))
    t = rng.standard_normal((M + extra, d_out)) * 0.3
    X = torch.as_tensor(x, dtype=torch.float32, device="cuda")
    Tn = torch.as_tensor(t / m.spec.out_scale, dtype=torch.float32, device="cuda")
    base = torch.tensor([extra], dtype=torch.int64, device="cuda")  # rows extra .. extra + M - 1
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    step = FusedTrainStep(m, X, Tn, M, base, loss, splits=4)
    step.grads()
    torch.cuda.synchronize()
    g_fused = m.grads.clone()
    assert int(base.item()) == extra + M
    xs, ts = x[extra:], t[extra:]
    # the oracle's output includes out_scale; the kernel fits T / out_scale with the unscaled output
    p_unscaled = dict(p, Wout=p["Wout"] / m.spec.out_scale, bout=p["bout"] / m.spec.out_scale)
    l_ref, g_ref = MO.loss_and_grads(p_unscaled, xs, ts / m.spec.out_scale)
    assert abs(loss.item() - l_ref) / l_ref < 3e-2
    for l in range(depth):
        o, shp, ob = m.offsets[l]
        assert _rel(g_fused[o:o + shp[0] * shp[1]].view(shp).cpu().numpy(), g_ref["W"][l]) < 3e-2, l
        assert _rel(g_fused[ob:ob + shp[0]].cpu().numpy(), g_ref["b"][l]) < 3e-2, l
    o, shp, ob = m.out_off
    assert _rel(g_fused[o:o + shp[0] * shp[1]].view(shp).cpu().numpy(), g_ref["Wout"]) < 3e-2
    assert _rel(g_fused[ob:ob + d_out].cpu().numpy(), g_ref["bout"]) < 3e-2
    # the layered backward pass on the same rows
    m.loss_and_backward(X[extra:], Tn[extra:] * m.spec.out_scale)
    torch.cuda.synchronize()
    rel = ((m.grads - g_fused).norm() / m.grads.norm()).item()
    print(f"[fused train {width}x{depth}] |g_fused - g_layered| / |g| = {rel:.2e}", flush=True)
    assert rel < 2e-2


def test_fused_training_reduces_loss_like_layered():
    from pyftle_amd.surrogate import train

    g = torch.Generator(device="cuda").manual_seed(0)
    X = torch.rand(32768, 3, device="cuda", generator=g) * torch.tensor([2.0, 1.0, 10.0], device="cuda")
    T = torch.stack([torch.sin(np.pi * X[:, 0]) * 0.3, torch.cos(np.pi * X[:, 1]) * 0.2 * torch.sin(X[:, 2])], 1)
    out = {}
    for fused in (True, False):
        m = _train_model(256, 3, 3, 2, seed=3)
        out[fused] = train(m, X, T, epochs=12, batch=2048, lr=3e-3, fused=fused)
    print(f"[gpu] fused loss {out[True][0]:.3e} -> {out[True][-1]:.3e}; layered {out[False][-1]:.3e}", flush=True)
    assert out[True][-1] < 0.1 * out[True][0]
    assert out[True][-1] < 2.0 * out[False][-1]


@pytest.mark.parametrize("width,depth,d_in,d_out", [(256, 5, 3, 2), (512, 3, 4, 3)])
def test_fused_train_step_equals_grads_then_adam(width, depth, d_in, d_out):
    """pyftle_mlp_train_step (the Adam update applied by the gradient reduction) against
    pyftle_mlp_train_grads followed by pyftle_mlp_train_adam, three steps from the same state:
    parameters, moments, packed weights, gradients, loss, batch offset and step index bit for bit."""
    from pyftle_amd.surrogate import FusedTrainStep

    rng = np.random.default_rng(5)
    M = 2048
    lo, hi = [0, 0, 0, 0][:d_in], [2, 1, 10, 3][:d_in]
    X = torch.as_tensor(rng.uniform(lo, hi, (4 * M, d_in)), dtype=torch.float32, device="cuda")
    Tn = torch.as_tensor(rng.standard_normal((4 * M, d_out)) * 0.3, dtype=torch.float32, device="cuda")
    state = []
    for fused in (False, True):
        m = _train_model(width, depth, d_in, d_out, seed=2)
        base = torch.zeros(1, dtype=torch.int64, device="cuda")
        loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        k = torch.full((1,), -1, dtype=torch.int32, device="cuda")
        st = FusedTrainStep(m, X, Tn, M, base, loss, splits=4, step=k)
        for _ in range(3):
            if fused:
                st.step(3e-3, 3e-5, 10, 0)
            else:
                st.grads()
                st.adam(3e-3, 3e-5, 10, 0, k)
        torch.cuda.synchronize()
        state.append((m.params.clone(), m.m.clone(), m.v.clone(), m.grads.clone(), st.packed.clone(),
                      loss.clone(), base.clone(), k.clone()))
    names = ("params", "m", "v", "grads", "packed", "loss", "base", "step")
    for name, a, b in zip(names, *state):
        assert torch.equal(a, b), name
    assert int(state[1][7].item()) == 2 and int(state[1][6].item()) == 3 * M
