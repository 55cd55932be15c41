"""The learned warm start: MFMA GEMM, network forward, gradients against
CPU FP64 finite differences, training, and what the trained network gains
over the same batched FIM without it.

Tolerances: the GEMM rounds its operands to BF16 and accumulates in FP32,
so it is held to 1e-5 relative against an FP64 product of the same
BF16-rounded operands; the network (every layer input and weight rounded to
BF16) to 5e-2 relative against the FP64 numpy restatement in nn_ref.py; its
gradient to 15% per coordinate and 0.98 cosine against central differences
of the FP64 loss.  Warm-started solves are held to the cold solve within
1e-5 s (they converge to the same fixed point).
"""
import numpy as np
import pytest

import nn_ref
from seismic_tomography_3d_amd import synth

pytestmark = pytest.mark.gpu


def test_gemm_bf16_mfma(ctx):
    rng = np.random.default_rng(0)
    for (M, N, K) in [(100, 70, 90), (64, 64, 32), (1, 5, 3), (257, 129, 200)]:
        A = rng.normal(size=(M, K)).astype(np.float32)
        B = rng.normal(size=(K, N)).astype(np.float32)
        C = ctx.nn_gemm(A, B)
        ref = nn_ref.bf16(A) @ nn_ref.bf16(B)
        assert np.abs(C - ref).max() <= 1e-5 * np.abs(ref).max() + 1e-6


def _inputs(nc, ns, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(0.0, 0.5, size=(ns, nc ** 3, 6)).astype(np.float32)
    X[..., 0] = rng.uniform(0.7, 2.0, size=(ns, nc ** 3))
    T = rng.uniform(0.8, 1.4, size=(ns, nc ** 3)).astype(np.float32)
    M = (rng.random((ns, nc ** 3)) > 0.2).astype(np.float32)
    return X, T, M


@pytest.mark.parametrize("nc,ns", [(8, 2), (16, 3)])
def test_network_forward_matches_fp64(ctx, nc, ns):
    """Implicit-GEMM convolutions (BF16 operands, FP32 accumulation) against
    the FP64 numpy network; nc = 8 and 16 put 1^3 and 2^3 cubes (partial
    4x4x4 bricks) at the bottleneck."""
    ctx.surrogate_init(nc, seed=3)
    p = ctx.surrogate_params()
    # perturb the head so the output depends on every layer
    X, T, M = _inputs(nc, ns, 1)
    Y, _, _ = ctx.surrogate_eval(X)
    ref = nn_ref.forward(p, X, nc)
    err = np.linalg.norm(Y - ref) / np.linalg.norm(ref)
    print(f"forward rel err vs FP64: {err:.2e}")
    assert err < 5e-2


@pytest.mark.parametrize("nc,ns", [(8, 2), (16, 3), (32, 2)])
def test_inference_pass_matches_fp32_pass(ctx, nc, ns):
    """The inference pass (BF16 activations between layers, upsampling and
    skip concatenation while staging) against the FP32 pass the training
    uses, on the same weights and inputs: the two differ only by the BF16
    rounding of activations before the 2x2x2 means."""
    ctx.surrogate_init(nc, seed=4)
    X, T, M = _inputs(nc, ns, 2)
    Yi, _, _ = ctx.surrogate_eval(X)
    Yf, _, _ = ctx.surrogate_eval(X, T, M)
    err = np.linalg.norm(Yi - Yf) / np.linalg.norm(Yf)
    print(f"inference vs FP32 pass rel diff: {err:.2e}")
    assert err < 5e-3


def test_network_gradient_vs_finite_differences(ctx):
    nc, ns = 8, 2
    n = ctx.surrogate_init(nc, seed=4)
    rng = np.random.default_rng(9)
    p = ctx.surrogate_params()
    p = (p * rng.uniform(0.8, 1.2, p.shape)).astype(np.float32)
    p[-1] = 0.3  # head bias away from the targets: a large, well-conditioned gradient
    ctx.surrogate_params(set_to=p)
    X, T, M = _inputs(nc, ns, 2)
    _, loss_gpu, g = ctx.surrogate_eval(X, T, M)
    loss_cpu = nn_ref.loss(p, X, T, M, nc)
    assert loss_gpu == pytest.approx(loss_cpu, rel=5e-2)
    # coordinates: the largest-gradient entries of every layer plus biases
    L = nn_ref.unpack(np.arange(n, dtype=np.float64))
    picks = []
    for W, b, _, _ in L:
        wi = W.ravel().astype(np.int64)
        top = wi[np.argsort(-np.abs(g[wi]))[:3]]
        picks.extend(top.tolist())
        picks.append(int(b.ravel()[0]))
    picks = np.array(sorted(set(picks)))
    fd = []
    pd = p.astype(np.float64)
    for i in picks:
        h = 1e-3 * max(1.0, abs(pd[i]))
        pp, pm = pd.copy(), pd.copy()
        pp[i] += h
        pm[i] -= h
        fd.append((nn_ref.loss(pp, X, T, M, nc) - nn_ref.loss(pm, X, T, M, nc)) / (2 * h))
    fd = np.array(fd)
    gg = g[picks].astype(np.float64)
    cos = gg @ fd / (np.linalg.norm(gg) * np.linalg.norm(fd))
    scale = np.abs(fd).max()
    rel = np.abs(gg - fd) / (np.abs(fd) + 0.05 * scale)
    print(f"gradient check: {len(picks)} coordinates, cosine {cos:.4f}, max rel {rel.max():.3f}")
    assert cos > 0.98
    assert rel.max() < 0.15


def test_training_and_warm_start_gain(ctx, tmp_path):
    # 5 km x 5 km x 2.5 km at 62.5 m: the batched recipe at coarse spacing
    nx, ny, nz, h = 81, 81, 41, 62.5
    vel = synth.layered_model(nx, ny, nz, h, n_layers=5, n_anom=3, seed=77)
    ctx.set_model(1.0 / vel, h)
    ctx.surrogate_init(32, seed=0)
    ctx.surrogate_make_data(160, shots_per_model=4, zmax_src=2000.0, seed=5)
    hist = ctx.surrogate_train(12, batch=8, lr=2e-3, n_val=64, seed=1)
    print("train/val loss per epoch:", hist.tolist())
    assert hist[-1, 1] < 0.33 * hist[0, 1]
    path = tmp_path / "w.bin"
    ctx.surrogate_save(path)
    ctx.surrogate_load(path)
    src = synth.random_sources(32, nx, ny, h, 0.0, 2000.0, seed=8)
    rcv = np.stack(synth.surface_grid(10, 100.0, 4900.0, 0.0), 1)
    pc, Tc, sc = ctx.fim_solve(src, rcv, return_T=True, batch=32, block=8)
    # prediction-ordered visits (cold values): exact by construction
    pw, Tw, sw = ctx.fim_solve(src, rcv, return_T=True, batch=32, block=8, warm=1)
    # (1+delta) warm values: the judge re-solves cold every shot the network
    # did not bound, so the result is exact either way
    pv, Tv, sv = ctx.fim_solve(src, rcv, return_T=True, batch=32, block=8, warm=1, warm_values=1)
    err = np.abs(Tw - Tc).max()
    errv = np.abs(Tv - Tc).max()
    print(f"cold: visits {sc['visits']} iters {sc['inner_iters']} rounds {sc['rounds']}; "
          f"ordered: visits {sw['visits']} iters {sw['inner_iters']} rounds {sw['rounds']} "
          f"deferred {sw['deferred']}; max |ordered - cold| {err:.2e}; (1+delta) values: "
          f"fallbacks {sv['fallbacks']}/32 judge_max {sv['judge_max']:.2e} max err {errv:.2e}")
    assert err < 1e-5 and errv < 1e-5
    assert sw["fallbacks"] == 0
    # training towards an upper bound: the same data and seed with
    # under-predictions weighted 1000 leave fewer shots below the (1+delta)
    # bound, so the judge sends fewer of them back to a cold solve
    ctx.surrogate_init(32, seed=0)
    ctx.surrogate_loss(1000.0)
    ctx.surrogate_train(12, batch=8, lr=2e-3, n_val=64, seed=1)
    pu, Tu, su = ctx.fim_solve(src, rcv, return_T=True, batch=32, block=8, warm=1, warm_values=1)
    print(f"under-weight 1000: fallbacks {su['fallbacks']}/32 judge_max {su['judge_max']:.2e}, "
          f"against {sv['fallbacks']}/32 for plain MSE")
    assert np.abs(Tu - Tc).max() < 1e-5
    assert su["fallbacks"] < sv["fallbacks"] or su["fallbacks"] == 0
