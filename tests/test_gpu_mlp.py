"""GPU BF16 MFMA network against the oracle MLP: the GEMM primitive (lane
layout, strided/transposed operands; exact BF16 products, FP32 sums),
forward outputs and loss gradients (norm-wise relative 1e-3: the device sums
in FP32 where the oracle sums in FP64, which can move a BF16 rounding by one
unit), one Adam step, bitwise determinism, and training that fits a
regression target."""

import numpy as np
import pytest
import torch

from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


def _bf16_bits(a):
    bits = (np.ascontiguousarray(omlp.to_bf16(a).astype(np.float32)).view(np.uint32) >> 16).astype(np.uint16)
    return torch.as_tensor(bits.view(np.int16)).cuda()


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def test_gemm_bf16_layout_and_strides():
    from simkit_amd import _lib
    from simkit_amd._device import ptr, stream
    rng = np.random.default_rng(0)
    M, N, K = 48, 64, 80
    A = omlp.to_bf16(rng.standard_normal((M, K)))
    B = omlp.to_bf16(rng.standard_normal((K, N)))
    ref = A @ B
    Ad, Bd = _bf16_bits(A), _bf16_bits(B)
    C = torch.zeros((M, N), dtype=torch.float32, device="cuda")
    _lib.call("sk_gemm_bf16", M, N, K, ptr(Ad), K, 1, ptr(Bd), N, 1, ptr(C), N, stream())
    assert np.abs(C.cpu().numpy() - ref).max() <= 1e-5 * np.abs(ref).max()
    # transposed storage of both operands: A given as A^T (K x M), B as B^T (N x K)
    AtD, BtD = _bf16_bits(A.T.copy()), _bf16_bits(B.T.copy())
    C2 = torch.zeros((M, N), dtype=torch.float32, device="cuda")
    _lib.call("sk_gemm_bf16", M, N, K, ptr(AtD), 1, M, ptr(BtD), 1, K, ptr(C2), N, stream())
    assert torch.equal(C, C2)


def _setup(dims, batch, seed):
    from simkit_amd.mlp import MLP
    net = MLP(dims, max_batch=256, seed=seed)
    rng = np.random.default_rng(seed + 100)
    X = rng.standard_normal((batch, dims[0])).astype(np.float32)
    T = rng.standard_normal((batch, dims[-1])).astype(np.float32)
    return net, X, T


@pytest.mark.parametrize("dims,batch", [([140, 256, 256, 256, 64], 200), ([7, 32, 5], 3)])
def test_forward_and_gradients_match_oracle(dims, batch):
    net, X, T = _setup(dims, batch, 1)
    layers = omlp.unflatten(net.get_params(), dims)
    Y = net.forward(torch.as_tensor(X).cuda()).cpu().numpy()
    Yo, _, _ = omlp.forward(layers, X)
    assert _rel(Y, Yo) < 1e-3
    loss = net.backward(torch.as_tensor(X).cuda(), torch.as_tensor(T).cuda())
    lo, go = omlp.loss_and_grad(layers, X, T)
    assert abs(loss - lo) <= 1e-3 * lo
    g = net.get_grad()
    go_flat = omlp.flatten(go)
    o = 0
    for (Wo, bo) in go:
        assert _rel(g[o:o + Wo.size], Wo.reshape(-1)) < 1e-3
        o += Wo.size
        assert _rel(g[o:o + bo.size], bo) < 1e-3
        o += bo.size
    assert o == go_flat.size == net.n_params


def test_adam_step_matches_oracle():
    dims = [16, 32, 8]
    net, X, T = _setup(dims, 40, 2)
    p0 = net.get_params().astype(np.float64)
    net.backward(torch.as_tensor(X).cuda(), torch.as_tensor(T).cuda(), want_loss=False)
    g = net.get_grad().astype(np.float64)
    net.adam_step(lr=1e-2)
    p1, m, v = omlp.adam(p0, g, 0.0, 0.0, 1, lr=1e-2)
    assert np.abs(net.get_params() - p1).max() < 1e-6


def test_training_is_deterministic_and_fits():
    dims = [8, 64, 64, 4]
    rng = np.random.default_rng(3)
    X = rng.standard_normal((2048, 8)).astype(np.float32)
    Wt = rng.standard_normal((8, 4)).astype(np.float32)
    T = np.tanh(0.5 * X @ Wt).astype(np.float32)
    Xd, Td = torch.as_tensor(X).cuda(), torch.as_tensor(T).cuda()
    from simkit_amd.mlp import MLP
    runs = []
    for _ in range(2):
        net = MLP(dims, max_batch=256, seed=4)
        losses = net.fit(Xd, Td, epochs=30, batch=128, lr=3e-3, seed=5)
        runs.append((losses, net.get_params()))
    assert runs[0][0] == runs[1][0] and np.array_equal(runs[0][1], runs[1][1])
    losses = runs[0][0]
    # the same schedule in the oracle (FP64 sums, BF16 rounding) reaches 0.0009 from 0.18
    assert losses[-1] < 0.05 * losses[0], losses


def test_normalized_predict_matches_forward_and_oracle():
    """sk_mlp_predict (FP64 in and out, input and output normalization in the library) against the FP32
    forward of the same normalized input (the same BF16 input rows, so the network's output bit for bit
    before the FP64 de-normalization) and against the oracle MLP."""
    from simkit_amd import _lib
    from simkit_amd._device import ptr, stream
    dims = [140, 256, 256, 256, 64]
    net, _, _ = _setup(dims, 4, 3)
    rng = np.random.default_rng(7)
    X = rng.standard_normal((5, dims[0])) * 3.0 + 1.0
    mu_i, sd_i = rng.standard_normal(dims[0]), rng.uniform(0.5, 4.0, dims[0])
    mu_o, sd_o = rng.standard_normal(dims[-1]), rng.uniform(1e-3, 2.0, dims[-1])
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float64)).cuda()
    Xd, mi, si, mo, so = d(X), d(mu_i), d(sd_i), d(mu_o), d(sd_o)
    Y = torch.empty((5, dims[-1]), dtype=torch.float64, device="cuda")
    _lib.call("sk_mlp_predict", net._h, ptr(Xd), 5, ptr(mi), ptr(si), ptr(mo), ptr(so), ptr(Y), stream())
    Xn = ((X - mu_i) / sd_i).astype(np.float32)
    Yf = net.forward(torch.as_tensor(Xn).cuda()).cpu().numpy().astype(np.float64)
    Yp = Y.cpu().numpy()
    # the network outputs are the same FP32 values; only the FP64 affine map can round differently
    assert np.abs(Yp - (Yf * sd_o + mu_o)).max() <= 1e-15 * np.abs(Yp).max()
    Yo, _, _ = omlp.forward(omlp.unflatten(net.get_params(), dims), Xn)
    assert _rel((Yp - mu_o) / sd_o, Yo) < 1e-3
