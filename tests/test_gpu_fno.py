"""GPU tests of the learned coarse-slice correction (bf16 MFMA kernels).

The GEMM kernel is checked against numpy on the same bf16-rounded operands
(float32 accumulation: tight tolerance) over every operand layout the
network uses; the network's forward pass against the oracle's bf16-emulating
forward (tight) and float64 forward (bf16 tolerance); its backward pass
against the oracle's exact gradients; the input packing, output assembly and
loss kernels against numpy.  All buffers are sized from the exact extents the
descriptors address.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import fno_oracle as fo

pytestmark = pytest.mark.gpu


def _dev(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a)).to(dtype).to("cuda").contiguous()


def _extent(batch, bs, dims):
    """Number of elements a strided operand spans: 1 + max offset."""
    return 1 + (batch - 1) * bs + sum((n - 1) * s for n, s in dims)


def _operand(rng, batch, bs, dims, dtype):
    """Random values over the operand's whole extent, rounded to bf16 when
    the operand is bf16 (so the numpy reference sees the same numbers)."""
    flat = rng.normal(size=_extent(batch, bs, dims)).astype(np.float32)
    t = _dev(flat, torch.bfloat16 if dtype == 1 else torch.float32)
    return t, t.to(torch.float32).cpu().numpy().astype(np.float64)


def _gather(flat, batch, bs, s0, s1, n0, n1):
    idx = (np.arange(batch)[:, None, None] * bs + np.arange(n0)[None, :, None] * s0 +
           np.arange(n1)[None, None, :] * s1)
    return flat[idx]


def _bf(x):
    return fo.bf16_round(x)


def _run_gemm(hip, d):
    hip.call("emsim_hip_gemm", ctypes.addressof(d),
             ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()


CASES = [
    # m, n, k, batch, a layout (sm, sk), b layout (sk, sn), a_type, b_type, c_type
    dict(m=100, n=70, k=50, batch=3, a=("row",), b=("row",), at=1, bt=0, ct=0),
    dict(m=64, n=32, k=512, batch=5, a=("row",), b=("row",), at=1, bt=1, ct=0),
    dict(m=33, n=130, k=77, batch=2, a=("col",), b=("col",), at=0, bt=1, ct=1),
    dict(m=257, n=2, k=32, batch=1, a=("row",), b=("row",), at=1, bt=1, ct=0),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['m']}x{c['n']}x{c['k']}b{c['batch']}")
def test_gemm_layouts(hip, case):
    rng = np.random.default_rng(1)
    m, n, k, B = case["m"], case["n"], case["k"], case["batch"]
    a_sm, a_sk = (k, 1) if case["a"][0] == "row" else (1, m)
    b_sk, b_sn = (n, 1) if case["b"][0] == "row" else (1, k)
    a_bs, b_bs, c_bs = m * k + 3, k * n + 5, m * n + 7
    A, a_np = _operand(rng, B, a_bs, [(m, a_sm), (k, a_sk)], case["at"])
    Bm, b_np = _operand(rng, B, b_bs, [(k, b_sk), (n, b_sn)], case["bt"])
    c_ext = _extent(B, c_bs, [(m, n), (n, 1)])
    Cm = torch.zeros(c_ext, dtype=torch.bfloat16 if case["ct"] else torch.float32, device="cuda")
    d = hip.GemmDesc(m=m, n=n, k=k, batch=B,
                     a=A.data_ptr(), a_bs=a_bs, a_sm=a_sm, a_sk=a_sk, a_type=case["at"],
                     b=Bm.data_ptr(), b_bs=b_bs, b_sk=b_sk, b_sn=b_sn, b_type=case["bt"],
                     c=Cm.data_ptr(), c_bs=c_bs, c_sm=n, c_sn=1, c_type=case["ct"])
    _run_gemm(hip, d)
    a = _bf(_gather(a_np, B, a_bs, a_sm, a_sk, m, k))
    b = _bf(_gather(b_np, B, b_bs, b_sk, b_sn, k, n))
    ref = np.einsum("bik,bkj->bij", a, b)
    got = _gather(Cm.to(torch.float32).cpu().numpy().astype(np.float64), B, c_bs, n, 1, m, n)
    tol = 8e-3 if case["ct"] else 1e-5
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()


def test_gemm_second_pair_bias_gelu_beta(hip):
    rng = np.random.default_rng(2)
    m, n, k, k2, B = 96, 32, 32, 32, 4
    A, a_np = _operand(rng, B, m * k, [(m, k), (k, 1)], 1)
    Bm, b_np = _operand(rng, 1, 0, [(k, n), (n, 1)], 1)
    A2, a2_np = _operand(rng, 1, 0, [(m, k2), (k2, 1)], 1)
    B2, b2_np = _operand(rng, B, k2 * n, [(k2, n), (n, 1)], 0)
    bias = _dev(rng.normal(size=n).astype(np.float32))
    Z = torch.zeros(B * m * n, dtype=torch.bfloat16, device="cuda")
    Cm = torch.zeros(B * m * n, dtype=torch.float32, device="cuda")
    d = hip.GemmDesc(m=m, n=n, k=k, batch=B,
                     a=A.data_ptr(), a_bs=m * k, a_sm=k, a_sk=1, a_type=1,
                     b=Bm.data_ptr(), b_bs=0, b_sk=n, b_sn=1, b_type=1,
                     k2=k2, a2=A2.data_ptr(), a2_bs=0, a2_sm=k2, a2_sk=1, a2_type=1,
                     b2=B2.data_ptr(), b2_bs=k2 * n, b2_sk=n, b2_sn=1, b2_type=0,
                     c=Cm.data_ptr(), c_bs=m * n, c_sm=n, c_sn=1, c_type=0,
                     bias=bias.data_ptr(), act=1, z=Z.data_ptr(), z_type=1)
    _run_gemm(hip, d)
    a = _bf(a_np.reshape(B, m, k))
    b = _bf(b_np.reshape(k, n))
    a2 = _bf(a2_np.reshape(m, k2))
    b2 = _bf(b2_np.reshape(B, k2, n))
    pre = a @ b + a2[None] @ b2 + bias.cpu().numpy()
    got = Cm.cpu().numpy().reshape(B, m, n)
    assert np.abs(got - fo.gelu(pre)).max() <= 1e-5 * np.abs(pre).max()
    z = Z.to(torch.float32).cpu().numpy().reshape(B, m, n)
    assert np.abs(z - pre).max() <= 8e-3 * np.abs(pre).max()
    # beta accumulates into the existing float32 output (no bias / act)
    prev = got.copy()
    d2 = hip.GemmDesc(m=m, n=n, k=k, batch=B,
                      a=A.data_ptr(), a_bs=m * k, a_sm=k, a_sk=1, a_type=1,
                      b=Bm.data_ptr(), b_bs=0, b_sk=n, b_sn=1, b_type=1,
                      c=Cm.data_ptr(), c_bs=m * n, c_sm=n, c_sn=1, c_type=0, beta=1.0)
    _run_gemm(hip, d2)
    got2 = Cm.cpu().numpy().reshape(B, m, n)
    assert np.abs(got2 - (prev + a @ b)).max() <= 1e-5 * np.abs(prev + a @ b).max()


def test_gemm_split_k(hip):
    rng = np.random.default_rng(3)
    m, n, K, chunk = 32, 32, 5000, 512
    A, a_np = _operand(rng, 1, 0, [(m, 1), (K, m)], 1)     # A^T layout: k rows of m
    Bm, b_np = _operand(rng, 1, 0, [(K, n), (n, 1)], 0)
    splits = (K + chunk - 1) // chunk
    part = torch.zeros(splits * m * n, dtype=torch.float32, device="cuda")
    d = hip.GemmDesc(m=m, n=n, k=K, batch=1,
                     a=A.data_ptr(), a_sm=1, a_sk=m, a_type=1,
                     b=Bm.data_ptr(), b_sk=n, b_sn=1, b_type=0,
                     c=part.data_ptr(), c_sm=n, c_sn=1, c_type=0,
                     split_k=chunk, c_split_stride=m * n)
    _run_gemm(hip, d)
    a = _bf(a_np.reshape(K, m)).T
    b = _bf(b_np.reshape(K, n))
    got = part.cpu().numpy().reshape(splits, m, n).sum(axis=0)
    ref = a @ b
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()


def test_gemm_rejects_bad_descriptors(hip):
    d = hip.GemmDesc(m=0, n=4, k=4, batch=1)
    with pytest.raises(ValueError):
        hip.call("emsim_hip_gemm", ctypes.addressof(d), None)


# ---------------------------------------------------------------------------
# the network
# ---------------------------------------------------------------------------
NN, CC, MM, LL, BB, CIN = 64, 32, 8, 3, 2, 6


@pytest.fixture(scope="module")
def net(hip):
    from emsim_amd.learned import FNO
    f = FNO(NN, CIN, CC, MM, LL)
    p = fo.random_params(CIN, CC, MM, LL, seed=7)
    f.set_params(p)
    rng = np.random.default_rng(8)
    x = np.zeros((BB, NN, NN, f.cin_pad))
    x[..., :CIN] = rng.normal(size=(BB, NN, NN, CIN))
    xt = _dev(x, torch.bfloat16)
    x_bf = xt.to(torch.float32).cpu().numpy().astype(np.float64)
    return f, f.get_params(), xt, x_bf


def test_param_roundtrip(net):
    f, p, _, _ = net
    p2 = fo.random_params(CIN, CC, MM, LL, seed=7)
    flat = f.flatten(p2)
    assert np.array_equal(f.flatten(f.unflatten(flat)), flat)
    assert np.array_equal(f.params_flat(), flat)


def test_forward_matches_oracle(net):
    f, p, xt, x_bf = net
    got = f.forward(xt).cpu().numpy().astype(np.float64)
    emu, _ = fo.forward(p, x_bf, MM, emulate_bf16=True)
    ref, _ = fo.forward(p, x_bf, MM)
    rel_emu = np.linalg.norm(got - emu) / np.linalg.norm(emu)
    rel_f64 = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"forward rel L2: vs bf16-emulating oracle {rel_emu:.2e}, vs float64 oracle {rel_f64:.2e}")
    assert rel_emu < 1e-3
    assert rel_f64 < 1.5e-2
    # training mode computes the same output
    got_t = f.forward(xt, training=True).cpu().numpy()
    assert np.array_equal(got_t, got.astype(np.float32))


def test_fused_rows_equal_generic_path(hip, net):
    """The fused row kernels and the all-GEMM sequence compute the same network."""
    f, _, xt, _ = net
    fused = f.forward(xt).cpu().numpy()
    hip.call("emsim_hip_fno_force_generic", 1)
    try:
        gen = f.forward(xt).cpu().numpy()
    finally:
        hip.call("emsim_hip_fno_force_generic", 0)
    rel = np.linalg.norm(fused - gen) / np.linalg.norm(gen)
    print(f"fused vs generic forward rel L2 {rel:.2e}")
    assert rel < 1e-3


def test_full_width_modes_match_oracle_and_generic(hip):
    """16 modes (2 M1 = 64 x-transform rows, M2 C = 512 columns: the shapes of
    the 512^2 network) through the fused path's x-transform kernels, against
    the oracle and against the all-GEMM sequence."""
    from emsim_amd.learned import FNO
    m, L = 16, 2
    f = FNO(NN, CIN, CC, m, L)
    p = fo.random_params(CIN, CC, m, L, seed=21)
    f.set_params(p)
    rng = np.random.default_rng(22)
    x = np.zeros((BB, NN, NN, f.cin_pad))
    x[..., :CIN] = rng.normal(size=(BB, NN, NN, CIN))
    xt = _dev(x, torch.bfloat16)
    x_bf = xt.to(torch.float32).cpu().numpy().astype(np.float64)
    got = f.forward(xt).cpu().numpy()
    emu, _ = fo.forward(f.get_params(), x_bf, m, emulate_bf16=True)
    rel_emu = np.linalg.norm(got - emu) / np.linalg.norm(emu)
    hip.call("emsim_hip_fno_force_generic", 1)
    try:
        gen = f.forward(xt).cpu().numpy()
    finally:
        hip.call("emsim_hip_fno_force_generic", 0)
    rel_gen = np.linalg.norm(got - gen) / np.linalg.norm(gen)
    print(f"16 modes: vs bf16-emulating oracle {rel_emu:.2e}, vs generic {rel_gen:.2e}")
    assert rel_emu < 1e-3
    assert rel_gen < 1e-3


def test_infer_equals_pack_then_forward(net):
    """Packing fused into the lifting gives the same output bit for bit."""
    f, _, _, _ = net
    rng = np.random.default_rng(12)
    ns = CIN - 2
    wave = _dev((rng.normal(size=(BB, NN, NN)) + 1j * rng.normal(size=(BB, NN, NN)))
                .astype(np.complex64), torch.complex64)
    sl = _dev(rng.normal(size=(BB, ns, NN, NN)).astype(np.float32))
    a = f.infer(wave, sl, 0.8, 0.035)
    b = f.forward(f.pack(wave, sl, 0.8, 0.035))
    assert torch.equal(a, b)


def test_backward_matches_oracle(net):
    f, p, xt, x_bf = net
    rng = np.random.default_rng(9)
    f.forward(xt, training=True)
    G = rng.normal(size=(BB, NN, NN, 2))
    f.backward(xt, _dev(G))
    torch.cuda.synchronize()
    g = f.grads()
    _, cache = fo.forward(p, x_bf, MM)
    ref = fo.backward(p, cache, G, MM)
    for key in ["lift_w", "lift_b", "proj_w", "proj_b", "w", "b", "r"]:
        pairs = [(g[key], ref[key])] if key not in ("w", "b", "r") else list(zip(g[key], ref[key]))
        for i, (a, b) in enumerate(pairs):
            if key == "lift_w":
                a, b = a[:CIN], b[:CIN]
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            print(f"grad {key}[{i}] rel L2 {rel:.2e}")
            assert rel < 2e-2, (key, i, rel)


def test_pack_apply_loss_kernels(hip, net):
    f, _, _, _ = net
    rng = np.random.default_rng(10)
    ns = CIN - 2
    wave = (rng.normal(size=(BB, NN, NN)) + 1j * rng.normal(size=(BB, NN, NN))).astype(np.complex64)
    sl = rng.normal(size=(BB, ns, NN, NN)).astype(np.float32)
    psi0, sigma = 0.75, 0.035
    x = f.pack(_dev(wave, torch.complex64), _dev(sl), psi0, sigma)
    xr = x.to(torch.float32).cpu().numpy()
    exp = np.concatenate([wave.real[..., None] / psi0, wave.imag[..., None] / psi0,
                          np.moveaxis(sl, 1, -1) * sigma], axis=-1)
    ref = fo.bf16_round(exp)
    assert (np.abs(xr[..., :CIN] - ref) <= 2.0 ** -7 * np.abs(ref)).all()   # 1 bf16 ulp
    assert not xr[..., CIN:].any()
    out = rng.normal(size=(BB, NN, NN, 2)).astype(np.float32)
    w = f.apply(_dev(wave, torch.complex64), _dev(out), psi0).cpu().numpy()
    assert np.allclose(w, wave + psi0 * (out[..., 0] + 1j * out[..., 1]), rtol=1e-6, atol=1e-6)
    target = (wave + 0.1 * rng.normal(size=wave.shape)).astype(np.complex64)
    loss, dout = f.loss(_dev(w, torch.complex64), _dev(target, torch.complex64), psi0)
    loss = loss.cpu().numpy()
    e = (np.abs(w - target) ** 2).sum(axis=(1, 2))
    t = (np.abs(target) ** 2).sum(axis=(1, 2))
    a = (np.abs(w) ** 2).sum(axis=(1, 2))
    assert np.allclose(loss[:, 0], e / t, rtol=1e-4)
    assert np.allclose(loss[:, 1], t, rtol=1e-4)
    assert np.allclose(loss[:, 2], a, rtol=1e-4)
    k = (2 * psi0 / BB) / t
    dref = np.stack([(w - target).real, (w - target).imag], -1) * k[:, None, None, None]
    assert np.allclose(dout.cpu().numpy(), dref, rtol=1e-4, atol=1e-7)


def test_training_reduces_loss_on_fixed_target(hip):
    """Adam steps on one fixed (input, target) pair lower the loss; the
    target (coarse + 0.1 x_0) is a pointwise linear function of the input."""
    from emsim_amd.learned import FNO
    f = FNO(NN, CIN, CC, MM, 2)
    f.set_params(f.init_params(seed=3))
    rng = np.random.default_rng(11)
    x = np.zeros((BB, NN, NN, f.cin_pad))
    x[..., :CIN] = rng.normal(size=(BB, NN, NN, CIN))
    xt = _dev(x, torch.bfloat16)
    x0 = xt[..., 0].to(torch.float32).cpu().numpy()
    coarse = _dev(np.ones((BB, NN, NN), np.complex64), torch.complex64)
    target = _dev((1.0 + 0.1 * x0).astype(np.complex64), torch.complex64)
    losses = []
    for step in range(60):
        out = f.forward(xt, training=True)
        pred = f.apply(coarse, out, 1.0)
        loss, dout = f.loss(pred, target, 1.0)
        losses.append(float(loss[:, 0].mean()))
        f.backward(xt, dout)
        f.adam_step(3e-3)
    print("loss", losses[0], "->", losses[-1])
    assert losses[-1] < 0.2 * losses[0]
