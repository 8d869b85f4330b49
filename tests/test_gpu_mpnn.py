"""GPU checks of the corrector's kernels through the C ABI: the MFMA GEMM against an fp64
product, aggregation and its backward, a full forward/backward against the CPU oracle
(float64 autograd), Adam, and that training lowers the loss."""
import numpy as np
import pytest
import torch

from fvm_solver_amd import _native as nat
from fvm_solver_amd.config_fvm import ConfigChannel
from fvm_solver_amd.fvm_stepping.facet_process import DeviceMesh
from fvm_solver_amd.learned.mpnn import MPNN, gemm
from fvm_solver_amd.mesh_gen.channel import channel_mesh_2d
from fvm_solver_amd.mesh_utils.fvm_mesh import mesh_from_tuple
from fvm_solver_amd.mesh_utils.mesh_store import channel_bc_tags
from oracle import mpnn_oracle as MO

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _rel(a, b):
    a = torch.as_tensor(a, dtype=torch.float64).cpu()
    b = torch.as_tensor(b, dtype=torch.float64).cpu()
    return float((a - b).norm() / max(float(b.norm()), 1e-300))


def _graph(n_cells=1500, seed=3):
    pts, cells, (fi, fb), tags = channel_mesh_2d(n_cells, seed=seed)
    mesh = mesh_from_tuple(pts, cells, fi, fb, device="cpu")
    dm = DeviceMesh(mesh, 4, channel_bc_tags(tags, 2), ConfigChannel(device="cpu"), "cpu")
    return np.ascontiguousarray(dm.host["nbr_s"].T)          # [NL, N], >= N on boundary facets


@pytest.mark.parametrize("M,N,K,splits,trans_a,trans_b,epi", [
    (300, 70, 50, 1, False, False, "bias_relu_res"),
    (129, 128, 256, 1, False, True, "plain"),
    (64, 128, 5000, 8, True, False, "plain"),
    (16, 128, 3001, 5, True, False, "bias"),
    (1000, 4, 128, 1, False, False, "bias"),
])
def test_gemm_against_fp64(M, N, K, splits, trans_a, trans_b, epi):
    g = torch.Generator().manual_seed(M + N + K)
    A = torch.randn(M, K, generator=g).to(torch.bfloat16)
    B = torch.randn(K, N, generator=g).to(torch.bfloat16)
    ref = A.double() @ B.double()
    bias = torch.randn(N, generator=g) if epi != "plain" else None
    if bias is not None:
        ref = ref + bias.double()
    res = torch.randn(M, N, generator=g).to(torch.bfloat16) if epi == "bias_relu_res" else None
    # operands stored transposed when asked, addressed through strides
    Ad = (A.t().contiguous() if trans_a else A).to(DEV)
    Bd = (B.t().contiguous() if trans_b else B).to(DEV)
    a_rs, a_cs = (1, M) if trans_a else (K, 1)
    b_rs, b_cs = (1, K) if trans_b else (N, 1)
    Cf = torch.empty(M, N, device=DEV)
    Cb = torch.empty(M, N, dtype=torch.bfloat16, device=DEV)
    part = torch.empty(splits * M * N, device=DEV) if splits > 1 else None
    gemm(Ad, a_rs, a_cs, Bd, b_rs, b_cs, M, N, K, bias=None if bias is None else bias.to(DEV), Cf=Cf, Cb=Cb,
         relu=epi == "bias_relu_res", res=None if res is None else res.to(DEV), splits=splits, part=part)
    torch.cuda.synchronize()
    err = _rel(Cf, ref)
    print(f"[gpu-test]   gemm {M}x{N}x{K} splits={splits} rel err {err:.2e}")
    assert err < 1e-6 * max(1.0, K ** 0.5)
    outb = torch.relu(ref) + res.double() if res is not None else ref
    assert _rel(Cb.float(), outb) < 8e-3


def test_aggregate_and_backward_against_oracle():
    nbr = _graph()
    NL, N = nbr.shape
    H = 64
    h = torch.randn(N, H).to(torch.bfloat16)
    A = torch.empty(N, 2 * H, dtype=torch.bfloat16, device=DEV)
    nb = torch.from_numpy(nbr.astype(np.int32)).to(DEV)
    lib, sp = nat.lib(), nat.stream_ptr(DEV)
    hd = h.to(DEV)
    nat.check(lib.fvm_mp_aggregate(nat.ptr(hd), H, nat.ptr(nb), N, NL, nat.ptr(A), sp), "agg")
    mean_ref = MO.neighbour_mean(h.double(), nbr)
    torch.cuda.synchronize()
    assert torch.equal(A[:, :H].cpu(), h)
    assert _rel(A[:, H:].float(), mean_ref) < 5e-3
    # backward: <dA, agg(h)> = <agg_bwd(dA), h> for the linear map (adjoint identity)
    dA = torch.randn(N, 2 * H, dtype=torch.float64)
    dh = torch.empty(N, H, device=DEV)
    nat.check(lib.fvm_mp_aggregate_bwd(nat.ptr(dA.float().to(DEV)), H, nat.ptr(nb), N, NL, None, nat.ptr(dh), sp),
              "agg_bwd")
    hx = torch.randn(N, H, dtype=torch.float64)
    lhs = (dA[:, :H] * hx).sum() + (dA[:, H:] * MO.neighbour_mean(hx, nbr)).sum()
    rhs = (dh.cpu().double() * hx).sum()
    torch.cuda.synchronize()
    assert abs(float(lhs - rhs)) < 1e-4 * float(dA.abs().sum() / N)


def _net_and_data(seed=0, n_cells=1500, kx=16, H=128, n_out=4):
    nbr = _graph(n_cells, seed + 3)
    N = nbr.shape[1]
    w = MO.init_params(kx, H, n_out, seed=seed)
    w = {k: v.float().to(torch.bfloat16).double() if k.startswith("W") else v for k, v in w.items()}
    g = torch.Generator().manual_seed(seed + 1)
    x = torch.randn(N, kx, generator=g).to(torch.bfloat16)
    t = torch.randn(N, n_out, generator=g) * 0.1
    net = MPNN(kx, n_out, H, device=DEV, weights={k: v.float() for k, v in w.items()})
    return nbr, w, x, t, net


def test_forward_backward_against_oracle():
    nbr, w, x, t, net = _net_and_data()
    nb = torch.from_numpy(nbr.astype(np.int32)).to(DEV)
    y = net.forward(x.to(DEV), nb, train=True).clone()
    loss = float(net.loss(t.to(DEV)))
    net.backward()
    torch.cuda.synchronize()
    lo, go, yo = MO.loss_and_grads(w, x.double(), t.double(), nbr, bf16_points=True)
    ey = _rel(y, yo)
    print(f"[gpu-test]   forward rel err {ey:.2e}, loss {loss:.6e} vs {lo:.6e}")
    # measured: forward 2.5e-5, gradients <= 7.3e-4 (bf16 storage points mirrored by
    # the oracle; the rest is fp32 accumulation against fp64)
    assert ey < 1e-3
    assert abs(loss - lo) < 1e-3 * lo
    for k in MO.param_names():
        gk = net.g[k].cpu().double()
        e = _rel(gk, go[k])
        cos = float((gk * go[k]).sum() / (gk.norm() * go[k].norm() + 1e-300))
        print(f"[gpu-test]   grad {k}: rel err {e:.2e} cos {cos:.6f}")
        assert e < 5e-3 and cos > 0.99999, k


def test_adam_against_oracle():
    nbr, w, x, t, net = _net_and_data(seed=5)
    ref = {k: (net.w[k].cpu().double(), torch.zeros_like(net.w[k].cpu().double()),
               torch.zeros_like(net.w[k].cpu().double())) for k in net.names}
    for step in range(1, 4):
        for k in net.names:
            net.g[k].copy_(torch.sin(net.w[k] * (step + 1)))
        gs = {k: net.g[k].cpu().double() for k in net.names}
        net.adam(lr=2e-3)
        for k in net.names:
            ref[k] = MO.adam_step(*ref[k][:1], gs[k], ref[k][1], ref[k][2], step, lr=2e-3)
    torch.cuda.synchronize()
    for k in net.names:
        assert _rel(net.w[k], ref[k][0]) < 1e-6, k
        if k in net.wb:
            assert torch.equal(net.wb[k].cpu(), net.w[k].cpu().to(torch.bfloat16))


def test_training_lowers_loss():
    """Fit a teacher network of the same shape from a different seed."""
    nbr, w, x, _, net = _net_and_data(seed=7, n_cells=3000)
    teacher = MO.init_params(16, 128, 4, seed=99)
    t = MO.forward(teacher, x.double(), nbr).float()
    t = (t - t.mean(0)) / t.std(0)
    nb = torch.from_numpy(nbr.astype(np.int32)).to(DEV)
    xd, td = x.to(DEV), t.to(DEV)
    losses = [float(net.train_step(xd, nb, td, lr=2e-3)) for _ in range(300)]
    print(f"[gpu-test]   loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < 0.3 * losses[0]


def test_fused_layer_matches_aggregate_then_gemm():
    """fvm_mp_layer (aggregation fused into the GEMM's operand loads, features x cells
    orientation) equals the two-kernel form: the aggregated operand bit for bit, the
    pre-activation to fp32 accumulation order, the bf16 output to one rounding step."""
    nbr = _graph(2500, 9)
    NL, N = nbr.shape
    H = 128
    g = torch.Generator().manual_seed(4)
    h = torch.randn(N, H, generator=g).to(torch.bfloat16).to(DEV)
    W = (torch.randn(2 * H, H, generator=g) * 0.05).to(DEV)
    b = (torch.randn(H, generator=g) * 0.1).to(DEV)
    nb = torch.from_numpy(nbr.astype(np.int32)).to(DEV)
    lib, sp = nat.lib(), nat.stream_ptr(DEV)
    WT = torch.empty(H, 2 * H, dtype=torch.bfloat16, device=DEV)
    nat.check(lib.fvm_transpose_bf16(nat.ptr(W), 2 * H, H, nat.ptr(WT), sp), "transpose")
    assert torch.equal(WT, W.t().contiguous().to(torch.bfloat16))
    h1 = torch.empty_like(h)
    Z1 = torch.empty(N, H, device=DEV)
    A1 = torch.empty(N, 2 * H, dtype=torch.bfloat16, device=DEV)
    nat.check(lib.fvm_mp_layer(nat.ptr(h), H, nat.ptr(nb), N, NL, nat.ptr(WT), nat.ptr(b), nat.ptr(h1), nat.ptr(Z1),
                               nat.ptr(A1), sp), "mp_layer")
    A2 = torch.empty_like(A1)
    nat.check(lib.fvm_mp_aggregate(nat.ptr(h), H, nat.ptr(nb), N, NL, nat.ptr(A2), sp), "agg")
    h2 = torch.empty_like(h)
    Z2 = torch.empty_like(Z1)
    gemm(A2, 2 * H, 1, W.to(torch.bfloat16), H, 1, N, H, 2 * H, bias=b, Cf=Z2, Cb=h2, relu=True, res=h)
    torch.cuda.synchronize()
    assert torch.equal(A1, A2)
    assert _rel(Z1, Z2) < 1e-6
    diff = (h1.float() - h2.float()).abs()
    assert float((diff > 0).float().mean()) < 1e-3
    assert float(diff.max()) <= float(h2.float().abs().max()) * 2 ** -7


def test_inference_forward_matches_training_forward():
    """The inference kernels (dedicated encoder, layers with permuted W^T rows, decoder
    fused into the last layer) against the training forward (generic GEMMs and the
    training layer) on the same network and inputs: hidden layers to one bf16 rounding
    step, the output to fp32 accumulation order."""
    nbr = _graph(3000, 12)
    NL, N = nbr.shape
    g = torch.Generator().manual_seed(8)
    x = torch.randn(N, 32, generator=g).to(torch.bfloat16).to(DEV)
    net = MPNN(32, 8, 128, device=DEV, seed=3)
    for k in net.names:                       # non-zero biases exercise the permuted bias rows
        if k.startswith("b"):
            net.w[k].copy_(torch.randn(net.w[k].shape, generator=g).to(DEV) * 0.1)
    nb = torch.from_numpy(nbr.astype(np.int32)).to(DEV)
    y_tr = net.forward(x, nb, train=True).clone()
    y_inf = net.forward(x, nb, train=False).clone()
    torch.cuda.synchronize()
    e = _rel(y_inf, y_tr)
    print(f"[gpu-test]   inference vs training forward: rel err {e:.2e}, bitwise equal {torch.equal(y_inf, y_tr)}")
    assert e < 1e-3
    # the encoder alone: equal to the training path's h0 (same MFMA products, same rounding)
    ws = net._work(N, True)
    h0 = torch.empty(N, 128, dtype=torch.bfloat16, device=DEV)
    lib, sp = nat.lib(), nat.stream_ptr(DEV)
    nat.check(lib.fvm_mp_encode(nat.ptr(x), N, 32, nat.ptr(net.wt["W_in"]), nat.ptr(net.w["b_in"]), 128, nat.ptr(h0),
                                sp), "mp_encode")
    net.forward(x, nb, train=True)
    torch.cuda.synchronize()
    d0 = (h0.float() - ws["Hs"][0].float()).abs()
    assert float(d0.max()) <= float(ws["Hs"][0].float().abs().max()) * 2 ** -7
    # one inference layer against the training layer
    h1 = torch.empty_like(h0)
    nat.check(lib.fvm_mp_layer_inf(nat.ptr(ws["Hs"][0]), 128, nat.ptr(nb), N, NL, nat.ptr(net.wt["W_0"]),
                                   nat.ptr(net.w["b_0"]), nat.ptr(h1), None, None, 0, None, sp), "mp_layer_inf")
    torch.cuda.synchronize()
    d1 = (h1.float() - ws["Hs"][1].float()).abs()
    print(f"[gpu-test]   layer: differing entries {float((d1 > 0).float().mean()):.2e}")
    assert float((d1 > 0).float().mean()) < 1e-3
    assert float(d1.max()) <= float(ws["Hs"][1].float().abs().max()) * 2 ** -7
