"""Message-passing corrector network on the mesh graph, trained and run on the GPU.

No reference counterpart: the reference defines no neural network.  The network is

    h0      = relu(x W_in + b_in)
    a_l     = [h_l | mean over the cell's interior face neighbours of h_l]
    h_{l+1} = h_l + relu(a_l W_l + b_l)                        l = 0 .. layers-1
    y       = h_L W_out + b_out

with bfloat16 weights and activations and fp32 accumulation on the MFMA units
(csrc/mpnn.hip); for hidden width 128 each message layer is one fused kernel (the
neighbour mean is formed while the GEMM stages its operand tile).  The graph is the
mesh's face adjacency (``FvmMeshDev.cell_nbr``), so a batch of meshes concatenated
block-diagonally is one graph.  Forward, backward and the
Adam update are the library's own kernels; fp32 master weights are kept beside their
bf16 copies.
"""
from __future__ import annotations

import math

import torch

from fvm_solver_amd import _native as nat

_BF = torch.bfloat16


def gemm(A, a_rs, a_cs, B, b_rs, b_cs, M, N, K, *, bias=None, Cf=None, c_rs=None, Cb=None, relu=False,
         res=None, splits=1, part=None):
    """z = A.B + bias on MFMA; Cf <- z (fp32), Cb <- res + relu(z) (bf16).  A, B are bf16
    tensors addressed through (row, col) strides in elements."""
    if c_rs is None:
        c_rs = N
    if splits > 1 and (part is None or part.numel() < splits * M * N):
        raise ValueError("split-K needs a partial buffer of splits * M * N floats")
    nat.check(nat.lib().fvm_gemm_bf16(nat.ptr(A), a_rs, a_cs, nat.ptr(B), b_rs, b_cs, M, N, K, splits,
                                      nat.ptr(bias), nat.ptr(Cf), c_rs, nat.ptr(Cb), int(relu), nat.ptr(res),
                                      nat.ptr(part), nat.stream_ptr(A.device)), "fvm_gemm_bf16")


def split_count(M_out: int, N_out: int, K: int, target_groups: int = 2048) -> int:
    """K splits for a weight-gradient GEMM (few output tiles, long K): enough
    workgroups to fill the chip, at least 256 rows of K per split."""
    tiles = math.ceil(M_out / 64) * math.ceil(N_out / 64)
    return max(1, min(math.ceil(target_groups / tiles), K // 256, 1024))


def init_weights(kx: int, hidden: int, n_out: int, layers: int = 3, seed: int = 0) -> dict:
    """Uniform(+-gain sqrt(6 / fan_in)) weights (gain 1 for the encoder, 0.5 for the
    message layers, 0.1 for the decoder), zero biases; fp32 on the CPU."""
    g = torch.Generator().manual_seed(seed)

    def u(fan_in, shape, gain):
        a = gain * math.sqrt(6.0 / fan_in)
        return (torch.rand(shape, generator=g, dtype=torch.float64) * 2 - 1).mul_(a).float()

    p = {"W_in": u(kx, (kx, hidden), 1.0), "b_in": torch.zeros(hidden)}
    for l in range(layers):
        p[f"W_{l}"] = u(2 * hidden, (2 * hidden, hidden), 0.5)
        p[f"b_{l}"] = torch.zeros(hidden)
    p["W_out"] = u(hidden, (hidden, n_out), 0.1)
    p["b_out"] = torch.zeros(n_out)
    return p


class MPNN:
    """The corrector network; all tensors on one GPU."""

    def __init__(self, kx: int, n_out: int, hidden: int = 128, layers: int = 3, device="cuda", seed: int = 0,
                 weights: dict | None = None):
        if hidden % 8:
            raise ValueError("hidden width must be a multiple of 8")
        self.kx, self.n_out, self.H, self.L = kx, n_out, hidden, layers
        self.device = torch.device(device)
        nat.require_device(torch.empty(0, device=self.device), "MPNN device")
        w = weights if weights is not None else init_weights(kx, hidden, n_out, layers, seed)
        self.names = list(w)
        self.w = {k: v.detach().to(self.device, torch.float32).contiguous() for k, v in w.items()}
        self.wb = {k: v.to(_BF) for k, v in self.w.items() if k.startswith("W")}
        # every weight is also kept transposed ([out][in], bf16): K-contiguous B operands
        # stage with 16-byte loads, and the fused layer reads W_l^T
        self.wt = {k: torch.empty(v.shape[1], v.shape[0], dtype=_BF, device=self.device)
                   for k, v in self.w.items() if k.startswith("W")}
        self._refresh_transposed()
        self.g = {k: torch.zeros_like(v) for k, v in self.w.items()}
        self.m = {k: torch.zeros_like(v) for k, v in self.w.items()}
        self.v = {k: torch.zeros_like(v) for k, v in self.w.items()}
        self.step_count = 0
        self._ws = {}
        self._cs_part = torch.empty(256 * 2 * hidden, device=self.device)

    # ------------------------------------------------------------------ parameters
    def state_dict(self) -> dict:
        return {k: v.detach().cpu().clone() for k, v in self.w.items()}

    def load_state_dict(self, d: dict):
        for k in self.names:
            self.w[k].copy_(d[k].to(self.device, torch.float32))
            if k in self.wb:
                self.wb[k].copy_(self.w[k].to(_BF))
        self._refresh_transposed()

    def _refresh_transposed(self):
        for k, t in self.wt.items():
            r, c = self.w[k].shape
            nat.check(nat.lib().fvm_transpose_bf16(nat.ptr(self.w[k]), r, c, nat.ptr(t), nat.stream_ptr(self.device)),
                      "fvm_transpose_bf16")

    # ------------------------------------------------------------------ workspace
    def _work(self, M: int, train: bool):
        key = (M, train)
        if key in self._ws:
            return self._ws[key]
        H, L, dev = self.H, self.L, self.device
        ws = {"Y": torch.empty(M, self.n_out, device=dev)}
        if train:
            ws["Z"] = [torch.empty(M, H, device=dev) for _ in range(L + 1)]
            ws["Hs"] = [torch.empty(M, H, dtype=_BF, device=dev) for _ in range(L + 1)]
            ws["A"] = [torch.empty(M, 2 * H, dtype=_BF, device=dev) for _ in range(L)]
            ws["dYb"] = torch.empty(M, self.n_out, dtype=_BF, device=dev)
            ws["dY"] = torch.empty(M, self.n_out, device=dev)
            ws["dH"] = [torch.empty(M, H, device=dev) for _ in range(2)]
            ws["dA"] = torch.empty(M, 2 * H, device=dev)
            ws["dZb"] = torch.empty(M, H, dtype=_BF, device=dev)
            ws["dZf"] = torch.empty(M, H, device=dev)
            ws["loss_part"] = torch.empty(256, dtype=torch.float64, device=dev)
            sp = [split_count(r, c, M) for r, c in ((self.kx, H), (2 * H, H), (H, self.n_out))]
            ws["splits"] = sp
            ws["part"] = torch.empty(max(s * r * c for s, (r, c) in
                                         zip(sp, ((self.kx, H), (2 * H, H), (H, self.n_out)))), device=dev)
        else:
            ws["Hs"] = [torch.empty(M, H, dtype=_BF, device=dev) for _ in range(2)]
            if H != 128:
                ws["A"] = [torch.empty(M, 2 * H, dtype=_BF, device=dev)]
        self._ws[key] = ws
        return ws

    # ------------------------------------------------------------------ forward
    def forward(self, X: torch.Tensor, nbr: torch.Tensor, train: bool = False) -> torch.Tensor:
        """X: bf16 [M, kx] node features; nbr: int32 [NL, M] face neighbours (>= M marks a
        boundary facet).  Returns y (fp32 [M, n_out], a workspace view overwritten by the
        next call with the same M)."""
        M = X.shape[0]
        if X.dtype != _BF or X.shape[1] != self.kx or not X.is_contiguous():
            raise ValueError(f"X must be contiguous bf16 [M, {self.kx}]")
        if nbr.dtype != torch.int32 or nbr.shape[1] != M:
            raise ValueError("nbr must be int32 [NL, M]")
        H, L = self.H, self.L
        ws = self._work(M, train)
        lib, sp = nat.lib(), nat.stream_ptr(self.device)
        Hs = ws["Hs"]
        if not train and H == 128 and self.kx == 32 and self.n_out == 8 and L >= 1:
            return self._forward_inference(X, nbr, ws)
        gemm(X, self.kx, 1, self.wt["W_in"], 1, self.kx, M, H, self.kx, bias=self.w["b_in"],
             Cf=ws["Z"][0] if train else None, Cb=Hs[0], relu=True)
        for l in range(L):
            h_in = Hs[l] if train else Hs[l % 2]
            h_out = Hs[l + 1] if train else Hs[(l + 1) % 2]
            if H == 128:
                nat.check(lib.fvm_mp_layer(nat.ptr(h_in), H, nat.ptr(nbr), M, nbr.shape[0], nat.ptr(self.wt[f"W_{l}"]),
                                           nat.ptr(self.w[f"b_{l}"]), nat.ptr(h_out),
                                           nat.ptr(ws["Z"][l + 1]) if train else None,
                                           nat.ptr(ws["A"][l]) if train else None, sp), "fvm_mp_layer")
                continue
            A = ws["A"][l] if train else ws["A"][0]
            nat.check(lib.fvm_mp_aggregate(nat.ptr(h_in), H, nat.ptr(nbr), M, nbr.shape[0], nat.ptr(A), sp),
                      "fvm_mp_aggregate")
            gemm(A, 2 * H, 1, self.wb[f"W_{l}"], H, 1, M, H, 2 * H, bias=self.w[f"b_{l}"],
                 Cf=ws["Z"][l + 1] if train else None, Cb=h_out, relu=True, res=h_in)
        h_last = Hs[L] if train else Hs[L % 2]
        gemm(h_last, H, 1, self.wt["W_out"], 1, H, M, self.n_out, H, bias=self.w["b_out"], Cf=ws["Y"])
        self._last = (X, nbr, ws)
        return ws["Y"]

    def _forward_inference(self, X, nbr, ws):
        """Inference forward with the fused kernels: encoder, L - 1 message layers, and
        the last layer with the decoder in its epilogue (h_L is never stored)."""
        M, H, L = X.shape[0], self.H, self.L
        lib, sp = nat.lib(), nat.stream_ptr(self.device)
        Hs = ws["Hs"]
        nat.check(lib.fvm_mp_encode(nat.ptr(X), M, self.kx, nat.ptr(self.wt["W_in"]), nat.ptr(self.w["b_in"]), H,
                                    nat.ptr(Hs[0]), sp), "fvm_mp_encode")
        for l in range(L):
            last = l == L - 1
            nat.check(lib.fvm_mp_layer_inf(nat.ptr(Hs[l % 2]), H, nat.ptr(nbr), M, nbr.shape[0],
                                           nat.ptr(self.wt[f"W_{l}"]), nat.ptr(self.w[f"b_{l}"]),
                                           None if last else nat.ptr(Hs[(l + 1) % 2]),
                                           nat.ptr(self.wt["W_out"]) if last else None,
                                           nat.ptr(self.w["b_out"]) if last else None, self.n_out if last else 0,
                                           nat.ptr(ws["Y"]) if last else None, sp), "fvm_mp_layer_inf")
        self._last = (X, nbr, ws)
        return ws["Y"]

    # ------------------------------------------------------------------ backward
    def loss(self, target: torch.Tensor) -> torch.Tensor:
        """MSE of the last training forward against target [M, n_out]; writes dL/dy and
        returns the loss as a device scalar (fp64)."""
        X, nbr, ws = self._last
        M = X.shape[0]
        nat.check(nat.lib().fvm_mse_loss(nat.ptr(ws["Y"]), self.n_out, nat.ptr(target), M, self.n_out,
                                         nat.ptr(ws["dY"]), nat.ptr(ws["loss_part"]), nat.stream_ptr(self.device)),
                  "fvm_mse_loss")
        return ws["loss_part"].sum() / (M * self.n_out)

    def backward(self, dY: torch.Tensor | None = None):
        """Gradients of the last training forward into self.g (dY defaults to the one
        ``loss`` wrote)."""
        X, nbr, ws = self._last
        if "Z" not in ws:
            raise RuntimeError("backward needs a forward with train=True")
        M, H, L, n_out = X.shape[0], self.H, self.L, self.n_out
        lib, sp = nat.lib(), nat.stream_ptr(self.device)
        dY = ws["dY"] if dY is None else dY.contiguous()
        s_in, s_mid, s_out = ws["splits"]
        part = ws["part"]
        nat.check(lib.fvm_cast_f32_bf16(nat.ptr(dY), nat.ptr(ws["dYb"]), M * n_out, sp), "fvm_cast_f32_bf16")
        dYb = ws["dYb"]
        gemm(ws["Hs"][L], 1, H, dYb, n_out, 1, H, n_out, M, Cf=self.g["W_out"], splits=s_out, part=part)
        self._colsum(dY, M, n_out, self.g["b_out"])
        dH = ws["dH"][0]
        gemm(dYb, n_out, 1, self.wb["W_out"], 1, n_out, M, H, n_out, Cf=dH)
        for l in reversed(range(L)):
            nat.check(lib.fvm_relu_bwd(nat.ptr(dH), nat.ptr(ws["Z"][l + 1]), nat.ptr(ws["dZb"]), nat.ptr(ws["dZf"]),
                                       M * H, sp), "fvm_relu_bwd")
            gemm(ws["A"][l], 1, 2 * H, ws["dZb"], H, 1, 2 * H, H, M, Cf=self.g[f"W_{l}"], splits=s_mid, part=part)
            self._colsum(ws["dZf"], M, H, self.g[f"b_{l}"])
            gemm(ws["dZb"], H, 1, self.wb[f"W_{l}"], 1, H, M, 2 * H, H, Cf=ws["dA"])  # B(k,j) = W_l[j][k]
            dH_new = ws["dH"][1] if dH is ws["dH"][0] else ws["dH"][0]
            nat.check(lib.fvm_mp_aggregate_bwd(nat.ptr(ws["dA"]), H, nat.ptr(nbr), M, nbr.shape[0], nat.ptr(dH),
                                               nat.ptr(dH_new), sp), "fvm_mp_aggregate_bwd")
            dH = dH_new
        nat.check(lib.fvm_relu_bwd(nat.ptr(dH), nat.ptr(ws["Z"][0]), nat.ptr(ws["dZb"]), nat.ptr(ws["dZf"]), M * H,
                                   sp), "fvm_relu_bwd")
        gemm(X, 1, self.kx, ws["dZb"], H, 1, self.kx, H, M, Cf=self.g["W_in"], splits=s_in, part=part)
        self._colsum(ws["dZf"], M, H, self.g["b_in"])

    def _colsum(self, Xf, M, N, out):
        nat.check(nat.lib().fvm_colsum_f32(nat.ptr(Xf), M, N, nat.ptr(out), nat.ptr(self._cs_part),
                                           nat.stream_ptr(self.device)), "fvm_colsum_f32")

    # ------------------------------------------------------------------ optimiser
    def adam(self, lr: float = 1e-3, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
        self.step_count += 1
        lib, sp = nat.lib(), nat.stream_ptr(self.device)
        for k in self.names:
            nat.check(lib.fvm_adam_f32(nat.ptr(self.w[k]), nat.ptr(self.g[k]), nat.ptr(self.m[k]), nat.ptr(self.v[k]),
                                       nat.ptr(self.wb.get(k)), self.w[k].numel(), lr, b1, b2, eps, self.step_count,
                                       sp), "fvm_adam_f32")
        self._refresh_transposed()

    def train_step(self, X, nbr, target, lr=1e-3) -> torch.Tensor:
        self.forward(X, nbr, train=True)
        loss = self.loss(target)
        self.backward()
        self.adam(lr)
        return loss
