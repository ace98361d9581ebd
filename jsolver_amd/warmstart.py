"""Learned warm start: a 4-level U-Net (16/32/64/64 channels, 3x3 convolutions)
mapping (log D_xx, log D_yy, lambda, rhs) to an initial phi for solve_2d.

Weights and activations are BF16, accumulation FP32; every convolution
(forward, input gradient, weight gradient) is the implicit-GEMM MFMA kernel
of jsolver_amd/csrc/jsv_unet.hip, as are pooling, upsampling + concatenation,
the ReLU backward and Adam.  torch tensors only hold the memory.

The network sees the (m x m) window of nodes 1..m of the ghost-padded
(m+3)^2 arrays (boundary node 1 and the unknowns 2..m), so every level
halves evenly for m a power of two >= 8.  Its prediction is placed on the
unknowns of an (m+3)^2 phi with a zero rim.  The solver then iterates to its
usual tolerance from that start; the evaluation (warmstart_train.evaluate)
falls back to a zero start for every problem whose guess has
||rhs - L phi0|| > ||rhs||.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib

BF16 = torch.bfloat16

# (name, cin, cout, ksize, relu); the decoder inputs are up(prev) ++ skip
LAYERS = [
    ("e1a", 4, 16, 3, 1), ("e1b", 16, 16, 3, 1),
    ("e2a", 16, 32, 3, 1), ("e2b", 32, 32, 3, 1),
    ("e3a", 32, 64, 3, 1), ("e3b", 64, 64, 3, 1),
    ("b1", 64, 64, 3, 1), ("b2", 64, 64, 3, 1),
    ("d3a", 128, 64, 3, 1), ("d3b", 64, 64, 3, 1),
    ("d2a", 96, 32, 3, 1), ("d2b", 32, 32, 3, 1),
    ("d1a", 48, 16, 3, 1), ("d1b", 16, 16, 3, 1),
    ("out", 16, 4, 1, 0),    # channel 0 is phi; 1-3 stay zero (keeps Cin of its input gradient a multiple of 4)
]


def _kp(k, cin):
    """Weight row length: k*k*cin padded to the 32-deep K step of the MFMA."""
    return int(math.ceil(k * k * cin / 32.0) * 32)


def _sp():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _chk(status):
    if status != 0:
        lib = _lib.load()
        lib.jsv_unet_last_error.restype = ctypes.c_char_p
        raise _lib.JsvError(lib.jsv_unet_last_error().decode())


def _lib_unet():
    lib = _lib.load()
    if not getattr(lib, "_unet_sigs", False):
        V, I, F, I64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_int64
        sig = {
            "jsv_conv_fwd": [V, V, V, V, V, I, I, I, I, I, I, I, I, V],
            "jsv_conv_wgrad": [V, V, V, V, V, I, I, I, I, I, I, I, I, V],
            "jsv_avgpool2": [V, V, I, I, I, I, I, V],
            "jsv_up_concat": [V, V, V, I, I, I, I, I, V],
            "jsv_up_concat_bwd": [V, V, V, I, I, I, I, I, V],
            "jsv_relu_bwd": [V, V, I64, V],
            "jsv_add_bf16": [V, V, I64, V],
            "jsv_f32_to_bf16": [V, V, I64, V],
            "jsv_adam": [V, V, V, V, V, I64, F, F, F, F, I, F, V],
        }
        for k, a in sig.items():
            fn = getattr(lib, k)
            fn.argtypes = a
            fn.restype = ctypes.c_int
        lib._unet_sigs = True
    return lib


class Layer:
    def __init__(self, name, cin, cout, k, relu, dev, gen):
        self.name, self.cin, self.cout, self.k, self.relu = name, cin, cout, k, relu
        self.kp = _kp(k, cin)
        fan_in = k * k * cin
        std = math.sqrt(2.0 / fan_in) if relu else math.sqrt(1.0 / fan_in)
        w = torch.zeros(cout, self.kp, dtype=torch.float32)
        w[:, :fan_in] = torch.randn(cout, fan_in, generator=gen) * std
        if name == "out":
            w[1:] = 0.0
        self.w = w.to(dev)                          # fp32 master, layout [co][tap*cin+ci] (padded)
        self.b = torch.zeros(cout, dtype=torch.float32, device=dev)
        self.w16 = self.w.to(BF16)
        self.m_w, self.v_w = torch.zeros_like(self.w), torch.zeros_like(self.w)
        self.m_b, self.v_b = torch.zeros_like(self.b), torch.zeros_like(self.b)
        self.gw = torch.zeros_like(self.w)
        self.gb = torch.zeros_like(self.b)

    def torch_weight(self):
        """[co][ci][kh][kw] fp32 view of the bf16 weights (for references)."""
        fan = self.k * self.k * self.cin
        w = self.w16.float()[:, :fan].reshape(self.cout, self.k, self.k, self.cin)
        return w.permute(0, 3, 1, 2).contiguous()

    def dgrad_weight16(self):
        """flipped, transposed weights of the input gradient: [ci][tap'*cout+co]."""
        fan = self.k * self.k * self.cin
        w = self.w16[:, :fan].reshape(self.cout, self.k, self.k, self.cin)
        wt = w.flip(1, 2).permute(3, 1, 2, 0).reshape(self.cin, self.k * self.k * self.cout)
        kp = _kp(self.k, self.cout)
        out = torch.zeros(self.cin, kp, dtype=BF16, device=w.device)
        out[:, :wt.shape[1]] = wt
        return out, kp


class UNet:
    """Forward/backward on NHWC bf16 activations, batch-major."""

    def __init__(self, seed=0, device=None):
        self.dev = device or torch.device("cuda", torch.cuda.current_device())
        g = torch.Generator().manual_seed(seed)
        self.layers = {n: Layer(n, ci, co, k, r, self.dev, g) for (n, ci, co, k, r) in LAYERS}
        self.step = 0
        # input / output normalisation (set from data)
        self.in_shift = torch.zeros(4)
        self.in_scale = torch.ones(4)
        self.out_scale = 1.0

    # ---- primitive launches --------------------------------------------
    def _conv(self, L, x, N, H, W, out32=False):
        lib = _lib_unet()
        y16 = torch.empty((N, H, W, L.cout), dtype=BF16, device=self.dev)
        y32 = torch.empty((N, H, W, L.cout), dtype=torch.float32, device=self.dev) if out32 else None
        _chk(lib.jsv_conv_fwd(_p(x), _p(L.w16), _p(L.b), _p(y16), _p(y32), N, H, W, L.cin, L.cout, L.k,
                              L.kp, L.relu, _sp()))
        return y16, y32

    def _conv_bwd(self, L, x, y, dy, N, H, W, need_dx=True):
        """dy: bf16 gradient of the layer output (pre-ReLU mask applied here)."""
        lib = _lib_unet()
        if L.relu:
            _chk(lib.jsv_relu_bwd(_p(dy), _p(y), dy.numel(), _sp()))
        M = N * H * W
        chunks = max(1, min(256, M // 4096))
        part = torch.empty(chunks * L.cout * (L.kp + 1), dtype=torch.float32, device=self.dev)
        _chk(lib.jsv_conv_wgrad(_p(x), _p(dy), _p(L.gw), _p(L.gb), _p(part), chunks, N, H, W, L.cin,
                                L.cout, L.k, L.kp, _sp()))
        if not need_dx:
            return None
        wd, kpd = L.dgrad_weight16()
        dx = torch.empty((N, H, W, L.cin), dtype=BF16, device=self.dev)
        _chk(lib.jsv_conv_fwd(_p(dy), _p(wd), None, _p(dx), None, N, H, W, L.cout, L.cin, L.k, kpd, 0, _sp()))
        return dx

    def _pool(self, x, N, H, W, C):
        y = torch.empty((N, H // 2, W // 2, C), dtype=BF16, device=self.dev)
        _chk(_lib_unet().jsv_avgpool2(_p(x), _p(y), N, H, W, C, 0, _sp()))
        return y

    def _pool_bwd(self, dy, N, H, W, C):
        dx = torch.empty((N, H, W, C), dtype=BF16, device=self.dev)
        _chk(_lib_unet().jsv_avgpool2(_p(dy), _p(dx), N, H, W, C, 1, _sp()))
        return dx

    def _upcat(self, x, skip, N, H, W, C1, C2):
        y = torch.empty((N, H, W, C1 + C2), dtype=BF16, device=self.dev)
        _chk(_lib_unet().jsv_up_concat(_p(x), _p(skip), _p(y), N, H, W, C1, C2, _sp()))
        return y

    def _upcat_bwd(self, dy, N, H, W, C1, C2):
        dx = torch.empty((N, H // 2, W // 2, C1), dtype=BF16, device=self.dev)
        ds = torch.empty((N, H, W, C2), dtype=BF16, device=self.dev)
        _chk(_lib_unet().jsv_up_concat_bwd(_p(dy), _p(dx), _p(ds), N, H, W, C1, C2, _sp()))
        return dx, ds

    # ---- network --------------------------------------------------------
    def forward(self, x, keep=False):
        """x: (N, H, W, 4) bf16 NHWC (normalised).  Returns fp32 (N, H, W)."""
        N, H, W, _ = x.shape
        Ls = self.layers
        a = {}
        a["x"] = x
        a["e1a"], _ = self._conv(Ls["e1a"], x, N, H, W)
        a["e1b"], _ = self._conv(Ls["e1b"], a["e1a"], N, H, W)
        a["p1"] = self._pool(a["e1b"], N, H, W, 16)
        H2, W2 = H // 2, W // 2
        a["e2a"], _ = self._conv(Ls["e2a"], a["p1"], N, H2, W2)
        a["e2b"], _ = self._conv(Ls["e2b"], a["e2a"], N, H2, W2)
        a["p2"] = self._pool(a["e2b"], N, H2, W2, 32)
        H3, W3 = H2 // 2, W2 // 2
        a["e3a"], _ = self._conv(Ls["e3a"], a["p2"], N, H3, W3)
        a["e3b"], _ = self._conv(Ls["e3b"], a["e3a"], N, H3, W3)
        a["p3"] = self._pool(a["e3b"], N, H3, W3, 64)
        H4, W4 = H3 // 2, W3 // 2
        a["b1"], _ = self._conv(Ls["b1"], a["p3"], N, H4, W4)
        a["b2"], _ = self._conv(Ls["b2"], a["b1"], N, H4, W4)
        a["u3"] = self._upcat(a["b2"], a["e3b"], N, H3, W3, 64, 64)
        a["d3a"], _ = self._conv(Ls["d3a"], a["u3"], N, H3, W3)
        a["d3b"], _ = self._conv(Ls["d3b"], a["d3a"], N, H3, W3)
        a["u2"] = self._upcat(a["d3b"], a["e2b"], N, H2, W2, 64, 32)
        a["d2a"], _ = self._conv(Ls["d2a"], a["u2"], N, H2, W2)
        a["d2b"], _ = self._conv(Ls["d2b"], a["d2a"], N, H2, W2)
        a["u1"] = self._upcat(a["d2b"], a["e1b"], N, H, W, 32, 16)
        a["d1a"], _ = self._conv(Ls["d1a"], a["u1"], N, H, W)
        a["d1b"], _ = self._conv(Ls["d1b"], a["d1a"], N, H, W)
        _, out = self._conv(Ls["out"], a["d1b"], N, H, W, out32=True)
        if keep:
            self._acts = a
        return out[..., 0].contiguous()

    def backward(self, dout):
        """dout: fp32 (N, H, W) gradient of the loss w.r.t. the output.
        Fills layer.gw / layer.gb."""
        a = self._acts
        N, H, W, _ = a["x"].shape
        H2, W2, H3, W3, H4, W4 = H // 2, W // 2, H // 4, W // 4, H // 8, W // 8
        Ls = self.layers
        d = torch.zeros((N, H, W, 4), dtype=BF16, device=self.dev)
        d[..., 0] = dout.to(BF16)
        d = self._conv_bwd(Ls["out"], a["d1b"], None, d, N, H, W)
        d = self._conv_bwd(Ls["d1b"], a["d1a"], a["d1b"], d, N, H, W)
        d = self._conv_bwd(Ls["d1a"], a["u1"], a["d1a"], d, N, H, W)
        d_d2b, d_e1b = self._upcat_bwd(d, N, H, W, 32, 16)
        d = self._conv_bwd(Ls["d2b"], a["d2a"], a["d2b"], d_d2b, N, H2, W2)
        d = self._conv_bwd(Ls["d2a"], a["u2"], a["d2a"], d, N, H2, W2)
        d_d3b, d_e2b = self._upcat_bwd(d, N, H2, W2, 64, 32)
        d = self._conv_bwd(Ls["d3b"], a["d3a"], a["d3b"], d_d3b, N, H3, W3)
        d = self._conv_bwd(Ls["d3a"], a["u3"], a["d3a"], d, N, H3, W3)
        d_b2, d_e3b = self._upcat_bwd(d, N, H3, W3, 64, 64)
        d = self._conv_bwd(Ls["b2"], a["b1"], a["b2"], d_b2, N, H4, W4)
        d = self._conv_bwd(Ls["b1"], a["p3"], a["b1"], d, N, H4, W4)
        d = self._pool_bwd(d, N, H3, W3, 64)
        _chk(_lib_unet().jsv_add_bf16(_p(d), _p(d_e3b), d.numel(), _sp()))
        d = self._conv_bwd(Ls["e3b"], a["e3a"], a["e3b"], d, N, H3, W3)
        d = self._conv_bwd(Ls["e3a"], a["p2"], a["e3a"], d, N, H3, W3)
        d = self._pool_bwd(d, N, H2, W2, 32)
        _chk(_lib_unet().jsv_add_bf16(_p(d), _p(d_e2b), d.numel(), _sp()))
        d = self._conv_bwd(Ls["e2b"], a["e2a"], a["e2b"], d, N, H2, W2)
        d = self._conv_bwd(Ls["e2a"], a["p1"], a["e2a"], d, N, H2, W2)
        d = self._pool_bwd(d, N, H, W, 16)
        _chk(_lib_unet().jsv_add_bf16(_p(d), _p(d_e1b), d.numel(), _sp()))
        d = self._conv_bwd(Ls["e1b"], a["e1a"], a["e1b"], d, N, H, W)
        self._conv_bwd(Ls["e1a"], a["x"], a["e1a"], d, N, H, W, need_dx=False)
        self._acts = None

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        self.step += 1
        lib = _lib_unet()
        for L in self.layers.values():
            _chk(lib.jsv_adam(_p(L.w), _p(L.gw), _p(L.m_w), _p(L.v_w), _p(L.w16), L.w.numel(), lr, b1, b2,
                              eps, self.step, wd, _sp()))
            _chk(lib.jsv_adam(_p(L.b), _p(L.gb), _p(L.m_b), _p(L.v_b), None, L.b.numel(), lr, b1, b2, eps,
                              self.step, 0.0, _sp()))

    # ---- (de)serialisation ---------------------------------------------
    def state_dict(self):
        sd = {f"{n}.w": L.w.cpu() for n, L in self.layers.items()}
        sd.update({f"{n}.b": L.b.cpu() for n, L in self.layers.items()})
        sd["in_shift"], sd["in_scale"] = self.in_shift, self.in_scale
        sd["out_scale"] = torch.tensor(self.out_scale)
        return sd

    def load_state_dict(self, sd):
        for n, L in self.layers.items():
            L.w.copy_(sd[f"{n}.w"].to(self.dev))
            L.b.copy_(sd[f"{n}.b"].to(self.dev))
            L.w16 = L.w.to(BF16)
        self.in_shift, self.in_scale = sd["in_shift"], sd["in_scale"]
        self.out_scale = float(sd["out_scale"])


# ---------------------------------------------------------------------------
# features / targets on the network window (nodes 1..m of each axis)
# ---------------------------------------------------------------------------
def features(net, lam, dxx, dyy, rhs):
    """Batched (B, m+3, m+3) float64 tensors -> (B, m, m, 4) bf16 normalised."""
    sl = (slice(None), slice(1, -2), slice(1, -2))
    x = torch.stack([torch.log(dxx[sl]), torch.log(dyy[sl]), lam[sl], rhs[sl]], dim=-1).float()
    x = (x - net.in_shift.to(x.device)) / net.in_scale.to(x.device)
    return x.to(BF16).contiguous()


def predict_phi(net, lam, dxx, dyy, rhs):
    """Initial phi (B, m+3, m+3) float64 with a zero rim."""
    out = net.forward(features(net, lam, dxx, dyy, rhs)).double() * net.out_scale
    phi = torch.zeros_like(rhs)
    phi[:, 2:-2, 2:-2] = out[:, 1:, 1:]
    return phi
