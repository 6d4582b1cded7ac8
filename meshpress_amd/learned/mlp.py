"""The proposal network on the GPU: 32 features -> 128 -> 128 -> 176 logits
with BF16 weights/activations and FP32 accumulation, driven through the
stride_mlp_* kernels of libstride_hip.so (forward, grouped softmax
cross-entropy, backward, Adam).  torch tensors only hold the device buffers.

Logit columns: predictor heads (one per axis, PRED_CLASSES each) at
0..50, coder heads (one per residual stream dx dy dz px py pz,
CODER_CLASSES each) at 51..170, padding at 171..175.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib

N_IN, N_HID, N_OUT = 32, 128, 176
N_HEADS = 9
HEAD_OFF = (0, 17, 34, 51, 71, 91, 111, 131, 151, 171)
PRED_CLASSES, CODER_CLASSES = 17, 20
N_W = N_IN * N_HID + N_HID * N_HID + N_HID * N_OUT
N_P = N_W + N_HID + N_HID + N_OUT
IGNORE = 255


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the proposal network runs on a ROCm GPU; torch.cuda is unavailable")
    return torch


def split_params(p: np.ndarray):
    """(W1, W2, W3, b1, b2, b3) views of a flat parameter vector."""
    o1 = N_IN * N_HID
    o2 = o1 + N_HID * N_HID
    o3 = o2 + N_HID * N_OUT
    return (p[:o1].reshape(N_IN, N_HID), p[o1:o2].reshape(N_HID, N_HID),
            p[o2:o3].reshape(N_HID, N_OUT), p[o3:o3 + N_HID], p[o3 + N_HID:o3 + 2 * N_HID],
            p[o3 + 2 * N_HID:])


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform hidden layers, Glorot-uniform output layer, zero biases;
    padding columns of the output layer stay zero."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_P, np.float32)
    W1, W2, W3, _, _, _ = split_params(p)
    W1[:] = rng.uniform(-1, 1, W1.shape) * np.sqrt(6.0 / N_IN)
    W2[:] = rng.uniform(-1, 1, W2.shape) * np.sqrt(6.0 / N_HID)
    W3[:, :HEAD_OFF[-1]] = rng.uniform(-1, 1, (N_HID, HEAD_OFF[-1])) * np.sqrt(6.0 / (N_HID + 20))
    return p


class ProposalMLP:
    """Device-resident parameters (float32 master + bf16 copy) and Adam state."""

    def __init__(self, params: np.ndarray | None = None, seed: int = 0):
        torch = _torch()
        self.torch = torch
        self.L = _lib.lib()
        sizes = [C.c_int32() for _ in range(5)]
        _lib.check(self.L.stride_mlp_sizes(*[C.byref(s) for s in sizes]))
        if tuple(s.value for s in sizes) != (N_IN, N_HID, N_OUT, N_W, N_P):
            raise RuntimeError("libstride_hip.so was built with another network shape")
        p = init_params(seed) if params is None else np.ascontiguousarray(params, np.float32)
        if p.shape != (N_P,):
            raise ValueError(f"expected {N_P} parameters, got {p.shape}")
        dev = "cuda"
        self.p = torch.from_numpy(p.copy()).to(dev)
        self.wb = torch.empty(N_W, dtype=torch.bfloat16, device=dev)
        _lib.check(self.L.stride_mlp_to_bf16(_lib.ptr(self.p), N_W, _lib.ptr(self.wb),
                                             _lib.stream_ptr()))
        self.m = torch.zeros(N_P, dtype=torch.float32, device=dev)
        self.v = torch.zeros(N_P, dtype=torch.float32, device=dev)
        self.g = torch.zeros(N_P, dtype=torch.float32, device=dev)
        self.step = 0

    # ------------------------------------------------------------ buffers
    def to_bf16(self, x):
        """float32 device tensor -> bf16 device tensor (round to nearest even)."""
        torch = self.torch
        x = x.contiguous()
        y = torch.empty(x.shape, dtype=torch.bfloat16, device=x.device)
        _lib.check(self.L.stride_mlp_to_bf16(_lib.ptr(x), x.numel(), _lib.ptr(y), _lib.stream_ptr()))
        return y

    def features_to_device(self, F: np.ndarray):
        torch = self.torch
        F = np.ascontiguousarray(F, np.float32).reshape(-1, N_IN)
        return self.to_bf16(torch.from_numpy(F).to("cuda"))

    # ------------------------------------------------------------ kernels
    def forward(self, x):
        """x: bf16 (n, 32) -> (h1 bf16, h2 bf16, z float32 (n, 176))."""
        torch = self.torch
        n = x.shape[0]
        h1 = torch.empty((n, N_HID), dtype=torch.bfloat16, device=x.device)
        h2 = torch.empty((n, N_HID), dtype=torch.bfloat16, device=x.device)
        z = torch.empty((n, N_OUT), dtype=torch.float32, device=x.device)
        _lib.check(self.L.stride_mlp_forward(_lib.ptr(x), n, _lib.ptr(self.wb), _lib.ptr(self.p),
                                             _lib.ptr(h1), _lib.ptr(h2), _lib.ptr(z),
                                             _lib.stream_ptr()))
        return h1, h2, z

    def xent(self, z, y, scale):
        """(dz bf16 (n, 176), per-sample loss float32 (n,))."""
        torch = self.torch
        n = z.shape[0]
        dz = torch.empty((n, N_OUT), dtype=torch.bfloat16, device=z.device)
        loss = torch.empty(n, dtype=torch.float32, device=z.device)
        _lib.check(self.L.stride_mlp_xent(_lib.ptr(z), _lib.ptr(y), n, float(scale), _lib.ptr(dz),
                                          _lib.ptr(loss), _lib.stream_ptr()))
        return dz, loss

    def backward(self, x, h1, h2, dz):
        """Fills self.g with the gradient of the loss that produced dz."""
        torch = self.torch
        n = x.shape[0]
        dh2 = torch.empty((n, N_HID), dtype=torch.bfloat16, device=x.device)
        dh1 = torch.empty((n, N_HID), dtype=torch.bfloat16, device=x.device)
        _lib.check(self.L.stride_mlp_backward(_lib.ptr(x), n, _lib.ptr(self.wb), _lib.ptr(h1),
                                              _lib.ptr(h2), _lib.ptr(dz), _lib.ptr(dh2),
                                              _lib.ptr(dh1), _lib.ptr(self.g), _lib.stream_ptr()))
        return dh2, dh1

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        _lib.check(self.L.stride_mlp_adam(_lib.ptr(self.p), _lib.ptr(self.m), _lib.ptr(self.v),
                                          _lib.ptr(self.g), _lib.ptr(self.wb), N_P, float(lr),
                                          float(beta1), float(beta2), float(eps), self.step,
                                          _lib.stream_ptr()))

    def train_step(self, x, y, lr):
        """One Adam step on a batch; returns the per-sample loss tensor."""
        h1, h2, z = self.forward(x)
        dz, loss = self.xent(z, y, 1.0 / x.shape[0])
        self.backward(x, h1, h2, dz)
        self.adam(lr)
        return loss

    def heads(self, z):
        """(predictor probabilities (n, 51), coder class (n, 6) uint8,
        coder class probability (n, 6)) as device tensors."""
        torch = self.torch
        n = z.shape[0]
        pp = torch.empty((n, 3 * PRED_CLASSES), dtype=torch.float32, device=z.device)
        cls = torch.empty((n, 6), dtype=torch.uint8, device=z.device)
        cp = torch.empty((n, 6), dtype=torch.float32, device=z.device)
        _lib.check(self.L.stride_mlp_heads(_lib.ptr(z), n, _lib.ptr(pp), _lib.ptr(cls), _lib.ptr(cp),
                                           _lib.stream_ptr()))
        return pp, cls, cp

    def infer(self, F: np.ndarray, want=("pred", "cls", "cp")):
        """Host features (n, 32) -> host (pred_prob, coder class, coder prob);
        outputs not named in `want` are returned as None (not copied back)."""
        n = len(F)
        if n == 0:
            return (np.zeros((0, 3 * PRED_CLASSES), np.float32), np.zeros((0, 6), np.uint8),
                    np.zeros((0, 6), np.float32))
        x = self.features_to_device(F)
        _, _, z = self.forward(x)
        pp, cls, cp = self.heads(z)
        return tuple(t.cpu().numpy() if w in want else None
                     for t, w in ((pp, "pred"), (cls, "cls"), (cp, "cp")))

    # ------------------------------------------------------------ I/O
    def params_host(self) -> np.ndarray:
        return self.p.cpu().numpy().copy()

    def save(self, path, **meta):
        np.savez(path, params=self.params_host(), step=np.int64(self.step),
                 **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as z:
            m = cls(params=z["params"])
            m.step = int(z["step"]) if "step" in z else 0
        return m
