"""Learned trace surrogate: a parameter MLP feeding a 1-D convolutional decoder.

No counterpart in the reference (which has no learned model).  The network
maps a query's encoded parameters to a 512-sample E_r trace, factored as a
log-amplitude and a unit-RMS shape:

    x (16) -> 4 x [Linear 256, SiLU] -> h
    h -> Linear -> log-amplitude (1)
    h -> Linear 32*32, SiLU -> z (length 32, 32 channels)
    z -> 4 x [upsample x2, Conv1d k=5, SiLU] -> (length 512, 16 channels)
      -> Conv1d k=5 -> shape (512)

Every matrix product of the forward and backward passes (the linear layers
and the convolutions as im2col products) runs in the library's BF16 MFMA
GEMM with FP32 accumulation (include/geosurrogate.h); activations and weights
enter the products in BF16, the optimiser keeps FP32 master weights (Adam).
torch only allocates device buffers.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native


@dataclass(frozen=True)
class SurrogateConfig:
    in_features: int = 16
    width: int = 256
    depth: int = 4
    seq0: int = 32
    channels: tuple[int, ...] = (32, 32, 32, 16, 16)  # z, then after each upsampling stage
    ksize: int = 5
    amp_weight: float = 1.0

    @property
    def stages(self) -> int:
        return len(self.channels) - 1

    @property
    def length(self) -> int:
        return self.seq0 * 2 ** self.stages


def bf16_bits(x: np.ndarray) -> np.ndarray:
    """FP32 -> BF16 bit patterns, round to nearest even (host side)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    r[nan] = ((u[nan] >> 16) | 0x40).astype(np.uint16)
    return r


def bf16_to_f32(bits: np.ndarray) -> np.ndarray:
    return (np.asarray(bits, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def parameter_shapes(cfg: SurrogateConfig) -> list[tuple[str, tuple[int, ...]]]:
    """(name, shape) of every parameter; matrices are (in, out) row-major."""
    out = []
    fan_in = cfg.in_features
    for i in range(cfg.depth):
        out += [(f"W{i}", (fan_in, cfg.width)), (f"b{i}", (cfg.width,))]
        fan_in = cfg.width
    out += [("Wa", (cfg.width, 1)), ("ba", (1,))]
    out += [("Wz", (cfg.width, cfg.seq0 * cfg.channels[0])), ("bz", (cfg.seq0 * cfg.channels[0],))]
    for s in range(cfg.stages):
        out += [(f"Wc{s}", (cfg.ksize * cfg.channels[s], cfg.channels[s + 1])),
                (f"bc{s}", (cfg.channels[s + 1],))]
    out += [("Wf", (cfg.ksize * cfg.channels[-1], 1)), ("bf", (1,))]
    return out


def initial_parameters(cfg: SurrogateConfig, seed: int = 0) -> dict[str, np.ndarray]:
    """Scaled-normal weights (variance 1/fan_in), zero biases."""
    rng = np.random.default_rng(seed)
    params = {}
    for name, shape in parameter_shapes(cfg):
        if name.startswith("W"):
            params[name] = (rng.standard_normal(shape) / np.sqrt(shape[0])).astype(np.float32)
        else:
            params[name] = np.zeros(shape, dtype=np.float32)
    return params


@dataclass
class _Cache:
    batch: int
    x16: object = None
    pre: list = field(default_factory=list)   # MLP pre-activations (FP32)
    act: list = field(default_factory=list)   # MLP activations (BF16), act[0] = input
    amp: object = None
    pz: object = None
    z: list = field(default_factory=list)     # decoder activations (BF16)
    cols: list = field(default_factory=list)  # im2col of every decoder conv (BF16)
    cpre: list = field(default_factory=list)  # decoder conv pre-activations (FP32)
    shape: object = None


class TraceSurrogate:
    """The network with its FP32 master weights, BF16 copies and Adam state on one GPU."""

    def __init__(self, cfg: SurrogateConfig | None = None, params: dict | None = None, *,
                 seed: int = 0, device: str = "cuda") -> None:
        import torch

        self.torch = torch
        self.cfg = cfg or SurrogateConfig()
        self.lib = _native.load_library()
        self.device = torch.device(device)
        if self.device.type != "cuda" or not torch.cuda.is_available():
            from .backends.base import BackendUnavailableError

            raise BackendUnavailableError("the trace surrogate runs on a ROCm GPU only")
        host = params if params is not None else initial_parameters(self.cfg, seed)
        self.params, self.params16, self.m, self.v, self.grads = {}, {}, {}, {}, {}
        for name, shape in parameter_shapes(self.cfg):
            value = np.asarray(host[name], dtype=np.float32).reshape(shape)
            p = torch.as_tensor(value.copy(), device=self.device)
            self.params[name] = p
            self.params16[name] = torch.empty(shape, dtype=torch.int16, device=self.device)
            self._check(self.lib.geo_f32_to_bf16(p.numel(), p.data_ptr(),
                                                 self.params16[name].data_ptr(), self._stream()))
            self.m[name] = torch.zeros_like(p)
            self.v[name] = torch.zeros_like(p)
            self.grads[name] = torch.zeros_like(p)
        self.step_count = 0
        self._loss_scratch = torch.zeros(1025, dtype=torch.float32, device=self.device)

    # ----------------------------------------------------------------- helpers
    def _stream(self):
        return C.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

    def _check(self, code):
        _native.check(code, self.lib)

    def _empty(self, shape, bf16=False):
        dtype = self.torch.int16 if bf16 else self.torch.float32
        return self.torch.empty(shape, dtype=dtype, device=self.device)

    def _gemm(self, M, N, K, a, a_rs, a_cs, b, b_rs, b_cs, *, bias=None, out32=None, out16=None,
              ldc=None, accumulate=False):
        self._check(self.lib.geo_gemm_bf16(
            M, N, K, a.data_ptr(), a_rs, a_cs, b.data_ptr(), b_rs, b_cs,
            None if bias is None else bias.data_ptr(),
            None if out32 is None else out32.data_ptr(),
            None if out16 is None else out16.data_ptr(), N if ldc is None else ldc,
            int(accumulate), self._stream()))

    def _silu(self, pre):
        out = self._empty(tuple(pre.shape), bf16=True)
        self._check(self.lib.geo_silu_forward(pre.numel(), pre.data_ptr(), out.data_ptr(),
                                              self._stream()))
        return out

    def _silu_back(self, pre, dout):
        d32 = self._empty(tuple(pre.shape))
        d16 = self._empty(tuple(pre.shape), bf16=True)
        self._check(self.lib.geo_silu_backward(pre.numel(), pre.data_ptr(), dout.data_ptr(),
                                               d32.data_ptr(), d16.data_ptr(), self._stream()))
        return d32, d16

    def _colsum(self, x, rows, cols, out):
        self._check(self.lib.geo_column_sums(rows, cols, x.data_ptr(), cols, out.data_ptr(),
                                             self._stream()))

    # ----------------------------------------------------------------- passes
    def forward(self, features) -> tuple:
        """features (B, in_features) FP32 device tensor -> (shape (B, L), log-amp (B,)) FP32."""
        cfg, P16 = self.cfg, self.params16
        x = features.to(device=self.device, dtype=self.torch.float32).contiguous()
        B = x.shape[0]
        if x.shape[1] != cfg.in_features:
            raise ValueError(f"features must have {cfg.in_features} columns")
        cache = _Cache(batch=B)
        x16 = self._empty((B, cfg.in_features), bf16=True)
        self._check(self.lib.geo_f32_to_bf16(x.numel(), x.data_ptr(), x16.data_ptr(),
                                             self._stream()))
        cache.act.append(x16)
        h, fan_in = x16, cfg.in_features
        for i in range(cfg.depth):
            pre = self._empty((B, cfg.width))
            self._gemm(B, cfg.width, fan_in, h, fan_in, 1, P16[f"W{i}"], cfg.width, 1,
                       bias=self.params[f"b{i}"], out32=pre)
            h = self._silu(pre)
            cache.pre.append(pre)
            cache.act.append(h)
            fan_in = cfg.width
        amp = self._empty((B, 1))
        self._gemm(B, 1, cfg.width, h, cfg.width, 1, P16["Wa"], 1, 1, bias=self.params["ba"],
                   out32=amp)
        nz = cfg.seq0 * cfg.channels[0]
        pz = self._empty((B, nz))
        self._gemm(B, nz, cfg.width, h, cfg.width, 1, P16["Wz"], nz, 1, bias=self.params["bz"],
                   out32=pz)
        z = self._silu(pz)
        cache.amp, cache.pz = amp, pz
        cache.z.append(z)
        length = cfg.seq0
        for s in range(cfg.stages + 1):
            last = s == cfg.stages
            up = 1 if last else 2
            c_in = cfg.channels[s]
            c_out = 1 if last else cfg.channels[s + 1]
            lo = length * up
            kc = cfg.ksize * c_in
            cols = self._empty((B * lo, kc), bf16=True)
            self._check(self.lib.geo_im2col_up(B, length, c_in, up, cfg.ksize, z.data_ptr(),
                                               cols.data_ptr(), self._stream()))
            pre = self._empty((B * lo, c_out))
            w, b = ("Wf", "bf") if last else (f"Wc{s}", f"bc{s}")
            self._gemm(B * lo, c_out, kc, cols, kc, 1, P16[w], c_out, 1, bias=self.params[b],
                       out32=pre)
            cache.cols.append(cols)
            if last:
                cache.shape = pre.view(B, lo)
            else:
                z = self._silu(pre)
                cache.cpre.append(pre)
                cache.z.append(z)
            length = lo
        self._cache = cache
        return cache.shape, amp.view(B)

    def backward(self, d_shape, d_amp) -> dict:
        """Gradients (FP32, in self.grads) of the last forward pass."""
        cfg, P16, G = self.cfg, self.params16, self.grads
        cache = getattr(self, "_cache", None)
        if cache is None:
            raise RuntimeError("backward needs a forward pass first (predict keeps no activations)")
        B = cache.batch
        ds = d_shape.contiguous().view(-1, 1)
        ds16 = self._empty(tuple(ds.shape), bf16=True)
        self._check(self.lib.geo_f32_to_bf16(ds.numel(), ds.data_ptr(), ds16.data_ptr(),
                                             self._stream()))
        # decoder, last conv first
        grad32, grad16 = ds, ds16
        length = cfg.length
        for s in range(cfg.stages, -1, -1):
            last = s == cfg.stages
            up = 1 if last else 2
            c_in = cfg.channels[s]
            c_out = 1 if last else cfg.channels[s + 1]
            lo = length
            li = lo // up
            kc = cfg.ksize * c_in
            w, b = ("Wf", "bf") if last else (f"Wc{s}", f"bc{s}")
            cols = cache.cols[s]
            # dW = cols^T grad, db = column sums, dcols = grad W^T
            self._gemm(kc, c_out, B * lo, cols, 1, kc, grad16, c_out, 1, out32=G[w])
            self._colsum(grad32, B * lo, c_out, G[b])
            dcols = self._empty((B * lo, kc))
            self._gemm(B * lo, kc, c_out, grad16, c_out, 1, P16[w], 1, c_out, out32=dcols)
            dz = self._empty((B, li, c_in))
            self._check(self.lib.geo_col2im_up(B, li, c_in, up, cfg.ksize, dcols.data_ptr(),
                                               dz.data_ptr(), self._stream()))
            pre = cache.cpre[s - 1] if s > 0 else cache.pz
            grad32, grad16 = self._silu_back(pre, dz)
            length = li
        # projection to the decoder and the amplitude head
        nz = cfg.seq0 * cfg.channels[0]
        h = cache.act[-1]
        self._gemm(cfg.width, nz, B, h, 1, cfg.width, grad16, nz, 1, out32=G["Wz"])
        self._colsum(grad32, B, nz, G["bz"])
        dh = self._empty((B, cfg.width))
        self._gemm(B, cfg.width, nz, grad16, nz, 1, P16["Wz"], 1, nz, out32=dh)
        da = d_amp.contiguous().view(B, 1)
        da16 = self._empty((B, 1), bf16=True)
        self._check(self.lib.geo_f32_to_bf16(B, da.data_ptr(), da16.data_ptr(), self._stream()))
        self._gemm(cfg.width, 1, B, h, 1, cfg.width, da16, 1, 1, out32=G["Wa"])
        self._colsum(da, B, 1, G["ba"])
        self._gemm(B, cfg.width, 1, da16, 1, 1, P16["Wa"], 1, 1, out32=dh, accumulate=True)
        # MLP
        for i in range(cfg.depth - 1, -1, -1):
            fan_in = cfg.in_features if i == 0 else cfg.width
            d32, d16 = self._silu_back(cache.pre[i], dh)
            self._gemm(fan_in, cfg.width, B, cache.act[i], 1, fan_in, d16, cfg.width, 1,
                       out32=G[f"W{i}"])
            self._colsum(d32, B, cfg.width, G[f"b{i}"])
            if i > 0:
                dh = self._empty((B, fan_in))
                self._gemm(B, fan_in, cfg.width, d16, cfg.width, 1, P16[f"W{i}"], 1, cfg.width,
                           out32=dh)
        return G

    def loss_and_grad(self, features, shape_t, amp_t) -> float:
        """Forward, trace loss and backward; returns the loss (synchronises)."""
        shape, amp = self.forward(features)
        B, L = shape.shape
        ds = self._empty((B, L))
        da = self._empty((B,))
        st = shape_t.to(device=self.device, dtype=self.torch.float32).contiguous()
        at = amp_t.to(device=self.device, dtype=self.torch.float32).contiguous()
        loss = self._loss_scratch[1024:]
        self._check(self.lib.geo_trace_loss(B, L, shape.data_ptr(), st.data_ptr(), amp.data_ptr(),
                                            1, at.data_ptr(), self.cfg.amp_weight, ds.data_ptr(),
                                            da.data_ptr(), self._loss_scratch.data_ptr(),
                                            loss.data_ptr(), self._stream()))
        self.backward(ds, da)
        return float(loss.item())

    def adam_step(self, lr: float, beta1=0.9, beta2=0.999, eps=1e-8) -> None:
        self.step_count += 1
        for name, p in self.params.items():
            self._check(self.lib.geo_adam_step(p.numel(), p.data_ptr(), self.grads[name].data_ptr(),
                                               self.m[name].data_ptr(), self.v[name].data_ptr(),
                                               self.params16[name].data_ptr(), lr, beta1, beta2,
                                               eps, self.step_count, self._stream()))

    def train_step(self, features, shape_t, amp_t, lr: float) -> float:
        loss = self.loss_and_grad(features, shape_t, amp_t)
        self.adam_step(lr)
        return loss

    def predict(self, features, batch: int = 8191) -> tuple[np.ndarray, np.ndarray]:
        """(shape (n, L), log-amp (n,)) as host arrays.  Batches of up to
        `batch` rows (a row's result does not depend on its batch: every
        GEMM tile and row chunk gives the same bits) are gathered on the
        device and copied to the host once.  The default keeps the decoder's
        last GEMM in one launch and the activations of a batch near 4 GB."""
        torch = self.torch
        x = torch.as_tensor(np.asarray(features, dtype=np.float32), device=self.device)
        n = x.shape[0]
        shapes = torch.empty((n, self.cfg.length), dtype=torch.float32, device=self.device)
        amps = torch.empty(n, dtype=torch.float32, device=self.device)
        for i in range(0, n, batch):
            s, a = self.forward(x[i:i + batch])
            shapes[i:i + s.shape[0]].copy_(s)
            amps[i:i + s.shape[0]].copy_(a)
        self._cache = None  # the activations of the last batch are not kept
        return shapes.cpu().numpy(), amps.cpu().numpy()

    def host_parameters(self) -> dict[str, np.ndarray]:
        return {k: v.cpu().numpy().copy() for k, v in self.params.items()}
