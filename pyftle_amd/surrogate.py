"""Learned flow-map surrogate: Fourier-feature MLP on the project's HIP kernels.

The network maps ``(x, t0)`` (seed position and window start) to the
displacement of the flow map over a fixed window ``T``:
``Phi(x, t0) = x + NN(x, t0)``.  Its Jacobian ``dPhi/dx`` comes from
forward-mode tangents (JVP) pushed through the same GEMMs, so the FTLE needs
one primal and ``dim`` tangent row blocks per seed instead of ``2*dim``
neighbour queries.  Training data are trajectories of the fused RK4 kernel;
the FP64 FTLE kernels finish the job.  An ensemble of members flags seeds
where the members disagree; those seeds fall back to RK4 (``FlowMapEnsemble``).

Kernels (``csrc/mlp.hip``): BF16 MFMA GEMM with bias+SiLU / SiLU' epilogues
and split-K weight gradients, Fourier features with analytic tangents, the
narrow output layer, MSE and Adam.  Master weights are FP32; the GEMMs read a
BF16 copy written by the Adam kernel.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import torch

from . import _native as N
from .field import stream_handle

EPI_F32, EPI_BIAS_SILU, EPI_DSILU, EPI_BIAS = 0, 1, 2, 3

_vp, _i32, _i64, _f32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float
_SIGS = {
    "pyftle_mlp_gemm": [_vp, _i64, _i32, _vp, _i64, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _i32,
                        _vp, _i64, _vp],
    "pyftle_mlp_fourier": [_vp, _i32, _vp, _vp, _i32, _i32, _i32, _i32, _vp, _vp],
    "pyftle_mlp_linear_out": [_vp, _i64, _i32, _vp, _vp, _i32, _i32, _i32, _vp, _vp],
    "pyftle_mlp_linear_out_backward": [_vp, _vp, _i32, _i32, _i32, _vp, _vp, _i64, _i32, _vp, _vp, _vp, _vp],
    "pyftle_mlp_reduce_splits": [_vp, _i32, _i64, _vp, _vp],
    "pyftle_mlp_colsum": [_vp, _i64, _i32, _i32, _i32, _vp, _vp],
    "pyftle_mlp_mse": [_vp, _vp, _i64, _f32, _vp, _vp, _vp],
    "pyftle_mlp_adam": [_vp, _vp, _vp, _vp, _i64, _f32, _f32, _f32, _f32, _i32, _f32, _vp, _vp],
    "pyftle_mlp_cast_bf16": [_vp, _i64, _vp, _vp],
    "pyftle_mlp_pack_fused": [_vp, _i32, _i32, _vp, _vp],
    "pyftle_mlp_adam_scheduled": [_vp, _vp, _vp, _vp, _i64, _f32, _f32, _f32, _f32, _f32, _f32, _i32, _i32, _vp,
                                  _vp, _vp],
    "pyftle_mlp_fused_forward": [_vp, _i32, _i32, _vp, _vp, _i32, _vp, _vp, _i32, _i32, _vp, _vp, _i32, _f32,
                                 _i32, _vp, _vp],
    "pyftle_mlp_train_work_bytes": [_i32, _i32, _i32, _i32, _i32],
    "pyftle_mlp_train_grads": [_vp, _vp],
    "pyftle_mlp_train_step": [_vp, _vp, _vp, _f32, _f32, _f32, _f32, _f32, _f32, _i32, _i32, _vp],
    "pyftle_mlp_pack_train": [_vp, _i32, _i32, _vp, _vp],
    "pyftle_mlp_train_adam": [_vp, _vp, _vp, _vp, _i32, _i32, _i32, _f32, _f32, _f32, _f32, _f32, _f32, _i32,
                              _i32, _vp, _vp, _vp],
}


class TrainArgs(ctypes.Structure):
    """``pyftle_mlp_train_args`` (include/pyftle_hip.h)."""

    _fields_ = [("x", _vp), ("targets", _vp), ("base_row", _vp), ("M", _i32), ("d_in", _i32), ("d_out", _i32),
                ("n_freq", _i32), ("width", _i32), ("depth", _i32), ("splits", _i32), ("scale", _vp), ("Bf", _vp),
                ("W_packed", _vp), ("params", _vp), ("grads", _vp), ("loss", _vp), ("work", _vp), ("step", _vp)]
# (outputs, tangents) covered by the one-kernel network, per hidden width
_FUSED_SHAPES = {256: ((2, 0), (2, 2), (3, 0)), 512: ((2, 0), (3, 0), (3, 3))}
_registered = False
_WORKSPACE: dict = {}


def _lib():
    global _registered
    if not _registered:
        N.extra_signatures(_SIGS)
        N.load().pyftle_mlp_train_work_bytes.restype = ctypes.c_int64
        _registered = True
    return N.load()


def pad128(n: int) -> int:
    return (n + 127) // 128 * 128


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def gemm(A, lda, a_kc, B, ldb, b_kc, M, Nn, K, C, ldc, epi=EPI_F32, splits=1, bias=None, Z=None,
         z_rows=0, stream=None):
    """``C = A B`` through ``pyftle_mlp_gemm`` (pointers may be tensor views)."""
    _lib()
    N.call("pyftle_mlp_gemm", A, int(lda), int(a_kc), B, int(ldb), int(b_kc), int(M), int(Nn), int(K),
           int(splits), int(epi), bias, Z, int(z_rows), C, int(ldc), stream_handle(stream))


def _off(t: torch.Tensor, elems: int) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr() + elems * t.element_size())


@dataclass
class MLPSpec:
    d_in: int = 3
    d_out: int = 2
    n_freq: int = 64
    width: int = 256
    depth: int = 5
    sigma: float = 1.0
    scale: tuple = (0.5, 1.0, 0.1)
    out_scale: float = 1.0

    def n_params(self) -> int:
        fan = 2 * self.n_freq
        n = 0
        for _ in range(self.depth):
            n += self.width * fan + self.width
            fan = self.width
        return n + self.d_out * self.width + self.d_out


class FlowMapMLP:
    """One network: parameters in one flat FP32 buffer (+ BF16 copy)."""

    def __init__(self, spec: MLPSpec, seed: int = 0, device: str = "cuda"):
        _lib()
        self.spec = spec
        g = torch.Generator().manual_seed(seed)
        dev = torch.device(device)
        self.B = (torch.randn(spec.n_freq, spec.d_in, generator=g, dtype=torch.float64) * spec.sigma).float().to(dev)
        self.scale = torch.tensor(spec.scale, dtype=torch.float32, device=dev)
        self.n_params = spec.n_params()
        self.params = torch.zeros(self.n_params, dtype=torch.float32, device=dev)
        self.offsets = []  # (W offset, W shape, b offset) per hidden layer
        o = 0
        fan = 2 * spec.n_freq
        for _ in range(spec.depth):
            w = torch.randn(spec.width, fan, generator=g, dtype=torch.float64) * math.sqrt(2.0 / fan)
            self.params[o:o + w.numel()] = w.reshape(-1).float().to(dev)
            self.offsets.append((o, (spec.width, fan), o + w.numel()))
            o += w.numel() + spec.width
            fan = spec.width
        wout = torch.randn(spec.d_out, spec.width, generator=g, dtype=torch.float64) * math.sqrt(1.0 / spec.width)
        self.params[o:o + wout.numel()] = wout.reshape(-1).float().to(dev)
        self.out_off = (o, (spec.d_out, spec.width), o + wout.numel())
        self.params_bf16 = torch.empty(self.n_params, dtype=torch.bfloat16, device=dev)
        self._fused = None
        self._fused_step = -1
        self.sync_bf16()
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grads = torch.zeros_like(self.params)
        self.step = 0
        self._ws = {}

    # -- parameter views ---------------------------------------------------
    def W(self, l):
        o, shp, _ = self.offsets[l]
        return self.params[o:o + shp[0] * shp[1]].view(shp)

    def Wb(self, l):
        o, shp, _ = self.offsets[l]
        return self.params_bf16[o:o + shp[0] * shp[1]].view(shp)

    def b(self, l):
        _, shp, ob = self.offsets[l]
        return self.params[ob:ob + shp[0]]

    def Wout(self):
        o, shp, _ = self.out_off
        return self.params[o:o + shp[0] * shp[1]].view(shp)

    def bout(self):
        _, shp, ob = self.out_off
        return self.params[ob:ob + shp[0]]

    def sync_bf16(self):
        """Refresh the BF16 copies after the FP32 parameters changed outside ``adam_step``."""
        N.call("pyftle_mlp_cast_bf16", _p(self.params), self.n_params, _p(self.params_bf16), stream_handle())
        self._fused = None

    def to_numpy(self):
        """Parameters as the oracle's dict (FP64 copies of the BF16 GEMM weights)."""
        s = self.spec
        return {
            "B": self.B.double().cpu().numpy(), "scale": self.scale.double().cpu().numpy(),
            "W": [self.Wb(l).double().cpu().numpy() for l in range(s.depth)],
            "b": [self.b(l).double().cpu().numpy() for l in range(s.depth)],
            "Wout": self.Wout().double().cpu().numpy() * s.out_scale,
            "bout": self.bout().double().cpu().numpy() * s.out_scale,
        }

    def fused_supported(self, n_tangents: int) -> bool:
        """Whether the one-kernel network (``mlp_fused.hip``) covers this shape."""
        s = self.spec
        return (s.n_freq == 64 and 2 <= s.depth <= 8 and s.d_in <= 4
                and (s.d_out, n_tangents) in _FUSED_SHAPES.get(s.width, ()))

    def pack_fused(self):
        """Weights in the fused kernel's order (bf16) and the hidden biases (depth, width)."""
        s = self.spec
        n = sum(shp[0] * shp[1] for _, shp, _ in self.offsets)
        Wp = torch.empty(n, dtype=torch.bfloat16, device=self.params.device)
        pos = 0
        for o, (rows, k), _ in self.offsets:
            N.call("pyftle_mlp_pack_fused", _off(self.params, o), rows, k, _off(Wp, pos), stream_handle())
            pos += rows * k
        bias = torch.stack([self.b(l) for l in range(s.depth)]).contiguous()
        # own allocations: the kernel reads bias and Wout with 16-byte vector loads
        self._fused = (Wp, bias, self.Wout().clone(), self.bout().clone())
        self._fused_step = self.step
        return self._fused

    def forward_fused(self, x: torch.Tensor, n_tangents: int = 0, out: torch.Tensor | None = None):
        """Inference through the one-kernel network; ``(1 + n_tangents, M, d_out)`` FP32."""
        s = self.spec
        if not self.fused_supported(n_tangents):
            raise ValueError(f"fused inference does not cover {s} with {n_tangents} tangents")
        if getattr(self, "_fused", None) is None or self._fused_step != self.step:
            self.pack_fused()
        Wp, bias, wout, bout = self._fused
        if x.dim() != 2 or x.shape[1] != s.d_in or x.dtype != torch.float32 or x.device != self.params.device:
            raise ValueError(f"x must be a float32 (M, {s.d_in}) tensor on {self.params.device}")
        x = x.contiguous()
        M = x.shape[0]
        shape = (1 + n_tangents, M, s.d_out)
        if out is None:
            Y = torch.empty(shape, dtype=torch.float32, device=x.device)
        elif tuple(out.shape) != shape or out.dtype != torch.float32 or not out.is_contiguous() \
                or out.device != x.device:
            raise ValueError(f"out must be a contiguous float32 tensor of shape {shape} on {x.device}")
        else:
            Y = out
        N.call("pyftle_mlp_fused_forward", _p(x), s.d_in, M, _p(self.scale), _p(self.B), s.n_freq, _p(Wp), _p(bias), s.depth,
               s.width, _p(wout), _p(bout), s.d_out, float(s.out_scale), int(n_tangents), _p(Y), stream_handle())
        return Y

    def state_dict(self):
        return {"spec": self.spec.__dict__, "B": self.B.cpu(), "params": self.params.cpu()}

    def load_state(self, sd):
        self.B.copy_(sd["B"].to(self.B.device))
        self.params.copy_(sd["params"].to(self.params.device))
        self.sync_bf16()
        self._fused = None

    # -- workspace ---------------------------------------------------------
    def _workspace(self, Mp: int, blocks: int):
        """Activation buffers, shared by every network of the same shape (the two most
        recently used sizes stay resident: an ensemble alternates tangent and primal-only
        members)."""
        s = self.spec
        key = (s.d_in, s.d_out, s.n_freq, s.width, s.depth, Mp, blocks, str(self.params.device))
        ws = _WORKSPACE.pop(key, None)
        if ws is not None:
            _WORKSPACE[key] = ws  # most recent last
        if ws is None:
            dev = self.params.device
            R = Mp * blocks
            while len(_WORKSPACE) >= 2:
                _WORKSPACE.pop(next(iter(_WORKSPACE)))
            ws = {
                "x": torch.zeros(Mp, s.d_in, dtype=torch.float32, device=dev),
                "F": torch.empty(R, 2 * s.n_freq, dtype=torch.bfloat16, device=dev),
                "H": [torch.empty(R, s.width, dtype=torch.bfloat16, device=dev) for _ in range(s.depth)],
                "Z": [torch.empty(Mp, s.width, dtype=torch.bfloat16, device=dev) for _ in range(s.depth)],
                "Y": torch.empty(R, s.d_out, dtype=torch.float32, device=dev),
            }
            _WORKSPACE[key] = ws
        self._ws = {(Mp, blocks): ws}
        return ws

    # -- forward / JVP -----------------------------------------------------
    def forward(self, x: torch.Tensor, n_tangents: int = 0, keep_z: bool = True):
        """Outputs ``(M, d_out)`` and, with ``n_tangents``, d out / d x_a for a < n_tangents.

        ``keep_z``: store every hidden layer's pre-activation (read by the tangent rows and by
        the backward pass); primal-only inference skips it and writes half the bytes."""
        s = self.spec
        M = x.shape[0]
        Mp = pad128(max(M, 1))
        blocks = 1 + n_tangents
        ws = self._workspace(Mp, blocks)
        ws["x"][:M].copy_(x)
        if M < Mp:
            ws["x"][M:].zero_()
        F2 = 2 * s.n_freq
        N.call("pyftle_mlp_fourier", _p(ws["x"]), s.d_in, _p(self.scale), _p(self.B), s.n_freq, Mp, F2,
               n_tangents, _p(ws["F"]), stream_handle())
        h, fan = ws["F"], F2
        for l in range(s.depth):
            H = ws["H"][l]
            zp = _p(ws["Z"][l]) if (keep_z or n_tangents) else ctypes.c_void_p(0)
            gemm(_p(h), fan, 1, _p(self.Wb(l)), fan, 1, Mp, s.width, fan, _p(H), s.width,
                 epi=EPI_BIAS_SILU, bias=_p(self.b(l)), Z=zp)
            if n_tangents:
                gemm(_off(h, Mp * fan), fan, 1, _p(self.Wb(l)), fan, 1, Mp * n_tangents, s.width, fan,
                     _off(H, Mp * s.width), s.width, epi=EPI_DSILU, Z=_p(ws["Z"][l]), z_rows=Mp)
            h, fan = H, s.width
        N.call("pyftle_mlp_linear_out", _p(h), s.width, s.width, _p(self.Wout()), _p(self.bout()), s.d_out,
               Mp * blocks, Mp, _p(ws["Y"]), stream_handle())
        Y = ws["Y"].view(blocks, Mp, s.d_out)[:, :M] * s.out_scale
        if n_tangents:
            return Y[0], [Y[1 + a] for a in range(n_tangents)]
        return Y[0]

    # -- training ----------------------------------------------------------
    def loss_and_backward(self, x: torch.Tensor, target: torch.Tensor, loss_out: torch.Tensor | None = None):
        """MSE over all outputs; fills ``self.grads``; returns the loss tensor (device)."""
        s = self.spec
        M = x.shape[0]
        Mp = pad128(M)
        self.forward(x)
        ws = self._ws[(Mp, 1)]
        dY = torch.zeros(Mp, s.d_out, dtype=torch.float32, device=x.device)
        loss = loss_out if loss_out is not None else torch.zeros(1, dtype=torch.float32, device=x.device)
        # the network output before out_scale is Y/out_scale; fit target/out_scale
        tgt = (target / s.out_scale).contiguous()
        N.call("pyftle_mlp_mse", _p(ws["Y"]), _p(tgt), M * s.d_out, 1.0 / (M * s.d_out), _p(dY), _p(loss),
               stream_handle())
        self.backward(dY, Mp)
        return loss

    def backward(self, dY: torch.Tensor, Mp: int):
        s = self.spec
        ws = self._ws[(Mp, 1)]
        dev = dY.device
        chunk = 32  # rows per partial sum of the narrow reductions (many partials: parallelism)
        chunks = (Mp + chunk - 1) // chunk
        g = self.grads
        o, shp, ob = self.out_off
        dWp = torch.empty(chunks, s.d_out, s.width, dtype=torch.float32, device=dev)
        dbp = torch.empty(chunks, s.d_out, dtype=torch.float32, device=dev)
        dZ = torch.empty(Mp, s.width, dtype=torch.bfloat16, device=dev)
        hlast = ws["H"][s.depth - 1]
        N.call("pyftle_mlp_linear_out_backward", _p(dY), _p(self.Wout()), s.d_out, s.width, Mp, _p(hlast),
               _p(ws["Z"][s.depth - 1]), s.width, chunk, _p(dWp), _p(dbp), _p(dZ), stream_handle())
        N.call("pyftle_mlp_reduce_splits", _p(dWp), chunks, s.d_out * s.width, _off(g, o), stream_handle())
        N.call("pyftle_mlp_reduce_splits", _p(dbp), chunks, s.d_out, _off(g, ob), stream_handle())
        # weight gradients are (out x in) GEMMs over K = Mp rows: split K so that the
        # 2-4 output tiles still give a few hundred work-groups
        splits = max(1, min(64, Mp // 256))
        while Mp % (splits * 64):
            splits -= 1
        F2 = 2 * s.n_freq
        for l in range(s.depth - 1, -1, -1):
            o, (wo, wi), ob = self.offsets[l]
            hin = ws["F"] if l == 0 else ws["H"][l - 1]
            part = torch.empty(splits, wo, wi, dtype=torch.float32, device=dev)
            gemm(_p(dZ), s.width, 0, _p(hin), wi, 0, wo, wi, Mp, _p(part), wi, epi=EPI_F32, splits=splits)
            N.call("pyftle_mlp_reduce_splits", _p(part), splits, wo * wi, _off(g, o), stream_handle())
            cp = torch.empty(chunks, wo, dtype=torch.float32, device=dev)
            N.call("pyftle_mlp_colsum", _p(dZ), s.width, Mp, wo, chunk, _p(cp), stream_handle())
            N.call("pyftle_mlp_reduce_splits", _p(cp), chunks, wo, _off(g, ob), stream_handle())
            if l > 0:
                dZp = torch.empty(Mp, wi, dtype=torch.bfloat16, device=dev)
                gemm(_p(dZ), s.width, 1, _p(self.Wb(l)), wi, 0, Mp, wi, wo, _p(dZp), wi,
                     epi=EPI_DSILU, Z=_p(ws["Z"][l - 1]), z_rows=Mp)
                dZ = dZp
        return g

    def adam_step(self, lr: float, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0):
        self.step += 1
        N.call("pyftle_mlp_adam", _p(self.params), _p(self.m), _p(self.v), _p(self.grads), self.n_params,
               float(lr), float(beta1), float(beta2), float(eps), self.step, float(weight_decay),
               _p(self.params_bf16), stream_handle())


def fused_train_supported(spec: MLPSpec, batch: int) -> bool:
    """Whether ``pyftle_mlp_train_grads`` covers this network and minibatch."""
    return (spec.n_freq == 64 and spec.width in (256, 512) and 2 <= spec.depth <= 8 and spec.d_in <= 4
            and spec.d_out <= 3 and batch % 128 == 0)


class FusedTrainStep:
    """One minibatch step on the fused training kernels: gradients of rows
    ``*base .. *base + batch - 1`` of (X, T / out_scale) (``pyftle_mlp_train_grads``, which
    also adds the loss to ``loss``, advances ``base`` and, given, the step index ``step``),
    then ``pyftle_mlp_train_adam`` (Adam at that step's learning rate, writing the packed
    bf16 weights the next step reads); ``step`` runs both with the update applied by the
    gradient reduction (``pyftle_mlp_train_step``)."""

    def __init__(self, model: FlowMapMLP, X: torch.Tensor, Tn: torch.Tensor, batch: int, base: torch.Tensor,
                 loss: torch.Tensor, splits: int = 0, step: torch.Tensor | None = None):
        _lib()
        s = model.spec
        if not fused_train_supported(s, batch):
            raise ValueError(f"the fused training step does not cover {s} with batch {batch}")
        if splits <= 0:  # measured best at 16384 rows: 16 at width 256, 8 at width 512
            splits = max(1, min(16 if s.width == 256 else 8, batch // 1024))
        while batch % (128 * splits):
            splits -= 1
        self.model, self.batch, self.splits = model, batch, splits
        dev = X.device
        nbytes = N.load().pyftle_mlp_train_work_bytes(batch, s.width, s.depth, s.d_out, splits)
        if nbytes <= 0:
            raise ValueError("unsupported fused training shape")
        self.work = torch.empty(int(nbytes), dtype=torch.uint8, device=dev)
        n_packed = s.width * 2 * s.n_freq + 2 * (s.depth - 1) * s.width * s.width
        self.packed = torch.empty(n_packed, dtype=torch.bfloat16, device=dev)
        self.repack()
        a = TrainArgs()
        a.x, a.targets, a.base_row = X.data_ptr(), Tn.data_ptr(), base.data_ptr()
        a.M, a.d_in, a.d_out, a.n_freq = batch, s.d_in, s.d_out, s.n_freq
        a.width, a.depth, a.splits = s.width, s.depth, splits
        a.scale, a.Bf = model.scale.data_ptr(), model.B.data_ptr()
        a.W_packed, a.params, a.grads = self.packed.data_ptr(), model.params.data_ptr(), model.grads.data_ptr()
        a.loss, a.work = loss.data_ptr(), self.work.data_ptr()
        a.step = step.data_ptr() if step is not None else None
        self.args = a
        self._keep = (X, Tn, base, loss, step)

    def repack(self):
        s = self.model.spec
        N.call("pyftle_mlp_pack_train", _p(self.model.params), s.width, s.depth, _p(self.packed), stream_handle())

    def grads(self):
        N.call("pyftle_mlp_train_grads", ctypes.byref(self.args), stream_handle())

    def step(self, lr_max: float, lr_min: float, total: int, step0: int):
        """``grads()`` then ``adam()`` with the update applied by the gradient reduction
        (``pyftle_mlp_train_step``: three launches instead of four, the same bits); needs the
        step index tensor given at construction."""
        if self.args.step is None:
            raise ValueError("the fused step needs the device step index")
        m, s = self.model, self.model.spec
        N.call("pyftle_mlp_train_step", ctypes.byref(self.args), _p(m.m), _p(m.v), 0.9, 0.999, 1e-8, 0.0,
               float(lr_max), float(lr_min), int(total), int(step0), stream_handle())

    def adam(self, lr_max: float, lr_min: float, total: int, step0: int, step: torch.Tensor):
        m, s = self.model, self.model.spec
        N.call("pyftle_mlp_train_adam", _p(m.params), _p(m.m), _p(m.v), _p(m.grads), s.width, s.depth, s.d_out,
               0.9, 0.999, 1e-8, 0.0, float(lr_max), float(lr_min), int(total), int(step0), _p(step),
               _p(self.packed), stream_handle())


class _Trainer:
    """One model's training run (``train``): shuffled copies of the data, the device-side row
    offset, step index and loss accumulator, and the step's kernels."""

    def __init__(self, model: FlowMapMLP, X: torch.Tensor, T: torch.Tensor, epochs: int, batch: int,
                 lr: float, lr_min: float, seed: int, graph: bool, fused: bool, splits: int):
        n = X.shape[0]
        if n < batch:
            raise ValueError(f"need at least one batch of {batch} rows, got {n}")
        dev = X.device
        self.model, self.X, self.n, self.batch = model, X, n, batch
        self.lr, self.lr_min, self.graph = lr, lr_min, graph
        self.steps_per_epoch = n // batch
        self.total = epochs * self.steps_per_epoch
        self.gen = torch.Generator(device=dev).manual_seed(seed)
        self.Xs, self.Ts = torch.empty_like(X), torch.empty_like(T)
        self.xb = torch.empty(batch, X.shape[1], dtype=X.dtype, device=dev)
        self.tb = torch.empty(batch, T.shape[1], dtype=T.dtype, device=dev)
        self.base = torch.zeros(1, dtype=torch.int64, device=dev)
        self.ar = torch.arange(batch, dtype=torch.int64, device=dev)
        self.k_dev = torch.zeros(1, dtype=torch.int32, device=dev)
        self.acc = torch.zeros(1, dtype=torch.float32, device=dev)
        self.step0 = model.step
        self.use_fused = fused and fused_train_supported(model.spec, batch)
        if self.use_fused:
            # targets of the network output, shuffled with the inputs each epoch; the step index
            # is advanced by the gradient call before Adam reads it
            self.Tn = (T / model.spec.out_scale).contiguous()
            self.k_dev.fill_(-1)
            self.fstep = FusedTrainStep(model, self.Xs, self.Ts, batch, self.base, self.acc, splits=splits,
                                        step=self.k_dev)
        else:
            self.Tn = T
        self.g = None
        self.done = 0

    def _adam(self):
        m = self.model
        N.call("pyftle_mlp_adam_scheduled", _p(m.params), _p(m.m), _p(m.v), _p(m.grads), m.n_params, 0.9, 0.999,
               1e-8, 0.0, float(self.lr), float(self.lr_min), self.total, self.step0, _p(self.k_dev),
               _p(m.params_bf16), stream_handle())

    def _body(self):
        if self.use_fused:
            self.fstep.step(self.lr, self.lr_min, self.total, self.step0)  # also advances base and the step index
            return
        idx = self.base + self.ar
        torch.index_select(self.Xs, 0, idx, out=self.xb)
        torch.index_select(self.Ts, 0, idx, out=self.tb)
        self.model.loss_and_backward(self.xb, self.tb, loss_out=self.acc)
        self._adam()
        self.base.add_(self.batch)

    def begin_epoch(self):
        perm = torch.randperm(self.n, generator=self.gen, device=self.X.device)
        torch.index_select(self.X, 0, perm, out=self.Xs)
        torch.index_select(self.Tn, 0, perm, out=self.Ts)
        self.base.zero_()
        self.acc.zero_()

    def step(self):
        if self.graph and self.g is None and self.done >= 2:
            self.g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.g):
                self._body()
        if self.g is not None:
            self.g.replay()
        else:
            self._body()
        self.done += 1

    def finish(self):
        self.model.step += self.total
        if self.use_fused:
            self.model.sync_bf16()  # the fused steps wrote the packed weights only


def train(model: FlowMapMLP, X: torch.Tensor, T: torch.Tensor, epochs: int = 20, batch: int = 32768,
          lr: float = 2e-3, lr_min: float = 2e-5, seed: int = 0, log=None, graph: bool = False,
          fused: bool = True, splits: int = 0):
    """Adam with a cosine learning-rate schedule over minibatches; returns per-epoch losses.

    Each epoch gathers a shuffled copy of the data once; a step takes the next ``batch``
    rows at a device-side offset, runs forward, backward and the device-scheduled Adam
    update (``pyftle_mlp_adam_scheduled``), so no launch argument changes between steps.
    ``fused`` (default, where ``fused_train_supported``): forward, backward and Adam in the
    fused training kernels (``FusedTrainStep``: one kernel per row tile for the activations,
    the split-K weight-gradient GEMMs, one reduction that also applies Adam), three launches
    a step; otherwise the layer-by-layer kernels (~45 launches).
    With ``graph`` the step is captured in a HIP graph after two eager steps and replayed;
    eager and replayed steps run the same kernels on the same buffers (bit-identical
    weights).  ``splits``: split-K of the fused step's weight-gradient GEMMs (0: its default).
    """
    tr = _Trainer(model, X, T, epochs, batch, lr, lr_min, seed, graph, fused, splits)
    losses = []
    for ep in range(epochs):
        tr.begin_epoch()
        for _ in range(tr.steps_per_epoch):
            tr.step()
        losses.append(tr.acc.item() / tr.steps_per_epoch)
        if log:
            log(f"epoch {ep} loss {losses[-1]:.3e}")
    tr.finish()
    return losses


def train_ensemble(models: list[FlowMapMLP], X: torch.Tensor, T: torch.Tensor, epochs: int = 20,
                   batch: int = 32768, lr: float = 2e-3, lr_min: float = 2e-5, seeds=None, fused: bool = True,
                   splits: int = 0) -> list[list[float]]:
    """``train`` of several models on the same data at once: model k's run is its own
    ``train(models[k], X, T, ..., seed=seeds[k])`` (the same weights bit for bit), with the
    runs' steps interleaved, each run on a stream of its own (default priority), so that one
    run's kernels fill what another's leave idle.  Per-epoch losses stay on the device until
    the end (no synchronization between epochs).  Returns each model's per-epoch losses."""
    seeds = list(range(len(models))) if seeds is None else list(seeds)
    if len(seeds) != len(models):
        raise ValueError("one seed per model")
    dev = X.device
    cur = torch.cuda.current_stream(dev)
    streams = [torch.cuda.Stream(device=dev) for _ in models]
    trainers, hist = [], []
    for m, st, sd in zip(models, streams, seeds):
        st.wait_stream(cur)
        with torch.cuda.stream(st):
            trainers.append(_Trainer(m, X, T, epochs, batch, lr, lr_min, sd, False, fused, splits))
            hist.append(torch.zeros(epochs, dtype=torch.float32, device=dev))
    for ep in range(epochs):
        for tr, st in zip(trainers, streams):
            with torch.cuda.stream(st):
                tr.begin_epoch()
        for _ in range(trainers[0].steps_per_epoch):
            for tr, st in zip(trainers, streams):
                with torch.cuda.stream(st):
                    tr.step()
        for tr, st, h in zip(trainers, streams, hist):
            with torch.cuda.stream(st):
                h[ep:ep + 1].copy_(tr.acc)
    out = []
    for tr, st, h in zip(trainers, streams, hist):
        with torch.cuda.stream(st):
            tr.finish()
            vals = h.cpu().tolist()
        cur.wait_stream(st)
        out.append([v / tr.steps_per_epoch for v in vals])
    return out


class FlowMapEnsemble:
    """Members trained on the same trajectories from different seeds."""

    def __init__(self, members: list[FlowMapMLP]):
        self.members = members

    def predict(self, x: torch.Tensor, dim: int, fused: bool = True, jacobian_members: int | None = None,
                per_member: bool = False):
        """Mean displacement, mean Jacobian of the displacement and member spread.

        The spread is the largest deviation of a member's displacement from the
        ensemble mean (any component).  ``jacobian_members`` (default: all) members
        also push ``dim`` tangents through the network and their Jacobians are
        averaged; the others run the primal rows only.  ``fused`` runs the
        one-kernel network where it covers the shape, the layer-by-layer kernels
        otherwise.
        """
        nj = len(self.members) if jacobian_members is None else max(1, min(jacobian_members, len(self.members)))
        outs, jacs = [], []
        for k, m in enumerate(self.members):
            nt = dim if k < nj else 0
            if fused and m.fused_supported(nt):
                Y = m.forward_fused(x, n_tangents=nt)
                outs.append(Y[0])
                if nt:
                    jacs.append(Y[1:].permute(1, 2, 0))  # (M, d_out, dim): d y_r / d x_c
            elif nt:
                y, tans = m.forward(x, n_tangents=nt)
                outs.append(y.clone())
                jacs.append(torch.stack([t.clone() for t in tans], dim=2))
            else:
                outs.append(m.forward(x, keep_z=False).clone())
        Y = torch.stack(outs)
        Jd = torch.stack(jacs)
        mean = Y.mean(0)
        spread = (Y - mean).abs().amax(dim=(0, 2))
        if per_member:  # the members' displacement Jacobians, (members, M, d_out, dim)
            return mean, Jd, spread
        return mean, Jd.mean(0), spread


def ftle_from_displacement_jacobian(Jd: torch.Tensor, T: float) -> torch.Tensor:
    """FTLE of the flow map x + d(x): J = I + dd/dx, FP64 closed-form kernels."""
    dim = Jd.shape[1]
    J = (Jd.double() + torch.eye(dim, dtype=torch.float64, device=Jd.device)).contiguous()
    out = torch.empty(J.shape[0], dtype=torch.float64, device=J.device)
    N.call("pyftle_ftle_from_jacobian", _p(J), dim, J.shape[0], float(T), _p(out), stream_handle())
    return out


def mlp_flops_per_query(spec: MLPSpec) -> int:
    fan = 2 * spec.n_freq
    f = 0
    for _ in range(spec.depth):
        f += 2 * fan * spec.width
        fan = spec.width
    return f + 2 * spec.width * spec.d_out


def mlp_train_flops_per_row(spec: MLPSpec) -> int:
    """Flops of one training row: the forward pass, the input gradients of every layer but
    the first (whose input, the Fourier features, needs none) and every weight gradient."""
    fwd = mlp_flops_per_query(spec)
    dgrad = fwd - 2 * 2 * spec.n_freq * spec.width
    return fwd + dgrad + fwd


__all__ = ["MLPSpec", "FlowMapMLP", "FlowMapEnsemble", "train", "ftle_from_displacement_jacobian",
           "gemm", "pad128", "mlp_flops_per_query", "mlp_train_flops_per_row"]
