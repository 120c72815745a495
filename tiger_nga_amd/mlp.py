"""Python side of the learned part: the BF16-MFMA perceptron of
``include/tiger_mlp.h`` (forward, gradients, Adam on the GPU) plus its
parameter layout, initialisation and file format.

Layers: nin -> hidden (GELU) x layers -> nout; every width padded to a
multiple of 64 with zero weights.  Inputs are bfloat16 tensors
[B, in_pad]; outputs float32 [B, out_pad]; targets float32 [B, nout].
"""
from __future__ import annotations

import ctypes

import numpy as np

from .finitechem import load_library


def _round64(x):
    return (x + 63) // 64 * 64


def _bind(lib):
    if getattr(lib, "_mlp_bound", False):
        return lib
    vp, i64 = ctypes.c_void_p, ctypes.c_int64
    fp = ctypes.POINTER(ctypes.c_float)
    lib.tgr_mlp_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
    lib.tgr_mlp_destroy.argtypes = [vp]
    lib.tgr_mlp_layout.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    lib.tgr_mlp_set_params.argtypes = [vp, fp]
    lib.tgr_mlp_get_params.argtypes = [vp, fp]
    lib.tgr_mlp_forward.argtypes = [vp, vp, i64, vp, vp]
    lib.tgr_mlp_gradients.argtypes = [vp, vp, vp, i64, fp, fp, vp]
    lib.tgr_mlp_train_step.argtypes = [vp, vp, vp, i64, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                       ctypes.c_float, fp, vp]
    lib.tgr_mlp_gemm_bf16.argtypes = [vp, i64, i64, vp, i64, i64, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, i64,
                                      vp]
    lib.tgr_mlp_last_error.restype = ctypes.c_char_p
    lib._mlp_bound = True
    return lib


class MLPError(RuntimeError):
    pass


def _check(lib, rc, what):
    if rc != 0:
        raise MLPError("%s failed (%d): %s" % (what, rc, lib.tgr_mlp_last_error().decode()))


def layer_shapes(nin, nout, hidden, layers):
    """[(out_pad, in_pad, out_valid, in_valid)] per linear layer."""
    sh = []
    for l in range(layers + 1):
        fi = _round64(nin) if l == 0 else hidden
        fo = _round64(nout) if l == layers else hidden
        sh.append((fo, fi, nout if l == layers else hidden, nin if l == 0 else hidden))
    return sh


def unflatten(flat, nin, nout, hidden, layers):
    """[(W[out_pad, in_pad], b[out_pad])] views of a flat parameter vector."""
    out, off = [], 0
    for fo, fi, _, _ in layer_shapes(nin, nout, hidden, layers):
        W = flat[off:off + fo * fi].reshape(fo, fi)
        off += fo * fi
        b = flat[off:off + fo]
        off += fo
        out.append((W, b))
    assert off == flat.size
    return out


def init_params(nin, nout, hidden, layers, seed=0):
    """He-normal weights for the GELU layers, a 0.1-scaled Glorot output
    layer, zero biases; padding zero."""
    shapes = layer_shapes(nin, nout, hidden, layers)
    flat = np.zeros(sum(fo * fi + fo for fo, fi, _, _ in shapes), dtype=np.float32)
    rng = np.random.default_rng(seed)
    for l, (W, _) in enumerate(unflatten(flat, nin, nout, hidden, layers)):
        _, _, no, ni = shapes[l]
        std = np.sqrt(2.0 / ni) if l < layers else 0.1 * np.sqrt(2.0 / (ni + no))
        W[:no, :ni] = rng.standard_normal((no, ni)) * std
    return flat


class MLP:
    """The network on the current HIP device."""

    def __init__(self, nin, nout, hidden=256, layers=4):
        self.lib = _bind(load_library())
        self.nin, self.nout, self.hidden, self.layers = nin, nout, hidden, layers
        h = ctypes.c_void_p()
        _check(self.lib, self.lib.tgr_mlp_create(nin, nout, hidden, layers, ctypes.byref(h)), "tgr_mlp_create")
        self._h = h
        n, ip, op = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
        _check(self.lib, self.lib.tgr_mlp_layout(h, ctypes.byref(n), ctypes.byref(ip), ctypes.byref(op)), "layout")
        self.nparam, self.in_pad, self.out_pad = n.value, ip.value, op.value

    def close(self):
        if getattr(self, "_h", None):
            self.lib.tgr_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        assert flat.size == self.nparam
        _check(self.lib, self.lib.tgr_mlp_set_params(self._h, flat.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
               "set_params")

    def get_params(self):
        flat = np.empty(self.nparam, dtype=np.float32)
        _check(self.lib, self.lib.tgr_mlp_get_params(self._h, flat.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
               "get_params")
        return flat

    def _x(self, x):
        import torch
        assert x.dtype == torch.bfloat16 and x.is_cuda and x.is_contiguous() and x.shape[1] == self.in_pad
        return x

    def forward(self, x, out=None, stream=None):
        import torch
        x = self._x(x)
        if out is None:
            out = torch.empty((x.shape[0], self.out_pad), dtype=torch.float32, device=x.device)
        _check(self.lib, self.lib.tgr_mlp_forward(self._h, x.data_ptr(), x.shape[0], out.data_ptr(), stream),
               "forward")
        return out

    def gradients(self, x, tgt, stream=None):
        x = self._x(x)
        assert tgt.is_contiguous() and tuple(tgt.shape) == (x.shape[0], self.nout)
        g = np.empty(self.nparam, dtype=np.float32)
        loss = ctypes.c_float()
        _check(self.lib, self.lib.tgr_mlp_gradients(self._h, x.data_ptr(), tgt.data_ptr(), x.shape[0],
                                                   g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                   ctypes.byref(loss), stream), "gradients")
        return g, loss.value

    def train_step(self, x, tgt, lr, beta1=0.9, beta2=0.999, eps=1e-8, want_loss=False, stream=None):
        x = self._x(x)
        assert tgt.is_contiguous() and tuple(tgt.shape) == (x.shape[0], self.nout)
        loss = ctypes.c_float()
        _check(self.lib, self.lib.tgr_mlp_train_step(self._h, x.data_ptr(), tgt.data_ptr(), x.shape[0], lr, beta1,
                                                    beta2, eps, ctypes.byref(loss) if want_loss else None, stream),
               "train_step")
        return loss.value if want_loss else None

    def gemm(self, A, sam, sak, B, sbk, sbn, M, N, K, C, ldc, stream=None):
        _check(self.lib, self.lib.tgr_mlp_gemm_bf16(A.data_ptr(), sam, sak, B.data_ptr(), sbk, sbn, M, N, K,
                                                   C.data_ptr(), ldc, stream), "gemm_bf16")
