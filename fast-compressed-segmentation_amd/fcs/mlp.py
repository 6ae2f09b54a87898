"""Learned brick-prefetch MLP (include/fcs/fcs_mlp.h) over ctypes."""
import ctypes

import numpy as np

from ._lib import FcsError, c_u32p, c_vp, check, lib

_f32p = ctypes.POINTER(ctypes.c_float)


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)


_sig("fcs_mlp_create", ctypes.c_int, c_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
     ctypes.c_uint64, ctypes.POINTER(c_vp))
_sig("fcs_mlp_destroy", None, c_vp)
_sig("fcs_mlp_num_params", ctypes.c_uint64, c_vp)
_sig("fcs_mlp_get_params", ctypes.c_int, c_vp, _f32p)
_sig("fcs_mlp_set_params", ctypes.c_int, c_vp, _f32p)
_sig("fcs_mlp_get_grads", ctypes.c_int, c_vp, _f32p)
_sig("fcs_mlp_forward", ctypes.c_int, c_vp, _f32p, ctypes.c_uint32, _f32p)
_sig("fcs_mlp_backward", ctypes.c_int, c_vp, _f32p, _f32p, ctypes.c_uint32, _f32p)
_sig("fcs_mlp_adam", ctypes.c_int, c_vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float)
_sig("fcs_mlp_train_masks", ctypes.c_int, c_vp, _f32p, ctypes.c_uint32, c_u32p, c_u32p, ctypes.c_uint32,
     ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_uint64, _f32p)


def _fp(a):
    return a.ctypes.data_as(_f32p)


class MLP:
    """n_in -> n_hidden -> n_hidden -> n_out FP32 network on the context's device."""

    def __init__(self, ctx, n_in, n_hidden, n_out, max_batch=64, seed=0):
        self.ctx = ctx
        self.n_in, self.n_hidden, self.n_out, self.max_batch = n_in, n_hidden, n_out, max_batch
        h = c_vp()
        check(lib.fcs_mlp_create(ctx.h, n_in, n_hidden, n_out, max_batch, seed, ctypes.byref(h)), ctx.h)
        self.h = h
        self.num_params = lib.fcs_mlp_num_params(h)

    def close(self):
        if self.h:
            lib.fcs_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(self.num_params, np.float32)
        check(lib.fcs_mlp_get_params(self.h, _fp(out)), self.ctx.h)
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.num_params:
            raise FcsError(-1, "parameter count mismatch")
        check(lib.fcs_mlp_set_params(self.h, _fp(p)), self.ctx.h)

    def grads(self):
        out = np.zeros(self.num_params, np.float32)
        check(lib.fcs_mlp_get_grads(self.h, _fp(out)), self.ctx.h)
        return out

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.n_in)
        out = np.zeros((x.shape[0], self.n_out), np.float32)
        check(lib.fcs_mlp_forward(self.h, _fp(x), x.shape[0], _fp(out)), self.ctx.h)
        return out

    def backward(self, x, y):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.n_in)
        y = np.ascontiguousarray(y, np.float32).reshape(-1, self.n_out)
        loss = ctypes.c_float()
        check(lib.fcs_mlp_backward(self.h, _fp(x), _fp(y), x.shape[0], ctypes.byref(loss)), self.ctx.h)
        return loss.value

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        check(lib.fcs_mlp_adam(self.h, lr, beta1, beta2, eps), self.ctx.h)

    def train_masks(self, poses, in_bits, tgt_bits, epochs=1, batch=64, lr=1e-3, seed=0):
        """poses: (N, n_pose) f32; in_bits / tgt_bits: (N, words) u32 LSB-first."""
        poses = np.ascontiguousarray(poses, np.float32)
        n, n_pose = poses.shape
        in_bits = np.ascontiguousarray(in_bits, np.uint32)
        tgt_bits = np.ascontiguousarray(tgt_bits, np.uint32)
        losses = np.zeros(epochs, np.float32)
        check(lib.fcs_mlp_train_masks(self.h, _fp(poses), n_pose, in_bits.ctypes.data_as(c_u32p),
                                      tgt_bits.ctypes.data_as(c_u32p), n, epochs, batch, lr, seed, _fp(losses)),
              self.ctx.h)
        return losses


def pack_bits(mask):
    """(N, bits) 0/1 -> (N, ceil(bits/32)) u32, LSB-first."""
    m = np.asarray(mask).astype(bool)
    n, nb = m.shape
    w = (nb + 31) // 32
    pad = np.zeros((n, w * 32), bool)
    pad[:, :nb] = m
    b = pad.reshape(n, w, 32).astype(np.uint64) << np.arange(32, dtype=np.uint64)
    return b.sum(-1).astype(np.uint32)
