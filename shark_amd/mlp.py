"""Multilayer perceptron of include/shark_mlp.h: configuration struct, parameter layout and the
GPU binding (forward, gradients, Adam training in libshark_gpu.so's own kernels)."""
import ctypes as C

import numpy as np

from . import gpu

SHARK_MLP_VERSION = 1


class MLPCfg(C.Structure):
    _fields_ = [("version", C.c_int32), ("n_in", C.c_int32), ("n_hidden", C.c_int32),
                ("n_layers", C.c_int32), ("n_out", C.c_int32), ("max_batch", C.c_int32),
                ("reserved_i", C.c_int32 * 2), ("beta1", C.c_double), ("beta2", C.c_double),
                ("adam_eps", C.c_double), ("reserved", C.c_double * 8)]


def make_cfg(n_in=69, n_hidden=256, n_layers=4, n_out=65, max_batch=4096, beta1=0.9, beta2=0.999,
             adam_eps=1e-8):
    c = MLPCfg()
    c.version = SHARK_MLP_VERSION
    c.n_in, c.n_hidden, c.n_layers, c.n_out, c.max_batch = n_in, n_hidden, n_layers, n_out, max_batch
    c.beta1, c.beta2, c.adam_eps = beta1, beta2, adam_eps
    return c


def layer_dims(cfg):
    dims = []
    for l in range(cfg.n_layers):
        di = cfg.n_in if l == 0 else cfg.n_hidden
        do = cfg.n_out if l == cfg.n_layers - 1 else cfg.n_hidden
        dims.append((di, do))
    return dims


def num_params(cfg):
    return sum(di * do + do for di, do in layer_dims(cfg))


def unpack(cfg, p):
    """[(W (dout x din), b (dout))] views of a flat parameter vector."""
    out, o = [], 0
    for di, do in layer_dims(cfg):
        W = p[o:o + di * do].reshape(do, di)
        o += di * do
        b = p[o:o + do]
        o += do
        out.append((W, b))
    return out


def init_params(cfg, seed=0):
    """He-normal weights (gain for SiLU ~ ReLU), zero biases, last layer scaled by 0.1."""
    rng = np.random.default_rng(seed)
    parts = []
    dims = layer_dims(cfg)
    for l, (di, do) in enumerate(dims):
        W = rng.standard_normal((do, di)) * np.sqrt(2.0 / di)
        if l == len(dims) - 1:
            W *= 0.1
        parts += [W.ravel(), np.zeros(do)]
    return np.concatenate(parts).astype(np.float32)


_P32 = C.POINTER(C.c_float)


def _lib():
    lib = gpu.load()
    if not getattr(lib, "_mlp_bound", False):
        lib.shark_mlp_create.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        lib.shark_mlp_num_params.restype = C.c_int64
        lib.shark_mlp_num_params.argtypes = [C.c_void_p]
        lib.shark_mlp_set_params.argtypes = [C.c_void_p, _P32]
        lib.shark_mlp_get_params.argtypes = [C.c_void_p, _P32]
        lib.shark_mlp_forward.argtypes = [C.c_void_p, _P32, C.c_int64, _P32]
        lib.shark_mlp_forward_layered.argtypes = [C.c_void_p, _P32, C.c_int64, _P32]
        lib.shark_mlp_grads.argtypes = [C.c_void_p, _P32, _P32, C.c_int64, _P32, _P32]
        lib.shark_mlp_set_data.argtypes = [C.c_void_p, _P32, _P32, C.c_int64]
        lib.shark_mlp_grads_weighted.argtypes = [C.c_void_p, _P32, _P32, _P32, C.c_int64, _P32, _P32]
        lib.shark_mlp_set_data_weighted.argtypes = [C.c_void_p, _P32, _P32, _P32, C.c_int64]
        lib.shark_mlp_train.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_uint64, _P32]
        lib.shark_mlp_kernel_time.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        lib.shark_mlp_destroy.argtypes = [C.c_void_p]
        lib._mlp_bound = True
    return lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class MLP:
    def __init__(self, cfg, params=None):
        self.L = _lib()
        self.cfg = cfg
        h = C.c_void_p()
        gpu.check(self.L.shark_mlp_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.np = self.L.shark_mlp_num_params(h)
        assert self.np == num_params(cfg)
        if params is not None:
            self.set_params(params)

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.np
        gpu.check(self.L.shark_mlp_set_params(self.h, p.ctypes.data_as(_P32)))

    def get_params(self):
        p = np.empty(self.np, dtype=np.float32)
        gpu.check(self.L.shark_mlp_get_params(self.h, p.ctypes.data_as(_P32)))
        return p

    def forward(self, x, layered=False, hilo=False):
        """Network outputs: the fused inference kernel (shark_mlp_forward), with hilo=True its
        variant with BF16 hi + lo activations (shark_mlp_forward_hilo, as the learned growth step
        runs it), or with layered=True the training path's one-product-per-layer forward
        (shark_mlp_forward_layered)."""
        x = _f32(x)
        n = x.shape[0]
        y = np.empty((n, self.cfg.n_out), dtype=np.float32)
        mb = self.cfg.max_batch
        if hilo:
            self.L.shark_mlp_forward_hilo.argtypes = [C.c_void_p, _P32, C.c_int64, _P32]
        fn = (self.L.shark_mlp_forward_layered if layered else
              self.L.shark_mlp_forward_hilo if hilo else self.L.shark_mlp_forward)
        for i in range(0, n, mb):
            xi = np.ascontiguousarray(x[i:i + mb])
            yi = np.empty((xi.shape[0], self.cfg.n_out), dtype=np.float32)
            gpu.check(fn(self.h, xi.ctypes.data_as(_P32), xi.shape[0], yi.ctypes.data_as(_P32)))
            y[i:i + mb] = yi
        return y

    def grads(self, x, t, w=None):
        """Loss and parameter gradient on one batch; w: per-element loss weights (n x n_out)."""
        x, t = _f32(x), _f32(t)
        loss = C.c_float()
        g = np.empty(self.np, dtype=np.float32)
        keep = None if w is None else _f32(w)
        wp = None if keep is None else keep.ctypes.data_as(_P32)
        gpu.check(self.L.shark_mlp_grads_weighted(self.h, x.ctypes.data_as(_P32), t.ctypes.data_as(_P32), wp,
                                                  x.shape[0], C.byref(loss), g.ctypes.data_as(_P32)))
        return loss.value, g

    def set_data(self, x, t, w=None):
        """Training set on the device; w: per-element loss weights (n x n_out) or None."""
        x, t = _f32(x), _f32(t)
        keep = None if w is None else _f32(w)
        wp = None if keep is None else keep.ctypes.data_as(_P32)
        gpu.check(self.L.shark_mlp_set_data_weighted(self.h, x.ctypes.data_as(_P32), t.ctypes.data_as(_P32),
                                                     wp, x.shape[0]))

    def train(self, epochs, batch, lr, seed=0):
        out = np.zeros(max(epochs, 1), dtype=np.float32)
        gpu.check(self.L.shark_mlp_train(self.h, int(epochs), int(batch), float(lr), int(seed),
                                         out.ctypes.data_as(_P32)))
        return out[:epochs]

    def kernel_ms(self):
        ms = C.c_double()
        gpu.check(self.L.shark_mlp_kernel_time(self.h, C.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            self.L.shark_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
