"""ctypes binding of the learned-guess network (lib/libpkmlp.so, include/pk_mlp.h).

Device buffers are torch tensors (float32, cuda) passed by data pointer; torch
only provides the memory and the stream.  There is no CPU path: loading fails
loudly when the library is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libpkmlp.so")
_lib = None


class PkMlpError(RuntimeError):
    pass


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PkMlpError("HIP network library not built: %s is missing (run __graft_entry__.build())" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    vp, ip, fp = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
    lib.pk_mlp_create.argtypes = [ip, ip, ip, ip, ctypes.c_uint64, ctypes.POINTER(vp)]
    lib.pk_mlp_destroy.argtypes = [vp]
    lib.pk_mlp_set_stream.argtypes = [vp, vp]
    lib.pk_mlp_nparams.argtypes = [vp]
    lib.pk_mlp_nparams.restype = ctypes.c_size_t
    lib.pk_mlp_get_params.argtypes = [vp, vp]
    lib.pk_mlp_set_params.argtypes = [vp, vp]
    lib.pk_mlp_get_grads.argtypes = [vp, vp]
    lib.pk_mlp_forward.argtypes = [vp, vp, ip, vp]
    lib.pk_mlp_loss_grad.argtypes = [vp, vp, vp, ip, ctypes.POINTER(fp)]
    lib.pk_mlp_adam_step.argtypes = [vp, fp, fp, fp, fp]
    lib.pk_mlp_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _check(rc, what):
    if rc != 0:
        raise PkMlpError("%s failed (%d): %s" % (what, rc, load().pk_mlp_last_error().decode(errors="replace")))


class Net:
    """n_in -> hidden -> hidden -> n_out, tanh, FP32, on the device."""

    def __init__(self, n_in, n_out, hidden=128, max_batch=65536, seed=0):
        lib = load()
        self.n_in, self.n_out, self.hidden, self.max_batch = n_in, n_out, hidden, max_batch
        self.h = ctypes.c_void_p()
        _check(lib.pk_mlp_create(n_in, hidden, n_out, max_batch, seed, ctypes.byref(self.h)), "pk_mlp_create")
        self.nparams = lib.pk_mlp_nparams(self.h)

    def close(self):
        if self.h:
            load().pk_mlp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_ptr):
        _check(load().pk_mlp_set_stream(self.h, ctypes.c_void_p(stream_ptr)), "pk_mlp_set_stream")

    def params(self):
        out = np.empty(self.nparams, dtype=np.float32)
        _check(load().pk_mlp_get_params(self.h, out.ctypes.data), "pk_mlp_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        _check(load().pk_mlp_set_params(self.h, p.ctypes.data), "pk_mlp_set_params")

    def grads(self):
        out = np.empty(self.nparams, dtype=np.float32)
        _check(load().pk_mlp_get_grads(self.h, out.ctypes.data), "pk_mlp_get_grads")
        return out

    def _device_rows(self, t, ncol, name, rows=None):
        """The kernels take raw device pointers: refuse anything but a
        contiguous float32 HIP tensor [rows][ncol] (a host pointer handed to
        a kernel would be dereferenced on the device)."""
        if not getattr(t, "is_cuda", False):
            raise PkMlpError("%s must be a device (HIP) tensor" % name)
        if str(t.dtype) != "torch.float32":
            raise PkMlpError("%s must be float32, got %s" % (name, t.dtype))
        if t.dim() != 2 or t.shape[1] != ncol:
            raise PkMlpError("%s must have shape [batch][%d], got %s" % (name, ncol, tuple(t.shape)))
        if not t.is_contiguous():
            raise PkMlpError("%s must be contiguous" % name)
        if rows is not None and t.shape[0] != rows:
            raise PkMlpError("%s has %d rows, expected %d" % (name, t.shape[0], rows))
        if t.shape[0] > self.max_batch:
            raise PkMlpError("%s has %d rows, above max_batch %d" % (name, t.shape[0], self.max_batch))

    def forward(self, X, Y):
        """X, Y: contiguous float32 cuda tensors [batch][n_in], [batch][n_out]."""
        self._device_rows(X, self.n_in, "X")
        self._device_rows(Y, self.n_out, "Y", X.shape[0])
        _check(load().pk_mlp_forward(self.h, ctypes.c_void_p(X.data_ptr()), X.shape[0],
                                     ctypes.c_void_p(Y.data_ptr())), "pk_mlp_forward")

    def loss_grad(self, X, T):
        self._device_rows(X, self.n_in, "X")
        self._device_rows(T, self.n_out, "T", X.shape[0])
        loss = ctypes.c_float()
        _check(load().pk_mlp_loss_grad(self.h, ctypes.c_void_p(X.data_ptr()), ctypes.c_void_p(T.data_ptr()),
                                       X.shape[0], ctypes.byref(loss)), "pk_mlp_loss_grad")
        return loss.value

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        _check(load().pk_mlp_adam_step(self.h, lr, beta1, beta2, eps), "pk_mlp_adam_step")
