"""The surrogate network through the C ABI (poe2d_mlp_*): an ensemble of bf16 fully connected
networks with GELU, trained with Adam on the GPU's matrix cores."""
import ctypes
import weakref

import numpy as np

from . import _abi


def _f(a):
    return a.ctypes.data_as(_abi.c_float_p)


class Network:
    def __init__(self, sizes, members=1, max_batch=256, seed=1):
        self._lib = _abi.load_library()
        self.sizes = [int(s) for s in sizes]
        self.members = int(members)
        self.max_batch = int(max_batch)
        arr = (ctypes.c_int * len(self.sizes))(*self.sizes)
        h = ctypes.c_void_p()
        _abi.check(self._lib, self._lib.poe2d_mlp_create(len(self.sizes) - 1, arr, self.members, self.max_batch,
                                                         ctypes.c_uint64(seed), ctypes.byref(h)), "mlp_create")
        self._h = h

    @property
    def nlayers(self):
        return len(self.sizes) - 1

    def get_params(self, member=0):
        out = []
        for l in range(self.nlayers):
            w = np.zeros((self.sizes[l + 1], self.sizes[l]), np.float32)
            b = np.zeros(self.sizes[l + 1], np.float32)
            _abi.check(self._lib, self._lib.poe2d_mlp_get_params(self._h, member, l, _f(w), _f(b)), "mlp_get_params")
            out.append((w, b))
        return out

    def get_bf16(self, member=0):
        """The bf16 weight copies the products read, per layer: (w16 [out][in], w16t [in][out]) as
        uint16 bit patterns."""
        out = []
        for l in range(self.nlayers):
            w16 = np.zeros((self.sizes[l + 1], self.sizes[l]), np.uint16)
            w16t = np.zeros((self.sizes[l], self.sizes[l + 1]), np.uint16)
            _abi.check(self._lib, self._lib.poe2d_mlp_get_bf16(self._h, member, l, w16.ctypes.data, w16t.ctypes.data),
                       "mlp_get_bf16")
            out.append((w16, w16t))
        return out

    def set_params(self, params, member=0):
        for l, (w, b) in enumerate(params):
            w = np.ascontiguousarray(w, np.float32)
            b = np.ascontiguousarray(b, np.float32)
            _abi.check(self._lib, self._lib.poe2d_mlp_set_params(self._h, member, l, _f(w), _f(b)), "mlp_set_params")

    def forward(self, x, member=0):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], self.sizes[-1]), np.float32)
        _abi.check(self._lib, self._lib.poe2d_mlp_forward(self._h, member, x.shape[0], _f(x), _f(y)), "mlp_forward")
        return y

    def gradients(self, x, t, member=0):
        """(loss, [(dW, db) per layer]) without updating the weights."""
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        gw = [np.zeros((self.sizes[l + 1], self.sizes[l]), np.float32) for l in range(self.nlayers)]
        gb = [np.zeros(self.sizes[l + 1], np.float32) for l in range(self.nlayers)]
        pw = (_abi.c_float_p * self.nlayers)(*[_f(g) for g in gw])
        pb = (_abi.c_float_p * self.nlayers)(*[_f(g) for g in gb])
        loss = ctypes.c_float()
        _abi.check(self._lib, self._lib.poe2d_mlp_gradients(self._h, member, x.shape[0], _f(x), _f(t),
                                                            ctypes.byref(loss), pw, pb), "mlp_gradients")
        return loss.value, list(zip(gw, gb))

    def set_dataset(self, x, t):
        self._dx = np.ascontiguousarray(x, np.float32)
        self._dt = np.ascontiguousarray(t, np.float32)
        _abi.check(self._lib, self._lib.poe2d_mlp_set_dataset(self._h, self._dx.shape[0], _f(self._dx), _f(self._dt)),
                   "mlp_set_dataset")

    def train(self, indices, batch, lr, member=0):
        """Adam steps on dataset rows indices.reshape(-1, batch); returns the losses."""
        idx = np.ascontiguousarray(indices, np.int32).ravel()
        nsteps = idx.size // batch
        losses = np.zeros(nsteps, np.float32)
        _abi.check(self._lib, self._lib.poe2d_mlp_train(self._h, member, nsteps, batch,
                                                        idx.ctypes.data_as(_abi.c_int_p), ctypes.c_float(lr),
                                                        _f(losses)), "mlp_train")
        return losses

    def set_output_scale(self, scale):
        """Per-output factors applied on the device to predicted means (None: none)."""
        if scale is None:
            _abi.check(self._lib, self._lib.poe2d_mlp_set_output_scale(self._h, None), "mlp_set_output_scale")
            return
        s = np.ascontiguousarray(scale, np.float32).ravel()
        if s.size != self.sizes[-1]:
            raise ValueError(f"output scale needs {self.sizes[-1]} values, got {s.size}")
        _abi.check(self._lib, self._lib.poe2d_mlp_set_output_scale(self._h, _f(s)), "mlp_set_output_scale")

    def train_many(self, indices, batch, lr, members):
        """Adam steps of several members at once (two streams), member j on dataset rows
        indices[j].reshape(-1, batch); returns their losses [len(members), steps].  The weights
        come out as train() per member leaves them."""
        members = np.ascontiguousarray(members, np.int32).ravel()
        idx = np.ascontiguousarray(indices, np.int32).reshape(len(members), -1)
        nsteps = idx.shape[1] // batch
        if idx.shape[1] != nsteps * batch:
            raise ValueError("every member needs a whole number of batches")
        losses = np.zeros((len(members), nsteps), np.float32)
        _abi.check(self._lib, self._lib.poe2d_mlp_train_many(
            self._h, len(members), members.ctypes.data_as(_abi.c_int_p), nsteps, batch,
            idx.ctypes.data_as(_abi.c_int_p), ctypes.c_float(lr), _f(losses)), "mlp_train_many")
        return losses

    def predict(self, x, want_mean=True, out=None):
        """(ensemble mean or None, spread per row).  out: where the mean goes, a C-contiguous float32
        [n, outputs] array (page-locked memory from pinned_empty copies fastest) or a DeviceArray of
        n rows of `outputs` values (the mean stays in HBM)."""
        x = np.ascontiguousarray(x, np.float32)
        n = x.shape[0]
        if isinstance(out, DeviceArray):
            if out.shape[0] != n or out.row_size != self.sizes[-1]:
                raise ValueError("out must hold n rows of the network's outputs")
            mean, mean_p = out, ctypes.cast(ctypes.c_void_p(out.ptr), _abi.c_float_p)
        elif out is not None:
            if out.dtype != np.float32 or out.shape != (n, self.sizes[-1]) or not out.flags.c_contiguous:
                raise ValueError("out must be a C-contiguous float32 array of shape (n, outputs)")
            mean, mean_p = out, _f(out)
        else:
            mean = np.zeros((n, self.sizes[-1]), np.float32) if want_mean else None
            mean_p = _f(mean) if want_mean else None
        spread = np.zeros(n, np.float32)
        _abi.check(self._lib, self._lib.poe2d_mlp_predict(self._h, n, _f(x), mean_p, _f(spread)), "mlp_predict")
        return mean, spread

    def time(self, batch, iters=5):
        f, t = ctypes.c_double(), ctypes.c_double()
        _abi.check(self._lib, self._lib.poe2d_mlp_time(self._h, batch, iters, ctypes.byref(f), ctypes.byref(t)),
                   "mlp_time")
        return f.value, t.value

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.poe2d_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def pinned_empty(shape, dtype=np.float32):
    """An uninitialised numpy array in page-locked host memory (poe2d_host_alloc), released when the
    array and every view of it are gone."""
    lib = _abi.load_library()
    dtype = np.dtype(dtype)
    nbytes = int(np.prod(shape)) * dtype.itemsize
    ptr = ctypes.c_void_p()
    _abi.check(lib, lib.poe2d_host_alloc(ctypes.c_size_t(max(nbytes, 1)), ctypes.byref(ptr)), "host_alloc")
    buf = (ctypes.c_char * max(nbytes, 1)).from_address(ptr.value)
    weakref.finalize(buf, lib.poe2d_host_free, ctypes.c_void_p(ptr.value))
    return np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)


class DeviceArray:
    """A C-contiguous float32 array in device memory (poe2d_device_alloc), for results that stay in
    HBM for a consumer on the GPU.  rows(lo, hi) is a view of rows [lo, hi) along the first axis;
    to_host and set_rows move whole rows between it and the host.  The memory is released when the
    array and every view of it are gone."""

    def __init__(self, shape, _base=None, _ptr=None):
        self._lib = _abi.load_library()
        self.shape = tuple(int(d) for d in shape)
        self.dtype = np.dtype(np.float32)
        self.row_size = int(np.prod(self.shape[1:], dtype=np.int64))
        self.nbytes = self.shape[0] * self.row_size * 4
        if _base is not None:
            self._base, self.ptr = _base, _ptr
            return
        ptr = ctypes.c_void_p()
        _abi.check(self._lib, self._lib.poe2d_device_alloc(ctypes.c_size_t(max(self.nbytes, 1)), ctypes.byref(ptr)),
                   "device_alloc")
        self._base, self.ptr = None, ptr.value
        weakref.finalize(self, self._lib.poe2d_device_free, ctypes.c_void_p(self.ptr))

    def __len__(self):
        return self.shape[0]

    def rows(self, lo, hi):
        if not 0 <= lo <= hi <= self.shape[0]:
            raise IndexError("rows out of range")
        return DeviceArray((hi - lo,) + self.shape[1:], _base=self if self._base is None else self._base,
                           _ptr=self.ptr + lo * self.row_size * 4)

    def _row_ptr(self, r):
        if not 0 <= r < self.shape[0]:
            raise IndexError("row out of range")
        return ctypes.c_void_p(self.ptr + int(r) * self.row_size * 4)

    def to_host(self, rows=None):
        """The array (rows None) or the listed rows, as a numpy array."""
        if rows is None:
            out = np.empty(self.shape, np.float32)
            _abi.check(self._lib, self._lib.poe2d_copy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr),
                                                       ctypes.c_size_t(self.nbytes)), "copy")
            return out
        rows = np.asarray(rows, np.int64).reshape(-1)
        out = np.empty((len(rows),) + self.shape[1:], np.float32)
        for i, r in enumerate(rows):
            _abi.check(self._lib, self._lib.poe2d_copy(out[i].ctypes.data_as(ctypes.c_void_p), self._row_ptr(r),
                                                       ctypes.c_size_t(self.row_size * 4)), "copy")
        return out

    def set_rows(self, rows, values):
        """Rows `rows` take values[i] (host arrays of a row's shape)."""
        rows = np.asarray(rows, np.int64).reshape(-1)
        values = np.ascontiguousarray(values, np.float32).reshape(len(rows), self.row_size)
        for i, r in enumerate(rows):
            _abi.check(self._lib, self._lib.poe2d_copy(self._row_ptr(r), values[i].ctypes.data_as(ctypes.c_void_p),
                                                       ctypes.c_size_t(self.row_size * 4)), "copy")


def gemm_bf16(a, b, ksplit=1):
    """a [M][K] @ b[N][K]^T with bf16 operands and FP32 accumulation on the matrix cores."""
    lib = _abi.load_library()
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    c = np.zeros((a.shape[0], b.shape[0]), np.float32)
    _abi.check(lib, lib.poe2d_gemm_bf16(a.shape[0], b.shape[0], a.shape[1], _f(a), _f(b), _f(c), ksplit), "gemm_bf16")
    return c


def gemm_bf16_noalloc(a, b, ksplit=1):
    """As gemm_bf16, through the entry point that allocates no device memory (its operands, product
    and partial sums must fit the library's fixed 32 MiB device scratch)."""
    lib = _abi.load_library()
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    c = np.zeros((a.shape[0], b.shape[0]), np.float32)
    _abi.check(lib, lib.poe2d_gemm_bf16_noalloc(a.shape[0], b.shape[0], a.shape[1], _f(a), _f(b), _f(c), ksplit),
               "gemm_bf16_noalloc")
    return c
