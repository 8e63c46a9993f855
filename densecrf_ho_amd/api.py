"""Host-side mirror of the reference's C++ interface (names, argument meaning
and exceptions of include/permutohedral.h, include/pairwise.h,
include/labelcompatibility.h and include/densecrf.h), backed by the C ABI.

Arrays are either numpy arrays (host; the call stages them to the device) or
torch tensors on the GPU (used in place).  Fields are (N, L) float32 - the
reference's L x N column-major MatrixXf - features (N, d) float32, labellings
(N,) int32.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

NO_NORMALIZATION, NORMALIZE_BEFORE, NORMALIZE_AFTER, NORMALIZE_SYMMETRIC = 0, 1, 2, 3


class EngineError(RuntimeError):
    """std::runtime_error of the reference (message kept) or an engine failure."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


def _check(rc: int):
    if rc != 0:
        raise EngineError(rc, _lib.load().dcrf_last_error().decode())


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


class DeviceArray:
    """A buffer in device memory allocated by the engine's own HIP runtime
    (dcrf_malloc): inputs kept resident in HBM without another GPU library."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _check(_lib.load().dcrf_malloc(C.byref(p), self.nbytes))
        self.ptr = p

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        _check(_lib.load().dcrf_memcpy(d.ptr, a.ctypes.data_as(C.c_void_p), d.nbytes, 0))
        return d

    def numpy(self):
        a = np.empty(self.shape, self.dtype)
        _check(_lib.load().dcrf_memcpy(a.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, 1))
        return a

    def copy_from(self, other: "DeviceArray"):
        _check(_lib.load().dcrf_memcpy(self.ptr, other.ptr, min(self.nbytes, other.nbytes), 2))

    def __del__(self):
        try:
            if self.ptr:
                _lib.load().dcrf_free(self.ptr)
                self.ptr = None
        except Exception:
            pass


def synchronize():
    _check(_lib.load().dcrf_device_synchronize())


def profile_enable(on: bool = True):
    _check(_lib.load().dcrf_profile_enable(int(on)))


def profile_filter(prefix: str = ""):
    """time only the launches whose name starts with `prefix` ("" for all)"""
    _check(_lib.load().dcrf_profile_filter(prefix.encode() if prefix else None))


def profile_report() -> dict:
    """{kernel name: (launch count, total ms)} since profile_enable()."""
    L = _lib.load()
    n = L.dcrf_profile_report(None, 0)
    buf = C.create_string_buffer(max(n, 1))
    L.dcrf_profile_report(buf, n)
    out = {}
    for line in buf.value.decode().splitlines():
        name, cnt, ms = line.rsplit(" ", 2)
        out[name] = (int(cnt), float(ms))
    return out


def set_reference_path(portable: bool):
    """Follow the reference's portable build (True: scalar Permutohedral::init
    and seqCompute, src/permutohedral.cpp:334-540) or its SSE2 build (False,
    the default).  Lattices built afterwards use it; built ones keep theirs."""
    _check(_lib.load().dcrf_set_reference_path(1 if portable else 0))


def reference_path() -> bool:
    """True while new lattices follow the reference's portable build."""
    return bool(_lib.load().dcrf_reference_path())


def autotune_enable(on: bool = True):
    _check(_lib.load().dcrf_autotune_enable(int(on)))


def autotune_report() -> dict:
    """{kernel@size_class: (grid, ms, {candidate grid: ms})} of the launch-shape autotuner."""
    L = _lib.load()
    n = L.dcrf_autotune_report(None, 0)
    buf = C.create_string_buffer(max(n, 1))
    L.dcrf_autotune_report(buf, n)
    out = {}
    for line in buf.value.decode().splitlines():
        head, trail = line.split("|", 1)
        key, grid, ms = head.split()
        cands = {}
        for t in trail.split():
            g, m = t.split(":")
            cands[int(g)] = float(m)
        out[key] = (int(grid), float(ms), cands)
    return out


class _Arg:
    """Pointer + on_device flag for an input or output array."""

    def __init__(self, x, dtype, out=False):
        if isinstance(x, DeviceArray):
            if x.dtype != np.dtype(dtype):
                raise TypeError(f"device array must be {np.dtype(dtype)}")
            self.arr = x
            self.ptr = x.ptr
            self.dev = 1
        elif _is_torch(x):
            import torch
            want = {np.float32: torch.float32, np.int32: torch.int32, np.uint8: torch.uint8}[dtype]
            if x.dtype != want or not x.is_contiguous() or not x.is_cuda:
                raise TypeError(f"device tensor must be a contiguous {want} CUDA tensor")
            self.arr = x
            self.ptr = C.c_void_p(x.data_ptr())
            self.dev = 1
        else:
            a = np.asarray(x)
            if out:
                if a.dtype != dtype or not a.flags.c_contiguous:
                    raise TypeError(f"output array must be a contiguous {np.dtype(dtype)} array")
            else:
                a = np.ascontiguousarray(a, dtype=dtype)
            self.arr = a
            self.ptr = a.ctypes.data_as(C.c_void_p)
            self.dev = 0


def _restricted_args(rQ, pI, Q, store, rout, nrows_out):
    """arguments of the restricted ordered filter, all on one side (host or device)"""
    a = _Arg(rQ, np.float32)
    ix = _Arg(pI, np.int32)
    e = _Arg(Q, np.float32) if (store and Q is not None) else None
    if rout is None:
        rout = _empty_like_field(rQ, nrows_out)
    o = _Arg(rout, np.float32, out=True)
    devs = {a.dev, ix.dev, o.dev} | ({e.dev} if e is not None else set())
    if len(devs) != 1:
        raise TypeError("rQ, pI, Q and rout must all be host arrays or all device arrays")
    return a, ix, e, o, rout


def _empty_like_field(x, shape, dtype=np.float32):
    if isinstance(x, DeviceArray):
        return DeviceArray(shape, dtype)
    if _is_torch(x):
        import torch
        tdt = {np.float32: torch.float32, np.int32: torch.int32}[dtype]
        return torch.empty(shape, dtype=tdt, device=x.device)
    return np.empty(shape, dtype)


class PottsCompatibility:
    """PottsCompatibility(w): mu(a,b) = -w [a==b] (include/labelcompatibility.h:51-63)."""

    def __init__(self, weight: float = 1.0):
        self.w = float(weight)

    kind = 0

    def params(self):
        return np.array([self.w], np.float32)

    def parameters(self):
        return np.array([self.w], np.float32)


class DiagonalCompatibility:
    """DiagonalCompatibility(v): out = diag(v) Q (src/labelcompatibility.cpp:78-96)."""

    kind = 1

    def __init__(self, v):
        self.v = np.ascontiguousarray(v, np.float32).reshape(-1)

    def params(self):
        return self.v


class MatrixCompatibility:
    """MatrixCompatibility(m): out = W Q with W = (m + m^T)/2
    (src/labelcompatibility.cpp:97-129); m is (L, L), m[i, j] row i column j."""

    kind = 2

    def __init__(self, m):
        self.m = np.asarray(m, np.float32)

    def params(self):
        return np.ascontiguousarray(self.m.T).reshape(-1)  # column-major, as Eigen stores it


def load_unary(path: str):
    """load_unary (src/file_storage.cpp:262-291): the .c_unary file as -log
    probabilities; returns (U (N, L) float32, W, H)."""
    L = _lib.load()
    W, H, M = C.c_int(), C.c_int(), C.c_int()
    _check(L.dcrf_io_load_unary(path.encode(), None, 0, C.byref(W), C.byref(H), C.byref(M)))
    U = np.empty((W.value * H.value, M.value), np.float32)
    _check(L.dcrf_io_load_unary(path.encode(), U.ctypes.data_as(C.c_void_p), U.size, C.byref(W), C.byref(H),
                                C.byref(M)))
    return U, W.value, H.value


def load_image(path: str):
    """load_image (src/file_storage.cpp:197-222) for uncompressed BMP: returns
    (RGB uint8 (N*3,), W, H), pixel (x, y) at (x + y W) * 3."""
    L = _lib.load()
    W, H = C.c_int(), C.c_int()
    _check(L.dcrf_io_load_image(path.encode(), None, 0, C.byref(W), C.byref(H)))
    im = np.empty(W.value * H.value * 3, np.uint8)
    _check(L.dcrf_io_load_image(path.encode(), im.ctypes.data_as(C.c_void_p), im.size, C.byref(W), C.byref(H)))
    return im, W.value, H.value


def save_map(Q, W: int, H: int, path: str):
    """save_map(Q, size, path, "MSRC"): argmax labels in the MSRC colours (PNG, or PPM)."""
    q = np.ascontiguousarray(Q, np.float32)
    _check(_lib.load().dcrf_io_save_map(q.ctypes.data_as(C.c_void_p), W, H, q.shape[1], path.encode()))


def less_confident_pixels(Q, tol: float = 0.99):
    """less_confident_pixels (src/densecrf.cpp:457-462): ascending indices of the
    pixels whose largest label value in Q (N, L) is <= tol"""
    q = np.ascontiguousarray(Q, np.float32)
    idx = np.zeros(q.shape[0], np.int32)
    n = C.c_int()
    _check(_lib.load().dcrf_less_confident_pixels(q.ctypes.data_as(C.c_void_p), q.shape[0], q.shape[1], tol,
                                                  idx.ctypes.data_as(C.c_void_p), C.byref(n), 0))
    return idx[: n.value]


def superpixel_terms(img, region, constant: float = 50.0, normaliser: float = 1000.0):
    """the per-region terms of DenseCRF2D::addSuperPixel for a given
    segmentation: (exp_of_superpixels (R,), constants (R,))"""
    im = np.ascontiguousarray(img, np.uint8).ravel()
    reg = np.ascontiguousarray(region, np.int32).ravel()
    R = int(reg.max()) + 1
    e = np.zeros(R, np.float32)
    c = np.zeros(R, np.float32)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    _check(_lib.load().dcrf_superpixel_terms(p(im), len(reg), p(reg), R, constant, normaliser, p(e), p(c)))
    return e, c


class Permutohedral:
    """Permutohedral lattice (include/permutohedral.h:50-88)."""

    def __init__(self):
        self._h = None
        self.N = self.d = self.M = 0

    def init(self, features):
        """Permutohedral::init(features); features (N, d)."""
        L = _lib.load()
        self.close()
        a = _Arg(features, np.float32)
        shp = tuple(a.arr.shape)
        if len(shp) != 2:
            raise ValueError("features must be (N, d)")
        h = C.c_void_p()
        _check(L.dcrf_lattice_create(a.ptr, shp[0], shp[1], a.dev, C.byref(h)))
        self._h = h
        self.N, self.d = shp
        M, Np = C.c_int(), C.c_int()
        _check(L.dcrf_lattice_info(self._h, C.byref(M), C.byref(Np)))
        self.M, self.Npad = M.value, Np.value
        return self

    def close(self):
        if self._h is not None:
            _lib.load().dcrf_lattice_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def compute(self, x, reverse: bool = False, out=None):
        """Permutohedral::compute(out, in, reverse); x (N, V)."""
        V = 1 if len(x.shape) == 1 else x.shape[1]
        a = _Arg(x, np.float32)
        if out is None:
            out = _empty_like_field(x, tuple(x.shape))
        o = _Arg(out, np.float32, out=True)
        _check(_lib.load().dcrf_lattice_compute(self._h, a.ptr, o.ptr, V, int(reverse), a.dev))
        return out

    def compute_upper_minus_lower_ord(self, Q, out=None):
        """Permutohedral::compute_upper_minus_lower_ord(out, Q); Q (N, V) in [0,1]."""
        a = _Arg(Q, np.float32)
        if out is None:
            out = _empty_like_field(Q, tuple(Q.shape))
        o = _Arg(out, np.float32, out=True)
        _check(_lib.load().dcrf_lattice_compute_upper_minus_lower_ord(self._h, a.ptr, o.ptr, Q.shape[1], a.dev))
        return out

    def compute_upper_minus_lower_ord_restricted(self, rout, rQ, pI, Q=None, store: bool = True):
        """Permutohedral::compute_upper_minus_lower_ord_restricted(rout, rQ, pI, Q, store):
        the ordered filter at the pixels pI (strictly ascending) with values rQ
        (len(pI), V); the other pixels take Q (N, V), read only with store.
        rout (len(pI), V) is added to (None: start from zeros)."""
        if rout is None:
            if _is_torch(rQ) or isinstance(rQ, DeviceArray):
                raise TypeError("pass rout for device inputs (it is added to)")
            rout = np.zeros(tuple(rQ.shape), np.float32)
        a, ix, e, o, rout = _restricted_args(rQ, pI, Q, store, rout, tuple(rQ.shape))
        _check(_lib.load().dcrf_lattice_compute_upper_minus_lower_ord_restricted(
            self._h, o.ptr, a.ptr, ix.ptr, len(pI), e.ptr if e else None, rQ.shape[1], int(store), a.dev))
        return rout

    def compute_upper_minus_lower_dc(self, out, low, middle_low, middle_high, high):
        """Permutohedral::compute_upper_minus_lower_dc: one value per pixel, out (N,) added to."""
        o = _Arg(out, np.float32, out=True)
        _check(_lib.load().dcrf_lattice_compute_upper_minus_lower_dc(self._h, o.ptr, low, middle_low, middle_high,
                                                                       high, o.dev))
        return out

    def gradient(self, a, b):
        """Permutohedral::gradient(df, a, b, V): (N, d) = d/df a^T K b for a, b (N, V)."""
        x = _Arg(a, np.float32)
        y = _Arg(b, np.float32)
        df = _empty_like_field(a, (self.N, self.d))
        o = _Arg(df, np.float32, out=True)
        _check(_lib.load().dcrf_lattice_gradient(self._h, o.ptr, x.ptr, y.ptr, a.shape[1], x.dev))
        return df

    def tables(self):
        """offset_, rank_, barycentric_ (Npad, d+1), keys (M, d), neighbours (d+1, M, 2)."""
        d1 = self.d + 1
        off = np.zeros((self.Npad, d1), np.int32)
        rk = np.zeros((self.Npad, d1), np.int32)
        bc = np.zeros((self.Npad, d1), np.float32)
        keys = np.zeros((max(self.M, 1), self.d), np.int16)
        nbr = np.zeros((d1, max(self.M, 1), 2), np.int32)
        p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        _check(_lib.load().dcrf_lattice_export(self._h, p(off), p(rk), p(bc), p(keys), p(nbr)))
        return dict(offset=off, rank=rk, bary=bc, keys=keys[: self.M], nbr=nbr[:, : self.M])


CONST_KERNEL, DIAG_KERNEL, FULL_KERNEL = 0, 1, 2


class PairwisePotential:
    """PairwisePotential (include/pairwise.h:63-95) with a Potts, Diagonal or
    Matrix compatibility, standalone (its own lattice); kernel_type defaults to
    CONST_KERNEL as in the reference's constructor."""

    def __init__(self, features, function: "PottsCompatibility", kernel_type=CONST_KERNEL,
                 normalization_type=3):
        L = _lib.load()
        a = _Arg(features, np.float32)
        N, d = tuple(a.arr.shape)
        h = C.c_void_p()
        w = float(function.w) if function.kind == 0 else 1.0
        _check(L.dcrf_pairwise_create(a.ptr, N, d, w, int(normalization_type), a.dev, C.byref(h)))
        self._h = h
        self.N, self.d = N, d
        if function.kind != 0:
            par = np.ascontiguousarray(function.params(), np.float32)
            nl = par.size if function.kind == 1 else int(round(np.sqrt(par.size)))
            _check(L.dcrf_pairwise_set_compatibility(self._h, function.kind, par.ctypes.data_as(C.c_void_p), nl))
        _check(L.dcrf_pairwise_set_kernel_type(self._h, int(kernel_type)))

    def close(self):
        if getattr(self, "_h", None) is not None:
            try:
                _lib.load().dcrf_pairwise_destroy(self._h)
            except Exception:
                pass
            self._h = None

    def __del__(self):
        self.close()

    def _field(self, Q, fn, *extra):
        a = _Arg(Q, np.float32)
        out = _empty_like_field(Q, tuple(Q.shape))
        o = _Arg(out, np.float32, out=True)
        _check(fn(self._h, a.ptr, o.ptr, Q.shape[1], *extra, a.dev))
        return out

    def apply(self, Q):
        return self._field(Q, _lib.load().dcrf_pairwise_apply, 0)

    def applyTranspose(self, Q):
        return self._field(Q, _lib.load().dcrf_pairwise_apply, 1)

    def apply_upper_minus_lower_ord(self, Q):
        return self._field(Q, _lib.load().dcrf_pairwise_apply_upper_minus_lower_ord)

    def apply_upper_minus_lower_ord_restricted(self, rQ, pI, Q=None, store: bool = True):
        """(len(pI), L) = -w * restricted ordered filter (src/pairwise.cpp:229-239)."""
        a, ix, e, o, out = _restricted_args(rQ, pI, Q, store, None, tuple(rQ.shape))
        _check(_lib.load().dcrf_pairwise_apply_upper_minus_lower_ord_restricted(
            self._h, o.ptr, a.ptr, ix.ptr, len(pI), e.ptr if e else None, rQ.shape[1], int(store), a.dev))
        return out

    def apply_bf(self, Q):
        return self._field(Q, _lib.load().dcrf_pairwise_apply_bf)

    def _sorted(self, fn, ind, *q):
        ix = _Arg(ind, np.int32)
        out = _empty_like_field(ind, tuple(ind.shape))
        o = _Arg(out, np.float32, out=True)
        qa = [_Arg(x, np.float32).ptr for x in q]
        _check(fn(self._h, ix.ptr, *qa, o.ptr, ind.shape[1], ix.dev))
        return out

    def apply_upper_minus_lower_bf(self, ind):
        return self._sorted(_lib.load().dcrf_pairwise_apply_upper_minus_lower_bf, ind)

    def apply_upper_minus_lower_bf_ord(self, ind, Q):
        return self._sorted(_lib.load().dcrf_pairwise_apply_upper_minus_lower_bf_ord, ind, Q)

    def apply_upper_minus_lower_dc(self, ind):
        return self._sorted(_lib.load().dcrf_pairwise_apply_upper_minus_lower_dc, ind)

    def kernelParameters(self):
        n = _lib.load().dcrf_pairwise_kernel_nparams(self._h)
        v = np.zeros(max(n, 1), np.float32)
        _check(_lib.load().dcrf_pairwise_kernel_parameters(self._h, v.ctypes.data_as(C.c_void_p)))
        return v[:n]

    def setKernelParameters(self, v):
        v = np.ascontiguousarray(v, np.float32)
        _check(_lib.load().dcrf_pairwise_set_kernel_parameters(self._h, v.ctypes.data_as(C.c_void_p), len(v)))

    def kernelGradient(self, b, Q):
        """PairwisePotential::kernelGradient(b, Q): gradient of b^T K C(Q) w.r.t. the kernel parameters."""
        n = _lib.load().dcrf_pairwise_kernel_nparams(self._h)
        x = _Arg(b, np.float32)
        q = _Arg(Q, np.float32)
        out = np.zeros(max(n, 1), np.float32)
        _check(_lib.load().dcrf_pairwise_kernel_gradient(self._h, x.ptr, q.ptr, Q.shape[1],
                                                         out.ctypes.data_as(C.c_void_p), x.dev))
        return out[:n]


class LP_inf_params:
    """LP_inf_params (include/densecrf.h:43-69); defaults of src/densecrf.cpp:287-302."""

    def __init__(self, prox_reg_const=1e-2, dual_gap_tol=1e3, prox_energy_tol=1e3, prox_max_iter=10,
                 fw_max_iter=10, qp_max_iter=1000, qp_tol=1e3, qp_const=1e-16, best_int=True, accel_prox=True,
                 work_set_size=10, approx_fw_iter=10, less_confident_percent=0.0, confidence_tol=0.95):
        self.__dict__.update(locals())
        del self.__dict__["self"]

    def _c(self):
        p = _lib.LPParams()
        for name, _ in _lib.LPParams._fields_:
            setattr(p, name, type(getattr(p, name))(getattr(self, name)))
        return p


class UnaryEnergy:
    """UnaryEnergy (include/unary.h:31-39).  get() is the unary as an (N, L)
    array (the memory of the reference's L x N matrix); parameters, setParameters
    and gradient are empty unless a subclass overrides them.  A DenseCRF puts
    the unary on the device with _upload(), again before each computation
    unless `versioned` says that `version` tells every change (then only when it
    moved): a subclass whose get() changes on its own is read afresh each time,
    as the reference calls get() in each solve."""

    versioned = False
    version = 0

    def get(self):
        raise NotImplementedError

    def parameters(self):
        return np.zeros(0, np.float32)

    def setParameters(self, v):
        pass

    def gradient(self, b):
        return np.zeros(0, np.float32)

    def _upload(self, crf):
        a = _Arg(self.get(), np.float32)
        _check(_lib.load().dcrf_crf_set_unary(crf._h, a.ptr, a.dev))


class ConstUnaryEnergy(UnaryEnergy):
    """ConstUnaryEnergy (include/unary.h:40-46): a fixed (N, L) unary (host
    array, DeviceArray or torch tensor)."""

    versioned = True

    def __init__(self, unary):
        self.unary = unary

    def get(self):
        return self.unary

    def _upload(self, crf):
        a = _Arg(self.unary, np.float32)
        _check(_lib.load().dcrf_crf_set_unary(crf._h, a.ptr, a.dev))


class LogisticUnaryEnergy(UnaryEnergy):
    """LogisticUnaryEnergy(L, f) (include/unary.h:47-57, src/unary.cpp:53-68):
    get() = L f on the GPU's matrix cores.  As every array of this API, the
    arguments are the transposes of the reference's matrices with the same
    memory: Lw is (F, L) (the reference's L x F), f is (N, F) (its F x N);
    get() returns (N, L).  parameters() is Lw flattened (the reference's L
    resized to L*F x 1) and gradient(b) is b f^T flattened likewise."""

    versioned = True

    def __init__(self, Lw, f):
        self.Lw = np.ascontiguousarray(Lw, np.float32)
        self.f = f
        F = self.Lw.shape[0]
        if tuple(self.f.shape)[1] != F:
            raise ValueError("LogisticUnaryEnergy: f must be (N, F) with F = Lw.shape[0]")
        self.version = 0

    def _dims(self):
        F, L = self.Lw.shape
        return L, F, int(tuple(self.f.shape)[0])

    def get(self, out=None):
        L, F, N = self._dims()
        x = _Arg(self.f, np.float32)
        out = _empty_like_field(self.f, (N, L)) if out is None else out
        o = _Arg(out, np.float32, out=True)
        w = self.Lw
        if x.dev:  # device features: the parameters go to the device as well
            w = DeviceArray.from_numpy(self.Lw)
        _check(_lib.load().dcrf_logistic_unary(w.ctypes.data_as(C.c_void_p) if not x.dev else w.ptr, x.ptr, L, F, N,
                                               o.ptr, x.dev))
        return out

    def parameters(self):
        return self.Lw.reshape(-1).copy()

    def setParameters(self, v):
        v = np.ascontiguousarray(v, np.float32).reshape(-1)
        if v.size != self.Lw.size:
            raise ValueError("LogisticUnaryEnergy.setParameters: wrong size")
        self.Lw = v.reshape(self.Lw.shape).copy()
        self.version += 1

    def gradient(self, b):
        L, F, N = self._dims()
        x = _Arg(b, np.float32)
        y = _Arg(self.f, np.float32)
        if x.dev != y.dev:
            raise TypeError("b and f must both be host arrays or both device buffers")
        if x.dev:
            g = DeviceArray((F * L,))
            _check(_lib.load().dcrf_logistic_unary_gradient(x.ptr, y.ptr, L, F, N, g.ptr, 1))
            return g.numpy()
        g = np.zeros(F * L, np.float32)
        _check(_lib.load().dcrf_logistic_unary_gradient(x.ptr, y.ptr, L, F, N, g.ctypes.data_as(C.c_void_p), 0))
        return g

    def _upload(self, crf):
        L, F, N = self._dims()
        x = _Arg(self.f, np.float32)
        w = DeviceArray.from_numpy(self.Lw) if x.dev else self.Lw
        _check(_lib.load().dcrf_crf_set_unary_logistic(crf._h, w.ptr if x.dev else w.ctypes.data_as(C.c_void_p),
                                                       x.ptr, F, x.dev))


class DenseCRF:
    """DenseCRF (include/densecrf.h:73-170) on the GPU engine."""

    def __init__(self, N: int, M: int):
        L = _lib.load()
        self.N, self.M = int(N), int(M)
        h = C.c_void_p()
        _check(L.dcrf_crf_create(self.N, self.M, C.byref(h)))
        self._h = h
        self.npairwise = 0
        self._pw_dims = []  # feature dimension of each pairwise term
        self.R = 0
        self._unary = None
        self._uploaded = (None, 0)  # (unary object, version) on the device

    def close(self):
        if getattr(self, "_h", None) is not None:
            _lib.load().dcrf_crf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- model -------------------------------------------------------------------
    def setUnaryEnergy(self, unary, f=None):
        """setUnaryEnergy (src/densecrf.cpp:255-266): an (N, L) array (a
        ConstUnaryEnergy), a UnaryEnergy object, or (Lw, f) for a
        LogisticUnaryEnergy(Lw, f)"""
        if f is not None:
            unary = LogisticUnaryEnergy(unary, f)
        elif not isinstance(unary, UnaryEnergy):
            unary = ConstUnaryEnergy(unary)
        self._unary = unary
        self._uploaded = (None, 0)
        self._sync_unary()

    def getUnaryEnergy(self):
        return self._unary

    def _sync_unary(self):
        u = self._unary
        if u is None:
            return
        if self._uploaded[0] is u and u.versioned and self._uploaded[1] == u.version:
            return
        u._upload(self)
        self._uploaded = (u, u.version)

    def addPairwiseEnergy(self, features, function=None, kernel_type=1, normalization_type=NORMALIZE_SYMMETRIC):
        """addPairwiseEnergy(features (N, d), LabelCompatibility, kernel type, normalisation),
        or addPairwiseEnergy(PairwisePotential) (src/densecrf.cpp:161-168): the
        potential's term is added and the object is consumed, as the reference
        takes ownership of it."""
        if isinstance(features, PairwisePotential):
            pw = features
            _check(_lib.load().dcrf_crf_add_pairwise_potential(self._h, pw._h))
            self.npairwise += 1
            self._pw_dims.append(pw.d)
            pw.close()
            return
        function = function or PottsCompatibility(1.0)
        a = _Arg(features, np.float32)
        d = a.arr.shape[1]
        if function.kind == 0:
            _check(_lib.load().dcrf_crf_add_pairwise(self._h, a.ptr, d, function.w, int(normalization_type), a.dev))
        else:
            par = np.ascontiguousarray(function.params(), np.float32)
            _check(_lib.load().dcrf_crf_add_pairwise_compat(self._h, a.ptr, d, function.kind,
                                                            par.ctypes.data_as(C.c_void_p),
                                                            int(normalization_type), a.dev))
        _check(_lib.load().dcrf_crf_pairwise_set_kernel_type(self._h, self.npairwise, int(kernel_type)))
        self.npairwise += 1
        self._pw_dims.append(d)

    def set_value_type(self, bf16: bool):
        """ordered-filter lattice rows of the terms added so far in bfloat16 (True) or float"""
        _check(_lib.load().dcrf_crf_set_value_type(self._h, int(bool(bf16))))

    def pairwise_parameters(self, k):
        """PairwisePotential::parameters() of potential k."""
        n = C.c_int()
        _check(_lib.load().dcrf_crf_pairwise_nparams(self._h, k, C.byref(n)))
        out = np.zeros(n.value, np.float32)
        _check(_lib.load().dcrf_crf_pairwise_parameters(self._h, k, out.ctypes.data_as(C.c_void_p)))
        return out

    def set_pairwise_parameters(self, k, v):
        """PairwisePotential::setParameters(v) of potential k."""
        v = np.ascontiguousarray(v, np.float32)
        if v.size != self.pairwise_parameters(k).size:
            raise ValueError("wrong number of parameters")
        _check(_lib.load().dcrf_crf_pairwise_set_parameters(self._h, k, v.ctypes.data_as(C.c_void_p)))

    def pairwise_gradient(self, k, b, Q):
        """PairwisePotential::gradient(b, Q) of potential k (parameters() layout)."""
        g = np.zeros(self.pairwise_parameters(k).size, np.float32)
        a, q = _Arg(b, np.float32), _Arg(Q, np.float32)
        if a.dev != q.dev:
            raise TypeError("b and Q must both be host arrays or both device buffers")
        _check(_lib.load().dcrf_crf_pairwise_gradient(self._h, k, a.ptr, q.ptr, g.ctypes.data_as(C.c_void_p), a.dev))
        return g

    # ---- unrolled mean field and its adjoint ------------------------------------
    def meanfield(self, iterations: int, out=None, device=None):
        """Q after `iterations` mean-field updates from softmax(-U); keeps the
        per-iteration checkpoints for meanfield_backward."""
        self._sync_unary()
        if out is None:
            out = np.empty((self.N, self.M), np.float32) if device is None else _torch_empty((self.N, self.M), device)
        o = _Arg(out, np.float32, out=True)
        _check(_lib.load().dcrf_crf_meanfield(self._h, int(iterations), o.ptr, o.dev))
        return out

    def meanfield_id(self) -> int:
        """number of the last meanfield forward on this CRF (0: none yet)"""
        n = C.c_longlong()
        _check(_lib.load().dcrf_crf_meanfield_id(self._h, C.byref(n)))
        return n.value

    def meanfield_backward(self, dQ, out=None):
        """dL/dQ_T -> (dL/dU, [dL/dparameters of each potential] (FP64)) of the
        last meanfield; raises if the unary or a pairwise term changed since."""
        a = _Arg(dQ, np.float32)
        out = _empty_like_field(dQ, (self.N, self.M)) if out is None else out
        o = _Arg(out, np.float32, out=True)
        sizes = [self.pairwise_parameters(k).size for k in range(self.npairwise)]
        g = np.zeros(max(sum(sizes), 1))
        _check(_lib.load().dcrf_crf_meanfield_backward(self._h, a.ptr, o.ptr, g.ctypes.data_as(C.c_void_p), a.dev))
        parts, off = [], 0
        for n in sizes:
            parts.append(g[off:off + n].copy())
            off += n
        return out, parts

    def cross_entropy(self, Q, labels, out=None):
        """(-sum_i log Q(y_i, i) over labelled pixels, dloss/dQ)."""
        a = _Arg(Q, np.float32)
        lab = _Arg(labels, np.int32)
        if a.dev != lab.dev:
            raise TypeError("Q and labels must both be host arrays or both device buffers")
        out = _empty_like_field(Q, (self.N, self.M)) if out is None else out
        o = _Arg(out, np.float32, out=True)
        e = C.c_double()
        _check(_lib.load().dcrf_crf_cross_entropy(self._h, a.ptr, lab.ptr, o.ptr, C.byref(e), a.dev))
        return e.value, out

    def addSuperPixelLayer(self, region, exp_of_superpixels, constants):
        """Superpixel layer from a region map (the state DenseCRF2D::addSuperPixel builds)."""
        r = np.ascontiguousarray(region, np.int32)
        R = int(r.max()) + 1
        e = np.ascontiguousarray(np.broadcast_to(np.asarray(exp_of_superpixels, np.float32), (R,)))
        c = np.ascontiguousarray(np.broadcast_to(np.asarray(constants, np.float32), (R,)))
        p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        _check(_lib.load().dcrf_crf_add_superpixel_layer(self._h, p(r), R, p(e), p(c)))
        self.R += int(np.count_nonzero(np.bincount(r, minlength=R)))

    # ---- potentials -------------------------------------------------------------
    def apply(self, k, Q, transpose=False, no_norm=False, out=None):
        a = _Arg(Q, np.float32)
        out = _empty_like_field(Q, (self.N, self.M)) if out is None else out
        o = _Arg(out, np.float32, out=True)
        _check(_lib.load().dcrf_crf_apply(self._h, k, a.ptr, o.ptr, int(transpose), int(no_norm), a.dev))
        return out

    def apply_upper_minus_lower_ord(self, k, Q, out=None):
        a = _Arg(Q, np.float32)
        out = _empty_like_field(Q, (self.N, self.M)) if out is None else out
        o = _Arg(out, np.float32, out=True)
        _check(_lib.load().dcrf_crf_apply_upper_minus_lower_ord(self._h, k, a.ptr, o.ptr, a.dev))
        return out

    def pairwise_kernel_gradient(self, k, b, Q):
        """PairwisePotential::kernelGradient(b, Q) of potential k: (0,), (d,) or
        (d*d,) values for a CONST, DIAG or FULL kernel"""
        x = _Arg(b, np.float32)
        q = _Arg(Q, np.float32)
        n = C.c_int()
        _check(_lib.load().dcrf_crf_pairwise_kernel_nparams(self._h, int(k), C.byref(n)))
        out = np.zeros(max(n.value, 1), np.float32)
        _check(_lib.load().dcrf_crf_pairwise_kernel_gradient(self._h, k, x.ptr, q.ptr,
                                                             out.ctypes.data_as(C.c_void_p), x.dev))
        return out[: n.value]

    def apply_bf(self, k, Q, out=None):
        """PairwisePotential::apply_bf of potential k (dense Gaussian, O(N^2 L))."""
        a = _Arg(Q, np.float32)
        out = _empty_like_field(Q, (self.N, self.M)) if out is None else out
        o = _Arg(out, np.float32, out=True)
        _check(_lib.load().dcrf_crf_apply_bf(self._h, k, a.ptr, o.ptr, a.dev))
        return out

    def apply_upper_minus_lower_bf(self, k, ind, out=None):
        """PairwisePotential::apply_upper_minus_lower_bf of potential k; ind (N, M):
        ind[c, l] is the pixel at sorted position c of label l; the result stays
        in that sorted order."""
        ix = _Arg(ind, np.int32)
        out = _empty_like_field(ind, (self.N, self.M)) if out is None else out
        o = _Arg(out, np.float32, out=True)
        _check(_lib.load().dcrf_crf_apply_upper_minus_lower_bf(self._h, k, ix.ptr, o.ptr, ix.dev))
        return out

    def apply_upper_minus_lower_dc(self, k, ind, out=None):
        """PairwisePotential::apply_upper_minus_lower_dc of potential k (divide and
        conquer over the sorted positions ind (N, M)); the result in sorted order."""
        ix = _Arg(ind, np.int32)
        out = _empty_like_field(ind, (self.N, self.M)) if out is None else out
        o = _Arg(out, np.float32, out=True)
        _check(_lib.load().dcrf_crf_apply_upper_minus_lower_dc(self._h, k, ix.ptr, o.ptr, ix.dev))
        return out

    def apply_upper_minus_lower_bf_ord(self, k, ind, Q, out=None):
        """PairwisePotential::apply_upper_minus_lower_bf_ord of potential k."""
        ix = _Arg(ind, np.int32)
        a = _Arg(Q, np.float32)
        out = _empty_like_field(Q, (self.N, self.M)) if out is None else out
        o = _Arg(out, np.float32, out=True)
        _check(_lib.load().dcrf_crf_apply_upper_minus_lower_bf_ord(self._h, k, ix.ptr, a.ptr, o.ptr, a.dev))
        return out

    def apply_upper_minus_lower_ord_restricted(self, k, rQ, pI, Q=None, store: bool = True, out=None):
        """PairwisePotential::apply_upper_minus_lower_ord_restricted of potential k:
        (len(pI), M) = compatibility(restricted ordered filter)."""
        a, ix, e, o, out = _restricted_args(rQ, pI, Q, store, out, (len(pI), self.M))
        _check(_lib.load().dcrf_crf_apply_upper_minus_lower_ord_restricted(
            self._h, k, o.ptr, a.ptr, ix.ptr, len(pI), e.ptr if e else None, int(store), a.dev))
        return out

    def unary_init(self, device=None):
        self._sync_unary()
        out = np.empty((self.N, self.M), np.float32) if device is None else _torch_empty((self.N, self.M), device)
        o = _Arg(out, np.float32, out=True)
        _check(_lib.load().dcrf_crf_unary_init(self._h, o.ptr, o.dev))
        return out

    # ---- solvers ---------------------------------------------------------------
    def qp_inference_super_pixels_non_convex(self, init, max_iter: int = 0, ref_stop: bool | None = None,
                                             stats: bool = False, out=None):
        """QP Frank-Wolfe.  Default: the reference's stopping rule; max_iter>0
        with ref_stop=False runs exactly max_iter iterations."""
        if ref_stop is None:
            ref_stop = max_iter <= 0
        self._sync_unary()
        a = _Arg(init, np.float32)
        out = _empty_like_field(init, (self.N, self.M)) if out is None else out
        o = _Arg(out, np.float32, out=True)
        st = None
        keep = None
        if stats:
            cap = max(max_iter, 1) if max_iter > 0 else 4096
            keep = [np.zeros(cap + 1), np.zeros(cap), np.zeros(cap), np.zeros(cap)]
            dp = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
            st = _lib.QPStats(0, 0, cap, dp(keep[0]), dp(keep[1]), dp(keep[2]), dp(keep[3]))
        _check(_lib.load().dcrf_crf_qp(self._h, a.ptr, o.ptr, int(max_iter), int(ref_stop), a.dev,
                                       C.byref(st) if st is not None else None))
        if stats:
            it, up = st.iterations, st.updates
            return out, dict(iters=it, updates=up, energy=keep[0][: it + 1], num=keep[1][:it], denom=keep[2][:it],
                             step=keep[3][:up])
        return out

    def lp_inference_prox_super_pixels(self, init, params: LP_inf_params | None = None, stats: bool = False,
                                       out=None):
        params = params or LP_inf_params()
        self._sync_unary()
        a = _Arg(init, np.float32)
        out = _empty_like_field(init, (self.N, self.M)) if out is None else out
        o = _Arg(out, np.float32, out=True)
        st = None
        keep = None
        if stats:
            cap = 4096
            keep = [np.zeros(cap, np.int32), np.zeros(cap), np.zeros(cap), np.zeros(cap)]
            dp = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
            st = _lib.LPStats(0, 0, cap, keep[0].ctypes.data_as(C.POINTER(C.c_int)), dp(keep[1]), dp(keep[2]),
                              dp(keep[3]), 0.0, 0.0)
        _check(_lib.load().dcrf_crf_lp(self._h, a.ptr, o.ptr, C.byref(params._c()), a.dev,
                                       C.byref(st) if st is not None else None))
        if stats:
            n = st.inner
            return out, dict(outer=st.outer, qp_sweeps=keep[0][:n], gap=keep[1][:n], delta=keep[2][:n],
                             int_energy=keep[3][: st.outer], init_lp_energy=st.init_lp_energy,
                             init_int_energy=st.init_int_energy)
        return out

    # ---- energies ---------------------------------------------------------------
    def currentMap(self, Q):
        a = _Arg(Q, np.float32)
        out = _empty_like_field(Q, (self.N,), np.int32)
        o = _Arg(out, np.int32, out=True)
        _check(_lib.load().dcrf_crf_current_map(self._h, a.ptr, o.ptr, a.dev))
        return out

    def _energy(self, fn, x, dtype):
        self._sync_unary()
        a = _Arg(x, dtype)
        e = C.c_double()
        _check(fn(self._h, a.ptr, a.dev, C.byref(e)))
        return e.value

    def unaryEnergy(self, labels):
        """DenseCRF::unaryEnergy(l): per-pixel U(l_i, i) (float32, (N,))"""
        self._sync_unary()
        lab = np.ascontiguousarray(labels, np.int32)
        out = np.zeros(self.N, np.float32)
        _check(_lib.load().dcrf_crf_unary_energy(self._h, lab.ctypes.data_as(C.c_void_p),
                                                 out.ctypes.data_as(C.c_void_p), 0))
        return out

    def pairwise_energy_true(self, labels, term: int = -1):
        """DenseCRF::pairwise_energy_true(l, term): per-pixel pairwise energy (float32, (N,))"""
        lab = np.ascontiguousarray(labels, np.int32)
        out = np.zeros(self.N, np.float32)
        _check(_lib.load().dcrf_crf_pairwise_energy_true(self._h, lab.ctypes.data_as(C.c_void_p), term,
                                                         out.ctypes.data_as(C.c_void_p), 0))
        return out

    def assignment_energy_true(self, labels):
        return self._energy(_lib.load().dcrf_crf_assignment_energy_true, labels, np.int32)

    def assignment_energy_higher_order(self, labels):
        return self._energy(_lib.load().dcrf_crf_assignment_energy_higher_order, labels, np.int32)

    def compute_energy_LP_higher_order(self, Q):
        return self._energy(_lib.load().dcrf_crf_lp_energy, Q, np.float32)

    def compute_energy_higher_order(self, Q):
        """DenseCRF::compute_energy_higher_order(Q): QP energy with the last QP's z_labels_"""
        return self._energy(_lib.load().dcrf_crf_energy_higher_order, Q, np.float32)

    def computeUCondGrad(self, Q, out=None):
        a = _Arg(Q, np.float32)
        out = _empty_like_field(Q, (self.N, self.M)) if out is None else out
        o = _Arg(out, np.float32, out=True)
        _check(_lib.load().dcrf_crf_ucondgrad(self._h, a.ptr, o.ptr, a.dev))
        return out

    def compute_LR_QP_value(self, Q):
        return self._energy(_lib.load().dcrf_crf_qp_value, Q, np.float32)


def _torch_empty(shape, device):
    import torch
    return torch.empty(shape, dtype=torch.float32, device=device)


class DenseCRF2D(DenseCRF):
    """DenseCRF2D (include/densecrf.h:172-194): pixel index y*W + x."""

    def __init__(self, W: int, H: int, M: int):
        super().__init__(W * H, M)
        self.W, self.H = int(W), int(H)

    def addSuperPixel(self, img, region, constant: float = 50.0, normaliser: float = 1000.0):
        """DenseCRF2D::addSuperPixel (src/densecrf.cpp:193-251) with the mean-shift
        segmentation given as a region map (one id per pixel): the layer's terms
        come from the image as in the reference."""
        e, c = superpixel_terms(img, region, constant, normaliser)
        self.addSuperPixelLayer(region, e, c)

    def addPairwiseGaussian(self, sx, sy, function: PottsCompatibility | None = None, kernel_type=1,
                            normalization_type=NORMALIZE_SYMMETRIC):
        if normalization_type != NORMALIZE_SYMMETRIC or getattr(function, "kind", 0) != 0:
            from .synthetic import gaussian_features
            return self.addPairwiseEnergy(gaussian_features(self.W, self.H, sx, sy), function, kernel_type,
                                          normalization_type)
        w = (function or PottsCompatibility(1.0)).w
        _check(_lib.load().dcrf_crf_add_pairwise_gaussian(self._h, self.W, self.H, float(sx), float(sy), w))
        _check(_lib.load().dcrf_crf_pairwise_set_kernel_type(self._h, self.npairwise, int(kernel_type)))
        self.npairwise += 1
        self._pw_dims.append(2)

    def addPairwiseBilateral(self, sx, sy, sr, sg, sb, im, function: PottsCompatibility | None = None,
                             kernel_type=1, normalization_type=NORMALIZE_SYMMETRIC):
        im = np.ascontiguousarray(im, np.uint8).reshape(-1)
        if normalization_type != NORMALIZE_SYMMETRIC or getattr(function, "kind", 0) != 0:
            from .synthetic import bilateral_features
            return self.addPairwiseEnergy(bilateral_features(self.W, self.H, sx, sy, sr, sg, sb, im), function,
                                          kernel_type, normalization_type)
        w = (function or PottsCompatibility(1.0)).w
        _check(_lib.load().dcrf_crf_add_pairwise_bilateral(self._h, self.W, self.H, float(sx), float(sy), float(sr),
                                                           float(sg), float(sb),
                                                           im.ctypes.data_as(C.c_void_p), w))
        _check(_lib.load().dcrf_crf_pairwise_set_kernel_type(self._h, self.npairwise, int(kernel_type)))
        self.npairwise += 1
        self._pw_dims.append(5)


class DenseCRFBatch:
    """Many images of uneven sizes solved together (no counterpart in the
    reference, which runs one DenseCRF per image): each image follows the
    per-image QP / LP algorithm.  Batch fields are the images' (N_b, L) fields
    concatenated in image order; ``offsets[b]`` is image b's first pixel."""

    def __init__(self, L: int, sizes):
        """sizes: list of (W, H) per image."""
        Lb = _lib.load()
        self.L = int(L)
        W = np.array([int(w) for w, _ in sizes], np.int32)
        H = np.array([int(h) for _, h in sizes], np.int32)
        self.B = len(W)
        h = C.c_void_p()
        _check(Lb.dcrf_batch_create_2d(self.L, self.B, W.ctypes.data_as(C.c_void_p), H.ctypes.data_as(C.c_void_p),
                                       C.byref(h)))
        self._h = h
        off = np.zeros(self.B + 1, np.int64)
        _check(Lb.dcrf_batch_pixel_offsets(self._h, off.ctypes.data_as(C.c_void_p)))
        self.offsets = off
        self.N = int(off[-1])
        self.sizes = list(zip(W.tolist(), H.tolist()))

    def close(self):
        if getattr(self, "_h", None) is not None:
            _lib.load().dcrf_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_graphs(self, on: bool):
        _check(_lib.load().dcrf_batch_set_graphs(self._h, int(on)))

    def set_value_type(self, bf16: bool):
        """ordered-filter lattice rows stored in bfloat16 (True) or float"""
        _check(_lib.load().dcrf_batch_set_value_type(self._h, int(bool(bf16))))

    def graph_captures(self) -> int:
        n = C.c_longlong()
        _check(_lib.load().dcrf_batch_graph_captures(self._h, C.byref(n)))
        return n.value

    def setUnaryEnergy(self, b: int, unary):
        a = _Arg(unary, np.float32)
        _check(_lib.load().dcrf_batch_set_unary(self._h, int(b), a.ptr, a.dev))

    def addPairwiseGaussian(self, sx, sy, function: PottsCompatibility | None = None):
        w = (function or PottsCompatibility(1.0)).w
        _check(_lib.load().dcrf_batch_add_pairwise_gaussian(self._h, float(sx), float(sy), w))

    def addPairwiseBilateral(self, sx, sy, sr, sg, sb, images, function: PottsCompatibility | None = None):
        w = (function or PottsCompatibility(1.0)).w
        ims = [np.ascontiguousarray(im, np.uint8).reshape(-1) for im in images]
        ptrs = (C.c_void_p * self.B)(*[im.ctypes.data_as(C.c_void_p).value for im in ims])
        _check(_lib.load().dcrf_batch_add_pairwise_bilateral(self._h, float(sx), float(sy), float(sr), float(sg),
                                                             float(sb), C.cast(ptrs, C.c_void_p), w))

    def addPairwiseEnergy(self, features, function: PottsCompatibility | None = None,
                          normalization_type=NORMALIZE_SYMMETRIC):
        """features: one (N_b, d) array per image (host) or one DeviceArray each."""
        w = (function or PottsCompatibility(1.0)).w
        args = [_Arg(f, np.float32) for f in features]
        d = args[0].arr.shape[1]
        ptrs = (C.c_void_p * self.B)(*[a.ptr.value if isinstance(a.ptr, C.c_void_p) else a.ptr for a in args])
        _check(_lib.load().dcrf_batch_add_pairwise(self._h, C.cast(ptrs, C.c_void_p), d, w, int(normalization_type),
                                                   args[0].dev))

    def image(self, field, b):
        """image b's rows of a batch field"""
        return field[self.offsets[b]:self.offsets[b + 1]]

    def unary_init(self, out=None):
        out = np.empty((self.N, self.L), np.float32) if out is None else out
        o = _Arg(out, np.float32, out=True)
        _check(_lib.load().dcrf_batch_unary_init(self._h, o.ptr, o.dev))
        return out

    def qp(self, init=None, max_iter: int = 0, ref_stop: bool | None = None, out=None):
        """QP of every image; returns (Q, iterations per image)."""
        if ref_stop is None:
            ref_stop = max_iter <= 0
        a = _Arg(init, np.float32) if init is not None else None
        proto = init if init is not None else out
        if out is None:
            out = _empty_like_field(proto, (self.N, self.L)) if proto is not None else \
                np.empty((self.N, self.L), np.float32)
        o = _Arg(out, np.float32, out=True)
        it = np.zeros(self.B, np.int32)
        _check(_lib.load().dcrf_batch_qp(self._h, a.ptr if a else None, o.ptr, int(max_iter), int(ref_stop), o.dev,
                                         it.ctypes.data_as(C.c_void_p)))
        return out, it

    def lp(self, init, params: LP_inf_params | None = None, out=None):
        """LP of every image; returns (Q, qp-gamma sweeps (B, inner steps))."""
        params = params or LP_inf_params()
        a = _Arg(init, np.float32)
        out = _empty_like_field(init, (self.N, self.L)) if out is None else out
        o = _Arg(out, np.float32, out=True)
        cap = int(params.prox_max_iter * params.fw_max_iter)
        sw = np.zeros((self.B, cap), np.int32)
        _check(_lib.load().dcrf_batch_lp(self._h, a.ptr, o.ptr, C.byref(params._c()), a.dev,
                                         sw.ctypes.data_as(C.c_void_p), cap))
        return out, sw

    def assignment_energy_true(self, labels):
        a = _Arg(labels, np.int32)
        e = np.zeros(self.B)
        _check(_lib.load().dcrf_batch_assignment_energy_true(self._h, a.ptr, a.dev, e.ctypes.data_as(C.c_void_p)))
        return e
