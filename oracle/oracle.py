"""ctypes front-end for the CPU oracle (liboracle.so) and, where it was built,
the compiled reference (oracle/_ref/libdcrf_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by bench.py's cpu_baseline leg, never by the engine itself.

Layouts are the reference's: features (N, d) pixel-major, fields (N, L)
label-contiguous, region maps (N,) int32.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB selects another build of the same sources (the sanitizer build of
# `make -C oracle check-asan`)
ORACLE_SO = os.environ.get("ORACLE_LIB", os.path.join(_HERE, "liboracle.so"))
REF_SO = os.path.join(_HERE, "_ref", "libdcrf_ref.so")
# the reference's Permutohedral compiled without its SSE code (portable build)
REFS_SO = os.path.join(_HERE, "_ref", "libdcrf_ref_scalar.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i16p = np.ctypeslib.ndpointer(np.int16, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_i = C.c_int
_f = C.c_float
_d = C.c_double

NTYPE = {"none": 0, "before": 1, "after": 2, "symmetric": 3}
COMPAT = {"potts": 0, "diagonal": 1, "matrix": 2}


def _sig(lib, name, res, *args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = list(args)
    return fn


_orc = None


def lib():
    """Load liboracle.so (build it with ``make -C oracle``)."""
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run `make -C oracle oracle`)")
        L = C.CDLL(ORACLE_SO)
        _sig(L, "orc_last_error", C.c_char_p)
        _sig(L, "orc_lattice_new", _vp, _f32p, _i, _i)
        _sig(L, "orc_lattice_free", None, _vp)
        _sig(L, "orc_lattice_M", _i, _vp)
        _sig(L, "orc_lattice_Npad", _i, _vp)
        _sig(L, "orc_lattice_get", None, _vp, _i32p, _i32p, _f32p, _i16p, _i32p)
        _sig(L, "orc_filter_f", None, _vp, _f32p, _f32p, _i, _i)
        _sig(L, "orc_filter_d", None, _vp, _f64p, _f64p, _i, _i)
        _sig(L, "orc_ordered_f", None, _vp, _f32p, _f32p, _i)
        _sig(L, "orc_ordered_d", None, _vp, _f32p, _f64p, _i)
        _sig(L, "orc_bf_apply", None, _f32p, _i, _i, _f32p, _i, _f32p)
        _sig(L, "orc_lattice_gradient", None, _vp, _f32p, _f32p, _i, _f32p)
        _sig(L, "orc_uml_dc", None, _f32p, _i, _i, _i32p, _i, _f32p)
        _sig(L, "orc_lattice_dc", None, _vp, _f32p, _i, _i, _i, _i)
        _sig(L, "orc_bf_uml", None, _f32p, _i, _i, _i32p, _vp, _i, _f32p)
        _sig(L, "orc_ordered_restricted", _i, _vp, _f32p, _i32p, _i, _f32p, _i, _i, _f32p)
        _sig(L, "orc_crf_new", _vp, _i, _i)
        _sig(L, "orc_crf_free", None, _vp)
        _sig(L, "orc_crf_set_unary", None, _vp, _f32p)
        _sig(L, "orc_crf_add_pairwise", None, _vp, _f32p, _i, _f, _i)
        _sig(L, "orc_crf_add_pairwise_compat", None, _vp, _f32p, _i, _i, _f32p, _i)
        _sig(L, "orc_crf_nparams", _i, _vp, _i)
        _sig(L, "orc_crf_pairwise_gradient_f", None, _vp, _i, _f32p, _f32p, _f64p)
        _sig(L, "orc_crf_pairwise_gradient_d", None, _vp, _i, _f64p, _f64p, _f64p)
        _sig(L, "orc_crf_meanfield_f", None, _vp, _i, _f32p)
        _sig(L, "orc_crf_meanfield_d", None, _vp, _i, _f64p)
        _sig(L, "orc_crf_meanfield_grad_d", None, _vp, _i, _f64p, _f64p, _f64p, _f64p)
        _sig(L, "orc_crf_lattice_M", _i, _vp, _i)
        _sig(L, "orc_crf_get_norm", None, _vp, _i, _f32p, _f64p)
        _sig(L, "orc_crf_add_layer", None, _vp, _i32p, _i, _f32p, _f32p)
        _sig(L, "orc_crf_apply_f", _i, _vp, _i, _f32p, _f32p, _i, _i)
        _sig(L, "orc_crf_apply_d", _i, _vp, _i, _f64p, _f64p, _i, _i)
        _sig(L, "orc_crf_apply_ord_f", _i, _vp, _i, _f32p, _f32p)
        _sig(L, "orc_crf_apply_ord_d", _i, _vp, _i, _f32p, _f64p)
        _sig(L, "orc_crf_qp", _i, _vp, _i, _f32p, _f32p, _i, _i, _f64p, _f64p, _f64p, _f64p, _i)
        _sig(L, "orc_crf_zlabels", None, _vp, _f32p)
        _sig(L, "orc_crf_lp", _i, _vp, _i, _f32p, _f32p, _f32p, _i32p, _i32p, _f64p, _f64p, _f64p, _i, _f64p)
        _sig(L, "orc_crf_assignment_energy_true", _d, _vp, _i, _i32p)
        _sig(L, "orc_crf_assignment_energy_ho", _d, _vp, _i, _i32p)
        _sig(L, "orc_crf_lp_energy", _d, _vp, _i, _f32p)
        _sig(L, "orc_crf_ucondgrad", None, _vp, _f32p, _f32p)
        _sig(L, "orc_crf_current_map", None, _vp, _f32p, _i32p)
        _sig(L, "orc_unary_init", None, _f32p, _f32p, _i, _i)
        _sig(L, "orc_feasible", None, _f32p, _f32p, _i, _i)
        _sig(L, "orc_renormalize", None, _f32p, _i, _i)
        _sig(L, "orc_set_mutation", None, _i)
        _sig(L, "orc_set_reference_path", None, _i)
        _orc = L
    return _orc


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class Lattice:
    """Oracle lattice built with the reference's SSE2 semantics, or its
    portable build's after set_reference_path(True)."""

    def __init__(self, feat: np.ndarray):
        feat = _c(feat, np.float32)
        self.N, self.d = feat.shape
        self._L = lib()
        self.h = self._L.orc_lattice_new(feat, self.N, self.d)
        self.M = self._L.orc_lattice_M(self.h)
        self.Npad = self._L.orc_lattice_Npad(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self._L.orc_lattice_free(self.h)
            self.h = None

    def arrays(self):
        d1 = self.d + 1
        off = np.zeros(self.Npad * d1, np.int32)
        rk = np.zeros(self.Npad * d1, np.int32)
        bc = np.zeros(self.Npad * d1, np.float32)
        keys = np.zeros(max(self.M * self.d, 1), np.int16)
        nbr = np.zeros(max(d1 * self.M * 2, 1), np.int32)
        self._L.orc_lattice_get(self.h, off, rk, bc, keys, nbr)
        return dict(offset=off.reshape(self.Npad, d1), rank=rk.reshape(self.Npad, d1),
                    bary=bc.reshape(self.Npad, d1), keys=keys[: self.M * self.d].reshape(self.M, self.d),
                    nbr=nbr[: d1 * self.M * 2].reshape(d1, self.M, 2))

    def filter(self, x: np.ndarray, reverse=False, fp64=False):
        x = np.asarray(x)
        V = x.shape[1] if x.ndim == 2 else 1
        if fp64:
            xi = _c(x, np.float64)
            out = np.zeros_like(xi)
            self._L.orc_filter_d(self.h, xi, out, V, int(reverse))
        else:
            xi = _c(x, np.float32)
            out = np.zeros_like(xi)
            self._L.orc_filter_f(self.h, xi, out, V, int(reverse))
        return out

    def ordered(self, q: np.ndarray, fp64=False):
        q = _c(q, np.float32)
        V = q.shape[1]
        out = np.zeros(q.shape, np.float64 if fp64 else np.float32)
        (self._L.orc_ordered_d if fp64 else self._L.orc_ordered_f)(self.h, q, out, V)
        return out

    def dc(self, out, low, ml, mh, high):
        """compute_upper_minus_lower_dc on one value per pixel; out (N,) added to"""
        out = _c(out, np.float32).copy()
        self._L.orc_lattice_dc(self.h, out, low, ml, mh, high)
        return out

    def gradient(self, a, b):
        """Permutohedral::gradient: (N, d) = d/df a^T K b, a and b (N, V)"""
        a = _c(a, np.float32)
        b = _c(b, np.float32)
        df = np.zeros((self.N, self.d), np.float32)
        self._L.orc_lattice_gradient(self.h, a, b, a.shape[1], df)
        return df

    def ordered_restricted(self, rq, pI, q, store, out=None):
        """seqCompute_upper_minus_lower_ord_restricted: rq (|pI| x V) for the
        pixels pI, q (N x V) for the others; out is added to (zeros if None)"""
        rq = _c(rq, np.float32)
        pI = _c(pI, np.int32)
        V = rq.shape[1]
        q = _c(q, np.float32) if q is not None else np.zeros((self.N, V), np.float32)
        out = np.zeros((len(pI), V), np.float32) if out is None else _c(out, np.float32).copy()
        if self._L.orc_ordered_restricted(self.h, rq, pI, len(pI), q, V, int(store), out) != 0:
            raise RuntimeError(self._L.orc_last_error().decode())
        return out


def bf_apply(feat, Q):
    """PairwisePotential::apply_bf before the compatibility (src/pairwise.cpp:368-383)"""
    feat = _c(feat, np.float32)
    Q = _c(Q, np.float32)
    out = np.zeros_like(Q)
    lib().orc_bf_apply(feat, feat.shape[0], feat.shape[1], Q, Q.shape[1], out)
    return out


def bf_uml(feat, ind, Q=None):
    """apply_upper_minus_lower_bf (Q None) / _bf_ord before the compatibility
    (src/pairwise.cpp:294-353); ind (N, L); the result in sorted order"""
    feat = _c(feat, np.float32)
    ind = _c(ind, np.int32)
    out = np.zeros(ind.shape, np.float32)
    q = None if Q is None else _c(Q, np.float32)
    lib().orc_bf_uml(feat, feat.shape[0], feat.shape[1], ind, None if q is None else q.ctypes.data_as(C.c_void_p),
                     ind.shape[1], out)
    return out


def uml_dc(feat, ind):
    """apply_upper_minus_lower_dc before the compatibility (src/pairwise.cpp:241-293);
    ind (N, L); the result in sorted order"""
    feat = _c(feat, np.float32)
    ind = _c(ind, np.int32)
    out = np.zeros(ind.shape, np.float32)
    lib().orc_uml_dc(feat, feat.shape[0], feat.shape[1], ind, ind.shape[1], out)
    return out


class LPParams:
    """``LP_inf_params`` (include/densecrf.h:43-69), defaults of src/densecrf.cpp:287-302."""

    def __init__(self, prox_reg_const=1e-2, dual_gap_tol=1e3, prox_energy_tol=1e3, prox_max_iter=10,
                 fw_max_iter=10, qp_max_iter=1000, qp_tol=1e3, qp_const=1e-16, best_int=True,
                 less_confident_percent=0.0, confidence_tol=0.95):
        self.prox_reg_const = prox_reg_const
        self.dual_gap_tol = dual_gap_tol
        self.prox_energy_tol = prox_energy_tol
        self.prox_max_iter = prox_max_iter
        self.fw_max_iter = fw_max_iter
        self.qp_max_iter = qp_max_iter
        self.qp_tol = qp_tol
        self.qp_const = qp_const
        self.best_int = best_int
        self.less_confident_percent = less_confident_percent
        self.confidence_tol = confidence_tol

    def packed(self):
        fp = np.array([self.prox_reg_const, self.dual_gap_tol, self.prox_energy_tol, self.qp_tol,
                       self.qp_const, self.less_confident_percent, self.confidence_tol], np.float32)
        ip = np.array([self.prox_max_iter, self.fw_max_iter, self.qp_max_iter, int(self.best_int)], np.int32)
        return fp, ip


class CRF:
    """Oracle dense CRF (Potts pairwise terms, optional Pn superpixel layers)."""

    def __init__(self, N: int, L: int):
        self.N, self.L = N, L
        self._L = lib()
        self.h = self._L.orc_crf_new(N, L)
        self.nk = 0

    def __del__(self):
        if getattr(self, "h", None):
            self._L.orc_crf_free(self.h)
            self.h = None

    def set_unary(self, U):
        self._L.orc_crf_set_unary(self.h, _c(U, np.float32))

    def add_pairwise(self, feat, w, ntype="symmetric"):
        feat = _c(feat, np.float32)
        self._L.orc_crf_add_pairwise(self.h, feat, feat.shape[1], float(w), NTYPE[ntype])
        self.nk += 1

    def add_layer(self, region, e, c):
        region = _c(region, np.int32)
        R = int(region.max()) + 1
        e = np.broadcast_to(np.asarray(e, np.float32), (R,)).copy()
        c = np.broadcast_to(np.asarray(c, np.float32), (R,)).copy()
        self._L.orc_crf_add_layer(self.h, region, R, e, c)

    def add_pairwise_compat(self, feat, kind, params, ntype="symmetric"):
        """Pairwise term with a Potts ("potts", [w]), Diagonal ("diagonal", v[L])
        or Matrix ("matrix", m (L, L), stored as (m + m^T)/2) compatibility."""
        feat = _c(feat, np.float32)
        ct = COMPAT[kind]
        # the matrix is passed column-major (Eigen); m[i, j] is row i, column j
        par = _c(np.asarray(params, np.float32).T.reshape(-1) if ct == 2 else np.ravel(params), np.float32)
        self._L.orc_crf_add_pairwise_compat(self.h, feat, feat.shape[1], ct, par, NTYPE[ntype])
        self.nk += 1

    def nparams(self, k):
        return self._L.orc_crf_nparams(self.h, k)

    def pairwise_gradient(self, k, b, Q, fp64=False):
        """PairwisePotential::gradient(b, Q) in the layout of parameters()."""
        g = np.zeros(self.nparams(k))
        if fp64:
            self._L.orc_crf_pairwise_gradient_d(self.h, k, _c(b, np.float64), _c(Q, np.float64), g)
        else:
            self._L.orc_crf_pairwise_gradient_f(self.h, k, _c(b, np.float32), _c(Q, np.float32), g)
        return g

    def meanfield(self, T, fp64=False):
        """Q after T unrolled mean-field updates from Q0 = softmax(-U)."""
        out = np.zeros((self.N, self.L), np.float64 if fp64 else np.float32)
        (self._L.orc_crf_meanfield_d if fp64 else self._L.orc_crf_meanfield_f)(self.h, int(T), out)
        return out

    def meanfield_grad(self, T, dQ):
        """FP64 forward and adjoint: (Q_T, dL/dU, [dL/dparams of each term])."""
        Q = np.zeros((self.N, self.L))
        dU = np.zeros((self.N, self.L))
        nps = [self.nparams(k) for k in range(self.nk)]
        g = np.zeros(max(sum(nps), 1))
        self._L.orc_crf_meanfield_grad_d(self.h, int(T), _c(dQ, np.float64), Q, dU, g)
        out, o = [], 0
        for n in nps:
            out.append(g[o:o + n].copy())
            o += n
        return Q, dU, out

    def lattice_M(self, k):
        return self._L.orc_crf_lattice_M(self.h, k)

    def norm(self, k):
        nf = np.zeros(self.N, np.float32)
        nd = np.zeros(self.N, np.float64)
        self._L.orc_crf_get_norm(self.h, k, nf, nd)
        return nf, nd

    def apply(self, k, x, fp64=False, no_norm=False, transpose=False):
        if fp64:
            xi = _c(x, np.float64)
            out = np.zeros_like(xi)
            self._L.orc_crf_apply_d(self.h, k, xi, out, int(no_norm), int(transpose))
        else:
            xi = _c(x, np.float32)
            out = np.zeros_like(xi)
            self._L.orc_crf_apply_f(self.h, k, xi, out, int(no_norm), int(transpose))
        return out

    def apply_ord(self, k, q, fp64=False):
        q = _c(q, np.float32)
        out = np.zeros(q.shape, np.float64 if fp64 else np.float32)
        rc = (self._L.orc_crf_apply_ord_d if fp64 else self._L.orc_crf_apply_ord_f)(self.h, k, q, out)
        if rc != 0:
            raise RuntimeError(self._L.orc_last_error().decode())
        return out

    def qp(self, init, max_iter=10, ref_stop=False, fp64=False):
        init = _c(init, np.float32)
        out = np.zeros_like(init)
        cap = max(max_iter, 1) if not ref_stop else 100000
        cap = min(cap, 100000)
        en = np.zeros(cap + 1)
        num = np.zeros(cap)
        den = np.zeros(cap)
        st = np.zeros(cap)
        it = self._L.orc_crf_qp(self.h, int(fp64), init, out, int(max_iter), int(ref_stop), en, num, den, st, cap)
        en = en[: it + 1]
        up = it - (1 if it > 0 and (num[it - 1] == 0 or den[it - 1] == 0) else 0)
        return out, dict(iters=it, updates=up, energy=en[~np.isnan(en)], num=num[:it], denom=den[:it],
                         step=st[:up])

    def zlabels(self, R):
        z = np.zeros(R * self.L, np.float32)
        self._L.orc_crf_zlabels(self.h, z)
        return z.reshape(R, self.L)

    def lp(self, init, params: LPParams | None = None, fp64=False):
        params = params or LPParams()
        fp, ip = params.packed()
        init = _c(init, np.float32)
        out = np.zeros_like(init)
        cap = 4096
        sw = np.zeros(cap, np.int32)
        gap = np.zeros(cap)
        dl = np.zeros(cap)
        ie = np.zeros(cap)
        e0 = np.zeros(2)
        it = self._L.orc_crf_lp(self.h, int(fp64), init, out, fp, ip, sw, gap, dl, ie, cap, e0)
        if it < 0:
            raise RuntimeError(self._L.orc_last_error().decode())
        nin = int(np.count_nonzero(sw))
        return out, dict(outer=it, qp_sweeps=sw[:nin], gap=gap[:nin], delta=dl[:nin], int_energy=ie[:it],
                         init_lp_energy=e0[0], init_int_energy=e0[1])

    def energy_true(self, lab, fp64=False):
        return self._L.orc_crf_assignment_energy_true(self.h, int(fp64), _c(lab, np.int32))

    def energy_ho(self, lab, fp64=False):
        return self._L.orc_crf_assignment_energy_ho(self.h, int(fp64), _c(lab, np.int32))

    def lp_energy(self, Q, fp64=False):
        return self._L.orc_crf_lp_energy(self.h, int(fp64), _c(Q, np.float32))

    def ucondgrad(self, Q):
        Q = _c(Q, np.float32)
        out = np.zeros_like(Q)
        self._L.orc_crf_ucondgrad(self.h, Q, out)
        return out

    def current_map(self, Q):
        Q = _c(Q, np.float32)
        lab = np.zeros(self.N, np.int32)
        self._L.orc_crf_current_map(self.h, Q, lab)
        return lab


def set_mutation(m: int):
    """Plant deliberate mistake `m` in the oracle (tests only; 0 restores it)."""
    lib().orc_set_mutation(int(m))


def set_reference_path(portable: bool):
    """Restate the reference's portable build (scalar init with ties rounded
    down, no padding lanes, seqCompute for every value size;
    src/permutohedral.cpp:334-540) instead of its SSE2 build for lattices
    built and filtered from now on."""
    lib().orc_set_reference_path(1 if portable else 0)


def unary_init(U):
    U = _c(U, np.float32)
    N, L = U.shape
    Q = np.zeros_like(U)
    lib().orc_unary_init(U, Q, N, L)
    return Q


def feasible(Q):
    Q = _c(Q, np.float32)
    out = np.zeros_like(Q)
    lib().orc_feasible(Q, out, Q.shape[0], Q.shape[1])
    return out


def renormalize(Q):
    Q = _c(Q, np.float32).copy()
    lib().orc_renormalize(Q, Q.shape[0], Q.shape[1])
    return Q


# ---------------------------------------------------------------------------
# compiled reference (development container only)
# ---------------------------------------------------------------------------
_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def reflib():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("reference build missing (oracle/_ref is only built where the reference exists)")
        R = C.CDLL(REF_SO)
        _sig(R, "ref_last_error", C.c_char_p)
        _sig(R, "ref_lattice_new", _vp, _f32p, _i, _i)
        _sig(R, "ref_lattice_free", None, _vp)
        _sig(R, "ref_lattice_M", _i, _vp)
        _sig(R, "ref_lattice_get", None, _vp, _i32p, _i32p, _f32p, _i32p, _i)
        _sig(R, "ref_lattice_compute", None, _vp, _f32p, _f32p, _i, _i, _i)
        _sig(R, "ref_lattice_ordered", None, _vp, _f32p, _f32p, _i, _i)
        _sig(R, "ref_probimage", None, C.c_char_p, _vp, _i32p)
        _sig(R, "ref_lattice_gradient", None, _vp, _f32p, _f32p, _i, _i, _f32p)
        _sig(R, "ref_lattice_dc", None, _vp, _f32p, _i, _i, _i, _i, _i)
        _sig(R, "ref_pairwise_kernel_gradient", _i, _f32p, _i, _i, _f, _i, _i, _vp, _i, _f32p, _f32p, _i, _f32p)
        _sig(R, "ref_crf_add_superpixel", _i, _vp, _vp, _i, _i, _i, _f, _f)
        _sig(R, "ref_crf_layers", None, _vp, _i32p, _f32p, _f32p)
        _sig(R, "ref_crf_apply_bf", None, _vp, _i, _i, _vp, _vp, _f32p)
        _sig(R, "ref_lattice_ordered_restricted", _i, _vp, _f32p, _i32p, _i, _f32p, _i, _i, _i, _f32p)
        _sig(R, "ref_logistic_unary", None, _f32p, _i, _i, _f32p, _i, _vp, _f32p, _f32p)
        _sig(R, "ref_crf_new", _vp, _i, _i, _i)
        _sig(R, "ref_crf_free", None, _vp)
        _sig(R, "ref_crf_set_unary", None, _vp, _f32p)
        _sig(R, "ref_crf_add_pairwise", None, _vp, _f32p, _i, _f, _i)
        _sig(R, "ref_crf_add_pairwise_compat", None, _vp, _f32p, _i, _i, _f32p, _i)
        _sig(R, "ref_crf_pairwise_gradient", _i, _vp, _i, _f32p, _f32p, C.c_void_p)
        _sig(R, "ref_crf_add_gaussian", None, _vp, _f, _f, _f)
        _sig(R, "ref_crf_add_bilateral", None, _vp, _f, _f, _f, _f, _f, _u8p, _f)
        _sig(R, "ref_crf_get_features", None, _vp, _i, _f32p)
        _sig(R, "ref_crf_add_layer", None, _vp, _i32p, _i, _f32p, _f32p)
        _sig(R, "ref_crf_R", _i, _vp)
        _sig(R, "ref_crf_apply", None, _vp, _i, _f32p, _f32p, _i)
        _sig(R, "ref_crf_apply_ord", _i, _vp, _i, _f32p, _f32p)
        _sig(R, "ref_crf_unary_init", None, _vp, _f32p)
        _sig(R, "ref_crf_qp", None, _vp, _f32p, _f32p)
        _sig(R, "ref_crf_zlabels", None, _vp, _f32p)
        _sig(R, "ref_crf_lp", _i, _vp, _f32p, _f32p, _f32p, _i32p)
        _sig(R, "ref_crf_energy_true", _d, _vp, _i32p)
        _sig(R, "ref_crf_energy_ho", _d, _vp, _i32p)
        _sig(R, "ref_crf_lp_energy", _d, _vp, _f32p)
        _sig(R, "ref_crf_qp_value", _d, _vp, _f32p)
        _sig(R, "ref_crf_energy_higher_order", _d, _vp, _f32p)
        _sig(R, "ref_crf_ucondgrad", None, _vp, _f32p, _f32p)
        _sig(R, "ref_crf_current_map", None, _vp, _f32p, _i32p)
        _sig(R, "ref_feasible", None, _f32p, _f32p, _i, _i)
        _sig(R, "ref_renormalize", None, _f32p, _i, _i)
        _ref = R
    return _ref


def ref_probimage(path):
    """ProbImage::decompress + boostToProb of the compiled reference: (N, D) probabilities, W, H"""
    R = reflib()
    dims = np.zeros(3, np.int32)
    R.ref_probimage(path.encode(), None, dims)
    W, H, D = (int(x) for x in dims)
    p = np.zeros((W * H, D), np.float32)
    R.ref_probimage(path.encode(), p.ctypes.data_as(C.c_void_p), dims)
    return p, W, H


def ref_logistic_unary(Lw, f, b=None):
    """the compiled reference's LogisticUnaryEnergy(L, f): (get() as (N, L),
    gradient(b) as L*F values or None); Lw is (F, L) and f (N, F), the memory
    of the reference's L x F and F x N matrices"""
    Lw, f = _c(Lw, np.float32), _c(f, np.float32)
    F, L = Lw.shape
    N = f.shape[0]
    U = np.zeros((N, L), np.float32)
    g = np.zeros(F * L, np.float32)
    bb = None if b is None else _c(b, np.float32)
    reflib().ref_logistic_unary(Lw, L, F, f, N, None if bb is None else bb.ctypes.data_as(C.c_void_p), U, g)
    return U, (None if b is None else g)


def ref_pairwise_kernel_gradient(feat, w, ktype, ntype, b, Q, kp=None):
    """standalone reference PairwisePotential(feat, Potts(w), ktype, ntype)
    [.setKernelParameters(kp)].kernelGradient(b, Q)"""
    feat = _c(feat, np.float32)
    b = _c(b, np.float32)
    Q = _c(Q, np.float32)
    kpa = None if kp is None else _c(kp, np.float32)
    out = np.zeros(feat.shape[1] ** 2 + 1, np.float32)
    n = reflib().ref_pairwise_kernel_gradient(feat, feat.shape[1], feat.shape[0], w, ktype, ntype,
                                             None if kpa is None else kpa.ctypes.data_as(C.c_void_p),
                                             0 if kpa is None else len(kpa), b, Q, b.shape[1], out)
    return out[:n]


class RefLattice:
    def __init__(self, feat):
        feat = _c(feat, np.float32)
        self.N, self.d = feat.shape
        self._R = reflib()
        self.h = self._R.ref_lattice_new(feat, self.N, self.d)
        self.M = self._R.ref_lattice_M(self.h)
        self.Npad = (self.N + 3) // 4 * 4

    def __del__(self):
        if getattr(self, "h", None):
            self._R.ref_lattice_free(self.h)
            self.h = None

    def arrays(self):
        d1 = self.d + 1
        off = np.zeros(self.Npad * d1, np.int32)
        rk = np.zeros(self.Npad * d1, np.int32)
        bc = np.zeros(self.Npad * d1, np.float32)
        nbr = np.zeros(max(d1 * self.M * 2, 1), np.int32)
        self._R.ref_lattice_get(self.h, off, rk, bc, nbr, self.Npad)
        return dict(offset=off.reshape(self.Npad, d1), rank=rk.reshape(self.Npad, d1),
                    bary=bc.reshape(self.Npad, d1), nbr=nbr[: d1 * self.M * 2].reshape(d1, self.M, 2))

    def filter(self, x, reverse=False):
        x = _c(x, np.float32)
        V = x.shape[1] if x.ndim == 2 else 1
        out = np.zeros_like(x)
        self._R.ref_lattice_compute(self.h, x, out, V, self.N, int(reverse))
        return out

    def ordered(self, q):
        q = _c(q, np.float32)
        out = np.zeros_like(q)
        self._R.ref_lattice_ordered(self.h, q, out, q.shape[1], self.N)
        return out

    def dc(self, out, low, ml, mh, high):
        out = _c(out, np.float32).copy()
        self._R.ref_lattice_dc(self.h, out, self.N, low, ml, mh, high)
        return out

    def gradient(self, a, b):
        """Permutohedral::gradient: (N, d) of a^T K b"""
        a = _c(a, np.float32)
        b = _c(b, np.float32)
        df = np.zeros((self.N, self.d), np.float32)
        self._R.ref_lattice_gradient(self.h, a, b, a.shape[1], self.N, df)
        return df

    def ordered_restricted(self, rq, pI, q, store, out=None):
        rq = _c(rq, np.float32)
        pI = _c(pI, np.int32)
        V = rq.shape[1]
        q = _c(q, np.float32) if q is not None else np.zeros((self.N, V), np.float32)
        out = np.zeros((len(pI), V), np.float32) if out is None else _c(out, np.float32).copy()
        self._R.ref_lattice_ordered_restricted(self.h, rq, pI, len(pI), q, V, self.N, int(store), out)
        return out


class RefCRF:
    def __init__(self, W, H, L):
        self.W, self.H, self.L, self.N = W, H, L, W * H
        self._R = reflib()
        self.h = self._R.ref_crf_new(W, H, L)

    def __del__(self):
        if getattr(self, "h", None):
            self._R.ref_crf_free(self.h)
            self.h = None

    def set_unary(self, U):
        self._R.ref_crf_set_unary(self.h, _c(U, np.float32))

    def add_pairwise(self, feat, w, ntype="symmetric"):
        feat = _c(feat, np.float32)
        self._R.ref_crf_add_pairwise(self.h, feat, feat.shape[1], float(w), NTYPE[ntype])

    def add_pairwise_compat(self, feat, kind, params, ntype="symmetric"):
        feat = _c(feat, np.float32)
        ct = COMPAT[kind]
        par = _c(np.asarray(params, np.float32).T.reshape(-1) if ct == 2 else np.ravel(params), np.float32)
        self._R.ref_crf_add_pairwise_compat(self.h, feat, feat.shape[1], ct, par, NTYPE[ntype])

    def pairwise_gradient(self, k, b, Q):
        b, Q = _c(b, np.float32), _c(Q, np.float32)
        n = self._R.ref_crf_pairwise_gradient(self.h, k, b, Q, None)
        g = np.zeros(n, np.float32)
        self._R.ref_crf_pairwise_gradient(self.h, k, b, Q, g.ctypes.data_as(C.c_void_p))
        return g

    def add_gaussian(self, sx, sy, w):
        self._R.ref_crf_add_gaussian(self.h, sx, sy, w)

    def add_bilateral(self, sx, sy, sr, sg, sb, img, w):
        self._R.ref_crf_add_bilateral(self.h, sx, sy, sr, sg, sb, _c(img, np.uint8), w)

    def features(self, k, d):
        out = np.zeros(self.N * d, np.float32)
        self._R.ref_crf_get_features(self.h, k, out)
        return out.reshape(self.N, d)

    def add_layer(self, region, e, c):
        region = _c(region, np.int32)
        R = int(region.max()) + 1
        e = np.broadcast_to(np.asarray(e, np.float32), (R,)).copy()
        c = np.broadcast_to(np.asarray(c, np.float32), (R,)).copy()
        self._R.ref_crf_add_layer(self.h, region, R, e, c)

    @property
    def R(self):
        return self._R.ref_crf_R(self.h)

    def apply(self, k, x, which=0):
        x = _c(x, np.float32)
        out = np.zeros_like(x)
        self._R.ref_crf_apply(self.h, k, x, out, which)
        return out

    def add_superpixel(self, img, sr, rr, minreg, const, norm):
        """DenseCRF2D::addSuperPixel (mean-shift); returns (region map, e_r, c_r) of all layers"""
        img = _c(img, np.uint8)
        R = self._R.ref_crf_add_superpixel(self.h, img.ctypes.data_as(C.c_void_p), sr, rr, minreg, const, norm)
        reg = np.full(self.N, -1, np.int32)
        e = np.zeros(R, np.float32)
        c = np.zeros(R, np.float32)
        self._R.ref_crf_layers(self.h, reg, e, c)
        return reg, e, c

    def energy_higher_order(self, Q):
        """compute_energy_higher_order(Q) (uses the z_labels_ of the last qp())"""
        return self._R.ref_crf_energy_higher_order(self.h, _c(Q, np.float32))

    def apply_bf(self, k, which, ind=None, Q=None):
        """0 apply_bf(Q), 1 apply_upper_minus_lower_bf(ind), 2 _bf_ord(ind, Q), 3 _dc(ind)"""
        out = np.zeros((self.N, self.L), np.float32)
        ind = None if ind is None else _c(ind, np.int32)
        Q = None if Q is None else _c(Q, np.float32)
        p = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)  # noqa: E731
        self._R.ref_crf_apply_bf(self.h, k, which, p(ind), p(Q), out)
        return out

    def apply_ord(self, k, q):
        q = _c(q, np.float32)
        out = np.zeros_like(q)
        if self._R.ref_crf_apply_ord(self.h, k, q, out) != 0:
            raise RuntimeError(self._R.ref_last_error().decode())
        return out

    def unary_init(self):
        Q = np.zeros((self.N, self.L), np.float32)
        self._R.ref_crf_unary_init(self.h, Q)
        return Q

    def qp(self, init):
        init = _c(init, np.float32)
        out = np.zeros_like(init)
        self._R.ref_crf_qp(self.h, init, out)
        return out

    def zlabels(self):
        z = np.zeros(self.R * self.L, np.float32)
        self._R.ref_crf_zlabels(self.h, z)
        return z.reshape(self.R, self.L)

    def lp(self, init, params: LPParams | None = None):
        params = params or LPParams()
        fp, ip = params.packed()
        init = _c(init, np.float32)
        out = np.zeros_like(init)
        if self._R.ref_crf_lp(self.h, init, out, fp, ip) != 0:
            raise RuntimeError(self._R.ref_last_error().decode())
        return out

    def energy_true(self, lab):
        return self._R.ref_crf_energy_true(self.h, _c(lab, np.int32))

    def energy_ho(self, lab):
        return self._R.ref_crf_energy_ho(self.h, _c(lab, np.int32))

    def lp_energy(self, Q):
        return self._R.ref_crf_lp_energy(self.h, _c(Q, np.float32))

    def qp_value(self, Q):
        return self._R.ref_crf_qp_value(self.h, _c(Q, np.float32))

    def ucondgrad(self, Q):
        Q = _c(Q, np.float32)
        out = np.zeros_like(Q)
        self._R.ref_crf_ucondgrad(self.h, Q, out)
        return out

    def current_map(self, Q):
        Q = _c(Q, np.float32)
        lab = np.zeros(self.N, np.int32)
        self._R.ref_crf_current_map(self.h, Q, lab)
        return lab


def ref_feasible(Q):
    Q = _c(Q, np.float32)
    out = np.zeros_like(Q)
    reflib().ref_feasible(Q, out, Q.shape[0], Q.shape[1])
    return out


def ref_renormalize(Q):
    Q = _c(Q, np.float32).copy()
    reflib().ref_renormalize(Q, Q.shape[0], Q.shape[1])
    return Q


_refs = None


def refs_available() -> bool:
    return os.path.exists(REFS_SO)


class RefScalarLattice(RefLattice):
    """the reference's Permutohedral as its portable build computes it"""

    def __init__(self, feat):
        global _refs
        if _refs is None:
            if not refs_available():
                raise RuntimeError("portable reference build missing (make -C oracle ref-scalar)")
            R = C.CDLL(REFS_SO)
            _sig(R, "refs_lattice_new", _vp, _f32p, _i, _i)
            _sig(R, "refs_lattice_free", None, _vp)
            _sig(R, "refs_lattice_M", _i, _vp)
            _sig(R, "refs_lattice_get", None, _vp, _i32p, _i32p, _f32p, _i32p, _i)
            _sig(R, "refs_lattice_compute", None, _vp, _f32p, _f32p, _i, _i, _i)
            _sig(R, "refs_lattice_ordered", None, _vp, _f32p, _f32p, _i, _i)
            _refs = R
        feat = _c(feat, np.float32)
        self.N, self.d = feat.shape
        self._S = _refs
        self.h = self._S.refs_lattice_new(feat, self.N, self.d)
        self.M = self._S.refs_lattice_M(self.h)
        self.Npad = self.N

    def __del__(self):
        if getattr(self, "h", None):
            self._S.refs_lattice_free(self.h)
            self.h = None

    def arrays(self):
        d1 = self.d + 1
        off = np.zeros(self.N * d1, np.int32)
        rk = np.zeros(self.N * d1, np.int32)
        bc = np.zeros(self.N * d1, np.float32)
        nbr = np.zeros(max(d1 * self.M * 2, 1), np.int32)
        self._S.refs_lattice_get(self.h, off, rk, bc, nbr, self.N)
        return dict(offset=off.reshape(self.N, d1), rank=rk.reshape(self.N, d1),
                    bary=bc.reshape(self.N, d1), nbr=nbr[: d1 * self.M * 2].reshape(d1, self.M, 2))

    def filter(self, x, reverse=False):
        x = _c(x, np.float32)
        V = x.shape[1] if x.ndim == 2 else 1
        out = np.zeros_like(x)
        self._S.refs_lattice_compute(self.h, x, out, V, self.N, int(reverse))
        return out

    def ordered(self, q):
        q = _c(q, np.float32)
        out = np.zeros_like(q)
        self._S.refs_lattice_ordered(self.h, q, out, q.shape[1], self.N)
        return out
