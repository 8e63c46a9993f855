"""Pins the oracle's label compatibilities, PairwisePotential::gradient and
the unrolled mean field with its adjoint (oracle/dcrf_oracle.cpp).

  Diagonal / Matrix compatibility (apply, applyTranspose, no-norm copy)
      fixtures from the compiled reference (compat_31x23_L5.npz, 1e-6).
  PairwisePotential::gradient for Potts, Diagonal and Matrix
      the same fixtures (the reference sums in float, the oracle in FP64).
  QP with a Matrix term under the reference stopping rule
      fixture Q (1e-5) and the labelling's true energy.
  mean field forward
      T = 0 is the pinned softmax of -U; one update equals softmax of -U minus
      the pinned potentials applied to Q0.
  mean-field adjoint (FP64)
      central finite differences of the FP64 forward, per parameter of every
      compatibility kind and along random directions in U.
"""

import numpy as np
import pytest

import oracle as O
from conftest import rel_err

COMPATS = ["potts", "diagonal", "matrix"]


def _params(g, kind):
    return {"potts": [3.0], "diagonal": g["v"], "matrix": g["m"]}[kind]


def _compat_crf(g, kind):
    N, L = g["x"].shape
    c = O.CRF(N, L)
    c.set_unary(g["U"])
    c.add_pairwise_compat(g["feat"], kind, _params(g, kind))
    return c


@pytest.mark.parametrize("kind", COMPATS)
def test_compat_potentials_match_reference(golden, kind):
    g = golden("compat_31x23_L5.npz")
    c = _compat_crf(g, kind)
    assert rel_err(c.apply(0, g["x"]), g[f"{kind}_apply"]) < 1e-6
    assert rel_err(c.apply(0, g["x"], transpose=True), g[f"{kind}_applyT"]) < 1e-6
    assert rel_err(c.apply(0, g["x"], no_norm=True), g[f"{kind}_nn"]) < 1e-6


@pytest.mark.parametrize("kind", COMPATS)
def test_pairwise_gradient_matches_reference(golden, kind):
    g = golden("compat_31x23_L5.npz")
    c = _compat_crf(g, kind)
    ref = g[f"{kind}_grad"].astype(np.float64)
    got = c.pairwise_gradient(0, g["b"], g["x"])
    assert got.shape == ref.shape
    assert rel_err(got, ref) < 1e-5
    # FP64 instantiation agrees with the float one
    got64 = c.pairwise_gradient(0, g["b"].astype(np.float64), g["x"].astype(np.float64), fp64=True)
    assert rel_err(got64, ref) < 1e-4


def test_matrix_gradient_layout_is_packed_upper_triangle(golden):
    """Matrix gradient = upper triangle (row by row) of b (KQ)^T + its
    transpose off the diagonal: checked against a dense product."""
    g = golden("compat_31x23_L5.npz")
    c = _compat_crf(g, "matrix")
    L = g["x"].shape[1]
    kq = c.apply(0, g["x"].astype(np.float64), fp64=True)  # W K x
    Wsym = 0.5 * (g["m"] + g["m"].T).astype(np.float64)
    kx = np.linalg.solve(Wsym, kq.T).T  # K x
    G = g["b"].astype(np.float64).T @ kx
    want = [G[i, j] + (G[j, i] if i != j else 0.0) for i in range(L) for j in range(i, L)]
    assert rel_err(c.pairwise_gradient(0, g["b"], g["x"]), want) < 1e-4


def test_qp_with_matrix_term_matches_reference(golden):
    g = golden("compat_31x23_L5.npz")
    N, L = g["x"].shape
    c = O.CRF(N, L)
    c.set_unary(g["U"])
    c.add_pairwise_compat(g["feat_g"], "potts", [3.0])
    c.add_pairwise_compat(g["feat"], "matrix", g["qp_m"])
    Q, st = c.qp(O.unary_init(g["U"]), max_iter=0, ref_stop=True)
    assert rel_err(Q, g["qp_Q"]) < 1e-5
    lab = c.current_map(Q)
    assert np.mean(lab == g["qp_lab"]) >= 0.999
    assert abs(c.energy_true(g["qp_lab"]) - float(g["qp_energy_true"])) <= 1e-5 * abs(float(g["qp_energy_true"]))


def _mf_crf(g, potts=3.0, mat=None, diag=None):
    N, L = g["x"].shape
    c = O.CRF(N, L)
    c.set_unary(g["U"])
    c.add_pairwise_compat(g["feat_g"], "potts", [potts])
    c.add_pairwise_compat(g["feat"], "matrix", mat if mat is not None else _sym(g))
    c.add_pairwise_compat(g["feat_g"], "diagonal", diag if diag is not None else 0.5 * g["v"])
    return c


def _sym(g):
    m = 0.3 * g["m"].astype(np.float32)
    return (0.5 * (m + m.T)).astype(np.float32)


def test_meanfield_forward_pieces(golden):
    g = golden("compat_31x23_L5.npz")
    c = _mf_crf(g)
    U = g["U"].astype(np.float64)
    e = np.exp(-U - (-U).max(1, keepdims=True))
    Q0 = e / e.sum(1, keepdims=True)
    np.testing.assert_allclose(c.meanfield(0, fp64=True), Q0, rtol=1e-12, atol=1e-15)
    a = -U - sum(c.apply(k, Q0, fp64=True) for k in range(3))
    e = np.exp(a - a.max(1, keepdims=True))
    np.testing.assert_allclose(c.meanfield(1, fp64=True), e / e.sum(1, keepdims=True), rtol=1e-12, atol=1e-15)
    # the float instantiation follows the FP64 one
    assert np.abs(c.meanfield(3) - c.meanfield(3, fp64=True)).max() < 1e-5


def _xent(Q, lab):
    return -np.log(Q[np.arange(len(lab)), lab]).sum()


def _xent_grad(Q, lab):
    d = np.zeros_like(Q)
    d[np.arange(len(lab)), lab] = -1.0 / Q[np.arange(len(lab)), lab]
    return d


def test_meanfield_adjoint_matches_finite_differences(golden):
    g = golden("compat_31x23_L5.npz")
    N, L = g["x"].shape
    T = 3
    rng = np.random.default_rng(11)
    lab = rng.integers(0, L, N)
    c = _mf_crf(g)
    Q = c.meanfield(T, fp64=True)
    _, dU, gs = c.meanfield_grad(T, _xent_grad(Q, lab))
    assert [x.size for x in gs] == [1, L * (L + 1) // 2, L]

    # U along random directions (the perturbation is rounded to float32 as the
    # unary is stored in float; the exact rounded difference is used)
    U0 = g["U"].astype(np.float64)
    for _ in range(3):
        d = rng.standard_normal((N, L))
        Up, Um = (U0 + 1e-3 * d).astype(np.float32), (U0 - 1e-3 * d).astype(np.float32)
        c.set_unary(Up)
        lp = _xent(c.meanfield(T, fp64=True), lab)
        c.set_unary(Um)
        lm = _xent(c.meanfield(T, fp64=True), lab)
        an = (dU * (Up.astype(np.float64) - Um.astype(np.float64))).sum()
        assert abs((lp - lm) - an) <= 1e-4 * abs(an)
    c.set_unary(g["U"])

    def loss_with(**kw):
        cc = _mf_crf(g, **kw)
        return _xent(cc.meanfield(T, fp64=True), lab)

    # Potts weight
    wp, wm = np.float32(3.0 + 1e-3), np.float32(3.0 - 1e-3)
    fd = (loss_with(potts=float(wp)) - loss_with(potts=float(wm))) / (float(wp) - float(wm))
    assert abs(fd - gs[0][0]) <= 1e-4 * abs(gs[0][0])
    # packed matrix entries (i, j) and (j, i) move together
    W = _sym(g)
    tri = [(i, j) for i in range(L) for j in range(i, L)]
    for k in (0, 3, 7, len(tri) - 1):
        i, j = tri[k]
        Wp, Wm = W.copy(), W.copy()
        Wp[i, j] = Wp[j, i] = np.float32(W[i, j] + 1e-3)
        Wm[i, j] = Wm[j, i] = np.float32(W[i, j] - 1e-3)
        fd = (loss_with(mat=Wp) - loss_with(mat=Wm)) / (float(Wp[i, j]) - float(Wm[i, j]))
        assert abs(fd - gs[1][k]) <= 1e-4 * max(abs(gs[1][k]), 1e-3 * np.abs(gs[1]).max())
    # diagonal entries
    v = (0.5 * g["v"]).astype(np.float32)
    for k in range(L):
        vp, vm = v.copy(), v.copy()
        vp[k] = np.float32(v[k] + 1e-3)
        vm[k] = np.float32(v[k] - 1e-3)
        fd = (loss_with(diag=vp) - loss_with(diag=vm)) / (float(vp[k]) - float(vm[k]))
        assert abs(fd - gs[2][k]) <= 1e-4 * max(abs(gs[2][k]), 1e-3 * np.abs(gs[2]).max())
