"""GPU label compatibilities, PairwisePotential::gradient and the unrolled
mean field with its adjoint, against the reference fixtures, the FP64 oracle
and finite differences.

Tolerances: potentials 1e-5 relative; gradients 1e-4 relative to the
reference's float result; mean-field Q within 1e-5 of the FP64 oracle; dU
within 1e-4 and parameter gradients within 1e-3 (relative to the largest
entry) of the FP64 oracle's adjoint; central finite differences (eps = 1e-3)
of the GPU forward within 1e-3 relative.
"""

import numpy as np
import pytest

import oracle as O
from conftest import rel_err

pytestmark = pytest.mark.gpu

COMPATS = ["potts", "diagonal", "matrix"]


def _fn(D, kind, par):
    return {"potts": lambda: D.PottsCompatibility(float(par[0])), "diagonal": lambda: D.DiagonalCompatibility(par),
            "matrix": lambda: D.MatrixCompatibility(par)}[kind]()


def _params(g, kind):
    return {"potts": [3.0], "diagonal": g["v"], "matrix": g["m"]}[kind]


def _gpu_compat(g, kind):
    import densecrf_ho_amd as D
    N, L = g["x"].shape
    c = D.DenseCRF(N, L)
    c.setUnaryEnergy(g["U"])
    c.addPairwiseEnergy(g["feat"], _fn(D, kind, _params(g, kind)))
    return c


@pytest.mark.parametrize("kind", COMPATS)
def test_compat_potentials_vs_reference(golden, kind):
    g = golden("compat_31x23_L5.npz")
    print("compat", kind, flush=True)
    c = _gpu_compat(g, kind)
    assert rel_err(c.apply(0, g["x"]), g[f"{kind}_apply"]) < 1e-5
    assert rel_err(c.apply(0, g["x"], transpose=True), g[f"{kind}_applyT"]) < 1e-5
    assert rel_err(c.apply(0, g["x"], no_norm=True), g[f"{kind}_nn"]) < 1e-5
    assert rel_err(c.pairwise_gradient(0, g["b"], g["x"]), g[f"{kind}_grad"]) < 1e-4


@pytest.mark.parametrize("kind", ["diagonal", "matrix"])
def test_compat_ordered_filter_vs_oracle(golden, kind):
    g = golden("compat_31x23_L5.npz")
    N, L = g["x"].shape
    c = _gpu_compat(g, kind)
    o = O.CRF(N, L)
    o.add_pairwise_compat(g["feat"], kind, _params(g, kind))
    q = g["x"]
    assert rel_err(c.apply_upper_minus_lower_ord(0, q), o.apply_ord(0, q)) < 1e-5


def test_compat_parameters_roundtrip(golden):
    g = golden("compat_31x23_L5.npz")
    L = g["x"].shape[1]
    c = _gpu_compat(g, "matrix")
    W = 0.5 * (g["m"] + g["m"].T)
    want = np.array([W[i, j] for i in range(L) for j in range(i, L)], np.float32)
    np.testing.assert_allclose(c.pairwise_parameters(0), want, rtol=1e-7)
    before = c.apply(0, g["x"])
    c.set_pairwise_parameters(0, 2 * want)
    np.testing.assert_allclose(c.apply(0, g["x"]), 2 * before, rtol=1e-5, atol=1e-5 * np.abs(before).max())
    # the no-norm copy keeps the weight it was created with
    np.testing.assert_allclose(c.apply(0, g["x"], no_norm=True), g["matrix_nn"], rtol=1e-5,
                               atol=1e-6 * np.abs(g["matrix_nn"]).max())


def test_qp_with_matrix_term_vs_reference(golden):
    import densecrf_ho_amd as D
    g = golden("compat_31x23_L5.npz")
    N, L = g["x"].shape
    c = D.DenseCRF(N, L)
    c.setUnaryEnergy(g["U"])
    c.addPairwiseEnergy(g["feat_g"], D.PottsCompatibility(3.0))
    c.addPairwiseEnergy(g["feat"], D.MatrixCompatibility(g["qp_m"]))
    Q = c.qp_inference_super_pixels_non_convex(c.unary_init())
    assert rel_err(Q, g["qp_Q"]) < 1e-4
    lab = c.currentMap(Q)
    assert np.mean(lab == g["qp_lab"]) >= 0.999
    e = c.assignment_energy_true(g["qp_lab"].astype(np.int32))
    assert abs(e - float(g["qp_energy_true"])) <= 1e-4 * abs(float(g["qp_energy_true"]))


def _sym(g):
    m = 0.3 * g["m"].astype(np.float32)
    return (0.5 * (m + m.T)).astype(np.float32)


def _mf_pair(g):
    import densecrf_ho_amd as D
    N, L = g["x"].shape
    c = D.DenseCRF(N, L)
    c.setUnaryEnergy(g["U"])
    c.addPairwiseEnergy(g["feat_g"], D.PottsCompatibility(3.0))
    c.addPairwiseEnergy(g["feat"], D.MatrixCompatibility(_sym(g)))
    c.addPairwiseEnergy(g["feat_g"], D.DiagonalCompatibility(0.5 * g["v"]))
    o = O.CRF(N, L)
    o.set_unary(g["U"])
    o.add_pairwise_compat(g["feat_g"], "potts", [3.0])
    o.add_pairwise_compat(g["feat"], "matrix", _sym(g))
    o.add_pairwise_compat(g["feat_g"], "diagonal", 0.5 * g["v"])
    return c, o


def test_meanfield_forward_and_adjoint_vs_fp64_oracle(golden):
    g = golden("compat_31x23_L5.npz")
    N, L = g["x"].shape
    T = 5
    c, o = _mf_pair(g)
    Q = c.meanfield(T)
    Qo = o.meanfield(T, fp64=True)
    assert np.abs(Q - Qo).max() < 1e-5
    lab = np.random.default_rng(3).integers(0, L, N).astype(np.int32)
    loss, dQ = c.cross_entropy(Q, lab)
    assert abs(loss + np.log(Qo[np.arange(N), lab]).sum()) < 1e-4 * abs(loss)
    dU, gs = c.meanfield_backward(dQ)
    dQo = np.zeros((N, L))
    dQo[np.arange(N), lab] = -1.0 / Qo[np.arange(N), lab]
    _, dUo, gso = o.meanfield_grad(T, dQo)
    assert rel_err(dU, dUo) < 1e-4
    for a, b in zip(gs, gso):
        assert rel_err(a, b) < 1e-3


def test_meanfield_gradients_vs_finite_differences(golden):
    """central differences (eps = 1e-3) of the GPU forward + loss"""
    g = golden("compat_31x23_L5.npz")
    N, L = g["x"].shape
    T = 5
    c, _ = _mf_pair(g)
    lab = np.random.default_rng(4).integers(0, L, N).astype(np.int32)

    def loss():
        return c.cross_entropy(c.meanfield(T), lab)[0]

    Q = c.meanfield(T)
    _, dQ = c.cross_entropy(Q, lab)
    dU, gs = c.meanfield_backward(dQ)
    # Potts weight of potential 0 and two packed entries of the matrix term
    for k, idx in ((0, 0), (1, 0), (1, 6), (2, 3)):
        p0 = c.pairwise_parameters(k)
        pp, pm = p0.copy(), p0.copy()
        pp[idx] = np.float32(p0[idx] + 1e-3)
        pm[idx] = np.float32(p0[idx] - 1e-3)
        c.set_pairwise_parameters(k, pp)
        lp = loss()
        c.set_pairwise_parameters(k, pm)
        lm = loss()
        c.set_pairwise_parameters(k, p0)
        fd = (lp - lm) / (float(pp[idx]) - float(pm[idx]))
        print("param", k, idx, fd, gs[k][idx], flush=True)
        assert abs(fd - gs[k][idx]) <= 1e-3 * max(abs(gs[k][idx]), 1e-2 * np.abs(gs[k]).max())
    # U along a random direction
    d = np.random.default_rng(5).standard_normal((N, L))
    U0 = g["U"].astype(np.float64)
    Up, Um = (U0 + 1e-3 * d).astype(np.float32), (U0 - 1e-3 * d).astype(np.float32)
    c.setUnaryEnergy(Up)
    lp = loss()
    c.setUnaryEnergy(Um)
    lm = loss()
    an = float((dU.astype(np.float64) * (Up.astype(np.float64) - Um.astype(np.float64))).sum())
    assert abs((lp - lm) - an) <= 1e-3 * abs(an)


def test_autograd_wrapper_matches_adjoint(golden):
    """the torch.autograd wrapper (CPU tensors) returns the engine's adjoint"""
    import densecrf_ho_amd as D
    D.load()  # the engine's HIP runtime is bound before torch is imported
    torch = pytest.importorskip("torch")
    from densecrf_ho_amd.autograd import meanfield
    g = golden("compat_31x23_L5.npz")
    N, L = g["x"].shape
    T = 3
    c, _ = _mf_pair(g)
    params = [torch.tensor(c.pairwise_parameters(k), requires_grad=True) for k in range(3)]
    U = torch.tensor(g["U"], requires_grad=True)
    lab = torch.from_numpy(np.random.default_rng(6).integers(0, L, N))
    Q = meanfield(U, params, c, T)
    loss = -torch.log(Q[torch.arange(N), lab]).sum()
    loss.backward()
    dQ = np.zeros((N, L), np.float32)
    Qn = Q.detach().numpy()
    dQ[np.arange(N), lab.numpy()] = -1.0 / Qn[np.arange(N), lab.numpy()]
    c.meanfield(T)
    dU, gs = c.meanfield_backward(dQ)
    np.testing.assert_allclose(U.grad.numpy(), dU, rtol=1e-5, atol=1e-6 * np.abs(dU).max())
    for p, gk in zip(params, gs):
        np.testing.assert_allclose(p.grad.numpy(), gk, rtol=1e-5, atol=1e-6 * np.abs(gk).max())


def test_backward_after_another_forward_or_model_change_raises(golden):
    """the engine keeps one forward's checkpoints: two forwards then one
    backward through the autograd wrapper raises, and so does the C ABI's
    backward after setUnaryEnergy or set_pairwise_parameters; a fresh forward
    makes backward work again"""
    import densecrf_ho_amd as D
    D.load()
    torch = pytest.importorskip("torch")
    from densecrf_ho_amd.autograd import meanfield
    g = golden("compat_31x23_L5.npz")
    N, L = g["x"].shape
    c, _ = _mf_pair(g)
    params = [torch.tensor(c.pairwise_parameters(k), requires_grad=True) for k in range(3)]
    U = torch.tensor(g["U"], requires_grad=True)
    Q1 = meanfield(U, params, c, 2)
    Q2 = meanfield(U, params, c, 3)
    with pytest.raises(RuntimeError, match="another meanfield ran"):
        (Q1.sum() + Q2.sum()).backward()
    dQ = np.ones((N, L), np.float32)
    c.meanfield(2)
    c.setUnaryEnergy(g["U"])
    with pytest.raises(D.EngineError, match="changed after the last meanfield"):
        c.meanfield_backward(dQ)
    c.meanfield(2)
    c.set_pairwise_parameters(0, c.pairwise_parameters(0))
    with pytest.raises(D.EngineError, match="changed after the last meanfield"):
        c.meanfield_backward(dQ)
    c.meanfield(2)
    dU, _ = c.meanfield_backward(dQ)
    assert np.isfinite(dU).all()
