"""The reference's unary interface (include/unary.h:31-57) and the other
pieces of the DenseCRF boundary: LogisticUnaryEnergy get() / gradient() on the
matrix cores against the compiled reference's fixture (1e-5 relative; exact on
the case whose sums are exact in float), setUnaryEnergy(L, f) and
setUnaryEnergy(UnaryEnergy) in the solvers, addPairwiseEnergy(PairwisePotential)
and the kernel type of addPairwiseEnergy."""

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("i", [0, 1, 2])
def test_logistic_unary_vs_reference(golden, i):
    import densecrf_ho_amd as D
    g = golden("logistic_unary.npz")
    Lw, f, b = g[f"Lw{i}"], g[f"f{i}"], g[f"b{i}"]
    print("logistic unary case", i, "L x F x N =", Lw.shape[1], Lw.shape[0], f.shape[0], flush=True)
    u = D.LogisticUnaryEnergy(Lw, f)
    U = u.get()
    G = u.gradient(b)
    want_U = f.astype(np.float64) @ Lw.astype(np.float64)
    want_G = (b.astype(np.float64).T @ f.astype(np.float64)).T.reshape(-1)
    print("get rel", rel_err(U, g[f"U{i}"]), "gradient rel", rel_err(G, g[f"g{i}"]), flush=True)
    if i == 2:  # multiples of 1/8: every product and sum is exact in float
        np.testing.assert_array_equal(U, g[f"U{i}"])
        np.testing.assert_array_equal(G, g[f"g{i}"])
    assert rel_err(U, g[f"U{i}"]) <= 1e-5 and rel_err(U, want_U) <= 1e-5
    assert rel_err(G, g[f"g{i}"]) <= 1e-5 and rel_err(G, want_G) <= 1e-5
    np.testing.assert_array_equal(u.parameters(), Lw.reshape(-1))
    # device buffers give the same bits
    fd = D.DeviceArray.from_numpy(f)
    ud = D.LogisticUnaryEnergy(Lw, fd)
    np.testing.assert_array_equal(ud.get().numpy(), U)
    np.testing.assert_array_equal(ud.gradient(D.DeviceArray.from_numpy(b)), G)


def test_logistic_unary_1080p_vs_numpy():
    """21 labels x 16 features at 1920x1080 (many pixel tiles) against an FP64 product"""
    import densecrf_ho_amd as D
    rng = np.random.default_rng(62)
    N, F, L = 1920 * 1080, 16, 21
    f = rng.standard_normal((N, F)).astype(np.float32)
    Lw = rng.standard_normal((F, L)).astype(np.float32)
    b = rng.standard_normal((N, L)).astype(np.float32)
    u = D.LogisticUnaryEnergy(Lw, f)
    assert rel_err(u.get(), f.astype(np.float64) @ Lw.astype(np.float64)) <= 1e-5
    want = (b.astype(np.float64).T @ f.astype(np.float64)).T.reshape(-1)
    assert rel_err(u.gradient(b), want) <= 1e-5


def _image_crf(D, wl, unary=None, f=None):
    c = D.DenseCRF2D(wl.W, wl.H, wl.L)
    c.setUnaryEnergy(wl.U if unary is None else unary, f)
    c.addPairwiseGaussian(3, 3, D.PottsCompatibility(3))
    c.addPairwiseBilateral(50, 50, 15, 15, 15, wl.img, D.PottsCompatibility(5))
    return c


def test_set_unary_logistic_in_the_solvers():
    """setUnaryEnergy(L, f) solves exactly as a ConstUnaryEnergy of its get();
    setParameters on the CRF's unary object reaches the next solve, as the
    reference re-reads unary_->get() in each solve (src/densecrf.cpp:532)"""
    import densecrf_ho_amd as D
    from densecrf_ho_amd import synthetic as S
    wl = S.Workload(64, 48, 6, 3)
    F = 6
    px = np.arange(wl.N)
    feat = np.stack([np.ones(wl.N), (px % wl.W) / wl.W, (px // wl.W) / wl.H] +
                    [wl.img[c::3] / 255.0 for c in range(3)], 1).astype(np.float32)
    Lw = (0.5 * (((np.arange(wl.L)[None, :] * 5 + np.arange(F)[:, None] * 3) % 7) - 3)).astype(np.float32)
    lc = _image_crf(D, wl, Lw, feat)
    assert isinstance(lc.getUnaryEnergy(), D.LogisticUnaryEnergy)
    cc = _image_crf(D, wl, D.ConstUnaryEnergy(lc.getUnaryEnergy().get()))
    q1 = lc.qp_inference_super_pixels_non_convex(lc.unary_init(), max_iter=5, ref_stop=False)
    q2 = cc.qp_inference_super_pixels_non_convex(cc.unary_init(), max_iter=5, ref_stop=False)
    np.testing.assert_array_equal(q1, q2)
    lab = np.ascontiguousarray(q1.argmax(1).astype(np.int32))
    assert lc.assignment_energy_true(lab) == cc.assignment_energy_true(lab)
    u = lc.getUnaryEnergy()
    u.setParameters(2 * u.parameters())
    e2 = lc.assignment_energy_true(lab)
    c2 = _image_crf(D, wl, D.ConstUnaryEnergy(D.LogisticUnaryEnergy(2 * Lw, feat).get()))
    assert e2 == c2.assignment_energy_true(lab) != cc.assignment_energy_true(lab)


def test_user_unary_energy_is_read_each_solve():
    """a UnaryEnergy subclass without a version is read again before every
    computation (the reference calls get() in each solve)"""
    import densecrf_ho_amd as D
    from densecrf_ho_amd import synthetic as S
    wl = S.Workload(32, 24, 4, 5)

    class Scaled(D.UnaryEnergy):
        def __init__(self, U):
            self.U, self.k = U, 1.0

        def get(self):
            return (self.k * self.U).astype(np.float32)

    u = Scaled(wl.U)
    c = _image_crf(D, wl, u)
    lab = np.zeros(wl.N, np.int32)
    e1 = c.assignment_energy_true(lab)
    u.k = 3.0
    e3 = c.assignment_energy_true(lab)
    ref = _image_crf(D, wl, (3.0 * wl.U).astype(np.float32)).assignment_energy_true(lab)
    assert e3 == ref and e3 != e1


def test_add_pairwise_potential_object(golden):
    """addPairwiseEnergy(PairwisePotential) adds the same term as
    addPairwiseEnergy(features, function): equal filters, parameters and mean
    field; the standalone potential applies its Matrix compatibility as the
    reference's fixture does"""
    import densecrf_ho_amd as D
    g = golden("compat_31x23_L5.npz")
    N, L = g["x"].shape
    a, b = D.DenseCRF(N, L), D.DenseCRF(N, L)
    for c in (a, b):
        c.setUnaryEnergy(g["U"])
    a.addPairwiseEnergy(g["feat"], D.MatrixCompatibility(g["m"]))
    pw = D.PairwisePotential(g["feat"], D.MatrixCompatibility(g["m"]), D.DIAG_KERNEL)
    assert rel_err(pw.apply(g["x"]), g["matrix_apply"]) < 1e-5
    b.addPairwiseEnergy(pw)
    assert pw._h is None  # consumed, as the reference takes ownership
    np.testing.assert_array_equal(a.apply(0, g["x"]), b.apply(0, g["x"]))
    np.testing.assert_array_equal(a.apply(0, g["x"], no_norm=True), b.apply(0, g["x"], no_norm=True))
    np.testing.assert_array_equal(a.pairwise_parameters(0), b.pairwise_parameters(0))
    np.testing.assert_array_equal(a.meanfield(2), b.meanfield(2))
    dpw = D.PairwisePotential(g["feat"], D.DiagonalCompatibility(g["v"]))
    assert rel_err(dpw.apply(g["x"]), g["diagonal_apply"]) < 1e-5


def test_kernel_type_of_add_pairwise_energy(golden):
    """addPairwiseEnergy's kernel_type fixes the kernel parameters of the term:
    CONST none, DIAG d, FULL d*d (the DIAG gradient equals the standalone
    potential's, pinned against the reference's kernel gradient fixture)"""
    import densecrf_ho_amd as D
    g = golden("compat_31x23_L5.npz")
    N, L = g["x"].shape
    c = D.DenseCRF(N, L)
    c.setUnaryEnergy(g["U"])
    d = g["feat"].shape[1]
    for kt in (D.CONST_KERNEL, D.DIAG_KERNEL, D.FULL_KERNEL):
        c.addPairwiseEnergy(g["feat"], D.PottsCompatibility(3.0), kernel_type=kt)
    sizes = [c.pairwise_kernel_gradient(k, g["x"], g["x"]).size for k in range(3)]
    assert sizes == [0, d, d * d]
    pw = D.PairwisePotential(g["feat"], D.PottsCompatibility(3.0), D.DIAG_KERNEL)
    np.testing.assert_allclose(c.pairwise_kernel_gradient(1, g["x"], g["x"]), pw.kernelGradient(g["x"], g["x"]),
                               rtol=1e-6)
