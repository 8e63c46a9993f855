"""GPU potentials, QP and energies against the reference fixtures and the
FP64 oracle.  Tolerances: potentials 1e-5 relative; QP energies 1e-4
relative to the FP64 oracle; label agreement >= 99.9%."""

import json
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, rel_err

pytestmark = pytest.mark.gpu


def gpu_crf(g):
    import densecrf_ho_amd as D
    W, H, L = int(g["W"]), int(g["H"]), int(g["L"])
    c = D.DenseCRF2D(W, H, L)
    c.setUnaryEnergy(g["U"])
    w = g["w"] if "w" in g else (3.0, 5.0)
    c.addPairwiseEnergy(g["feat0"], D.PottsCompatibility(float(w[0])))
    c.addPairwiseEnergy(g["feat1"], D.PottsCompatibility(float(w[1])))
    return c


def test_potentials_vs_reference(golden):
    g = golden("crf_31x23_L5.npz")
    c = gpu_crf(g)
    for k in range(2):
        print("potential", k, flush=True)
        assert rel_err(c.apply(k, g["x"]), g[f"apply{k}"]) < 1e-5
        assert rel_err(c.apply(k, g["x"], transpose=True), g[f"applyT{k}"]) < 1e-5
        assert rel_err(c.apply(k, g["x"], no_norm=True), g[f"applyNN{k}"]) < 1e-5
        assert rel_err(c.apply_upper_minus_lower_ord(k, g[f"ordq{k}"]), g[f"ord{k}"]) < 1e-5


def test_unary_init(golden):
    g = golden("crf_31x23_L5.npz")
    c = gpu_crf(g)
    np.testing.assert_allclose(c.unary_init(), g["Q0"], rtol=2e-6, atol=1e-7)


@pytest.mark.parametrize("case", ["crf_31x23_L5.npz", "qp_convex_31x23_L5.npz", "qp_ties_31x23_L5.npz"])
def test_qp_reference_stopping_vs_reference(golden, case):
    g = golden(case)
    print("case", case, flush=True)
    c = gpu_crf(g)
    Q = c.qp_inference_super_pixels_non_convex(g["Q0"])
    assert rel_err(Q, g["Qqp"]) < 1e-4
    assert (Q.argmax(1) == g["Qqp"].argmax(1)).mean() >= 0.999


def test_energies_vs_reference(golden):
    g = golden("crf_31x23_L5.npz")
    c = gpu_crf(g)
    lab = c.currentMap(g["Qqp"])
    np.testing.assert_array_equal(lab, g["lab_qp"])
    e = c.assignment_energy_true(lab)
    assert abs(e - float(g["e_true"])) <= 1e-5 * abs(float(g["e_true"]))
    v = c.compute_LR_QP_value(g["Qqp"])
    assert abs(v - float(g["e_qpvalue"])) <= 1e-5 * abs(float(g["e_qpvalue"]))


def _c0():
    from densecrf_ho_amd import synthetic as S
    return S.Workload(320, 213, 40, 0)


def test_config0_qp_fixed_iterations_vs_fp64_oracle():
    """320x213, 21 labels, 10 Frank-Wolfe iterations."""
    import densecrf_ho_amd as D
    wl = _c0()
    orc = O.CRF(wl.N, 21)
    orc.set_unary(wl.U)
    c = D.DenseCRF2D(320, 213, 21)
    c.setUnaryEnergy(wl.U)
    for f, w in wl.kernels:
        orc.add_pairwise(f, w)
        c.addPairwiseEnergy(f, D.PottsCompatibility(w))
    Q0 = O.unary_init(wl.U)
    Qd, std = orc.qp(Q0, max_iter=10, fp64=True)
    Qf, stf = orc.qp(Q0, max_iter=10)
    Qg, stg = c.qp_inference_super_pixels_non_convex(Q0, max_iter=10, ref_stop=False, stats=True)
    print("energies gpu", stg["energy"], "fp64", std["energy"], flush=True)
    # the line search stops early (num == 0 once the argmin no longer moves)
    assert stg["iters"] == std["iters"] <= 10
    np.testing.assert_allclose(stg["energy"], std["energy"], rtol=1e-4)
    assert (Qg.argmax(1) == Qd.argmax(1)).mean() >= 0.999
    assert (Qg.argmax(1) == Qf.argmax(1)).mean() >= 0.999
    eg = c.assignment_energy_true(c.currentMap(Qg))
    ed = orc.energy_true(orc.current_map(Qd), fp64=True)
    assert abs(eg - ed) <= 1e-4 * abs(ed)


def test_config0_reference_stopping_rule(golden):
    import densecrf_ho_amd as D
    with open(os.path.join(GOLDEN, "c0_320x213.json")) as f:
        ref = json.load(f)
    g = golden("c0_320x213.npz")
    wl = _c0()
    c = D.DenseCRF2D(320, 213, 21)
    c.setUnaryEnergy(wl.U)
    c.addPairwiseGaussian(3, 3, D.PottsCompatibility(3))
    c.addPairwiseBilateral(50, 50, 15, 15, 15, wl.img, D.PottsCompatibility(5))
    Q = c.qp_inference_super_pixels_non_convex(c.unary_init())
    lab = c.currentMap(Q)
    assert (lab == g["lab_qp"]).mean() >= 0.999
    e = c.assignment_energy_true(lab)
    assert abs(e - ref["e_true_qp"]) <= 1e-4 * abs(ref["e_true_qp"])


@pytest.mark.parametrize("L", [10, 11])
def test_potentials_wide_rows_vs_oracle(L):
    """Potentials at label counts whose rows end in a partial quad (L % 4 = 2
    or 3): normalised, transposed and no-norm filters, accumulating over two
    terms in the QP, against the FP64 oracle."""
    import densecrf_ho_amd as D
    from densecrf_ho_amd import synthetic as S
    wl = S.Workload(80, 60, 12, 9, L=L)
    c = D.DenseCRF2D(80, 60, L)
    c.setUnaryEnergy(wl.U)
    orc = O.CRF(wl.N, L)
    orc.set_unary(wl.U)
    for f, w in wl.kernels:
        c.addPairwiseEnergy(f, D.PottsCompatibility(w))
        orc.add_pairwise(f, w)
    x = np.random.default_rng(L).standard_normal((wl.N, L)).astype(np.float32)
    for k in range(2):
        print("L", L, "potential", k, flush=True)
        assert rel_err(c.apply(k, x), orc.apply(k, x, fp64=True)) < 1e-4
        assert rel_err(c.apply(k, x, transpose=True), orc.apply(k, x, fp64=True, transpose=True)) < 1e-4
        assert rel_err(c.apply(k, x, no_norm=True), orc.apply(k, x, fp64=True, no_norm=True)) < 1e-4
    Q0 = O.unary_init(wl.U)
    Qd, std = orc.qp(Q0, max_iter=5, fp64=True)
    Qg, stg = c.qp_inference_super_pixels_non_convex(Q0, max_iter=5, ref_stop=False, stats=True)
    np.testing.assert_allclose(stg["energy"], std["energy"], rtol=1e-4)
    assert (Qg.argmax(1) == Qd.argmax(1)).mean() >= 0.999


@pytest.mark.parametrize("compat", ["potts", "matrix"])
def test_restricted_ordered_potential_vs_oracle(golden, compat):
    """PairwisePotential::apply_upper_minus_lower_ord_restricted
    (src/pairwise.cpp:229-239): rout = 0, the restricted filter of the
    potential's lattice, then its compatibility (Potts -w, or the Matrix form)"""
    import densecrf_ho_amd as D
    g = golden("crf_31x23_L5.npz")
    W, H, L = int(g["W"]), int(g["H"]), int(g["L"])
    c = D.DenseCRF2D(W, H, L)
    c.setUnaryEnergy(g["U"])
    rng = np.random.default_rng(17)
    m = rng.standard_normal((L, L)).astype(np.float32)
    fn = D.PottsCompatibility(5.0) if compat == "potts" else D.MatrixCompatibility(m)
    c.addPairwiseEnergy(g["feat1"], fn)
    N = W * H
    q = rng.random((N, L)).astype(np.float32)
    pI = np.sort(rng.choice(N, N // 3, replace=False)).astype(np.int32)
    rq = rng.random((len(pI), L)).astype(np.float32)
    out = c.apply_upper_minus_lower_ord_restricted(0, rq, pI, q, True)
    r = O.Lattice(g["feat1"]).ordered_restricted(rq, pI, q, True)
    if compat == "potts":
        ref = np.float32(-5.0) * r
    else:
        Wm = ((m + m.T) / 2).astype(np.float32)
        ref = (r.astype(np.float64) @ Wm.astype(np.float64).T).astype(np.float32)
    assert rel_err(out, ref) < 1e-5
    rq2 = rng.random((len(pI), L)).astype(np.float32)
    out2 = c.apply_upper_minus_lower_ord_restricted(0, rq2, pI, None, False)
    lat = O.Lattice(g["feat1"])
    lat.ordered_restricted(rq, pI, q, True)
    r2 = lat.ordered_restricted(rq2, pI, None, False)
    ref2 = np.float32(-5.0) * r2 if compat == "potts" else \
        (r2.astype(np.float64) @ Wm.astype(np.float64).T).astype(np.float32)
    assert rel_err(out2, ref2) < 1e-5


def _bf_crf(g):
    import densecrf_ho_amd as D
    W, H, L = int(g["W"]), int(g["H"]), int(g["L"])
    c = D.DenseCRF2D(W, H, L)
    c.setUnaryEnergy(g["U"])
    c.addPairwiseEnergy(g["feat0"], D.PottsCompatibility(3.0))
    c.addPairwiseEnergy(g["feat1"], D.PottsCompatibility(5.0))
    return c


@pytest.mark.parametrize("k", [0, 1])
def test_brute_force_filters_vs_reference(golden, k):
    """apply_bf / apply_upper_minus_lower_bf / _bf_ord of potential k against
    the compiled reference's outputs (within 1e-5 of the largest value)"""
    g = golden("bf_12x9_L4.npz")
    c = _bf_crf(g)
    print("potential", k, flush=True)
    assert rel_err(c.apply_bf(k, g["Q"]), g[f"bf{k}"]) < 1e-5
    assert rel_err(c.apply_upper_minus_lower_bf(k, g["ind"]), g[f"uml{k}"]) < 1e-5
    assert rel_err(c.apply_upper_minus_lower_bf_ord(k, g["ind"], g["Q"]), g[f"umlord{k}"]) < 1e-5
    assert rel_err(c.apply_upper_minus_lower_bf_ord(k, g["ind_q"], g["Q"]), g[f"umlord_q{k}"]) < 1e-5


def test_brute_force_filters_vs_oracle_larger():
    """40x30 image, 21 labels, bilateral kernel: GPU brute-force filters
    against the oracle (every sum over 1200 pixels)"""
    import densecrf_ho_amd as D
    from densecrf_ho_amd import synthetic as S
    wl = S.Workload(40, 30, 6, 3, L=21)
    c = D.DenseCRF2D(40, 30, 21)
    c.setUnaryEnergy(wl.U)
    f = wl.kernels[1][0]
    c.addPairwiseEnergy(f, D.PottsCompatibility(5.0))
    rng = np.random.default_rng(8)
    Q = rng.random((wl.N, 21)).astype(np.float32)
    ind = np.argsort(Q, axis=0).astype(np.int32)
    w = np.float32(-5.0)
    assert rel_err(c.apply_bf(0, Q), w * O.bf_apply(f, Q)) < 1e-5
    assert rel_err(c.apply_upper_minus_lower_bf(0, ind), w * O.bf_uml(f, ind)) < 1e-5
    assert rel_err(c.apply_upper_minus_lower_bf_ord(0, ind, Q), w * O.bf_uml(f, ind, Q)) < 1e-5
    bad = ind.copy()
    bad[3, 2] = wl.N
    with pytest.raises(D.EngineError, match="out of range"):
        c.apply_upper_minus_lower_bf(0, bad)


@pytest.mark.parametrize("tag", ["12x9", "31x23"])
def test_divide_and_conquer_filter_vs_reference(golden, tag):
    """apply_upper_minus_lower_dc of both potentials and the lattice step
    compute_upper_minus_lower_dc (even and odd split) against the compiled
    reference (within 1e-5 of the largest value)"""
    import densecrf_ho_amd as D
    g = golden("dc_L4.npz")
    W, H = (int(x) for x in tag.split("x"))
    c = D.DenseCRF2D(W, H, 4)
    c.setUnaryEnergy(g[f"U_{tag}"])
    c.addPairwiseEnergy(g[f"feat0_{tag}"], D.PottsCompatibility(3.0))
    c.addPairwiseEnergy(g[f"feat1_{tag}"], D.PottsCompatibility(5.0))
    for k in range(2):
        print("potential", k, flush=True)
        out = c.apply_upper_minus_lower_dc(k, g[f"ind_{tag}"])
        assert rel_err(out, g[f"dc{k}_{tag}"]) < 1e-5
        lat = D.Permutohedral().init(g[f"feat{k}_{tag}"])
        N = lat.N
        row = g[f"row0_{k}_{tag}"]
        a = lat.compute_upper_minus_lower_dc(row.copy(), 0, N // 2, N // 2, N)
        assert rel_err(a, g[f"lat_even{k}_{tag}"]) < 1e-5
        b = lat.compute_upper_minus_lower_dc(row.copy(), 3, (N - 3) // 2 + 3, (N - 3) // 2 + 4, N - 2)
        assert rel_err(b, g[f"lat_odd{k}_{tag}"]) < 1e-5


def test_divide_and_conquer_filter_vs_oracle_larger():
    """80x60 image (4800 pixels: nodes of up to 2400 pixels in one
    workgroup), 21 labels, orders sorted by random values"""
    import densecrf_ho_amd as D
    from densecrf_ho_amd import synthetic as S
    wl = S.Workload(80, 60, 10, 5, L=21)
    c = D.DenseCRF2D(80, 60, 21)
    c.setUnaryEnergy(wl.U)
    f = wl.kernels[1][0]
    c.addPairwiseEnergy(f, D.PottsCompatibility(5.0))
    ind = np.argsort(np.random.default_rng(6).random((wl.N, 21)), axis=0).astype(np.int32)
    out = c.apply_upper_minus_lower_dc(0, ind)
    assert rel_err(out, np.float32(-5.0) * O.uml_dc(f, ind)) < 1e-5


def test_divide_and_conquer_size_guard():
    """N * L above 2^20 (one launch per 15 pixels and label) is refused
    before any launch"""
    import densecrf_ho_amd as D
    from densecrf_ho_amd import synthetic as S
    wl = S.Workload(256, 200, 10, 5, L=21)  # 51,200 pixels x 21 labels > 2^20
    c = D.DenseCRF2D(256, 200, 21)
    c.setUnaryEnergy(wl.U)
    c.addPairwiseEnergy(wl.kernels[0][0], D.PottsCompatibility(3.0))
    ind = np.tile(np.arange(wl.N, dtype=np.int32)[:, None], (1, 21))
    with pytest.raises(D.EngineError, match="2\\^20"):
        c.apply_upper_minus_lower_dc(0, ind)


def test_standalone_pairwise_potential_filters(golden):
    """the standalone PairwisePotential (its own lattice and handle): the
    brute-force, divide-and-conquer and restricted filters against the
    reference fixtures and the oracle"""
    import densecrf_ho_amd as D
    g = golden("bf_12x9_L4.npz")
    p = D.PairwisePotential(g["feat1"], D.PottsCompatibility(5.0))
    assert rel_err(p.apply_bf(g["Q"]), g["bf1"]) < 1e-5
    assert rel_err(p.apply_upper_minus_lower_bf(g["ind"]), g["uml1"]) < 1e-5
    assert rel_err(p.apply_upper_minus_lower_bf_ord(g["ind"], g["Q"]), g["umlord1"]) < 1e-5
    dc = p.apply_upper_minus_lower_dc(g["ind"])
    assert rel_err(dc, np.float32(-5.0) * O.uml_dc(g["feat1"], g["ind"])) < 1e-5
    rng = np.random.default_rng(21)
    N = g["Q"].shape[0]
    pI = np.sort(rng.choice(N, N // 2, replace=False)).astype(np.int32)
    rq = rng.random((len(pI), 4)).astype(np.float32)
    got = p.apply_upper_minus_lower_ord_restricted(rq, pI, g["Q"], True)
    ref = np.float32(-5.0) * O.Lattice(g["feat1"]).ordered_restricted(rq, pI, g["Q"], True)
    assert rel_err(got, ref) < 1e-5


def test_add_superpixel_from_region_map(golden):
    """DenseCRF2D.addSuperPixel with the reference's mean-shift regions of its
    own data/img.bmp: the QP with that layer ends at the reference's
    higher-order energy (qp_sp of example_inference)"""
    import densecrf_ho_amd as D
    with open(os.path.join(GOLDEN, "data_img.json")) as f:
        ref = json.load(f)
    g = np.load(os.path.join(GOLDEN, "data_img_labels.npz"))
    U, W, H = D.load_unary(os.path.join(GOLDEN, "data", "img.c_unary"))
    im, _, _ = D.load_image(os.path.join(GOLDEN, "data", "img.bmp"))
    c = D.DenseCRF2D(W, H, U.shape[1])
    c.setUnaryEnergy(U)
    c.addPairwiseGaussian(2, 2, D.PottsCompatibility(3))
    c.addPairwiseBilateral(30, 30, 8, 8, 8, im, D.PottsCompatibility(2))
    c.addSuperPixel(im, g["regions"], 50.0, 1000.0)
    Q = c.qp_inference_super_pixels_non_convex(c.unary_init())
    lab = Q.argmax(1).astype(np.int32)
    e = c.assignment_energy_higher_order(lab)
    print("qp_sp energy", e, "reference", ref["e_qp_sp"], flush=True)
    assert abs(e - ref["e_qp_sp"]) <= 1e-4 * abs(ref["e_qp_sp"])
    assert (lab == g["qp_sp"]).mean() >= 0.999


def test_per_pixel_energies_add_up(golden):
    """unaryEnergy + pairwise_energy_true (the reference's debugging
    functions) summed in FP64 give assignment_energy_true; term -1 is the sum
    of the single terms in order; pinned to the reference's total energy"""
    import densecrf_ho_amd as D
    g = golden("crf_31x23_L5.npz")
    W, H, L = int(g["W"]), int(g["H"]), int(g["L"])
    c = D.DenseCRF2D(W, H, L)
    c.setUnaryEnergy(g["U"])
    c.addPairwiseEnergy(g["feat0"], D.PottsCompatibility(3.0))
    c.addPairwiseEnergy(g["feat1"], D.PottsCompatibility(5.0))
    lab = g["lab_qp"]
    u = c.unaryEnergy(lab)
    p = c.pairwise_energy_true(lab)
    np.testing.assert_array_equal(p, c.pairwise_energy_true(lab, 0) + c.pairwise_energy_true(lab, 1))
    np.testing.assert_array_equal(u, g["U"][np.arange(W * H), lab])
    e = float(np.sum((u + p).astype(np.float64)))
    assert abs(e - c.assignment_energy_true(lab)) <= 1e-9 * abs(e)
    assert abs(e - float(g["e_true"])) <= 1e-5 * abs(float(g["e_true"]))
