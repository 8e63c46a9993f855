"""GPU LP solver and superpixel (Pn) terms against the reference fixtures and
the oracle.  Bars: integer positions (ucondgrad) exact; energies within 1e-4
relative; label agreement >= 99.9%."""

import json
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, rel_err

pytestmark = pytest.mark.gpu


def gpu_crf(g, layers=True):
    import densecrf_ho_amd as D
    W, H, L = int(g["W"]), int(g["H"]), int(g["L"])
    c = D.DenseCRF2D(W, H, L)
    c.setUnaryEnergy(g["U"])
    c.addPairwiseEnergy(g["feat0"], D.PottsCompatibility(3.0))
    c.addPairwiseEnergy(g["feat1"], D.PottsCompatibility(5.0))
    i = 0
    while layers and f"region{i}" in g:
        c.addSuperPixelLayer(g[f"region{i}"], 0.5, 0.2)
        i += 1
    return c


def orc_crf(g):
    c = O.CRF(int(g["W"]) * int(g["H"]), int(g["L"]))
    c.set_unary(g["U"])
    c.add_pairwise(g["feat0"], 3.0)
    c.add_pairwise(g["feat1"], 5.0)
    i = 0
    while f"region{i}" in g:
        c.add_layer(g[f"region{i}"], 0.5, 0.2)
        i += 1
    return c


@pytest.mark.parametrize("case", ["crf_31x23_L5.npz", "crf_32x24_L21_ho.npz"])
def test_lp_vs_reference(golden, case):
    import densecrf_ho_amd as D
    g = golden(case)
    print("case", case, flush=True)
    c = gpu_crf(g)
    Q, st = c.lp_inference_prox_super_pixels(g["Qqp"], D.LP_inf_params(prox_max_iter=int(g["lp_outer_cap"])),
                                             stats=True)
    o = orc_crf(g)
    Qo, sto = o.lp(g["Qqp"], O.LPParams(prox_max_iter=int(g["lp_outer_cap"])))
    print("gpu sweeps", st["qp_sweeps"][:20], "oracle", sto["qp_sweeps"][:20], flush=True)
    print("gpu int energies", st["int_energy"], "oracle", sto["int_energy"], flush=True)
    # the float oracle is the reference's LP bit for bit on these fixtures
    # (test_oracle.py::test_lp_matches_reference), so its counts are the reference's
    np.testing.assert_array_equal(Qo, g["Qlp"])
    np.testing.assert_array_equal(st["qp_sweeps"], sto["qp_sweeps"])
    assert st["outer"] == sto["outer"]
    np.testing.assert_allclose(st["int_energy"], sto["int_energy"], rtol=1e-4)
    lab, labr = Q.argmax(1), g["Qlp"].argmax(1)
    assert (lab == labr).mean() >= 0.999
    e, er = o.energy_ho(lab, fp64=True), o.energy_ho(labr, fp64=True)
    assert abs(e - er) <= 1e-4 * abs(er)
    assert rel_err(Q, g["Qlp"]) < 1e-3


def test_portable_build_qp_lp_vs_oracle(golden):
    """QP and LP on lattices of the reference's portable build (engine and
    oracle both switched): the same iteration and sweep counts, energies within
    1e-4, labels >= 99.9 %; 21 labels, so every filter takes seqCompute's
    double blur"""
    import densecrf_ho_amd as D
    g = golden("crf_32x24_L21_ho.npz")
    D.set_reference_path(True)
    O.set_reference_path(True)
    try:
        c = gpu_crf(g)
        o = orc_crf(g)
        Q0 = O.unary_init(g["U"])
        Qq, stq = c.qp_inference_super_pixels_non_convex(Q0, max_iter=10, ref_stop=False, stats=True)
        Qqo, stqo = o.qp(Q0, max_iter=10)
        print("qp energies gpu", stq["energy"], "oracle", stqo["energy"], flush=True)
        assert stq["iters"] == stqo["iters"]
        np.testing.assert_allclose(stq["energy"], stqo["energy"], rtol=1e-4)
        assert (Qq.argmax(1) == Qqo.argmax(1)).mean() >= 0.999
        P = D.LP_inf_params(prox_max_iter=int(g["lp_outer_cap"]))
        Q, st = c.lp_inference_prox_super_pixels(Qqo, P, stats=True)
        Qo, sto = o.lp(Qqo, O.LPParams(prox_max_iter=int(g["lp_outer_cap"])))
        print("lp sweeps gpu", st["qp_sweeps"][:20], "oracle", sto["qp_sweeps"][:20], flush=True)
        np.testing.assert_array_equal(st["qp_sweeps"], sto["qp_sweeps"])
        assert st["outer"] == sto["outer"]
        np.testing.assert_allclose(st["int_energy"], sto["int_energy"], rtol=1e-4)
        assert (Q.argmax(1) == Qo.argmax(1)).mean() >= 0.999
        assert rel_err(Q, Qo) < 1e-3
    finally:
        D.set_reference_path(False)
        O.set_reference_path(False)


def test_lp_less_confident_stop_vs_reference(golden):
    """the less-confident stop (src/densecrf.cpp:784-794) ends the GPU LP at
    the oracle's outer step for each setting, with the reference's result"""
    import densecrf_ho_amd as D
    g = golden("lp_lc_31x23_L5.npz")
    c = gpu_crf(g, layers=False)
    o = orc_crf(g)
    for i, (pct, tol) in enumerate(g["settings"]):
        print("setting", i, pct, tol, flush=True)
        kw = dict(prox_max_iter=int(g["prox_max_iter"]), less_confident_percent=float(pct), confidence_tol=float(tol))
        Q, st = c.lp_inference_prox_super_pixels(g["Qqp"], D.LP_inf_params(**kw), stats=True)
        _, sto = o.lp(g["Qqp"], O.LPParams(**kw))
        assert st["outer"] == sto["outer"]
        lab, labr = Q.argmax(1), g[f"Qlp{i}"].argmax(1)
        assert (lab == labr).mean() >= 0.999
        assert rel_err(Q, g[f"Qlp{i}"]) < 1e-3


def test_higher_order_pieces_vs_reference(golden):
    g = golden("crf_32x24_L21_ho.npz")
    c = gpu_crf(g)
    lab = g["lab_qp"]
    e = c.assignment_energy_higher_order(lab)
    assert abs(e - float(g["e_ho"])) <= 1e-6 * abs(float(g["e_ho"]))
    e = c.compute_energy_LP_higher_order(g["Qqp"])
    assert abs(e - float(g["e_lp"])) <= 1e-5 * abs(float(g["e_lp"]))


def test_qp_with_superpixels_vs_reference(golden):
    g = golden("crf_32x24_L21_ho.npz")
    c = gpu_crf(g)
    Q = c.qp_inference_super_pixels_non_convex(g["Q0"])
    assert rel_err(Q, g["Qqp"]) < 1e-4
    assert (Q.argmax(1) == g["Qqp"].argmax(1)).mean() >= 0.999


def test_ucondgrad_positions_exact(golden):
    """computeUCondGrad with the reference's int truncation: same positions."""
    g = golden("crf_32x24_L21_ho.npz")
    c = gpu_crf(g)
    np.testing.assert_array_equal(c.computeUCondGrad(g["Qpert"]), g["ucond"])
    # values spanning several integers exercise the truncating comparisons
    rng = np.random.default_rng(4)
    Q = (rng.standard_normal(g["Qpert"].shape) * 2.5).astype(np.float32)
    np.testing.assert_array_equal(c.computeUCondGrad(Q), orc_crf(g).ucondgrad(Q))


def test_config0_lp_vs_reference(golden):
    """320x213, 21 labels: QP (reference stopping) then the LP with 2 outer steps."""
    import densecrf_ho_amd as D
    from densecrf_ho_amd import synthetic as S
    with open(os.path.join(GOLDEN, "c0_320x213.json")) as f:
        ref = json.load(f)
    g = golden("c0_320x213.npz")
    wl = S.Workload(320, 213, 40, 0)
    c = D.DenseCRF2D(320, 213, 21)
    c.setUnaryEnergy(wl.U)
    for feat, w in wl.kernels:
        c.addPairwiseEnergy(feat, D.PottsCompatibility(w))
    Qqp = c.qp_inference_super_pixels_non_convex(c.unary_init())
    Qlp = c.lp_inference_prox_super_pixels(Qqp, D.LP_inf_params(prox_max_iter=2))
    lab = c.currentMap(Qlp)
    agree = (lab == g["lab_lp"]).mean()
    e = c.assignment_energy_true(lab)
    print("C0 LP: agreement", agree, "energy", e, "reference", ref["e_true_lp"], flush=True)
    assert agree >= 0.999
    assert abs(e - ref["e_true_lp"]) <= 1e-4 * abs(ref["e_true_lp"])


def test_qp_lp_with_superpixels_vs_fp64_oracle():
    """qp_sp then lp_sp at 128x96 with three Voronoi superpixel layers (12, 24
    and 48 regions, Pn weight 0.5, LP multiplier 0.2) against the FP64 oracle:
    labels >= 99.9 %, higher-order integer energies within 1e-4"""
    import densecrf_ho_amd as D
    from densecrf_ho_amd import synthetic as S
    W, H, L = 128, 96, 21
    wl = S.Workload(W, H, 12, 7, L=L)
    layers = S.superpixel_layers(W, H, cells=(12, 24, 48), seeds=(31, 32, 33))
    c = D.DenseCRF2D(W, H, L)
    c.setUnaryEnergy(wl.U)
    o = O.CRF(wl.N, L)
    o.set_unary(wl.U)
    for f, w in wl.kernels:
        c.addPairwiseEnergy(f, D.PottsCompatibility(w))
        o.add_pairwise(f, w)
    for region, _ in layers:
        c.addSuperPixelLayer(region, 0.5, 0.2)
        o.add_layer(region, 0.5, 0.2)
    q0 = O.unary_init(wl.U)
    qg = c.qp_inference_super_pixels_non_convex(q0, max_iter=10, ref_stop=False)
    qo, _ = o.qp(q0, max_iter=10, fp64=True)
    print("qp label agreement", (qg.argmax(1) == qo.argmax(1)).mean(), flush=True)
    assert (qg.argmax(1) == qo.argmax(1)).mean() >= 0.999
    P = D.LP_inf_params(prox_max_iter=3, fw_max_iter=5)
    lg = c.lp_inference_prox_super_pixels(qg, P).argmax(1).astype(np.int32)
    lo = o.lp(qo, O.LPParams(prox_max_iter=3, fw_max_iter=5), fp64=True)[0].argmax(1).astype(np.int32)
    eg, eo = o.energy_ho(lg, fp64=True), o.energy_ho(lo, fp64=True)
    print("lp label agreement", (lg == lo).mean(), "energies", eg, eo, flush=True)
    assert (lg == lo).mean() >= 0.999
    assert abs(eg - eo) <= 1e-4 * abs(eo)


def test_ucondgrad_large_regions_vs_oracle():
    """Regions of hundreds to thousands of members (many 64-member chunks per
    label, carries across chunks and labels): positions equal the oracle's
    for values spanning several integers and for values mostly in [0, 1]
    with a few negatives."""
    import densecrf_ho_amd as D
    from densecrf_ho_amd import synthetic as S
    W, H, L = 160, 120, 21
    wl = S.Workload(W, H, 8, 3, L=L)
    c = D.DenseCRF2D(W, H, L)
    c.setUnaryEnergy(wl.U)
    o = O.CRF(wl.N, L)
    o.set_unary(wl.U)
    for f, w in wl.kernels:
        c.addPairwiseEnergy(f, D.PottsCompatibility(w))
        o.add_pairwise(f, w)
    for region, _ in S.superpixel_layers(W, H, cells=(4, 9, 30), seeds=(41, 42, 43)):
        c.addSuperPixelLayer(region, 0.5, 0.2)
        o.add_layer(region, 0.5, 0.2)
    rng = np.random.default_rng(11)
    Qa = (rng.standard_normal((wl.N, L)) * 2.5).astype(np.float32)
    Qb = rng.random((wl.N, L)).astype(np.float32)
    Qb[rng.random((wl.N, L)) < 0.001] *= -1.0
    for Q in (Qa, Qb):
        np.testing.assert_array_equal(c.computeUCondGrad(Q), o.ucondgrad(Q))


def test_compute_energy_higher_order_vs_reference(golden):
    """compute_energy_higher_order after the QP (the z_labels_ it leaves), at
    the reference QP's result and at another field, within 1e-4 of the
    compiled reference; without a QP run the region term is refused"""
    import densecrf_ho_amd as D
    g = golden("crf_32x24_L21_ho.npz")
    ge = golden("eho_32x24_L21.npz")
    c = gpu_crf(g)
    with pytest.raises(D.EngineError, match="run the QP first"):
        c.compute_energy_higher_order(g["Qqp"])
    c.qp_inference_super_pixels_non_convex(g["Q0"])
    e = c.compute_energy_higher_order(g["Qqp"])
    print("energy", e, "reference", float(ge["e_qp"]), flush=True)
    assert abs(e - float(ge["e_qp"])) <= 1e-4 * abs(float(ge["e_qp"]))
    e2 = c.compute_energy_higher_order(ge["Qo"])
    assert abs(e2 - float(ge["e_other"])) <= 1e-4 * abs(float(ge["e_other"]))
