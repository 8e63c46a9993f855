"""Pins the CPU oracle (oracle/dcrf_oracle.cpp).

How each oracle function is pinned:

  lattice init (offset/rank/barycentric/neighbours, ghost lanes)
      golden fixtures from the compiled reference (bit-exact), plus the
      invariants below; planted mistakes 1 and 2 must fail the fixtures.
  filter (seqCompute V<=2 and sseCompute V>2, both blur orders)
      fixtures (bit-exact), linearity, adjointness of the reverse blur, the
      constant field.  (The V<=2 blur's double evaluation equals float
      arithmetic unless the operands' exponents are >28 apart, so no fixture
      can tell the two apart; the oracle keeps the double form.)
  ordered filter (10-bin upper minus lower)
      fixtures with q at exact bin edges (bit-exact), equivalence with plain
      filters of bin-indicator fields; planted mistake 3 fails.
  PairwisePotential apply / applyTranspose / no-norm copy, Potts
      fixtures (bit-exact).
  unary_init (softmax)
      numpy softmax (1e-6; the reference's vectorised exp differs by an ulp).
  QP (reference stopping rule, bug-compatible superpixel sums)
      fixtures Q and z_labels (bit-exact) at 31x23, 32x24 with 3 layers and
      the 320x213 config; a convex case (negative Potts weights, so the
      line-search step is not clamped) and an exact-tie case (duplicated
      unary columns) pin the step formula and the first-minimum rule;
      planted mistakes 5, 6, 7 fail.
  LP (prox, qp-gamma, ordered filter, projection, best-integer tracking)
      fixtures (bit-exact final Q) and 320x213 energies; feasible_Q against a
      textbook simplex projection; renormalize; ucondgrad (planted 8 fails);
      on the reference's own data/img.* (with and without its mean-shift
      layer) runs cut short, started from the reference's own unary_init,
      give its labels and energies exactly.
  energies (assignment true / higher order, LP energy, QP value)
      fixtures (exact doubles).
  LP less-confident stop (less_confident_percent, confidence_tol)
      fixtures of four settings that stop after 4, 1, 2 and 3 outer steps
      (bit-exact final Q, pairwise different).
  restricted ordered filter (store / reuse / another pixel set, accumulation)
      fixtures at d=2 and d=5 with values at bin edges (bit-exact); planted
      mistake 9 (blur every point) fails; with every pixel in pI it equals
      the ordered filter exactly.
  brute-force filters (apply_bf, upper_minus_lower_bf, _bf_ord)
      fixtures (bit-exact, including ties in the value split), the dense
      FP64 product K Q, the closed form of the position split, and
      bf_ord = -bf for the ascending order of distinct values.
  lattice feature gradient (Permutohedral::gradient)
      fixtures at d=2 and d=5 (bit-exact), bilinearity.
  divide-and-conquer filter (apply_upper_minus_lower_dc, compute_..._dc)
      fixtures at 12x9 and 31x23, both kernels, even and odd splits
      (bit-exact).

Not restated in the oracle (the GPU is checked against reference fixtures
directly): DenseKernel::featureGradient / gradient and kernel parameters
(kgrad_31x23.npz), the file formats and addSuperPixel terms (test_io.py).

The compiled reference (oracle/_ref) is used in addition where it was built.
"""

import json
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, rel_err

LATTICES = ["g2_32x24", "b5_32x24", "b5_31x23", "r3_101", "r1_50", "ties_d1"]


@pytest.fixture(autouse=True)
def _no_mutation():
    O.set_mutation(0)
    yield
    O.set_mutation(0)


def _lat(golden, name):
    g = golden(f"lattice_{name}.npz")
    return g, O.Lattice(g["feat"])


@pytest.mark.parametrize("name", LATTICES)
def test_lattice_tables_match_reference(golden, name):
    g, lat = _lat(golden, name)
    A = lat.arrays()
    assert lat.M == int(g["M"])
    np.testing.assert_array_equal(A["offset"], g["offset"])
    np.testing.assert_array_equal(A["rank"], g["rank"])
    np.testing.assert_array_equal(A["bary"], g["bary"])
    np.testing.assert_array_equal(A["nbr"], g["nbr"])


@pytest.mark.parametrize("name", LATTICES)
def test_filter_matches_reference(golden, name):
    g, lat = _lat(golden, name)
    np.testing.assert_array_equal(lat.filter(g["x"]), g["out"])
    np.testing.assert_array_equal(lat.filter(g["x"], reverse=True), g["out_rev"])
    np.testing.assert_array_equal(lat.filter(g["x1"]), g["out1"])
    np.testing.assert_array_equal(lat.ordered(g["q"]), g["ord"])


PORTABLE = ["ties_d1", "b5_31x23", "g2_31x23", "r3_101"]


@pytest.mark.parametrize("name", PORTABLE)
def test_portable_build_matches_reference(golden, name):
    """the oracle's restatement of the reference's portable build (no __SSE__:
    ties rounded down, no padding lanes, seqCompute at every value size)
    against lattice_portable.npz, written by that build of the reference"""
    g = golden("lattice_portable.npz")
    O.set_reference_path(True)
    try:
        lat = O.Lattice(g[f"{name}_feat"])
        A = lat.arrays()
        assert lat.M == int(g[f"{name}_M"]) and lat.Npad == lat.N
        for k in ("offset", "rank", "bary", "nbr"):
            np.testing.assert_array_equal(A[k], g[f"{name}_{k}"], err_msg=k)
        np.testing.assert_array_equal(lat.filter(g[f"{name}_x"]), g[f"{name}_out"])
        np.testing.assert_array_equal(lat.filter(g[f"{name}_x"], reverse=True), g[f"{name}_out_rev"])
        np.testing.assert_array_equal(lat.filter(g[f"{name}_x1"]), g[f"{name}_out1"])
        np.testing.assert_array_equal(lat.filter(g[f"{name}_x1"], reverse=True), g[f"{name}_out1_rev"])
        np.testing.assert_array_equal(lat.ordered(g[f"{name}_q"]), g[f"{name}_ord"])
    finally:
        O.set_reference_path(False)


def test_portable_build_differs_where_it_should(golden):
    """the two builds differ exactly where the fixtures say: exact .5 ties of
    ties_d1 round the other way (different tables), and at V=21 the SSE
    build's float blur differs from seqCompute's double blur; the SSE-mode
    oracle still reproduces the SSE fixture afterwards"""
    g = golden("lattice_portable.npz")
    sse = O.Lattice(g["ties_d1_feat"])
    assert not np.array_equal(sse.arrays()["offset"][: sse.N], g["ties_d1_offset"]) or \
        not np.array_equal(sse.arrays()["bary"][: sse.N], g["ties_d1_bary"])
    lat = O.Lattice(g["b5_31x23_feat"])
    assert not np.array_equal(lat.filter(g["b5_31x23_x"]), g["b5_31x23_out"])
    ref = golden("lattice_ties_d1.npz")
    np.testing.assert_array_equal(O.Lattice(ref["feat"]).arrays()["offset"], ref["offset"])


@pytest.mark.parametrize("name", LATTICES)
def test_lattice_invariants(golden, name):
    g, lat = _lat(golden, name)
    A = lat.arrays()
    d = lat.d
    N = lat.N
    b = A["bary"][:N]
    # barycentric weights of a pixel sum to 1 and lie in [0,1]
    np.testing.assert_allclose(b.sum(1), 1.0, atol=2e-6)
    assert b.min() >= -1e-6 and b.max() <= 1 + 1e-6
    # the remainder-r vertex has every coordinate congruent to r mod d+1
    keys = A["keys"].astype(np.int64)
    off = A["offset"][:N]
    for r in range(d + 1):
        k = keys[off[:, r]]
        assert np.all((k - r) % (d + 1) == 0)
    # each vertex's coordinates sum (with the implied last coordinate) to 0:
    # neighbours are symmetric: n2 of n1 along axis j is the point itself
    nbr = A["nbr"]
    for j in range(d + 1):
        n1 = nbr[j, :, 0]
        ok = n1 >= 0
        back = nbr[j, n1[ok], 1]
        np.testing.assert_array_equal(back, np.nonzero(ok)[0])
    # every pixel's d+1 vertices are distinct points
    assert np.all(np.sort(off, 1)[:, 1:] != np.sort(off, 1)[:, :-1])


def test_lattice_translation_invariance():
    rng = np.random.default_rng(0)
    f = (rng.standard_normal((64, 2)) * 3).astype(np.float32)
    a = O.Lattice(f)
    # shifting the elevated point by a lattice vector keeps the key set
    # structure: same number of points and the same per-pixel partition
    f2 = f.copy()
    scale = np.sqrt(2 / 3) * 3 / np.sqrt(2)
    f2[:, 0] += np.float32(3 / scale * 2)  # a multiple of (d+1) in elevated units
    b = O.Lattice(f2)
    assert a.M == b.M


def test_filter_linearity_adjoint_and_constant(golden):
    g, lat = _lat(golden, "b5_32x24")
    rng = np.random.default_rng(1)
    a = rng.standard_normal((lat.N, 3))
    b = rng.standard_normal((lat.N, 3))
    Ka = lat.filter(a, fp64=True)
    Kb = lat.filter(b, fp64=True)
    np.testing.assert_allclose(lat.filter(2 * a - 3 * b, fp64=True), 2 * Ka - 3 * Kb, rtol=1e-12, atol=1e-10)
    # <a, K b> = <K^T a, b> with K^T the reverse blur order
    lhs = (a * Kb).sum()
    rhs = (lat.filter(a, reverse=True, fp64=True) * b).sum()
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    # a constant field filters to the norm input K.1 in every column
    ones = np.ones((lat.N, 4))
    k1 = lat.filter(np.ones((lat.N, 1)), fp64=True)
    np.testing.assert_allclose(lat.filter(ones, fp64=True), np.repeat(k1, 4, 1), rtol=1e-12)


def test_filter_close_to_brute_force_gaussian(golden):
    """The lattice approximates sum_j exp(-|f_i-f_j|^2/2) x_j up to a few %."""
    g, lat = _lat(golden, "g2_32x24")
    f = g["feat"].astype(np.float64)
    x = np.ones((lat.N, 1))
    d2 = ((f[:, None, :] - f[None, :, :]) ** 2).sum(-1)
    bf = np.exp(-0.5 * d2).sum(1)
    k = lat.filter(x, fp64=True)[:, 0]
    corr = np.corrcoef(k, bf)[0, 1]
    assert corr > 0.99
    ratio = np.median(k / bf)
    assert 0.3 < ratio < 3.0


def test_ordered_equals_binned_plain_filters(golden):
    """out_i = alpha-weighted sum over bins: lower minus upper indicator filters."""
    g, lat = _lat(golden, "b5_31x23")
    q = g["q"][:, :1].astype(np.float32)
    lo = np.minimum(np.floor(q * np.float32(10)).astype(int), 9)
    up = np.maximum(np.floor((q.astype(np.float64) - 1e-9) * 10).astype(int), 0)
    out = np.zeros((lat.N, 1))
    for c in range(10):
        # lower: pixels j with bin_lo(j) >= c contribute to i when bin_lo(i) == c
        kl = lat.filter((lo >= c).astype(np.float64), fp64=True)
        ku = lat.filter((up <= c).astype(np.float64), fp64=True)
        out += np.where(lo == c, -kl, 0) + np.where(up == c, ku, 0)
    np.testing.assert_allclose(lat.ordered(q, fp64=True), out, rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("mut,name", [(1, "b5_31x23"), (2, "ties_d1"), (3, "b5_31x23")])
def test_planted_lattice_mistakes_are_caught(golden, mut, name):
    g = golden(f"lattice_{name}.npz")
    O.set_mutation(mut)
    lat = O.Lattice(g["feat"])
    A = lat.arrays()
    same = (lat.M == int(g["M"]) and A["offset"].shape == g["offset"].shape
            and np.array_equal(A["offset"], g["offset"]) and np.array_equal(A["nbr"], g["nbr"])
            and np.array_equal(lat.filter(g["x1"]), g["out1"]) and np.array_equal(lat.ordered(g["q"]), g["ord"]))
    assert not same, f"mutation {mut} went unnoticed"


# ---------------------------------------------------------------------------
# CRF level
# ---------------------------------------------------------------------------
def _crf(g):
    W, H, L = int(g["W"]), int(g["H"]), int(g["L"])
    c = O.CRF(W * H, L)
    c.set_unary(g["U"])
    c.add_pairwise(g["feat0"], 3.0)
    c.add_pairwise(g["feat1"], 5.0)
    i = 0
    while f"region{i}" in g:
        c.add_layer(g[f"region{i}"], 0.5, 0.2)
        i += 1
    return c


def test_features_match_reference(golden):
    from densecrf_ho_amd import synthetic as S
    g = golden("crf_31x23_L5.npz")
    W, H = int(g["W"]), int(g["H"])
    np.testing.assert_array_equal(S.gaussian_features(W, H, 3, 3), g["feat0"])
    np.testing.assert_array_equal(S.bilateral_features(W, H, 50, 50, 15, 15, 15, g["img"]), g["feat1"])
    wl = S.Workload(W, H, int(g["K"]), int(g["seed"]), L=int(g["L"]))
    np.testing.assert_array_equal(wl.U, g["U"])


@pytest.mark.parametrize("case", ["crf_31x23_L5.npz", "crf_32x24_L21_ho.npz"])
def test_potentials_match_reference(golden, case):
    g = golden(case)
    c = _crf(g)
    for k in range(2):
        np.testing.assert_array_equal(c.apply(k, g["x"]), g[f"apply{k}"])
        np.testing.assert_array_equal(c.apply(k, g["x"], transpose=True), g[f"applyT{k}"])
        np.testing.assert_array_equal(c.apply(k, g["x"], no_norm=True), g[f"applyNN{k}"])
        np.testing.assert_array_equal(c.apply_ord(k, g[f"ordq{k}"]), g[f"ord{k}"])


def test_ordered_rejects_out_of_range(golden):
    g = golden("crf_31x23_L5.npz")
    c = _crf(g)
    q = g["ordq0"].copy()
    q[3, 1] = 1.5
    with pytest.raises(RuntimeError, match="Value out of bounds"):
        c.apply_ord(0, q)


def test_unary_init_is_softmax(golden):
    g = golden("crf_32x24_L21_ho.npz")
    U = g["U"].astype(np.float64)
    e = np.exp(-U - (-U).max(1, keepdims=True))
    np.testing.assert_allclose(O.unary_init(g["U"]), e / e.sum(1, keepdims=True), rtol=2e-6)
    np.testing.assert_allclose(O.unary_init(g["U"]), g["Q0"], rtol=1e-6)


@pytest.mark.parametrize("case", ["crf_31x23_L5.npz", "crf_32x24_L21_ho.npz"])
def test_qp_matches_reference(golden, case):
    g = golden(case)
    c = _crf(g)
    Q, st = c.qp(g["Q0"], max_iter=0, ref_stop=True)
    np.testing.assert_array_equal(Q, g["Qqp"])
    if "zlabels" in g:
        np.testing.assert_array_equal(c.zlabels(g["zlabels"].shape[0]), g["zlabels"])
    assert st["iters"] >= 1


def _qp_crf(g):
    c = O.CRF(int(g["W"]) * int(g["H"]), int(g["L"]))
    c.set_unary(g["U"])
    c.add_pairwise(g["feat0"], float(g["w"][0]))
    c.add_pairwise(g["feat1"], float(g["w"][1]))
    return c


@pytest.mark.parametrize("case", ["qp_convex_31x23_L5.npz", "qp_ties_31x23_L5.npz"])
def test_qp_step_and_ties_match_reference(golden, case):
    g = golden(case)
    Q, st = _qp_crf(g).qp(g["Q0"], max_iter=0, ref_stop=True)
    np.testing.assert_array_equal(Q, g["Qqp"])


@pytest.mark.parametrize("mut,case", [(5, "qp_convex_31x23_L5.npz"), (6, "qp_ties_31x23_L5.npz"),
                                      (7, "crf_32x24_L21_ho.npz")])
def test_planted_qp_mistakes_are_caught(golden, mut, case):
    g = golden(case)
    O.set_mutation(mut)
    c = _crf(g) if "zlabels" in g else _qp_crf(g)
    Q, _ = c.qp(g["Q0"], max_iter=0, ref_stop=True)
    same = np.array_equal(Q, g["Qqp"])
    if "zlabels" in g:
        same = same and np.array_equal(c.zlabels(g["zlabels"].shape[0]), g["zlabels"])
    assert not same


@pytest.mark.parametrize("case", ["crf_31x23_L5.npz", "crf_32x24_L21_ho.npz"])
def test_energies_match_reference(golden, case):
    g = golden(case)
    c = _crf(g)
    lab = g["lab_qp"]
    np.testing.assert_array_equal(c.current_map(g["Qqp"]), lab)
    assert c.energy_true(lab) == float(g["e_true"])
    assert c.energy_ho(lab) == float(g["e_ho"])
    assert c.lp_energy(g["Qqp"]) == float(g["e_lp"])


@pytest.mark.parametrize("case", ["crf_31x23_L5.npz", "crf_32x24_L21_ho.npz"])
def test_lp_pieces_match_reference(golden, case):
    g = golden(case)
    c = _crf(g)
    np.testing.assert_array_equal(O.feasible(g["Qpert"]), g["feasible"])
    np.testing.assert_array_equal(O.renormalize(np.abs(g["Qpert"])), g["renorm"])
    np.testing.assert_array_equal(c.ucondgrad(g["Qpert"]), g["ucond"])


def test_planted_ucondgrad_mistake_is_caught(golden):
    g = golden("crf_32x24_L21_ho.npz")
    O.set_mutation(8)
    c = _crf(g)
    assert not np.array_equal(c.ucondgrad(g["Qpert"]), g["ucond"])


def test_feasible_is_simplex_projection(golden):
    """feasible_Q is the Euclidean projection onto the probability simplex."""
    g = golden("crf_31x23_L5.npz")
    Q = g["Qpert"].astype(np.float64)
    u = -np.sort(-Q, axis=1)
    css = np.cumsum(u, 1) - 1
    k = np.arange(1, Q.shape[1] + 1)
    rho = ((u - css / k) > 0).sum(1)
    theta = css[np.arange(len(Q)), rho - 1] / rho
    proj = np.clip(Q - theta[:, None], 0, 1)
    np.testing.assert_allclose(O.feasible(g["Qpert"]), proj, atol=2e-6)


@pytest.mark.parametrize("case", ["crf_31x23_L5.npz", "crf_32x24_L21_ho.npz"])
def test_lp_matches_reference(golden, case):
    g = golden(case)
    c = _crf(g)
    Q, st = c.lp(g["Qqp"], O.LPParams(prox_max_iter=int(g["lp_outer_cap"])))
    np.testing.assert_array_equal(Q, g["Qlp"])
    assert len(st["qp_sweeps"]) >= 1 and st["qp_sweeps"].min() >= 2  # first sweep never stops


def _restricted_calls(lat, g, k):
    out1 = lat.ordered_restricted(g[f"rq1_{k}"], g[f"pI{k}"], g[f"q{k}"], True)
    out2 = lat.ordered_restricted(g[f"rq2_{k}"], g[f"pI{k}"], None, False)
    out3 = lat.ordered_restricted(g[f"rq3_{k}"], g[f"pI2_{k}"], None, False, out=g[f"o3_{k}"])
    return out1, out2, out3


@pytest.mark.parametrize("k", [0, 1])
def test_restricted_ordered_matches_reference(golden, k):
    """compute_upper_minus_lower_ord_restricted: store, reuse of the stored
    splat and active set, a second pixel set, accumulation into out; bit for
    bit against the compiled reference"""
    g = golden("ordered_restricted_31x23.npz")
    lat = O.Lattice(g[f"feat{k}"])
    for i, o in enumerate(_restricted_calls(lat, g, k)):
        np.testing.assert_array_equal(o, g[f"out{i + 1}_{k}"])


def test_restricted_ordered_planted_mistake_is_caught(golden):
    g = golden("ordered_restricted_31x23.npz")
    try:
        O.set_mutation(9)  # blur every lattice point instead of the active ones
        lat = O.Lattice(g["feat1"])
        outs = _restricted_calls(lat, g, 1)
    finally:
        O.set_mutation(0)
    assert any(not np.array_equal(o, g[f"out{i + 1}_1"]) for i, o in enumerate(outs))


def test_restricted_ordered_with_every_pixel_is_the_ordered_filter(golden):
    """with every pixel in pI and N % 4 == 0 every lattice point is active,
    so the restricted filter is the ordered filter (same float order)"""
    g = golden("lattice_b5_32x24.npz")
    lat = O.Lattice(g["feat"])
    q = np.random.default_rng(2).random((lat.N, 5)).astype(np.float32)
    pI = np.arange(lat.N, dtype=np.int32)
    np.testing.assert_array_equal(lat.ordered_restricted(q, pI, q, True), lat.ordered(q))


@pytest.mark.parametrize("k", [0, 1])
def test_brute_force_filters_match_reference(golden, k):
    """apply_bf, apply_upper_minus_lower_bf and _bf_ord (src/pairwise.cpp:294-383)
    before the Potts weight: bit for bit against the compiled reference"""
    g = golden("bf_12x9_L4.npz")
    w = np.float32(3.0 if k == 0 else 5.0)
    f = g[f"feat{k}"]
    np.testing.assert_array_equal(-w * O.bf_apply(f, g["Q"]), g[f"bf{k}"])
    np.testing.assert_array_equal(-w * O.bf_uml(f, g["ind"]), g[f"uml{k}"])
    np.testing.assert_array_equal(-w * O.bf_uml(f, g["ind"], g["Q"]), g[f"umlord{k}"])
    np.testing.assert_array_equal(-w * O.bf_uml(f, g["ind_q"], g["Q"]), g[f"umlord_q{k}"])


def test_brute_force_filters_closed_forms(golden):
    """apply_bf is K Q with K = exp(-|f_c - f_b|^2) (FP64 dense product); in
    the ascending order of distinct values the value split equals the
    position split with the opposite sign (upper_bf_ord = -upper_bf); a
    permutation of one label's order permutes its sorted output"""
    g = golden("bf_12x9_L4.npz")
    f = g["feat1"].astype(np.float64)
    K = np.exp(-((f[:, None, :] - f[None, :, :]) ** 2).sum(-1))
    np.testing.assert_allclose(O.bf_apply(g["feat1"], g["Q"]), K @ g["Q"].astype(np.float64), rtol=1e-5)
    rng = np.random.default_rng(3)
    Q = rng.random(g["Q"].shape).astype(np.float32)  # distinct values
    ind = np.argsort(Q, axis=0).astype(np.int32)
    a, b = O.bf_uml(g["feat1"], ind, Q), O.bf_uml(g["feat1"], ind)
    assert np.abs(a + b).max() <= 1e-5 * np.abs(b).max()
    Ks = np.stack([K[np.ix_(ind[:, l], ind[:, l])] for l in range(ind.shape[1])])
    up = np.stack([np.triu(Ks[l], 1).sum(1) - np.tril(Ks[l], -1).sum(1) for l in range(ind.shape[1])], 1)
    np.testing.assert_allclose(b, up, rtol=1e-5, atol=1e-5 * np.abs(up).max())


@pytest.mark.parametrize("k", [0, 1])
def test_lattice_gradient_matches_reference(golden, k):
    """Permutohedral::gradient (src/permutohedral.cpp:1095-1178) bit for bit"""
    g = golden("kgrad_31x23.npz")
    lat = O.Lattice(g[f"feat{k}"])
    np.testing.assert_array_equal(lat.gradient(g[f"a{k}"], g[f"b{k}"]), g[f"df{k}"])


def test_lattice_gradient_is_bilinear_and_swaps_with_transpose(golden):
    """d/df a^T K b is linear in a and in b, and grad(a, b) of K equals
    grad(b, a) of the transposed filter (the two passes swap roles)"""
    g = golden("kgrad_31x23.npz")
    lat = O.Lattice(g["feat1"])
    a, b = g["a1"], g["b1"]
    a2 = np.random.default_rng(4).standard_normal(a.shape).astype(np.float32)
    ga = lat.gradient(a, b) + lat.gradient(a2, b)
    np.testing.assert_allclose(lat.gradient(a + a2, b), ga, rtol=1e-4, atol=1e-4 * np.abs(ga).max())
    gb = lat.gradient(a, 2 * b)
    np.testing.assert_allclose(gb, 2 * lat.gradient(a, b), rtol=1e-5, atol=1e-5 * np.abs(gb).max())


@pytest.mark.parametrize("tag", ["12x9", "31x23"])
@pytest.mark.parametrize("k", [0, 1])
def test_divide_and_conquer_filter_matches_reference(golden, tag, k):
    """apply_upper_minus_lower_dc (per-label sorted lattices, brute-force
    leaves, lattice steps) and compute_upper_minus_lower_dc for an even and
    an odd split, bit for bit"""
    g = golden("dc_L4.npz")
    w = np.float32(3.0 if k == 0 else 5.0)
    f = g[f"feat{k}_{tag}"]
    np.testing.assert_array_equal(-w * O.uml_dc(f, g[f"ind_{tag}"]), g[f"dc{k}_{tag}"])
    lat = O.Lattice(f)
    N = lat.N
    row = g[f"row0_{k}_{tag}"]
    np.testing.assert_array_equal(lat.dc(row, 0, N // 2, N // 2, N), g[f"lat_even{k}_{tag}"])
    np.testing.assert_array_equal(lat.dc(row, 3, (N - 3) // 2 + 3, (N - 3) // 2 + 4, N - 2), g[f"lat_odd{k}_{tag}"])


def test_lp_less_confident_stop_matches_reference(golden):
    """less_confident_percent / confidence_tol (src/densecrf.cpp:457-462,
    784-794): four settings stop after 4 (no check), 1, 2 and 3 outer steps;
    each final Q equals the compiled reference's bit for bit, and the four
    results differ, so each fixture pins where the stop came."""
    g = golden("lp_lc_31x23_L5.npz")
    c = _crf(g)
    outs = []
    for i, (pct, tol) in enumerate(g["settings"]):
        P = O.LPParams(prox_max_iter=int(g["prox_max_iter"]), less_confident_percent=float(pct),
                       confidence_tol=float(tol))
        Q, st = c.lp(g["Qqp"], P)
        np.testing.assert_array_equal(Q, g[f"Qlp{i}"])
        outs.append(st["outer"])
    assert outs == [4, 1, 2, 3]
    for i in range(1, 4):
        assert not np.array_equal(g[f"Qlp{i}"], g[f"Qlp{i - 1}"])


def test_config0_end_to_end(golden):
    """320x213, 40 cells, 21 labels: QP (reference stopping) then a 2-step LP."""
    from densecrf_ho_amd import synthetic as S
    with open(os.path.join(GOLDEN, "c0_320x213.json")) as f:
        ref = json.load(f)
    g = golden("c0_320x213.npz")
    wl = S.Workload(320, 213, 40, 0)
    c = O.CRF(wl.N, 21)
    c.set_unary(wl.U)
    for f, w in wl.kernels:
        c.add_pairwise(f, w)
    Q0 = O.unary_init(wl.U)
    Qqp, st = c.qp(Q0, max_iter=0, ref_stop=True)
    lab = c.current_map(Qqp)
    np.testing.assert_array_equal(lab, g["lab_qp"])
    assert abs(c.energy_true(lab) - ref["e_true_qp"]) <= 1e-9 * abs(ref["e_true_qp"])
    Qlp, _ = c.lp(Qqp, O.LPParams(prox_max_iter=2))
    lab2 = c.current_map(Qlp)
    assert (lab2 == g["lab_lp"]).mean() >= 0.999
    assert abs(c.energy_true(lab2) - ref["e_true_lp"]) <= 1e-4 * abs(ref["e_true_lp"])


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not built here")
@pytest.mark.parametrize("d,N", [(4, 203), (6, 97), (2, 1000)])
def test_random_lattices_against_compiled_reference(d, N):
    rng = np.random.default_rng(d * 1000 + N)
    f = (rng.standard_normal((N, d)) * 4).astype(np.float32)
    a, b = O.Lattice(f), O.RefLattice(f)
    A, B = a.arrays(), b.arrays()
    assert a.M == b.M
    for k in ("offset", "rank", "bary", "nbr"):
        np.testing.assert_array_equal(A[k], B[k])
    x = rng.standard_normal((N, 7)).astype(np.float32)
    np.testing.assert_array_equal(a.filter(x), b.filter(x))
    q = rng.random((N, 3)).astype(np.float32)
    np.testing.assert_array_equal(a.ordered(q), b.ordered(q))


@pytest.mark.parametrize("method,prox,fw", [("lp", 1, 10), ("lp_sp", 1, 1), ("lp_sp", 1, 2)])
def test_lp_first_steps_on_reference_data(method, prox, fw):
    """the float oracle's LP on data/img.* from the reference's own initial
    field (rebuilt bit for bit), cut short: the compiled reference's labels and
    discretised energy exactly"""
    import densecrf_ho_amd as D
    from densecrf_ho_amd import synthetic as S
    g = np.load(os.path.join(GOLDEN, "data_img_lp_steps.npz"), allow_pickle=False)
    labs = np.load(os.path.join(GOLDEN, "data_img_labels.npz"), allow_pickle=False)
    U, W, H = D.load_unary(os.path.join(GOLDEN, "data", "img.c_unary"))
    im, _, _ = D.load_image(os.path.join(GOLDEN, "data", "img.bmp"))
    Q0 = O.unary_init(U).reshape(-1)
    Q0[g["q0_idx"]] = g["q0_val"]
    c = O.CRF(W * H, U.shape[1])
    c.set_unary(U)
    c.add_pairwise(S.gaussian_features(W, H, 2, 2), 3.0)
    c.add_pairwise(S.bilateral_features(W, H, 30, 30, 8, 8, 8, im), 2.0)
    if method == "lp_sp":
        c.add_layer(labs["regions"], labs["exp_of_superpixels"], labs["constants"])
    Q, _ = c.lp(Q0.reshape(U.shape), O.LPParams(prox_max_iter=prox, fw_max_iter=fw))
    lab = c.current_map(Q)
    np.testing.assert_array_equal(lab, g[f"lab_{method}_{prox}x{fw}"])
    assert c.energy_true(lab) == float(g[f"e_{method}_{prox}x{fw}"])
