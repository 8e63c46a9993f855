"""Batched solving (many images of uneven sizes in one union lattice per
pairwise term, per-image solver control on the device, HIP-graph replay)
against the one-image GPU path and the FP64 oracle.

Tolerances: QP fields within 1e-5 of the one-image path and energies within
1e-4 (relative) of the FP64 oracle; LP labellings agree on >= 99.9% of the
pixels with the one-image path and their energies within 1e-4; graph replay
is bit-identical to eager execution.
"""

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

# uneven sizes, two with N % 4 != 0 (ghost lanes per image)
SIZES = [(31, 23), (48, 40), (17, 13), (64, 36)]


def _workloads(L):
    from densecrf_ho_amd import synthetic as S
    return [S.Workload(W, H, max(2, W * H // 300), 50 + i, L=L) for i, (W, H) in enumerate(SIZES)]


def _batch(wls, L, graphs=True):
    import densecrf_ho_amd as D
    b = D.DenseCRFBatch(L, SIZES)
    b.set_graphs(graphs)
    for i, wl in enumerate(wls):
        b.setUnaryEnergy(i, wl.U)
    b.addPairwiseGaussian(3, 3, D.PottsCompatibility(3))
    b.addPairwiseBilateral(50, 50, 15, 15, 15, [wl.img for wl in wls], D.PottsCompatibility(5))
    return b


def _single(wl, L):
    import densecrf_ho_amd as D
    c = D.DenseCRF2D(wl.W, wl.H, L)
    c.setUnaryEnergy(wl.U)
    c.addPairwiseGaussian(3, 3, D.PottsCompatibility(3))
    c.addPairwiseBilateral(50, 50, 15, 15, 15, wl.img, D.PottsCompatibility(5))
    return c


@pytest.mark.parametrize("L", [5, 21])
def test_batch_qp_matches_one_image_path_and_oracle(L):
    wls = _workloads(L)
    b = _batch(wls, L)
    Q, it = b.qp(max_iter=10, ref_stop=False)
    for i, wl in enumerate(wls):
        print("image", i, wl.W, wl.H, flush=True)
        c = _single(wl, L)
        q1 = c.qp_inference_super_pixels_non_convex(c.unary_init(), max_iter=10, ref_stop=False)
        np.testing.assert_array_equal(b.image(Q, i), q1)
        o = O.CRF(wl.N, L)
        o.set_unary(wl.U)
        for f, w in wl.kernels:
            o.add_pairwise(f, w)
        qo, _ = o.qp(O.unary_init(wl.U), max_iter=10, fp64=True)
        assert np.abs(b.image(Q, i) - qo).max() < 1e-3
        e = o.energy_true(b.image(Q, i).argmax(1).astype(np.int32), fp64=True)
        eo = o.energy_true(qo.argmax(1).astype(np.int32), fp64=True)
        assert abs(e - eo) <= 1e-4 * abs(eo)


def test_batch_qp_reference_stopping_matches_one_image_path():
    L = 5
    wls = _workloads(L)
    b = _batch(wls, L)
    Q, it = b.qp(max_iter=0, ref_stop=True)
    for i, wl in enumerate(wls):
        c = _single(wl, L)
        q1, st = c.qp_inference_super_pixels_non_convex(c.unary_init(), stats=True)
        assert it[i] == st["iters"]
        assert np.abs(b.image(Q, i) - q1).max() < 1e-5


def test_batch_graph_replay_equals_eager():
    """every replayed iteration (QP and LP pieces) equals eager execution;
    two solves in a row reuse the captured graphs"""
    L = 21
    wls = _workloads(L)
    be = _batch(wls, L, graphs=False)
    bg = _batch(wls, L, graphs=True)
    from densecrf_ho_amd import LP_inf_params
    P = LP_inf_params(prox_max_iter=2, fw_max_iter=4)
    for rep in range(2):
        print("solve", rep, flush=True)
        Qe, ite = be.qp(max_iter=6, ref_stop=False)
        Qg, itg = bg.qp(max_iter=6, ref_stop=False)
        np.testing.assert_array_equal(Qe, Qg)
        np.testing.assert_array_equal(ite, itg)
        Le, swe = be.lp(Qe, P)
        Lg, swg = bg.lp(Qg, P)
        np.testing.assert_array_equal(swe, swg)
        np.testing.assert_array_equal(Le, Lg)


def test_batch_graphs_recaptured_after_buffers_grow():
    """the LP's ordered filter widens the lattices' value rows (the buffers
    move); a QP graph captured before must not be replayed afterwards"""
    import densecrf_ho_amd as D
    L = 5
    wls = _workloads(L)
    b = _batch(wls, L, graphs=True)
    Q1, _ = b.qp(max_iter=4, ref_stop=False)
    c1 = b.graph_captures()
    assert c1 >= 1
    b.lp(Q1, D.LP_inf_params(prox_max_iter=1, fw_max_iter=2))
    c2 = b.graph_captures()
    Q2, _ = b.qp(max_iter=4, ref_stop=False)
    assert b.graph_captures() > c2  # captured again with the new buffers
    np.testing.assert_array_equal(Q1, Q2)


@pytest.mark.parametrize("prox,fw", [(3, 5), (5, 10)])
def test_batch_lp_matches_one_image_path(prox, fw):
    import densecrf_ho_amd as D
    L = 21
    wls = _workloads(L)
    b = _batch(wls, L)
    P = D.LP_inf_params(prox_max_iter=prox, fw_max_iter=fw)
    Q, _ = b.qp(max_iter=10, ref_stop=False)
    Ql, sw = b.lp(Q, P)
    lab_b = Ql.argmax(1).astype(np.int32)
    eb = b.assignment_energy_true(lab_b)
    for i, wl in enumerate(wls):
        print("image", i, flush=True)
        c = _single(wl, L)
        q0 = np.ascontiguousarray(b.image(Q, i))
        q1, st = c.lp_inference_prox_super_pixels(q0, P, stats=True)
        print("sweeps batch", sw[i][: len(st["qp_sweeps"])].tolist(), flush=True)
        print("sweeps one  ", list(st["qp_sweeps"]), "outer", st["outer"], flush=True)
        np.testing.assert_array_equal(sw[i][: len(st["qp_sweeps"])], st["qp_sweeps"])
        # same reduction partition per image: the batched LP is the one-image LP bit for bit
        np.testing.assert_array_equal(b.image(Ql, i), q1)
        l1 = q1.argmax(1).astype(np.int32)
        lb = b.image(lab_b, i)
        assert (lb == l1).mean() >= 0.999
        e1 = c.assignment_energy_true(l1)
        assert abs(eb[i] - c.assignment_energy_true(np.ascontiguousarray(lb))) <= 1e-9 * abs(eb[i]) + 1e-9
        assert abs(eb[i] - e1) <= 1e-4 * abs(e1)


def test_batch_lp_less_confident_stop_matches_one_image_path():
    """the less-confident stop is decided per image on the device; each image
    ends where it ends alone, bit for bit"""
    import densecrf_ho_amd as D
    L = 21
    wls = _workloads(L)
    b = _batch(wls, L)
    P = D.LP_inf_params(prox_max_iter=4, fw_max_iter=4, less_confident_percent=30.0, confidence_tol=0.6)
    Q, _ = b.qp(max_iter=10, ref_stop=False)
    Ql, _ = b.lp(Q, P)
    outer = []
    for i, wl in enumerate(wls):
        print("image", i, flush=True)
        c = _single(wl, L)
        q1, st = c.lp_inference_prox_super_pixels(np.ascontiguousarray(b.image(Q, i)), P, stats=True)
        outer.append(st["outer"])
        np.testing.assert_array_equal(b.image(Ql, i), q1)
    print("outer steps per image", outer, flush=True)


def test_batch_lp_with_bin_edges_matches_one_image_path():
    """each image's initial Q puts 0.5 / 0.5 on its own label pairs (0.5 lies
    on a bin edge): 6 (label, bin) pairs per image, 10 over the batch.  Every
    image must take the same edge mode as alone (correction of the upper
    pass), so the batched LP equals the one-image LP bit for bit."""
    import densecrf_ho_amd as D
    L = 21
    wls = _workloads(L)
    b = _batch(wls, L)
    rng = np.random.default_rng(9)
    parts = []
    for i, wl in enumerate(wls):
        q = np.zeros((wl.N, L), np.float32)
        q[np.arange(wl.N), rng.integers(0, L, wl.N)] = 1.0
        for j in range(3):
            pix = rng.choice(wl.N, max(4, wl.N // 20), replace=False)
            q[pix] = 0.0
            q[pix, 2 * i + 2 * j] = 0.5
            q[pix, 2 * i + 2 * j + 1] = 0.5
        parts.append(q)
    Q0 = np.ascontiguousarray(np.concatenate(parts))
    P = D.LP_inf_params(prox_max_iter=2, fw_max_iter=3)
    Ql, _ = b.lp(Q0, P)
    for i, wl in enumerate(wls):
        print("image", i, flush=True)
        c = _single(wl, L)
        q1 = c.lp_inference_prox_super_pixels(parts[i], P)
        np.testing.assert_array_equal(b.image(Ql, i), q1)


def test_bf16_rows_ordered_filter_error():
    """bfloat16 storage of the ordered filter's lattice rows: the output moves
    by the rows' rounding (relative error below 1e-2 of the FP64 oracle's
    largest value; the float rows stay within 1e-5)"""
    import densecrf_ho_amd as D
    from densecrf_ho_amd import synthetic as S
    L = 21
    wl = S.Workload(64, 48, 8, 7, L=L)
    q = np.random.default_rng(1).random((wl.N, L)).astype(np.float32)
    o = O.CRF(wl.N, L)
    for f, w in wl.kernels:
        o.add_pairwise(f, w)
    for k in range(2):
        ref = o.apply_ord(k, q, fp64=True)
        c = D.DenseCRF(wl.N, L)
        for f, w in wl.kernels:
            c.addPairwiseEnergy(f, D.PottsCompatibility(w))
        e32 = np.abs(c.apply_upper_minus_lower_ord(k, q) - ref).max() / np.abs(ref).max()
        c.set_value_type(True)
        e16 = np.abs(c.apply_upper_minus_lower_ord(k, q) - ref).max() / np.abs(ref).max()
        print("potential", k, "fp32", e32, "bf16", e16, flush=True)
        assert e32 < 1e-5
        assert 1e-5 < e16 < 1e-2


def test_batch_bf16_lp_close_to_float():
    import densecrf_ho_amd as D
    L = 21
    wls = _workloads(L)
    P = D.LP_inf_params(prox_max_iter=2, fw_max_iter=4)
    b = _batch(wls, L)
    Q, _ = b.qp(max_iter=10, ref_stop=False)
    Q32, _ = b.lp(Q, P)
    e32 = b.assignment_energy_true(Q32.argmax(1).astype(np.int32))
    b.set_value_type(True)
    Q16, _ = b.lp(Q, P)
    e16 = b.assignment_energy_true(Q16.argmax(1).astype(np.int32))
    rel = np.abs(e16 - e32) / np.abs(e32)
    print("bf16 vs float LP energies", rel, flush=True)
    assert rel.max() < 2e-2
    assert (Q16.argmax(1) == Q32.argmax(1)).mean() > 0.98
