"""GPU lattice and filters against the oracle and the reference fixtures.

Integer tables (offset, rank, neighbours, M) and barycentric weights must be
bit-identical to the reference; filters are compared within 1e-5 relative
(they are bit-identical wherever a lattice point's splat fits one chunk)."""

import numpy as np
import pytest

import oracle as O
from conftest import rel_err

pytestmark = pytest.mark.gpu

LATTICES = ["g2_32x24", "b5_32x24", "b5_31x23", "r3_101", "r1_50", "ties_d1"]


def gpu_lattice(feat):
    import densecrf_ho_amd as D
    return D.Permutohedral().init(feat)


@pytest.mark.parametrize("name", LATTICES)
def test_tables_bit_exact_vs_reference(golden, name):
    g = golden(f"lattice_{name}.npz")
    print("case", name, flush=True)
    lat = gpu_lattice(g["feat"])
    T = lat.tables()
    assert lat.M == int(g["M"])
    np.testing.assert_array_equal(T["offset"], g["offset"])
    np.testing.assert_array_equal(T["rank"], g["rank"])
    np.testing.assert_array_equal(T["bary"], g["bary"])
    np.testing.assert_array_equal(T["nbr"], g["nbr"])


@pytest.mark.parametrize("name", LATTICES)
def test_filters_vs_reference(golden, name):
    g = golden(f"lattice_{name}.npz")
    print("case", name, flush=True)
    lat = gpu_lattice(g["feat"])
    assert rel_err(lat.compute(g["x"]), g["out"]) < 1e-5
    assert rel_err(lat.compute(g["x"], reverse=True), g["out_rev"]) < 1e-5
    assert rel_err(lat.compute(g["x1"]), g["out1"]) < 1e-5
    assert rel_err(lat.compute_upper_minus_lower_ord(g["q"]), g["ord"]) < 1e-5


@pytest.mark.parametrize("name", ["ties_d1", "b5_31x23", "g2_31x23", "r3_101"])
def test_portable_build_vs_reference(golden, name):
    """set_reference_path(True): tables bit-identical to the reference's
    portable build, filters at V=1 and V=21 both ways within 1e-5 (the splat
    and the double blur of seqCompute), the ordered filter within 1e-5; the
    default path returns afterwards"""
    import densecrf_ho_amd as D
    g = golden("lattice_portable.npz")
    print("case", name, flush=True)
    D.set_reference_path(True)
    try:
        assert D.reference_path()
        lat = gpu_lattice(g[f"{name}_feat"])
    finally:
        D.set_reference_path(False)
    assert not D.reference_path()
    T = lat.tables()
    assert lat.M == int(g[f"{name}_M"])
    for k in ("offset", "rank", "bary", "nbr"):
        np.testing.assert_array_equal(T[k], g[f"{name}_{k}"], err_msg=k)
    for x, o in (("x", "out"), ("x1", "out1")):
        assert rel_err(lat.compute(g[f"{name}_{x}"]), g[f"{name}_{o}"]) < 1e-5
        assert rel_err(lat.compute(g[f"{name}_{x}"], reverse=True), g[f"{name}_{o}_rev"]) < 1e-5
    assert rel_err(lat.compute_upper_minus_lower_ord(g[f"{name}_q"]), g[f"{name}_ord"]) < 1e-5


def test_reference_path_rejects_other_values():
    import densecrf_ho_amd as D
    from densecrf_ho_amd import _lib
    assert _lib.load().dcrf_set_reference_path(2) != 0
    assert not D.reference_path()


@pytest.mark.parametrize("W,H,K,d", [(320, 213, 40, 5), (641, 479, 120, 2), (641, 479, 120, 5)])
def test_image_lattice_vs_oracle(W, H, K, d):
    """Voronoi images (641x479 has N % 4 == 3: padding lanes)."""
    from densecrf_ho_amd import synthetic as S
    print("case", W, H, d, flush=True)
    wl = S.Workload(W, H, K, 1)
    feat = wl.kernels[0 if d == 2 else 1][0]
    ref = O.Lattice(feat)
    A = ref.arrays()
    lat = gpu_lattice(feat)
    T = lat.tables()
    assert lat.M == ref.M
    for k in ("offset", "rank", "bary", "nbr", "keys"):
        np.testing.assert_array_equal(T[k], A[k], err_msg=k)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((wl.N, 21)).astype(np.float32)
    out = lat.compute(x)
    assert rel_err(out, ref.filter(x.astype(np.float64), fp64=True)) < 1e-4
    assert rel_err(out, ref.filter(x)) < 1e-5
    q = rng.random((wl.N, 21)).astype(np.float32)
    o2 = lat.compute_upper_minus_lower_ord(q)
    assert rel_err(o2, ref.ordered(q, fp64=True)) < 1e-4


@pytest.mark.parametrize("ncombo", [1, 3, 8, 9])
def test_ordered_filter_bin_edge_modes(ncombo):
    """q exactly on a bin edge (0.5, 0.7f and 0.9f give bin_up = bin_lo - 1)
    for `ncombo` (label, bin) pairs, 40 pixels each: up to 8 pairs take the
    upper pass as Tb - lower + a filtered correction, 9 pairs the full upper
    pass; both within 1e-5 of the float oracle and 1e-4 of the FP64 oracle"""
    from densecrf_ho_amd import synthetic as S
    print("combos", ncombo, flush=True)
    wl = S.Workload(96, 64, 12, 4)
    feat = wl.kernels[1][0]
    ref = O.Lattice(feat)
    lat = gpu_lattice(feat)
    rng = np.random.default_rng(11)
    q = rng.random((wl.N, 21)).astype(np.float32)
    combos = [(k, v) for k in range(3) for v in (0.5, 0.7, 0.9)][:ncombo]
    for k, v in combos:
        q[rng.choice(wl.N, 40, replace=False), k] = np.float32(v)
    out = lat.compute_upper_minus_lower_ord(q)
    assert rel_err(out, ref.ordered(q)) < 1e-5
    assert rel_err(out, ref.ordered(q, fp64=True)) < 1e-4


def _ordered_fields(kind, N, L, rng, cells):
    """rescaled fields in [0, 1] of the forms the LP produces: one-hot rows
    (every label of a point uniform), a background level in one bin with the
    dominant label near 1 and a few spread values, dense random values, and
    values on bin edges in 1 or 9 (label, bin) pairs (correction and full
    upper-pass modes)"""
    lab = cells % L
    if kind == "onehot":
        q = np.zeros((N, L), np.float32)
        q[np.arange(N), lab] = 1.0
    elif kind == "background":
        q = (0.1 + 0.08 * rng.random((N, L))).astype(np.float32)
        q[np.arange(N), lab] = (0.85 + 0.15 * rng.random(N)).astype(np.float32)
        m = rng.random((N, L)) < 0.03
        q[m] = rng.random(m.sum()).astype(np.float32)
    else:
        q = rng.random((N, L)).astype(np.float32)
    if kind.startswith("edges"):
        n = int(kind[5:])
        combos = [(k, v) for k in range(3) for v in (0.5, 0.7, 0.9)][:n]
        for k, v in combos:
            q[rng.choice(N, 40, replace=False), k] = np.float32(v)
    return q


@pytest.mark.parametrize("L", [5, 21, 33])
@pytest.mark.parametrize("kind", ["onehot", "background", "random", "edges1", "edges9"])
@pytest.mark.parametrize("d", [2, 5])
def test_coded_ordered_filter_equals_dense_rows(kind, d, L, monkeypatch):
    """The coded ordered filter (per-(point, label) codes, rows of mixed
    labels only) gives the dense rows' results bit for bit, in every bin-edge
    mode; both against the float oracle within 1e-5."""
    from densecrf_ho_amd import synthetic as S
    print("case", kind, "d", d, "L", L, flush=True)
    wl = S.Workload(320, 213, 40, 6)
    feat = wl.kernels[0 if d == 2 else 1][0]
    rng = np.random.default_rng(L * 100 + d)
    cells = np.asarray(wl.region).reshape(-1)
    q = _ordered_fields(kind, wl.N, L, rng, cells)
    lat = gpu_lattice(feat)
    monkeypatch.delenv("DCRF_ORDERED_DENSE", raising=False)
    coded = lat.compute_upper_minus_lower_ord(q)
    coded2 = lat.compute_upper_minus_lower_ord(q)  # a second call over the first call's buffers
    monkeypatch.setenv("DCRF_ORDERED_DENSE", "1")
    dense = lat.compute_upper_minus_lower_ord(q)
    np.testing.assert_array_equal(coded.view(np.uint32), dense.view(np.uint32))
    np.testing.assert_array_equal(coded2.view(np.uint32), dense.view(np.uint32))
    assert rel_err(coded, O.Lattice(feat).ordered(q)) < 1e-5


def test_ordered_filter_rejects_out_of_range(golden):
    import densecrf_ho_amd as D
    g = golden("lattice_b5_31x23.npz")
    lat = gpu_lattice(g["feat"])
    q = g["q"].copy()
    q[5, 2] = -0.25
    with pytest.raises(D.EngineError, match="Value out of bounds"):
        lat.compute_upper_minus_lower_ord(q)


@pytest.mark.parametrize("V", [8, 9, 10, 11, 12, 24, 64])
@pytest.mark.parametrize("d", [2, 5])
def test_wide_rows_vs_oracle(V, d):
    """Value widths V >= 8 take the kernels with one lane per quad of values:
    whole quads and a last quad of 1, 2 or 3 values (V % 4), forward and
    reverse blur order, against the oracle's float and FP64 filters."""
    from densecrf_ho_amd import synthetic as S
    print("case V", V, "d", d, flush=True)
    wl = S.Workload(96, 64, 16, 5)
    feat = wl.kernels[0 if d == 2 else 1][0]
    ref = O.Lattice(feat)
    lat = gpu_lattice(feat)
    rng = np.random.default_rng(V * 10 + d)
    x = rng.standard_normal((wl.N, V)).astype(np.float32)
    for rev in (False, True):
        out = lat.compute(x, reverse=rev)
        assert rel_err(out, ref.filter(x, reverse=rev)) < 1e-5
        assert rel_err(out, ref.filter(x.astype(np.float64), reverse=rev, fp64=True)) < 1e-4


def _restricted_gpu_calls(lat, g, k):
    out1 = lat.compute_upper_minus_lower_ord_restricted(None, g[f"rq1_{k}"], g[f"pI{k}"], g[f"q{k}"], True)
    out2 = lat.compute_upper_minus_lower_ord_restricted(None, g[f"rq2_{k}"], g[f"pI{k}"], None, False)
    out3 = lat.compute_upper_minus_lower_ord_restricted(g[f"o3_{k}"].copy(), g[f"rq3_{k}"], g[f"pI2_{k}"], None,
                                                        False)
    return out1, out2, out3


@pytest.mark.parametrize("k", [0, 1])
def test_restricted_ordered_vs_reference(golden, k):
    """compute_upper_minus_lower_ord_restricted (src/permutohedral.cpp:702-880):
    a call with store, one reusing the stored splat and active points, one
    with another pixel set added into a nonzero output; within 1e-5 of the
    compiled reference (the kernels keep its float order)"""
    g = golden("ordered_restricted_31x23.npz")
    print("case d =", 2 if k == 0 else 5, flush=True)
    lat = gpu_lattice(g[f"feat{k}"])
    for i, o in enumerate(_restricted_gpu_calls(lat, g, k)):
        ref = g[f"out{i + 1}_{k}"]
        print("call", i + 1, "max abs diff", float(np.abs(o - ref).max()), flush=True)
        assert rel_err(o, ref) < 1e-5


def test_restricted_ordered_errors(golden):
    import densecrf_ho_amd as D
    g = golden("ordered_restricted_31x23.npz")
    lat = gpu_lattice(g["feat1"])
    pI, rq = g["pI1"], g["rq1_1"]
    with pytest.raises(D.EngineError, match="store first"):
        lat.compute_upper_minus_lower_ord_restricted(None, rq, pI, None, False)
    with pytest.raises(D.EngineError, match="ascending"):
        lat.compute_upper_minus_lower_ord_restricted(None, rq, pI[::-1].copy(), g["q1"], True)
    bad = rq.copy()
    bad[3, 1] = 1.5
    with pytest.raises(D.EngineError, match="Value out of bounds"):
        lat.compute_upper_minus_lower_ord_restricted(None, bad, pI, g["q1"], True)
    # no pixel in pI with store: the reference reads pI[0] there (undefined), so it is refused
    with pytest.raises(D.EngineError, match="at least one pixel"):
        lat.compute_upper_minus_lower_ord_restricted(None, np.zeros((0, 5), np.float32), np.zeros(0, np.int32),
                                                     g["q1"], True)
    # after a call with store, an empty list without store slices nothing
    first = lat.compute_upper_minus_lower_ord_restricted(None, rq, pI, g["q1"], True)
    empty = lat.compute_upper_minus_lower_ord_restricted(None, np.zeros((0, 5), np.float32),
                                                         np.zeros(0, np.int32), None, False)
    assert empty.shape == (0, 5)
    # and the stored state still serves the next call as it did the first
    again = lat.compute_upper_minus_lower_ord_restricted(None, rq, pI, None, False)
    np.testing.assert_array_equal(again, first)


@pytest.mark.parametrize("d", [2, 5])
def test_restricted_ordered_every_pixel_is_ordered_filter(d):
    """640x480 (N % 4 == 0): with every pixel in pI every lattice point is
    active and the restricted filter is the ordered filter"""
    from densecrf_ho_amd import synthetic as S
    wl = S.Workload(640, 480, 120, 1, L=21)
    feat = wl.kernels[0 if d == 2 else 1][0]
    lat = gpu_lattice(feat)
    q = np.random.default_rng(11).random((wl.N, 21)).astype(np.float32)
    pI = np.arange(wl.N, dtype=np.int32)
    a = lat.compute_upper_minus_lower_ord_restricted(None, q, pI, q, True)
    b = lat.compute_upper_minus_lower_ord(q)
    print("max abs diff", float(np.abs(a - b).max()), flush=True)
    assert rel_err(a, b) < 1e-5


@pytest.mark.parametrize("k", [0, 1])
def test_lattice_gradient_vs_reference(golden, k):
    """Permutohedral::gradient: d/df a^T K b against the compiled reference
    (within 1e-5 of the largest entry)"""
    g = golden("kgrad_31x23.npz")
    lat = gpu_lattice(g[f"feat{k}"])
    df = lat.gradient(g[f"a{k}"], g[f"b{k}"])
    print("max abs diff", float(np.abs(df - g[f"df{k}"]).max()), flush=True)
    assert rel_err(df, g[f"df{k}"]) < 1e-5


@pytest.mark.parametrize("k", [0, 1])
def test_pairwise_kernel_gradient_vs_reference(golden, k):
    """PairwisePotential::kernelGradient for DIAG and FULL kernels under the
    four normalisations, and after setKernelParameters, against the compiled
    reference (1e-3 of the largest entry: the parameter sums run over every
    pixel, in FP64 here and in float in the reference)"""
    import densecrf_ho_amd as D
    g = golden("kgrad_31x23.npz")
    f, b, q = g[f"feat{k}"], g[f"b{k}"], g[f"q{k}"]
    for kt in (1, 2):
        for nt in (0, 1, 2, 3):
            print("kernel type", kt, "normalisation", nt, flush=True)
            p = D.PairwisePotential(f, D.PottsCompatibility(3.0), kt, nt)
            got = p.kernelGradient(b, q)
            assert rel_err(got, g[f"kg{k}_{kt}_{nt}"]) < 1e-3
    p = D.PairwisePotential(f, D.PottsCompatibility(3.0), D.DIAG_KERNEL, 3)
    np.testing.assert_array_equal(p.kernelParameters(), np.ones(f.shape[1], np.float32))
    p.setKernelParameters(g[f"kp_diag{k}"])
    np.testing.assert_array_equal(p.kernelParameters(), g[f"kp_diag{k}"])
    assert rel_err(p.kernelGradient(b, q), g[f"kgp{k}_1"]) < 1e-3
    p = D.PairwisePotential(f, D.PottsCompatibility(3.0), D.FULL_KERNEL, 3)
    p.setKernelParameters(g[f"kp_full{k}"])
    assert rel_err(p.kernelGradient(b, q), g[f"kgp{k}_2"]) < 1e-3
    c = D.PairwisePotential(f, D.PottsCompatibility(3.0))  # CONST_KERNEL: no parameters
    assert c.kernelParameters().size == 0 and c.kernelGradient(b, q).size == 0


@pytest.mark.parametrize("d", [2, 5])
def test_gradient_and_restricted_at_320x213_vs_oracle(d):
    """the lattice gradient (V = 21) and the restricted ordered filter (a third
    of the pixels) at 320x213 against the oracle"""
    from densecrf_ho_amd import synthetic as S
    wl = S.Workload(320, 213, 40, 0, L=21)
    feat = wl.kernels[0 if d == 2 else 1][0]
    lat = gpu_lattice(feat)
    ref = O.Lattice(feat)
    rng = np.random.default_rng(d)
    a = rng.standard_normal((wl.N, 21)).astype(np.float32)
    b = rng.standard_normal((wl.N, 21)).astype(np.float32)
    g = lat.gradient(a, b)
    assert rel_err(g, ref.gradient(a, b)) < 1e-4
    q = rng.random((wl.N, 21)).astype(np.float32)
    pI = np.sort(rng.choice(wl.N, wl.N // 3, replace=False)).astype(np.int32)
    rq = rng.random((len(pI), 21)).astype(np.float32)
    out = lat.compute_upper_minus_lower_ord_restricted(None, rq, pI, q, True)
    assert rel_err(out, ref.ordered_restricted(rq, pI, q, True)) < 1e-5


def test_less_confident_pixels():
    """less_confident_pixels (src/densecrf.cpp:457-462): ascending indices of
    the pixels whose largest value is <= tol (ties at tol included)"""
    import densecrf_ho_amd as D
    rng = np.random.default_rng(13)
    Q = rng.random((50000, 21)).astype(np.float32) ** 8
    Q[7] = np.float32(0.5)  # a pixel exactly at the tolerance
    for tol in (0.5, 0.9, 0.99):
        got = D.less_confident_pixels(Q, tol)
        np.testing.assert_array_equal(got, np.nonzero(Q.max(1) <= np.float32(tol))[0])
