"""Write the golden fixtures of tests/golden/ from the compiled reference.

The reference core is built from its own sources by ``make -C oracle ref``
into oracle/_ref/ (development container only).  This script drives it
through ctypes on seeded synthetic inputs and stores inputs and outputs as
compressed .npz files.  Run from the repository root:

    python tests/golden/make_golden.py          # every fixture
    python tests/golden/make_golden.py lp_lc    # only lp_lc_31x23_L5.npz
    python tests/golden/make_golden.py restricted  # only ordered_restricted_31x23.npz
    python tests/golden/make_golden.py bf       # only bf_12x9_L4.npz
    python tests/golden/make_golden.py data     # only data_img.json / data_img_labels.npz
    python tests/golden/make_golden.py data_lp_steps  # only data_img_lp_steps.npz (after data)
    python tests/golden/make_golden.py kgrad    # only kgrad_31x23.npz
    python tests/golden/make_golden.py dc       # only dc_L4.npz
    python tests/golden/make_golden.py eho      # only eho_32x24_L21.npz
    python tests/golden/make_golden.py logistic # only logistic_unary.npz
    python tests/golden/make_golden.py portable # only lattice_portable.npz (make -C oracle ref-scalar)
"""

from __future__ import annotations

import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from densecrf_ho_amd import synthetic as S  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def edge_q(N, V, seed):
    """values in [0,1] with exact bin edges k/10, 0 and 1 mixed in"""
    rng = np.random.default_rng(seed)
    q = rng.random((N, V)).astype(np.float32)
    m = rng.random((N, V)) < 0.25
    q[m] = (rng.integers(0, 11, m.sum()) / np.float32(10)).astype(np.float32)
    q[0, :] = 0.0
    q[1, :] = 1.0
    return q


def tie_features(n, seed):
    """d=1 features whose elevated coordinate is an exact odd integer, so that
    rounding to the nearest multiple of d+1=2 meets exact .5 ties
    (src/permutohedral.cpp:224-229)."""
    rng = np.random.default_rng(seed)
    s0 = np.float32(1.0 / np.sqrt(2.0) * np.float32(np.sqrt(2.0 / 3.0) * 2))
    out = []
    while len(out) < n:
        m = int(rng.integers(-40, 40)) * 2 + 1
        f = np.float32(m / float(s0))
        for _ in range(8):
            if np.float32(f * s0) == np.float32(m):
                out.append(f)
                break
            f = np.nextafter(f, np.float32(np.inf) if np.float32(f * s0) < m else np.float32(-np.inf))
        else:
            out.append(np.float32(rng.standard_normal() * 5))
    return np.array(out, np.float32).reshape(n, 1)


def lattice_case(name, feat, V=5, seed=0):
    lat = O.RefLattice(feat)
    A = lat.arrays()
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((lat.N, V)).astype(np.float32)
    x1 = (rng.random((lat.N, 1)) + 0.5).astype(np.float32)
    q = edge_q(lat.N, V, seed + 1)
    np.savez_compressed(os.path.join(OUT, f"lattice_{name}.npz"), feat=feat, M=np.int32(lat.M),
                        offset=A["offset"], rank=A["rank"], bary=A["bary"], nbr=A["nbr"],
                        x=x, out=lat.filter(x), out_rev=lat.filter(x, reverse=True), x1=x1, out1=lat.filter(x1),
                        q=q, ord=lat.ordered(q))
    print(f"lattice_{name}: N={lat.N} d={feat.shape[1]} M={lat.M}")


def crf_case(name, W, H, K, seed, L, layers=None, lp_outer=3):
    wl = S.Workload(W, H, K, seed, L=L)
    r = O.RefCRF(W, H, L)
    r.set_unary(wl.U)
    r.add_gaussian(3, 3, 3)
    r.add_bilateral(50, 50, 15, 15, 15, wl.img, 5)
    regions = []
    if layers:
        for i, (reg, _) in enumerate(S.superpixel_layers(W, H, layers, [21 + i for i in range(len(layers))])):
            r.add_layer(reg, 0.5, 0.2)
            regions.append(reg)
    rng = np.random.default_rng(seed + 7)
    x = rng.random((wl.N, L)).astype(np.float32)
    d = dict(W=np.int32(W), H=np.int32(H), K=np.int32(K), seed=np.int32(seed), L=np.int32(L), img=wl.img, U=wl.U,
             x=x)
    for i, reg in enumerate(regions):
        d[f"region{i}"] = reg
    for k in range(2):
        d[f"feat{k}"] = r.features(k, 2 if k == 0 else 5)
        d[f"apply{k}"] = r.apply(k, x, 0)
        d[f"applyT{k}"] = r.apply(k, x, 1)
        d[f"applyNN{k}"] = r.apply(k, x, 2)
        q = edge_q(wl.N, L, seed + 11 + k)
        d[f"ordq{k}"] = q
        d[f"ord{k}"] = r.apply_ord(k, q)
    Q0 = r.unary_init()
    d["Q0"] = Q0
    Qqp = r.qp(Q0)
    d["Qqp"] = Qqp
    if layers:
        d["zlabels"] = r.zlabels()
    lab = r.current_map(Qqp)
    d["lab_qp"] = lab.astype(np.int32)
    d["e_true"] = np.float64(r.energy_true(lab))
    d["e_ho"] = np.float64(r.energy_ho(lab))
    d["e_lp"] = np.float64(r.lp_energy(Qqp))
    d["e_qpvalue"] = np.float64(r.qp_value(Qqp))
    Qpert = (Qqp * 1.3 - 0.1).astype(np.float32)
    d["Qpert"] = Qpert
    d["feasible"] = O.ref_feasible(Qpert)
    d["renorm"] = O.ref_renormalize(np.abs(Qpert))
    d["ucond"] = r.ucondgrad(Qpert)
    P = O.LPParams(prox_max_iter=lp_outer)
    d["Qlp"] = r.lp(Qqp, P)
    d["lp_outer_cap"] = np.int32(lp_outer)
    np.savez_compressed(os.path.join(OUT, f"crf_{name}.npz"), **d)
    print(f"crf_{name}: N={wl.N} L={L} R={r.R}")


def qp_case(name, W, H, K, seed, L, weights=(3.0, 5.0), dup_label=None):
    """QP end-to-end on a small image.  Negative Potts weights make the
    problem convex so the line-search step is not clamped to 1; duplicated
    unary columns create exact ties for the first-minimum rule."""
    wl = S.Workload(W, H, K, seed, L=L)
    U = wl.U.copy()
    if dup_label is not None:
        U[:, dup_label[1]] = U[:, dup_label[0]]
    r = O.RefCRF(W, H, L)
    r.set_unary(U)
    r.add_gaussian(3, 3, weights[0])
    r.add_bilateral(50, 50, 15, 15, 15, wl.img, weights[1])
    Q0 = r.unary_init()
    Qqp = r.qp(Q0)
    np.savez_compressed(os.path.join(OUT, f"qp_{name}.npz"), W=np.int32(W), H=np.int32(H), L=np.int32(L), U=U,
                        img=wl.img, w=np.array(weights, np.float32), Q0=Q0, Qqp=Qqp,
                        feat0=r.features(0, 2), feat1=r.features(1, 5))
    print(f"qp_{name}: N={wl.N} L={L}")


def c0_case():
    """Config 0 (320x213, 40 cells, 21 labels): end-to-end QP with the
    reference's stopping rule and a 2-iteration LP, energies and labels."""
    W, H, K, L = 320, 213, 40, 21
    wl = S.Workload(W, H, K, 0, L=L)
    r = O.RefCRF(W, H, L)
    r.set_unary(wl.U)
    r.add_gaussian(3, 3, 3)
    r.add_bilateral(50, 50, 15, 15, 15, wl.img, 5)
    Q0 = r.unary_init()
    Qqp = r.qp(Q0)
    lab_qp = r.current_map(Qqp)
    P = O.LPParams(prox_max_iter=2)
    Qlp = r.lp(Qqp, P)
    lab_lp = r.current_map(Qlp)
    res = dict(e_true_qp=r.energy_true(lab_qp), qp_value=r.qp_value(Qqp), lp_energy_qp=r.lp_energy(Qqp),
               e_true_lp=r.energy_true(lab_lp), lp_energy_lp=r.lp_energy(Qlp),
               unary_sum=float(wl.U.astype(np.float64).sum()))
    np.savez_compressed(os.path.join(OUT, "c0_320x213.npz"), lab_qp=lab_qp.astype(np.uint8),
                        lab_lp=lab_lp.astype(np.uint8), Qqp_rowsum=Qqp.astype(np.float64).sum(1))
    with open(os.path.join(OUT, "c0_320x213.json"), "w") as f:
        json.dump(res, f, indent=1)
    print("c0:", res)


def compat_case(name, W, H, K, seed, L):
    """Potts / Diagonal / Matrix compatibilities on the bilateral kernel:
    apply, applyTranspose, the no-norm copy and PairwisePotential::gradient;
    plus a reference-rule QP with a Matrix term."""
    wl = S.Workload(W, H, K, seed, L=L)
    rng = np.random.default_rng(seed + 100)
    v = rng.uniform(-3, 3, L).astype(np.float32)
    m = rng.normal(0, 2, (L, L)).astype(np.float32)
    x = rng.random((wl.N, L)).astype(np.float32)
    b = rng.standard_normal((wl.N, L)).astype(np.float32)
    feat = wl.kernels[1][0]
    d = dict(U=wl.U, feat=feat, feat_g=wl.kernels[0][0], v=v, m=m, x=x, b=b)
    for kind, par in (("potts", [3.0]), ("diagonal", v), ("matrix", m)):
        r = O.RefCRF(W, H, L)
        r.set_unary(wl.U)
        r.add_pairwise_compat(feat, kind, par)
        d[f"{kind}_apply"] = r.apply(0, x, 0)
        d[f"{kind}_applyT"] = r.apply(0, x, 1)
        d[f"{kind}_nn"] = r.apply(0, x, 2)
        d[f"{kind}_grad"] = r.pairwise_gradient(0, b, x)
    r = O.RefCRF(W, H, L)
    r.set_unary(wl.U)
    r.add_pairwise_compat(wl.kernels[0][0], "potts", [3.0])
    r.add_pairwise_compat(feat, "matrix", -np.abs(m) - 1.0)
    q0 = r.unary_init()
    d["qp_m"] = -np.abs(m) - 1.0
    d["qp_Q"] = r.qp(q0)
    lab = np.zeros(wl.N, np.int32)
    r._R.ref_crf_current_map(r.h, O._c(d["qp_Q"], np.float32), lab)
    d["qp_lab"] = lab
    d["qp_energy_true"] = np.float64(r._R.ref_crf_energy_true(r.h, lab))
    np.savez_compressed(os.path.join(OUT, f"compat_{name}.npz"), **d)
    print(f"compat_{name}: N={wl.N} L={L}")


LC_SETTINGS = ((0.0, 0.95), (10.0, 0.6), (30.0, 0.6), (40.0, 0.6))


def lp_less_confident_case():
    """LP with the less-confident stop (src/densecrf.cpp:784-794): the same
    31x23, 5-label problem as crf_31x23_L5, four (percent, confidence_tol)
    settings that stop after 4 (no check), 1, 2 and 3 outer steps."""
    W, H, L = 31, 23, 5
    wl = S.Workload(W, H, 8, 1, L=L)
    r = O.RefCRF(W, H, L)
    r.set_unary(wl.U)
    r.add_gaussian(3, 3, 3)
    r.add_bilateral(50, 50, 15, 15, 15, wl.img, 5)
    Qqp = r.qp(r.unary_init())
    d = dict(W=np.int32(W), H=np.int32(H), L=np.int32(L), U=wl.U, feat0=r.features(0, 2), feat1=r.features(1, 5),
             Qqp=Qqp, settings=np.array(LC_SETTINGS, np.float32), prox_max_iter=np.int32(4))
    for i, (pct, tol) in enumerate(LC_SETTINGS):
        d[f"Qlp{i}"] = r.lp(Qqp, O.LPParams(prox_max_iter=4, less_confident_percent=pct, confidence_tol=tol))
    np.savez_compressed(os.path.join(OUT, "lp_lc_31x23_L5.npz"), **d)
    print("lp_lc_31x23_L5")


def restricted_case():
    """Permutohedral::compute_upper_minus_lower_ord_restricted
    (src/permutohedral.cpp:702-880) on the 31x23 lattices (d=2 and d=5),
    V=5, values with exact bin edges: (1) store with the pixels pI, (2) no
    store, same pI, new values, (3) no store, another pI (points outside the
    stored active set), output added to a nonzero field."""
    wl = S.Workload(31, 23, 8, 1)
    N, V = wl.N, 5
    d = {}
    for k, (f, _) in enumerate(wl.kernels):
        lat = O.RefLattice(f)
        rng = np.random.default_rng(40 + k)
        q = edge_q(N, V, 41 + k)
        pI = np.sort(rng.choice(N, N // 5, replace=False)).astype(np.int32)
        pI2 = np.sort(rng.choice(N, N // 4, replace=False)).astype(np.int32)
        rq1 = edge_q(len(pI), V, 42 + k)
        rq2 = edge_q(len(pI), V, 43 + k)
        rq3 = edge_q(len(pI2), V, 44 + k)
        o3 = rng.standard_normal((len(pI2), V)).astype(np.float32)
        d.update({f"feat{k}": f, f"q{k}": q, f"pI{k}": pI, f"pI2_{k}": pI2, f"rq1_{k}": rq1, f"rq2_{k}": rq2,
                  f"rq3_{k}": rq3, f"o3_{k}": o3})
        d[f"out1_{k}"] = lat.ordered_restricted(rq1, pI, q, True)
        d[f"out2_{k}"] = lat.ordered_restricted(rq2, pI, None, False)
        d[f"out3_{k}"] = lat.ordered_restricted(rq3, pI2, None, False, out=o3)
    np.savez_compressed(os.path.join(OUT, "ordered_restricted_31x23.npz"), **d)
    print("ordered_restricted_31x23")


def bf_case():
    """brute-force filters of the reference's PairwisePotential
    (src/pairwise.cpp:294-383) on a 12x9 image, 4 labels, both kernels:
    apply_bf(Q), apply_upper_minus_lower_bf(ind) with random orders, and
    apply_upper_minus_lower_bf_ord(ind, Q) with random orders and with the
    ascending order of Q (which has ties)"""
    W, H, L = 12, 9, 4
    wl = S.Workload(W, H, 3, 1, L=L)
    r = O.RefCRF(W, H, L)
    r.set_unary(wl.U)
    r.add_gaussian(3, 3, 3)
    r.add_bilateral(50, 50, 15, 15, 15, wl.img, 5)
    rng = np.random.default_rng(12)
    Q = rng.random((wl.N, L)).astype(np.float32)
    Q[5:9] = Q[1]  # equal values: ties in the value split
    ind = np.stack([rng.permutation(wl.N) for _ in range(L)], 1).astype(np.int32)
    ind_q = np.argsort(Q, axis=0, kind="stable").astype(np.int32)
    d = dict(W=np.int32(W), H=np.int32(H), L=np.int32(L), U=wl.U, img=wl.img, Q=Q, ind=ind, ind_q=ind_q)
    for k in range(2):
        d[f"feat{k}"] = r.features(k, 2 if k == 0 else 5)
        d[f"bf{k}"] = r.apply_bf(k, 0, Q=Q)
        d[f"uml{k}"] = r.apply_bf(k, 1, ind=ind)
        d[f"umlord{k}"] = r.apply_bf(k, 2, ind=ind, Q=Q)
        d[f"umlord_q{k}"] = r.apply_bf(k, 2, ind=ind_q, Q=Q)
    np.savez_compressed(os.path.join(OUT, "bf_12x9_L4.npz"), **d)
    print("bf_12x9_L4")


def data_case():
    """the reference's own example input (data/img.bmp, data/img.c_unary, kept
    in tests/golden/data/) through its example_inference defaults
    (src/example_inference.cpp:99-128): the unaries as the compiled
    ProbImage decodes them (checksum), and the discretised energies and
    labellings of the unary, qp and lp methods"""
    import hashlib
    import densecrf_ho_amd as D
    src = os.path.join(OUT, "data")
    p, W, H = O.ref_probimage(os.path.join(src, "img.c_unary"))
    U = (-np.log(p.astype(np.float64))).astype(np.float32)  # load_unary's -log (src/file_storage.cpp:285)
    im, Wi, Hi = D.load_image(os.path.join(src, "img.bmp"))
    assert (Wi, Hi) == (W, H)
    r = O.RefCRF(W, H, U.shape[1])
    r.set_unary(U)
    r.add_gaussian(2, 2, 3)
    r.add_bilateral(30, 30, 8, 8, 8, im, 2)
    Q0 = r.unary_init()
    out = dict(W=W, H=H, L=int(U.shape[1]), unary_sha256=hashlib.sha256(U.tobytes()).hexdigest(),
               unary_samples=[float(U[i, k]) for i, k in [(0, 0), (1234, 5), (68159, 20)]],
               image_sha256=hashlib.sha256(im.tobytes()).hexdigest())
    labs = {}
    for method, Q in [("unary", Q0), ("qp", r.qp(Q0)), ("lp", r.lp(Q0, O.LPParams()))]:
        lab = r.current_map(Q)
        out[f"e_{method}"] = float(r.energy_true(lab))
        labs[method] = lab.astype(np.uint8)
    # qp_sp / lp_sp: the reference's mean-shift superpixels (addSuperPixel(img,
    # 8, 4, 100, 50, 1000), src/example_inference.cpp:48,67) kept as a region
    # map with the layer's exp_of_superpixels_; qp_sp reports the higher-order
    # energy, lp_sp the pairwise one (src/example_inference.cpp:76-87)
    reg, e, c = r.add_superpixel(im, 8, 4, 100, 50.0, 1000.0)
    labs["regions"] = reg.astype(np.int32)
    labs["exp_of_superpixels"] = e
    labs["constants"] = c
    lab = r.current_map(r.qp(Q0))
    out["e_qp_sp"] = float(r.energy_ho(lab))
    labs["qp_sp"] = lab.astype(np.uint8)
    lab = r.current_map(r.lp(Q0, O.LPParams()))
    out["e_lp_sp"] = float(r.energy_true(lab))
    labs["lp_sp"] = lab.astype(np.uint8)
    # The LP methods amplify rounding: run the reference again from initial
    # fields one ulp away from its own unary_init (the oracle's softmax, which
    # differs from Eigen's vectorised exp by an ulp, and two random 1-ulp
    # perturbations) and keep the spread of its final energies and of the
    # label agreement with the unperturbed run.
    rng = np.random.default_rng(77)
    inits = [O.unary_init(U)]
    for _ in range(2):
        q = Q0.copy()
        m = rng.random(q.shape) < 0.02
        q[m] = np.nextafter(q[m], np.where(rng.random(m.sum()) < 0.5, 0, 1).astype(np.float32))
        inits.append(q)
    for method, layers in [("lp", False), ("lp_sp", True)]:
        rr = O.RefCRF(W, H, U.shape[1])
        rr.set_unary(U)
        rr.add_gaussian(2, 2, 3)
        rr.add_bilateral(30, 30, 8, 8, 8, im, 2)
        if layers:
            rr.add_layer(reg, e, c)
        es, ag = [out[f"e_{method}"]], []
        for q in inits:
            lab = rr.current_map(rr.lp(q, O.LPParams()))
            es.append(float(rr.energy_true(lab)))
            ag.append(float((lab == labs[method]).mean()))
        out[f"e_{method}_spread"] = [min(es), max(es)]
        out[f"agree_{method}_min"] = min(ag)
    with open(os.path.join(OUT, "data_img.json"), "w") as f:
        json.dump(out, f, indent=1)
    np.savez_compressed(os.path.join(OUT, "data_img_labels.npz"), **labs)
    print("data_img", {k: v for k, v in out.items() if k.startswith("e_")})


def data_lp_steps_case():
    """The LP methods on the reference's own example input from the
    reference's own unary_init, cut short (src/densecrf.cpp:526-799 with
    prox_max_iter / fw_max_iter): lp after one and two proximal steps, lp_sp
    after the first one and two Frank-Wolfe steps of its first proximal step.
    Those come before the rounding spread of the full runs sets in (with the
    superpixel layer, computeUCondGrad picks the last value above and below
    zero per region, src/densecrf.cpp:800-826, and values within rounding of
    zero decide it from the third step on).  The initial field is kept as its
    difference from the oracle's unary_init (the entries where Eigen's
    vectorised exp rounds differently), so a test rebuilds the reference's
    field bit for bit; the discretised energy and labels of every cut run are
    kept."""
    import densecrf_ho_amd as D
    src = os.path.join(OUT, "data")
    p, W, H = O.ref_probimage(os.path.join(src, "img.c_unary"))
    U = (-np.log(p.astype(np.float64))).astype(np.float32)
    im, _, _ = D.load_image(os.path.join(src, "img.bmp"))
    labs = dict(np.load(os.path.join(OUT, "data_img_labels.npz"), allow_pickle=False))
    out = {}
    for method, layers in [("lp", False), ("lp_sp", True)]:
        r = O.RefCRF(W, H, U.shape[1])
        r.set_unary(U)
        r.add_gaussian(2, 2, 3)
        r.add_bilateral(30, 30, 8, 8, 8, im, 2)
        if layers:
            r.add_layer(labs["regions"], labs["exp_of_superpixels"], labs["constants"])
        Q0 = r.unary_init()
        if method == "lp":
            q0o = O.unary_init(U)
            idx = np.flatnonzero(Q0.reshape(-1) != q0o.reshape(-1)).astype(np.int32)
            out["q0_idx"], out["q0_val"] = idx, Q0.reshape(-1)[idx]
        for prox, fw in ((1, 10), (2, 10)) if method == "lp" else ((1, 1), (1, 2)):
            lab = r.current_map(r.lp(Q0, O.LPParams(prox_max_iter=prox, fw_max_iter=fw)))
            out[f"e_{method}_{prox}x{fw}"] = np.float64(r.energy_true(lab))
            out[f"lab_{method}_{prox}x{fw}"] = lab.astype(np.uint8)
    np.savez_compressed(os.path.join(OUT, "data_img_lp_steps.npz"), **out)
    print("data_img_lp_steps", {k: float(v) for k, v in out.items() if k.startswith("e_")},
          "init entries patched", out["q0_idx"].size)


def kgrad_case():
    """kernel (feature) gradients: Permutohedral::gradient on the 31x23
    lattices (d=2 and d=5, V=5), and a standalone PairwisePotential's
    kernelGradient(b, Q) for DIAG and FULL kernels under the four
    normalisation types, and after setKernelParameters
    (src/permutohedral.cpp:1095-1178, src/pairwise.cpp:90-181, :396-407)"""
    wl = S.Workload(31, 23, 8, 1)
    rng = np.random.default_rng(31)
    d = {}
    for k, (f, _) in enumerate(wl.kernels):
        a = rng.standard_normal((wl.N, 5)).astype(np.float32)
        b = rng.standard_normal((wl.N, 5)).astype(np.float32)
        q = rng.random((wl.N, 5)).astype(np.float32)
        d.update({f"feat{k}": f, f"a{k}": a, f"b{k}": b, f"q{k}": q})
        d[f"df{k}"] = O.RefLattice(f).gradient(a, b)
        for kt in (1, 2):
            for nt in (0, 1, 2, 3):
                d[f"kg{k}_{kt}_{nt}"] = O.ref_pairwise_kernel_gradient(f, 3.0, kt, nt, b, q)
        dd = f.shape[1]
        kp_diag = (0.5 + rng.random(dd)).astype(np.float32)
        kp_full = (np.eye(dd) + 0.1 * rng.standard_normal((dd, dd))).astype(np.float32).ravel(order="F")
        d[f"kp_diag{k}"] = kp_diag
        d[f"kp_full{k}"] = kp_full
        d[f"kgp{k}_1"] = O.ref_pairwise_kernel_gradient(f, 3.0, 1, 3, b, q, kp=kp_diag)
        d[f"kgp{k}_2"] = O.ref_pairwise_kernel_gradient(f, 3.0, 2, 3, b, q, kp=kp_full)
    np.savez_compressed(os.path.join(OUT, "kgrad_31x23.npz"), **d)
    print("kgrad_31x23")


def dc_case():
    """the divide-and-conquer filter: PairwisePotential::apply_upper_minus_lower_dc
    (src/pairwise.cpp:241-293) on 12x9 and 31x23 images, 4 labels, random
    orders, both kernels; Permutohedral::compute_upper_minus_lower_dc
    (src/permutohedral.cpp:881-997) for even and odd splits added to a
    nonzero row"""
    d = {}
    for W, H in [(12, 9), (31, 23)]:
        L = 4
        wl = S.Workload(W, H, 3, 1, L=L)
        r = O.RefCRF(W, H, L)
        r.set_unary(wl.U)
        r.add_gaussian(3, 3, 3)
        r.add_bilateral(50, 50, 15, 15, 15, wl.img, 5)
        rng = np.random.default_rng(W)
        ind = np.stack([rng.permutation(wl.N) for _ in range(L)], 1).astype(np.int32)
        tag = f"{W}x{H}"
        d[f"U_{tag}"] = wl.U
        d[f"img_{tag}"] = wl.img
        d[f"ind_{tag}"] = ind
        for k in range(2):
            f = r.features(k, 2 if k == 0 else 5)
            d[f"feat{k}_{tag}"] = f
            d[f"dc{k}_{tag}"] = r.apply_bf(k, 3, ind=ind)
            lat = O.RefLattice(f)
            out0 = rng.standard_normal(wl.N).astype(np.float32)
            d[f"row0_{k}_{tag}"] = out0
            N = wl.N
            d[f"lat_even{k}_{tag}"] = lat.dc(out0, 0, N // 2, N // 2, N)
            d[f"lat_odd{k}_{tag}"] = lat.dc(out0, 3, (N - 3) // 2 + 3, (N - 3) // 2 + 4, N - 2)
    np.savez_compressed(os.path.join(OUT, "dc_L4.npz"), **d)
    print("dc_L4")


def eho_case():
    """DenseCRF::compute_energy_higher_order (src/densecrf.cpp:960-980) after the
    QP of crf_32x24_L21_ho (its z_labels_): at the QP's result and at another Q"""
    W, H, L = 32, 24, 21
    wl = S.Workload(W, H, 10, 2, L=L)
    r = O.RefCRF(W, H, L)
    r.set_unary(wl.U)
    r.add_gaussian(3, 3, 3)
    r.add_bilateral(50, 50, 15, 15, 15, wl.img, 5)
    for i, (reg, _) in enumerate(S.superpixel_layers(W, H, (5, 9, 17), [21 + i for i in range(3)])):
        r.add_layer(reg, 0.5, 0.2)
    Qqp = r.qp(r.unary_init())
    Qo = np.random.default_rng(9).random((W * H, L)).astype(np.float32)
    np.savez_compressed(os.path.join(OUT, "eho_32x24_L21.npz"), Qo=Qo, e_qp=np.float64(r.energy_higher_order(Qqp)),
                        e_other=np.float64(r.energy_higher_order(Qo)))
    print("eho_32x24_L21")


def logistic_case():
    """LogisticUnaryEnergy(L, f).get() and gradient(b) of the compiled reference
    (src/unary.cpp:53-68) on three shapes: 21 labels x 37 features x 713 pixels
    (feature and pixel counts off the kernels' tiles) and 5 x 3 x 1000 with
    normal values, and 21 x 70 x 4200 (several feature chunks, feature blocks
    and pixel blocks) with multiples of 1/8 in [-2, 2], whose products and sums
    are exact in float in any order"""
    rng = np.random.default_rng(61)
    out = {}
    for i, (L, F, N) in enumerate([(21, 37, 713), (5, 3, 1000), (21, 70, 4200)]):
        if i < 2:
            draw = lambda shape: rng.standard_normal(shape).astype(np.float32)  # noqa: E731
        else:
            draw = lambda shape: (rng.integers(-16, 17, shape) / 8).astype(np.float32)  # noqa: E731
        Lw, f, b = draw((F, L)), draw((N, F)), draw((N, L))
        U, g = O.ref_logistic_unary(Lw, f, b)
        out.update({f"Lw{i}": Lw, f"f{i}": f, f"b{i}": b, f"U{i}": U, f"g{i}": g})
    np.savez_compressed(os.path.join(OUT, "logistic_unary.npz"), **out)
    print("logistic_unary")


def portable_case():
    """lattices of the reference's portable build (no __SSE__): tables, the
    filter at V=1 and V=21 both ways and the ordered filter, for features with
    exact rounding ties and two image kernels"""
    wl = S.Workload(31, 23, 8, 1)
    rng = np.random.default_rng(21)
    cases = {"ties_d1": tie_features(61, 7), "b5_31x23": wl.kernels[1][0], "g2_31x23": wl.kernels[0][0],
             "r3_101": (rng.standard_normal((101, 3)) * 3).astype(np.float32)}
    d = {}
    for k, feat in cases.items():
        lat = O.RefScalarLattice(feat)
        A = lat.arrays()
        x = rng.standard_normal((lat.N, 21)).astype(np.float32)
        x1 = (rng.random((lat.N, 1)) + 0.5).astype(np.float32)
        q = edge_q(lat.N, 21, 22)
        d.update({f"{k}_feat": feat, f"{k}_M": np.int32(lat.M), f"{k}_offset": A["offset"], f"{k}_rank": A["rank"],
                  f"{k}_bary": A["bary"], f"{k}_nbr": A["nbr"], f"{k}_x": x, f"{k}_out": lat.filter(x),
                  f"{k}_out_rev": lat.filter(x, reverse=True), f"{k}_x1": x1, f"{k}_out1": lat.filter(x1),
                  f"{k}_out1_rev": lat.filter(x1, reverse=True), f"{k}_q": q, f"{k}_ord": lat.ordered(q)})
        print(f"portable {k}: N={lat.N} d={feat.shape[1]} M={lat.M}")
    np.savez_compressed(os.path.join(OUT, "lattice_portable.npz"), names=np.array(list(cases)), **d)


def main():
    if sys.argv[1:] == ["portable"]:
        portable_case()
        return
    if sys.argv[1:] == ["logistic"]:
        logistic_case()
        return
    if sys.argv[1:] == ["eho"]:
        eho_case()
        return
    if sys.argv[1:] == ["dc"]:
        dc_case()
        return
    if sys.argv[1:] == ["kgrad"]:
        kgrad_case()
        return
    if sys.argv[1:] == ["data"]:
        data_case()
        return
    if sys.argv[1:] == ["data_lp_steps"]:
        data_lp_steps_case()
        return
    if sys.argv[1:] == ["bf"]:
        bf_case()
        return
    if sys.argv[1:] == ["lp_lc"]:
        lp_less_confident_case()
        return
    if sys.argv[1:] == ["restricted"]:
        restricted_case()
        return
    W, H = 32, 24
    wl = S.Workload(W, H, 10, 0)
    lattice_case("g2_32x24", wl.kernels[0][0])
    lattice_case("b5_32x24", wl.kernels[1][0])
    wl2 = S.Workload(31, 23, 8, 1)
    lattice_case("b5_31x23", wl2.kernels[1][0], seed=3)
    rng = np.random.default_rng(5)
    lattice_case("r3_101", (rng.standard_normal((101, 3)) * 3).astype(np.float32), seed=4)
    lattice_case("r1_50", (rng.standard_normal((50, 1)) * 2).astype(np.float32), seed=6)
    lattice_case("ties_d1", tie_features(61, 7), seed=8)
    crf_case("31x23_L5", 31, 23, 8, 1, 5)
    crf_case("32x24_L21_ho", 32, 24, 10, 2, 21, layers=(5, 9, 17), lp_outer=2)
    qp_case("convex_31x23_L5", 31, 23, 8, 1, 5, weights=(-3.0, -5.0))
    qp_case("ties_31x23_L5", 31, 23, 8, 1, 5, dup_label=(1, 3))
    compat_case("31x23_L5", 31, 23, 8, 1, 5)
    lp_less_confident_case()
    restricted_case()
    bf_case()
    kgrad_case()
    dc_case()
    eho_case()
    data_case()
    data_lp_steps_case()
    c0_case()
    portable_case()


if __name__ == "__main__":
    main()
