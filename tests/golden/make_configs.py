"""Fixtures of the benchmark configurations at their stated sizes.

Writes compact results (labels as uint8, FP64 energies, sweep and iteration
counts, gradient projections) of the float oracle, the FP64 oracle and, where
it was built, the compiled reference (oracle/_ref) for:

  c1   LP at 640x480 (120 Voronoi cells, seed 1): QP 10 iterations from
       unary_init, then the LP with 5 proximal x 10 Frank-Wolfe steps; also
       the reference's QP + LP with the unaries moved one ulp per entry
       (4 seeds), the spread of its own result under rounding
  c2   qp_sp / lp_sp at 1280x720 (230 cells, seed 2) with three Voronoi
       superpixel layers of 300 / 600 / 1200 cells (seeds 21 / 22 / 23),
       Pn weight 0.5 and LP multiplier 0.2: the QP with the reference's
       stopping rule from unary_init, then the LP 5 x 10 (and 4 runs with
       one-ulp perturbed unaries, as for c1)
  c3   the three smallest images of the 64-image batch (seeds 3000+i):
       QP 10 iterations + LP 5 x 10 (and 3 runs with one-ulp perturbed unaries each)
  c4   the unrolled mean field (5 updates) and its adjoint on the 8 images of
       512x512 (seeds 4000+i), spatial Potts (w=3) and bilateral Matrix
       (m = -5 I) terms, cross-entropy against labels of seed 4100+i: FP64
       loss, parameter gradients, dU and Q projected on 3 seeded directions
       and at 4096 seeded entries

Development container only (the compiled reference and minutes of host time
are needed).  Run from the repository root:

    python tests/golden/make_configs.py            # every configuration
    python tests/golden/make_configs.py c1 c4      # some of them
    MAKE_CONFIGS_ONLY=s MAKE_CONFIGS_MERGE=1 python tests/golden/make_configs.py c1 c2
                                                   # only the short-LP runs, merged in

The short LP (one proximal step of 5 Frank-Wolfe steps from the QP result,
kind 's') is what the GPU is held to at 1e-4: the proximal LP amplifies
rounding (its 10-bin ordered filter is discontinuous in Q), so a full 5 x 10
run of any implementation that rounds differently, the FP64 oracle included,
ends up to several percent from the reference's, and the p-runs measure how
far the reference itself moves under one-ulp changes of its unaries.
"""

from __future__ import annotations

import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from densecrf_ho_amd import synthetic as S  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
L = 21
QP_IT = 10
LP = dict(prox_max_iter=5, fw_max_iter=10)
LP_SHORT = dict(prox_max_iter=1, fw_max_iter=5)  # one proximal step of 5 inner steps

C1 = dict(W=640, H=480, K=120, seed=1)
C2 = dict(W=1280, H=720, K=230, seed=2, cells=(300, 600, 1200), seeds=(21, 22, 23), e=0.5, c=0.2)
C3_IMAGES = 64
C4 = dict(B=8, W=512, T=5)


def log(msg):
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)


def orc(wl, layers=(), U=None):
    c = O.CRF(wl.N, L)
    c.set_unary(wl.U if U is None else U)
    for f, w in wl.kernels:
        c.add_pairwise(f, w)
    for region in layers:
        c.add_layer(region, C2["e"], C2["c"])
    return c


def ref(wl, layers=(), U=None):
    c = O.RefCRF(wl.W, wl.H, L)
    c.set_unary(wl.U if U is None else U)
    for f, w in wl.kernels:
        c.add_pairwise(f, w)
    for region in layers:
        c.add_layer(region, C2["e"], C2["c"])
    return c


def u8(lab):
    return np.asarray(lab, np.uint8)


# ---------------------------------------------------------------- c1 -------
def ulp_perturbed(Q, seed):
    """every entry of Q moved one ulp up or down (seeded signs); applied to the
    unaries, whose every inner LP step reads them (a perturbed starting field
    would be erased when the LP restarts from max_rounding,
    src/densecrf.cpp:544-547)"""
    sgn = np.random.default_rng(seed).integers(0, 2, Q.shape).astype(bool)
    return np.where(sgn, np.nextafter(Q, np.float32(np.inf)), np.nextafter(Q, np.float32(-np.inf))).astype(np.float32)


def lp_stats(out, kind, st):
    out[f"sweeps_{kind}"] = np.asarray(st["qp_sweeps"], np.int32)
    out[f"outer_{kind}"] = np.int32(st["outer"])
    out[f"int_energy_{kind}"] = np.asarray(st["int_energy"], np.float64)
    out[f"gap_{kind}"] = np.asarray(st["gap"], np.float64)
    out[f"delta_{kind}"] = np.asarray(st["delta"], np.float64)


def c1_run(kind):
    """kind: 'f' float oracle, 'd' FP64 oracle, 'r' compiled reference (LP from
    the float oracle's QP, which is the reference's QP body bit for bit), 'p<k>'
    the compiled reference's QP + LP with the unaries moved one ulp per entry
    (seed k), energies of the unperturbed model: the spread of the reference's
    own result under rounding"""
    wl = S.Workload(C1["W"], C1["H"], C1["K"], C1["seed"], L=L)
    o = orc(wl)
    q0 = O.unary_init(wl.U)
    fp64 = kind == "d"
    t0 = time.time()
    qqp, sq = o.qp(q0, max_iter=QP_IT, fp64=fp64)
    lab_qp = o.current_map(qqp)
    if kind == "s":  # the reference's short LP from the same QP result, and the oracle's stats
        qs = ref(wl).lp(qqp, O.LPParams(**LP_SHORT))
        _, st = o.lp(qqp, O.LPParams(**LP_SHORT))
        lab = o.current_map(qs)
        log(f"c1 s: {time.time() - t0:.1f} s")
        return {"lab_short_r": u8(lab), "e_short_r": o.energy_true(lab, fp64=True),
                "gap_short_f": np.asarray(st["gap"], np.float64), "sweeps_short_f": np.asarray(st["qp_sweeps"], np.int32)}
    if kind.startswith("p"):
        Up = ulp_perturbed(wl.U, int(kind[1:]))
        op, r = orc(wl, U=Up), ref(wl, U=Up)
        qp_p, _ = op.qp(O.unary_init(Up), max_iter=QP_IT)
        lab = o.current_map(r.lp(qp_p, O.LPParams(**LP)))
        e = o.energy_true(lab, fp64=True)
        log(f"c1 {kind}: {time.time() - t0:.1f} s, lp energy {e:.6f}")
        return {f"e_lp_{kind}": e, f"lab_lp_{kind}": u8(lab)}
    out = {f"lab_qp_{kind}": u8(lab_qp), f"e_qp_{kind}": o.energy_true(lab_qp, fp64=True)}
    if kind == "r":
        r = ref(wl)
        qlp = r.lp(qqp, O.LPParams(**LP))
        st = None
    else:
        qlp, st = o.lp(qqp, O.LPParams(**LP), fp64=fp64)
    lab = o.current_map(qlp)
    out[f"lab_lp_{kind}"] = u8(lab)
    out[f"e_lp_{kind}"] = o.energy_true(lab, fp64=True)
    if st is not None:
        lp_stats(out, kind, st)
    if kind == "f":
        out["lp_rowsum_f"] = qlp.astype(np.float64).sum(1)
    if kind == "r":
        out["lp_rowsum_r"] = qlp.astype(np.float64).sum(1)
    log(f"c1 {kind}: {time.time() - t0:.1f} s, qp energy {out[f'e_qp_{kind}']:.6f}, lp energy {out[f'e_lp_{kind}']:.6f}")
    return out


# ---------------------------------------------------------------- c2 -------
def c2_layers():
    return [r for r, _ in S.superpixel_layers(C2["W"], C2["H"], cells=C2["cells"], seeds=C2["seeds"])]


def c2_run(kind):
    """kinds as for c1 ('p<k>': the compiled reference's qp_sp + lp_sp with the
    unaries moved one ulp per entry)"""
    wl = S.Workload(C2["W"], C2["H"], C2["K"], C2["seed"], L=L)
    layers = c2_layers()
    o = orc(wl, layers)
    fp64 = kind == "d"
    t0 = time.time()
    q0 = O.unary_init(wl.U)
    if kind == "s":
        r = ref(wl, layers)
        qqp = r.qp(q0)
        qs = r.lp(qqp, O.LPParams(**LP_SHORT))
        _, st = o.lp(qqp, O.LPParams(**LP_SHORT))
        lab = o.current_map(qs)
        log(f"c2 s: {time.time() - t0:.1f} s")
        return {"lab_short_r": u8(lab), "e_short_r": o.energy_ho(lab, fp64=True),
                "gap_short_f": np.asarray(st["gap"], np.float64), "sweeps_short_f": np.asarray(st["qp_sweeps"], np.int32)}
    if kind.startswith("p"):
        Up = ulp_perturbed(wl.U, int(kind[1:]))
        r = ref(wl, layers, U=Up)
        qqp = r.qp(O.unary_init(Up))
        lab = o.current_map(r.lp(qqp, O.LPParams(**LP)))
        e = o.energy_ho(lab, fp64=True)
        log(f"c2 {kind}: {time.time() - t0:.1f} s, lp energy {e:.6f}")
        return {f"e_lp_{kind}": e, f"lab_lp_{kind}": u8(lab)}
    if kind == "r":
        r = ref(wl, layers)
        qqp = r.qp(q0)
        it = -1
    else:
        qqp, sq = o.qp(q0, max_iter=0, ref_stop=True, fp64=fp64)
        it = sq["iters"]
    lab_qp = o.current_map(qqp)
    out = {f"lab_qp_{kind}": u8(lab_qp), f"e_qp_{kind}": o.energy_ho(lab_qp, fp64=True)}
    if it >= 0:
        out[f"qp_iters_{kind}"] = np.int32(it)
    log(f"c2 {kind}: qp {time.time() - t0:.1f} s, {it} iterations, energy {out[f'e_qp_{kind}']:.6f}")
    if kind == "r":
        qlp = r.lp(qqp, O.LPParams(**LP))
        st = None
    else:
        qlp, st = o.lp(qqp, O.LPParams(**LP), fp64=fp64)
    lab = o.current_map(qlp)
    out[f"lab_lp_{kind}"] = u8(lab)
    out[f"e_lp_{kind}"] = o.energy_ho(lab, fp64=True)
    if st is not None:
        lp_stats(out, kind, st)
    log(f"c2 {kind}: total {time.time() - t0:.1f} s, lp energy {out[f'e_lp_{kind}']:.6f}")
    return out


# ---------------------------------------------------------------- c3 -------
def c3_sizes():
    sizes = []
    for i in range(C3_IMAGES):
        rng = np.random.default_rng(3000 + i)
        H = int(rng.integers(240, 1081))
        W = int(rng.integers(320, 1921))
        sizes.append((W, H))
    return sizes


def c3_pick():
    sizes = c3_sizes()
    return sorted(range(C3_IMAGES), key=lambda i: (sizes[i][0] * sizes[i][1], i))[:3]


def c3_run(job):
    """job: image index, or (image index, k) for the float oracle's QP + LP
    (equal to the reference's on these inputs) with the unaries moved one ulp
    per entry"""
    i, pk = (job, 0) if isinstance(job, int) else job
    W, H = c3_sizes()[i]
    wl = S.Workload(W, H, max(1, round(W * H / 4000)), 3000 + i, L=L)
    o = orc(wl)
    t0 = time.time()
    q0 = O.unary_init(wl.U)
    out = {}
    if pk:
        Up = ulp_perturbed(wl.U, pk)
        op = orc(wl, U=Up)
        qqp, _ = op.qp(O.unary_init(Up), max_iter=QP_IT)
        qlp, _ = op.lp(qqp, O.LPParams(**LP))
        e = o.energy_true(o.current_map(qlp), fp64=True)
        log(f"c3 image {i} p{pk}: {time.time() - t0:.1f} s, lp energy {e:.6f}")
        return {f"img{i}_e_lp_p{pk}": e}
    for kind, fp64 in (("f", False), ("d", True)):
        qqp, _ = o.qp(q0, max_iter=QP_IT, fp64=fp64)
        lab_qp = o.current_map(qqp)
        qlp, st = o.lp(qqp, O.LPParams(**LP), fp64=fp64)
        lab = o.current_map(qlp)
        out[f"img{i}_lab_qp_{kind}"] = u8(lab_qp)
        out[f"img{i}_e_qp_{kind}"] = o.energy_true(lab_qp, fp64=True)
        out[f"img{i}_lab_lp_{kind}"] = u8(lab)
        out[f"img{i}_e_lp_{kind}"] = o.energy_true(lab, fp64=True)
        out[f"img{i}_sweeps_{kind}"] = np.asarray(st["qp_sweeps"], np.int32)
        out[f"img{i}_gap_{kind}"] = np.asarray(st["gap"], np.float64)
        if kind == "f":  # the short LP (the float oracle's equals the reference's on these inputs)
            qs, _ = o.lp(qqp, O.LPParams(**LP_SHORT))
            ls = o.current_map(qs)
            out[f"img{i}_lab_short_f"] = u8(ls)
            out[f"img{i}_e_short_f"] = o.energy_true(ls, fp64=True)
    log(f"c3 image {i} ({W}x{H}): {time.time() - t0:.1f} s")
    return out


# ---------------------------------------------------------------- c4 -------
def c4_directions(i, N):
    rng = np.random.default_rng(4200 + i)
    return rng.standard_normal((3, N, L)).astype(np.float32)


def c4_sample(i, N):
    """4096 seeded (pixel, label) entries, as flat indices pixel * L + label"""
    return np.sort(np.random.default_rng(4300 + i).choice(N * L, 4096, replace=False))


def c4_run(i):
    Wd, T = C4["W"], C4["T"]
    wl = S.Workload(Wd, Wd, round(Wd * Wd / 4000), 4000 + i, L=L)
    lab = np.random.default_rng(4100 + i).integers(0, L, wl.N).astype(np.int32)
    o = O.CRF(wl.N, L)
    o.set_unary(wl.U)
    o.add_pairwise_compat(wl.kernels[0][0], "potts", [3.0])
    o.add_pairwise_compat(wl.kernels[1][0], "matrix", -5.0 * np.eye(L, dtype=np.float32))
    t0 = time.time()
    Q = o.meanfield(T, fp64=True)
    p = Q[np.arange(wl.N), lab]
    loss = float(-np.log(p).sum())
    dQ = np.zeros((wl.N, L))
    dQ[np.arange(wl.N), lab] = -1.0 / p
    Q2, dU, grads = o.meanfield_grad(T, dQ)
    dirs = c4_directions(i, wl.N)
    pick = c4_sample(i, wl.N)
    out = {f"img{i}_loss": loss, f"img{i}_gpotts": grads[0], f"img{i}_gmatrix": grads[1],
           f"img{i}_dU_proj": np.array([float((dU * d).sum()) for d in dirs]),
           f"img{i}_dU_sample": dU.reshape(-1)[pick], f"img{i}_Q_sample": Q.reshape(-1)[pick],
           f"img{i}_Q_proj": np.array([float((Q * d).sum()) for d in dirs]),
           f"img{i}_lab": u8(Q.argmax(1))}
    log(f"c4 image {i}: {time.time() - t0:.1f} s, loss {loss:.6f}")
    return out


def _call(job):
    fn, arg = job
    return globals()[fn](arg)


def main(which):
    jobs = []
    spread = ["p1", "p2", "p3", "p4", "s"] if O.ref_available() else []
    if "c2" in which:  # longest jobs first
        jobs += [("c2_run", k) for k in ["f", "d"] + (["r"] if O.ref_available() else []) + spread]
    if "c1" in which:
        jobs += [("c1_run", k) for k in ["f", "d"] + (["r"] if O.ref_available() else []) + spread]
    if "c3" in which:
        jobs += [("c3_run", i) for i in c3_pick()] + [("c3_run", (i, k)) for i in c3_pick() for k in (1, 2, 3)]
    if "c4" in which:
        jobs += [("c4_run", i) for i in range(C4["B"])]
    only = os.environ.get("MAKE_CONFIGS_ONLY")
    if only:
        jobs = [j for j in jobs if j[1] == only]
    workers = int(os.environ.get("MAKE_CONFIGS_WORKERS", "6"))
    with mp.get_context("spawn").Pool(min(workers, len(jobs))) as pool:
        res = pool.map(_call, jobs, chunksize=1)
    merged = {}
    for (fn, _), r in zip(jobs, res):
        merged.setdefault(fn[:2], {}).update(r)
    names = {"c1": "cfg_c1_640x480.npz", "c2": "cfg_c2_1280x720_sp.npz", "c3": "cfg_c3_batch_small3.npz",
             "c4": "cfg_c4_meanfield_8x512.npz"}
    for cfg, d in merged.items():
        path = os.path.join(OUT, names[cfg])
        if os.environ.get("MAKE_CONFIGS_MERGE") == "1" and os.path.exists(path):
            old = dict(np.load(path, allow_pickle=False))
            old.update(d)
            d = old
        if "lp_rowsum_f" in d and "lp_rowsum_r" in d:
            # the float oracle's LP against the compiled reference's, row by row
            d["lp_rowsum_equal_f_r"] = np.bool_(np.array_equal(d["lp_rowsum_f"], d["lp_rowsum_r"]))
        d.pop("lp_rowsum_f", None)
        d.pop("lp_rowsum_r", None)
        if cfg == "c3":
            d["images"] = np.asarray(c3_pick(), np.int32)
        np.savez_compressed(os.path.join(OUT, names[cfg]), **d)
        log(f"wrote {names[cfg]} ({os.path.getsize(os.path.join(OUT, names[cfg]))} bytes)")


if __name__ == "__main__":
    main(sys.argv[1:] or ["c1", "c2", "c3", "c4"])
