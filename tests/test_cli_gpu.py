"""The reference's command line (example_inference) on the reference's own
example input (data/img.bmp, data/img.c_unary, kept in tests/golden/data/)
with its default parameters: each method's discretised energy against the
compiled reference's, and the saved MSRC map's labels equal to the
reference's on >= 99.9% of the pixels.  The superpixel methods take the
reference's own mean-shift segmentation as a region map.

Energy bars: 1e-4 relative for unary, qp and qp_sp, with >= 99.9% of the
labels equal.  The LP methods run ten proximal steps from the unaries on a
real image, and their final energy follows float rounding: the reference's
own LP, started from fields one ulp from its unary_init, ends at different
energies (data_img.json records the spread of those runs and their lowest
label agreement with the unperturbed run).  The LP methods are held to that
spread, widened by 1e-3, and to that agreement less 0.5%."""

import json
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from test_io import MSRC, _read_png

pytestmark = pytest.mark.gpu

DATA = os.path.join(GOLDEN, "data")


@pytest.mark.parametrize("method", ["unary", "qp", "lp", "qp_sp", "lp_sp"])
def test_example_inference_on_reference_data(tmp_path, method):
    exe = os.path.join(ROOT, "densecrf_ho_amd", "example_inference")
    assert os.path.exists(exe), "built by __graft_entry__.build()"
    with open(os.path.join(GOLDEN, "data_img.json")) as f:
        ref = json.load(f)
    g = np.load(os.path.join(GOLDEN, "data_img_labels.npz"))
    out = tmp_path / "seg.png"
    cmd = [exe, method, "--image", os.path.join(DATA, "img.bmp"), "--unary", os.path.join(DATA, "img.c_unary"),
           "--output", str(out)]
    if method.endswith("_sp"):
        g["regions"].astype("<i4").tofile(tmp_path / "regions.i32")
        cmd += ["--regions", str(tmp_path / "regions.i32")]
    print("running", method, flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    line = [x for x in r.stdout.splitlines() if x.startswith("# method:")][-1]
    e = float(line.split("\t")[1])
    er = ref[f"e_{method}"]
    print("energy", e, "reference", er, flush=True)
    lp = method.startswith("lp")
    if lp:
        lo, hi = ref[f"e_{method}_spread"]
        print("reference spread", lo, hi, flush=True)
        assert lo * (1 - 1e-3) <= e <= hi * (1 + 1e-3)
    else:
        assert abs(e - er) <= 1e-4 * abs(er)
    img = _read_png(out)
    W, H = ref["W"], ref["H"]
    want = MSRC[g[method]].reshape(H, W, 3)
    agree = float((img == want).all(-1).mean())
    print("label agreement", agree, flush=True)
    assert agree >= (ref[f"agree_{method}_min"] - 0.005 if lp else 0.999)


def _reference_data_lp_case():
    """the reference's data/img.* with its example_inference pairwise terms,
    its mean-shift layer, and its own unary_init rebuilt bit for bit (the
    oracle's softmax with the entries where Eigen's exp rounds differently
    patched from the fixture)"""
    import densecrf_ho_amd as D
    import oracle as O
    g = np.load(os.path.join(GOLDEN, "data_img_lp_steps.npz"), allow_pickle=False)
    labs = np.load(os.path.join(GOLDEN, "data_img_labels.npz"), allow_pickle=False)
    U, W, H = D.load_unary(os.path.join(DATA, "img.c_unary"))
    im, _, _ = D.load_image(os.path.join(DATA, "img.bmp"))
    Q0 = O.unary_init(U).reshape(-1)
    Q0[g["q0_idx"]] = g["q0_val"]
    return g, labs, U, W, H, im, Q0.reshape(U.shape)


@pytest.mark.parametrize("method,prox,fw", [("lp", 1, 10), ("lp", 2, 10), ("lp_sp", 1, 1), ("lp_sp", 1, 2)])
def test_lp_first_steps_on_reference_data(method, prox, fw):
    """The LP methods on the reference's own input from the reference's own
    initial field, cut short before the rounding spread of the full runs sets
    in (lp: one and two proximal steps; lp_sp: the first one and two
    Frank-Wolfe steps, since from the third on computeUCondGrad's choice of the
    last value above and below zero per region follows values within rounding
    of zero): the discretised energy within 1e-4 of the compiled reference's
    and >= 99.9% of the labels equal."""
    import densecrf_ho_amd as D
    g, labs, U, W, H, im, Q0 = _reference_data_lp_case()
    print("case", method, prox, fw, flush=True)
    c = D.DenseCRF2D(W, H, U.shape[1])
    c.setUnaryEnergy(U)
    c.addPairwiseGaussian(2, 2, D.PottsCompatibility(3))
    c.addPairwiseBilateral(30, 30, 8, 8, 8, im, D.PottsCompatibility(2))
    if method == "lp_sp":
        c.addSuperPixelLayer(labs["regions"], labs["exp_of_superpixels"], labs["constants"])
    Q = c.lp_inference_prox_super_pixels(Q0, D.LP_inf_params(prox_max_iter=prox, fw_max_iter=fw))
    lab = np.ascontiguousarray(Q.argmax(1).astype(np.int32))
    e, er = c.assignment_energy_true(lab), float(g[f"e_{method}_{prox}x{fw}"])
    agree = float((lab == g[f"lab_{method}_{prox}x{fw}"]).mean())
    print("energy", e, "reference", er, "label agreement", agree, flush=True)
    assert abs(e - er) <= 1e-4 * abs(er)
    assert agree >= 0.999
    c.close()
