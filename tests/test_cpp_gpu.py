"""The C++ interface (include/densecrf_ho_amd.hpp) driven like the reference's
example_inference: QP then LP on a small synthetic image, compared with the
oracle, plus the reference's "Value out of bounds" exception."""

import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_cpp_example_matches_oracle(tmp_path):
    from densecrf_ho_amd import synthetic as S
    exe = os.path.join(ROOT, "tests", "cpp", "api_example")
    assert os.path.exists(exe), "built by __graft_entry__.build()"
    W, H, L = 64, 48, 21
    wl = S.Workload(W, H, 6, 3)
    wl.U.tofile(tmp_path / "u.f32")
    wl.img.tofile(tmp_path / "im.u8")
    r = subprocess.run([exe, str(W), str(H), str(L), str(tmp_path / "u.f32"), str(tmp_path / "im.u8"),
                        str(tmp_path / "qp.f32"), str(tmp_path / "lp.f32")], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    vals = dict(line.split(" ", 1) for line in r.stdout.strip().splitlines())
    assert vals["exception"].strip() == "Value out of bounds"
    Qqp = np.fromfile(tmp_path / "qp.f32", np.float32).reshape(wl.N, L)
    orc = O.CRF(wl.N, L)
    orc.set_unary(wl.U)
    for f, w in wl.kernels:
        orc.add_pairwise(f, w)
    Qo, _ = orc.qp(O.unary_init(wl.U), max_iter=0, ref_stop=True)
    assert (Qqp.argmax(1) == Qo.argmax(1)).mean() >= 0.999
    e = float(vals["qp_energy"])
    eo = orc.energy_true(orc.current_map(Qo), fp64=True)
    assert abs(e - eo) <= 1e-4 * abs(eo)
    Qlp = np.fromfile(tmp_path / "lp.f32", np.float32).reshape(wl.N, L)
    Ql, _ = orc.lp(Qo, O.LPParams(prox_max_iter=2))
    assert (Qlp.argmax(1) == Ql.argmax(1)).mean() >= 0.999
    # adjoint through the C++ layer: Potts (1) + Matrix (L(L+1)/2) parameters
    assert int(vals["meanfield_params"]) == 1 + L * (L + 1) // 2
    assert np.isfinite(float(vals["meanfield_dU00"]))
    # a batch of the same image twice equals the one-image QP
    assert float(vals["batch_qp_diff"]) < 1e-5
    # the other filters and gradients through the C++ layer
    assert float(vals["restricted_vs_ordered"]) < 1e-5
    assert np.isfinite(float(vals["bf_mid"])) and np.isfinite(float(vals["dc_mid"]))
    assert int(vals["kernel_gradient_size"]) == 2
    assert np.isfinite(float(vals["lattice_gradient_abs"])) and float(vals["lattice_gradient_abs"]) > 0
    # UnaryEnergy objects: LogisticUnaryEnergy through setUnaryEnergy(UnaryEnergy*)
    assert float(vals["logistic_get_max"]) > 0
    assert int(vals["logistic_params"]) == L * 6 and int(vals["logistic_gradient_size"]) == L * 6
    assert float(vals["logistic_vs_const_qp_diff"]) == 0.0
    assert int(vals["logistic_set_parameters_changes_unary"]) == 1
    # addPairwiseEnergy(PairwisePotential*) with a Matrix compatibility
    assert float(vals["potential_object_vs_features_diff"]) == 0.0
    assert int(vals["potential_object_params"]) == L * (L + 1) // 2
