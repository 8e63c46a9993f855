"""CPU timings of the compiled reference (oracle/_ref, built from the
reference's own sources) and of the oracle on the small configurations, one
thread each: config 0 (QP at 320x213, 40 cells) and config 1 (QP then LP
5x10 at 640x480, 120 cells).  Test infrastructure: run in the development
container, where the reference sources exist.

    python oracle/time_configs.py            # prints one JSON line
"""

from __future__ import annotations

import json
import os
import platform
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import oracle as O  # noqa: E402
from densecrf_ho_amd import synthetic as S  # noqa: E402

L = 21


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def timed(fn):
    t = time.perf_counter()
    r = fn()
    return r, time.perf_counter() - t


def config(W, H, K, seed, lp):
    wl = S.Workload(W, H, K, seed, L=L)
    ref = O.RefCRF(W, H, L)
    ref.set_unary(wl.U)
    ref.add_gaussian(3, 3, 3)
    ref.add_bilateral(50, 50, 15, 15, 15, wl.img, 5)
    orc = O.CRF(wl.N, L)
    orc.set_unary(wl.U)
    for f, w in wl.kernels:
        orc.add_pairwise(f, w)
    Q0 = ref.unary_init()
    out = {}
    qr, out["reference_qp_s"] = timed(lambda: ref.qp(Q0))
    (qo, _), out["oracle_qp_s"] = timed(lambda: orc.qp(Q0, ref_stop=True))
    if lp:
        P = O.LPParams(prox_max_iter=5, fw_max_iter=10)
        _, out["reference_lp_s"] = timed(lambda: ref.lp(qr, P))
        _, out["oracle_lp_s"] = timed(lambda: orc.lp(qo, P))
    return {k: round(v, 3) for k, v in out.items()}


def main():
    res = {"cpu_model": cpu_model(), "threads": 1,
           "c0_qp_320x213": config(320, 213, 40, 0, lp=False),
           "c1_qp_lp_640x480": config(640, 480, 120, 1, lp=True)}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
