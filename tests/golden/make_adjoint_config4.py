"""Writes tests/golden/adjoint_config4.npz: the FP64 oracle's discrete RK4 adjoint of the
differentiation workload (bench.adjoint_problem: 128x128x4 film of 4 nm cells, Gaussian-blob
Ms map in [5e5, 1.1e6] A/m, A 1.3e-11 J/m, alpha 0.05, 1 ns sinc pulse of 20 mT in a seeded
direction, smooth seeded m0 and m_target), 4000 steps of 250 fs with checkpoints every 100.

The oracle runs on exactly the inputs the GPU path receives (the FP32 Ms map, m0 and m_target
rounded to FP32).  Stored: the loss, dL/dMs per cell (float64) and dL/d amp.  About 25 minutes
on 8 host cores:  python tests/golden/make_adjoint_config4.py
"""
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import bench  # noqa: E402
import oracle_lib as O  # noqa: E402

STEPS, DT, CKPT = 4000, 250e-15, 100


def main():
    n, h, p, Ms, m0, mt = bench.adjoint_problem()
    Ms64 = np.asarray(Ms, np.float32).astype(np.float64)
    m064 = np.asarray(m0, np.float32).astype(np.float64)
    mt64 = np.asarray(mt, np.float32).astype(np.float64)
    t0 = time.time()
    a = O.OracleAdjoint(n, h, p, Ms64, DT, STEPS, Nz_override=8, ckpt=CKPT)
    L, Mf = a.forward(m064, mt64)
    print(f"forward {time.time() - t0:.0f} s, L = {L:.12f}", flush=True)
    g, ga = a.backward(m064, mt64)
    print(f"backward done at {time.time() - t0:.0f} s, dL/damp = {ga:.9e}", flush=True)
    out = Path(__file__).resolve().parent / "adjoint_config4.npz"
    np.savez_compressed(out, loss=np.float64(L), gMs=np.asarray(g, np.float64), gamp=np.float64(ga),
                        m_avg=np.asarray(Mf, np.float64).reshape(-1, 3).mean(axis=0) / Ms64.mean(),
                        steps=STEPS, dt=DT, ckpt=CKPT)
    print(f"wrote {out}")


if __name__ == "__main__":
    main()
