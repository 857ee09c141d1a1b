"""Generate grid257_tol200.npz: the oracle's 200-step trajectory of the
257 x 257 cloth in tolerance mode (FP64, 1e-6 relative residual).

The workload: 257 x 257 vertices, 0.5 m, two top corners pinned, seeded
+-1 mm jitter (seed 2), dt 0.005 s, 200 steps.  The oracle (oracle/, the C
restatement of the reference's plain method, bit-exact to the compiled
reference in reference mode) runs it twice: from x0, and from x0 with every
z scaled by (1 + 1e-15) (about 1e-18 m).  Solves stopped at a 1e-6 residual
follow different iterates once their inputs differ at the rounding level, so
the second run measures how far two correct FP64 trajectories drift apart;
the GPU's trajectory is held to a multiple of that drift.

The fixture keeps a strided vertex subset (every 17th vertex) of the
positions at steps 1, 10, 20, 50, 100, 150 and 200, the full-mesh norms and
coordinate sums, the iteration counts, and the drift between the two runs at
each kept step (max |x_a - x_b| / max |x_a - x0|).  Takes about an hour:

    make -C oracle
    python tests/golden/make_oracle_trajectory.py
"""
import concurrent.futures as cf
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle_ffi as of  # noqa: E402
from acousticsim_amd import workloads as wl  # noqa: E402

STEPS = 200
KEEP = (1, 10, 20, 50, 100, 150, 200)
STRIDE = 17
PERTURB = 1e-15


def trajectory(x0, conf, tag):
    oc = of.OracleCloth(conf)
    oc.set_state(x0, x0)
    kept, iters = {}, []
    t0 = time.time()
    for k in range(1, STEPS + 1):
        st = oc.step(of.SOLVE_TOLERANCE, 1e-6)
        iters.append(st["iterations"])
        if k in KEEP:
            kept[k] = oc.pos().copy()
        print("%s step %d: %d iterations, residual %.3e, %.0f s" % (
            tag, k, st["iterations"], st["true_rel_residual"], time.time() - t0), flush=True)
    return kept, np.array(iters, dtype=np.int32)


def main():
    conf = wl.grid_config(257, dt=wl.SINGLE_LARGE["dt"])
    x0 = wl.jittered_positions(conf, wl.SINGLE_LARGE["seed"])
    x0b = x0.copy()
    x0b[2::3] *= 1.0 + PERTURB
    with cf.ThreadPoolExecutor(2) as ex:
        fa = ex.submit(trajectory, x0, conf, "a")
        fb = ex.submit(trajectory, x0b, conf, "b")
        (ka, ia), (kb, _) = fa.result(), fb.result()
    nv = len(x0) // 3
    sel = np.arange(0, nv, STRIDE)
    out = {"config_json": np.array(__import__("json").dumps(conf)),
           "seed": np.array(wl.SINGLE_LARGE["seed"]), "steps": np.array(KEEP, dtype=np.int32),
           "vertex_subset": sel.astype(np.int32),
           "x0_subset": x0.reshape(-1, 3)[sel].reshape(-1), "iterations": ia,
           "tolerance": np.array(1e-6), "perturbation": np.array(PERTURB)}
    for k in KEEP:
        xa, xb = ka[k], kb[k]
        out["x_step%d" % k] = xa.reshape(-1, 3)[sel].reshape(-1)
        out["norm_step%d" % k] = np.array(np.linalg.norm(xa))
        out["sum_step%d" % k] = xa.reshape(-1, 3).sum(axis=0)
        out["chaos_step%d" % k] = np.array(np.abs(xa - xb).max() / np.abs(xa - x0).max())
        print("step %d: drift %.3e of the displacement" % (k, float(out["chaos_step%d" % k])))
    path = os.path.join(HERE, "grid257_tol200.npz")
    np.savez_compressed(path, **out)
    print("wrote %s (%d KB)" % (path, os.path.getsize(path) // 1024))


if __name__ == "__main__":
    main()
