"""Writes tests/golden/mlem_panel128_rd.npz: the FP64 oracle's MLEM at the full benchmark size, for the GPU
reference-match test of that configuration.

Configuration: two planograms (psi 0 and 90 degrees) of 128 x 128 crystals at 2 mm, D = 200 mm, shape
(101, 101, 128, 228) each; a 128^3 volume at 2 mm; the seed-2 20-ellipsoid phantom; data = Poisson (numpy
PCG64, seed 102) of its fp32 oracle forward projection scaled to 1e8 counts; 20 MLEM iterations from ones
with the relative-difference prior (beta 0.01, delta 2), PETL-mode back projection, in FP64.

The data are regenerated from the same seeds by the test (the fp32 oracle forward is deterministic: every
bin is summed by one thread in a fixed order), and the fixture keeps the per-v1-row sums of the data so
that the test can confirm it reconstructs from the same counts.

Stored: per-iteration FP64 log-likelihood of the image each iteration starts from, plus that of the final
image; per-iteration L2 norm of the image; the final image and the image after iteration 1 on the
every-fourth-voxel grid f[::4, ::4, ::4] (32^3); the data's per-row sums.  Runtime: about an hour on 8
host cores.

Usage: python tests/golden/make_mlem_panel128_golden.py [--state DIR] [out.npz]
With --state, the data, the sensitivity and every iterate are saved in DIR as they are formed, and a run
started again with the same DIR resumes after the last saved iteration (the iterations run one at a time
through oracle.mlem with the saved sensitivity, which is the same arithmetic as one 20-iteration call).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from petl_amd.workloads import ellipsoid_phantom, poisson_data  # noqa: E402

ITERS, BETA, DELTA, COUNTS = 20, 0.01, 2.0, 1.0e8


def geometry():
    plans = [O.planogram(psi, 200.0, 256.0, 256.0, 2.0) for psi in (0.0, 90.0)]
    return plans, O.volume(128, 128, 128, 2.0)


def data(nthreads=None):
    """(plans, vol, phantom, y): the configuration's inputs, from the seeds."""
    plans, vol = geometry()
    f = ellipsoid_phantom(128, 2.0, 20, seed=2)
    lam = O.project(plans, vol, f, nthreads=nthreads)
    y = poisson_data(lam, COUNTS, seed=102)
    return plans, vol, f, y


def row_sums(y):
    return np.stack([yi.reshape(yi.shape[0], -1).sum(axis=1, dtype=np.float64) for yi in y])


def main(out, state=None):
    t0 = time.time()
    st = None
    if state:
        os.makedirs(state, exist_ok=True)
        st = os.path.join(state, "progress.json")
    plans, vol = geometry()
    ypaths = [os.path.join(state, f"y{i}.npy") for i in range(2)] if state else None
    if state and all(os.path.exists(p) for p in ypaths):
        y = [np.load(p) for p in ypaths]
    else:
        _, _, _, y = data()
        if state:
            for p, yi in zip(ypaths, y):
                np.save(p, yi)
    print(f"data: {time.time() - t0:.1f} s, {sum(float(yi.sum(dtype=np.float64)) for yi in y):.6e} counts",
          flush=True)
    spath = os.path.join(state, "sens.npy") if state else None
    if state and os.path.exists(spath):
        S = np.load(spath)
    else:
        S = O.sensitivity_image(plans, vol, "petl", np.float64)
        if state:
            np.save(spath, S)
    print(f"sensitivity: {time.time() - t0:.0f} s", flush=True)
    prog = dict(ll=[], norms=[])
    f = np.ones(vol.shape, np.float64)
    first = None
    if st and os.path.exists(st):
        prog = json.load(open(st))
        done = len(prog["norms"])
        if done:
            f = np.load(os.path.join(state, f"f{done}.npy"))
            first = np.load(os.path.join(state, "f1.npy"))[::4, ::4, ::4]
    for n in range(len(prog["norms"]), ITERS):
        ll = []
        f = O.mlem(plans, vol, y, 1, f=f, delta=DELTA, beta=BETA, prior="rd", dtype=np.float64, mode="petl",
                   loglik=ll, sens=S)
        prog["ll"].append(ll[0])
        prog["norms"].append(float(np.linalg.norm(f.ravel())))
        if n == 0:
            first = f[::4, ::4, ::4].copy()
        if state:
            np.save(os.path.join(state, f"f{n + 1}.npy"), f)
            if n >= 2:  # keep f1 and the latest iterate
                old = os.path.join(state, f"f{n}.npy")
                if os.path.exists(old):
                    os.remove(old)
            json.dump(prog, open(st, "w"))
        print(f"iteration {n + 1}: {time.time() - t0:.0f} s, |f| {prog['norms'][-1]:.10e}, LL {prog['ll'][-1]:.12e}",
              flush=True)
    ll_final = sum(O._loglik_terms(O.project([p], vol, f, dtype=np.float64)[0], yi) for p, yi in zip(plans, y))
    np.savez_compressed(out, loglik=np.array(prog["ll"] + [ll_final]), image_norm=np.array(prog["norms"]),
                        image_sub=f[::4, ::4, ::4], image_sub_iter1=first, y_row_sums=row_sums(y),
                        config=np.array([ITERS, BETA, DELTA, COUNTS]))
    print(f"wrote {out} after {time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("out", nargs="?", default=os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                           "mlem_panel128_rd.npz"))
    ap.add_argument("--state", default=None)
    a = ap.parse_args()
    main(a.out, a.state)
