"""Write the golden vectors: the reference's own trajectories on tiny grids.

Runs ``tse.py`` from the reference checkout (read-only, at
``/root/reference``) on each ``configs/g*.py`` with a pure-Python ``numba``
stand-in on ``sys.path`` (``transportse_amd/numba_stub``), so the reference's
own Python source executes unchanged: ``njit`` is the identity and
``prange`` is ``range``. One difference from plain Python is removed
deliberately: Numba lowers ``numpy.sum`` to a sequential C-order loop
(``for v in np.nditer(arr): c += v``), while numpy uses pairwise summation.
A ``sitecustomize`` module replaces ``numpy.sum`` with that sequential loop so
``l2_norm`` (tse.py:273-281) reproduces Numba's rounding.

What these vectors pin that closed forms cannot: the exact iteration counts,
norm schedule, accept/grow/halve/reset control of compute_P / compute_C /
compute_T (tse.py:699-742, 877-916, 1045-1081) and the SIA driver order
(tse.py:1181-1494), bit for bit.

Outputs: ``tests/golden/<case>.npz`` (numpy arrays only) with one leading
step axis per field; nothing from the reference is stored but the numbers it
produced. Usage: ``python tests/golden/make_golden.py [case ...]``.
"""
import glob
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REFERENCE = "/root/reference/tse.py"
STUB = os.path.join(REPO, "transportse_amd", "numba_stub")
CASES = ["g1_salt", "g2_sia_heat", "g3_halving", "g4_smolar", "g5_nlte",
         "g6_xz", "g7_column", "g8_xy"]

SITECUSTOMIZE = '''
import numpy
_orig_sum = numpy.sum
def _numba_sum(a, *args, **kw):
    if args or kw:
        return _orig_sum(a, *args, **kw)
    arr = numpy.asarray(a)
    if arr.dtype.kind != "f":
        return _orig_sum(arr)
    c = 0.0
    for v in numpy.ascontiguousarray(arr, dtype=numpy.float64).ravel().tolist():
        c += v
    return numpy.float64(c)
numpy.sum = _numba_sum
'''


def run_case(case: str) -> str:
    cfg = os.path.join(HERE, "configs", case + ".py")
    with tempfile.TemporaryDirectory() as tmp:
        dump = os.path.join(tmp, "dump")
        site = os.path.join(tmp, "site")
        os.makedirs(dump)
        os.makedirs(site)
        with open(os.path.join(site, "sitecustomize.py"), "w") as f:
            f.write(SITECUSTOMIZE)
        env = dict(os.environ)
        env["PYTHONPATH"] = os.pathsep.join([site, STUB])
        env["TSE_GOLDEN_DUMP"] = dump
        proc = subprocess.run([sys.executable, "-u", REFERENCE, cfg], cwd=tmp,
                              env=env, capture_output=True, text=True,
                              timeout=1800)
        log = proc.stdout + proc.stderr
        if proc.returncode != 0:
            print(log[-4000:])
            raise RuntimeError(f"reference run of {case} failed")
        files = sorted(glob.glob(os.path.join(dump, "step*.npz")))
        if not files:
            raise RuntimeError("no steps dumped")
        recs = [dict(np.load(f)) for f in files]
        out = {}
        for key in recs[0]:
            out[key] = np.stack([np.asarray(r[key]) for r in recs])
        path = os.path.join(HERE, case + ".npz")
        np.savez_compressed(path, **out)
        with open(os.path.join(HERE, case + ".log.txt"), "w") as f:
            # strip ANSI colour codes and the run-time line (non-deterministic)
            import re
            clean = re.sub(r"\x1b\[[0-9;]*m", "", log)
            clean = "\n".join(l for l in clean.splitlines()
                              if not l.startswith("[Run time]"))
            f.write(clean + "\n")
        return path


def main(argv):
    cases = argv[1:] or CASES
    for c in cases:
        p = run_case(c)
        z = np.load(p)
        print(c, "->", os.path.relpath(p, REPO), "steps:", z["P"].shape[0],
              "numitP:", z["numitP"].tolist(), "numitC:", z["numitC"].tolist(),
              "numitT:", z["numitT"].tolist())


if __name__ == "__main__":
    main(sys.argv)
