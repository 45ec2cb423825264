"""Write the golden vectors: the reference's own trajectories on tiny grids.

Runs ``tse.py`` from the reference checkout (read-only, at
``/root/reference``) on each ``configs/g*.py`` with a pure-Python ``numba``
stand-in on ``sys.path`` (``transportse_amd/numba_stub``), so the reference's
own Python source executes unchanged: ``njit`` is the identity and
``prange`` is ``range``. One difference from plain Python is removed
deliberately: Numba lowers ``numpy.sum`` to a sequential C-order loop
(``for v in np.nditer(arr): c += v``), while numpy uses pairwise summation.
A ``sitecustomize`` module replaces ``numpy.sum`` with that sequential loop for
calls from tse.py itself, so ``l2_norm`` (tse.py:273-281) reproduces Numba's
rounding while the config's own Python callbacks keep numpy's.

What these vectors pin that closed forms cannot: the exact iteration counts,
norm schedule, accept/grow/halve/reset control of compute_P / compute_C /
compute_T (tse.py:699-742, 877-916, 1045-1081) and the SIA driver order
(tse.py:1181-1494), bit for bit.

Output snapshots: with PLOTSAVE on, the reference writes its 34-tuple
(tse.py:312-329) with ``joblib.dump``.  The same ``sitecustomize`` replaces
``joblib.dump`` with a writer of the tuple's elements as numpy arrays
(``numpy.savez``), so no pickle is ever read back; the config runs from a
temporary copy, so those files land there and not next to the committed
config.

Outputs: ``tests/golden/<case>.npz`` (numpy arrays only) with one leading
step axis per field, plus ``plot<index>_<element>`` arrays for each output
snapshot; nothing from the reference is stored but the numbers it produced.
Usage: ``python tests/golden/make_golden.py [case ...]``.
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
         "g6_xz", "g7_column", "g8_xy", "g9_plot"]

SITECUSTOMIZE = '''
import sys
import numpy
_orig_sum = numpy.sum
def _numba_sum(a, *args, **kw):
    # only the reference's own (jitted) code gets Numba's sequential sum; a
    # config's plain-Python callbacks (mass_balance) keep numpy's pairwise sum
    if args or kw or not sys._getframe(1).f_code.co_filename.endswith("tse.py"):
        return _orig_sum(a, *args, **kw)
    arr = numpy.asarray(a)
    if arr.dtype.kind != "f":
        return _orig_sum(arr)
    c = 0.0
    for v in numpy.ascontiguousarray(arr, dtype=numpy.float64).ravel().tolist():
        c += v
    return numpy.float64(c)
numpy.sum = _numba_sum

import joblib
def _arrays_dump(value, filename, compress=0, **kw):
    # the output tuple's elements as numpy arrays instead of a pickle
    items = {"e%02d" % n: numpy.asarray(v) for n, v in enumerate(value)}
    if hasattr(filename, "write"):
        numpy.savez(filename, **items)
    else:
        with open(filename, "wb") as f:
            numpy.savez(f, **items)
    return [filename]
joblib.dump = _arrays_dump
'''


def run_case(case: str) -> str:
    import shutil
    with tempfile.TemporaryDirectory() as tmp:
        dump = os.path.join(tmp, "dump")
        site = os.path.join(tmp, "site")
        cfgdir = os.path.join(tmp, "cfg")
        os.makedirs(dump)
        os.makedirs(site)
        os.makedirs(cfgdir)
        for name in (case + ".py", "tiny_base.py"):
            shutil.copy(os.path.join(HERE, "configs", name), cfgdir)
        cfg = os.path.join(cfgdir, case + ".py")
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
        # output snapshots, in PLOTINDEX order (numpy.savez files)
        plots = sorted(glob.glob(os.path.join(cfgdir, case + "-[0-9]*.p")))
        for n, fn in enumerate(plots):
            with np.load(fn, allow_pickle=False) as z:
                for key in z.files:
                    out[f"plot{n}_{key}"] = z[key]
        out["plot_count"] = np.array(len(plots))
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
