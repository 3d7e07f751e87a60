"""Process exit with library objects of every kind still alive: the ordered
release at interpreter exit (captured batches, servers, gradient runners,
texture brushes, brushes, canvases, layers, contexts) ends the process with
status 0, with no finalizer running against a torn-down runtime."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r'''
import numpy as np
from painty_amd import hip as H, serve as SV, workload as W

wl = W.small(64, 72, 20, 6, 3, pickup=True)
other = H.Context(0)
keep = []
for ctx in (None, other):
    cv = H.Canvas(wl.rows, wl.cols, ctx)
    cv.getPaintLayer().setBuffers(wl.wet.K, wl.wet.S, wl.wet.V)
    cv.setR0(wl.R0)
    br = H.FootprintBrush(wl.radii[0], wl.cores[0], ctx)
    out = H.Canvas(wl.rows, wl.cols, ctx)
    keep.append((cv, br, out, H.PaintLayer(8, 8, ctx)))
jobs = [(cv, br, wl.stamps, out) for cv, br, out, _ in keep]
H.batch_set_config(16, 1, 1, 4)  # waves (the persistent launch is covered elsewhere)
H.batch_run(jobs)
g = H.BatchGraph(jobs)          # a captured batch, never closed
g.launch()
H.batch_submit(jobs)            # still in flight at exit
srv = SV.Server(n_slots=2, hbm_budget_bytes=1 << 28)
t = srv.submit(1, 64, 64, wl.stamps, wl.radii[0], wl.cores[0])  # never waited for
tb = H.TextureBrush(np.ones((8, 16)))
gl = H.TextureBrushGpu()
del other                        # the objects made on it outlive the caller's reference
print("__EXIT_READY__", flush=True)
'''


def test_exit_with_live_objects():
    r = subprocess.run([sys.executable, "-c", PROGRAM], cwd=ROOT, capture_output=True,
                       text=True, timeout=120)
    assert "__EXIT_READY__" in r.stdout, r.stdout + r.stderr
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
