"""Runs the CPU execution of the device BVH build (hipcpu_build.py) on a mesh and reads its result."""
import os
import subprocess
import tempfile
import time

import numpy as np

from photon_amd.types import NODE_DTYPE


def run(exe, positions, indices, threads=8, timeout=None):
    """Returns (nodes, order, stats bytes, wall seconds) of the build of `positions` (V, 3) float32 and
    `indices` (T, 3) int32."""
    p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    i = np.ascontiguousarray(indices, np.int32).reshape(-1, 3)
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.array([len(p), len(i)], np.int32).tobytes())
            f.write(p.tobytes())
            f.write(i.tobytes())
        env = dict(os.environ, HIPCPU_THREADS=str(threads),
                   ASAN_OPTIONS="detect_leaks=0:detect_stack_use_after_return=0:abort_on_error=1",
                   UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
        t = time.perf_counter()
        r = subprocess.run([exe, fin, fout], env=env, capture_output=True, text=True, timeout=timeout)
        wall = time.perf_counter() - t
        if r.returncode != 0:
            raise RuntimeError(f"CPU execution of the device BVH build failed ({r.returncode}):\n{r.stderr[-4000:]}")
        data = open(fout, "rb").read()
    count = int(np.frombuffer(data[:8], np.int64)[0])
    off = 8
    nodes = np.frombuffer(data[off:off + count * NODE_DTYPE.itemsize], NODE_DTYPE)
    off += count * NODE_DTYPE.itemsize
    order = np.frombuffer(data[off:off + 4 * len(i)], np.int32)
    off += 4 * len(i)
    return nodes, order, data[off:], wall
