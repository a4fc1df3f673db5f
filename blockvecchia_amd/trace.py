"""Launch occupancy of the factorization from a wavefront trace.

A library built with `python -m blockvecchia_amd.build --trace`
(lib/libbv_trace.so) records, for every wavefront of the fused, stepped,
lookahead and paired factorization kernels, its start and end on the device's
100 MHz real-time counter (bv_trace_read in include/bv.h).  From one
evaluation's records this module derives, per launch (one panel step on one
stream): its window, the gap since the previous launch ended, the mean number
of resident wavefronts against the kernel's capacity, and the drain phase
after its last wavefront started, where the launch holds fewer wavefronts
than the GPU has slots for.  Their sum over the evaluation is what a single
persistent launch could at best recover.

    python -m blockvecchia_amd.trace --workload uneven   (n = 200,000 clustered, cs = 400)
    python -m blockvecchia_amd.trace --workload bench    (n = 20,000 uniform, cs = 300)

The factorization form follows the autotuner unless BV_PATH / BV_CHUNKS /
BV_CHUNK_STREAMS fix it; with several chunks on several streams the launches
of one step overlap and are reported together.
"""
import argparse
import ctypes
import os
import sys

import numpy as np

import blockvecchia_amd as bv

TRACE_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libbv_trace.so")
KERNELS = {0: "fused", 1: "stepped", 2: "lookahead", 3: "paired"}
# resident wavefronts per SIMD of each kernel (its __launch_bounds__ occupancy)
WAVES_PER_SIMD = {0: 2, 1: 3, 2: 3, 3: 2}
CAPACITY = 1 << 22


def use_trace_library():
    """Load lib/libbv_trace.so as the package's library (raises OSError when
    it is missing)."""
    if not os.path.exists(TRACE_LIB):
        raise OSError("lib/libbv_trace.so not built: run `python -m blockvecchia_amd.build --trace`")
    lib = bv.load_library(TRACE_LIB)
    lib.bv_trace_reset.argtypes = [ctypes.c_int]
    lib.bv_trace_reset.restype = ctypes.c_int
    lib.bv_trace_read.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int64,
                                  ctypes.POINTER(ctypes.c_int64)]
    lib.bv_trace_read.restype = ctypes.c_int
    bv._lib = lib
    return lib


def read(lib, device=0):
    """The records since the last reset: int64 array (n, 5) of start, end
    (microseconds from the earliest start), kernel, step, workgroup."""
    buf = np.zeros(3 * CAPACITY, dtype=np.uint64)
    count = ctypes.c_int64(0)
    rc = lib.bv_trace_read(device, buf.ctypes.data, CAPACITY, ctypes.byref(count))
    if rc != 0:
        raise RuntimeError("bv_trace_read failed (%d)" % rc)
    n = min(count.value, CAPACITY)
    if count.value > CAPACITY:
        raise RuntimeError("trace overflow: %d records, %d kept" % (count.value, CAPACITY))
    r = buf[:3 * n].reshape(n, 3)
    ident = r[:, 2]
    kernel = (ident >> np.uint64(56)).astype(np.int64)
    step = ((ident >> np.uint64(48)) & np.uint64(0xff)).astype(np.int64) - 1
    wg = ((ident >> np.uint64(8)) & np.uint64(0xffffffff)).astype(np.int64)
    t = r[:, :2].astype(np.int64)
    t0 = t[:, 0].min() if n else 0
    us = (t - t0) * 0.01
    return np.column_stack([us[:, 0], us[:, 1], kernel, step, wg])


def launch_table(rec, cus=256):
    """Per launch (kernel, step): dict rows plus evaluation totals."""
    rows = []
    order = sorted({(int(k), int(s)) for k, s in rec[:, 2:4]}, key=lambda ks: rec[
        (rec[:, 2] == ks[0]) & (rec[:, 3] == ks[1]), 0].min())
    busy_slot_us, span_lo, span_hi, prev_end = 0.0, None, None, None
    gaps = drain_idle = 0.0
    for k, s in order:
        sel = (rec[:, 2] == k) & (rec[:, 3] == s)
        st, en = rec[sel, 0], rec[sel, 1]
        slots = WAVES_PER_SIMD[k] * 4 * cus
        lo, hi, last_start = st.min(), en.max(), st.max()
        dur = hi - lo
        busy = float((en - st).sum())
        busy_slot_us += busy / slots
        grid = np.linspace(last_start, hi, 256)
        resident = np.searchsorted(np.sort(st), grid, side="right") - \
            np.searchsorted(np.sort(en), grid, side="right")
        idle_share = float(np.mean(1.0 - np.minimum(resident, slots) / slots))
        gap = 0.0 if prev_end is None else max(0.0, lo - prev_end)
        gaps += gap
        drain_idle += idle_share * (hi - last_start)
        prev_end = hi if prev_end is None else max(prev_end, hi)
        span_lo = lo if span_lo is None else min(span_lo, lo)
        span_hi = hi if span_hi is None else max(span_hi, hi)
        rows.append(dict(kernel=KERNELS.get(k, str(k)), step=s, waves=int(sel.sum()), start_us=lo,
                         duration_us=dur, gap_before_us=gap,
                         mean_occupancy=busy / (slots * dur) if dur > 0 else 0.0,
                         drain_us=hi - last_start, drain_idle_share=idle_share))
    span = (span_hi - span_lo) if rows else 0.0
    totals = dict(span_us=span, slot_utilisation=busy_slot_us / span if span else 0.0,
                  launch_gaps_us=gaps, launch_gap_share=gaps / span if span else 0.0,
                  drain_idle_us=drain_idle, drain_idle_share=drain_idle / span if span else 0.0)
    return rows, totals


def _workload(name):
    from blockvecchia_amd import datagen
    if name == "uneven":
        x, y = datagen.clustered_locations(200000, 0)
        return x, y, 10000, 400, [1.0, 0.05, 1.5]
    x, y = datagen.uniform_locations(20000, 0)
    return x, y, 1500, 300, [1.5, 0.1, 0.5]


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", choices=["uneven", "bench"], default="uneven")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cus", type=int, default=256)
    args = ap.parse_args(argv)
    lib = use_trace_library()
    x, y, bc, cs, theta = _workload(args.workload)
    n = len(x)
    blocks = bv.partition(x, y, np.zeros(n), bc, cs, seed=0, device=0)
    obs = np.random.default_rng(1).standard_normal(n)   # timing does not depend on the values
    b = bv.Blocks(blocks.x, blocks.y, obs, blocks.batch_num, blocks.nbr, cs)
    eng = bv.Engine(b)
    for _ in range(args.warmup):
        eng.llh(theta)
    cfg = eng.factor_config()
    lib.bv_trace_reset(0)
    eng.llh(theta)
    kernel_ms = eng.last_kernel_ms()
    rec = read(lib)
    eng.close()
    rows, tot = launch_table(rec, args.cus)
    print("workload %s: n=%d bc=%d cs=%d blocks=%d theta=%s, factor config %s, kernel time %.3f ms"
          % (args.workload, n, bc, cs, b.batch_count, theta, list(cfg), kernel_ms))
    print("| launch | wavefronts | start us | duration us | gap before us | mean occupancy "
          "| drain us | idle slots in drain |")
    print("|---|---|---|---|---|---|---|---|")
    for r in rows:
        print("| %s step %d | %d | %.1f | %.1f | %.1f | %.2f | %.1f | %.2f |" % (
            r["kernel"], r["step"], r["waves"], r["start_us"], r["duration_us"], r["gap_before_us"],
            r["mean_occupancy"], r["drain_us"], r["drain_idle_share"]))
    print("span %.1f us; wavefront time / slots %.3f of the span; gaps between launches %.1f us "
          "(%.1f%%); idle slots while launches drain %.1f us (%.1f%%)" % (
              tot["span_us"], tot["slot_utilisation"], tot["launch_gaps_us"],
              100 * tot["launch_gap_share"], tot["drain_idle_us"], 100 * tot["drain_idle_share"]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
