"""Launch-shape autotuner for the gfx950 kernels, and its lookup table.

Tunables (each a key of the library's tuning interface):

* ``flat`` -- ``dpw``: data rows each workgroup of the tiled flat kernel
  walks while its template stays in registers (more rows: fewer template
  reads and fewer, longer workgroups; fewer rows: more workgroups in flight);
* ``big`` -- ``big_chunk_mb``: for the four-step lengths (n >= 32768), the
  pairs whose n-point intermediate is held at once (column launch, then row
  launch, per chunk): larger chunks mean fewer, fuller launches until the
  intermediate no longer stays in the memory-side cache between the two;
* ``hier`` -- ``refine_grid``: workgroups of the list-driven refine launch,
  which drains however many pairs the coarse gate passed;
* ``coarse`` -- ``coarse_lds`` x ``coarse_rpw`` x ``coarse_tpw`` x
  ``coarse_fuse``: where the matrix-core coarse screen keeps its DFT
  constants (registers, or read from a per-workgroup LDS copy: fewer VGPRs,
  more waves per SIMD), the data rows and templates each wave takes (more of
  either: the per-wave set-up is shared by more pairs; fewer: more waves in
  flight), and whether (n = 4096, one bin, constants in LDS) the screen's
  workgroups refine their own survivors in the same launch.  Timed as whole
  hierarchical calls back to back (wall clock per call, fp16 screen), at a
  gate no noise pair passes and at one about 1% pass: single launches
  bracketed by events rank these splits differently from a stream of calls;
* ``queue`` -- ``grid_per_cu``: persistent workgroups per CU of the
  mixed-length refine queue;
* ``band`` -- not a launch knob but the measured costs the hierarchical
  filter chooses its coarse band from: per length, the cost of the screen
  alone and of refining every pair, per band, at two batch sizes
  (HierarchicalFilter's automatic configuration).

``python -m apogee_amd.autotune --out apogee_amd/tuning_gfx950.json`` times
every candidate on the device with HIP events (median of repeated launches,
inputs resident) over a fixed set of shapes and writes the winners.  At run
time :func:`lookup` answers with the entry of the same operation and length
whose (ndata, ntemplates) is nearest in log scale; with no table or no entry
the library's built-in defaults stay.

The reference deliberately does no autotuning (tools/hmf_design.py:4); its
knobs are environment variables read once (src/matchfilt.c:93-107).
"""
import json
import math
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
TABLE_PATH = os.path.join(_HERE, "tuning_gfx950.json")

CANDIDATES = {
    "flat": {"dpw": [1, 2, 4, 8, 16]},
    "big": {"big_chunk_mb": [64, 96, 128, 160, 192, 256]},
    "hier": {"refine_grid": [256, 512, 1024, 2048, 4096]},
    "coarse": {"coarse_lds": [0, 1], "coarse_rpw": [1, 2], "coarse_tpw": [1, 2, 4, 8, 16], "coarse_fuse": [0, 1]},
    "queue": {"grid_per_cu": [1, 2]},
}

# shapes the table covers: (n, ndata, ntemplates, binsize)
FLAT_SHAPES = [
    (1024, 16, 1024, 1024), (4096, 16, 1024, 4096), (4096, 64, 256, 512), (4096, 4, 4096, 4096),
    (16384, 16, 256, 16384), (65536, 4, 128, 1024),
]
BIG_SHAPES = [(32768, 4, 256, 1024), (65536, 4, 128, 1024)]
HIER_SHAPES = [(4096, 16, 1024, 4096, 512)]
COARSE_SHAPES = [(4096, 16, 1024, 512), (4096, 64, 256, 512), (4096, 4, 4096, 512)]
QUEUE_WORKLOADS = [("uneven-4000", 4000)]
# hierarchical band costs: (n, ndata, ntemplates) of the full batch, timed
# also at a small batch (4 x 64 pairs) so cost can be interpolated in pairs
BAND_SHAPES = [(256, 16, 256), (512, 16, 256), (1024, 16, 1024), (2048, 16, 512), (4096, 16, 1024),
               (8192, 16, 256), (16384, 16, 256), (32768, 4, 128), (65536, 4, 128)]
BAND_MAX = 16384      # the largest band the one-pass coarse screen runs

_table = None


def _load():
    global _table
    if _table is None:
        try:
            with open(TABLE_PATH) as fh:
                _table = json.load(fh)
        except (OSError, ValueError):
            _table = {"entries": []}
    return _table


def lookup(op, n=None, ndata=None, ntemplates=None):
    """Best known settings {key: value} for an operation and shape (may be {})."""
    entries = [e for e in _load().get("entries", []) if e.get("op") == op]
    if n is not None:
        same = [e for e in entries if e.get("n") == n]
        entries = same or entries
    if not entries:
        return {}

    def dist(e):
        d = 0.0
        for key, v in (("ndata", ndata), ("ntemplates", ntemplates), ("n", n)):
            if v and e.get(key):
                d += abs(math.log2(v) - math.log2(e[key]))
        return d
    return dict(min(entries, key=dist)["best"])


def apply(target, op, n=None, ndata=None, ntemplates=None):
    """Set the looked-up values on anything with ``set_tuning(key, value)``."""
    best = lookup(op, n, ndata, ntemplates)
    for k, v in best.items():
        target.set_tuning(k, int(v))
    return best


# ------------------------------------------------------------- measuring

def _median_ms(fn, reps, ctx, name):
    import statistics
    times = []
    for _ in range(reps):
        ctx.prof_enable(True)
        fn()
        ctx.sync()
        ms, cnt = ctx.prof_read(name)
        times.append(ms)
    ctx.prof_enable(False)
    return statistics.median(times)


def _tune_flat(device, n, nd, nt, binsize, reps, rng):
    import numpy as np
    from . import _hip
    ctx = _hip.Context(device, n, nd, nt)
    ctx.set_rows("data", 0, (rng.standard_normal((nd, n)) + 1j * rng.standard_normal((nd, n))).astype(np.complex64))
    ctx.set_rows("template", 0, (rng.standard_normal((nt, n)) + 1j * rng.standard_normal((nt, n))).astype(np.complex64))
    import torch
    nb = -(-n // binsize)
    idx = torch.empty(nd * nt * nb, dtype=torch.int64, device="cuda:%d" % device)
    val = torch.empty(nd * nt * nb, dtype=torch.complex64, device="cuda:%d" % device)
    ms = {}
    for dpw in CANDIDATES["flat"]["dpw"]:
        if dpw > nd:
            continue
        ctx.set_tuning("dpw", dpw)
        run = lambda: ctx.peaks_async(0, nd, 0, nt, binsize, 0.0, 0, n, idx.data_ptr(), val.data_ptr())
        run()
        ctx.sync()
        ms[dpw] = _median_ms(run, reps, ctx, "flat")
    best = min(ms, key=ms.get)
    ctx.destroy()
    return {"op": "flat", "n": n, "ndata": nd, "ntemplates": nt, "binsize": binsize,
            "best": {"dpw": best}, "ms": {str(k): round(v, 5) for k, v in ms.items()}}


def _tune_big(device, n, nd, nt, binsize, reps, rng):
    import numpy as np
    from . import _hip
    ctx = _hip.Context(device, n, nd, nt)
    ctx.set_rows("data", 0, (rng.standard_normal((nd, n)) + 1j * rng.standard_normal((nd, n))).astype(np.complex64))
    ctx.set_rows("template", 0, (rng.standard_normal((nt, n)) + 1j * rng.standard_normal((nt, n))).astype(np.complex64))
    import torch
    nb = -(-n // binsize)
    idx = torch.empty(nd * nt * nb, dtype=torch.int64, device="cuda:%d" % device)
    val = torch.empty(nd * nt * nb, dtype=torch.complex64, device="cuda:%d" % device)
    ms = {}
    for mb in CANDIDATES["big"]["big_chunk_mb"]:
        ctx.set_tuning("big_chunk_mb", mb)
        run = lambda: ctx.peaks_async(0, nd, 0, nt, binsize, 0.0, 0, n, idx.data_ptr(), val.data_ptr())
        run()
        ctx.sync()
        ms[mb] = _median_ms(run, reps, ctx, "flat")
    best = min(ms, key=ms.get)
    ctx.destroy()
    return {"op": "big", "n": n, "ndata": nd, "ntemplates": nt, "binsize": binsize,
            "best": {"big_chunk_mb": best}, "ms": {str(k): round(v, 5) for k, v in ms.items()}}


def _tune_hier(device, n, nd, nt, binsize, band, reps, rng):
    import numpy as np
    import torch
    from . import _hip
    ctx = _hip.Context(device, n, nd, nt)
    ctx.set_rows("data", 0, (rng.standard_normal((nd, n)) + 1j * rng.standard_normal((nd, n))).astype(np.complex64))
    ctx.set_rows("template", 0, (rng.standard_normal((nt, n)) + 1j * rng.standard_normal((nt, n))).astype(np.complex64))
    ctx.hier_set_band(band, -1.0)
    nb = -(-n // binsize)
    idx = torch.empty(nd * nt * nb, dtype=torch.int64, device="cuda:%d" % device)
    val = torch.empty(nd * nt * nb, dtype=torch.complex64, device="cuda:%d" % device)
    ms = {}
    # gates that pass about 1% and about 25% of noise pairs: the refine
    # launch sees a short or a long uneven list, which is what its grid size
    # trades against; the sum of the two medians decides
    cmax = np.sort(ctx.hier_coarse_max(0, nd, 0, nt, 0, n, 0).ravel())
    gates = [float(cmax[int(f * cmax.size)]) for f in (0.99, 0.75)]
    for grid in CANDIDATES["hier"]["refine_grid"]:
        ctx.set_tuning("refine_grid", grid)
        ms[grid] = 0.0
        for gate in gates:
            run = lambda: ctx.hier_peaks_async(0, nd, 0, nt, binsize, 0.0, 0, n, gate, idx.data_ptr(),
                                               val.data_ptr())
            run()
            ctx.sync()
            times = []
            for _ in range(reps):
                ctx.prof_enable(True)
                run()
                ctx.sync()
                times.append(sum(ctx.prof_read(k)[0] for k in ("coarse", "gate", "refine")))
            ctx.prof_enable(False)
            ms[grid] += sorted(times)[len(times) // 2]
    best = min(ms, key=ms.get)
    ctx.destroy()
    return {"op": "hier", "n": n, "ndata": nd, "ntemplates": nt, "binsize": binsize, "band": band,
            "best": {"refine_grid": best}, "ms": {str(k): round(v, 5) for k, v in ms.items()}}


def _tune_coarse(device, n, nd, nt, band, reps, rng, calls=200):
    import itertools
    import time
    import numpy as np
    import torch
    from . import _hip
    ctx = _hip.Context(device, n, nd, nt)
    ctx.set_rows("data", 0, (rng.standard_normal((nd, n)) + 1j * rng.standard_normal((nd, n))).astype(np.complex64))
    ctx.set_rows("template", 0, (rng.standard_normal((nt, n)) + 1j * rng.standard_normal((nt, n))).astype(np.complex64))
    ctx.hier_set_band(band, -1.0)
    ctx.hier_set_coarse_precision(1, 2e-3)   # the fp16 screen
    idx = torch.empty(nd * nt, dtype=torch.int64, device="cuda:%d" % device)
    val = torch.empty(nd * nt, dtype=torch.complex64, device="cuda:%d" % device)
    # coarse |z| of every pair: a gate just above the largest passes none,
    # one at the 99th percentile about 1%
    cmax = np.sort(ctx.hier_coarse_max(0, nd, 0, nt, 0, n, 1).ravel())
    gates = [float(cmax[-1]) * 1.01, float(cmax[int(0.99 * cmax.size)])]
    ms = {}
    c = CANDIDATES["coarse"]
    for lds, rpw, tpw, fuse in itertools.product(c["coarse_lds"], c["coarse_rpw"], c["coarse_tpw"],
                                                 c["coarse_fuse"]):
        if fuse and not lds:
            continue   # the one-launch form reads its constants from LDS
        ctx.set_tuning("coarse_lds", lds)
        ctx.set_tuning("coarse_rpw", rpw)
        ctx.set_tuning("coarse_tpw", tpw)
        ctx.set_tuning("coarse_fuse", fuse)
        total = 0.0
        for gate in gates:
            run = lambda: ctx.hier_peaks_async(0, nd, 0, nt, n, 0.0, 0, n, gate, idx.data_ptr(), val.data_ptr())
            for _ in range(20):
                run()
            ctx.sync()
            times = []
            for _ in range(reps):
                t0 = time.perf_counter()
                for _ in range(calls):
                    run()
                ctx.sync()
                times.append((time.perf_counter() - t0) * 1e3 / calls)
            total += sorted(times)[len(times) // 2]
        ms[(lds, rpw, tpw, fuse)] = total
    best = min(ms, key=ms.get)
    ctx.destroy()
    return {"op": "coarse", "n": n, "ndata": nd, "ntemplates": nt, "band": band,
            "best": {"coarse_lds": best[0], "coarse_rpw": best[1], "coarse_tpw": best[2], "coarse_fuse": best[3]},
            "ms": {"%d:%dx%d%s" % (k[0], k[1], k[2], "f" if k[3] else ""): round(v, 5) for k, v in ms.items()},
            "timed": "ms per hierarchical call, back to back, summed over a gate passing none and one passing ~1%"}


def band_entry(n):
    """The measured band-cost entry of length n, or None."""
    for e in _load().get("entries", []):
        if e.get("op") == "band" and e.get("n") == n:
            return e
    return None


def _stage_ms(ctx, run, reps):
    """Median over reps of the summed event time of one hierarchical call."""
    import statistics
    run()
    ctx.sync()
    times = []
    for _ in range(reps):
        ctx.prof_enable(True)
        run()
        ctx.sync()
        times.append(sum(ctx.prof_read(k)[0] for k in ("coarse", "gate", "refine")))
    ctx.prof_enable(False)
    return statistics.median(times)


def _measure_band(device, n, nd, nt, reps, rng):
    """Hierarchical call costs per coarse band at length n: the screen alone
    (a gate nothing passes: coarse screen, gate and an empty refine launch)
    and with every pair refined, each at a small (4 x 64) and the full
    (nd x nt) batch, FP32 screen for every band and the fp16 matrix-core
    screen at band 512.  Unit-norm white templates, complex N(0,1) data."""
    import numpy as np
    import torch
    from . import _hip
    ctx = _hip.Context(device, n, nd, nt)
    ctx.set_rows("data", 0, (rng.standard_normal((nd, n)) + 1j * rng.standard_normal((nd, n))).astype(np.complex64))
    T = rng.standard_normal((nt, n)) + 1j * rng.standard_normal((nt, n))
    ctx.set_rows("template", 0, (T / np.linalg.norm(T, axis=1, keepdims=True)).astype(np.complex64))
    idx = torch.empty(nd * nt, dtype=torch.int64, device="cuda:%d" % device)
    val = torch.empty(nd * nt, dtype=torch.complex64, device="cuda:%d" % device)
    shapes = [(min(nd, 4), min(nt, 64)), (nd, nt)]

    def hier(snd, snt, gate):
        return lambda: ctx.hier_peaks_async(0, snd, 0, snt, n, 0.0, 0, n, gate, idx.data_ptr(), val.data_ptr())

    flat = []
    for snd, snt in shapes:
        run = lambda: ctx.peaks_async(0, snd, 0, snt, n, 0.0, 0, n, idx.data_ptr(), val.data_ptr())
        run()
        ctx.sync()
        flat.append(round(_median_ms(run, reps, ctx, "flat"), 5))
    bands = {}
    b = 64
    while b < n and b <= BAND_MAX:
        ctx.hier_set_band(b, -1.0)
        ctx.hier_set_coarse_precision(0, 0.0)
        row = {"screen_ms": [], "refine_ms_all": []}
        for snd, snt in shapes:
            screen = _stage_ms(ctx, hier(snd, snt, 3.0e38), reps)
            every = _stage_ms(ctx, hier(snd, snt, 0.0), reps)
            row["screen_ms"].append(round(screen, 5))
            row["refine_ms_all"].append(round(max(every - screen, 0.0), 5))
        if b == 512:
            ctx.hier_set_coarse_precision(1, 2e-3)
            row["screen_ms_fp16"] = [round(_stage_ms(ctx, hier(snd, snt, 3.0e38), reps), 5)
                                     for snd, snt in shapes]
            ctx.hier_set_coarse_precision(0, 0.0)
        bands[str(b)] = row
        b *= 2
    ctx.destroy()
    return {"op": "band", "n": n, "ndata": nd, "ntemplates": nt,
            "pairs": [snd * snt for snd, snt in shapes], "flat_ms": flat, "bands": bands}


def uneven_workload(nreq, seed, device="cuda"):
    """The uneven-refine plan: n in {1024, 4096, 16384, 65536} with weights
    .4/.3/.2/.1, binsize n/8, candidates ~ geometric(1/3); spectra are seeded
    Gaussian complex64 generated on the device.  Returns (pool, plan kwargs)."""
    import numpy as np
    import torch
    rng = np.random.default_rng(seed)
    n = rng.choice([1024, 4096, 16384, 65536], size=nreq, p=[0.4, 0.3, 0.2, 0.1]).astype(np.int64)
    cand = rng.geometric(1.0 / 3.0, size=nreq).astype(np.int64)
    sizes = n * (1 + cand)
    doff = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    pool = torch.randn(int(sizes.sum()), dtype=torch.complex64, device=device, generator=g)
    return pool, dict(n=n, data_offset=doff, template_offset=doff + n, candidates=cand, binsize=n // 8)


def _tune_queue(device, name, nreq, reps):
    from .refine import RefineQueue
    import torch
    pool, plan = uneven_workload(nreq, 20261204, device="cuda:%d" % device)
    q = RefineQueue(device="gpu:%d" % device)
    q.plan(pool, **plan)
    idx = torch.empty(q.total, dtype=torch.int64, device=pool.device)
    val = torch.empty(q.total, dtype=torch.complex64, device=pool.device)
    ms = {}
    for gpc in CANDIDATES["queue"]["grid_per_cu"]:
        q.set_tuning("grid_per_cu", gpc)
        q.run_async(idx, val)
        q.sync()
        times = []
        for _ in range(reps):
            q._q.prof_enable(True)
            q.run_async(idx, val)
            q.sync()
            times.append(q._q.prof_read()[0])
        q._q.prof_enable(False)
        ms[gpc] = sorted(times)[len(times) // 2]
    best = min(ms, key=ms.get)
    return {"op": "queue", "workload": name, "requests": nreq, "best": {"grid_per_cu": best},
            "ms": {str(k): round(v, 5) for k, v in ms.items()}}


def main(argv=None):
    import argparse
    import numpy as np
    from . import _hip
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=TABLE_PATH)
    ap.add_argument("--reps", type=int, default=7)
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--only", default="",
                    help="comma-separated operations to re-measure (flat, big, hier, coarse, band, queue); "
                         "the other entries of the existing table are kept")
    args = ap.parse_args(argv)
    only = {o for o in args.only.split(",") if o}
    name, arch = _hip.device_info(args.device)
    rng = np.random.default_rng(7)
    keep = []
    if only:
        try:
            with open(TABLE_PATH) as fh:
                keep = [e for e in json.load(fh).get("entries", []) if e.get("op") not in only]
        except (OSError, ValueError):
            keep = []
    want = (lambda op: not only or op in only)
    entries = []

    def add(e):
        print(json.dumps(e), flush=True)
        entries.append(e)

    if want("flat"):
        for n, nd, nt, bs in FLAT_SHAPES:
            add(_tune_flat(args.device, n, nd, nt, bs, args.reps, rng))
    if want("big"):
        for n, nd, nt, bs in BIG_SHAPES:
            add(_tune_big(args.device, n, nd, nt, bs, args.reps, rng))
    if want("hier"):
        for n, nd, nt, bs, band in HIER_SHAPES:
            add(_tune_hier(args.device, n, nd, nt, bs, band, args.reps, rng))
    if want("coarse"):
        for n, nd, nt, band in COARSE_SHAPES:
            add(_tune_coarse(args.device, n, nd, nt, band, args.reps, rng))
    if want("band"):
        for n, nd, nt in BAND_SHAPES:
            add(_measure_band(args.device, n, nd, nt, args.reps, rng))
    if want("queue"):
        for wname, nreq in QUEUE_WORKLOADS:
            add(_tune_queue(args.device, wname, nreq, args.reps))
    order = ["flat", "big", "hier", "coarse", "band", "queue"]
    entries = sorted(keep + entries, key=lambda e: order.index(e["op"]) if e["op"] in order else 99)
    table = {"device": name, "arch": arch,
             "method": "median of %d event-timed launches (coarse: of %d wall-clock runs of back-to-back calls)"
                       % (args.reps, args.reps),
             "entries": entries}
    with open(args.out, "w") as fh:
        json.dump(table, fh, indent=1)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
