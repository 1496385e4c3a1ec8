"""Benchmark: JSON GB/s indexed + queried on one MI355X.

One step = build the structural index of the whole input (string mask,
leveled rows, newline index) and run every query of the workload over it, all
stream-ordered on one HIP stream, with the input already resident in HBM.

Default workload (the throughput configuration): 8 GiB of seeded synthetic
NDJSON, ~1 KiB records of depth <= 4, four queries over one shared 3-level
index: $.user.lang, $.e[0].f, $.h[*].t and $.h[?(@.k == "v")].id.

Prints one JSON line.  Timing uses HIP events recorded on the stream the
kernels run on; each kernel's launch time gives its roofline figure
(`roofline` is the kernel that takes longest, `roofline_kernels` lists
both).  The CPU baseline is the CPU oracle (a port of the reference
algorithm) on a bounded prefix sample of the same data, at all cores of the
process's CPU share and at one thread.  The default run also summarizes, in
the same line, the batched configuration (10,000 files: one graph launch
against plain per-file launches, `batched`) and the serving configuration
(p50/p99 latency and sustained GB/s per arrival rate, `serving`); they run
after the timed steps.  `--workload serving` prints the serving line alone.
"""
from __future__ import annotations

import argparse
import ctypes
import gc
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
import __graft_entry__ as ge  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "kernel_traffic.json")


KERNEL_SOURCES = ["csrc/gpj_index.hip", "csrc/gpj_query.hip", "csrc/gpj_query_wave.hip",
                  "csrc/gpj_bytecode.h", "csrc/gpj_kernels.h"]


def kernel_sources_hash() -> str:
    """sha256 over the kernel sources the traffic profile was measured on."""
    import hashlib
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "15418-project_amd", rel), "rb") as f:
            h.update(rel.encode() + b"\0" + f.read())
    return h.hexdigest()


def profiled_traffic(workload: str, kernel: str):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE + WRITE_SIZE
    (scaled by the copy-kernel calibration), or None when not profiled or
    when the kernel sources changed since the profile was taken."""
    if not os.path.exists(PROFILE_TRAFFIC):
        return None
    with open(PROFILE_TRAFFIC) as f:
        prof = json.load(f)
    if prof.get("kernel_sources_sha256") != kernel_sources_hash():
        return None
    return prof.get("workloads", {}).get(workload, {}).get(kernel, {}).get("hbm_bytes_per_launch")


def cpu_description() -> dict:
    """Host CPU of this run: model name, logical CPUs and the CPUs this
    process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable}


def baseline_threads() -> int:
    """All cores of this process's CPU share (at most 16: one GPU's share of
    the host)."""
    return max(1, min(16, cpu_description()["usable_cpus"]))


def roofline_entry(kernel: str, algo_bytes: float, ms: float, traffic):
    achieved = algo_bytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kernel,
            "kernel_ms": round(ms, 4), "algorithmic_bytes": int(algo_bytes)}

DEFAULT_WORKLOAD = "ndjson_8GiB_4q"
WORKLOADS = {
    # name: (schema, bytes, seed, queries, extension queries allowed)
    "ndjson_8GiB_4q": (2, 8 << 30, 0x15418002,
                       ["$.user.lang", "$.e[0].f", "$.h[*].t", '$.h[?(@.k == "v")].id']),
    "ndjson_1GiB_4q": (2, 1 << 30, 0x15418002,
                       ["$.user.lang", "$.e[0].f", "$.h[*].t", '$.h[?(@.k == "v")].id']),
    "ndjson_64MiB_ab": (0, 64 << 20, 0x15418000, ["$.a.b"]),
    # the throughput schema at smaller sizes (where the query walks cross over)
    "ndjson_64MiB_4q": (2, 64 << 20, 0x15418002,
                        ["$.user.lang", "$.e[0].f", "$.h[*].t", '$.h[?(@.k == "v")].id']),
    "ndjson_256MiB_4q": (2, 256 << 20, 0x15418002,
                         ["$.user.lang", "$.e[0].f", "$.h[*].t", '$.h[?(@.k == "v")].id']),
    # one 1 GiB document: a top-level array of 2M objects; its elements are
    # the records (queries start with $[*])
    "array_1GiB_items_sku": (1, 2_000_000, 0x15418001, ["$[*].items[*].sku"]),
    # 10,000 NDJSON files, log-normal sizes (median 256 KiB, sigma 1.5,
    # clamped to 1 KiB .. 256 MiB), schemas 0/2/3 mixed, 2 queries per file
    "batch_10k_files": (-1, 10_000, 0x15418003, []),
}
BATCH_QUERIES = ["$.a.b", "$.a", "$.user.lang", "$.h[*].t", "$.entities.hashtags[*].text"]
BATCH_SCHEMA_MASKS = {0: 0b00011, 2: 0b01100, 3: 0b10100}


class Hip:
    """The few HIP runtime calls the timer needs (events + one stream)."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                               ctypes.c_void_p]

    def _ok(self, st, what):
        if st != 0:
            raise RuntimeError(f"{what} failed with HIP error {st}")

    def stream(self):
        s = ctypes.c_void_p()
        self._ok(self.L.hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        return s.value

    def event(self):
        e = ctypes.c_void_p()
        self._ok(self.L.hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        return e.value

    def record(self, e, s):
        self._ok(self.L.hipEventRecord(ctypes.c_void_p(e), ctypes.c_void_p(s)), "hipEventRecord")

    def sync_event(self, e):
        self._ok(self.L.hipEventSynchronize(ctypes.c_void_p(e)), "hipEventSynchronize")

    def elapsed_ms(self, a, b) -> float:
        ms = ctypes.c_float()
        self._ok(self.L.hipEventElapsedTime(ctypes.byref(ms), ctypes.c_void_p(a),
                                            ctypes.c_void_p(b)), "hipEventElapsedTime")
        return float(ms.value)


def cut_at_newline(data: np.ndarray, limit: int) -> np.ndarray:
    if data.size <= limit:
        return data
    nl = np.flatnonzero(data[:limit] == 10)
    return data[: int(nl[-1]) + 1] if nl.size else data[:limit]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD,
                    choices=sorted(WORKLOADS) + ["serving"])
    ap.add_argument("--no-extras", action="store_true",
                    help="default workload only: leave out the batched (10,000 files) and "
                         "serving summaries")
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-verify", type=int, default=100,
                    help="requests per rate the replay compares with solo runs (-1 = all)")
    ap.add_argument("--cpu-sample-mib", type=int, default=2048)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--results", choices=["dense", "slots"], default="slots",
                    help="slots: per-record result arrays ([records][capacity][2], the "
                         "reference's layout; capacity 8 for extension queries); dense: each "
                         "query's spans in one array with CSR offsets (counted, scanned and "
                         "emitted on the device).  The other layout is timed afterwards on "
                         "the same index and compared in full")
    ap.add_argument("--no-layout-check", action="store_true")
    ap.add_argument("--walk", choices=["auto", "lane", "wave"], default="auto",
                    help="query walk: one record per lane, one record per wave (staged in "
                         "LDS), or auto (the wave walk for few records, or few long ones)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain copy kernel over MIB MiB (profiler counter scale)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1 "
              f"(got {args.gpus}).  The sharding design is described in DESIGN.md.",
              file=sys.stderr)
        sys.exit(2)

    gpj = ge.load_package()
    gpj.lib()
    if not gpj.device_available():
        raise SystemExit("bench.py: no HIP device")
    hip = Hip()
    if args.copy_calibration:
        nb = args.copy_calibration << 20
        src, dst = gpj.DeviceBuffer(nb), gpj.DeviceBuffer(nb)
        for _ in range(3):
            gpj.check(gpj.lib().gpj_debug_copy(dst.ptr, src.ptr, nb, None))
        gpj.check(gpj.lib().gpj_stream_sync(None))
        print(json.dumps({"copy_calibration_bytes": nb, "launches": 3}))
        return
    if args.workload == "serving":
        print(json.dumps(serving_line(args)), flush=True)
        return
    gpj.check(gpj.lib().gpj_query_set_walk({"auto": gpj.WALK_AUTO, "lane": gpj.WALK_LANE,
                                            "wave": gpj.WALK_WAVE}[args.walk]), "query walk")
    schema, nbytes, seed, queries = WORKLOADS[args.workload]
    if schema == -1:
        print(json.dumps(run_file_batch(args, gpj, hip, nbytes, seed)), flush=True)
        return
    out = run_single(args, gpj, hip, schema, nbytes, seed, queries)
    if args.workload == DEFAULT_WORKLOAD and not args.no_extras:
        # the batched and serving configurations, summarized in the same line
        gc.collect()
        b = run_file_batch(args, gpj, hip, WORKLOADS["batch_10k_files"][1],
                           WORKLOADS["batch_10k_files"][2], summary=True)
        out["batched"] = {k: b[k] for k in ("config", "ms_per_step", "value", "unit",
                                            "per_file_launches", "batched_over_per_file",
                                            "identical_results", "kernels_ms", "autotune")}
        gc.collect()
        out["serving"] = run_serving(args)
    print(json.dumps(out), flush=True)


def run_single(args, gpj, hip, schema, nbytes, seed, queries):
    """One input indexed and queried per step (the single-input workloads)."""
    t0 = time.time()
    data = gpj.generate(schema, nbytes, seed)
    gen_s = time.time() - t0
    n = int(data.size)
    print(f"[bench] generated {n} bytes in {gen_s:.1f}s", file=sys.stderr, flush=True)

    array_mode = schema == 1
    if array_mode:
        compiled = [gpj.compile_element_query(q) for q in queries]
    else:
        compiled = [gpj.compile_query(q, extensions=True) for q in queries]
    # bytecode with extension opcodes ([*], array filters) runs in variable mode
    variable = [gpj.uses_extension_opcodes(c[0]) for c in compiled]
    ext_cap = 24 if array_mode else 8
    levels = max(1, max(c[1] for c in compiled)) + (1 if array_mode else 0)
    words = (n + 63) // 64
    nl_cap = n // 256 + 2

    stream = hip.stream()
    d_in = gpj.DeviceBuffer(n)
    e0, e1 = hip.event(), hip.event()
    hip.record(e0, stream)
    gpj.check(gpj.lib().gpj_h2d(d_in.ptr, data.ctypes.data, n, stream))
    hip.record(e1, stream)
    hip.sync_event(e1)
    h2d_gbs = n / (hip.elapsed_ms(e0, e1) * 1e-3) / 1e9

    d_string = gpj.DeviceBuffer(words * 8)
    d_leveled = gpj.DeviceBuffer(words * 8 * levels)
    d_newline = gpj.DeviceBuffer(nl_cap * 8)
    d_numlines = gpj.DeviceBuffer(8)
    d_scratch = gpj.DeviceBuffer(gpj.lib().gpj_index_scratch_bytes(n))
    idx = gpj.IndexT(d_newline.ptr, d_numlines.ptr, nl_cap, d_string.ptr, d_leveled.ptr,
                     words, levels, 0)
    d_irs = [gpj.DeviceBuffer.from_host(np.frombuffer(c[0], dtype=np.uint8), stream)
             for c in compiled]
    caps = [ext_cap if v else c[2] for c, v in zip(compiled, variable)]
    L = gpj.lib()
    grid_lines = nl_cap
    nq = len(compiled)
    dense = args.results == "dense"
    descs = (gpj.QueryDescT * nq)()
    slot_bufs = []

    def slot_results(rows):
        # per-record result arrays ([rows][cap][2] spans, counts for extension
        # bytecode), the reference's layout extended by a capacity
        nonlocal slot_bufs
        slot_bufs = [(gpj.DeviceBuffer(max(rows * k * 16, 16)),
                      gpj.DeviceBuffer(max(rows * 8, 8)) if v else None)
                     for k, v in zip(caps, variable)]
        for qi, c in enumerate(compiled):
            sp, ct = slot_bufs[qi]
            descs[qi] = gpj.QueryDescT(d_irs[qi].ptr, len(c[0]), caps[qi], int(variable[qi]),
                                       c[1], sp.ptr, ct.ptr if ct else None)

    dense_res = (gpj.DenseResultT * nq)()
    dense_bufs = {}

    def dense_results(rows, spans_each):
        # one CSR offsets array and one span array per query; scratch
        dense_bufs.clear()
        dense_bufs["offsets"] = [gpj.DeviceBuffer((rows + 1) * 8) for _ in range(nq)]
        dense_bufs["spans"] = [gpj.DeviceBuffer(max(k * 16, 16)) for k in spans_each]
        dense_bufs["scratch"] = gpj.DeviceBuffer(
            L.gpj_query_dense_scratch_bytes(rows, nq, int(sum(spans_each))))
        dense_bufs["rows"] = rows
        for qi, c in enumerate(compiled):
            dense_res[qi] = gpj.DenseResultT(dense_bufs["offsets"][qi].ptr,
                                             dense_bufs["spans"][qi].ptr, int(spans_each[qi]))
            descs[qi] = gpj.QueryDescT(d_irs[qi].ptr, len(c[0]), caps[qi], int(variable[qi]),
                                       c[1], None, None)

    if dense:
        dense_results(nl_cap, [nl_cap] * nq)
    else:
        slot_results(nl_cap)

    bounds_cap = nl_cap
    d_bounds = gpj.DeviceBuffer(bounds_cap * 8) if array_mode else None
    d_nbounds = gpj.DeviceBuffer(8) if array_mode else None
    d_bscratch = (gpj.DeviceBuffer(L.gpj_row_bounds_scratch_bytes(words))
                  if array_mode else None)

    def run_queries(qdescs, count, first=0):
        if array_mode:
            gpj.check(L.gpj_row_bounds(ctypes.byref(idx), d_bounds.ptr, bounds_cap,
                                       d_nbounds.ptr, d_bscratch.ptr, stream), "row bounds")
        if dense:
            res = ctypes.cast(ctypes.addressof(dense_res) + first * ctypes.sizeof(gpj.DenseResultT),
                              ctypes.POINTER(gpj.DenseResultT))
            if array_mode:
                gpj.check(L.gpj_query_run_records_dense(d_in.ptr, n, ctypes.byref(idx),
                                                        d_bounds.ptr, d_nbounds.ptr, qdescs,
                                                        count, grid_lines, res,
                                                        dense_bufs["scratch"].ptr, stream),
                          "dense record query")
            else:
                gpj.check(L.gpj_query_run_dense(d_in.ptr, n, ctypes.byref(idx), qdescs, count,
                                                grid_lines, res, dense_bufs["scratch"].ptr,
                                                stream), "dense query")
        elif array_mode:
            gpj.check(L.gpj_query_run_records(d_in.ptr, n, ctypes.byref(idx), d_bounds.ptr,
                                              d_nbounds.ptr, qdescs, count, grid_lines, stream),
                      "query records")
        else:
            gpj.check(L.gpj_query_run_batch(d_in.ptr, n, ctypes.byref(idx), qdescs, count,
                                            grid_lines, stream), "query batch")

    def step(ev=None):
        if ev is not None:
            hip.record(ev[0], stream)
        gpj.check(L.gpj_index_build_into(d_in.ptr, n, levels, d_string.ptr, d_leveled.ptr,
                                         d_newline.ptr, nl_cap, d_numlines.ptr, d_scratch.ptr,
                                         stream), "index")
        if ev is not None:
            hip.record(ev[1], stream)
        # all queries of the workload in one launch over the shared index
        run_queries(descs, len(compiled))
        if ev is not None:
            hip.record(ev[2], stream)

    def single_query(qi):
        # one query alone, through the same batched entry point
        run_queries(ctypes.byref(descs[qi]), 1, qi)

    # first build: learn the line count (the reference reads it back the same
    # way, orig_index_builder.cu:325-330) and size the query grids to it
    step()
    gpj.check(L.gpj_stream_sync(stream))
    gpj.index_status(d_scratch, n, stream)  # raises if a tile's carry wait gave up
    num_lines = int(d_numlines.to_host(np.int64, 1)[0])
    nl_entries = num_lines
    if num_lines > nl_cap:
        raise SystemExit("bench.py: newline capacity too small for this workload")
    if array_mode:
        # records are the top-level array's elements: bounds - 1 of them
        m = int(d_nbounds.to_host(np.int64, 1)[0])
        if m > bounds_cap:
            raise SystemExit("bench.py: record-bound capacity too small for this workload")
        num_lines = max(m - 1, 0)
    grid_lines = num_lines
    if dense:
        # exact outputs: each query's total from the first run
        totals = [int(o.to_host(np.int64, dense_bufs["rows"] + 1)[min(num_lines,
                                                                      dense_bufs["rows"])])
                  for o in dense_bufs["offsets"]]
        dense_results(num_lines, [max(t, 1) for t in totals])
        step()
    else:
        slot_results(num_lines)
    for _ in range(args.warmup):
        step()
    gpj.check(L.gpj_stream_sync(stream))

    evs = [[hip.event() for _ in range(3)] for _ in range(args.steps)]
    t_start, t_end = hip.event(), hip.event()
    gpj.check(L.gpj_stream_sync(stream))
    hip.record(t_start, stream)
    for k in range(args.steps):
        step(evs[k])
    hip.record(t_end, stream)
    hip.sync_event(t_end)
    gpj.check(L.gpj_stream_sync(stream))
    total_ms = hip.elapsed_ms(t_start, t_end)
    ms_per_step = total_ms / args.steps
    index_ms = float(np.mean([hip.elapsed_ms(e[0], e[1]) for e in evs]))
    batch_ms = float(np.mean([hip.elapsed_ms(e[1], e[2]) for e in evs]))
    # diagnostic only (outside the timed region): each query alone
    query_ms = []
    for qi in range(len(compiled)):
        a, b = hip.event(), hip.event()
        hip.record(a, stream)
        single_query(qi)
        hip.record(b, stream)
        hip.sync_event(b)
        query_ms.append(hip.elapsed_ms(a, b))
    run_queries(descs, len(compiled))
    gpj.check(L.gpj_stream_sync(stream))
    # the last timed build succeeded (each build clears and re-arms the flag)
    gpj.index_status(d_scratch, n, stream)

    # roofline per kernel: algorithmic bytes per launch over the launch time
    # measured with HIP events on the launch stream
    records = num_lines
    algo_index = n + n / 8 + levels * n / 8 + 8 * nl_entries

    # each query's results in CSR form (offsets [records + 1], spans [total, 2])
    def slot_form(qi):
        sp, ct = slot_bufs[qi]
        spans = sp.to_host(np.int64, num_lines * caps[qi] * 2).reshape(num_lines, caps[qi], 2)
        counts = ct.to_host(np.int64, num_lines) if ct else (spans[:, :, 0] >= 0).sum(axis=1)
        if int(counts.max(initial=0)) > caps[qi]:
            raise SystemExit("bench.py: a record exceeded the inline result capacity")
        return gpj.dense_from_slots(spans, counts)

    def csr_results():
        if not dense:
            return [slot_form(qi) for qi in range(nq)]
        offs = [o.to_host(np.int64, num_lines + 1) for o in dense_bufs["offsets"]]
        return [(o, sp.to_host(np.int64, 2 * int(o[-1])).reshape(-1, 2))
                for sp, o in zip(dense_bufs["spans"], offs)]

    got = csr_results()
    matches = [int(o[-1]) for o, _ in got]
    # the other result layout, timed over the same index (queries only) and
    # compared with the timed layout in full
    other = None
    if num_lines and not args.no_layout_check:
        dense = not dense
        if dense:
            dense_results(num_lines, [max(m, 1) for m in matches])
        else:
            dense_bufs.clear()
            slot_results(num_lines)
        run_queries(descs, nq)
        a, b = hip.event(), hip.event()
        gpj.check(L.gpj_stream_sync(stream))
        hip.record(a, stream)
        for _ in range(args.steps):
            run_queries(descs, nq)
        hip.record(b, stream)
        hip.sync_event(b)
        other_ms = hip.elapsed_ms(a, b) / args.steps
        alt = csr_results()
        same = all(np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])
                   for x, y in zip(got, alt))
        if not same:
            raise SystemExit("bench.py: dense and per-record results differ")
        result_bytes = {
            "dense": int(sum((num_lines + 1) * 8 + 16 * m for m in matches)),
            "slots": int(sum(num_lines * 16 * k + (8 * num_lines if v else 0)
                             for k, v in zip(caps, variable)))}
        other = {"results": "dense" if dense else "slots", "query_batch_ms": round(other_ms, 4),
                 "identical_to_timed": same, "result_bytes": result_bytes}
        dense = not dense

    cpu = None
    if not args.no_cpu_baseline:
        orc = ge.load_oracle()

        def cpu_run(sample_bytes, threads):
            if array_mode:
                # a prefix of the array: its complete elements are the records
                sample = data[: min(data.size, sample_bytes >> 3)]
            else:
                sample = cut_at_newline(data, sample_bytes)
            outs = []
            t0 = time.time()
            ix = orc.build_index(sample, levels, threads)
            if array_mode:
                b = orc.row_bounds(ix)
                for c, v in zip(compiled, variable):
                    outs.append(orc.query_records(sample, ix, b, c[0], ext_cap if v else c[2], v))
            else:
                for c, v in zip(compiled, variable):
                    if v:
                        outs.append(orc.query_ext(sample, ix, c[0], ext_cap, threads))
                    else:
                        outs.append((orc.query(sample, ix, c[0], c[2], threads), None))
            secs = time.time() - t0
            return {"value": round(sample.size / secs / 1e9, 4), "threads": threads,
                    "sample": f"first {sample.size} bytes of the same data"
                              f"{'' if array_mode else ' (cut at a newline)'}, index + all "
                              f"{len(compiled)} queries, {secs:.2f}s wall"}, outs

        def sample_agrees(outs) -> int:
            # the timed results of the sample's records equal the oracle's
            # (all but the sample's last record, which the cut changes)
            checked = 0
            for (sp, ct), v, (g_off, g_sp) in zip(outs, variable, got):
                o_off, o_sp = gpj.dense_from_slots(sp, ct if v else None)
                r = o_off.size - 2
                if r <= 0:
                    continue
                if not (np.array_equal(g_off[: r + 1], o_off[: r + 1]) and
                        np.array_equal(g_sp[: int(g_off[r])], o_sp[: int(o_off[r])])):
                    raise SystemExit("bench.py: results differ from the CPU oracle's on the "
                                     "CPU baseline sample")
                checked = r
            return checked

        all_threads = args.cpu_threads or baseline_threads()
        t_all, outs = cpu_run(args.cpu_sample_mib << 20, all_threads)
        checked = sample_agrees(outs)
        del outs
        t_one, _ = cpu_run(args.cpu_sample_mib << 17, 1)  # 1/8 of the sample on one thread
        cpu = {"value": t_all["value"], "unit": "GB/s", "cores": all_threads, "kind": "port",
               "sample": t_all["sample"] + "; CPU oracle (port of the reference algorithm, "
                         "index and line queries split over contiguous line ranges"
                         f"{', element queries on one thread' if array_mode else ''})",
               "threads_all": t_all, "threads_1": t_one,
               "records_equal_to_gpu": checked, **cpu_description()}

    # per query: newline index read, one pass over the first leveled row,
    # 8 bytes per emitted span (index-walk and text reads are data dependent)
    algo_queries = sum(8 * (records + 1) + n / 8 + 8 * m for m in matches)
    # the walk the query calls took (gpj_query_set_walk; the input buffer is
    # 16-byte aligned)
    walk_used = args.walk if args.walk != "auto" else (
        "wave" if gpj.auto_walk_is_wave(n, grid_lines) else "lane")
    qkernel = "gpj::query_wave_kernel" if walk_used == "wave" else "gpj::query_batch_kernel"
    kernels = [
        roofline_entry("gpj::index_kernel", algo_index, index_ms,
                       profiled_traffic(args.workload, "gpj::index_kernel")),
        roofline_entry(qkernel, algo_queries, batch_ms, profiled_traffic(args.workload, qkernel)),
    ]
    dominant = max(kernels, key=lambda k: k["kernel_ms"])
    out = {
        "metric": "JSON GB/s indexed+queried (fraction of HBM peak); p50/p99 request latency",
        "value": round(n / (ms_per_step * 1e-3) / 1e9, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": args.workload, "bytes": n,
                   "records": num_lines if array_mode else num_lines - 1,
                   "queries": queries, "levels": levels, "matches": matches,
                   "results": args.results, "other_results": other, "walk": walk_used,
                   "parallelism": "single-gpu"},
        "roofline": dominant,
        "roofline_kernels": kernels,
        "cpu_baseline": cpu,
        "kernels_ms": {"index": round(index_ms, 4), "query_batch": round(batch_ms, 4),
                       **{f"query_alone[{q}]": round(t, 4) for q, t in zip(queries, query_ms)}},
        "h2d_GBps_pageable": round(h2d_gbs, 2),
        "end_to_end_hbm_frac": round(
            (algo_index + algo_queries) / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
    }
    return out


def run_file_batch(args, gpj, hip, num_files: int, seed: int, summary: bool = False):
    """Many files: one graph launch (batched index + batched query) per step,
    against plain per-file launches over the same device-resident files."""
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.exp(rng.normal(np.log(256 << 10), 1.5, num_files)), 1 << 10, 256 << 20)
    schemas = rng.choice([0, 2, 3], num_files).astype(np.int32)
    masks = np.array([BATCH_SCHEMA_MASKS[int(x)] for x in schemas], dtype=np.uint32)
    t0 = time.time()
    buf, off, sz = gpj.generate_files(schemas, sizes.astype(np.uint64), seed)
    print(f"[bench] generated {num_files} files, {int(sz.sum())} bytes in "
          f"{time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    compiled = [gpj.compile_query(q, extensions=True) for q in BATCH_QUERIES]
    stream = gpj.new_stream()
    b = gpj.FileBatch(buf, off, sz, masks, compiled, capacity=8, stream=stream)
    L = gpj.lib()
    total = int(sz.sum())

    def timed(fn, reps):
        a, z = hip.event(), hip.event()
        gpj.check(L.gpj_stream_sync(stream))
        hip.record(a, stream)
        for _ in range(reps):
            fn()
        hip.record(z, stream)
        hip.sync_event(z)
        return hip.elapsed_ms(a, z) / reps

    def one_batched(batch):
        batch.launch()
        return timed(batch.launch, 2)

    tune = gpj.autotune_batch(b, one_batched,
                              os.path.join(ROOT, "15418-project_amd", "autotune_cache.json"))
    # per-kernel split of one eager batched run (roofline of the index kernel)
    e = [hip.event() for _ in range(3)]
    gpj.check(L.gpj_stream_sync(stream))
    hip.record(e[0], stream)
    gpj.check(L.gpj_batch_index(b.d_in.ptr, b.n, b.d_files.ptr, b.num_files, b.d_tile_file.ptr,
                                b.levels, b.d_string.ptr, b.d_leveled.ptr, b.d_newline.ptr,
                                b.d_prefix.ptr, b.d_scratch.ptr, stream), "batch index")
    hip.record(e[1], stream)
    gpj.check(L.gpj_batch_query_shaped(b.d_in.ptr, b.n, b.d_files.ptr, b.num_files, b.levels,
                                       b.d_string.ptr, b.d_leveled.ptr, b.d_newline.ptr,
                                       b.d_prefix.ptr, b.grid_records, b.per_wave, b.descs,
                                       len(compiled), ctypes.byref(b.results[0]), stream),
              "batch query")
    hip.record(e[2], stream)
    hip.sync_event(e[2])
    index_ms, query_ms = hip.elapsed_ms(e[0], e[1]), hip.elapsed_ms(e[1], e[2])

    b.capture()
    for _ in range(args.warmup):
        b.launch_graph()
    batched_ms = timed(b.launch_graph, args.steps)
    b.per_file_setup()
    b.launch_per_file()
    per_file_ms = timed(b.launch_per_file, 2 if summary else max(1, args.steps // 4))
    bs, bc = b.fetch()
    ps, pc = b.fetch(per_file=True)
    identical = all(np.array_equal(bs[k], ps[k]) and np.array_equal(bc[k], pc[k])
                    for k in range(2))
    if not identical:
        raise SystemExit("bench.py: batched and per-file results differ")
    b.release_graph()

    lines = b.total_lines
    algo_index = b.n + b.n / 8 + b.levels * b.n / 8 + 8 * lines
    # each file runs two queries: its newline slice, one pass over its first
    # leveled row, 8 bytes per emitted span
    matched = int(sum(int(np.minimum(bc[k], b.capacity).sum()) for k in range(2)))
    algo_queries = 2 * (8 * lines + total / 8) + 8 * matched
    kernels = [roofline_entry("gpj::index_kernel (multi-file)", algo_index, index_ms,
                              profiled_traffic(args.workload, "gpj::index_kernel")),
               roofline_entry("gpj::query_files_kernel", algo_queries, query_ms,
                              profiled_traffic(args.workload, "gpj::query_files_kernel"))]
    dominant = max(kernels, key=lambda k: k["kernel_ms"])
    cpu = None
    if not args.no_cpu_baseline and not summary:
        orc = ge.load_oracle()
        threads = args.cpu_threads or baseline_threads()
        budget, done, t0 = args.cpu_sample_mib << 19, 0, time.time()
        nfiles = 0
        for f in range(num_files):
            if done >= budget:
                break
            data = buf[int(off[f]):int(off[f] + sz[f])]
            ix = orc.build_index(data, b.levels, threads)
            for q in range(len(compiled)):
                if (int(masks[f]) >> q) & 1:
                    ir, _d, slots = compiled[q]
                    if gpj.uses_extension_opcodes(ir):
                        orc.query_ext(data, ix, ir, 8, threads)
                    else:
                        orc.query(data, ix, ir, 8, threads)
            done += int(sz[f])
            nfiles += 1
        cpu_s = time.time() - t0
        cpu = {"value": round(done / cpu_s / 1e9, 4), "unit": "GB/s", "cores": threads,
               "kind": "port",
               "sample": f"first {nfiles} files ({done} bytes), index + their 2 queries each; "
                         f"multi-threaded CPU oracle, {cpu_s:.2f}s wall", **cpu_description()}
    out = {
        "metric": "JSON GB/s indexed+queried (fraction of HBM peak); p50/p99 request latency",
        "value": round(total / (batched_ms * 1e-3) / 1e9, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(batched_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "batch_10k_files", "files": num_files, "bytes": total,
                   "buffer_bytes": b.n, "records": lines - num_files, "queries": BATCH_QUERIES,
                   "levels": b.levels, "parallelism": "single-gpu"},
        "roofline": dominant,
        "roofline_kernels": kernels,
        "cpu_baseline": cpu,
        "kernels_ms": {"batch_index": round(index_ms, 4), "batch_query": round(query_ms, 4)},
        "per_file_launches": {"ms_per_step": round(per_file_ms, 4),
                              "GBps": round(total / (per_file_ms * 1e-3) / 1e9, 3)},
        "batched_over_per_file": round(per_file_ms / batched_ms, 2),
        "identical_results": identical,
        "autotune": tune,
        "line_counts": "on the device (gpj_batch_line_counts), no host pass over the bytes",
    }
    return out


SERVE_BIN = os.path.join(ROOT, "15418-project_amd", "_build", "gpjson_serve")
SERVING_RATES = "500,1000,2000,5000"


def run_serving(args) -> dict:
    """The serving configuration through the gpjson_serve replay (a child
    process): Poisson arrivals at each rate, documents of 4 KiB .. 64 MiB
    (log-uniform) from the three record schemas, 1..8 queries each from that
    schema's configuration, 30% re-queries, an 80 GiB HBM budget, 8 slots."""
    cmd = [SERVE_BIN, "--requests", str(args.serve_requests), "--rates", SERVING_RATES,
           "--min-doc", "4KiB", "--max-doc", "64MiB", "--requery", "0.3", "--budget-gib", "80",
           "--slots", "8", "--verify", str(args.serve_verify), "--seed", "0x15418004"]
    print(f"[bench] serving: {' '.join(cmd[1:])}", file=sys.stderr, flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
    if p.returncode != 0:
        raise SystemExit("bench.py: serving replay failed: " + p.stderr[-1500:])
    rows = [json.loads(x) for x in p.stdout.splitlines() if x.startswith("{")]
    keep = ("rate_per_s", "p50_ms", "p99_ms", "mean_ms", "sustained_GBps", "new_document_GBps",
            "makespan_s", "resident_hits", "evictions", "retries", "peak_device_bytes",
            "verified", "identical")
    return {"workload": f"serving: {args.serve_requests} requests per rate, Poisson arrivals, "
                        "documents 4 KiB-64 MiB log-uniform from the 64 MiB / 1 GiB / 8 GiB "
                        "record schemas, 1-8 queries each, 30% re-queries, 80 GiB HBM budget, "
                        "8 slots; latency = arrival to results in host memory",
            "documents": rows[0]["documents"] if rows else 0,
            "rates": [{k: r[k] for k in keep} for r in rows]}


def serving_line(args) -> dict:
    """The bench line of the serving workload: value = sustained GB/s at the
    highest arrival rate, with p50/p99 latency at every rate."""
    srv = run_serving(args)
    top = max(srv["rates"], key=lambda r: r["rate_per_s"])
    return {
        "metric": "JSON GB/s indexed+queried (fraction of HBM peak); p50/p99 request latency",
        "value": top["sustained_GBps"], "unit": "GB/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(top["makespan_s"] * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": "serving", "requests_per_rate": args.serve_requests,
                   "rates_per_s": [r["rate_per_s"] for r in srv["rates"]],
                   "parallelism": "single-gpu"},
        "p50_ms": top["p50_ms"], "p99_ms": top["p99_ms"], "serving": srv,
        "roofline": None, "cpu_baseline": None,
    }


if __name__ == "__main__":
    main()
