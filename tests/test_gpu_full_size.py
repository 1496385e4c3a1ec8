"""GPU parity at the full configured sizes, through properties that hold at
any size:

* 8 GiB NDJSON of ~1 KiB records, four queries over one 3-level index: the
  newline index is checked exhaustively (count, order, every entry a '\\n');
  the string mask and leveled rows of random record-aligned windows equal
  the oracle's index of the window alone (at a record start the carried
  state is the initial one); each query's spans on 2,000 random records
  equal the oracle's on those records alone; and each query's total equals
  a count of byte patterns the generator writes only for matching values
  (inside a string every '"' is escaped, so a pattern that starts with an
  unescaped '"' after '{' or ',' cannot occur there).
* The 1 GiB top-level array of 2 M objects, $[*].items[*].sku: element
  records on 2,000 random elements equal the oracle's line mode on those
  elements' bytes, and the total equals the number of '"sku":' keys.
* The 64 MiB plain input, compared whole with the oracle.
* The 10,000-file batch (7.9 GB): line prefix, and every structure and
  result of 31 files (30 random ones and the largest) equal the oracle run on
  each file alone; plain per-file launches equal the batched launch for all
  files.
"""
import numpy as np
import pytest

from batchdata import BATCH_QUERIES, plan

pytestmark = pytest.mark.gpu
SAMPLE = 2000
THROUGHPUT_QUERIES = ["$.user.lang", "$.e[0].f", "$.h[*].t", '$.h[?(@.k == "v")].id']


def count_pattern(data: np.ndarray, pat: bytes, chunk: int = 1 << 28) -> int:
    """Occurrences of `pat` in `data`, by chunks that overlap by len(pat) - 1."""
    total, n, k = 0, data.size, len(pat)
    for lo in range(0, n, chunk):
        hi = min(n, lo + chunk + k - 1)
        total += data[lo:hi].tobytes().count(pat)
    return total


def count_byte(data: np.ndarray, b: int, chunk: int = 1 << 28) -> int:
    return sum(int(np.count_nonzero(data[lo:lo + chunk] == b))
               for lo in range(0, data.size, chunk))


def bits(words: np.ndarray) -> np.ndarray:
    return np.unpackbits(np.ascontiguousarray(words).view(np.uint8), bitorder="little")


def masked(spans: np.ndarray, counts) -> np.ndarray:
    out = spans.copy()
    if counts is not None:
        out[np.arange(out.shape[1])[None, :] >= counts[:, None]] = -1
    return out


def relative(spans: np.ndarray, starts: np.ndarray) -> np.ndarray:
    """Spans made relative to each record's first byte (-1 stays -1)."""
    out = spans.copy()
    hit = out >= 0
    out[hit] = (out - starts[:, None, None])[hit]
    return out


@pytest.fixture(scope="module")
def ndjson_8gib(gpu):
    data = gpu.generate(gpu.SCHEMA_NDJSON_1K, 8 << 30, 0x15418002)
    print(f"[full] ndjson {data.size} bytes", flush=True)
    compiled = [gpu.compile_query(q, extensions=True) for q in THROUGHPUT_QUERIES]
    levels = max(c[1] for c in compiled)
    dev = gpu.DeviceBuffer.from_host(data)
    idx = gpu.build_index(dev, data.size, levels)
    results = gpu.run_query_batch(dev, data.size, idx, compiled, capacity=8)
    yield data, compiled, levels, idx, results
    idx.free()
    dev.free()


def test_ndjson_8gib_newline_index(ndjson_8gib, gpu):
    data, _c, _l, idx, _r = ndjson_8gib
    nl = idx.newline()
    assert nl[0] == 0
    assert nl.size == 1 + count_byte(data, 10)
    assert np.all(np.diff(nl[1:]) > 0)
    assert np.all(data[nl[1:]] == 10)


def test_ndjson_8gib_index_windows(ndjson_8gib, gpu, orc):
    data, _c, levels, idx, _r = ndjson_8gib
    nl = idx.newline()
    rng = np.random.default_rng(7)
    for i in rng.integers(0, nl.size - 40, 12):
        s = 0 if i == 0 else int(nl[i]) + 1
        e = int(nl[i + 32]) + 1
        w0, w1 = s // 64, (e - 1) // 64
        print(f"[full] window {s}..{e}", flush=True)
        os_, ol, onl = orc.build_index(data[s:e].tobytes(), levels)
        np.testing.assert_array_equal(onl[1:] + s, nl[i + 1:i + 33])
        lo = s - 64 * w0
        got = bits(gpu.d2h(idx.raw.d_string + 8 * w0, np.uint64, w1 - w0 + 1))
        np.testing.assert_array_equal(got[lo:lo + e - s], bits(os_)[:e - s])
        for k in range(levels):
            ptr = idx.raw.d_leveled + 8 * (k * idx.level_size + w0)
            got = bits(gpu.d2h(ptr, np.uint64, w1 - w0 + 1))
            np.testing.assert_array_equal(got[lo:lo + e - s], bits(ol[k])[:e - s])


def test_ndjson_8gib_query_sample(ndjson_8gib, gpu, orc):
    data, compiled, levels, idx, results = ndjson_8gib
    nl = idx.newline()
    rng = np.random.default_rng(11)
    recs = np.sort(rng.choice(nl.size - 1, SAMPLE, replace=False))
    starts = np.where(recs == 0, 0, nl[recs] + 1)
    ends = nl[recs + 1]
    lines = [data[a:b].tobytes() for a, b in zip(starts, ends)]
    mini = b"\n".join(lines) + b"\n"
    mstarts = np.cumsum([0] + [len(x) + 1 for x in lines[:-1]])
    oix = orc.build_index(mini, levels)
    for (ir, _d, slots), (spans, counts), q in zip(compiled, results, THROUGHPUT_QUERIES):
        print(f"[full] sample {q}", flush=True)
        got = relative(masked(spans[recs], None if counts is None else counts[recs]), starts)
        if counts is None:
            want = orc.query(mini, oix, ir, slots)[:SAMPLE]
        else:
            ws, wc = orc.query_ext(mini, oix, ir, spans.shape[1])
            np.testing.assert_array_equal(counts[recs], wc[:SAMPLE])
            want = masked(ws[:SAMPLE], wc[:SAMPLE])
        np.testing.assert_array_equal(got, relative(want, mstarts))


def test_ndjson_8gib_query_totals(ndjson_8gib):
    data, _c, _l, idx, results = ndjson_8gib
    records = idx.num_lines - 1
    (lang, _), (f0, _), (t_sp, t_ct), (id_sp, id_ct) = results
    # every record has user.lang and a first e element with "f"
    assert int((lang[:, 0, 0] >= 0).sum()) == records
    assert int((f0[:, 0, 0] >= 0).sum()) == records
    # "t" keys appear only in h elements, k == "v" only as '"k":"v","id":'
    assert int(t_ct.sum()) == count_pattern(data, b'{"t":')
    assert int(id_ct.sum()) == count_pattern(data, b',"k":"v","id":')
    assert int(id_ct.sum()) > 0


def test_array_1gib_items_sku(gpu, orc):
    data = gpu.generate(gpu.SCHEMA_BIG_ARRAY, 2_000_000, 0x15418001)
    print(f"[full] array {data.size} bytes", flush=True)
    query = "$[*].items[*].sku"
    compiled = gpu.compile_element_query(query)
    levels = compiled[1] + 1
    dev = gpu.DeviceBuffer.from_host(data)
    idx = gpu.build_index(dev, data.size, levels)
    bounds, m = gpu.array_bounds(idx)
    assert m == 2_000_001
    (spans, counts), = gpu.run_query_records(dev, data.size, idx, bounds, m, [compiled], 24)
    b = bounds.to_host(np.int64, m)
    assert int(counts.sum()) == count_pattern(data, b'"sku":')
    rng = np.random.default_rng(13)
    recs = np.sort(rng.choice(m - 1, SAMPLE, replace=False))
    starts = b[recs] + 1
    lines = [data[a:z].tobytes() for a, z in zip(starts, b[recs + 1])]
    mini = b"\n".join(lines) + b"\n"
    mstarts = np.cumsum([0] + [len(x) + 1 for x in lines[:-1]])
    ir, depth, _slots = gpu.compile_query("$" + query[4:], extensions=True)
    assert ir == compiled[0]
    oix = orc.build_index(mini, depth)
    ws, wc = orc.query_ext(mini, oix, ir, spans.shape[1])
    np.testing.assert_array_equal(counts[recs], wc[:SAMPLE])
    np.testing.assert_array_equal(relative(spans[recs], starts),
                                  relative(masked(ws[:SAMPLE], wc[:SAMPLE]), mstarts))
    idx.free()
    dev.free()


def test_ndjson_64mib_ab_exact(gpu, orc):
    # the plain 64 MiB configuration is small enough to compare whole
    data = gpu.generate(gpu.SCHEMA_NDJSON_2_4K, 64 << 20, 0x15418000)
    ir, depth, slots = gpu.compile_query("$.a.b")
    dev = gpu.DeviceBuffer.from_host(data)
    idx = gpu.build_index(dev, data.size, depth)
    got = gpu.run_query(dev, data.size, idx, ir, slots)
    s, lv, nl = orc.build_index(data, depth, 8)
    np.testing.assert_array_equal(idx.newline(), nl)
    np.testing.assert_array_equal(idx.string(), s)
    np.testing.assert_array_equal(idx.leveled(), lv)
    want = orc.query(data, (s, lv, nl), ir, slots, 8)
    np.testing.assert_array_equal(got, want)
    assert int((got[:, 0, 0] >= 0).sum()) > 0


def test_batch_10k_files_sample(gpu, orc):
    from test_gpu_batch_files import CAP, oracle_file
    schemas, targets, masks = plan(10000, 256 << 10, 1.5, 1 << 10, 256 << 20, 0x15418003)
    buf, off, sz = gpu.generate_files(schemas, targets, 0x15418003)
    print(f"[full] batch {int(sz.sum())} bytes", flush=True)
    compiled = [gpu.compile_query(q, extensions=True) for q in BATCH_QUERIES]
    b = gpu.FileBatch(buf, off, sz, masks, compiled, capacity=CAP, stream=None)
    b.launch()
    b.buf = buf
    np.testing.assert_array_equal(b.device_line_prefix(), b.line_prefix)
    # the device line counts FileBatch sized its buffers with, against the host
    host = np.array([np.count_nonzero(buf[int(o):int(o) + int(z)] == 10) + 1
                     for o, z in zip(off, sz)], dtype=np.int64)
    np.testing.assert_array_equal(b.line_counts, host)
    spans, counts = b.fetch()
    rng = np.random.default_rng(17)
    files = set(rng.choice(b.num_files, 30, replace=False).tolist())
    files.add(int(np.argmax(b.sizes)))
    for f in sorted(files):
        print(f"[full] file {f} bytes={int(b.sizes[f])}", flush=True)
        data, (os_, ol, onl), per_slot = oracle_file(orc, gpu, b, f)
        w0, w = int(b.offsets[f]) // 64, (int(b.sizes[f]) + 63) // 64
        np.testing.assert_array_equal(gpu.d2h(b.d_string.ptr + 8 * w0, np.uint64, w), os_)
        for k in range(b.levels):
            np.testing.assert_array_equal(
                gpu.d2h(b.d_leveled.ptr + 8 * (k * b.words + w0), np.uint64, w), ol[k])
        a, z = int(b.line_prefix[f]), int(b.line_prefix[f + 1])
        np.testing.assert_array_equal(gpu.d2h(b.d_newline.ptr + 8 * a, np.int64, z - a), onl)
        for k, (sp, ct) in enumerate(per_slot):
            np.testing.assert_array_equal(counts[k][a:z], ct)
            np.testing.assert_array_equal(spans[k][a:z], sp)
    # plain per-file launches give the same results for every file
    b.per_file_setup()
    b.launch_per_file()
    ps, pc = b.fetch(per_file=True)
    for k in range(2):
        np.testing.assert_array_equal(pc[k], counts[k])
        np.testing.assert_array_equal(ps[k], spans[k])
