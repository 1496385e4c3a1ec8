"""GPU parity of multi-file batches: one index launch + one query launch over
many files equals plain per-file launches and the oracle run on each file
alone; graph replays equal the eager run."""
import numpy as np
import pytest

from batchdata import BATCH_QUERIES, plan

pytestmark = pytest.mark.gpu
CAP = 8


@pytest.fixture(scope="module")
def batch(gpu):
    schemas, targets, masks = plan(120, 12288, 1.5, 1024, 1 << 19, 0x15418003)
    buf, off, sz = gpu.generate_files(schemas, targets, 0x15418003)
    compiled = [gpu.compile_query(q, extensions=True) for q in BATCH_QUERIES]
    b = gpu.FileBatch(buf, off, sz, masks, compiled, capacity=CAP, stream=None)
    b.launch()
    b.buf = buf
    return b


def oracle_file(orc, gpu, b, f):
    lo, n = int(b.offsets[f]), int(b.sizes[f])
    data = bytes(b.buf[lo:lo + n])
    ix = orc.build_index(data, b.levels)
    per_slot = []
    for q in range(len(BATCH_QUERIES)):
        if not (int(b.masks[f]) >> q) & 1:
            continue
        ir, _depth, _slots = b.compiled[q]
        if gpu.uses_extension_opcodes(ir):
            sp, ct = orc.query_ext(data, ix, ir, CAP)
            sp = sp.copy()
            sp[np.arange(CAP)[None, :] >= np.minimum(ct, CAP)[:, None]] = -1
        else:
            sp = orc.query(data, ix, ir, CAP)
            ct = (sp[:, :, 0] >= 0).sum(axis=1)
        per_slot.append((sp, ct))
    return data, ix, per_slot


def test_line_prefix(batch):
    np.testing.assert_array_equal(batch.device_line_prefix(), batch.line_prefix)


def test_device_line_counts_equal_host_count(batch):
    # gpj_batch_line_counts (the sizes FileBatch uses) against a host count of
    # each file's newlines
    want = np.array([np.count_nonzero(batch.buf[int(o):int(o) + int(z)] == 10) + 1
                     for o, z in zip(batch.offsets, batch.sizes)], dtype=np.int64)
    np.testing.assert_array_equal(batch.line_counts, want)
    assert batch.line_prefix[0] == 0 and batch.total_lines == int(want.sum())


def test_batched_equals_oracle_per_file(gpu, orc, batch):
    spans, counts = batch.fetch()
    string = batch.d_string.to_host(np.uint64, batch.words)
    leveled = batch.d_leveled.to_host(np.uint64, batch.words * batch.levels).reshape(
        batch.levels, batch.words)
    newline = batch.d_newline.to_host(np.int64, batch.total_lines)
    for f in range(batch.num_files):
        print(f"[batch] file {f} bytes={int(batch.sizes[f])}", flush=True)
        data, (os_, ol, onl), per_slot = oracle_file(orc, gpu, batch, f)
        w0, w = int(batch.offsets[f]) // 64, (int(batch.sizes[f]) + 63) // 64
        np.testing.assert_array_equal(string[w0:w0 + w], os_)
        np.testing.assert_array_equal(leveled[:, w0:w0 + w], ol)
        a, z = int(batch.line_prefix[f]), int(batch.line_prefix[f + 1])
        np.testing.assert_array_equal(newline[a:z], onl)
        for k, (sp, ct) in enumerate(per_slot):
            np.testing.assert_array_equal(counts[k][a:z], ct)
            np.testing.assert_array_equal(spans[k][a:z], sp)
        for k in range(len(per_slot), 2):
            assert not counts[k][a:z].any()


def test_per_file_launches_equal_batched(batch):
    batch.per_file_setup()
    batch.launch_per_file()
    ps, pc = batch.fetch(per_file=True)
    bs, bc = batch.fetch()
    for k in range(2):
        np.testing.assert_array_equal(pc[k], bc[k])
        np.testing.assert_array_equal(ps[k], bs[k])


def test_graph_replays_equal_eager(gpu, batch):
    eager = batch.fetch()
    stream = gpu.new_stream()
    old = batch.stream
    batch.stream = stream
    try:
        batch.capture()
        for rep in range(3):
            print(f"[graph] replay {rep}", flush=True)
            batch.launch_graph()
            got = batch.fetch()
            for k in range(2):
                np.testing.assert_array_equal(got[1][k], eager[1][k])
                np.testing.assert_array_equal(got[0][k], eager[0][k])
            # eager work between replays
            batch.launch()
            gpu.check(gpu.lib().gpj_stream_sync(stream))
    finally:
        batch.release_graph()
        batch.stream = old


def test_launch_shape_does_not_change_results(batch):
    # records per wave and the grid size the autotuner picks from
    want = batch.fetch()
    old = (batch.grid_records, batch.per_wave)
    try:
        for per_wave, grid in ((16, 256 * 4 * 16), (32, batch.total_lines), (64, 256 * 8 * 64)):
            print(f"[batch] per_wave={per_wave} grid={grid}", flush=True)
            batch.per_wave, batch.grid_records = per_wave, min(grid, batch.total_lines)
            batch.launch()
            got = batch.fetch()
            for k in range(2):
                np.testing.assert_array_equal(got[1][k], want[1][k])
                np.testing.assert_array_equal(got[0][k], want[0][k])
    finally:
        batch.grid_records, batch.per_wave = old
        batch.launch()
