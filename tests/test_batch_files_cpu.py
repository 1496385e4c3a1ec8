"""Host side of multi-file batches: the generator's layout and the tile map's
validation (no device needed)."""
import numpy as np
import pytest

from batchdata import plan


def test_generated_layout(gpj):
    schemas, targets, _ = plan(60, 8192, 1.5, 1024, 1 << 20, 7)
    buf, off, sz = gpj.generate_files(schemas, targets, 0x15418003)
    assert (off % gpj.BATCH_ALIGN == 0).all()
    assert (np.diff(off.astype(np.int64)) >= sz[:-1].astype(np.int64)).all()
    for f in range(60):
        body = bytes(buf[int(off[f]):int(off[f] + sz[f])])
        assert body.endswith(b"\n") and body.startswith(b"{")
        assert sz[f] <= targets[f] or body.count(b"\n") == 1
        end = int(off[f + 1]) if f + 1 < 60 else buf.size
        assert not buf[int(off[f] + sz[f]):end].any()  # zero padding
    again = gpj.generate_files(schemas, targets, 0x15418003)
    assert (again[0] == buf).all()


def make_files(gpj, spec):
    files = (gpj.BatchFileT * len(spec))()
    for i, (o, b, m) in enumerate(spec):
        files[i] = gpj.BatchFileT(o, b, 0, 4, m, 0)
    return files


def test_tile_map(gpj):
    A = gpj.BATCH_ALIGN
    files = make_files(gpj, [(0, 100, 1), (A, 3 * A, 3), (5 * A, 1, 2)])
    out = np.zeros(6, dtype=np.uint32)
    gpj.check(gpj.lib().gpj_batch_tile_map(files, 3, 6 * A, out.ctypes.data))
    assert out.tolist() == [0, 1, 1, 1, 1, 2]


@pytest.mark.parametrize("spec", [
    [(0, 0, 1)],                      # empty file
    [(100, 10, 1)],                   # misaligned
    [(0, 20000, 1), (16384, 10, 1)],  # overlap
    [(0, 10, 0b111)],                 # three queries
    [(0, 70000, 1)],                  # beyond the buffer
])
def test_tile_map_rejects(gpj, spec):
    files = make_files(gpj, spec)
    out = np.zeros(8, dtype=np.uint32)
    st = gpj.lib().gpj_batch_tile_map(files, len(spec), 4 * gpj.BATCH_ALIGN, out.ctypes.data)
    assert st == gpj.ERR_INVALID_ARGUMENT
