"""Shared multi-file batch workload for tests: seeded log-normal file sizes,
three record schemas, two queries per file."""
import numpy as np

BATCH_QUERIES = ["$.a.b", "$.a", "$.user.lang", "$.h[*].t", "$.entities.hashtags[*].text"]
# schema -> batch queries (bit q = BATCH_QUERIES[q])
SCHEMA_MASKS = {0: 0b00011, 2: 0b01100, 3: 0b10100}


def plan(num_files: int, median: int, sigma: float, lo: int, hi: int, seed: int):
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.exp(rng.normal(np.log(median), sigma, num_files)), lo, hi)
    schemas = rng.choice([0, 2, 3], num_files).astype(np.int32)
    masks = np.array([SCHEMA_MASKS[int(s)] for s in schemas], dtype=np.uint32)
    return schemas, sizes.astype(np.uint64), masks
