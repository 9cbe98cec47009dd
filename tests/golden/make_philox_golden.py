"""Writes philox4x64_10.json: known-answer blocks of Philox4x64-10 taken
from numpy.random.Philox (numpy 2.2).  numpy advances the 256-bit counter
before producing a block, so Philox(counter=c - 1, key=k).random_raw(4) is
the block for counter c."""
import json
from pathlib import Path

import numpy as np
from numpy.random import Philox

CASES = [
    ([1, 0, 0, 0], [0, 0]),
    ([2, 0, 0, 0], [0, 0]),
    ([5, 7, 1, 0], [123, 0]),
    ([0, 1, 0, 0], [1, 0]),            # carry into word 1 (counter - 1 = [2^64-1, 0, 0, 0])
    ([12345, 3, 1, 0], [20260101, 0]),
    ([2**63 + 11, 200, 0, 0], [2**64 - 1, 0]),
    ([99, 2**40, 1, 0], [7, 2**63]),
    ([1, 1, 1, 1], [3, 5]),
]


def block(counter, key):
    c = [int(x) for x in counter]
    # subtract one from the 256-bit counter (little-endian words)
    val = sum(w << (64 * i) for i, w in enumerate(c)) - 1
    prev = [(val >> (64 * i)) & (2**64 - 1) for i in range(4)]
    bg = Philox(counter=np.array(prev, dtype=np.uint64), key=np.array(key, dtype=np.uint64))
    return [int(x) for x in bg.random_raw(4)]


def main():
    out = [{"counter": [str(x) for x in c], "key": [str(x) for x in k],
            "block": [str(x) for x in block(c, k)]} for c, k in CASES]
    # a consecutive stream from one generator, to pin the counter convention
    bg = Philox(counter=np.zeros(4, np.uint64), key=np.array([42, 0], np.uint64))
    stream = [str(int(x)) for x in bg.random_raw(12)]
    doc = {"source": f"numpy {np.__version__} numpy.random.Philox", "cases": out,
           "stream_key": ["42", "0"], "stream_first_counter": ["1", "0", "0", "0"], "stream": stream}
    Path(__file__).with_name("philox4x64_10.json").write_text(json.dumps(doc, indent=1) + "\n")


if __name__ == "__main__":
    main()
