"""The stm_test command line (reference-compatible text op files) against the oracle.

Keys and values are signed, as the reference's Data_Collector<int, int> reads them
(src/input-helper.h:70,136-195): the library sees each key with its sign bit flipped, which
keeps signed order.  "G a b" ops return the pairs with a <= key <= b before the op."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CLI = Path(__file__).resolve().parent.parent / "eirene_amd" / "stm_test"
LINE = re.compile(r"key: (-?\d+)\t operator (\d+)\t ans: (-1|\d+)(?:\t vals: (-?\d+))?$")
PAIR = re.compile(r"pair: (-?\d+)\t vals: (-?\d+)$")
SIGN = 1 << 63


def mapped(k):
    """The library's key for signed key k (sign bit flipped)."""
    return (int(k) & (2 ** 64 - 1)) ^ SIGN


def unmapped(u):
    x = int(u) ^ SIGN
    return x - 2 ** 64 if x >= SIGN else x


def write_ops(path, rng, n, domain, warehouse):
    """Writes n ops; returns the signed ops as (type, key, value or end key) tuples."""
    out = []
    with open(path, "w") as f:
        for _ in range(n):
            k = int(rng.choice(domain))
            r = rng.random()
            if r < 0.45:
                v = int(rng.integers(-(1 << 62), 1 << 62))
                f.write(f"P {k} {v}\n")
                out.append(("P", k, len(warehouse)))
                warehouse.append(v)
            elif r < 0.95:
                f.write(f"R {k}\n")
                out.append(("R", k, 0))
            else:
                b = k + int(rng.integers(-5, 70000))  # some empty, some wider than one scan
                f.write(f"G {k} {b}\n")
                out.append(("G", k, b))
        f.write("garbage line\n")
    return out


@pytest.mark.parametrize("method", ["38", "learned", "index"])
def test_cli_matches_oracle(tmp_path, oracle_mod, method):
    assert CLI.exists(), "build the package first (python -m eirene_amd.build)"
    rng = np.random.default_rng(11)
    domain = np.concatenate([rng.integers(-(1 << 40), 1 << 40, size=3000),
                             np.arange(-200, 200)])  # negative keys around zero too
    warehouse = []
    files, batches = [], []
    for b in range(3):
        p = tmp_path / f"ops{b}.txt"
        batches.append(write_ops(p, rng, 5000, domain, warehouse))
        files += ["-i", str(p)]
    res = subprocess.run([str(CLI), "-t", "-v", method, *files], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    lines = [l for l in res.stdout.splitlines() if l.startswith(("key: ", "pair: "))]
    assert res.stdout.count("total gpu exec time") == 3
    for line in ("total working num: ", "total roll back times: ", "total traversal_steps: ",
                 "total leaf traversal_steps: ", "average traversal_steps: "):
        assert res.stdout.count(line) == 3  # the reference's ABORT_COUNT totals, per file
    o = oracle_mod.Oracle()
    i = 0
    n_pairs = 0
    for ops in batches:
        for typ, k, v in ops:
            m = LINE.match(lines[i])
            i += 1
            assert m, lines[i - 1]
            assert int(m.group(1)) == k
            if typ == "G":
                # the pairs a..b of the state before the op, from the oracle's scan replay
                assert int(m.group(2)) == 3
                a, z = mapped(k), mapped(v)
                want = []
                if z >= a:
                    L = min(z - a + 1, 65536)
                    start = a
                    while True:
                        cnt, _, sk, sv = o.run_batch_scan(np.array([start], dtype=np.uint64),
                                                          np.array([L], dtype=np.uint64),
                                                          np.array([3], dtype=np.uint8))
                        c = int(cnt[0])  # pairs the scan found (the area beyond stays zero)
                        hi = min(z, start + L - 1)  # this chunk's key interval
                        want += [(int(x), int(y)) for x, y in zip(sk[:c], sv[:c]) if int(x) <= hi]
                        if z - start + 1 <= 65536:
                            break
                        start += 65536
                        L = min(z - start + 1, 65536)
                assert int(m.group(3)) == len(want)
                for kk, vv in want:
                    pm = PAIR.match(lines[i])
                    i += 1
                    assert pm and int(pm.group(1)) == unmapped(kk)
                    assert int(pm.group(2)) == warehouse[vv]
                n_pairs += len(want)
                continue
            op = 0 if typ == "P" else 1
            assert int(m.group(2)) == op
            ev, ef = o.run_batch(np.array([mapped(k)], dtype=np.uint64),
                                 np.array([v], dtype=np.uint64), np.array([op], dtype=np.uint8))
            if ef[0]:
                assert int(m.group(3)) == int(ev[0])
                assert int(m.group(4)) == warehouse[int(ev[0])]
            else:
                assert m.group(3) == "-1"
    assert i == len(lines) and n_pairs > 0
    # the sign-bit flip keeps signed order: the oracle's u64 order of mapped keys is signed order
    keys_in_order = [unmapped(x) for x in o.dump()[0]]
    assert keys_in_order == sorted(keys_in_order) and min(keys_in_order) < 0


def test_cli_rejects_input_before_method(tmp_path):
    p = tmp_path / "x.txt"
    p.write_text("R 1\n")
    res = subprocess.run([str(CLI), "-i", str(p)], capture_output=True, text=True, timeout=60)
    assert res.returncode == 2
