"""The stm_test command line (reference-compatible text op files) against the oracle."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CLI = Path(__file__).resolve().parent.parent / "eirene_amd" / "stm_test"
LINE = re.compile(r"key: (\d+)\t operator (\d+)\t ans: (-1|\d+)(?:\t vals: (\d+))?")


def write_ops(path, rng, n, domain, warehouse):
    """Writes n ops; returns (keys, vals-as-offsets, ops) the tree will see."""
    keys, vals, ops = [], [], []
    with open(path, "w") as f:
        for _ in range(n):
            k = int(rng.choice(domain))
            r = rng.random()
            if r < 0.45:
                v = int(rng.integers(0, 1 << 62))
                f.write(f"P {k} {v}\n")
                keys.append(k), vals.append(len(warehouse)), ops.append(0)
                warehouse.append(v)
            elif r < 0.95:
                f.write(f"R {k}\n")
                keys.append(k), vals.append(0), ops.append(1)
            else:
                f.write(f"G {k} {k + 10}\n")  # RANGE runs with GET semantics on this path
                keys.append(k), vals.append(k + 10), ops.append(3)
        f.write("garbage line\n")
    return (np.array(keys, dtype=np.uint64), np.array(vals, dtype=np.uint64),
            np.array(ops, dtype=np.uint8))


@pytest.mark.parametrize("method", ["38", "learned", "index"])
def test_cli_matches_oracle(tmp_path, oracle_mod, method):
    assert CLI.exists(), "build the package first (python -m eirene_amd.build)"
    rng = np.random.default_rng(11)
    domain = rng.integers(0, 1 << 40, size=3000)
    warehouse = []
    files, batches = [], []
    for b in range(3):
        p = tmp_path / f"ops{b}.txt"
        batches.append(write_ops(p, rng, 5000, domain, warehouse))
        files += ["-i", str(p)]
    res = subprocess.run([str(CLI), "-t", "-v", method, *files], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    got = [LINE.match(l) for l in res.stdout.splitlines() if l.startswith("key: ")]
    assert all(got) and len(got) == 15000
    assert res.stdout.count("total gpu exec time") == 3
    assert res.stdout.count("total traversal_steps: ") == 3  # the reference's ABORT_COUNT totals
    o = oracle_mod.Oracle()
    i = 0
    for keys, vals, ops in batches:
        ev, ef = o.run_batch(keys, vals, ops)
        for j in range(keys.size):
            m = got[i]
            i += 1
            assert int(m.group(1)) == int(keys[j]) and int(m.group(2)) == int(ops[j])
            if ef[j]:
                assert int(m.group(3)) == int(ev[j])
                assert int(m.group(4)) == warehouse[int(ev[j])]
            else:
                assert m.group(3) == "-1"


def test_cli_rejects_input_before_method(tmp_path):
    p = tmp_path / "x.txt"
    p.write_text("R 1\n")
    res = subprocess.run([str(CLI), "-i", str(p)], capture_output=True, text=True, timeout=60)
    assert res.returncode == 2
