"""The committed traffic profile names the kernel bench.py's roofline looks up (CPU only)."""
import json
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_traffic_profile_has_the_value_route_lookup():
    prof = json.loads((ROOT / "profiles" / "hbm_traffic_gamma_2p26_batch_2p22.json").read_text())
    tuned = json.loads((ROOT / "eirene_amd" / "tuning" / "gamma_2p26_batch_2p22.json").read_text())["best"]
    kname = "k_lookup<%d,%d,true>" % (tuned["fanout"], tuned["lanes_per_op"])
    keys = prof["routes"]["learned_tuned"]
    assert any(k.startswith(kname) for k in keys), (kname, list(keys))
    for k in keys:
        assert keys[k]["read_bytes"] > 0 or keys[k]["dispatches"] >= 1
    assert abs(prof["scale"]["WRITE_SIZE"] - 1.0) < 0.05 and abs(prof["scale"]["FETCH_SIZE"] - 2.0) < 0.05
