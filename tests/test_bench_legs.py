"""bench.py's auxiliary legs that run in a process of their own: a failed leg is reported in its field of the
result line, not raised (the child here fails on its own arguments, before it touches a device)."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_failed_leg_process_is_reported(monkeypatch):
    import bench

    monkeypatch.setattr(sys, "argv", ["bench.py", "--serve-requests", "not-a-number"])
    res = bench.leg_in_child("serve")
    assert set(res) == {"error"} and "status 2" in res["error"]
