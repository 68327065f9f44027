"""The random-gather probe (tests/gather_probe.hip) runs on the GPU and
reports every load form.  Its rocprofv3 request-size counters are what the
gather-fold's line-size figure (128 bytes per random gather) rests on."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE = os.path.join(ROOT, "tests", "gather_probe")


@pytest.mark.gpu
def test_gather_probe_runs_every_form():
    if not os.path.exists(PROBE):
        pytest.fail("tests/gather_probe is not built (run __graft_entry__.build())")
    r = subprocess.run([PROBE], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [ln for ln in r.stdout.splitlines() if "GB/s" in ln]
    assert len(lines) == 28 and r.stdout.strip().endswith("done"), r.stdout
    # every random-gather form and array size moves its samples at a
    # positive, finite rate, and so do both scatter forms
    for ln in lines:
        rate = float(ln.split("as 128-byte lines")[1].split()[0])
        assert 0 < rate < 1e6, ln
    scat = [ln for ln in r.stdout.splitlines() if "G samples/s" in ln]
    assert len(scat) == 2, r.stdout
    for ln in scat:
        assert 0 < float(ln.split("us")[-1].split()[0]) < 1e4, ln
