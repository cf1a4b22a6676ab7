"""The C ABI from a compiled host program (tests/c_client/mcint_client.cpp, built by build()):
integrate, unweight with compaction, the device event record and LHEF output through
include/mcint.h alone give the same results, bit for bit, as the same calls through the Python
layer."""
import json
import os
import subprocess

import pytest

import cupythia_amd as mi

pytestmark = pytest.mark.gpu

CLIENT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cupythia_amd", "_lib",
                      "mcint_client")


def test_c_client_matches_python_layer(ctx, tmp_path):
    if not os.path.exists(CLIENT):
        pytest.fail("the C-ABI client is not built (run __graft_entry__.build())")
    n, key = 1 << 20, 5
    path = str(tmp_path / "client.lhe")
    out = subprocess.run([CLIENT, str(n), str(key), path], check=True, capture_output=True, text=True, timeout=120)
    got = json.loads(out.stdout.strip().splitlines()[-1])
    phys = mi.physics_default(1)
    run = mi.make_run(phys, n, key=key)
    est = ctx.integrate(run)
    assert got["sigma_mb"] == est.sigmaGen() and got["sigma_err_mb"] == est.sigmaErr()
    ev = ctx.unweight_compact(run)
    assert got["n_acc"] == ev.n and got["wmax"] == ev.wmax
    rec = ctx.build_events(phys, ev, key=key)
    assert got["pdg_sum"] == int(rec.pdg.sum())
    assert got["e_sum"] == sum(rec.e.cpu().numpy().reshape(-1).tolist())     # the client's order
    ref = str(tmp_path / "python.lhe")
    ctx.write_lhe(ref, phys, ev, key=key)
    text = open(path).read()
    assert text.count("<event>") == ev.n and text == open(ref).read()
