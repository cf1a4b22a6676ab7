"""The oracle against every known answer in tests/golden/reference_kats.json."""
import json
import os

import pytest

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


def test_philox_vectors(O):
    for v in GOLD["philox4x32_10"]["vectors"]:
        assert O.philox_block(v["ctr"], v["key"]) == [int(x, 16) for x in v["out"]]


def test_cub_compaction_count(O):
    g = GOLD["cub_compaction"]
    L = O.lib()
    s = g["sqrt_s"] ** 2
    wmax = L.orc_c1_wmax_scan(g["wmax_scan_points"], s, g["alphaS"], g["cos_max"])
    assert L.orc_c1_unweight_compat(int(g["seed"], 16), g["candidates"], s, g["alphaS"], g["cos_max"], wmax,
                                    None, None) == g["accepted"]


def test_fused_gg2gg_numbers(O):
    g = GOLD["fused_gg2gg"]
    L = O.lib()
    assert L.orc_c1_simpson(2000000, 1e4, 0.118, 0.9, O.CONV_CUPYTHIA) == pytest.approx(g["simpson_mb"], abs=5e-11)
    assert L.orc_c1_cpumc_compat(int(g["cpumc_seed"], 16), g["cpumc_trials"], 1e4, 0.118, 0.9,
                                 O.CONV_CUPYTHIA) == pytest.approx(g["cpumc_mb"], abs=5e-11)


def test_reference_vegas_1d(O):
    """The reference's own 1-D VEGAS (oracle restatement of vegas.cu:40-124): eta 10.0% -> 76.0%,
    integral within 1.2e-5 of Simpson (to the printed digits)."""
    import ctypes as C
    g = GOLD["vegas_1d"]
    out = (C.c_double * 3)()
    conv = O.CONV_CUPYTHIA * 1e9
    O.lib().orc_vegas1d_compat(128, 256, 256, 2000, 12, 1e4, 0.118, 0.9, conv, out)
    ref = O.lib().orc_c1_simpson(2000000, 1e4, 0.118, 0.9, conv)
    assert round(out[0], 3) == g["eta_first"] and round(out[1], 3) == g["eta_last"]
    assert round(abs(out[2] - ref) / ref, 6) == g["relerr"]
