"""LHEF output of compacted events: structure, four-momentum conservation, masslessness,
consistency with the compacted record, flavour content of each process's flux list and colour-flow
validity (every colour tag connects exactly one colour end and one anticolour end)."""
import ctypes
import math
from collections import Counter

import pytest

import cupythia_amd as mi

pytestmark = pytest.mark.gpu


def parse(path):
    lines = open(path).read().split("\n")
    assert lines[0].startswith("<LesHouchesEvents")
    i = lines.index("<init>")
    beam = lines[i + 1].split()
    nproc = int(beam[9])
    procs = [lines[i + 2 + k].split() for k in range(nproc)]
    events = []
    k = i + 2 + nproc
    while k < len(lines):
        if lines[k] == "<event>":
            head = lines[k + 1].split()
            parts = [lines[k + 2 + j].split() for j in range(int(head[0]))]
            events.append((head, parts))
            k += 2 + int(head[0])
        else:
            k += 1
    return beam, procs, events


def colour_ok(parts):
    ends = Counter()
    for p in parts:
        st, c, a = int(p[1]), int(p[4]), int(p[5])
        # incoming colour behaves as an outgoing anticolour and vice versa
        if c:
            ends[(c, "c" if st == 1 else "a")] += 1
        if a:
            ends[(a, "a" if st == 1 else "c")] += 1
    tags = {t for (t, _) in ends}
    return all(ends[(t, "c")] == 1 and ends[(t, "a")] == 1 for t in tags)


def test_lhe_hadronic(ctx, O, tmp_path):
    phys = mi.physics_default(1, proc_mask=0x3F)
    run = mi.make_run(phys, 1 << 18, key=21, counter0=99)
    ev = ctx.unweight_compact(run)
    assert ev.n > 100
    path = str(tmp_path / "events.lhe")
    ctx.write_lhe(path, phys, ev, key=21)
    beam, procs, events = parse(path)
    assert beam[0] == "2212" and float(beam[2]) == pytest.approx(6800.0)
    assert len(events) == ev.n
    assert {int(p[3]) for p in procs} <= {111, 112, 113, 114, 115, 116}
    sH = ev.sHat.cpu().numpy()
    pid = ev.pid.cpu().numpy()
    flux = {111: "gg", 112: "gg", 113: "qg", 114: "qq", 115: "qqbar", 116: "qqbar"}
    # unweighted events: one weight, sigma_tot / N in pb; alphaEM(MZ); alphaS at the event's scale
    xw = ev.estimate.sigmaGen() * 1e9 / ev.n
    cfg = O.hadcfg_default(procMask=0x3F)
    for n, (head, parts) in enumerate(events):
        assert int(head[0]) == 4 and int(head[1]) == pid[n]
        assert float(head[2]) == pytest.approx(xw, rel=1e-10) and float(head[4]) == pytest.approx(0.00781751, rel=1e-10)
        scale = float(head[3])
        assert float(head[5]) == pytest.approx(O.lib().orc_alphas_eval(ctypes.byref(cfg.as_), scale * scale), rel=1e-9)
        p = [[float(x) for x in q[6:10]] for q in parts]
        tot = [p[2][j] + p[3][j] - p[0][j] - p[1][j] for j in range(4)]
        scale = max(abs(x) for x in p[0] + p[1])
        assert max(abs(x) for x in tot) < 1e-9 * scale
        for q in p:
            m2 = q[3] ** 2 - q[0] ** 2 - q[1] ** 2 - q[2] ** 2
            assert abs(m2) < 1e-8 * q[3] ** 2
        shat = (p[0][3] + p[1][3]) ** 2 - (p[0][2] + p[1][2]) ** 2
        assert shat == pytest.approx(sH[n], rel=2e-6)          # the record's FP32 kinematics
        ids = [int(q[0]) for q in parts]
        kind = flux[pid[n]]
        if kind == "gg":
            assert ids[0] == ids[1] == 21
        elif kind == "qg":
            assert (ids[0] == 21) != (ids[1] == 21)
        elif kind == "qq":
            assert ids[0] != 21 and ids[1] != 21
        else:
            assert ids[0] == -ids[1] and ids[0] != 21
        assert sum(1 for i in ids[:2] if i != 21) - sum(1 for i in ids[2:] if i != 21) in (0, -2, 2)
        assert colour_ok(parts), parts


def test_lhe_partonic_reference_events(ctx, O, tmp_path):
    """The reference's kernel-10 events (gg -> gg at sqrt(s) = 100 GeV) as LHEF."""
    L = O.lib()
    wmax = L.orc_c1_wmax_scan(200000, 1e4, 0.118, 0.9) * O.CONVERT2MB * 0.9 * 1e4
    run = mi.make_run(mi.physics_default(0), 1 << 16, precision=mi.MI_PREC_FP64, rng=mi.MI_RNG_SPLITMIX_COMPAT,
                      key=0xC0B0)
    ev = ctx.unweight_compact(run, wmax=wmax)
    path = str(tmp_path / "gg.lhe")
    ctx.write_lhe(path, mi.physics_default(0), ev, key=0xC0B0)
    _, procs, events = parse(path)
    assert len(events) == ev.n and [int(p[3]) for p in procs] == [111]
    for head, parts in events:
        e = [float(q[9]) for q in parts]
        assert all(x == pytest.approx(50.0, rel=1e-12) for x in e)
        assert colour_ok(parts)
