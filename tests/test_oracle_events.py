"""Pins of the event-record oracle (oracle_events.c), CPU only.

(1) kinematics: four-momentum conservation, on-shell masses, (p1 + p2)^2 = x1 x2 S and
    (p1 - p3)^2 = tHat for every code, massless and massive;
(2) colour lines: every tag joins exactly one colour end and one anticolour end (an incoming
    colour counts as an outgoing anticolour), quarks carry colour only, antiquarks anticolour only,
    gluons both; flavour is conserved and each pair belongs to the code's flux list;
(3) colour-topology frequencies at fixed kinematics against forms independent of Pythia's
    expressions: g g -> g g splits as u^2 : t^2 : s^2 (Pythia's sigTS = (9/8)(s^4 + t^4 + u^4) /
    (s^2 t^2), likewise sigUS, sigTU, so the weights are proportional to 1/(s^2 t^2), 1/(s^2 u^2),
    1/(t^2 u^2)); the two-flow codes are mirror images under t <-> u (P_TS(t, u) + P_TS(u, t) = 1)
    and the t-channel flow dominates as t -> 0; identical quarks take the u-channel flow with the
    weight of the textbook u-channel term;
(4) incoming-pair frequencies at fixed (x1, x2) against xf1 xf2 times the textbook (ESW) bracket
    of the pair, and the new flavour of 116 uniform over nQuarkNew."""
import math

import numpy as np

S = 13600.0 ** 2


def kin(code, x1, x2, cth, mQ=0.0):
    s = x1 * x2 * S
    beta = math.sqrt(1 - 4 * mQ * mQ / s)
    t = mQ * mQ - 0.5 * s * (1 - beta * cth)
    return s, t


def events(O, code, n, x1=0.01, x2=0.003, cth=0.3, key=5, cfg=None, **kw):
    cfg = cfg or O.hadcfg_default(procMask=0x3FF)
    mQ = cfg.mQ[0] if code in (121, 122) else (cfg.mQ[1] if code in (123, 124) else 0.0)
    s, t = kin(code, x1, x2, cth, mQ)
    idx = np.arange(n, dtype=np.uint64) * 7 + 3
    return O.build_events(cfg, key, idx, np.full(n, code), np.full(n, s), np.full(n, t), np.full(n, x1),
                          np.full(n, x2), **kw), s, t


def random_events(O, n, seed=1):
    rng = np.random.default_rng(seed)
    cfg = O.hadcfg_default(procMask=0x3FF)
    codes = rng.choice([111, 112, 113, 114, 115, 116, 121, 122, 123, 124], n)
    x1 = np.exp(rng.uniform(-9, -0.05, n))
    x2 = np.exp(rng.uniform(-9, -0.05, n))
    s = x1 * x2 * S
    m = np.where(np.isin(codes, (121, 122)), 1.5, np.where(np.isin(codes, (123, 124)), 4.8, 0.0))
    ok = s > 4 * m * m * 1.01
    codes, x1, x2, s, m = codes[ok], x1[ok], x2[ok], s[ok], m[ok]
    beta = np.sqrt(1 - 4 * m * m / s)
    c = rng.uniform(-0.99, 0.99, codes.size)
    t = m * m - 0.5 * s * (1 - beta * c)
    idx = rng.integers(0, 2 ** 40, codes.size).astype(np.uint64)
    return cfg, O.build_events(cfg, 11, idx, codes, s, t, x1, x2), codes, s, t, x1, x2, m


def mdot(a, b):
    return a[..., 3] * b[..., 3] - a[..., 0] * b[..., 0] - a[..., 1] * b[..., 1] - a[..., 2] * b[..., 2]


def test_kinematics_invariants(O):
    cfg, r, codes, s, t, x1, x2, m = random_events(O, 20000)
    p = r["p"]
    assert np.all(r["n_part"] == 4)
    etot = p[:, 0, 3] + p[:, 1, 3]
    bal = p[:, 2] + p[:, 3] - p[:, 0] - p[:, 1]
    assert np.max(np.abs(bal) / etot[:, None]) < 1e-15
    for k in range(4):
        m2 = mdot(p[:, k], p[:, k])
        assert np.max(np.abs(m2 - r["m"][:, k] ** 2) / etot ** 2) < 1e-12
    assert np.all(r["m"][:, 2] == m) and np.all(r["m"][:, :2] == 0)
    sk = mdot(p[:, 0] + p[:, 1], p[:, 0] + p[:, 1])
    assert np.max(np.abs(sk - s) / s) < 1e-12
    q = p[:, 0] - p[:, 2]
    assert np.max(np.abs(mdot(q, q) - t) / s) < 1e-11          # the record's tHat, sign of cos theta
    pT2 = p[:, 2, 0] ** 2 + p[:, 2, 1] ** 2
    assert np.max(np.abs(r["scale"] ** 2 - (pT2 + m * m)) / s) < 1e-11
    assert np.all(r["status"] == [-21, -21, 23, 23]) and np.all(r["mo1"] == [-1, -1, 0, 0])
    assert np.all(r["mo2"] == [-1, -1, 1, 1]) and np.all(r["d1"] == [2, 2, -1, -1])


def colour_ok(pdg, st, col, acol):
    ends = {}
    for k in range(4):
        inc = st[k] < 0
        for tag, kind in ((col[k], "a" if inc else "c"), (acol[k], "c" if inc else "a")):
            if tag:
                ends.setdefault(tag, []).append(kind)
        q = pdg[k]
        if q == 21:
            if not (col[k] and acol[k] and col[k] != acol[k]):
                return False
        elif (q > 0) != (col[k] > 0 and acol[k] == 0) or (q < 0) != (acol[k] > 0 and col[k] == 0):
            return False
    return all(sorted(v) == ["a", "c"] for v in ends.values())


def test_colour_lines_and_flavours(O):
    cfg, r, codes, *_ = random_events(O, 6000, seed=2)
    flux = {111: "gg", 112: "gg", 121: "gg", 123: "gg", 113: "qg", 114: "qq", 115: "qqbar", 116: "qqbar",
            122: "qqbar", 124: "qqbar"}
    for i, code in enumerate(codes):
        pdg = r["pdg"][i]
        assert colour_ok(pdg, r["status"][i], r["col"][i], r["acol"][i]), (code, pdg, r["col"][i], r["acol"][i])
        a, b = pdg[0], pdg[1]
        kind = flux[code]
        if kind == "gg":
            assert a == b == 21
        elif kind == "qg":
            assert (a == 21) != (b == 21) and 0 < abs(a if b == 21 else b) <= 5
        elif kind == "qq":
            assert a != 21 and b != 21
        else:
            assert a == -b and a != 21
        nq = lambda ids: np.bincount([abs(x) for x in ids if x != 21], weights=[np.sign(x) for x in ids if x != 21],
                                     minlength=6) if any(x != 21 for x in ids) else np.zeros(6)
        assert np.array_equal(nq(pdg[:2]), nq(pdg[2:])) or code in (112, 116, 121, 122, 123, 124)
        if code in (121, 122):
            assert sorted(pdg[2:]) == [-4, 4]
        if code in (123, 124):
            assert sorted(pdg[2:]) == [-5, 5]


def connected(r, i, j):
    """Per event: do partons i and j (scalars or per-event index arrays) share a colour tag?"""
    rows = np.arange(r["col"].shape[0])
    ci = np.stack([r["col"][rows, i], r["acol"][rows, i]], 1)
    cj = np.stack([r["col"][rows, j], r["acol"][rows, j]], 1)
    hit = np.zeros(rows.size, dtype=bool)
    for a in range(2):
        for b in range(2):
            hit |= (ci[:, a] == cj[:, b]) & (ci[:, a] != 0)
    return hit


def within(f, p, n, k=4.5):
    return abs(f - p) <= k * math.sqrt(p * (1 - p) / n) + 1e-12


def test_gg2gg_topologies_closed_form(O):
    n = 40000
    for cth in (-0.7, 0.1, 0.8):
        r, s, t = events(O, 111, n, cth=cth)
        u = -s - t
        tot = s * s + t * t + u * u
        tu = ~connected(r, 0, 1)                        # p1, p2 unconnected: t-u flow
        ts = ~tu & connected(r, 0, 2)
        us = ~tu & ~ts
        for frac, p in ((ts.mean(), u * u / tot), (us.mean(), t * t / tot), (tu.mean(), s * s / tot)):
            assert within(frac, p, n), (cth, frac, p)
        # orientation: colour of parton 1 is tag 501 or its anticolour is (the swapped flow)
        assert within((r["col"][:, 0] == 501).mean(), 0.5, n)


def test_two_flow_codes_mirror_and_limits(O):
    n = 40000
    for code in (112, 115, 121, 123):
        # parton 3 colour-connected to parton 1: the t-channel flow; t <-> u mirrors the two flows
        fr = [connected(events(O, code, n, cth=c)[0], 0, 2).mean() for c in (0.35, -0.35)]
        assert abs(fr[0] + fr[1] - 1.0) < 5 * math.sqrt(0.5 / n), (code, fr)
        assert abs(fr[0] - 0.5) > 0.05, (code, fr)            # the point is not a symmetric one
        r, s, t = events(O, code, n, x1=0.2, x2=0.1, cth=0.9999)   # t -> 0
        assert connected(r, 0, 2).mean() > 0.9, code
    # q g -> q g as u -> 0: sigTU dominates, where the incoming quark's colour line runs to the
    # outgoing gluon, not to the incoming gluon (sigTS)
    r, s, t = events(O, 113, n, x1=0.2, x2=0.1, cth=-0.9999)
    q_in = np.where(r["pdg"][:, 0] != 21, 0, 1)
    assert (~connected(r, q_in, 1 - q_in)).mean() > 0.9
    r, s, t = events(O, 113, n, x1=0.2, x2=0.1, cth=0.9999)   # t -> 0: both weights grow as 1/t^2
    q_in = np.where(r["pdg"][:, 0] != 21, 0, 1)
    u = -s - t
    pts = (u * u / (t * t) - 4 / 9 * u / s) / (u * u / (t * t) - 4 / 9 * u / s + s * s / (t * t) - 4 / 9 * s / u)
    assert abs(connected(r, q_in, 1 - q_in).mean() - pts) < 0.02


def test_identical_quarks_u_channel_weight(O):
    n = 200000
    for cth in (0.5, -0.2):
        r, s, t = events(O, 114, n, cth=cth)
        u = -s - t
        same = r["pdg"][:, 0] == r["pdg"][:, 1]
        ns = int(same.sum())
        assert ns > 2000
        sigT, sigU = (s * s + u * u) / (t * t), (s * s + t * t) / (u * u)
        # default flow for q q: p3 takes p2's colour; the u-channel flow gives p3 p1's colour
        alt = connected(r, 0, 2)[same]
        assert within(alt.mean(), sigU / (sigT + sigU), ns), (cth, alt.mean(), sigU / (sigT + sigU))


def test_incoming_pair_frequencies(O):
    cfg = O.hadcfg_default(procMask=0x3FF)
    x1, x2, n = 0.02, 0.005, 200000
    xf = lambda i, x: O.lib().orc_toypdf_xf(__import__("ctypes").byref(cfg.pdf), i, x)
    # q g -> q g: P(parton 1 is the gluon)
    r, s, t = events(O, 113, n, x1=x1, x2=x2, cfg=cfg)
    qs = [q for q in range(-5, 6) if q]
    wq1 = sum(xf(q, x1) * xf(21, x2) for q in qs)
    wg1 = sum(xf(21, x1) * xf(q, x2) for q in qs)
    assert within((r["pdg"][:, 0] == 21).mean(), wg1 / (wq1 + wg1), n)
    # q q' -> q q': categories weighted by xf xf and the textbook bracket of the pair
    r, s, t = events(O, 114, n, x1=x1, x2=x2, cfg=cfg, cth=0.4)
    u = -s - t
    W = {}
    for a in qs:
        for b in qs:
            W[(a, b)] = xf(a, x1) * xf(b, x2) * O.lib().orc_bracket_textbook(114, s, t, u, a, b, 3)
    tot = sum(W.values())
    a, b = r["pdg"][:, 0], r["pdg"][:, 1]
    for name, sel, pred in (("same", a == b, lambda p: p[0] == p[1]), ("opp", a == -b, lambda p: p[0] == -p[1]),
                            ("uu", (a == 2) & (b == 2), lambda p: p == (2, 2)),
                            ("qbar qbar", (a < 0) & (b < 0), lambda p: p[0] < 0 and p[1] < 0)):
        p = sum(w for k, w in W.items() if pred(k)) / tot
        assert within(sel.mean(), p, n), (name, sel.mean(), p)
    # q qbar -> q' qbar': the new flavour uniform over nQuarkNew = 3, sign following parton 1
    r, s, t = events(O, 116, n, cfg=cfg)
    new = np.abs(r["pdg"][:, 2])
    for k in (1, 2, 3):
        assert within((new == k).mean(), 1 / 3, n)
    assert np.all(np.sign(r["pdg"][:, 2]) == np.sign(r["pdg"][:, 0]))


def test_partonic_and_unknown_codes(O):
    cfg = O.hadcfg_default()
    r = O.build_events(cfg, 9, np.arange(4, dtype=np.uint64), [111, 111, 999, 113], [1e4] * 4,
                       [-2e3, -7e3, -5e3, -5e3], partonic=True, sHat_partonic=1e4)
    assert list(r["n_part"]) == [4, 4, 0, 0]
    assert np.all(r["p"][:2, :, 3] == 50.0)
    assert np.all(r["pdg"][:2] == 21)
