"""Pins of the oracle's ERI engines at high angular momentum (l = 4..6),
where the 40-digit McMurchie-Davidson vectors of md_eri.json stop:

- independent McMurchie-Davidson blocks in x86 extended precision
  (tests/golden/md_eri_high.json, written by gen_md_ld.py, which first
  reproduces md_eri.json to 1e-18) for g, h and i shells, contracted and
  uncontracted, up to l_total 17 and the full (gg|gg) block;
- OS/HGP against Rys on four distinct centres up to l_total 24, the
  reference's own guard crates/integral/tests/eri_maxima.rs:77-110 with the
  metrics and bounds of :14-27, :66-71 (significant-element relative 1e-9,
  absolute 1e-9 peak + 1e-10), and cross_engine.rs:40-50 (1e-11 + 1e-10 rel)
  on the smaller corners.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(HERE, name)) as f:
        return json.load(f)


@pytest.mark.parametrize("engine", [O.ENGINE_OS, O.ENGINE_RYS, O.ENGINE_AUTO])
def test_high_l_blocks_match_mcmurchie_davidson(engine):
    g = load("md_eri_high.json")
    shells = [(s[0], tuple(s[1]), s[2], s[3], False) for s in g["shells"]]
    ob = O.Basis(shells)
    for q in g["quartets"]:
        idx = np.array(q["index"])
        ref = np.array(q["value"])
        v = ob.eri_block(*q["shells"], engine=engine, cartesian=True).ravel()
        assert v.size == q["n"]
        peak = np.max(np.abs(ref))
        err = np.max(np.abs(v[idx] - ref))
        assert err < 1e-12 * max(1.0, peak), (q["shells"], engine, err, peak)


def maxima_basis(ls):
    """eri_maxima.rs:30-44: four distinct centres, one primitive each."""
    cs = [(0.0, 0.0, 0.0), (0.5, -0.3, 0.2), (-0.4, 0.6, -0.1), (0.2, 0.4, 0.8)]
    es = [0.9, 1.3, 0.7, 1.1]
    return O.Basis([(l, c, [e], [1.0], False) for l, c, e in zip(ls, cs, es)])


def metrics(os_, rys):
    """eri_maxima.rs:14-27: worst absolute, worst relative over elements of at
    least 1e-3 of the peak, peak."""
    peak = np.max(np.abs(rys))
    d = np.abs(os_ - rys)
    sig = np.abs(rys) >= 1e-3 * peak
    return float(np.max(d)), float(np.max(d[sig] / np.abs(rys[sig]))), float(peak)


@pytest.mark.parametrize("ls", [(4, 4, 4, 4), (6, 0, 6, 0), (6, 6, 0, 0), (0, 0, 6, 6),
                                (0, 3, 4, 6), (6, 4, 3, 0), (5, 5, 5, 5), (6, 6, 6, 0),
                                (2, 6, 6, 6), (6, 6, 6, 6)])
def test_os_matches_rys_at_the_maxima(ls):
    b = maxima_basis(ls)
    os_ = b.eri_block(0, 1, 2, 3, engine=O.ENGINE_OS, cartesian=True).ravel()
    rys = b.eri_block(0, 1, 2, 3, engine=O.ENGINE_RYS, cartesian=True).ravel()
    wabs, wsig, peak = metrics(os_, rys)
    assert wsig <= 1e-9 and wabs <= 1e-9 * max(peak, 1.0) + 1e-10, (ls, wabs, wsig, peak)
    if sum(ls) <= 16:  # cross_engine.rs:40-50 on the corners its suite runs
        assert np.all(np.abs(os_ - rys) <= 1e-11 + 1e-10 * np.abs(rys)), ls
