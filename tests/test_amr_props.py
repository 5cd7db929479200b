"""The full-size AMR property checker (tests/amr_props.py) on hierarchies
the oracle builds: every property holds, and each planted mistake (a
density off by one part in 10^4, a particle moved to another node, a
refined octet dropped, a particle listed twice, a domain particle listed in
a neighbouring cell) is caught."""
import os

import numpy as np
import pytest

from ahf_amd import dumps

import amr_props


def _oracle_levels(oracle, snapgen, tmp_path, cfg, side, seed):
    snap = str(tmp_path / "p.snap")
    assert snapgen.ahf_snapgen_write(cfg, side, seed, snap.encode()) == 0
    out = str(tmp_path / "p.grid")
    assert oracle.oracle_amr_dump(snap.encode(), out.encode()) > 0
    s = dumps.read_snapshot(snap)
    return s, amr_props.levels_from_dump(dumps.read_grids(out))


CASES = [(1, 48, 7), (3, 48, 5), (4, 64, 1004), (4, 112, 1004), (0, 48, 1000)]


@pytest.mark.parametrize("cfg,side,seed", CASES)
def test_oracle_hierarchy_has_the_properties(oracle, snapgen, tmp_path, cfg, side, seed):
    s, levels = _oracle_levels(oracle, snapgen, tmp_path, cfg, side, seed)
    assert len(levels) >= 3
    fails = amr_props.check_hierarchy(levels, s.pos, np.random.default_rng(1), sample=400)
    assert not fails, fails


def _sampled_level(levels, rng):
    """a level with sampled nodes, and one of those nodes"""
    for l in range(len(levels) - 1):
        nodes = amr_props._sample_nodes(levels[l], rng, 50)
        if nodes.size:
            return l, nodes
    raise AssertionError("no sampled node")


def test_planted_mistakes_are_caught(oracle, snapgen, tmp_path):
    s, levels = _oracle_levels(oracle, snapgen, tmp_path, 1, 48, 7)
    rng = np.random.default_rng(2)
    l, nodes = _sampled_level(levels, rng)

    def fails_with(mutate, **kw):
        lv = [dict(x) for x in levels]
        mutate(lv)
        return amr_props.check_hierarchy(lv, s.pos, np.random.default_rng(1), sample=400, **kw)

    # a density off by one part in 10^4 on every node of one level
    def dens(lv):
        lv[l]["dens"] = (lv[l]["dens"].astype(np.float64) * 1.0001 + 1e-4).astype(np.float32)
    assert any("densities differ" in f for f in fails_with(dens))

    # a refined octet dropped from the next level
    def drop(lv):
        fin = lv[l + 1]
        x, y, z = amr_props._nodes_xyz(lv[l], nodes)
        test = np.zeros(nodes.size, bool)
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (0, 1):
                    nb = amr_props.node_index(lv[l], x + dx, y + dy, z + dz)
                    test |= lv[l]["dens"][nb].astype(np.float64) >= lv[l]["critdens"] - 1.0
        keys = (x | (y << 21) | (z << 42)).astype(np.uint64)[test]
        assert keys.size
        fin["oct_key"] = np.setdiff1d(fin["oct_key"], keys)
    assert any("not refined" in f for f in fails_with(drop, dens_levels={l}))

    big = int(np.argmax([x["list"].size for x in levels]))

    # a particle listed twice
    def dup(lv):
        lv[big]["list"] = lv[big]["list"].copy()
        lv[big]["list"][1] = lv[big]["list"][0]
    assert any("exactly once" in f for f in fails_with(dup))

    # the lists of the most populated level rotated by one node: particles
    # listed in neighbouring cells
    def shift(lv):
        d = lv[big]
        cnt = np.roll(np.diff(d["list0"]), 1)
        d["list0"] = np.concatenate([[0], np.cumsum(cnt)])
    assert any("ll()'s cell" in f or "outside their node's cell" in f for f in fails_with(shift))

    # a fine level's particle swapped into another node
    def move(lv):
        for j in range(1, len(lv)):
            d = lv[j]
            if d["list"].size > 100:
                d["list"] = d["list"].copy()
                d["list"][[0, -1]] = d["list"][[-1, 0]]
                return
        raise AssertionError("no fine list")
    assert any("outside their node's cell" in f for f in fails_with(move))
