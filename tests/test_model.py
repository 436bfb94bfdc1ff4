"""The reference's state pytrees and the packed vector (cells_amd.model.pack_states /
unpack_states) on the CPU: a round trip per world is exact, the leaves carry the reference's names
and axes, and malformed pytrees are refused."""

import numpy as np
import pytest

from cells_amd.model import pack_states, unpack_states
from cells_amd.world import StateLayout, build, config0, config1


@pytest.mark.parametrize("cfg", [config1(worlds=4, cells=(20, 60), seed=3), config0(worlds=3, cells=(9, 16))],
                         ids=["calcium", "water"])
def test_round_trip_per_world_is_exact(cfg):
    p = build(cfg)
    L = StateLayout.of(p)
    y = np.random.default_rng(2).normal(size=L.total)
    states = unpack_states(p, y)
    assert len(states) == p.W
    assert np.array_equal(pack_states(p, states), y)
    for k, st in enumerate(states):
        n = int(p.cell_off[k + 1] - p.cell_off[k])
        E = int(p.edge_off[k + 1] - p.edge_off[k])
        c0, e0 = int(p.cell_off[k]), int(p.edge_off[k])
        assert st["amount_increment"].shape == (p.S, n)
        assert np.array_equal(st["amount_increment"][1], L.cell(y, L.dn + 1)[c0:c0 + n])
        assert st["junction_gate"].shape == (E,)
        assert np.array_equal(st["junction_gate"], L.edge(y, L.gate)[e0:e0 + E])
        assert st["reservoir_supplied"].shape == (p.S, 1)


def test_malformed_pytrees_are_refused():
    p = build(config1(worlds=2, cells=(20, 30), seed=4))
    states = unpack_states(p, np.zeros(StateLayout.of(p).total))
    with pytest.raises(ValueError, match="one state per world"):
        pack_states(p, states[:1])
    bad = [dict(s) for s in states]
    del bad[1]["pump_cycles"]
    with pytest.raises(ValueError, match="missing state leaves"):
        pack_states(p, bad)
    bad = [dict(s) for s in states]
    bad[0]["junction_gate"] = bad[0]["junction_gate"][:-1]
    with pytest.raises(ValueError, match="shape"):
        pack_states(p, bad)
    bad = [dict(s) for s in states]
    bad[0]["extracellular_amount_increment"] = np.ones_like(bad[0]["extracellular_amount_increment"])
    with pytest.raises(ValueError, match="no slots"):
        pack_states(p, bad)
