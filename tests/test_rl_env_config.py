"""Host logic of the batched environment: the stacked configuration
(rl_env.stack_config) and the action scaling of rl_env.py:132-144."""
import pytest

from rl_env import _apply_action_to_wells, stack_config


def _cfg():
    return {"reservoir": {"dimensions": [5, 4, 2], "grid_size": [10.0, 10.0, 5.0]},
            "fluid": {"pressure": 20.0},
            "wells": [{"name": "INJ", "type": "injector", "i": 0, "j": 1, "k": 0, "radius": 0.1,
                       "control_type": "rate", "control_value": 10.0},
                      {"name": "PROD", "type": "producer", "coordinates": [4, 3, 1], "radius": 0.1,
                       "control": {"type": "bhp", "value": 15.0}}],
            "simulation": {"time_step_days": 1.0}}


def test_stack_config_repeats_wells_per_copy():
    cfg = _cfg()
    out = stack_config(cfg, 3)
    assert out["reservoir"]["dimensions"] == [15, 4, 2]
    assert [w["name"] for w in out["wells"]] == ["INJ#0", "PROD#0", "INJ#1", "PROD#1", "INJ#2", "PROD#2"]
    assert [w["i"] for w in out["wells"][0::2]] == [0, 5, 10]
    assert [w["coordinates"] for w in out["wells"][1::2]] == [[4, 3, 1], [9, 3, 1], [14, 3, 1]]
    assert out["fluid"] == cfg["fluid"] and out["simulation"] == cfg["simulation"]
    assert cfg["reservoir"]["dimensions"] == [5, 4, 2] and len(cfg["wells"]) == 2  # input untouched
    one = stack_config(cfg, 1)
    assert one["reservoir"]["dimensions"] == [5, 4, 2] and one["wells"][0]["name"] == "INJ#0"
    with pytest.raises(ValueError):
        stack_config(cfg, 0)
    bad = _cfg()
    del bad["wells"][0]["i"]
    with pytest.raises(ValueError):
        stack_config(bad, 2)


class _W:
    def __init__(self, ct, cv):
        self.control_type, self.control_value = ct, cv


def test_action_scaling():
    wells = [_W("rate", 100.0), _W("bhp", 15.0), _W("other", 3.0)]
    _apply_action_to_wells([1.0, -0.5, 1.0], wells)
    assert wells[0].control_value == pytest.approx(120.0)
    assert wells[1].control_value == pytest.approx(12.5)
    assert wells[2].control_value == 3.0
    _apply_action_to_wells([-1.0, 0.0, 0.0], wells)  # compounds on the current control
    assert wells[0].control_value == pytest.approx(96.0)
