"""The default discrete action table (RLGym DefaultAction, DefaultAction.cpp:3-59)
built by the library: pinned by the properties its enumeration fixes."""
import itertools

import pytest

import rsxlib as R


@pytest.fixture(scope="module")
def lut():
    import os
    if not os.path.exists(R.engine.LIB_PATH):
        pytest.skip("librsx.so is not built")
    t = R.engine.default_action_lut()
    return [(c.throttle, c.steer, c.pitch, c.yaw, c.roll, c.jump, c.boost, c.handbrake) for c in t]


def test_size_and_uniqueness(lut):
    assert len(lut) == 90
    assert len(set(lut)) == 90


def test_ground_block(lut):
    ground = lut[:24]
    for thr, steer, pitch, yaw, roll, jump, boost, hb in ground:
        assert pitch == 0 and roll == 0 and jump == 0 and yaw == steer
        assert not (boost and thr != 1)
    # throttle-major order over (-1, 0, 1), then steer, boost, handbrake
    assert [g[0] for g in ground] == [-1.0] * 6 + [0.0] * 6 + [1.0] * 12
    assert ground[0] == (-1, -1, 0, -1, 0, 0, 0, 0) and ground[1] == (-1, -1, 0, -1, 0, 0, 0, 1)
    # every allowed ground combination is present
    want = {(t, s, b, h) for t, s, b, h in itertools.product((-1, 0, 1), (-1, 0, 1), (0, 1), (0, 1))
            if not (b and t != 1)}
    assert {(g[0], g[1], g[6], g[7]) for g in ground} == want


def test_air_block(lut):
    air = lut[24:]
    assert len(air) == 66
    for thr, steer, pitch, yaw, roll, jump, boost, hb in air:
        assert thr == boost and steer == yaw
        assert not (jump and yaw != 0)
        assert not (pitch == 0 and roll == 0 and jump == 0)
        assert hb == (1 if jump and (pitch or yaw or roll) else 0)
    # pitch-major order
    assert [a[2] for a in air][:1] == [-1.0] and [a[2] for a in air][-1] == 1.0
    assert air == sorted(air, key=lambda a: (a[2], a[3], a[4], a[5], a[6]))


def test_mask_tables_counts(lut):
    """Counts the enumeration fixes: 24 ground entries; air: the 48 aerial
    entries without jump, less the first aerial entry (the reference compares
    the index strictly), plus the 6 ground entries with throttle = boost and
    handbrake = (yaw != 0); 18 jump entries (yaw 0, any pitch, roll, boost);
    boost: 6 ground + 33 aerial."""
    spec = R.load_spec()
    ground, air, boost, jump = spec.default_action_masks(lut)
    assert ground == [1] * 24 + [0] * 66
    assert sum(air) == 47 + 6 and air[24] == 0 and air[25] == 1
    assert sum(jump) == 18 and all(lut[i][5] for i in range(90) if jump[i])
    assert sum(boost) == 39


def test_mask_rules(lut):
    """GetActionMask on hand-made records: grounded without boost drops the
    boost entries; airborne with a flip gets every jump entry; a turtled car
    (world contact normal up) gets jump entries with no flip left."""
    spec = R.load_spec()
    masks = spec.default_action_masks(lut)
    c = R.CarState()
    c.is_on_ground, c.boost = 1, 0
    m = spec.action_mask(c, masks)
    assert m == [int(i < 24 and not lut[i][6]) | int(lut[i][5]) for i in range(90)]
    c.is_on_ground, c.boost, c.has_flipped, c.air_time_since_jump = 0, 50, 0, 0.1
    m = spec.action_mask(c, masks)
    assert m == [a | j for a, j in zip(masks[1], masks[3])]
    c.has_flipped = 1
    assert spec.action_mask(c, masks) == list(masks[1])
    c.world_contact, c.world_contact_normal.z = 1, 0.95
    assert spec.action_mask(c, masks) == [a | j for a, j in zip(masks[1], masks[3])]
