"""Host side of the automatic tap changer: the ranking of regulated transformers.

The grids and expected ranks are the reference's own unit cases
(tests/cpp_unit_tests/optimizer/test_tap_position_optimizer.cpp:201-783, "Test Transformer ranking"), the
ones built from two-winding transformers and lines: the meshed grid whose middle transformer is reached
through a closer regulated transformer (last rank), the transformer into itself, the multi-source chain
(its edge weights), and the two configurations controlled towards the source (AutomaticTapInputError).
The ranking runs on the host only; no device is needed.
"""

import numpy as np
import pytest

pgm = pytest.importorskip("power_grid_model_amd")
from power_grid_model_amd import PowerGridModel, initialize_array, tap_regulator_ranks  # noqa: E402

FROM, TO = 0, 1


def _grid(nodes, transformers, regulators, sources, lines=(), links=()):
    """transformers: (id, from, to, tap_side); regulators: (id, transformer, control_side); sources: (id, node)."""
    node = initialize_array("input", "node", len(nodes))
    node["id"] = [n for n, _ in nodes]
    node["u_rated"] = [u for _, u in nodes]
    tr = initialize_array("input", "transformer", len(transformers))
    for k, (i, f, t, side) in enumerate(transformers):
        tr[k]["id"], tr[k]["from_node"], tr[k]["to_node"] = i, f, t
        tr[k]["from_status"] = tr[k]["to_status"] = 1
        tr[k]["winding_from"] = tr[k]["winding_to"] = 1
        tr[k]["clock"] = 0
        tr[k]["tap_side"] = side
        tr[k]["tap_min"], tr[k]["tap_max"], tr[k]["tap_nom"] = -128, 127, 0
    reg = initialize_array("input", "transformer_tap_regulator", len(regulators))
    for k, (i, t, side) in enumerate(regulators):
        reg[k]["id"], reg[k]["regulated_object"], reg[k]["status"], reg[k]["control_side"] = i, t, 1, side
    src = initialize_array("input", "source", len(sources))
    for k, (i, n) in enumerate(sources):
        src[k]["id"], src[k]["node"], src[k]["status"], src[k]["u_ref"] = i, n, 1, 1.0
    data = {"node": node, "transformer": tr, "transformer_tap_regulator": reg, "source": src}
    if lines:
        ln = initialize_array("input", "line", len(lines))
        for k, (i, f, t) in enumerate(lines):
            ln[k]["id"], ln[k]["from_node"], ln[k]["to_node"] = i, f, t
            ln[k]["from_status"] = ln[k]["to_status"] = 1
        data["line"] = ln
    if links:
        lk = initialize_array("input", "link", len(links))
        for k, (i, f, t) in enumerate(links):
            lk[k]["id"], lk[k]["from_node"], lk[k]["to_node"] = i, f, t
            lk[k]["from_status"] = lk[k]["to_status"] = 1
        data["link"] = lk
    return PowerGridModel(data)


def test_meshed_grid_low_priority_rank():
    """test_tap_position_optimizer.cpp:572-618: 0 -(11)- 1 -(12)- 2 -(13)- 3 with a line 0 - 2. The control
    side of 12 is also reached through the line, so 12 gets the last rank after {11, 13}."""
    model = _grid([(0, 10e3), (1, 10e3), (2, 10e3), (3, 10e3)],
                  [(11, 0, 1, TO), (12, 1, 2, FROM), (13, 2, 3, FROM)],
                  [(41, 11, TO), (42, 12, TO), (43, 13, TO)], [(31, 0)], lines=[(21, 0, 2)])
    np.testing.assert_array_equal(tap_regulator_ranks(model), [0, 1, 0])


def test_meshed_grid_through_a_link():
    """The meshed grid above with the line 0 - 2 a link: every branch other than a regulated transformer is
    an edge of weight 0 (tap_position_optimizer.hpp:240-251, 289), so the ranks are the same."""
    model = _grid([(0, 10e3), (1, 10e3), (2, 10e3), (3, 10e3)],
                  [(11, 0, 1, TO), (12, 1, 2, FROM), (13, 2, 3, FROM)],
                  [(41, 11, TO), (42, 12, TO), (43, 13, TO)], [(31, 0)], links=[(21, 0, 2)])
    np.testing.assert_array_equal(tap_regulator_ranks(model), [0, 1, 0])
    upd = initialize_array("update", "link", 1)
    upd["id"] = 21
    upd["from_status"] = 0
    model.update(update_data={"link": upd})
    # without the link the chain is 11, 12, 13 in turn
    np.testing.assert_array_equal(tap_regulator_ranks(model), [0, 1, 2])
    upd["id"] = 12
    with pytest.raises(pgm.PowerGridError, match="Wrong type"):
        model.update(update_data={"link": upd})


def test_transformer_into_itself():
    """test_tap_position_optimizer.cpp:748-783: ranks {10}, {12}, {11}."""
    model = _grid([(0, 150e3), (1, 10e3), (2, 10e3)],
                  [(10, 0, 1, FROM), (11, 1, 1, FROM), (12, 1, 2, FROM)],
                  [(30, 10, TO), (31, 11, TO), (32, 12, TO)], [(20, 0)])
    np.testing.assert_array_equal(tap_regulator_ranks(model), [0, 2, 1])


def test_multiple_source_chain():
    """test_tap_position_optimizer.cpp:282-309: regulated arcs 0->1, 1->2, 3->2, 4->3 with sources at 0
    and 4 weigh 1, 2, 2, 1."""
    model = _grid([(0, 10e3), (1, 10e3), (2, 10e3), (3, 10e3), (4, 10e3)],
                  [(10, 0, 1, FROM), (11, 1, 2, FROM), (12, 3, 2, FROM), (13, 4, 3, FROM)],
                  [(20, 10, TO), (21, 11, TO), (22, 12, TO), (23, 13, TO)], [(30, 0), (31, 4)])
    np.testing.assert_array_equal(tap_regulator_ranks(model), [0, 1, 1, 0])


def test_regulator_off_or_unreached():
    """Regulators that are off, or whose transformer no source reaches, are not ranked."""
    model = _grid([(0, 10e3), (1, 10e3), (2, 10e3), (3, 10e3)],
                  [(10, 0, 1, FROM), (11, 2, 3, FROM)],
                  [(20, 10, TO), (21, 11, TO)], [(30, 0)])
    np.testing.assert_array_equal(tap_regulator_ranks(model), [0, -1])


def test_controlled_towards_source():
    """test_tap_position_optimizer.cpp:620-637."""
    model = _grid([(0, 150e3), (1, 10e3)], [(2, 0, 1, FROM)], [(4, 2, FROM)], [(3, 0)])
    with pytest.raises(pgm.AutomaticTapInputError):
        tap_regulator_ranks(model)


def test_several_controlled_towards_source():
    """test_tap_position_optimizer.cpp:639-702: 20 and 30 are listed, once each, in ascending order."""
    model = _grid([(0, 150e3), (1, 10e3), (2, 10e3), (3, 10e3), (4, 10e3), (5, 10e3)],
                  [(10, 0, 1, FROM), (20, 1, 2, TO), (30, 1, 3, TO), (40, 2, 4, FROM), (50, 3, 5, FROM)],
                  [(70, 10, TO), (71, 20, FROM), (72, 30, FROM), (73, 40, TO), (74, 50, TO)], [(60, 0)])
    with pytest.raises(pgm.AutomaticTapInputError) as info:
        tap_regulator_ranks(model)
    msg = str(info.value)
    assert "controlled from non-source side towards source side" in msg
    assert msg.endswith("Transformer IDs: 20, 30")


def test_regulator_input_validation():
    """main_core/input.hpp:168-241: a regulator must regulate a transformer, at most one per transformer."""
    good = [(0, 150e3), (1, 10e3)]
    with pytest.raises(pgm.PowerGridError, match="regulated by more than one regulator"):
        _grid(good, [(2, 0, 1, FROM)], [(4, 2, TO), (5, 2, TO)], [(3, 0)])
    with pytest.raises(pgm.PowerGridError, match="regulator is not supported for object with ID 1"):
        _grid(good, [(2, 0, 1, FROM)], [(4, 1, TO)], [(3, 0)])
    with pytest.raises(pgm.PowerGridError, match="cannot be found"):
        _grid(good, [(2, 0, 1, FROM)], [(4, 99, TO)], [(3, 0)])


def test_model_update_of_regulators():
    """PGM_update_model applies transformer_tap_regulator updates (transformer_tap_regulator.hpp:38-46: a null
    attribute keeps its value): switching a regulator off removes it from the ranking, and switching it
    back on restores the ranks; an unknown id is rejected."""
    model = _grid([(0, 10e3), (1, 10e3), (2, 10e3), (3, 10e3)],
                  [(11, 0, 1, TO), (12, 1, 2, FROM), (13, 2, 3, FROM)],
                  [(41, 11, TO), (42, 12, TO), (43, 13, TO)], [(31, 0)], lines=[(21, 0, 2)])
    np.testing.assert_array_equal(tap_regulator_ranks(model), [0, 1, 0])
    upd = initialize_array("update", "transformer_tap_regulator", 1)
    upd["id"] = 41
    upd["status"] = 0
    model.update(update_data={"transformer_tap_regulator": upd})
    # without 41, transformer 11 joins nodes 0 and 1 at distance 0: 12 runs between two buses at distance
    # 0 (last rank), 13 from distance 0 to 1 (first rank)
    np.testing.assert_array_equal(tap_regulator_ranks(model), [-1, 1, 0])
    upd["status"] = 1
    model.update(update_data={"transformer_tap_regulator": upd})
    np.testing.assert_array_equal(tap_regulator_ranks(model), [0, 1, 0])
    upd["id"] = 99
    with pytest.raises(pgm.PowerGridError, match="cannot be found"):
        model.update(update_data={"transformer_tap_regulator": upd})
