"""World painting: carrier spawn and removal (crates/embloks-core/src/world.rs:74-238).

Restates the reference's world tests (world.rs:296-420) on `embloks_amd.world`,
which builds the scenes every backend (the HIP library and the oracle alike) is
given. Painting over, erasing and repainting decide which electrons stay, so
an edited scene uploaded to the device is the scene the reference would have.
"""
import numpy as np

from embloks_amd import world as W


def row(c0, c1, r):
    return [(c, r) for c in range(c0, c1)]


def make_world(n, cols, rows):
    return W.World(cols, rows, dict(W.abi.DEFAULTS, carriers_per_cell=n), seed=0)


def electrons_per_cell(w):
    counts = np.zeros((w.cols, w.rows), np.int64)
    cell = np.floor(w.pos).astype(np.int64)
    np.add.at(counts, (cell[:, 0], cell[:, 1]), 1)
    return counts


def assert_electrons_inside_conductors(w):
    cell = np.floor(w.pos).astype(np.int64)
    assert np.all(w.kind[cell[:, 0], cell[:, 1]] > 0)


def test_painting_spawns_n_electrons_per_cell_inside_it():
    # world.rs:324-337
    for n in (1, 4):
        w = make_world(n, 12, 8)
        w.paint(row(2, 7, 3), W.CONDUCTOR, resistance=1.0)
        assert w.n_electrons == 5 * n
        counts = electrons_per_cell(w)
        assert all(counts[c, 3] == n for c in range(2, 7))
        assert abs(w.electron_charge + w.settings["cell_charge"] / n) < 1e-12


def test_repainting_conductors_keeps_their_electrons():
    # world.rs:339-346
    w = make_world(1, 12, 8)
    w.paint(row(2, 7, 3), W.CONDUCTOR, resistance=1.0)
    before = w.pos.copy()
    w.paint(row(2, 7, 3), W.CONDUCTOR, resistance=5.0)
    assert np.array_equal(w.pos, before)


def test_painting_a_battery_over_a_wire_keeps_its_charges_and_adds_more():
    # world.rs:348-363
    w = make_world(1, 12, 8)
    w.paint(row(2, 7, 3), W.CONDUCTOR, resistance=1.0)
    before = w.pos.copy()
    w.paint([(4, 3)], W.BATTERY, emf=5.0)
    s = w.settings
    assert np.array_equal(w.pos[: len(before)], before)
    assert w.n_electrons == len(before) + (s["battery_density"] - 1) * s["carriers_per_cell"]
    assert w.resistance[4, 3] == s["battery_resistance"]
    # And back to a wire: the extra electrons leave again.
    w.paint(row(4, 5, 3), W.CONDUCTOR, resistance=1.0)
    assert w.n_electrons == len(before)


def test_a_painted_battery_gets_a_dense_ion_and_electrons():
    # world.rs:365-375
    w = make_world(1, 12, 8)
    w.paint([(4, 3)], W.BATTERY, emf=5.0)
    s = w.settings
    assert w.density[4, 3] == s["battery_density"]
    assert w.n_electrons == s["battery_density"] * s["carriers_per_cell"]
    assert_electrons_inside_conductors(w)
    w.paint([(4, 3)], W.VACUUM)
    assert w.n_electrons == 0


def test_erasing_removes_electrons_of_erased_cells():
    # world.rs:377-384
    w = make_world(2, 12, 8)
    w.paint(row(2, 7, 3), W.CONDUCTOR, resistance=1.0)
    w.paint([(4, 3)], W.VACUUM)
    assert w.n_electrons == 8
    assert_electrons_inside_conductors(w)


def test_erasing_an_empty_cell_takes_nearest_electron_from_its_island():
    # world.rs:386-394
    w = make_world(1, 12, 8)
    w.paint(row(2, 7, 3), W.CONDUCTOR, resistance=1.0)
    w.pos[:] = [2.5, 3.5]   # every electron in the leftmost cell, cell 6 empty
    w.paint([(6, 3)], W.VACUUM)
    assert w.n_electrons == 4


def test_extra_electrons_in_erased_cell_move_to_nearest_cell_of_same_island():
    # world.rs:396-408
    w = make_world(1, 12, 8)
    w.paint(row(2, 7, 3), W.CONDUCTOR, resistance=1.0)
    w.paint([(10, 6)], W.CONDUCTOR, resistance=1.0)   # a separate island far away
    w.pos[:5] = [6.5, 3.5]   # the whole row crowds the end cell
    w.paint([(6, 3)], W.VACUUM)
    assert w.n_electrons == 5
    counts = electrons_per_cell(w)
    assert counts[5, 3] == 4   # leftovers landed in the nearest remaining cell...
    assert counts[10, 6] == 1  # ...not in the other island
    assert_electrons_inside_conductors(w)
    moved = np.floor(w.pos[:4]).astype(np.int64)
    assert np.all(moved == [5, 3]) and np.all(w.vel[:4] == 0.0)


def test_erasing_whole_island_removes_all_its_electrons_even_if_charged():
    # world.rs:410-418
    w = make_world(1, 12, 8)
    w.paint(row(2, 4, 3), W.CONDUCTOR, resistance=1.0)
    w.paint([(9, 6)], W.CONDUCTOR, resistance=1.0)
    w.pos[2] = [2.5, 3.5]   # the other island's electron moves in (net -1)
    w.paint([(2, 3), (3, 3)], W.VACUUM)
    assert w.n_electrons == 0


def test_nearest_electrons_leave_first_when_the_lost_cells_run_short():
    # The quota's shortfall comes from the electrons nearest the lost cells'
    # centres (world.rs:205-216), not from anywhere on the island.
    w = make_world(1, 12, 8)
    w.paint(row(1, 11, 3), W.CONDUCTOR, resistance=1.0)
    w.pos[:] = [[c + 0.5, 3.5] for c in range(1, 11)]
    w.pos[5] = [1.5, 3.5]   # cell 6 is empty; its neighbours 5 and 7 hold the nearest
    w.paint([(6, 3)], W.BATTERY, emf=1.0)    # no loss: a battery adds density
    assert w.n_electrons == 10 + (w.settings["battery_density"] - 1)
    w.paint([(6, 3)], W.CONDUCTOR, resistance=1.0)   # back to wire: loses 3 carriers
    assert w.n_electrons == 10
    w2 = make_world(1, 12, 8)
    w2.paint(row(1, 11, 3), W.CONDUCTOR, resistance=1.0)
    w2.pos[:] = [[c + 0.5, 3.5] for c in range(1, 11)]
    w2.pos[5] = [1.5, 3.5]
    w2.paint([(6, 3)], W.VACUUM)   # quota 1, no electron inside: one of the nearest leaves
    assert w2.n_electrons == 9
    left = sorted(np.floor(w2.pos[:, 0]).astype(int).tolist())
    # Cells 5 and 7 tie at distance 1; the stable sort takes the lower index.
    assert left == [1, 1, 2, 3, 4, 7, 8, 9, 10]


def test_labels_join_edges_not_corners_in_scan_order():
    # connectivity.rs:20-62: islands numbered by their first cell in scan order.
    w = W.World(6, 5)
    w.paint([(0, 0), (1, 1)], W.CONDUCTOR, resistance=1.0)   # diagonal: two islands
    w.paint([(3, 0), (3, 1), (4, 1)], W.CONDUCTOR, resistance=1.0)
    w.paint([(1, 4)], W.BATTERY, emf=1.0)
    lab = w.labels()
    assert lab[0, 0] == 0 and lab[1, 1] == 1 and lab[1, 4] == 2
    assert lab[3, 0] == lab[3, 1] == lab[4, 1] == 3
    assert (lab == -1).sum() == 30 - 6
