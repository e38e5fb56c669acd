"""The oracle's node sets: the wall acting on a subset of nodes (the slave
set) and held nodes (the fixed set), both switched off in the reference,
which collects the sets (mesh.h:270-329) and applies neither.  Checked
through properties that pin their meaning: an empty contact set is a run
without the wall, a full one is the unmasked run, held nodes never move,
and the energy balance still closes (reactions at held nodes do no work)."""
import numpy as np

from oracle import pyoracle as O

from common import cube, rel, run_oracle, wall_k


def test_empty_contact_set_is_a_run_without_the_wall():
    x, conn = cube(3, gap=1e-6)
    k = wall_k(0.25, 3)
    a = run_oracle(x, conn, (0, 0, -20.0), 2e-7, 60, O.EP, k, contact_nodes=[])
    b = run_oracle(x, conn, (0, 0, -20.0), 2e-7, 60, O.EP)
    assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["v"], b["v"])
    assert a["energy"][-1][O.E_CONTACTF] == 0.0


def test_full_contact_set_is_the_unmasked_run():
    x, conn = cube(3, gap=1e-6)
    k = wall_k(0.25, 3)
    a = run_oracle(x, conn, (0, 0, -20.0), 2e-7, 60, O.EP, k, contact_nodes=np.arange(len(x)))
    b = run_oracle(x, conn, (0, 0, -20.0), 2e-7, 60, O.EP, k)
    assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["energy"], b["energy"])
    assert b["energy"][-1][O.E_CONTACTF] > 0.0


def test_contact_subset_lands_between():
    # only the bottom nodes with even x index feel the wall: the others
    # sink through it
    x, conn = cube(3, gap=1e-6)
    k = wall_k(0.25, 3)
    bottom = np.where(x[:, 2] < 1e-5)[0]
    sub = bottom[np.round(x[bottom, 0] / (0.25 / 3)).astype(int) % 2 == 0]
    o = run_oracle(x, conn, (0, 0, -20.0), 2e-7, 80, O.EP, k, contact_nodes=sub)
    other = np.setdiff1d(bottom, sub)
    assert o["x"][sub, 2].min() > o["x"][other, 2].max()


def test_held_nodes_stay_and_energy_balances():
    # a block swinging on its clamped face (x = 0) under gravity and a
    # lateral initial velocity
    x, conn = cube(3, gap=0.0)
    fixed = np.where(x[:, 0] < 1e-9)[0]
    o = run_oracle(x, conn, (0.0, 8.0, 0.0), 2e-7, 200, O.EP, fixed_nodes=fixed)
    assert np.array_equal(o["x"][fixed], x[fixed])
    assert not o["v"][fixed].any() and not o["a"][fixed].any()
    free = np.setdiff1d(np.arange(len(x)), fixed)
    assert np.abs(o["x"][free] - x[free]).max() > 1e-5
    e = o["energy"]
    ke0 = e[0][O.E_KINETIC]
    # the held nodes start at rest: the initial kinetic energy counts only
    # the free nodes
    assert ke0 < 0.5 * o["m"].sum() * 64.0
    assert abs(e[-1][O.E_BALANCE]) < 2e-3 * ke0
    assert e[-1][O.E_INTERNAL] > 0.05 * ke0


def test_all_held_nothing_moves():
    x, conn = cube(2, gap=0.0)
    o = run_oracle(x, conn, (0.0, 0.0, -10.0), 2e-7, 20, O.EP, fixed_nodes=np.arange(len(x)))
    assert np.array_equal(o["x"], x) and not o["v"].any()
    assert o["energy"][-1][O.E_INTERNAL] == 0.0
