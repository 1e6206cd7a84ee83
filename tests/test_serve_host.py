"""Host logic of the serving loop (no GPU): grid routing, slot planning by
HBM footprint, the synthetic Poisson request stream."""
import math

from eff_pot_amd import serve


def test_grid_routing():
    assert serve.grid_for(100) == 64
    assert serve.grid_for(3000) == 128
    for n in (100, 300, 1000, 2000, 3000):
        g = serve.grid_for(n)
        assert g in serve.GRID_CLASSES
        assert 2 * serve.radius(n) + 20 <= serve.SPACING * g
        smaller = [c for c in serve.GRID_CLASSES if c < g]
        assert all(2 * serve.radius(n) + 20 > serve.SPACING * c for c in smaller)


def test_slot_plan_fits_budget():
    budget = 100 * 2 ** 30
    plan = serve.plan_slots(budget)
    used = sum(serve.footprint_bytes(serve.class_params(n, k)) for n, k in plan.items())
    assert used <= budget * 1.01
    assert all(k >= 1 for k in plan.values())


def test_stream_statistics():
    reqs = serve.make_stream(2000, rate=5.0, seed=3)
    gaps = [b.arrival - a.arrival for a, b in zip(reqs, reqs[1:])]
    assert abs(sum(gaps) / len(gaps) - 0.2) < 0.02
    ns = [r.natoms for r in reqs]
    assert min(ns) >= 100 and max(ns) <= 3000
    assert abs(sum(math.log(n) for n in ns) / len(ns) - (math.log(100) + math.log(3000)) / 2) < 0.1
    for r in reqs:
        lim = serve.SPACING * serve.grid_for(r.natoms) / 2 - 8.0
        assert all(abs(c) <= lim + 1e-12 for c in r.pos)
        assert 10 <= r.D <= 60 and 3.5 <= r.re <= 5.0 and 0.8 <= r.a <= 1.4


def test_budget_reserves_and_marks_starved_classes():
    """Admission accounting: a block's footprint is reserved only while the
    budget holds it; a class refused is marked waiting until it gets budget,
    which is what makes other classes hand back their idle blocks."""
    b = serve.Budget(10)
    assert b.take(6, 64) and b.used == 6
    assert not b.take(6, 128) and b.used == 6
    assert b.others_waiting(64) and not b.others_waiting(128)
    b.give(6)
    assert b.take(6, 128) and not b.others_waiting(64)
    assert b.take(4, 96) and b.used == 10 and b.peak == 10
    b.not_waiting(96)
    assert not b.take(1, 96) and b.others_waiting(128)
    b.not_waiting(96)
    assert not b.others_waiting(128)


def test_block_footprint_is_what_a_block_reserves():
    for n, k in serve.BLOCK_SLOTS.items():
        one = serve.footprint_bytes(serve.class_params(n, 1))
        blk = serve.footprint_bytes(serve.class_params(n, k))
        assert one < blk <= k * one
