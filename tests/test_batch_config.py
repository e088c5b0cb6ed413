"""The batched damage sweep's case draws (host logic, no GPU)."""
from semhip import batch as B


def test_damage_sweep_draws_are_seeded_and_in_range():
    a, b = B.damage_sweep(n_cases=8, seed=3), B.damage_sweep(n_cases=8, seed=3)
    assert [c.n for c in a.cases] == [c.n for c in b.cases]
    assert all(60 <= c.n <= 120 and 0.5 <= c.scale <= 0.9 for c in a.cases)
    assert all(5e-3 <= c.semi[0] <= 20e-3 and 0 <= c.centre[0] <= 0.5 for c in a.cases)
