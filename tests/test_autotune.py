"""The autotuner's crowd-threshold stage (prrtc_amd.autotune.tune_crowd) on
the CPU, with a stand-in for the library whose batch times depend on the
threshold: a count replaces the default only when it is >= 3% faster, and
launch shapes without crowd claims are left alone."""
from prrtc_amd import autotune


class FakeLib:
    def __init__(self, seconds_by_crowd):
        self.seconds = seconds_by_crowd
        self.crowd = 0
        self.calls = []

    def set_tuning(self, robot, wpc, waves, bwaves, crowd=0):
        self.crowd = crowd
        self.calls.append((robot, wpc, waves, bwaves, crowd))

    def solve_batch(self, probs, st):
        return [{"solved": True} for _ in probs], None, None


def run(seconds, launch=None, monkeypatch=None):
    monkeypatch.setattr(autotune, "problems", lambda robot, n: list(range(n)))
    launch = launch or {"waves_per_cu": 8, "waves_per_task": 4, "batch_waves_per_task": 4, "crowd_above": 0}
    lib = FakeLib(seconds)
    # the batch time of the threshold set last, exactly
    monkeypatch.setattr(autotune, "timed", lambda fn: (lib.seconds[lib.crowd], fn()))
    return autotune.tune_crowd(lib, "baxter", 10, launch, 128, 3000), lib


def test_clear_gain_moves_off_the_default(monkeypatch):
    (crowd, rows), lib = run({0: 0.10, 8: 0.099, 16: 0.07, 32: 0.08}, monkeypatch=monkeypatch)
    assert crowd == 16 and [r["crowd_above"] for r in rows] == list(autotune.CROWD_CANDIDATES)
    assert lib.calls[-1] == ("baxter", 0, 0, 0, 0)   # the tuning is reset afterwards


def test_gain_below_three_percent_keeps_the_default(monkeypatch):
    (crowd, _), _ = run({0: 0.100, 8: 0.099, 16: 0.0985, 32: 0.11}, monkeypatch=monkeypatch)
    assert crowd == 0


def test_no_crowd_claims_is_left_alone(monkeypatch):
    one_wave = {"waves_per_cu": 4, "waves_per_task": 1, "batch_waves_per_task": 1, "crowd_above": -1}
    (crowd, rows), lib = run({0: 0.1}, launch=one_wave, monkeypatch=monkeypatch)
    assert crowd == -1 and rows == [] and lib.calls == []
