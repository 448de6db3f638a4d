"""emt6ro.tune.autotune's bookkeeping, on a stand-in batch that records the
calls (the timings are made up): the first timing runs with the window pass
off, a refinement with the caller's window setting, and the caller's setting
and the chosen variant are what the batch is left with."""
import emt6ro.tune as T


class FakeBatch:
    """Times are a function of (variant, window on, steps since reset)."""

    def __init__(self, n=8):
        self.n, self.window, self.variant, self.t = n, -1, None, 0
        self.log = []

    def set_window(self, rows):
        self.window = rows

    def set_variant(self, rs, v):
        self.variant = v

    def reset(self, epoch=0):
        self.t = 0

    def run(self, steps, path="persistent"):
        self.log.append((self.variant, self.window, self.t, steps))
        early = {1: 100.0, 2: 101.0, 3: 150.0}[self.variant]
        late = {1: 120.0, 2: 110.0, 3: 150.0}[self.variant]
        ms = late if (self.t > 0 and self.window != 0) else early
        self.t += steps
        return ms


def _variants():
    return [dict(id=i, kind=3, nt=256, l=4, kmax=2, kf=0, ps=2, rs=56) for i in (1, 2, 3)]


def test_first_timing_without_window_and_refinement_with_it(monkeypatch):
    monkeypatch.setattr(T, "variants", _variants)
    monkeypatch.setattr(T, "fits", lambda v, h, w: True)
    b = FakeBatch()
    b.set_window(-1)
    out = T.autotune(b, {56: (53, 53)}, steps=10, use_cache=False, refine_warm=50)
    first = [e for e in b.log if e[3] == 10 and e[2] == 0]
    assert {e[0] for e in first} == {1, 2, 3} and all(e[1] == 0 for e in first)
    warms = [e for e in b.log if e[3] == 50]
    assert len(warms) == 2 and all(e[0] == 1 and e[1] == -1 for e in warms)
    refined = [e for e in b.log if e[3] == 10 and e[2] == 50]
    assert sorted(e[0] for e in refined) == [1, 2] and all(e[1] == -1 for e in refined)
    assert out[56]["variant"] == 2 and out[56]["ms_later"] == {1: 120.0, 2: 110.0}
    assert b.variant == 2 and b.window == -1


def test_window_setting_restored_without_refinement(monkeypatch):
    monkeypatch.setattr(T, "variants", _variants)
    monkeypatch.setattr(T, "fits", lambda v, h, w: True)
    b = FakeBatch()
    b.set_window(21)
    out = T.autotune(b, {56: (53, 53)}, steps=10, use_cache=False)
    assert out[56]["variant"] == 1 and b.variant == 1 and b.window == 21
    assert all(e[1] == 0 for e in b.log)
