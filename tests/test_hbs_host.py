"""CPU: host-side bookkeeping of the HBS engine that needs no GPU -- the leaf-permutation
cache (identity orders need no device copy and no gathers) and the event-based stage
clocks of HBSMAT (pending intervals fold into the total when it is read or reset)."""
import numpy as np


def test_identity_permutation_needs_no_device_copy():
    from sslablu_amd import hbs
    hbs._PERMS.clear()
    assert hbs.device_perm(None, "cuda") is None
    assert hbs.device_perm(np.arange(9216), "cuda") is None
    assert hbs.device_perm(list(range(5)), "cuda:0") is None
    # cached by content and device: a second identity of the same length is a hit
    n = len(hbs._PERMS)
    assert hbs.device_perm(np.arange(9216, dtype=np.int32), "cuda") is None
    assert len(hbs._PERMS) == n


class _Ev:
    """Stands in for a timing event pair: elapsed_time in milliseconds."""

    def __init__(self, t):
        self.t = t
        self.waited = False

    def synchronize(self):
        self.waited = True

    def elapsed_time(self, other):
        return other.t - self.t


def test_stage_clock_folds_pending_intervals():
    from sslablu_amd.matAssembly.HBS.HBSnew import _StageClock
    c = _StageClock()
    e1, e2, e3, e4 = _Ev(0.0), _Ev(250.0), _Ev(1000.0), _Ev(1500.0)
    c.open += [(e1, e2), (e3, e4)]
    assert abs(c.value() - 0.75) < 1e-12 and e2.waited and e4.waited
    assert c.open == [] and abs(c.value() - 0.75) < 1e-12
    c.open.append((_Ev(0.0), _Ev(100.0)))
    c.set(2.0)  # assignment drops what is pending
    assert c.value() == 2.0
