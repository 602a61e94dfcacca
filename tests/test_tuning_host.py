"""CPU: the per-thread policy cache of sslablu_amd.tuning (no GPU, the library call is
replaced by a recorder): equal settings skip the library, scopes restore what was set,
a failed call drops the cached value, and every thread starts unknown."""
import threading

import pytest


@pytest.fixture
def rec(monkeypatch):
    from sslablu_amd import tuning
    calls = []

    def fake_call(name, *args):
        if name == "sslab_gemm_set_policy" and getattr(tuning, "_fail", False):
            raise RuntimeError("rejected")
        calls.append(name)
        return 0

    monkeypatch.setattr(tuning, "call", fake_call)
    monkeypatch.setattr(tuning, "_tls", threading.local())
    yield tuning, calls
    tuning._fail = False


def test_equal_policy_skips_library(rec):
    tuning, calls = rec
    p = tuning.Policy(512, 1024, 64, 256, 2, 12, 0)
    tuning.set_policy(p)
    tuning.set_policy(p)
    assert calls == ["sslab_gemm_set_policy"]
    with tuning.policy(p):  # already standing: no calls on entry or exit
        pass
    assert calls == ["sslab_gemm_set_policy"]
    q = tuning.Policy(1024, 512, 64, 256, 3, 14, 1)
    with tuning.policy(q):
        assert tuning.get_policy() == q
    assert tuning.get_policy() == p
    assert calls.count("sslab_gemm_set_policy") == 3  # q in, p back


def test_backward_table_scope_and_clear(rec):
    tuning, calls = rec
    t = [tuning.Policy(), tuning.Policy(cfg_split=12)]
    with tuning.backward_policies(t):
        with tuning.backward_policies(t):  # same table: no call
            pass
    assert calls == ["sslab_rb_set_backward_policies"] * 2  # set, then cleared on exit
    tuning.set_backward_policies([])  # already clear
    assert len(calls) == 2
    with tuning.backward_policies(None):
        pass
    assert len(calls) == 2


def test_failed_call_drops_cache(rec):
    tuning, calls = rec
    p = tuning.Policy()
    tuning.set_policy(p)
    tuning._fail = True
    with pytest.raises(RuntimeError):
        tuning.set_policy(tuning.Policy(cfg_mid=2))
    assert getattr(tuning._tls, "policy", "unset") is None
    tuning._fail = False
    tuning.set_policy(p)  # unknown again: the library is called
    assert calls.count("sslab_gemm_set_policy") == 2


def test_threads_start_unknown(rec):
    tuning, calls = rec
    tuning.set_policy(tuning.Policy())
    seen = []
    th = threading.Thread(target=lambda: seen.append(getattr(tuning._tls, "policy", None)))
    th.start()
    th.join()
    assert seen == [None]
