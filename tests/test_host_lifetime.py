"""Host-side lifetime rules of the Python front-end (no GPU): a context
destroys its batches and surrogates before itself, and a child closed after
its context never reaches the library (which would read the freed context)."""
import ctypes as C

import pytest

import piro_amd.core as core


class FakeLib:
    """records the library calls the lifetime logic makes"""

    def __init__(self):
        self.calls = []
        self.next = 100

    def _handle(self, out):
        self.next += 1
        C.cast(out, C.POINTER(C.c_void_p))[0] = self.next

    def piro_ctx_create(self, device, out):
        self._handle(out)
        self.calls.append("ctx_create")
        return 0

    def piro_ctx_destroy(self, h):
        self.calls.append("ctx_destroy")
        return 0

    def piro_surrogate_create(self, ctx, seed, out):
        self._handle(out)
        self.calls.append("sur_create")
        return 0

    def piro_surrogate_num_params(self):
        return 1

    def piro_surrogate_destroy(self, h):
        self.calls.append("sur_destroy")
        return 0

    def piro_batch_create(self, ctx, n, dims, alpha, dt, out):
        self._handle(out)
        self.calls.append("batch_create")
        return 0

    def piro_batch_info(self, h, tot, nnz, offs):
        return 0

    def piro_batch_destroy(self, h):
        self.calls.append("batch_destroy")
        return 0


@pytest.fixture
def fake(monkeypatch):
    f = FakeLib()
    monkeypatch.setattr(core, "lib", f)
    return f


class _Alpha:
    ptr = None


def test_context_closes_children_first(fake, monkeypatch):
    monkeypatch.setattr(core, "_p", lambda a: None)
    ctx = core.Context(0)
    sur = core.Surrogate(ctx, seed=1)
    b = core.Batch(ctx, [(4, 4, 3)], _Alpha(), [0.1])
    ctx.close()
    assert fake.calls[-1] == "ctx_destroy"
    assert fake.calls.count("sur_destroy") == 1 and fake.calls.count("batch_destroy") == 1
    # later closes (e.g. from __del__ at interpreter exit) do nothing
    sur.close()
    b.close()
    del sur, b
    assert fake.calls.count("sur_destroy") == 1 and fake.calls.count("batch_destroy") == 1


def test_child_closed_before_context(fake):
    ctx = core.Context(0)
    sur = core.Surrogate(ctx, seed=1)
    sur.close()
    ctx.close()
    assert fake.calls == ["ctx_create", "sur_create", "sur_destroy", "ctx_destroy"]
