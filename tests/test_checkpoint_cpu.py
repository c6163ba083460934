"""Checkpoint writer of the least-squares solver, on host tensors (no GPU):
writes land in path order, a write left in flight by one call is finished
before the next call's first write on the same path, wait_checkpoints()
drains it, and the file is only ever a whole checkpoint (written to
path.tmp, then renamed)."""
import os

import torch

from poststackmig3d_amd import lsm


def _state(it, n=1000):
    g = torch.Generator().manual_seed(it)
    return {"m": torch.randn(n, generator=g), "r": torch.randn(n, generator=g),
            "p": torch.randn(n, generator=g), "gamma": float(it), "iteration": it,
            "residual_norms": [float(i) for i in range(it + 1)],
            "gradient_norms": [float(i) for i in range(it + 1)]}


def _load(path):
    return torch.load(path, map_location="cpu", weights_only=True)


def test_writes_in_order_and_drain(tmp_path):
    ck = str(tmp_path / "cg.pt")
    w = lsm._CheckpointWriter(ck)
    for it in (1, 2, 3):
        w.submit(_state(it))
    w.detach()                      # the last write may still be in flight
    w2 = lsm._CheckpointWriter(ck)  # takes it over: finished before w2's first write
    assert w2.prev is w or w2.prev is None
    s5 = _state(5)
    w2.submit(s5)
    s5["m"].zero_()                 # the writer copied the tensors
    w2.detach()
    lsm.wait_checkpoints(ck)
    st = _load(ck)
    assert st["iteration"] == 5
    assert torch.equal(st["m"], _state(5)["m"])
    assert st["residual_norms"] == _state(5)["residual_norms"]
    assert not os.path.exists(ck + ".tmp")
    lsm.wait_checkpoints()          # nothing pending: a no-op


def test_detach_without_writes_releases(tmp_path):
    ck = str(tmp_path / "cg.pt")
    w = lsm._CheckpointWriter(ck)
    w.submit(_state(1))
    w.detach()
    w2 = lsm._CheckpointWriter(ck)
    w2.detach()                     # no write of its own: waits for the earlier one
    assert ck not in lsm._pending
    assert _load(ck)["iteration"] == 1


def test_write_error_is_raised_by_wait(tmp_path):
    ck = str(tmp_path / "missing" / "cg.pt")   # no such directory
    w = lsm._CheckpointWriter(ck)
    w.submit(_state(1))
    w.detach()
    try:
        lsm.wait_checkpoints(ck)
    except (OSError, RuntimeError):   # (torch.save reports a missing directory as RuntimeError)
        pass
    else:
        raise AssertionError("the failed write was not reported")
    assert ck not in lsm._pending
