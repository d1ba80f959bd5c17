"""Device-side failure reporting.  The wide immersed-boundary solve's grid
barrier waits are bounded; a wait that gives up is reported as an error by
the next call on the plan instead of returning wrong forces silently.  The
condition is forced from the host: MC_IB_WIDE_POLL_LIMIT=0 makes every wait
that is not already complete give up at once, so the kernels exit normally."""
from __future__ import annotations

import sys

import pytest

from microcosmos_amd import workloads as W

from .helpers import device_config, device_state

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


def test_wide_barrier_timeout_is_reported(dev, monkeypatch):
    print("wide IB solve with the barrier poll bound at 0 (forced timeout)", flush=True)
    sys.stdout.flush()
    from microcosmos_amd import Rollout
    from microcosmos_amd._lib import MicrocosmosError
    from microcosmos_amd._runtime import clear_plans
    # 48 x 40 nodes on 64^2 chain into one cluster above 768 nodes: the wide
    # solve runs on 10 workgroups with grid barriers
    wl = W.small(grid=(64, 64), n_fil=48, n_nodes=40, seed=3)
    monkeypatch.setenv("MC_IB_WIDE_POLL_LIMIT", "0")
    clear_plans()
    try:
        nodes, edges, fields = device_state(wl, dev)
        ro = Rollout(nodes, edges, fields, device_config(wl), wl.dt,
                     actuation=(wl.act_amp, wl.act_omega, wl.act_phase))
        ro.run(3, use_graph=False)
        with pytest.raises(MicrocosmosError, match="barrier"):
            ro.check()
        # reported once; the idle IB state was restored on the stream
        ro.check()
        assert torch.isfinite(ro.position).all()
        # a timeout in later work is reported again
        ro.run(2, use_graph=False)
        with pytest.raises(MicrocosmosError, match="barrier"):
            ro.check()
        del ro
    finally:
        monkeypatch.delenv("MC_IB_WIDE_POLL_LIMIT", raising=False)
        clear_plans()
        torch.cuda.synchronize()


def test_default_poll_bound_reports_nothing(dev):
    """The same crowded case with the default bound completes without error."""
    print("wide IB solve with the default barrier poll bound", flush=True)
    from microcosmos_amd import Rollout
    from microcosmos_amd._runtime import clear_plans
    clear_plans()
    wl = W.small(grid=(64, 64), n_fil=48, n_nodes=40, seed=3)
    nodes, edges, fields = device_state(wl, dev)
    ro = Rollout(nodes, edges, fields, device_config(wl), wl.dt,
                 actuation=(wl.act_amp, wl.act_omega, wl.act_phase))
    ro.run(3, use_graph=True)
    ro.check()
