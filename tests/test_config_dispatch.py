"""The rod constraint follows ConstraintSolverConfig.bending_constraint and
solver_type only sets the Cosserat inertial target (simulate.py:71-77)."""
from __future__ import annotations

from dataclasses import replace

import numpy as np
import pytest

from oracle import step_ref as R
from microcosmos_amd import (COSSERAT_SCHEME, PBD_SCHEME, pbd_rod_constraint,
                             stable_cosserat_constraint)
from microcosmos_amd import workloads as W
from microcosmos_amd._runtime import (CONSTRAINT_COSSERAT, CONSTRAINT_PBD, params_from_config)

from .helpers import oracle_params, run_oracle


@pytest.mark.parametrize("bc,st,kern,target", [
    (pbd_rod_constraint, "pbd", CONSTRAINT_PBD, 0),
    (stable_cosserat_constraint, "cosserat", CONSTRAINT_COSSERAT, 1),
    (stable_cosserat_constraint, "pbd", CONSTRAINT_COSSERAT, 0),   # Cosserat with y = 0
    (pbd_rod_constraint, "cosserat", CONSTRAINT_PBD, 1),           # PBD, debug_vector = prediction
])
def test_params_follow_bending_constraint(bc, st, kern, target):
    p = params_from_config(replace(PBD_SCHEME, bending_constraint=bc, solver_type=st))
    assert p.bending_constraint == kern
    assert p.solver_type == target


def test_custom_constraint_refused():
    with pytest.raises(NotImplementedError):
        params_from_config(replace(PBD_SCHEME, bending_constraint=lambda *a, **k: a[1]))


def test_oracle_mismatched_pairs_differ_from_matched():
    """In the restatement the Cosserat constraint under solver_type "pbd" pulls
    toward y = 0 (debug_vector stays zero), so it differs from the matched
    Cosserat scheme; PBD under solver_type "cosserat" moves exactly like PBD
    but reports the prediction in debug_vector."""
    wl = W.small(seed=11)
    ov = dict(enable_fluid=False, cycles_per_step=4)
    base_c = oracle_params(wl, R.COSSERAT_SCHEME, **ov)
    a = run_oracle(wl, 2, base_c)
    b = run_oracle(wl, 2, replace(base_c, solver_type="pbd", bending_constraint="cosserat"))
    assert np.abs(a.position - b.position).max() > 1e-3
    assert np.all(b.debug_vector == 0) and np.any(a.debug_vector != 0)
    base_p = oracle_params(wl, R.PBD_SCHEME, **ov)
    c = run_oracle(wl, 2, base_p)
    d = run_oracle(wl, 2, replace(base_p, solver_type="cosserat", bending_constraint="pbd"))
    assert np.array_equal(c.position, d.position) and np.array_equal(c.theta, d.theta)
    assert np.all(c.debug_vector == 0) and np.any(d.debug_vector != 0)
