"""Pins of the oracle's swimmer fitness (group centre, displacement tracker)
by closed forms, float64 and invariants.  CPU only."""
from __future__ import annotations

import numpy as np

from oracle import step_ref as R


def test_lane_tree_mean_small_exact():
    # three rows: lanes 0..2 hold one row each; the tree adds them as
    # ((r0 + r2) + r1) in float32 -- equal to the exact mean here
    d = np.array([[1.0, 2.0], [3.0, 4.0], [5.0, 9.0]], np.float32)
    np.testing.assert_array_equal(R._lane_tree_mean(d), np.array([3.0, 5.0], np.float32))


def test_lane_tree_mean_matches_float64_mean():
    rng = np.random.default_rng(1)
    for n in (1, 7, 255, 256, 257, 1000, 4097):
        d = rng.uniform(-30, 30, (n, 2))
        m32 = R._lane_tree_mean(d.astype(np.float32)).astype(np.float64)
        np.testing.assert_allclose(m32, d.mean(axis=0), rtol=0, atol=2e-5)
        np.testing.assert_allclose(R._lane_tree_mean(d), d.mean(axis=0), rtol=1e-13, atol=1e-12)


def test_periodic_com_straddling_boundary_closed_form():
    # nodes at x = W - 1 and x = 0.5: minimum-image mean sits at W - 0.25
    grid = (32, 64)
    pos = np.array([[63.0, 5.0], [0.5, 5.0]], np.float32)
    com = R.group_com(pos, grid)[0]
    np.testing.assert_array_equal(com, np.array([63.75, 5.0], np.float32))
    # the plain mean is the naive midpoint
    np.testing.assert_array_equal(R.group_com(pos, grid, periodic=False)[0],
                                  np.array([31.75, 5.0], np.float32))


def test_periodic_com_equals_mean_for_compact_blob_and_is_translation_invariant():
    rng = np.random.default_rng(2)
    grid = (48, 80)
    blob = rng.uniform(-3, 3, (300, 2)) + np.array([70.0, 40.0])  # crosses y = 48 and x = 80
    wrapped = np.mod(blob, np.array([80.0, 48.0]))
    com = R.group_com(wrapped, grid)[0]
    expect = np.mod(blob.mean(axis=0), np.array([80.0, 48.0]))
    np.testing.assert_allclose(com, expect, atol=1e-10)
    # shifting every node by a lattice period changes nothing
    shifted = wrapped + np.array([80.0, -48.0])
    np.testing.assert_allclose(R.group_com(shifted, grid)[0], com, atol=1e-10)


def test_group_offsets_split_groups():
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, 16, (10, 2))
    com = R.group_com(pos, (16, 16), [0, 4, 10])
    np.testing.assert_allclose(com[0], R.group_com(pos[:4], (16, 16))[0])
    np.testing.assert_allclose(com[1], R.group_com(pos[4:], (16, 16))[0])


def test_swim_track_rigid_translation_closed_form():
    # a rigid body moving by v per step across the periodic boundary:
    # cumulative displacement after k updates is k v (float64)
    grid = (20, 30)
    v = np.array([0.7, -0.45])
    body = np.array([[1.0, 1.0], [2.0, 1.5], [3.0, 2.5]]) + np.array([25.0, 3.0])
    frames = [np.mod(body + k * v, np.array([30.0, 20.0])) for k in range(40)]
    hist = R.swim_track(frames, grid)
    for k in range(1, 40):
        np.testing.assert_allclose(hist[k - 1][0], k * v, atol=1e-9)
    # displacement_start_step: the first 10 updates do not accumulate
    hist_s = R.swim_track(frames, grid, start_step=10)
    np.testing.assert_allclose(hist_s[9][0], 0.0, atol=0)
    np.testing.assert_allclose(hist_s[-1][0], (39 - 10) * v, atol=1e-9)
