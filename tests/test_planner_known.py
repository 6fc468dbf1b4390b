"""The oracle's planner (or_plan) against hand-derived known answers of the
reference planner at num_new_configs = 1 (tests/planner_known.py: the
reference's one deterministic case), in both of its modes (one slot makes
them equal), and the planted mistakes each of those answers must catch."""
import numpy as np
import pytest

import planner_known as K


@pytest.fixture(scope="module")
def point_robot(oracle_robots):
    return oracle_robots["sphere"]


def oracle_plan(R, mode=0):
    def plan(scene, st):
        _, res, path = R.plan(scene, K.START, K.GOAL[None], st, mode=mode)
        return bool(res.solved), res.iters, res.start_tree_size, res.goal_tree_size, path, res.cost
    return plan


@pytest.mark.parametrize("mode", [0, 1])
def test_oracle_known_answers(point_robot, mode):
    bad = K.check_all(oracle_plan(point_robot, mode))
    assert not bad, "\n".join(bad)


def test_known_answer_samples_are_the_closed_form():
    """The side table of the docstring follows from phi_3(k) alone."""
    for k, x in [(1, -1.6667), (4, -0.5556), (10, -1.2963), (25, 1.2963), (30, -3.7654)]:
        assert abs(10 * K.radical_inverse(3, k) - 5 - x) < 1e-4


@pytest.mark.parametrize("plant", ["iter1_tree", "b1_ratio", "b1_split", "b2_rule", "dd_sign", "dd_skip",
                                   "chain_drop", "steer", "path_order", "n_ext"])
def test_planted_mistakes_are_caught(point_robot, plant):
    from oracle import oracle
    oracle.set_plant(plant)
    try:
        bad = K.check_all(oracle_plan(point_robot))
    finally:
        oracle.set_plant(None)
    assert bad, f"planted mistake {plant} passes every known answer"
    assert not K.check_all(oracle_plan(point_robot))
