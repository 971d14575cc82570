"""Host-side logic of the learned per-level weights (no GPU): which stored
model a benchmark configuration's learned mode uses, the smoother options
of each kind, the training feature range the inference clips to, and the
constant policy the network is compared with."""
import json

import numpy as np

import gpusolver as gs
import learned


def test_config_kind_prefers_the_hybrid_models(tmp_path):
    p = tmp_path / "w.json"
    p.write_text(json.dumps({"bgs": {}, "sgs": {}}))
    assert learned.config_kind("hetero_elasticity_96", p) == "bgs"
    assert learned.config_kind("varcoef3d_160", p) == "sgs"
    p.write_text(json.dumps({"bgs": {}, "sgs": {}, "hbgs": {}, "hsgs": {}}))
    assert learned.config_kind("elasticity_64", p) == "hbgs"
    assert learned.config_kind("poisson3d_128", p) == "hsgs"
    p.write_text(json.dumps({"hbgs": {}}))
    assert learned.config_kind("poisson2d_256", p) is None
    assert learned.config_kind("hetero_elasticity_96", tmp_path / "missing.json") is None


def test_solver_options_of_each_kind():
    assert learned.solver_options("bgs") == {"smoother": gs.BLOCK_GS}
    assert learned.solver_options("sgs") == {"smoother": gs.SGS}
    # The reference's hybrid block smoother: block GS fine, block Jacobi coarse
    # (src/Preconditioners/MultigridHybridBlockPreconditioner.cc:9-31).
    assert learned.solver_options("hbgs") == {"smoother": gs.BLOCK_GS,
                                              "coarse_smoother": gs.BLOCK_JACOBI}
    assert learned.solver_options("hsgs") == {"smoother": gs.SGS,
                                              "coarse_smoother": gs.DAMPED_JACOBI}
    for kind, spec in learned.KINDS.items():
        if "reference" in spec:  # a hybrid is compared with its all-Gauss-Seidel kind
            assert learned.KINDS[spec["reference"]]["smoother"] == spec["smoother"]
            assert "coarse_smoother" not in learned.KINDS[spec["reference"]]


def test_feature_range_and_constant_policy():
    rows = [{"features": [[40.0, 0.3, 1.0, 4.0, 8.0], [140.0, 0.4, 0.1, 4.5, 23.0]],
             "omega": [0.8, 1.0], "sweeps": [1, 1]},
            {"features": [[44.0, 0.35, 1.0, 5.0, 8.0], [150.0, 0.42, 0.12, 3.5, 22.0]],
             "omega": [1.0, 0.5], "sweeps": [1, 0]},
            {"features": [[43.0, 0.36, 1.0, 4.4, 9.0], [141.0, 0.41, 0.11, 4.2, 24.0]],
             "omega": [0.9, 0.9], "sweeps": [1, 1]}]
    lo, hi = learned.feature_range(rows)
    np.testing.assert_array_equal(lo, [40.0, 0.3, 0.1, 3.5, 8.0])
    np.testing.assert_array_equal(hi, [150.0, 0.42, 1.0, 5.0, 24.0])
    c = learned.constant_policy(rows, learned.KINDS["hbgs"])
    # Most frequent sweep count per level; mean damping of the labels that
    # smooth the level.
    assert c["sweeps"] == [1, 1]
    np.testing.assert_allclose(c["omega"], [0.9, 0.95])


def test_decode_keeps_the_finest_level_smoothed():
    Y = np.array([[0.0, -5.0], [0.4, 5.0], [-2.0, -1.0]])
    omega, sweeps = learned.decode(Y, min_sweeps=0)
    np.testing.assert_allclose(omega, [1.0, 1.2, 0.2])
    assert sweeps.tolist() == [1, 3, 1]
