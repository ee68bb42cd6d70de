"""CPU-side checks of libbrownian: the library loads, exports every entry point
declared in include/brownian.h, and its host start-up reproduces the
reference's initial state (golden vectors) without a GPU."""
import ast
import ctypes
import os

import numpy as np
import pytest

import brownian_amd as B

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module", autouse=True)
def _lib():
    if not os.path.exists(B.LIB_PATH):
        pytest.fail("libbrownian.so not built (run __graft_entry__.build())")
    return B.lib()


def test_exports_every_declared_symbol():
    names = B.declared_symbols()
    assert len(names) >= 40
    so = ctypes.CDLL(B.LIB_PATH)
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing
    # the reference's entry-point names (cpp/parameters.hpp:12-14)
    for n in ["initialize", "initialize_colloid", "initialize_fluid", "create_box", "neighbour_list_mpcd",
              "neighbour_list_md", "compute_force_md", "tumble", "rotation_mpcd", "run", "update_pos_md",
              "update_pos_mpcd", "fluid_colloid_collision", "update_activity_direction",
              "update_velocity_colloid"]:
        assert n in names


def test_default_params_are_reference_defaults():
    p = B.default_params()
    assert (p.L, p.density, p.n_colloid, p.n_sub, p.seed) == (30, 10, 10, 10, 77777)
    assert (p.ndt, p.mass_colloid, p.sig_colloid, p.theta_deg) == (0.1, 654.1, 5.0, 130.0)
    assert p.sigma == 0.8 * 5.0


@pytest.mark.parametrize("case", ["small", "c0like"])
def test_host_startup_matches_reference(case):
    g = dict(np.load(os.path.join(GOLD, case + ".npz")))
    p = dict(ast.literal_eval(str(g["params_json"])))
    out = B.host_initialize(**{k: v for k, v in p.items()})
    np.testing.assert_array_equal(out["pos"], g["start_pos"])
    np.testing.assert_array_equal(out["vel"], g["start_vel"])
    np.testing.assert_array_equal(out["c_pos"], g["start_c_pos"])
    np.testing.assert_array_equal(out["c_vel"], g["start_c_vel"])
    np.testing.assert_array_equal(out["c_omega"], g["start_c_omega"])
    np.testing.assert_array_equal(out["c_ra"], g["start_c_ra"])


def test_host_generator_matches_reference_stream():
    """The library's host generator state after start-up equals the oracle's
    (which is pinned to the reference stream)."""
    import oracle as O
    g = dict(np.load(os.path.join(GOLD, "small.npz")))
    p = dict(ast.literal_eval(str(g["params_json"])))
    out = B.host_initialize(**p)
    o = O.Oracle(**p)
    o.startup()
    np.testing.assert_array_equal(out["rng"], o.rng_state())


def test_maxwell_and_sphere_laws():
    out = B.host_initialize(L=10, density=10, n_colloid=20, sig_colloid=2.0, sigma=2.0, vel_law=1, orient_law=1,
                            seed=99)
    v = out["vel"]
    assert np.abs(v.mean(axis=0)).max() < 1e-12
    assert abs(v.var() - 1.0) < 0.03          # kT/m = 1
    k = ((v - v.mean(0)) ** 4).mean() / v.var() ** 2
    assert abs(k - 3.0) < 0.15                # Gaussian kurtosis (uniform law gives 1.8)
    ra = out["c_ra"]
    np.testing.assert_allclose(np.linalg.norm(ra, axis=1), 1.0, atol=1e-12)


def test_create_rejects_bad_parameters_without_gpu_state():
    # L < 7 would give duplicate stencil cells (cpp/neighbour_list.cpp:12-25)
    with pytest.raises(B.BrownianError):
        B.System(L=6, density=2, n_colloid=0)


def test_autotuner_candidate_shapes_divide_every_box_edge():
    """The tile shapes the autotuners try: each divides its box edge (one
    shape for all the edges of an ensemble), and with `even` the ensemble
    search refuses edge sets with no even-edged common shape before it starts
    any system (the surrogate's 2x2x2 coarse cells need even edges)."""
    from brownian_amd.autotune import autotune_ensemble, candidates, common_candidates
    for L in (16, 24, 32, 40, 48):
        for shape in candidates(L, 10):
            assert all(L % e == 0 for e in shape), (L, shape)
    common = common_candidates([16, 24, 32, 40, 48], 10)
    assert common and all(all(L % e == 0 for e in s) for s in common for L in (16, 24, 32, 40, 48))
    assert all(e % 2 == 0 for s in common for e in s)
    assert common_candidates([20, 30], 10) and any(e % 2 for s in common_candidates([20, 30], 10) for e in s)
    with pytest.raises(ValueError):
        autotune_ensemble([dict(L=20), dict(L=30)], even=True)
