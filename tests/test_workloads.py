"""The benchmark's workload generator (mr_aha_amd.workloads, which the
benchmark and the serving leg use) produces the same inputs as the oracle's
generator that wrote the golden fixtures (oracle.generators), at the
fixtures' recipes and seeds: identical random draws (x bit for bit, the
basis and its Gram blocks to double rounding) and the same closed-form
trajectories and coil maps to float32 rounding."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import generators as gen  # noqa: E402


def _cmp(w, ref):
    g, traj, phi2, x, S, _ = ref
    assert w["traj"].shape == traj.shape and w["phi2"].shape == phi2.shape
    assert np.max(np.abs(w["traj"].numpy() - traj)) <= 4e-7
    # the same products in complex128 before the cast: torch and numpy round
    # the imaginary part of a diagonal a conj(a) differently (0 vs ~1e-21)
    assert np.max(np.abs(w["phi2"].numpy() - phi2)) <= 1e-12 * np.max(np.abs(phi2))
    np.testing.assert_array_equal(w["x"].numpy(), x)
    assert np.max(np.abs(w["SEs"].numpy() - S)) <= 1e-6


def test_radial_matches_fixture_generator():
    from mr_aha_amd import workloads
    _cmp(workloads.radial_2d(device="cpu"), gen.radial_workload())


def test_stack_matches_fixture_generator():
    from mr_aha_amd import workloads
    kw = dict(n=40, nz=8, L=16, C=12, n_spokes=48, n_frames=200)
    w = workloads.stack_of_stars(nkx=80, device="cpu", **kw)
    _cmp(w, gen.stack_workload(n_kx=80, **kw))


def test_kooshball_matches_fixture_generator():
    from mr_aha_amd import workloads
    kw = dict(n=24, L=8, C=8, n_spokes=400, n_frames=100)
    w = workloads.kooshball(nkx=48, device="cpu", **kw)
    _cmp(w, gen.kooshball_workload(n_kx=48, **kw))
