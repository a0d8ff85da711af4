"""The serving request stream (seed 4) keeps the stated ranges: Poisson arrivals at 20 req/s,
2-216 atoms, C fraction in [0, 0.5] with the lattice constant interpolated between Si and C,
ecutwfc in [15, 25] Ha, 8-40 k-points on a straight line, bands = 2 N_atoms + 8."""
import numpy as np

from q3_amd import workloads as W


def test_stream_ranges_and_rate():
    reqs = W.cfg5_requests(2000)
    t = np.array([r.arrival_s for r in reqs])
    assert np.all(np.diff(t) > 0)
    assert abs(len(reqs) / t[-1] - 20.0) < 1.5            # Poisson rate
    gaps = np.diff(np.r_[0.0, t])
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.1       # exponential gaps
    nat = np.array([len(r.crystal.species) for r in reqs])
    assert nat.min() == 2 and nat.max() == 216
    for r in reqs[:200]:
        n = len(r.crystal.species)
        assert r.nbands == 2 * n + 8 and r.noccupied == 2 * n
        assert 0.0 <= r.carbon_fraction <= 0.5
        a_expect = W.A_SI + (W.A_C - W.A_SI) * r.carbon_fraction
        assert abs(r.crystal.alat - a_expect) < 1e-12
        assert 15.0 <= r.ecutwfc_ry / 2 <= 25.0
        assert 8 <= len(r.kpoints) <= 40
        d = np.diff(r.kpoints, axis=0)
        assert np.allclose(d, d[0])                        # straight, evenly spaced path


def test_stream_is_deterministic_and_capped():
    a = W.cfg5_requests(30, seed=4)
    b = W.cfg5_requests(30, seed=4)
    assert all(x.arrival_s == y.arrival_s and np.array_equal(x.kpoints, y.kpoints) for x, y in zip(a, b))
    small = W.cfg5_requests(200, max_atoms=16)
    assert max(len(r.crystal.species) for r in small) <= 16
