"""k-path requests on the band-structure server: the reference's non-self-consistent k-loop
(forward propagation along the path, no propagation out of a segment boundary, the backward fix
over the edge k-points; iterator/kohn_sham/planewave/pseudopotential/init.hpp:1036-1106,
1599-1815) run by one worker on device-resident wavefunctions.

* Against the oracle's k-loop (oracle.nscf_band_path over the oracle's Davidson) on nine points of
  a 2-atom cell's line in two segments: every band within 1e-8 Ha (every band held to 1e-11 Ry),
  as many solves as the oracle's schedule, every k-point solved.
* Against the same k-loop driven from the host through the diagonalizer (the host generator's
  starting blocks, propagation by numpy, stored eigenvalues passed in): within 1e-10 Ry.
* The same request with independent k-points reaches the same eigenvalues.
* Cancelling while k-path requests run: they stop before their next solve and end with an error;
  the server then serves a new request.
"""
import time

import numpy as np
import pytest

from oracle import oracle as O
from q3_amd import BandRequestSpec, BandStructureServer, PlanewaveBasis, PlanewaveDavidsonDiagonalizer
from q3_amd import PlanewaveHamiltonianApplicator
from q3_amd import workloads as W
from tests.helpers import oracle_solver, small_path_problem

pytestmark = pytest.mark.gpu

HA = 2.0
NK, SEGMENTS, PPS = 9, 2, 4
ECUT = 16.0


def _spec(r, ks, tol, kpath=1, segments=SEGMENTS, pps=PPS, max_iter=400):
    sp = W.serving_species()
    basis = PlanewaveBasis(r.crystal, ECUT, 0.0, True, (1, 1, 1))
    V = W.local_potential(r.crystal, basis, sp)
    bs, db = W.d_blocks_for(r.crystal, sp)
    return BandRequestSpec(basis, V, np.asarray(ks), sp, bs, db, r.nbands, r.nbands, tolerance_ry=tol,
                           max_iterations=max_iter, kpath_propagation=kpath, band_path_segments=segments,
                           band_path_points_per_segment=pps)


@pytest.fixture(scope="module")
def path_problem():
    return small_path_problem(NK, ECUT)


def test_kpath_request_against_the_oracle_kloop(path_problem):
    r, ob, V, bs, db, ks, per_k = path_problem
    nb = r.nbands
    solve, initial = oracle_solver(ob, V, bs, db, per_k, nb, 1e-11)
    oe, reps, steps = O.nscf_band_path(solve, initial, [len(p[0]) for p in per_k], nb, SEGMENTS, PPS)
    assert all(rep["unconverged"] == 0 for rep in reps)
    srv = BandStructureServer(workers=2)
    eig, st = srv.wait(srv.submit(_spec(r, ks, 1e-11)))
    srv.close()
    err = np.abs(eig - oe).max()
    print(f"k-path vs oracle k-loop: {err / HA:.2e} Ha; {st['solves']} solves, {st['iterations']} iterations "
          f"(oracle {sum(rep['iterations'] for rep in reps)})", flush=True)
    assert st["solves"] == len(steps) == NK + 6 + 5
    assert st["kpoints_done"] == NK
    assert err <= 1e-8 * HA


def test_kpath_request_equals_the_host_driven_kloop(gpu_context, path_problem):
    r, _ob, _V, _bs, _db, ks, _per_k = path_problem
    spec = _spec(r, ks, 2e-9)
    srv = BandStructureServer(workers=2)
    eig, st = srv.wait(srv.submit(spec))
    srv.close()
    b = spec.basis
    h = PlanewaveHamiltonianApplicator(gpu_context, b)
    h.update_potential(spec.v_r)
    kps = [b.compute_gk(k) for k in ks]

    def solve(k, psi0, stored):
        h.prepare_kpoint(kps[k])
        h.build_projectors(b.crystal, spec.species, spec.block_sizes, spec.d_blocks)
        d = PlanewaveDavidsonDiagonalizer(h)
        d.set_tolerance(spec.tolerance_ry)
        d.set_maximum_iterations(spec.max_iterations)
        e, p = d.solve(spec.nbands, spec.noccupied, psi0, stored_eigenvalues=stored)
        return e, p, d.report_dict()

    he, reps, steps = O.nscf_band_path(solve, lambda k: kps[k].random_psi(spec.nbands, k),
                                       [kp.npw for kp in kps], spec.nbands, SEGMENTS, PPS)
    d = np.abs(eig - he).max()
    print(f"server vs host-driven k-loop: {d:.1e} Ry; iterations {st['iterations']} vs "
          f"{sum(rep['iterations'] for rep in reps)}", flush=True)
    assert d <= 1e-10


def test_independent_kpoints_reach_the_same_eigenvalues(path_problem):
    r, _ob, _V, _bs, _db, ks, _per_k = path_problem
    srv = BandStructureServer(workers=4)
    tp = srv.submit(_spec(r, ks, 2e-9))
    ti = srv.submit(_spec(r, ks, 2e-9, kpath=0))
    (ep, sp), (ei, si) = srv.wait(tp), srv.wait(ti)
    srv.close()
    print(f"k-path {sp['solves']} solves / {sp['iterations']} iterations; independent {si['solves']} / "
          f"{si['iterations']}; max difference {np.abs(ep - ei).max():.1e} Ry", flush=True)
    assert si["solves"] == NK and si["kpoints_done"] == NK
    assert np.abs(ep - ei).max() <= 1e-8 * HA


def test_cancel_stops_running_kpaths():
    r = next(q for q in W.cfg5_requests(80, seed=4) if len(q.crystal.species) == 2)
    ks = r.kpoints[0][None, :] + np.linspace(0.0, 1.0, 40)[:, None] * (r.kpoints[-1] - r.kpoints[0])[None, :]
    n_steps = len(O.nscf_schedule(40, 1, 39))
    srv = BandStructureServer(workers=1)
    tickets = [srv.submit(_spec(r, ks, 1e-10, segments=1, pps=39)) for _ in range(2)]
    time.sleep(0.05)
    srv.cancel_pending()
    results = [srv.wait_status(t) for t in tickets]
    for status, _eig, st in results:
        print(f"status {status}: {st['solves']} of {n_steps} solves, {st['kpoints_done']} k-points", flush=True)
        if status != 0:
            assert status == 3 and st["solves"] < n_steps
    assert any(status == 3 for status, _e, _s in results)
    # the server still serves
    eig, st = srv.wait(srv.submit(_spec(r, ks[:3], 1e-8, segments=1, pps=2)))
    srv.close()
    assert st["solves"] == len(O.nscf_schedule(3, 1, 2)) and np.all(np.isfinite(eig))


def test_server_estimate_is_the_itemised_footprint(path_problem):
    from q3_amd.planewave import band_request_footprint
    r, _ob, _V, _bs, _db, ks, _per_k = path_problem
    srv = BandStructureServer(workers=3)
    for kpath in (0, 1):
        spec = _spec(r, ks, 1e-8, kpath=kpath)
        assert srv.estimate(spec) == band_request_footprint(spec, 1 if kpath else 3)["bytes"]
    srv.close()
