"""The band-structure server at the serving stream's stated sizes (cfg5: Si/C alloy cells of 2 to
216 atoms, 15-25 Ha, bands = 2 N_atoms + 8).

* The largest classes, a 64-atom and a 216-atom cell of the stream at 25 Ha (the top of the cutoff
  range; 440 bands, subspace n = 880 for the 216-atom cell), one k-point each, served together:
  each equals a direct Davidson solve of the same k-point from the same starting block
  (q3x_kpoint_random_psi of the k-point's index, projectors built on the device).
* A 2-atom request of the stream against the FP64 oracle's eigenvalues
  (tests/golden/cfg5_small.npz, tests/golden/make_cfg5_small.py): every band within 1e-8 Ha.
* Cancelling pending work: requests not yet started end with an error, running ones finish.
"""
import os

import numpy as np
import pytest

from q3_amd import BandRequestSpec, BandStructureServer, PlanewaveBasis, PlanewaveDavidsonDiagonalizer
from q3_amd import PlanewaveHamiltonianApplicator
from q3_amd import workloads as W
from q3_amd._lib import StateError

pytestmark = pytest.mark.gpu

HA = 2.0
FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cfg5_small.npz")


def _first_with_atoms(n, count=80):
    for r in W.cfg5_requests(count, seed=4):
        if len(r.crystal.species) == n:
            return r
    raise AssertionError(f"no {n}-atom request among the first {count}")


def _spec(r, ecut_ry, nk, tol, max_iter=80, every_band=False):
    """every_band: every band held to the tolerance (n_occupied = nb), as the oracle fixture is."""
    sp = W.serving_species()
    basis = PlanewaveBasis(r.crystal, ecut_ry, 0.0, True, (1, 1, 1))
    V = W.local_potential(r.crystal, basis, sp)
    bs, db = W.d_blocks_for(r.crystal, sp)
    return BandRequestSpec(basis, V, r.kpoints[:nk], sp, bs, db, r.nbands, r.nbands if every_band else r.noccupied,
                           tolerance_ry=tol, max_iterations=max_iter)


def _direct(ctx, spec, k_index):
    kp = spec.basis.compute_gk(spec.kpoints[k_index])
    h = PlanewaveHamiltonianApplicator(ctx, spec.basis)
    h.update_potential(spec.v_r)
    h.prepare_kpoint(kp)
    h.build_projectors(spec.basis.crystal, spec.species, spec.block_sizes, spec.d_blocks)
    d = PlanewaveDavidsonDiagonalizer(h)
    d.set_tolerance(spec.tolerance_ry)
    d.set_maximum_iterations(spec.max_iterations)
    e, _ = d.solve(spec.nbands, spec.noccupied, kp.random_psi(spec.nbands, k_index))
    return e


def test_largest_cell_classes_at_25_ha_equal_direct_solves(gpu_context):
    specs = [_spec(_first_with_atoms(n), 25.0 * HA, 1, 2e-6) for n in (64, 216)]
    for s in specs:
        print(f"request: {len(s.basis.crystal.species)} atoms, {s.nbands} bands, box {s.basis.wave_dims}",
              flush=True)
    srv = BandStructureServer(workers=2)
    tickets = [srv.submit(s) for s in specs]
    res = [srv.wait(t) for t in tickets]
    srv.close()
    for s, (eig, st) in zip(specs, res):
        print(f"{len(s.basis.crystal.species)} atoms: served in {st['done_s'] - st['submit_s']:.1f} s, "
              f"{st['iterations']} iterations", flush=True)
        direct = _direct(gpu_context, s, 0)
        assert np.all(np.isfinite(eig)) and np.all(np.diff(eig[0]) >= -1e-12)
        assert np.abs(eig[0] - direct).max() <= 1e-10


def test_small_request_against_the_oracle():
    fx = np.load(FIX)
    r = W.cfg5_requests(int(fx["request"]) + 1, seed=4)[int(fx["request"])]
    spec = _spec(r, float(fx["ecutwfc_ry"]), 2, 1e-11, max_iter=500, every_band=True)
    srv = BandStructureServer(workers=2)
    eig, _ = srv.wait(srv.submit(spec))
    srv.close()
    assert eig.shape == fx["eigenvalues_ry"].shape
    assert np.abs(eig - fx["eigenvalues_ry"]).max() <= 1e-8 * HA


def test_cancel_pending_ends_waiting_requests():
    reqs = [r for r in W.cfg5_requests(40, seed=4) if len(r.crystal.species) <= 8][:6]
    specs = [_spec(r, 12.0 * HA, 6, 1e-8) for r in reqs]
    srv = BandStructureServer(workers=1)
    tickets = [srv.submit(s) for s in specs]
    n = srv.cancel_pending()
    assert 1 <= n <= len(specs)
    outcomes = []
    for t in tickets:
        try:
            e, _ = srv.wait(t)
            outcomes.append("done")
            assert np.all(np.isfinite(e))
        except StateError as ex:
            assert "cancelled" in str(ex)
            outcomes.append("cancelled")
    srv.close()
    assert outcomes.count("cancelled") == n
    # the server still serves after a cancel
    srv = BandStructureServer(workers=1)
    e, _ = srv.wait(srv.submit(specs[0]))
    srv.close()
    assert np.all(np.isfinite(e))
