"""The HIP path against the committed dense cfg1 fixture (tests/golden/cfg1_dense.npz, made by
tests/golden/make_cfg1_dense.py from a dense Hamiltonian diagonalised with scipy): H|psi> of a
seeded block, and Davidson eigenvalues at three k-points in FP64 and in mixed precision."""
import os

import numpy as np
import pytest

from q3_amd import FP64, MIXED, PlanewaveDavidsonDiagonalizer, PlanewaveHamiltonianApplicator
from q3_amd import workloads as W
from tests.helpers import check_eigenvalues, kcase, rel_err

pytestmark = pytest.mark.gpu

HA = 2.0
FX = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cfg1_dense.npz"))


def _case(j):
    case = kcase(W.cfg1(), int(FX["kindices"][j]))
    np.testing.assert_array_equal(case.kp.g_vector_indices, FX[f"k{j}_g_index"])
    np.testing.assert_array_equal(case.kp.fft_map, FX[f"k{j}_fft_map"])
    return case


def _applicator(ctx, case):
    h = PlanewaveHamiltonianApplicator(ctx, case.basis)
    h.update_potential(case.V)
    h.prepare_kpoint(case.kp)
    h.set_projectors(case.beta, case.bsizes, case.dblocks)
    return h


def test_hpsi_matches_dense_fixture(gpu_context):
    h = _applicator(gpu_context, _case(0))
    psi = np.asfortranarray(FX["k0_psi"])
    want = FX["k0_hpsi"]
    print("cfg1 k-point 6, H|psi> of 2 bands, FP64 and mixed", flush=True)
    assert rel_err(h.apply_batch(psi, FP64), want) <= 1e-10
    assert rel_err(h.apply_batch(psi, MIXED), want) <= 2e-6


@pytest.mark.parametrize("j", [0, 1, 2])
def test_davidson_matches_dense_fixture(gpu_context, j):
    case = _case(j)
    work = case.w
    exact = FX[f"k{j}_eigenvalues_ry"]
    h = _applicator(gpu_context, case)
    print(f"cfg1 fixture k-point {j}: FP64 Davidson to 1e-10 Ry", flush=True)
    d = PlanewaveDavidsonDiagonalizer(h)
    d.set_tolerance(1e-10)
    d.set_maximum_iterations(200)
    eig, _ = d.solve(work.nbands, work.noccupied, precision=FP64)
    assert d.number_of_unconverged_bands() == 0
    check_eigenvalues(eig, exact, work.noccupied, work.nbands, j)
    print(f"cfg1 fixture k-point {j}: mixed Davidson with the FP64 sweep", flush=True)
    d2 = PlanewaveDavidsonDiagonalizer(h)
    d2.set_tolerance(1e-10)
    d2.set_maximum_iterations(200)
    e32, _ = d2.solve(work.nbands, work.noccupied, precision=MIXED)
    assert d2.number_of_unconverged_bands() == 0
    check_eigenvalues(e32, exact, work.noccupied, work.nbands, j, occupied_bar=1e-6 * HA)
