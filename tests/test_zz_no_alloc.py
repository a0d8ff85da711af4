"""q3x_hamiltonian_apply_noalloc (no device allocation; scratch in the library's fixed per-GPU
array) against q3x_hamiltonian_apply_device on the same inputs, through the C ABI.

Two calls in a row on each workload, the second with fewer bands and other vectors, so that the
second call runs over the scratch contents the first left behind.  The output buffers start
filled with NaN, so a value the product did not write shows.  Bars: FP64 relative error <= 1e-12
(the FP64 H|psi> bar against the oracle is 1e-10), mixed precision <= 2e-6.  cfg2 (160 bands,
2.6 MB of stick buffer per band) runs in band chunks that fit the 64 MiB array.
"""
import numpy as np
import pytest

from q3_amd import FP64, MIXED, Context, PlanewaveHamiltonianApplicator
from q3_amd import workloads as W
from q3_amd._lib import lib
from q3_amd.planewave import DeviceArray
from tests.helpers import kcase, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def _run(ctx, h, psi, precision, noalloc):
    npw, nb = psi.shape
    d_psi = DeviceArray.from_host(ctx, psi)
    d_h = DeviceArray.from_host(ctx, np.full((npw, nb), np.nan + 1j * np.nan, np.complex128))
    if noalloc:
        h.apply_noalloc(d_psi.ptr, d_h.ptr, nb, precision)
    else:
        h.apply_device(d_psi.ptr, d_h.ptr, nb, precision)
    ctx.synchronize()
    out = d_h.download((npw, nb))
    d_psi.free()
    d_h.free()
    return out


def test_scratch_size_is_stated():
    assert lib().q3x_noalloc_scratch_bytes() == 64 << 20


@pytest.mark.parametrize("make,kindex,counts", [(W.cfg4, 1, (48, 17)), (W.cfg1, 2, (8, 3)), (W.cfg2, 0, (160, 40))])
@pytest.mark.parametrize("precision,tol", [(FP64, 1e-12), (MIXED, 2e-6)])
def test_noalloc_equals_apply_device(ctx, make, kindex, counts, precision, tol):
    case = kcase(make(), kindex)
    h = PlanewaveHamiltonianApplicator(ctx, case.basis)
    h.update_potential(case.V)
    h.prepare_kpoint(case.kp)
    h.set_projectors(case.beta, case.bsizes, case.dblocks)
    for call, nb in enumerate(counts):
        psi = W.random_psi(case.kp.kinetic_energy_squared, nb, seed=40 + call)
        print(f"{case.w.name} k{kindex} call {call}: {nb} bands, npw {case.kp.npw}, precision {precision}",
              flush=True)
        got = _run(ctx, h, psi, precision, noalloc=True)
        want = _run(ctx, h, psi, precision, noalloc=False)
        assert np.isfinite(got).all()
        assert rel_err(got, want) <= tol
