"""Device generalized Hermitian eigensolver (q3x_hegv_device) against scipy.linalg.eigh:
eigenvalues, residuals H X - S X w and S-orthonormality, on the sizes the Davidson subspace
takes (LDS-resident and L2-resident tridiagonalisation), with exactly degenerate spectra, an
S far from the identity (Newton-Schulz branch), only the upper triangles read, a leading
dimension above n, and a non-positive-definite S reported as Q3X_ERR_NUMERIC."""
import numpy as np
import pytest
import scipy.linalg as sl

from q3_amd._lib import Q3XError
from q3_amd.planewave import DeviceArray, hegv_device

pytestmark = pytest.mark.gpu


def _herm(n, rng, spectrum=None):
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    if spectrum is None:
        return A + A.conj().T
    Q, _ = np.linalg.qr(A)
    return (Q * spectrum) @ Q.conj().T


def _solve(ctx, H, S, nvec, ld=None, garbage_lower=False):
    n = H.shape[0]
    ld = ld or n
    rng = np.random.default_rng(99)
    Hb = np.zeros((ld, n), complex, order="F")
    Sb = np.zeros((ld, n), complex, order="F")
    Hb[:n, :n] = H
    Sb[:n, :n] = S
    if garbage_lower:  # only the upper triangles may be read
        il = np.tril_indices(n, -1)
        Hb[il] = rng.standard_normal(len(il[0])) * 1e3
        Sb[il] = rng.standard_normal(len(il[0])) * 1e3
    dH, dS = DeviceArray.from_host(ctx, Hb), DeviceArray.from_host(ctx, Sb)
    dX = DeviceArray(ctx, max(ld * max(nvec, 1) * 16, 16))
    print(f"hegv_device n={n} nvec={nvec} ld={ld}", flush=True)
    w = hegv_device(ctx, n, dH, ld, dS, ld, nvec, dX, ld)
    X = dX.download((ld, max(nvec, 1)))[:n, :nvec]
    return w, X


def _check(H, S, w, X, nvec, tol=1e-12):
    ref = sl.eigh(H, S, eigvals_only=True)[:nvec]
    scale = max(np.abs(ref).max(), 1.0)
    assert np.abs(w - ref).max() <= tol * scale, np.abs(w - ref).max()
    assert np.all(np.diff(w) >= -tol * scale)
    res = H @ X - (S @ X) * w
    assert np.abs(res).max() <= 50 * tol * scale * np.sqrt(H.shape[0]), np.abs(res).max()
    G = X.conj().T @ S @ X
    assert np.abs(G - np.eye(nvec)).max() <= 50 * tol * np.sqrt(H.shape[0]), np.abs(G - np.eye(nvec)).max()


@pytest.mark.parametrize("n,nvec", [(1, 1), (2, 2), (3, 3), (5, 2), (17, 17), (64, 32), (96, 48), (99, 99),
                                    (100, 50), (160, 160), (250, 125), (320, 160), (450, 100),
                                    (700, 64), (880, 440)])
def test_hegv_device_near_identity_overlap(gpu_context, n, nvec):
    rng = np.random.default_rng(n)
    H = _herm(n, rng)
    E = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    S = np.eye(n) + 1e-12 * (E + E.conj().T)
    w, X = _solve(gpu_context, H, S, nvec)
    _check(H, S, w, X, nvec)


@pytest.mark.parametrize("n", [40, 120, 320])
def test_hegv_device_degenerate_spectrum(gpu_context, n):
    rng = np.random.default_rng(7 + n)
    spec = np.repeat(np.round(rng.standard_normal(n // 4), 2), 4)[:n]  # 4-fold degenerate levels
    spec = np.concatenate([spec, rng.standard_normal(n - len(spec))])
    H = _herm(n, rng, spec)
    w, X = _solve(gpu_context, H, np.eye(n), n // 2)
    _check(H, np.eye(n), w, X, n // 2)


@pytest.mark.parametrize("n", [30, 150])
def test_hegv_device_general_overlap(gpu_context, n):
    rng = np.random.default_rng(3 * n)
    H = _herm(n, rng)
    B = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    S = B @ B.conj().T / n + 0.05 * np.eye(n)  # condition number ~1e2..1e3
    w, X = _solve(gpu_context, H, S, n // 3)
    _check(H, S, w, X, n // 3, tol=1e-10)


def test_hegv_device_upper_triangle_and_leading_dimension(gpu_context):
    n, ld = 77, 101
    rng = np.random.default_rng(5)
    H = _herm(n, rng)
    S = np.eye(n)
    w, X = _solve(gpu_context, H, S, 40, ld=ld, garbage_lower=True)
    _check(H, S, w, X, 40)


def test_hegv_device_matches_host_solver(gpu_context):
    from q3_amd._lib import lib
    import ctypes
    n, nvec = 90, 45
    rng = np.random.default_rng(11)
    H = np.asfortranarray(_herm(n, rng))
    S = np.asfortranarray(np.eye(n, dtype=complex))
    wd, _ = _solve(gpu_context, H, S, nvec)
    wh = np.zeros(n)
    Xh = np.zeros((n, nvec), complex, order="F")
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    assert lib().q3x_hegv(n, p(H), p(S), p(wh), p(Xh), nvec) == 0
    assert np.abs(wd - wh[:nvec]).max() <= 1e-12 * np.abs(wh).max()


def test_hegv_device_indefinite_overlap(gpu_context):
    n = 20
    rng = np.random.default_rng(1)
    H = _herm(n, rng)
    S = np.diag(np.r_[np.ones(n - 1), -0.5]).astype(complex)
    with pytest.raises(Q3XError) as e:
        _solve(gpu_context, H, S, 5)
    assert e.value.status == 4


@pytest.mark.parametrize("n,nvec", [(320, 160), (880, 440)])
def test_forced_barrier_giveup_is_redone_on_one_workgroup(gpu_context, n, nvec):
    """The multi-workgroup tridiagonalisation started with its give-up flag set (the host-side
    test hook) runs to its end without waiting; the solve is redone on one workgroup, counted,
    and still meets scipy.  n = 880 is the subspace of a 216-atom, 440-band request."""
    from q3_amd._lib import lib
    rng = np.random.default_rng(n + 1)
    H = _herm(n, rng)
    E = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    S = np.eye(n) + 1e-9 * (E + E.conj().T)
    before = lib().q3x_eig_barrier_redos_total()
    lib().q3x_debug_force_barrier_giveups(1)
    w, X = _solve(gpu_context, H, S, nvec)
    lib().q3x_debug_force_barrier_giveups(0)
    assert lib().q3x_eig_barrier_redos_total() == before + 1
    _check(H, S, w, X, nvec)


def test_hegv_device_stage_profile(gpu_context):
    """q3x_hegv_device times its stages on the context's profiler (q3x_profile_*, ids 6-11:
    prepare, GEMMs, tridiagonalisation, multisection, inverse iteration, back-transformation),
    as the Davidson solve does; the solve is unchanged by the timing."""
    import ctypes
    from q3_amd._lib import lib
    rng = np.random.default_rng(17)
    n, nvec = 320, 160  # the multi-workgroup tridiagonalisation
    H = _herm(n, rng)
    B = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * (0.3 / n)
    S = np.eye(n) + B @ B.conj().T
    tot, cnt = (ctypes.c_double * 12)(), (ctypes.c_long * 12)()
    h = gpu_context.h
    assert lib().q3x_profile_enable(h, 1) == 0
    try:
        assert lib().q3x_profile_read(h, 12, tot, cnt, 1) == 0
        w, X = _solve(gpu_context, H, S, nvec)
        assert lib().q3x_profile_read(h, 12, tot, cnt, 1) == 0
    finally:
        lib().q3x_profile_enable(h, 0)
    _check(H, S, w, X, nvec)
    names = [lib().q3x_profile_kernel_name(i).decode() for i in range(12)]
    assert names[8] == "eig_tridiagonal", names
    for i in range(6, 12):
        assert cnt[i] >= 1 and tot[i] > 0, (names[i], list(cnt), list(tot))
    assert all(cnt[i] == 0 for i in range(6)), list(cnt)  # no H|psi> kernel ran
