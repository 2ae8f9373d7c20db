"""GPU checks of the shifted block preconditioner of the local eigensolver
(through the C ABI hook ttf_eig_block_prec).

The blocks are the diagonal blocks of the dense local matrix built column by
column from the oracle's bfun3 (Kronecker-pinned): block k holds the rows and
columns of right interface index k.  The device inverses of the shifted
blocks must match numpy's inverses of those blocks to 1e-10 relative, for
block orders handled in one piece (<= 128) and through the Schur split (up
to 256, including an uneven split), and the positive-definiteness flag must
be 0 for a shift above the smallest block eigenvalue.
"""
import numpy as np
import pytest

from oracle import kernels as K

pytestmark = pytest.mark.gpu


def _sym_problem(rng, r1, n, r2, ra1, ra2):
    P1 = rng.standard_normal((r1, r1, ra1))
    P2 = rng.standard_normal((r2, ra2, r2))
    P1 = 0.5 * (P1 + np.transpose(P1, (1, 0, 2)))
    P2 = 0.5 * (P2 + np.transpose(P2, (2, 1, 0)))
    A = rng.standard_normal((ra1, n, n, ra2))
    A = 0.5 * (A + np.transpose(A, (0, 2, 1, 3)))
    return K._flat(P1), K._flat(A), K._flat(P2)


def _dense_local(r1, n, r2, ra1, ra2, phi1, A, phi2):
    N = r1 * n * r2
    B = np.empty((N, N))
    for j in range(N):
        e = np.zeros(N)
        e[j] = 1.0
        B[:, j] = K.dbfun3(r1, n, r2, r1, n, r2, ra1, ra2, phi1, A, phi2, e)
    return B


@pytest.mark.parametrize("r1,n,r2", [(12, 2, 5), (64, 2, 3), (100, 2, 2), (85, 3, 2), (128, 2, 2)])
def test_shifted_block_inverse(ttlib, r1, n, r2):
    import ttfort
    rng = np.random.default_rng(r1 * 7 + n)
    ra1, ra2 = 3, 2
    phi1, A, phi2 = _sym_problem(rng, r1, n, r2, ra1, ra2)
    B = _dense_local(r1, n, r2, ra1, ra2, phi1, A, phi2)
    assert np.allclose(B, B.T, atol=1e-12 * np.abs(B).max())
    nb = r1 * n
    lam = np.linalg.eigvalsh(B)
    spread = lam[-1] - lam[0]
    sigma = lam[0] - 0.05 * spread     # below the spectrum: every block positive definite
    inv, pd = ttfort.eig_block_prec(r1, n, r2, ra1, ra2, phi1, A, phi2, sigma)
    assert pd
    J = np.reshape(inv, (nb, nb, r2), order="F")
    for k in range(r2):
        Mk = B[k * nb:(k + 1) * nb, k * nb:(k + 1) * nb] - sigma * np.eye(nb)
        ref = np.linalg.inv(Mk)
        err = np.abs(J[:, :, k] - ref).max() / np.abs(ref).max()
        assert err < 1e-10, (k, err)


@pytest.mark.parametrize("r1,n,r2", [(12, 2, 5), (100, 2, 2)])
def test_shift_inside_a_block_spectrum_is_flagged(ttlib, r1, n, r2):
    import ttfort
    rng = np.random.default_rng(r1 + 3 * r2)
    ra1, ra2 = 3, 2
    phi1, A, phi2 = _sym_problem(rng, r1, n, r2, ra1, ra2)
    B = _dense_local(r1, n, r2, ra1, ra2, phi1, A, phi2)
    nb = r1 * n
    bmin = [np.linalg.eigvalsh(B[k * nb:(k + 1) * nb, k * nb:(k + 1) * nb])[0] for k in range(r2)]
    sigma = min(bmin) + 1e-3 * (max(bmin) - min(bmin) + 1.0)   # one block indefinite
    _, pd = ttfort.eig_block_prec(r1, n, r2, ra1, ra2, phi1, A, phi2, sigma)
    assert not pd
    sigma = min(bmin) - 1e-3 * (max(bmin) - min(bmin) + 1.0)   # all blocks positive definite
    _, pd = ttfort.eig_block_prec(r1, n, r2, ra1, ra2, phi1, A, phi2, sigma)
    assert pd
