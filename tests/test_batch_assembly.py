"""Setup of the batch of Q1 diffusion problems: the variable-coefficient operator
reduces to the Q1 Laplace matrix for c == 1, and its energy equals the
coefficient-weighted Gauss sum of |grad u_h|^2 computed independently here."""
import numpy as np


def _grad_energy(u, level, coef, level_c):
    """sum over cells of sum_q w_q c(x_q) |grad u_h(x_q)|^2 (2x2 Gauss), numpy."""
    N = 2 ** level
    m = N + 1
    U = u.reshape(m, m)                     # [y, x]
    h = 1.0 / N
    g = 1.0 / np.sqrt(3.0)
    Nc = 2 ** level_c
    C = coef.reshape(Nc, Nc)
    e = 0.0
    for t in (-g, g):
        for s in (-g, g):
            u00, u10, u01, u11 = U[:-1, :-1], U[:-1, 1:], U[1:, :-1], U[1:, 1:]
            du_dt = 0.5 * ((1 - s) / 2 * (u10 - u00) + (1 + s) / 2 * (u11 - u01))
            du_ds = 0.5 * ((1 - t) / 2 * (u01 - u00) + (1 + t) / 2 * (u11 - u10))
            xs = (np.arange(N) + 0.5 + 0.5 * t) * h
            ys = (np.arange(N) + 0.5 + 0.5 * s) * h
            cx = np.minimum((xs * Nc).astype(int), Nc - 1)
            cy = np.minimum((ys * Nc).astype(int), Nc - 1)
            cq = C[np.ix_(cy, cx)]
            e += np.sum(cq * (du_dt ** 2 + du_ds ** 2))
    return e


def test_constant_coefficient_is_q1_laplace(feat):
    for level in (1, 3, 5):
        q = feat.q1_problem(2, level)
        v = feat.q1_diffusion(level, level, np.ones(4 ** level))
        assert np.allclose(v, q["val"], rtol=0, atol=8 * np.finfo(float).eps)


def test_energy_identity_and_kernel(feat):
    rng = np.random.default_rng(0)
    for level, level_c in ((3, 3), (3, 5), (4, 6)):
        coef = feat.lognormal_cells(77 + level_c, 4 ** level_c, 1.0)
        q = feat.q1_problem(2, level)
        v = feat.q1_diffusion(level, level_c, coef)
        n = q["n"]
        A = np.zeros((n, n))
        for r in range(n):
            for i in range(int(q["rowptr"][r]), int(q["rowptr"][r + 1])):
                A[r, int(q["col"][i])] = v[i]
        u = rng.standard_normal(n)
        assert np.isclose(u @ A @ u, _grad_energy(u, level, coef, level_c), rtol=1e-12)
        assert np.allclose(A, A.T, atol=1e-13)
        assert np.max(np.abs(A.sum(axis=1))) < 1e-12


def test_batch_levels_and_coefficients(feat):
    lv = feat.batch_levels(28054777172512, 64, 3, 8)
    assert lv.min() >= 3 and lv.max() <= 8 and len(set(lv.tolist())) >= 5
    c = feat.lognormal_cells(1234, 200000, 1.0)
    z = np.log(c)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01
