"""write_out of a device-resident system matrix (the MG-PCG hierarchy's levels,
assembled on the device) through the C-ABI: the fm_csr image holds exactly the
host assembly's arrays, the MatrixMarket file its values to 7 digits."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,lf,lc", [(2, 5, 1), (3, 3, 1)])
def test_write_device_level_matrices(feat, torch_gpu, tmp_path, dim, lf, lc):
    print(f"write level matrices dim={dim} levels {lf}..{lc}", flush=True)
    cfg = feat.mgpcg_cfg(nu_pre=4, nu_post=4, omega_jac=0.5, omega_rich=1.0, coarse="direct",
                         tol_rel=1e-8, max_iter=100)
    s = feat.MgPcg(cfg, q1=(dim, lf, lc))
    try:
        for lvl, level in enumerate(range(lf, lc - 1, -1)):
            q = feat.q1_problem(dim, level)
            s.write_matrix(lvl, "csr", tmp_path / f"A{level}.csr")
            got = feat.csr_read_file("csr", tmp_path / f"A{level}.csr")
            assert (got["rows"], got["cols"]) == (q["n"], q["n"])
            assert np.array_equal(got["rowptr"], q["rowptr"]) and np.array_equal(got["col"], q["col"])
            assert np.array_equal(got["val"].view(np.uint64), q["val"].view(np.uint64))
            s.write_matrix(lvl, "mtx", tmp_path / f"A{level}.mtx", symmetric=True)
            m = feat.csr_read_file("mtx", tmp_path / f"A{level}.mtx")
            assert np.array_equal(m["rowptr"], q["rowptr"]) and np.array_equal(m["col"], q["col"])
            assert np.array_equal(m["val"], np.array([float(f"{v:.6e}") for v in q["val"]]))
    finally:
        s.close()
