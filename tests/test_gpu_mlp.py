"""The learned smoother on the GPU: the BF16-MFMA network (forward, gradient,
Adam), its input features and the smoothing step against the CPU oracle; the
Jacobi fallback; training on residual/error pairs of the FP64 multigrid and
what the trained smoother gains over Jacobi."""
import ctypes as C

import numpy as np
import pytest

from tests.test_gpu_lafem import dev, host
from tests.test_mlp_oracle import NP, NPP, OW1, OB1, random_problem, from_bf16, bf16

pytestmark = pytest.mark.gpu


def d16(torch, a):
    return dev(torch, np.asarray(a, dtype=np.uint16).view(np.int16))


def params_dev(feat, handle, torch, p):
    k, _ = feat.libs()
    pd = dev(torch, np.concatenate([p, np.zeros(NPP - NP, np.float32)]).astype(np.float32))
    wbf = torch.zeros(2 * 4096, dtype=torch.int16, device="cuda")
    feat.check(k.feat_mlp_to_bf16(handle.h, feat.ptr(pd), feat.ptr(wbf)), "to_bf16")
    return pd, wbf


def test_mfma_forward_vs_oracle(feat, handle, torch_gpu, oracle):
    torch = torch_gpu
    k, _ = feat.libs()
    for n, seed in ((16, 10), (1000, 11), (5003, 12)):
        p, Xb, t = random_problem(n, seed)
        pd, wbf = params_dev(feat, handle, torch, p)
        Xd = d16(torch, Xb)
        y = torch.zeros(n, dtype=torch.float32, device="cuda")
        feat.check(k.feat_mlp_infer(handle.h, feat.ptr(pd), feat.ptr(wbf), feat.ptr(Xd), n, feat.ptr(y)), "infer")
        yo = oracle.mlp_forward(p, Xb, round_bf16=True)
        yg = host(y)
        assert np.linalg.norm(yg - yo) <= 5e-3 * np.linalg.norm(yo), n
        assert np.max(np.abs(yg - yo)) <= 2e-2 * np.max(np.abs(yo))


def test_mfma_gradient_vs_oracle(feat, handle, torch_gpu, oracle):
    torch = torch_gpu
    k, _ = feat.libs()
    for n, seed in ((64, 20), (3000, 21)):
        p, Xb, t = random_problem(n, seed)
        pd, wbf = params_dev(feat, handle, torch, p)
        Xd, td = d16(torch, Xb), dev(torch, t)
        cap = 4 * 2048
        part = torch.zeros(cap * NPP, dtype=torch.float32, device="cuda")
        g = torch.zeros(NPP, dtype=torch.float32, device="cuda")
        feat.check(k.feat_mlp_grad(handle.h, feat.ptr(pd), feat.ptr(wbf), feat.ptr(Xd), feat.ptr(td), n,
                                   feat.ptr(part), cap, feat.ptr(g)), "grad")
        go = oracle.mlp_grad(p, Xb, t, round_bf16=True)
        gg = host(g)
        assert np.linalg.norm(gg[:NP] - go[:NP]) <= 1e-2 * np.linalg.norm(go[:NP]), n
        for a, b in ((0, 4096), (4096, 4160), (4160, 8256), (8256, 8320), (8320, 8385)):
            assert np.linalg.norm(gg[a:b] - go[a:b]) <= 3e-2 * np.linalg.norm(go[a:b]) + 1e-7, (n, a)
        assert np.isclose(gg[NPP - 1], go[NPP - 1], rtol=1e-3)
        # padded inputs get no gradient
        assert not np.any(gg[OW1:OB1].reshape(64, 64)[:, 54:])
        # deterministic: a second evaluation gives the same bits
        g2 = torch.zeros_like(g)
        feat.check(k.feat_mlp_grad(handle.h, feat.ptr(pd), feat.ptr(wbf), feat.ptr(Xd), feat.ptr(td), n,
                                   feat.ptr(part), cap, feat.ptr(g2)), "grad")
        assert np.array_equal(host(g2), gg)


def test_adam_step(feat, handle, torch_gpu):
    torch = torch_gpu
    k, _ = feat.libs()
    p, _, _ = random_problem(1, 30)
    pd, wbf = params_dev(feat, handle, torch, p)
    rng = np.random.default_rng(31)
    g = rng.standard_normal(NPP).astype(np.float32)
    g[OW1:OB1].reshape(64, 64)[:, 54:] = 0.0
    m = rng.standard_normal(NPP).astype(np.float32) * 0.1
    v = np.abs(rng.standard_normal(NPP)).astype(np.float32) * 0.1
    gd, md, vd = dev(torch, g), dev(torch, m), dev(torch, v)
    lr, b1, b2, eps, step = 1e-3, 0.9, 0.999, 1e-8, 3
    feat.check(k.feat_mlp_adam(handle.h, feat.ptr(pd), feat.ptr(wbf), feat.ptr(gd), feat.ptr(md), feat.ptr(vd), step,
                               lr, b1, b2, eps), "adam")
    mm = b1 * m[:NP] + (1 - b1) * g[:NP]
    vv = b2 * v[:NP] + (1 - b2) * g[:NP] ** 2
    pe = p - lr * (mm / (1 - b1 ** step)) / (np.sqrt(vv / (1 - b2 ** step)) + eps)
    pe[OW1:OB1].reshape(64, 64)[:, 54:] = 0.0
    assert np.allclose(host(pd)[:NP], pe, rtol=1e-5, atol=1e-7)
    assert np.array_equal(host(wbf).view(np.uint16)[:4096], bf16(host(pd)[OW1:OB1]))


def test_features_and_smoothing_step_vs_oracle(feat, handle, torch_gpu, oracle):
    torch = torch_gpu
    k, _ = feat.libs()
    m = 9
    n = m ** 3
    rng = np.random.default_rng(40)
    r, e = rng.standard_normal(n), rng.standard_normal(n)
    dinv = rng.uniform(0.5, 2.0, n)
    cfeat = rng.standard_normal(n).astype(np.float32)
    nodes = rng.integers(0, n, 500).astype(np.uint64)
    Xo, to = oracle.mlp_features(m, r, e, dinv, cfeat, nodes)
    Xd = torch.zeros((nodes.size, 64), dtype=torch.int16, device="cuda")
    td = torch.zeros(nodes.size, dtype=torch.float32, device="cuda")
    snap = np.zeros(nodes.size, np.uint32)
    rd, ed, dd, cd = dev(torch, r), dev(torch, e), dev(torch, dinv), dev(torch, cfeat)
    nd, sd = dev(torch, nodes.view(np.int64)), dev(torch, snap.view(np.int32))
    feat.check(k.feat_mlp_features(handle.h, m, feat.ptr(rd), feat.ptr(ed), feat.ptr(dd), feat.ptr(cd), feat.ptr(nd),
                                   feat.ptr(sd), nodes.size, feat.ptr(Xd), feat.ptr(td)), "features")
    assert np.array_equal(host(Xd).view(np.uint16), Xo)
    assert np.array_equal(host(td), to)
    p, _, _ = random_problem(1, 41)
    pd, wbf = params_dev(feat, handle, torch, p)
    x0 = rng.standard_normal(n)
    xd = dev(torch, x0)
    feat.check(k.feat_mlp_smooth_f64(handle.h, m, feat.ptr(pd), feat.ptr(wbf), feat.ptr(rd), feat.ptr(dd), feat.ptr(cd),
                                     0.8, 0.5, feat.ptr(xd)), "smooth")
    xo = oracle.mlp_smooth(m, p, r, dinv, cfeat, 0.8, 0.5, x0)
    xg = host(xd)
    idx = np.arange(n)
    bnd = (idx % m == 0) | (idx % m == m - 1) | ((idx // m) % m == 0) | ((idx // m) % m == m - 1) | \
        (idx // (m * m) == 0) | (idx // (m * m) == m - 1)
    assert np.array_equal(xg[bnd], x0[bnd])
    assert np.linalg.norm((xg - x0) - (xo - x0)) <= 1e-2 * np.linalg.norm(xo - x0)


@pytest.fixture(scope="module")
def lmg(feat, torch_gpu):
    s = feat.LearnedMG(level_fine=5, level_coarse=1, batch=4096, lr=2e-3)
    yield s
    s.close()


def test_training_and_gain_over_jacobi(lmg):
    rho = lmg.generate()
    assert 0 < rho < 1
    jac = lmg.solve(learned=False)
    assert jac["status"] == "success"
    e0 = lmg.energy()
    loss = lmg.train(300, seed=1)
    assert np.all(np.isfinite(loss))
    # the relative energy one learned step leaves on the training fields falls below
    # the untrained (damped Jacobi) value
    e1 = lmg.energy()
    print(f"relative one-step energy: Jacobi {e0:.4f}, learned {e1:.4f}")
    assert e1 < 0.95 * e0, (e0, e1)
    lrn = lmg.solve(learned=True)
    raw = lmg.solve(learned=True, allow_fallback=False)
    print(f"jacobi-MG {jac['iters']} iterations, learned-MG {lrn['iters']} iterations "
          f"(fallback {lrn['fell_back']} at {lrn['fallback_iter']}), without fallback {raw['status']} "
          f"{raw['iters']}, rho_jacobi {rho:.3f}")
    # the plain FP64 defect decides: the trained smoother reaches 1e-8 in fewer
    # iterations than the Jacobi multigrid (training is deterministic: seeded data,
    # index-ordered gradient reductions)
    assert lrn["status"] == "success"
    assert lrn["def_final"] <= 1e-8 * lrn["def_init"]
    assert lrn["iters"] < jac["iters"], (lrn["iters"], jac["iters"])


def test_untrained_is_jacobi_and_fallback(lmg):
    saved = lmg.params()
    # zero output layer: the learned smoother steps along the Jacobi direction (with the
    # energy-optimal length), no fallback, about Jacobi's iteration count
    zero = saved.copy()
    zero[8320:8385] = 0.0
    lmg.params(zero)
    jac = lmg.solve(learned=False)
    r0 = lmg.solve(learned=True)
    assert r0["status"] == "success" and not r0["fell_back"]
    assert abs(r0["iters"] - jac["iters"]) <= 3, (r0["iters"], jac["iters"])
    # a network that outputs garbage (a huge constant) makes the smoothing useless: the
    # solver falls back to Jacobi and still converges, judged by the FP64 defect
    bad = saved.copy()
    bad[8384] = 1e3
    lmg.params(bad)
    r = lmg.solve(learned=True)
    lmg.params(saved)
    assert r["fell_back"] and r["fallback_iter"] >= 1
    assert r["status"] == "success"


def test_energy_optimal_step(feat, handle, torch_gpu):
    # x += alpha c, d -= alpha q with alpha = a / b; alpha = 0 when b <= 0
    torch = torch_gpu
    k, _ = feat.libs()
    P = feat.ptr
    rng = np.random.default_rng(5)
    n = 1000
    x, d, c, q = (rng.standard_normal(n) for _ in range(4))
    for a, b in [(0.7, 2.0), (1.0, 0.0), (1.0, -1.0)]:
        xd, dd, cd, qd = (torch.from_numpy(v.copy()).cuda() for v in (x, d, c, q))
        ab = torch.tensor([a, b], dtype=torch.float64, device="cuda")
        feat.check(k.feat_mlp_step_f64(handle.h, n, P(xd), P(dd), P(cd), P(qd), P(ab), P(ab[1:])), "step")
        torch.cuda.synchronize()
        alpha = a / b if b > 0 else 0.0
        # one fma per element: within one rounding of the separately rounded numpy form
        assert np.allclose(xd.cpu().numpy(), x + alpha * c, rtol=4e-16, atol=1e-15)
        assert np.allclose(dd.cpu().numpy(), d - alpha * q, rtol=4e-16, atol=1e-15)
        if alpha == 0.0:
            assert np.array_equal(xd.cpu().numpy(), x) and np.array_equal(dd.cpu().numpy(), d)
