"""Exact adjoint, exact attenuation gradient and the torch autograd binding on
the GPU, checked against the double-precision oracle (whose adjoint and
gradient are pinned by dense brute force and finite differences in
tests/test_oracle_pins.py), by the dot-product identity and by FP64 central
finite differences of the Poisson negative log-likelihood."""
import numpy as np
import pytest

from conftest import rel_l2, report_parity
from niftyrec_amd import phantoms
from oracle import orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _case(shape=(32, 24, 32), V=12, K=5, seed=0):
    rng = np.random.default_rng(seed)
    act, mu = phantoms.phantom(shape, 3, seed, lesion_axes=(1.5, 4.0))
    act = act + 0.1 * rng.random(shape, dtype=np.float32)
    mu = (mu / 0.15 * 0.1).astype(np.float32)
    cams = rng.uniform(0, 2 * np.pi, V).astype(np.float32)
    psf = phantoms.gaussian_psf(shape[2], K, 0.5, 1.6) if K else None
    if psf is not None:
        psf[:, :, 0] *= 1.2  # asymmetric: correlation differs from convolution
    return act, mu, cams, psf


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _plan(shape, cams, psf, bg=0.0):
    from niftyrec_amd.api import Plan
    return Plan(shape, cams, psf, background=bg)


@pytest.mark.parametrize("K", [0, 5, 11])
def test_adjoint_matches_fp64_and_dot_identity(K):
    from niftyrec_amd.api import ADJOINT
    act, mu, cams, psf = _case(K=K)
    plan = _plan(act.shape, cams, psf)
    y = np.random.default_rng(1).random((len(cams),) + act.shape[1:], dtype=np.float32)
    Aty = plan.backproject(_t(y), _t(mu), mode=ADJOINT, truncate=False).cpu().numpy()
    ref = orc.backproject(y, mu, cams, psf, 0.0, 0.0, 0, mode=1, dtype=np.float64)
    e = rel_l2(Aty, ref)
    Ax = plan.project(_t(act), _t(mu), truncate=False).cpu().numpy()
    lhs = np.vdot(Ax.astype(np.float64), y.astype(np.float64))
    rhs = np.vdot(act.astype(np.float64), Aty.astype(np.float64))
    mism = abs(lhs - rhs) / abs(lhs)
    print("adjoint K=%d: relL2 vs FP64 adjoint %.2e, dot-product mismatch %.2e" % (K, e, mism))
    assert e <= 1e-5
    assert mism <= 1e-6


def test_adjoint_general_rotations():
    from niftyrec_amd.api import ADJOINT
    act, mu, _, psf = _case(V=5, seed=2)
    c3 = np.random.default_rng(2).uniform(-0.6, 0.6, (3, 5)).astype(np.float32)
    plan = _plan(act.shape, c3, psf)
    y = np.random.default_rng(3).random((5,) + act.shape[1:], dtype=np.float32)
    Aty = plan.backproject(_t(y), _t(mu), mode=ADJOINT, truncate=False).cpu().numpy()
    ref = orc.backproject(y, mu, c3, psf, 0.0, 0.0, 0, mode=1, dtype=np.float64)
    Ax = plan.project(_t(act), _t(mu), truncate=False).cpu().numpy()
    mism = abs(np.vdot(Ax.astype(np.float64), y) - np.vdot(act.astype(np.float64), Aty)) / abs(np.vdot(Ax, y))
    print("general adjoint: relL2 %.2e dot mismatch %.2e" % (rel_l2(Aty, ref), mism))
    assert rel_l2(Aty, ref) <= 1e-5
    assert mism <= 1e-6


@pytest.mark.parametrize("K", [0, 5])
def test_exact_attenuation_gradient_matches_fp64(K):
    act, mu, cams, psf = _case(K=K, seed=4)
    plan = _plan(act.shape, cams, psf)
    s = np.random.default_rng(5).random((len(cams),) + act.shape[1:], dtype=np.float32)
    g = plan.gradient_attenuation(_t(s), _t(act), _t(mu), exact=True, truncate=False).cpu().numpy()
    ref = orc.gradient_mu(s, act, mu, cams, psf, dtype=np.float64)
    e = rel_l2(g, ref)
    print("exact mu gradient K=%d: relL2 vs FP64 %.2e" % (K, e))
    assert e <= 1e-5


def test_reference_gradient_through_plan():
    act, mu, cams, psf = _case(seed=6)
    plan = _plan(act.shape, cams, psf)
    s = np.random.default_rng(7).random((len(cams),) + act.shape[1:], dtype=np.float32)
    g = plan.gradient_attenuation(_t(s), _t(act), _t(mu), exact=False, truncate=True).cpu().numpy()
    ref = orc.gradient_ref(s, act, mu, cams, psf)
    assert rel_l2(g, ref) <= 1e-5


def test_autograd_poisson_nll_finite_differences():
    """d/d activity and d/d mu of the Poisson NLL through torch autograd against
    FP64 central differences (step 1e-3) on 64 random voxels each."""
    from niftyrec_amd.torch_ops import poisson_nll, project
    act, mu, cams, psf = _case(shape=(16, 12, 16), V=8, K=5, seed=8)
    plan = _plan(act.shape, cams, psf)
    y = orc.project(act * 1.3, mu * 0.8, cams, psf, dtype=np.float64)
    y = np.random.default_rng(9).poisson(y * 20).astype(np.float32) / 20.0
    A = _t(act).requires_grad_(True)
    M = _t(mu).requires_grad_(True)
    loss = poisson_nll(project(A, M, plan), _t(y))
    loss.backward()
    ga, gm = A.grad.cpu().numpy(), M.grad.cpu().numpy()

    def f(a, m):
        p = orc.project(a, m, cams, psf, 0.0, 0.0, 0, dtype=np.float64)
        return float((p - y * np.log(np.maximum(p, 1e-8))).sum())

    rng = np.random.default_rng(10)
    h = 1e-3
    for name, grad, which in (("activity", ga, 0), ("mu", gm, 1)):
        base = [act.astype(np.float64), mu.astype(np.float64)]
        idx = rng.choice(act.size, 64, replace=False)
        fd = []
        for i in idx:
            p, m = [b.copy() for b in base], [b.copy() for b in base]
            p[which].ravel()[i] += h
            m[which].ravel()[i] -= h
            fd.append((f(*p) - f(*m)) / (2 * h))
        fd = np.array(fd)
        got = grad.ravel()[idx].astype(np.float64)
        err = np.abs(got - fd).max() / np.abs(fd).max()
        print("autograd d/d%s: max |grad - FD| / max|FD| = %.2e" % (name, err))
        assert err <= 2e-4


def test_plan_osem_matches_sequential_mlem_on_subsets():
    """et_plan_osem with S ordered subsets equals S reference MLEM updates on
    views {s, s+S, ...} applied in turn (oracle, single precision)."""
    act, mu, cams, psf = _case(seed=11, V=16)
    plan = _plan(act.shape, cams, psf)
    y = orc.project(act, mu, cams, psf)
    X = torch.ones(act.shape, dtype=torch.float32, device="cuda")
    plan.osem(X, _t(y), _t(mu), subsets=4, iterations=2)
    x, x64 = np.ones_like(act), np.ones(act.shape)
    for _ in range(2):
        for s in range(4):
            x = orc.mlem_step(x, y[s::4], mu, cams[s::4], psf)
            x64 = orc.mlem_step(x64, y[s::4], mu, cams[s::4], psf, dtype=np.float64)
    report_parity("plan osem (4 subsets x 2 iterations)", X.cpu().numpy(), x, x64)


def test_plan_osem_fixed_depth_chunks():
    """et_plan_set_tuning fixes the sweeps' depth chunking and et_plan_get_tuning
    reports the chunking of the cached OSEM data ((0, 0) before the first OSEM
    call, the timed choice after it); every chunking gives the oracle's result."""
    act, mu, cams, psf = _case(seed=14, V=8)
    plan = _plan(act.shape, cams, psf)
    assert plan.tuning() == (0, 0) and plan.projector() == 0
    y = orc.project(act, mu, cams, psf)
    x, x64 = np.ones_like(act), np.ones(act.shape)
    for s in range(2):
        x = orc.mlem_step(x, y[s::2], mu, cams[s::2], psf)
        x64 = orc.mlem_step(x64, y[s::2], mu, cams[s::2], psf, dtype=np.float64)
    for D in (0, 1, 3):
        plan.set_tuning(D, 0)
        assert plan.tuning() == (0, 0)      # the cache is dropped until the next OSEM call
        X = torch.ones(act.shape, dtype=torch.float32, device="cuda")
        plan.osem(X, _t(y), _t(mu), subsets=2, iterations=1)
        dp, db = plan.tuning()
        if D:
            assert (dp, db) == (D, D)
        else:
            assert dp >= 1 and db >= 1
        report_parity("plan osem, depth chunks (%d, %d)" % (dp, db), X.cpu().numpy(), x, x64)


def test_plan_profile_sampling():
    """et_plan_profile_select + et_plan_profile_sample: only the selected
    category is timed, and of its launches every `every`-th one, starting with
    the first."""
    act, mu, cams, psf = _case(seed=15, V=8)
    plan = _plan(act.shape, cams, psf)
    y = _t(orc.project(act, mu, cams, psf))
    X = torch.ones(act.shape, dtype=torch.float32, device="cuda")
    M = _t(mu)
    plan.osem(X, y, M, subsets=4, iterations=1)     # caches built outside the count
    for every, want in ((1, 12), (3, 4), (5, 3)):
        plan.profile(True, categories=["project"], every=every)
        plan.osem(X, y, M, subsets=4, iterations=3)  # 12 projection launches
        r = plan.profile_read()
        plan.profile(False)
        assert r["project"][0] == want and r["project"][1] > 0
        assert r["bp_rotated_frame"][0] == 0 and r["bp_pull"][0] == 0


@pytest.mark.parametrize("shape, V, K", [((96, 20, 96), 12, 7), ((160, 9, 160), 8, 3), ((65, 20, 65), 8, 5),
                                          ((130, 24, 130), 8, 11), ((190, 12, 190), 6, 15)])
def test_plan_osem_wide_volumes(shape, V, K):
    """OSEM through the plan at widths that leave the last wave of a block
    partly idle (96: 2 waves, 160: 3 waves) and ragged axial sizes: cached
    attenuation factors in the padded layout, coalesced stores, tuned chunks.
    Wider than 64 columns with a PSF, the projector works in 64-column tiles
    with R halo columns on either side: 65 leaves one column in the last tile,
    130 two in the third, 190 sixty-two in the third (R = 7)."""
    act, mu, cams, psf = _case(shape=shape, V=V, K=K, seed=12)
    plan = _plan(act.shape, cams, psf)
    y = orc.project(act, mu, cams, psf)
    X = torch.ones(act.shape, dtype=torch.float32, device="cuda")
    plan.osem(X, _t(y), _t(mu), subsets=2, iterations=1)
    tiles = plan.projector()  # timed choice between the tiled and the full-width projector
    print("plan osem %s: projector tiles %d (1: full width)" % (shape, tiles))
    assert tiles == 1 or (K > 1 and shape[2] > 64 and tiles == -(-shape[2] // 64))
    x, x64 = np.ones_like(act), np.ones(act.shape)
    for s in range(2):
        x = orc.mlem_step(x, y[s::2], mu, cams[s::2], psf)
        x64 = orc.mlem_step(x64, y[s::2], mu, cams[s::2], psf, dtype=np.float64)
    report_parity("plan osem %s, V=%d, K=%d" % (shape, V, K), X.cpu().numpy(), x, x64)


@pytest.mark.parametrize("shape, V, K", [((24, 140, 24), 8, 5), ((200, 10, 200), 4, 7), ((40, 140, 40), 6, 17)])
def test_plan_osem_sweep_kernels(shape, V, K):
    """OSEM through a plan the row-lane kernels do not take (more than 128
    axial rows, or more than 192 columns): the sweep kernels with cached
    attenuation factors, timed depth chunks."""
    act, mu, cams, psf = _case(shape=shape, V=V, K=K, seed=16)
    plan = _plan(act.shape, cams, psf)
    y = orc.project(act, mu, cams, psf)
    X = torch.ones(act.shape, dtype=torch.float32, device="cuda")
    plan.osem(X, _t(y), _t(mu), subsets=2, iterations=1)
    assert plan.projector() == 0  # the sweep kernels
    x, x64 = np.ones_like(act), np.ones(act.shape)
    for s in range(2):
        x = orc.mlem_step(x, y[s::2], mu, cams[s::2], psf)
        x64 = orc.mlem_step(x64, y[s::2], mu, cams[s::2], psf, dtype=np.float64)
    report_parity("plan osem on the sweep kernels %s, V=%d, K=%d" % (shape, V, K), X.cpu().numpy(), x, x64)


def test_plan_osem_general_path():
    """A plan with a non-zero background takes the materialising path: OSEM
    with one subset equals the oracle's MLEM steps; more subsets are refused
    with a parameter status (the general path has no subset caches)."""
    from niftyrec_amd import _lib
    act, mu, cams, psf = _case(seed=13, V=8)
    plan = _plan(act.shape, cams, psf, bg=0.2)
    y = orc.project(act, mu, cams, psf, bg=0.2)
    X = torch.ones(act.shape, dtype=torch.float32, device="cuda")
    plan.osem(X, _t(y), _t(mu), subsets=1, iterations=2)
    x, x64 = np.ones_like(act), np.ones(act.shape)
    for _ in range(2):
        x = orc.mlem_step(x, y, mu, cams, psf, bg=0.2)
        x64 = orc.mlem_step(x64, y, mu, cams, psf, bg=0.2, dtype=np.float64)
    report_parity("plan osem, general path", X.cpu().numpy(), x, x64)
    with pytest.raises(_lib.NiftyRecError):
        plan.osem(X, _t(y), _t(mu), subsets=2, iterations=1)
