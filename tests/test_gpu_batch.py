"""Batched reconstructions (et_batch_osem): uneven problems (volume shapes,
view counts, PSF sizes, with and without attenuation, ragged subsets) run as
grouped launches captured in one hipGraph.  Each problem must match the same
plan run alone through et_plan_osem bit for bit (same depth chunking, same
kernels), graph replays must match eager issue, and one problem is checked
against the oracle's sequential MLEM updates."""
import numpy as np
import pytest

from conftest import rel_l2, report_parity
from niftyrec_amd import phantoms
from oracle import orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

# (nz, ny, nx), views, PSF size (0: none), attenuation
PROBLEMS = [((32, 24, 32), 16, 5, True), ((48, 20, 48), 22, 7, True), ((32, 28, 32), 13, 3, False),
            ((40, 16, 40), 20, 5, True), ((24, 30, 24), 9, 0, True)]
SUBSETS = 4


def _problem(i, shape, V, K, has_mu):
    rng = np.random.default_rng(100 + i)
    act, mu = phantoms.phantom(shape, 2, 100 + i, lesion_axes=(1.5, 3.0))
    act = act + 0.1
    mu = (mu / 0.15 * 0.08).astype(np.float32) if has_mu else None
    cams = np.sort(rng.uniform(0, 2 * np.pi, V)).astype(np.float32)
    psf = phantoms.gaussian_psf(shape[2], K, 0.5, 1.5) if K else None
    y = orc.project(act, mu, cams, psf)
    y = rng.poisson(y * 50).astype(np.float32) / 50.0
    return act, mu, cams, psf, y


def _t(a):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _plans(cases):
    from niftyrec_amd.api import Plan
    plans = []
    for (act, mu, cams, psf, y) in cases:
        p = Plan(act.shape, cams, psf)
        p.set_tuning(4, 0)
        plans.append(p)
    return plans


def test_batch_matches_single_plans_and_replays():
    from niftyrec_amd.api import Batch
    cases = [_problem(i, *pr) for i, pr in enumerate(PROBLEMS)]
    solo = _plans(cases)
    grouped = _plans(cases)
    batch = Batch(grouped)
    Y = [_t(c[4]) for c in cases]
    M = [_t(c[1]) for c in cases]
    Xs = [torch.ones(c[0].shape, dtype=torch.float32, device="cuda") for c in cases]
    Xb = [x.clone() for x in Xs]
    Xe = [x.clone() for x in Xs]

    for p, x, y, m in zip(solo, Xs, Y, M):
        p.osem(x, y, m, subsets=SUBSETS, iterations=2)
    batch.osem(Xb, Y, M, subsets=SUBSETS, iterations=2, use_graph=True)
    torch.cuda.synchronize()
    for i, (a, b) in enumerate(zip(Xs, Xb)):
        same = torch.equal(a, b)
        print("problem %d %s: graph batch vs single plan identical=%s relL2 %.2e"
              % (i, tuple(a.shape), same, rel_l2(b.cpu().numpy(), a.cpu().numpy())))
        assert same

    # eager work on the device between replays, then two more replays
    junk = torch.randn(1 << 20, device="cuda").cumsum(0)
    for rep in range(2):
        for p, x, y, m in zip(solo, Xs, Y, M):
            p.osem(x, y, m, subsets=SUBSETS, iterations=1)
        batch.osem(Xb, Y, M, subsets=SUBSETS, iterations=1, use_graph=True)
        junk = junk * 0.5 + 1.0
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(zip(Xs, Xb)):
            assert torch.equal(a, b), "replay %d problem %d" % (rep, i)

    # eager issue of the same launch list
    batch.osem(Xe, Y, M, subsets=SUBSETS, iterations=4, use_graph=False)
    torch.cuda.synchronize()
    for i, (a, b) in enumerate(zip(Xs, Xe)):
        assert torch.equal(a, b), "eager batch problem %d" % i
    assert torch.isfinite(junk).all()


def test_batch_against_oracle():
    from niftyrec_amd.api import Batch
    cases = [_problem(i, *pr) for i, pr in enumerate(PROBLEMS[:3])]
    plans = _plans(cases)
    batch = Batch(plans)
    X = [torch.ones(c[0].shape, dtype=torch.float32, device="cuda") for c in cases]
    batch.osem(X, [_t(c[4]) for c in cases], [_t(c[1]) for c in cases], subsets=SUBSETS, iterations=1)
    torch.cuda.synchronize()
    for i, (act, mu, cams, psf, y) in enumerate(cases):
        x, x64 = np.ones_like(act), np.ones(act.shape)
        for s in range(SUBSETS):
            x = orc.mlem_step(x, y[s::SUBSETS], mu, cams[s::SUBSETS], psf)
            x64 = orc.mlem_step(x64, y[s::SUBSETS], mu, cams[s::SUBSETS], psf, dtype=np.float64)
        report_parity("problem %d: batch OSEM (%d subsets)" % (i, SUBSETS), X[i].cpu().numpy(), x, x64)
