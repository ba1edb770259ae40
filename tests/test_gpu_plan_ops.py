"""Single operations of a created plan with attenuation, which run on the
row-lane kernels: projection, reference and exact-adjoint backprojection and
the exact attenuation gradient, against the oracle (single and double
precision) at widths that take the full-width and the tiled projector; the
cached attenuation factors follow the map's content; and OSEM (a plan's or a
batch's) after operations that reallocate a plan's work buffers (the cached
launch table or graph is then rebuilt) matches runs without them."""
import numpy as np
import pytest

from conftest import report_parity
from niftyrec_amd import phantoms
from oracle import orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _case(shape, V, K, seed):
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


@pytest.mark.parametrize("shape, V, K", [((40, 20, 40), 10, 5), ((32, 24, 32), 8, 0), ((96, 20, 96), 8, 7),
                                          ((130, 24, 130), 6, 11), ((17, 1, 17), 3, 3), ((65, 7, 65), 1, 5),
                                          ((23, 45, 23), 5, 7)])
def test_plan_operations_on_row_lane_kernels(shape, V, K):
    from niftyrec_amd.api import ADJOINT, Plan
    act, mu, cams, psf = _case(shape, V, K, seed=90)
    plan = Plan(act.shape, cams, psf)
    A, M = _t(act), _t(mu)
    tag = "%dx%dx%d, %d views, K=%d" % (shape[2], shape[1], shape[0], V, K)
    p = plan.project(A, M).cpu().numpy()
    report_parity("plan project (%s)" % tag, p, orc.project(act, mu, cams, psf),
                  orc.project(act, mu, cams, psf, dtype=np.float64))
    y = np.random.default_rng(91).random((V,) + act.shape[1:], dtype=np.float32)
    b = plan.backproject(_t(y), M).cpu().numpy()
    report_parity("plan backproject, reference (%s)" % tag, b, orc.backproject(y, mu, cams, psf),
                  orc.backproject(y, mu, cams, psf, dtype=np.float64))
    ba = plan.backproject(_t(y), M, mode=ADJOINT, truncate=False).cpu().numpy()
    report_parity("plan backproject, exact adjoint (%s)" % tag, ba,
                  orc.backproject(y, mu, cams, psf, 0.0, 0.0, 0, mode=1),
                  orc.backproject(y, mu, cams, psf, 0.0, 0.0, 0, mode=1, dtype=np.float64))
    g = plan.gradient_attenuation(_t(y), A, M, exact=True, truncate=False).cpu().numpy()
    report_parity("plan exact attenuation gradient (%s)" % tag, g,
                  orc.gradient_mu(y, act, mu, cams, psf, dtype=np.float32),
                  orc.gradient_mu(y, act, mu, cams, psf, dtype=np.float64))


def test_plan_operations_in_view_batches():
    """Backprojections and the gradient in batches of 3 views (set_tuning's
    view batch): the pulls accumulate over the batches and truncate after the
    last, and the launches of every batch go in one table.  The gradient takes
    non-negative weights, as the other gradient tests: with signed ones its
    tails cancel, and there the FP32 results (this library's and the oracle's
    alike) are up to 3e-3 from FP64 element-wise."""
    from niftyrec_amd.api import ADJOINT, Plan
    act, mu, cams, psf = _case((40, 20, 40), 10, 5, seed=96)
    plan = Plan(act.shape, cams, psf)
    plan.set_tuning(0, 3)
    A, M = _t(act), _t(mu)
    y = np.random.default_rng(97).random((10,) + act.shape[1:], dtype=np.float32) - np.float32(0.3)
    b = plan.backproject(_t(y), M).cpu().numpy()  # negative values: truncation at the end matters
    report_parity("plan backproject, reference, 3-view batches", b, orc.backproject(y, mu, cams, psf),
                  orc.backproject(y, mu, cams, psf, dtype=np.float64))
    ba = plan.backproject(_t(y), M, mode=ADJOINT, truncate=False).cpu().numpy()
    report_parity("plan backproject, exact adjoint, 3-view batches", ba,
                  orc.backproject(y, mu, cams, psf, 0.0, 0.0, 0, mode=1),
                  orc.backproject(y, mu, cams, psf, 0.0, 0.0, 0, mode=1, dtype=np.float64))
    s = np.abs(y)
    g = plan.gradient_attenuation(_t(s), A, M, exact=True, truncate=False).cpu().numpy()
    report_parity("plan exact attenuation gradient, 3-view batches", g,
                  orc.gradient_mu(s, act, mu, cams, psf, dtype=np.float32),
                  orc.gradient_mu(s, act, mu, cams, psf, dtype=np.float64))


def test_plan_operation_factors_follow_the_map():
    """The factors cached for the single operations are keyed on the map's
    content: a map rewritten in place (same address) is caught."""
    from niftyrec_amd.api import Plan
    act, mu, cams, psf = _case((48, 20, 48), 8, 5, seed=92)
    plan = Plan(act.shape, cams, psf)
    A, M = _t(act), _t(mu)
    plan.project(A, M)
    M.mul_(0.5)
    mu2 = mu * np.float32(0.5)
    p = plan.project(A, M).cpu().numpy()
    report_parity("plan project after the map was halved in place", p, orc.project(act, mu2, cams, psf),
                  orc.project(act, mu2, cams, psf, dtype=np.float64))


def test_osem_after_operations_that_grow_the_buffers():
    """OSEM caches a launch table that refers to the plan's work buffers; a
    backprojection of all views in between grows (reallocates) the
    rotated-frame buffer, and the next OSEM call must rebuild the table: its
    result is bit-identical to a plan that ran the same two OSEM calls alone."""
    from niftyrec_amd.api import Plan
    act, mu, cams, psf = _case((48, 20, 48), 16, 5, seed=93)
    y = orc.project(act, mu, cams, psf)
    Y, M = _t(y), _t(mu)
    plans = [Plan(act.shape, cams, psf) for _ in range(2)]
    for p in plans:
        p.set_tuning(4, 0)  # one chunking for both (the timed choice may differ)
    X = [torch.ones(act.shape, dtype=torch.float32, device="cuda") for _ in range(2)]
    for p, x in zip(plans, X):
        p.osem(x, Y, M, subsets=4, iterations=1)
    plans[0].backproject(Y, M)  # 16 views of rotated-frame volumes, against OSEM's 4
    plans[0].project(_t(act), M)
    for p, x in zip(plans, X):
        p.osem(x, Y, M, subsets=4, iterations=1)
    torch.cuda.synchronize()
    assert torch.equal(X[0], X[1])
    x32, x64 = np.ones_like(act), np.ones(act.shape)
    for _ in range(2):
        for s in range(4):
            x32 = orc.mlem_step(x32, y[s::4], mu, cams[s::4], psf)
            x64 = orc.mlem_step(x64, y[s::4], mu, cams[s::4], psf, dtype=np.float64)
    report_parity("osem after operations that grow the buffers", X[0].cpu().numpy(), x32, x64)


def test_batch_after_operations_that_grow_the_buffers():
    """A batch replays a captured graph that refers to its plans' work buffers;
    after a single operation of one plan reallocates them, the next batch call
    captures again: bit-identical to a batch whose plans ran nothing else."""
    from niftyrec_amd.api import Batch, Plan
    cases = [_case((48, 20, 48), 16, 5, seed=94), _case((40, 24, 40), 12, 3, seed=95)]
    ys = [orc.project(act, mu, cams, psf) for act, mu, cams, psf in cases]
    runs = []
    for grow in (True, False):
        plans = [Plan(act.shape, cams, psf) for act, mu, cams, psf in cases]
        for p in plans:
            p.set_tuning(4, 0)
        batch = Batch(plans)
        X = [torch.ones(c[0].shape, dtype=torch.float32, device="cuda") for c in cases]
        Y = [_t(y) for y in ys]
        M = [_t(c[1]) for c in cases]
        batch.osem(X, Y, M, subsets=4, iterations=1)
        if grow:
            plans[0].backproject(Y[0], M[0])  # 16 views of rotated-frame volumes, against OSEM's 4
        batch.osem(X, Y, M, subsets=4, iterations=1)
        torch.cuda.synchronize()
        runs.append([x.clone() for x in X])
    for i in range(len(cases)):
        assert torch.equal(runs[0][i], runs[1][i]), "problem %d" % i


def test_batch_of_sweep_kernel_plans():
    """Plans the row-lane kernels do not take (140 axial rows), PSF radius 8
    and 2, in one batch: its grouped launches size their grids from the
    sweeps' slab heights; each result is bit-identical to the plan's own
    et_plan_osem with the same chunking."""
    from niftyrec_amd.api import Batch, Plan
    cases = [_case((40, 140, 40), 6, 17, seed=98), _case((32, 140, 32), 8, 5, seed=99)]
    ys = [_t(orc.project(act, mu, cams, psf)) for act, mu, cams, psf in cases]
    M = [_t(c[1]) for c in cases]

    def plans():
        ps = [Plan(act.shape, cams, psf) for act, mu, cams, psf in cases]
        for p in ps:
            p.set_tuning(4, 0)
        return ps

    Xb = [torch.ones(c[0].shape, dtype=torch.float32, device="cuda") for c in cases]
    Batch(plans()).osem(Xb, ys, M, subsets=2, iterations=1)
    for i, p in enumerate(plans()):
        Xs = torch.ones(cases[i][0].shape, dtype=torch.float32, device="cuda")
        p.osem(Xs, ys[i], M[i], subsets=2, iterations=1)
        assert p.projector() == 0
        torch.cuda.synchronize()
        assert torch.equal(Xs, Xb[i]), "problem %d" % i


@pytest.mark.parametrize("shape, V, K, truncate", [((40, 20, 40), 10, 5, False), ((96, 20, 96), 8, 7, True),
                                                    ((130, 24, 130), 6, 11, False), ((17, 1, 17), 3, 3, False)])
def test_project_backward_matches_the_two_operations(shape, V, K, truncate):
    """et_plan_project_backward right after a projection of the same pair (one
    pass: both rotated-frame sweeps, one exact adjoint pull for both outputs):
    bit-identical to backproject(ADJOINT) and gradient_attenuation(exact) with
    all views in one batch, and against the oracle.  Then with another
    activity (no projection of it: the two operations in turn)."""
    from niftyrec_amd.api import ADJOINT, Plan
    act, mu, cams, psf = _case(shape, V, K, seed=101)
    plan = Plan(act.shape, cams, psf)
    plan.set_tuning(0, V)
    A, M = _t(act), _t(mu)
    s = np.random.default_rng(102).random((V,) + act.shape[1:], dtype=np.float32)
    if truncate:
        s -= np.float32(0.3)  # negative sums: the truncation matters
    S = _t(s)
    plan.project(A, M, truncate=False)
    ga, gm = plan.project_backward(S, A, M, truncate=truncate)
    ba = plan.backproject(S, M, mode=ADJOINT, truncate=truncate)
    g = plan.gradient_attenuation(S, A, M, exact=True, truncate=truncate)
    torch.cuda.synchronize()
    assert torch.equal(ga, ba)
    assert torch.equal(gm, g)
    tag = "%dx%dx%d, %d views, K=%d%s" % (shape[2], shape[1], shape[0], V, K, ", truncated" if truncate else "")
    tr = 1 if truncate else 0
    report_parity("project_backward, activity (%s)" % tag, ga.cpu().numpy(),
                  orc.backproject(s, mu, cams, psf, 0.0, 0.0, tr, mode=1),
                  orc.backproject(s, mu, cams, psf, 0.0, 0.0, tr, mode=1, dtype=np.float64))
    if not truncate:  # (the oracle's gradient is untruncated)
        report_parity("project_backward, attenuation (%s)" % tag, gm.cpu().numpy(),
                      orc.gradient_mu(s, act, mu, cams, psf, dtype=np.float32),
                      orc.gradient_mu(s, act, mu, cams, psf, dtype=np.float64))
    act2 = act * np.float32(0.5)
    ga2, gm2 = plan.project_backward(S, _t(act2), M, truncate=truncate)
    assert torch.equal(ga2, ba)  # (the activity's adjoint does not depend on the activity)
    if not truncate:
        report_parity("project_backward without the projection (%s)" % tag, gm2.cpu().numpy(),
                      orc.gradient_mu(s, act2, mu, cams, psf, dtype=np.float32),
                      orc.gradient_mu(s, act2, mu, cams, psf, dtype=np.float64))


def test_project_backward_on_sweep_kernel_plans():
    """A plan the row-lane kernels do not take (140 axial rows):
    et_plan_project_backward runs the two operations in turn, bit-identical to
    them, and torch autograd through it matches the oracle."""
    from niftyrec_amd.api import ADJOINT, Plan
    from niftyrec_amd.torch_ops import project
    act, mu, cams, psf = _case((24, 140, 24), 4, 5, seed=103)
    plan = Plan(act.shape, cams, psf)
    A, M = _t(act), _t(mu)
    s = np.random.default_rng(104).random((4,) + act.shape[1:], dtype=np.float32)
    S = _t(s)
    plan.project(A, M, truncate=False)
    ga, gm = plan.project_backward(S, A, M)
    assert plan.projector() == 0
    assert torch.equal(ga, plan.backproject(S, M, mode=ADJOINT, truncate=False))
    assert torch.equal(gm, plan.gradient_attenuation(S, A, M, exact=True, truncate=False))
    Av, Mv = A.clone().requires_grad_(True), M.clone().requires_grad_(True)
    (project(Av, Mv, plan) * S).sum().backward()
    report_parity("autograd on the sweep kernels, d/d activity", Av.grad.cpu().numpy(),
                  orc.backproject(s, mu, cams, psf, 0.0, 0.0, 0, mode=1),
                  orc.backproject(s, mu, cams, psf, 0.0, 0.0, 0, mode=1, dtype=np.float64))
    report_parity("autograd on the sweep kernels, d/d mu", Mv.grad.cpu().numpy(),
                  orc.gradient_mu(s, act, mu, cams, psf, dtype=np.float32),
                  orc.gradient_mu(s, act, mu, cams, psf, dtype=np.float64))


def test_repeated_operations_with_new_content():
    """Repeated single operations over the plan's same work buffers reuse the
    launch table on the device (not uploaded again); each call still sees its
    own inputs: two activities, then two sinograms, each against the oracle."""
    from niftyrec_amd.api import ADJOINT, Plan
    act, mu, cams, psf = _case((48, 20, 48), 6, 5, seed=105)
    plan = Plan(act.shape, cams, psf)
    M = _t(mu)
    rng = np.random.default_rng(106)
    for k in range(2):
        a = act * np.float32(1.0 + k) + rng.random(act.shape, dtype=np.float32) * np.float32(0.1)
        p = plan.project(_t(a), M).cpu().numpy()
        report_parity("repeated plan project, call %d" % k, p, orc.project(a, mu, cams, psf),
                      orc.project(a, mu, cams, psf, dtype=np.float64))
    for k in range(2):
        y = rng.random((6,) + act.shape[1:], dtype=np.float32)
        b = plan.backproject(_t(y), M, mode=ADJOINT, truncate=False).cpu().numpy()
        report_parity("repeated plan backproject, call %d" % k, b,
                      orc.backproject(y, mu, cams, psf, 0.0, 0.0, 0, mode=1),
                      orc.backproject(y, mu, cams, psf, 0.0, 0.0, 0, mode=1, dtype=np.float64))


def test_profile_of_sweep_kernel_plans():
    """Profiling a plan the row-lane kernels do not take: its projections do
    not take the events offered to a row-lane projector launch and are timed
    by events recorded around them; every category holds its launches."""
    from niftyrec_amd.api import Plan
    act, mu, cams, psf = _case((24, 140, 24), 8, 5, seed=107)
    plan = Plan(act.shape, cams, psf)
    plan.set_tuning(4, 0)
    Y, M = _t(orc.project(act, mu, cams, psf)), _t(mu)
    X = torch.ones(act.shape, dtype=torch.float32, device="cuda")
    plan.osem(X, Y, M, subsets=2, iterations=1)  # caches built outside the count
    plan.profile(True)
    plan.osem(X, Y, M, subsets=2, iterations=2)
    r = plan.profile_read()
    plan.profile(False)
    assert plan.projector() == 0
    for cat in ("project", "bp_rotated_frame", "bp_pull"):
        assert r[cat][0] >= 4 and r[cat][1] > 0, (cat, r[cat])  # (at least one launch per subset update)
