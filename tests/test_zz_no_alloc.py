"""et_plan_project_noalloc through the C ABI: the projection whose temporaries
live in the library's static device scratch array, against et_plan_project on
the same plan (bit-identical with the same depth chunking where the plan also
projects with the sweep kernels, i.e. without attenuation, and within the
parity tolerance of the plan's own path), on calls in a row whose later
inputs are smaller (the scratch keeps the last call's contents), and its
refusals (a plan whose copies exceed the array, a non-axial plan)."""
import ctypes

import numpy as np
import pytest

from conftest import max_rel, rel_l2
from niftyrec_amd import phantoms

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _problem(shape, V, K, seed):
    rng = np.random.default_rng(seed)
    act, mu = phantoms.phantom(shape, 3, seed)
    act = act + 0.1 * rng.random(shape, dtype=np.float32)
    mu = (mu / 0.15 * 0.1).astype(np.float32)
    cams = rng.uniform(0, 2 * np.pi, V).astype(np.float32)
    psf = phantoms.gaussian_psf(shape[2], K, 0.5, 1.6) if K else None
    return act, mu, cams, psf


def _t(a):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _noalloc(plan, act, mu, truncate=1):
    out = torch.full(plan.sino_shape, -7.0, dtype=torch.float32, device="cuda")
    st = plan._L.et_plan_project_noalloc(plan._h, _ptr(act), _ptr(mu), _ptr(out), truncate,
                                         ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return st, out


@pytest.mark.parametrize("with_mu_second", [False, True])
def test_noalloc_matches_plan_project_on_calls_in_a_row(with_mu_second):
    from niftyrec_amd.api import Plan
    big = _problem((128, 128, 128), 24, 11, seed=70)
    small = _problem((48, 40, 48), 8, 5, seed=71)
    runs = [(big, True), (small, with_mu_second)]
    for (act_np, mu_np, cams, psf), use_mu in runs:
        print("no-alloc projection: shape %s, %d views, mu %s" % (act_np.shape, len(cams), use_mu), flush=True)
        plan = Plan(act_np.shape, cams, psf)
        act, mu = _t(act_np), _t(mu_np) if use_mu else None
        st, got = _noalloc(plan, act, mu)
        assert st == 0
        if mu is None:
            # without attenuation the plan projects with the same sweep kernels
            plan.set_tuning(1, 0)  # one depth chunk, as the no-alloc path
            ref1 = plan.project(act, mu)
            plan.set_tuning(0, 0)
            torch.cuda.synchronize()
            assert torch.equal(got, ref1)
            print("  bit-identical to et_plan_project with one depth chunk")
        # the plan's own path (with attenuation the row-lane kernels, cached
        # factors; otherwise the sweep kernels with its chunking)
        ref = plan.project(act, mu)
        torch.cuda.synchronize()
        g, r = got.cpu().numpy(), ref.cpu().numpy()
        e, m = rel_l2(g, r), max_rel(g, r)
        print("  relL2 vs et_plan_project %.2e, max rel %.2e" % (e, m))
        assert e <= 1e-6 and m <= 1e-5
        plan.close()


def test_noalloc_refusals():
    from niftyrec_amd.api import Plan
    # 256^3: one padded copy alone exceeds the 32 MiB array
    cams = np.linspace(0, np.pi, 4, dtype=np.float32)
    plan = Plan((256, 256, 256), cams, None)
    act = torch.ones((256, 256, 256), dtype=torch.float32, device="cuda")
    st, out = _noalloc(plan, act, None)
    assert st == 2
    assert bool((out == -7.0).all())  # nothing written
    plan.close()
    # rotations about more than one axis take the general path: refused
    act_np, mu_np, _, psf = _problem((32, 24, 32), 5, 5, seed=72)
    c3 = np.random.default_rng(72).uniform(-0.6, 0.6, (3, 5)).astype(np.float32)
    plan = Plan(act_np.shape, c3, psf)
    st, _ = _noalloc(plan, _t(act_np), _t(mu_np))
    assert st == 2
    # the activity is required
    plan2 = Plan(act_np.shape, np.zeros(5, np.float32), psf)
    st, _ = _noalloc(plan2, None, _t(mu_np))
    assert st == 2
    plan.close()
    plan2.close()
