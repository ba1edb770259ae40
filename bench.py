"""Benchmark: OSEM iterations/s for SPECT at 128^3 x 120 views, depth-dependent
PSF 11x11 and attenuation, 8 ordered subsets of 15 views, on one MI355X with
every input resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-batch]
                    [--workload spect128|batch48|joint256|autograd128|hbm_copy] [--depth-chunks D]

A step is one OSEM iteration: 8 subset updates, each = projection of the
subset's 15 views -> ratio -> reference backprojection -> multiplicative
update (sensitivities cached).  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
METRIC = "OSEM iters/s at 128³×120 views PSF+atten; GVoxel-views/s; % HBM peak"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batch48 speed-ups of the default line")
    ap.add_argument("--workload", default="spect128")
    ap.add_argument("--depth-chunks", type=int, default=0,
                    help="fix the sweeps' depth chunking (0: timed by the plan); profiler runs fix it "
                         "because counter collection perturbs the timing")
    return ap.parse_args()


# Kernels of the profiled categories in the OSEM subset loop, as named in the
# rocprofv3 summaries: the row-lane sweeps (PSF radius R, RPL rows per lane of
# a 16-lane row holding a column's whole axial extent, NR column rounds), the
# tiled projector when the plan chose it (Plan.projector() > 1), and the pull
# with the OSEM update.
def kernel_of(cat, R, ny, nx, tiled):
    rpl, nr = -(-ny // 32) * 2, -(-nx // 64)
    project = ("void (anonymous namespace)::k_yl_project_tile<%d, %d>" % (R, rpl) if tiled else
               "void (anonymous namespace)::k_yl_project<%d, %d, %d>" % (R, rpl, nr))
    return {"project": project,
            "bp_rotated_frame": "void (anonymous namespace)::k_yl_bpu<%d, %d, %d>" % (R, rpl, nr),
            "bp_pull": "void (anonymous namespace)::k_yl_pull<%d, 2>" % rpl}.get(cat)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "osem_spect128_hbm_traffic.txt")


def measured_traffic(prefix, grid=None):
    """HBM bytes per launch of the kernel named `prefix` from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary: the row of that launch grid
    (work-items), else the row with the most dispatches; None when the summary
    lacks the kernel."""
    if prefix is None or not os.path.exists(TRAFFIC_FILE):
        return None
    best = None
    for ln in open(TRAFFIC_FILE):
        if not ln.startswith(prefix):
            continue
        f = ln.split()
        g, n, mb = int(f[-5]), int(f[-4]), float(f[-1])
        if grid is not None and g == grid:
            return mb * 1e6
        if best is None or n > best[0]:
            best = (n, mb)
    return None if best is None else best[1] * 1e6


def sweep_grid(cat, nx, nz, views, depth_chunks, tiled=False):
    """Work-items of one subset launch: a sweep runs views x depth chunks blocks
    of 16 W threads (W slots of at most 3 column rounds, a multiple of 4), the
    tiled projector views x depth chunks x ceil(nx / 64) blocks of 512; the
    pull ceil(nx nz / 16) blocks of 256."""
    if cat == "bp_pull":
        return -(-nx * nz // 16) * 256
    if tiled and cat == "project":
        return views * depth_chunks * -(-nx // 64) * 512
    nr = -(-nx // 64)
    w = (-(-nx // nr) + 3) // 4 * 4
    return views * depth_chunks * 16 * w


def cpu_baseline(w, subsets, y_sub, x_gpu):
    """Reference CPU path (oracle/_ref, single thread as built by the reference)
    on one subset update of the same workload (the first ordered subset, from
    x = 1, measured counts y_sub); falls back to the oracle port.  The same
    update is checked here: x_gpu (the library's update) against the
    reference's result and against the double-precision oracle.  Returns
    (cpu_baseline record, error record)."""
    from oracle import refabi, orc
    a, m, p, c = w["activity"], w["attenuation"], w["psf"], w["cameras"]
    views = c[0::subsets]  # the first ordered subset: 15 views
    if os.path.exists(refabi.REF_PATH):
        R = refabi.load_ref()
        x32 = np.ones_like(a)
        t = time.perf_counter()
        st, x32 = refabi.mlem_step(R, x32, y_sub, m, views, p)
        dt = time.perf_counter() - t
        kind, cores = "reference", 1
        assert st == 0
    else:
        t = time.perf_counter()
        x32 = orc.mlem_step(np.ones_like(a), y_sub, m, views, p)
        dt = time.perf_counter() - t
        kind, cores = "port", os.cpu_count()
    it_time = dt * subsets  # one iteration = `subsets` subset updates of this size
    base = {"value": 1.0 / it_time, "unit": "OSEM it/s", "cores": cores, "kind": kind,
            "sample": "one subset update (et_array_mlem_step on %d of %d views, 128^3, PSF 11, attenuation; "
                      "the reference recomputes the sensitivity) took %.2f s; iteration = %d such updates"
                      % (len(views), len(c), dt, subsets)}
    x64 = orc.mlem_step(np.ones(a.shape), y_sub, m, views, p, dtype=np.float64)

    def rel(u, v):
        u, v = u.astype(np.float64), v.astype(np.float64)
        return float(np.linalg.norm(u - v) / np.linalg.norm(v))

    def mrel(u, v):
        u, v = u.astype(np.float64), v.astype(np.float64)
        return float(np.abs(u - v).max() / np.abs(v).max())

    err = {"update": "first ordered subset (%d views) from x = 1, measured counts" % len(views),
           "relL2_vs_reference": rel(x_gpu, x32), "max_abs_over_peak_vs_reference": mrel(x_gpu, x32),
           "relL2_vs_fp64": rel(x_gpu, x64), "reference_relL2_vs_fp64": rel(x32, x64),
           "reference": kind}
    return base, err


def run_batch48(steps, warmup, per_call=True):
    """48 uneven problems (phantoms.batch_problem), 8 ordered subsets: one OSEM
    iteration of every problem through et_batch_osem (grouped launches, one
    hipGraph replay), timed over `steps`; the same plans run one after another
    (et_plan_osem); and the per-call ABI path (et_array_mlem_step on host
    buffers, one call per subset; sample: subset 0 of every problem, times 8).
    Returns a dict of the timings and speed-ups."""
    import ctypes
    import torch
    from niftyrec_amd import api, phantoms, _lib

    S, N = 8, 48
    dev = torch.device("cuda:0")
    probs = [phantoms.batch_problem(i) for i in range(N)]
    plans, X, Y, M = [], [], [], []
    for w in probs:
        p = api.Plan(w["activity"].shape, w["cameras"], w["psf"])
        A = torch.from_numpy(w["activity"]).to(dev)
        Mu = torch.from_numpy(w["attenuation"]).to(dev)
        y = p.project(A, Mu).cpu().numpy()
        plans.append(p)
        Y.append(torch.from_numpy(phantoms.poisson_data(y, w["counts"], seed=2)).to(dev))
        M.append(Mu)
        X.append(torch.ones_like(A))
    batch = api.Batch(plans)
    batch.osem(X, Y, M, subsets=S, iterations=max(warmup, 1))
    torch.cuda.synchronize()
    for x in X:
        x.fill_(1.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.osem(X, Y, M, subsets=S, iterations=1)
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / steps
    assert all(torch.isfinite(x).all() for x in X)

    # plans one after another (same kernels, one problem per launch, no graph)
    for p, x, y, m in zip(plans, X, Y, M):
        p.osem(x, y, m, subsets=S, iterations=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for p, x, y, m in zip(plans, X, Y, M):
        p.osem(x, y, m, subsets=S, iterations=1)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    out = {"problems": N, "subsets": S, "ms_per_iteration": t_batch * 1e3,
           "plans_sequential_ms_per_iteration": t_seq * 1e3, "speedup_vs_sequential_plans": t_seq / t_batch,
           "gvoxel_views_per_s": sum(w["activity"].size * len(w["cameras"]) for w in probs) / t_batch / 1e9}
    if per_call:
        L = _lib.load()
        fp = ctypes.POINTER(ctypes.c_float)
        t_call = 0.0
        for w, y in zip(probs, Y):
            a = np.ones_like(w["activity"])
            ysub = np.ascontiguousarray(y.cpu().numpy()[0::S])
            cams = np.ascontiguousarray(np.stack([np.zeros_like(w["cameras"][0::S]), w["cameras"][0::S],
                                                  np.zeros_like(w["cameras"][0::S])]))
            psf = np.ascontiguousarray(w["psf"])
            nz, ny, nx = a.shape
            mu = np.ascontiguousarray(w["attenuation"])
            K = psf.shape[2]
            t1 = time.perf_counter()
            st = L.et_array_mlem_step(a.ctypes.data_as(fp), nx, ny, ysub.ctypes.data_as(fp), ysub.shape[0],
                                      cams.ctypes.data_as(fp), 1, mu.ctypes.data_as(fp), 1, psf.ctypes.data_as(fp),
                                      K, K, 0.0, 0.0, 1e-6, 1)
            t_call += time.perf_counter() - t1
            assert st == 0, st
        out.update(per_call_abi_ms_per_iteration=t_call * S * 1e3, speedup_vs_per_call=t_call * S / t_batch,
                   per_call_sample="subset 0 of all 48 problems through et_array_mlem_step (host buffers), times 8")
    for p in plans:
        p.close()
    return out


def main_batch(args):
    """The batch48 workload on its own (see run_batch48): a step is one OSEM
    iteration of all 48 problems."""
    from niftyrec_amd import phantoms
    b = run_batch48(args.steps, args.warmup)
    line = {
        "metric": "batched OSEM iterations/s, 48 uneven SPECT problems",
        "value": 1e3 / b["ms_per_iteration"],
        "unit": "OSEM it/s (all 48 problems)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": b["ms_per_iteration"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantoms, Poisson counts)",
        "config": {"workload": "batch48: 48 problems, volumes from %s, 60-180 views, PSF 7/11/15, attenuation, "
                               "8 ordered subsets x 1 iteration per step" % phantoms.BATCH_SHAPES,
                   "problems": b["problems"], "subsets": b["subsets"]},
    }
    line.update({k: v for k, v in b.items() if k not in ("problems", "subsets", "ms_per_iteration")})
    print(json.dumps(line), flush=True)


def main_joint(args):
    """Joint activity + attenuation estimation at 256^3, 180 views, PSF 21x21,
    attenuation 0.15/0.05/0.30: a step is one OSEM iteration of the activity
    (10 ordered subsets of 18 views; the subset sensitivities are rebuilt because
    mu changed) followed by one projected gradient step on mu through
    torch.autograd (forward projection of all views, Poisson NLL, exact
    attenuation-gradient kernel; rotated volumes recomputed per view)."""
    import torch
    from niftyrec_amd import api, phantoms
    from niftyrec_amd.torch_ops import poisson_nll, project

    S = 10
    w = phantoms.workload("spect256")
    act, mu, psf, cams = w["activity"], w["attenuation"], w["psf"], w["cameras"]
    dev = torch.device("cuda:0")
    plan = api.Plan(act.shape, cams, psf)
    plan.set_tuning(16, 0)
    A = torch.from_numpy(act).to(dev)
    Mtrue = torch.from_numpy(mu).to(dev)
    Y = torch.from_numpy(phantoms.poisson_data(plan.project(A, Mtrue).cpu().numpy(), w["counts"], seed=3)).to(dev)
    X = torch.ones_like(A)
    M = (Mtrue * 0.8).contiguous()

    def step():
        # M changed in place (its version advanced): osem rebuilds the cached
        # sensitivities and attenuation factors
        plan.osem(X, Y, M, subsets=S, iterations=1)
        Mv = M.clone().requires_grad_(True)
        loss = poisson_nll(project(X, Mv, plan), Y)
        loss.backward()
        g = Mv.grad
        with torch.no_grad():
            M.sub_(0.01 * g / g.abs().max().clamp_min(1e-30)).clamp_(min=0.0)
        return loss

    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    assert torch.isfinite(X).all() and torch.isfinite(M).all() and torch.isfinite(loss)
    line = {
        "metric": "joint activity+attenuation steps/s at 256^3 x 180 views, PSF 21x21",
        "value": 1.0 / dt, "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantom, attenuation 0.15/0.05/0.30, Poisson counts %.0e)" % w["counts"],
        "config": {"workload": "joint256: 256^3, 180 views, PSF 21x21; step = OSEM iteration (10 subsets, "
                               "sensitivities rebuilt) + autograd gradient step on mu", "subsets": S},
    }
    print(json.dumps(line), flush=True)


def main_autograd(args):
    """The differentiable projector at 128^3, 120 views, PSF 11x11, with
    attenuation: a step is one forward projection of all views through
    torch.autograd, the Poisson negative log-likelihood of the measured counts,
    and its backward pass to both the activity (the exact adjoint
    backprojection) and the attenuation map (the exact attenuation-gradient
    kernel and its adjoint pull; rotated volumes recomputed per view)."""
    import torch
    from niftyrec_amd import api, phantoms
    from niftyrec_amd.torch_ops import poisson_nll, project

    w = phantoms.workload("spect128")
    act, mu, psf, cams = w["activity"], w["attenuation"], w["psf"], w["cameras"]
    dev = torch.device("cuda:0")
    plan = api.Plan(act.shape, cams, psf)
    A0 = torch.from_numpy(act).to(dev)
    M0 = torch.from_numpy(mu).to(dev)
    Y = torch.from_numpy(phantoms.poisson_data(plan.project(A0, M0).cpu().numpy(), w["counts"], seed=3)).to(dev)
    A = (A0 * 0.9 + 0.05).contiguous().requires_grad_(True)
    M = (M0 * 0.8).contiguous().requires_grad_(True)

    def step():
        A.grad = None
        M.grad = None
        loss = poisson_nll(project(A, M, plan), Y)
        loss.backward()
        return loss

    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    assert torch.isfinite(loss) and torch.isfinite(A.grad).all() and torch.isfinite(M.grad).all()
    nx, ny, nz = act.shape[2], act.shape[1], act.shape[0]
    V = len(cams)
    line = {
        "metric": "autograd steps/s (projection + Poisson NLL + gradients in activity and mu) at 128^3 x 120 "
                  "views, PSF 11x11",
        "value": 1.0 / dt, "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantom, Poisson counts %.0e)" % w["counts"],
        "config": {"workload": "autograd128: %dx%dx%d activity+attenuation, %d views, PSF %dx%d; step = forward "
                               "projection, Poisson NLL, backward to activity and mu" % (nx, ny, nz, V,
                                                                                        psf.shape[-1],
                                                                                        psf.shape[-1])},
        "gvoxel_views_per_s": V * nx * ny * nz / dt / 1e9,
    }
    print(json.dumps(line), flush=True)


def main_copy(args):
    """Device-memory bandwidth of plain copies (torch clone, read + write) and
    reads (torch row sums) of 64 MiB to 2 GiB: the ceilings the sweep kernels'
    measured traffic is read against, and the calibration run for the
    FETCH_SIZE / WRITE_SIZE counters (a copy of known size).  `value` is the
    256 MiB copy."""
    import torch

    def rate(fn, nbytes):
        for _ in range(max(args.warmup, 1)):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        return nbytes / dt / 1e9, dt

    table = {}
    value, ms = None, None
    for mib in (64, 256, 1024, 2048):
        x = torch.ones(mib * 256 * 1024, device="cuda:0")
        nb = mib * 1024 * 1024
        cp, dt = rate(lambda: x.clone(), 2 * nb)
        rd, _ = rate(lambda: x.view(-1, 1024).sum(1), nb)  # row sums: a streaming read, small write
        table["%d MiB" % mib] = {"copy_GBps": round(cp, 1), "read_GBps": round(rd, 1)}
        if mib == 256:
            value, ms = cp, dt * 1e3
        del x
        torch.cuda.empty_cache()
    line = {"metric": "device copy bandwidth (read + write)", "value": value, "unit": "GB/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "hbm_copy: 256 MiB torch clone (counter calibration); copies and reads of "
                                   "64 MiB to 2 GiB", "sizes": table}}
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (use --gpus 1)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch48":
        return main_batch(args)
    if args.workload == "joint256":
        return main_joint(args)
    if args.workload == "autograd128":
        return main_autograd(args)
    if args.workload == "hbm_copy":
        return main_copy(args)
    import torch
    from niftyrec_amd import api, phantoms

    w = phantoms.workload(args.workload)
    S = w["subsets"] if args.workload == "spect128" else 1
    act, mu, psf, cams = w["activity"], w["attenuation"], w["psf"], w["cameras"]
    nz, ny, nx = act.shape
    V = len(cams)
    dev = torch.device("cuda:0")
    A = torch.from_numpy(act).to(dev)
    M = torch.from_numpy(mu).to(dev)
    plan = api.Plan(act.shape, cams, psf)
    if args.depth_chunks > 0:
        plan.set_tuning(args.depth_chunks, 0)
    noiseless = plan.project(A, M).cpu().numpy()
    Y = torch.from_numpy(phantoms.poisson_data(noiseless, w["counts"], seed=1)).to(dev)
    X = torch.ones_like(A)

    # warmup (first call builds the cached sensitivities and attenuation factors)
    plan.osem(X, Y, M, subsets=S, iterations=max(args.warmup, 1))
    torch.cuda.synchronize()

    dchunks = plan.tuning()  # depth chunks chosen by the plan's timed tuning (projection, backprojection)

    # untimed pass with every launch timed: per-kernel split, dominant kernel
    plan.profile(True)
    for _ in range(min(args.steps, 5)):
        plan.osem(X, Y, M, subsets=S, iterations=1)
    split = plan.profile_read()
    n_split = min(args.steps, 5)
    cat = max((k for k in split if split[k][0] > 0), key=lambda k: split[k][1])

    # timed region: HIP events (on the plan's stream) around every (S+1)-th launch
    # of the dominant kernel: one per step, rotating over the subsets
    X.fill_(1.0)
    plan.profile(True, categories=[cat], every=S + 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.osem(X, Y, M, subsets=S, iterations=1)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prof = plan.profile_read()
    plan.profile(False)
    assert torch.isfinite(X).all()

    ms = elapsed / args.steps * 1e3
    its = args.steps / elapsed
    n = nx * ny * nz
    P = nx * ny
    bytes_iter = V * (16.0 * n + 8.0 * P) + S * 16.0 * n  # per view: act+mu read, sino read+write; per subset: x, sens read, x write
    launches, tot_ms, tot_bytes = prof[cat]
    kname = kernel_of(cat, (psf.shape[-1] - 1) // 2, ny, nx, plan.projector() > 1)
    achieved = tot_bytes / (tot_ms * 1e-3)
    line = {
        "metric": METRIC,
        "value": its,
        "unit": "OSEM it/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantom, Poisson counts %.0e)" % w["counts"],
        "config": {"workload": "%s: %dx%dx%d activity+attenuation, %d views over 360 deg, depth-dependent "
                               "Gaussian PSF %dx%d, %d ordered subsets x 1 iteration per step"
                               % (args.workload, nx, ny, nz, V, psf.shape[-1], psf.shape[-1], S),
                   "volume": [nx, ny, nz], "views": V, "psf": int(psf.shape[-1]), "subsets": S},
        "gvoxel_views_per_s": V * n / (ms * 1e-3) / 1e9,
        "hbm_frac_of_peak": bytes_iter / (ms * 1e-3) / HBM_PEAK,
        "roofline": {"bound": "hbm", "kernel": cat, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": measured_traffic(kname, sweep_grid(cat, nx, nz, V // S,
                                                                   dchunks[cat == "bp_rotated_frame"],
                                                                   "project_tile" in (kname or ""))),
                     "kernel_name": kname,
                     "avg_launch_ms": tot_ms / launches, "algorithmic_bytes_per_launch": tot_bytes / launches,
                     "depth_chunks": dchunks[cat == "bp_rotated_frame"]},
        "kernel_ms_per_step": {k: v[1] / n_split for k, v in split.items() if v[0]},
        "kernel_split_note": "per-kernel times from an untimed pass with every launch timed; the timed "
                             "region times every %d-th launch of the dominant kernel (%d launches timed)"
                             % (S + 1, launches),
    }
    if not args.no_cpu_baseline:
        # the first ordered subset's update from x = 1 on the GPU (a plan over
        # its views), checked in the CPU-baseline leg
        sub = api.Plan(act.shape, cams[0::S], psf)
        Xs = torch.ones_like(A)
        Ys = Y[0::S].contiguous()
        sub.osem(Xs, Ys, M, subsets=1, iterations=1)
        torch.cuda.synchronize()
        line["cpu_baseline"], line["subset_update_error"] = cpu_baseline(w, S, Ys.cpu().numpy(), Xs.cpu().numpy())
        sub.close()
    if not args.no_batch:
        # batches of uneven problems (outside the timed region): speed-ups of one
        # hipGraph replay of grouped launches over the plans run one after
        # another and over the per-call ABI
        line["batch48"] = run_batch48(3, 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
