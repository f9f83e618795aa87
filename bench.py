"""Benchmark: polarized pixels/s of the per-pixel ray pipeline on one MI355X.

Workload (the configuration the headline metric is quoted on): a seeded
synthetic iharm-format FMKS dump of 288x128x128 cells (a = 0.9375), resident
in HBM after on-device derivation, imaged at 160x160 pixels, thetacam 163,
230 GHz, R_high = 40, M_unit = 3e26, FP64.  One step = one full image
(geodesic pass + polarized transfer pass + output layout) with the inputs
already on the device; outputs stay on the device.

Prints one JSON line: the metric and its value (pixels/s of the image), the
roofline entry of the geodesic launches, the CPU baseline, and the batch,
mixed-precision, adjoint and adaptive measurements.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "polarized pixels/s on 160^2 FMKS dump: batched/mixed vs plain FP64, dFtot<=1e-4"
PEAK_FP64 = 78.6  # TFLOP/s, MI355X FP64 vector
C1 = dict(N=(288, 128, 128), a=0.9375, seed=1, M_unit=3e26, nx=160, ny=160, thetacam=163.,
          freqcgs=230e9, trat_large=40.)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(h, prims, cfg, sample_stride, nthreads):
    """The oracle (C + OpenMP restatement of the reference) on the host cores,
    on a bounded, evenly strided sample of the same image's pixels."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    t0 = time.time()
    # the build with multiply-add contraction (GCC's default, as in the
    # reference's makefile) rather than the bit-exact checking build
    fma = os.path.join(os.path.dirname(O.LIB), "liboracle_fma.so")
    lib = O.load(fma) if os.path.exists(fma) else None
    build = ("gcc -std=gnu99 -O3 -march=x86-64-v3 -funroll-loops -fopenmp %s (oracle/Makefile; the reference's "
             "makefile: -O3 -march=native -mtune=native -flto -fopenmp -funroll-loops)"
             % ("-ffp-contract=fast" if lib is not None else "-ffp-contract=off"))
    orc, _ = O.iharm_oracle(h, prims, M_unit=cfg["M_unit"], trat_large=cfg["trat_large"],
                            **({"L": lib} if lib is not None else {}))
    log("cpu baseline: derivation %.1fs" % (time.time() - t0))
    p = O.default_params(nx=cfg["nx"], ny=cfg["ny"], thetacam=cfg["thetacam"], freqcgs=cfg["freqcgs"])
    cam = orc.camera(p)
    pix = np.array([(i, j) for i in range(0, cfg["nx"], sample_stride)
                    for j in range(cfg["ny"])], np.int32)
    t0 = time.time()
    rc, out7, ns, fl = orc.render_pixels(p, cam, pix, nthreads=nthreads)
    dt = time.time() - t0
    assert rc == 0
    log("cpu baseline: %d pixels in %.1fs on %d threads" % (len(pix), dt, nthreads))
    return dict(value=len(pix) / dt, unit="pixels/s", cores=nthreads, kind="port", build=build,
                sample="%d of %d pixels (every %d-th pixel row, all columns) of the same 160x160 image, "
                       "oracle C+OpenMP, %d threads" % (len(pix), cfg["nx"] * cfg["ny"], sample_stride,
                                                       nthreads))


def roofline(stage_ms, n_geo, streamed):
    """Roofline of the dominant kernel (longest stage): algorithmic FLOP per
    launch = per-unit count (op-counting build, ipole_amd/flops_per_step.json)
    x units (geodesic steps = transfer steps per image); traffic = measured
    HBM bytes per launch (ipole_amd/hbm_traffic.json).  In the streamed
    render the geodesic pass is one launch per phase (k_trace_budget, then
    k_trace_resume_mask on the rays still live), timed together; the
    transfer kernels run beside the later phases and are not timed one by
    one."""
    fl = json.load(open(os.path.join(ROOT, "ipole_amd", "flops_per_step.json")))["pol"]
    tr = json.load(open(os.path.join(ROOT, "ipole_amd", "hbm_traffic.json")))["bytes_per_launch"]
    per_unit = dict(k_trace=fl["F_trace_per_step"], k_prep=fl["F_prep_per_step"],
                    k_chain=fl["F_chain_per_step"])
    if streamed:
        t = dict(k_trace=stage_ms[0])
    else:
        t = dict(k_trace=stage_ms[0], k_prep=stage_ms[1], k_chain=stage_ms[2])
    ach = {k: per_unit[k] * n_geo / (t[k] * 1e-3) / 1e12 for k in t}
    dom = max(t, key=t.get)
    name = "k_trace_budget+k_trace_resume_mask" if streamed and dom == "k_trace" else dom
    return dict(bound="fp64-valu", achieved=round(ach[dom], 4), peak=PEAK_FP64, unit="TFLOP/s",
                frac=round(ach[dom] / PEAK_FP64, 5), traffic=tr.get(dom), kernel=name,
                note="FP64 compute roof (MI355X FP64 peak 78.6 TFLOP/s, which the FP64 vector ALU "
                     "also reaches; these kernels use the vector ALU, no matrix instructions). "
                     "FLOP per unit from the op-counting build (%s: %.0f per step); per-kernel "
                     "TFLOP/s: %s; traffic = HBM bytes per launch from FETCH_SIZE/WRITE_SIZE "
                     "counters scaled by a plain-copy calibration"
                     % (dom, per_unit[dom], ", ".join("%s %.3f" % (k, v) for k, v in ach.items())))


def batch_measure(args, h=None, prims=None):
    """Batched render of the 64-job set in one call (FP64, then mixed
    precision) against the same jobs rendered one by one with the plain path
    (dump re-derived per M_unit).  Returns a dict of the figures."""
    from ipole_amd import ipgpu as G
    from ipole_amd import synth
    cfg = dict(C1)
    if h is None:
        h, prims = synth.make_dump(*cfg["N"], a=cfg["a"], seed=cfg["seed"])
    jobs = synth.batch_jobs()
    if args.small:
        jobs = jobs[:8]
    M_ref = 1e25
    fl = json.load(open(os.path.join(ROOT, "ipole_amd", "flops_per_step.json")))["pol"]
    ctx = G.Context(0)
    g, m = G.iharm_geom(h), G.iharm_model()
    ctx.set_geometry(g)
    ctx.set_model(m)
    ims, ims32 = [], []
    for jb in jobs:
        im, _, _ = G.make_image(g, m.L_unit, nx=jb["n"], ny=jb["n"], thetacam=jb["thetacam"],
                                freqcgs=jb["freqcgs"])
        ims.append(im)
        im, _, _ = G.make_image(g, m.L_unit, nx=jb["n"], ny=jb["n"], thetacam=jb["thetacam"],
                                freqcgs=jb["freqcgs"], transfer_fp32=1)
        ims32.append(im)
    npix = sum(jb["n"] ** 2 for jb in jobs)
    # plain path, one job at a time
    plain_ms, Fplain = [], []
    for jb, im in zip(jobs, ims):
        ctx.upload_prims(prims, G.iharm_units(jb["M_unit"], trat_large=cfg["trat_large"]))
        # warm: the record pool and arenas sized for this job (a render that
        # grows them counts the allocation in its device time)
        for _ in range(2):
            ctx.render(im, fetch=False)
        t = []
        for _ in range(2):
            ctx.render(im, fetch=False)
            t.append(ctx.kernel_times()[2])
        plain_ms.append(min(t))
        r = ctx.render(im)
        Fplain.append(r["imageS"][..., 0].sum())
    log("plain: %d jobs, %d pixels, %.1f ms device" % (len(jobs), npix, sum(plain_ms)))
    ctx.upload_prims(prims, G.iharm_units(M_ref, trat_large=cfg["trat_large"]))
    batch = [(im, jb["M_unit"] / M_ref) for jb, im in zip(jobs, ims)]
    nrep = max(1, min(args.steps, 3))
    tuned = ctx.autotune_batch(batch)
    log("batch FP64 autotuned: %s" % tuned)
    t_dev = []
    for s in range(nrep):
        ctx.render_batch(batch, fetch=False)
        t_dev.append(ctx.kernel_times())
    res = ctx.render_batch(batch)
    n_rad = ctx.step_counts()[1]
    bc = ctx.batch_counts()
    n_xfer = int(sum(int(r["nstep"].sum()) for r in res))
    dF = max(abs(r["imageS"][..., 0].sum() / f - 1) for r, f in zip(res, Fplain))
    ms = float(np.mean([t[2] for t in t_dev]))
    ms_x = float(np.mean([t[1] for t in t_dev]))
    log("batch FP64: %.1f ms device (geodesic %.1f, transfer %.1f); max |dFtot|/Ftot vs plain %.2e" %
        (ms, np.mean([t[0] for t in t_dev]), ms_x, dF))
    # the transfer's algorithmic FLOP at the op-counting build's per-unit
    # counts: with several jobs of a ray in one lane, the shared part per
    # item (its steps, radiating steps, the item) and the per-job part per
    # (job, ray) pair; each job on its own, the single-job fused transfer's
    flop_single = (fl["F_fused_per_radiating_step"] * n_rad + fl["F_fused_per_nonradiating_step"] * (n_xfer - n_rad)
                   + fl["F_fused_per_ray"] * npix)
    if bc["items"] < bc["job_items"]:
        flop = (fl["F_multi_shared_per_radiating_step"] * bc["item_radiating_steps"]
                + fl["F_multi_shared_per_nonradiating_step"] * (bc["item_steps"] - bc["item_radiating_steps"])
                + fl["F_multi_shared_per_ray"] * bc["items"]
                + fl["F_multi_job_per_radiating_step"] * n_rad
                + fl["F_multi_job_per_nonradiating_step"] * (bc["job_steps"] - n_rad)
                + fl["F_multi_job_per_ray"] * bc["job_items"])
        kname = "k_xfer_multi<FP64>"
        fnote = ("FLOP = per item: radiating steps x %.0f + non-radiating x %.0f + %.0f (the geometry its jobs "
                 "share), plus per (job, ray): radiating steps x %.0f + non-radiating x %.0f + %.0f "
                 "(op-counting build, ipole_amd/flops_per_step.json); the jobs one by one would need %.3g"
                 % (fl["F_multi_shared_per_radiating_step"], fl["F_multi_shared_per_nonradiating_step"],
                    fl["F_multi_shared_per_ray"], fl["F_multi_job_per_radiating_step"],
                    fl["F_multi_job_per_nonradiating_step"], fl["F_multi_job_per_ray"], flop_single))
    else:
        flop = flop_single
        kname = "k_xfer_queue<FP64>" if tuned["queue"] else "k_xfer_batch<FP64>"
        fnote = ("FLOP = radiating steps x %.0f + non-radiating x %.0f + rays x %.0f (op-counting build, "
                 "ipole_amd/flops_per_step.json)" % (fl["F_fused_per_radiating_step"],
                                                     fl["F_fused_per_nonradiating_step"], fl["F_fused_per_ray"]))
    tflops = flop / (ms_x * 1e-3) / 1e12
    # the M_unit search: the same jobs at new scales on the resident records
    # (the transfer alone, replayed from its captured graph)
    batch_b = [(im, 1.5 * jb["M_unit"] / M_ref) for jb, im in zip(jobs, ims)]
    ctx.rerender_batch(batch_b, fetch=False)
    t_rr = []
    for s in range(nrep):
        ctx.rerender_batch(batch if s % 2 else batch_b, fetch=False)
        t_rr.append(ctx.kernel_times()[2])
    ms_rr = float(np.mean(t_rr))
    log("batch FP64 transfer: %d steps (%d radiating), %.3g FLOP, %.2f TFLOP/s (%.1f%% of FP64 peak); "
        "rerender at new M_unit scales %.1f ms" % (n_xfer, n_rad, flop, tflops, 100 * tflops / PEAK_FP64, ms_rr))
    batch32 = [(im, jb["M_unit"] / M_ref) for jb, im in zip(jobs, ims32)]
    tuned32 = ctx.autotune_batch(batch32)
    log("batch mixed autotuned: %s" % tuned32)
    t32 = []
    for s in range(nrep):
        ctx.render_batch(batch32, fetch=False)
        t32.append(ctx.kernel_times())
    res32 = ctx.render_batch(batch32)
    ctx.close()
    dF32 = max(abs(r["imageS"][..., 0].sum() / f - 1) for r, f in zip(res32, Fplain))
    ms32 = float(np.mean([t[2] for t in t32]))
    log("batch mixed: %.1f ms device (geodesic %.1f, transfer %.1f); max |dFtot|/Ftot vs plain FP64 %.2e" %
        (ms32, np.mean([t[0] for t in t32]), np.mean([t[1] for t in t32]), dF32))
    return dict(workload="64 jobs on the 288x128x128 FMKS dump (thetacam {17,50,90,130,163}, "
                         "{86,230,345} GHz, M_unit log-uniform 1e24-1e26, 64^2-256^2 px, seed 0)",
                pixels=npix, plain_ms_device=round(sum(plain_ms), 3),
                plain_pixels_per_s=round(npix / (sum(plain_ms) * 1e-3), 1),
                fp64_ms_device=round(ms, 3), fp64_pixels_per_s=round(npix / (ms * 1e-3), 1),
                fp64_speedup_vs_plain=round(sum(plain_ms) / ms, 3), fp64_max_dFtot_vs_plain=dF,
                fp64_autotuned=tuned, transfer_steps=n_xfer, radiating_steps=int(n_rad),
                fp64_transfer_ms_device=round(ms_x, 3),
                transfer_items=bc,
                fp64_transfer_roofline=dict(
                    bound="fp64-valu", kernel=kname, achieved=round(tflops, 3), peak=PEAK_FP64, unit="TFLOP/s",
                    frac=round(tflops / PEAK_FP64, 4), flop=flop, note=fnote),
                fp64_rerender_ms_device=round(ms_rr, 3), mixed_autotuned=tuned32,
                mixed_ms_device=round(ms32, 3), mixed_pixels_per_s=round(npix / (ms32 * 1e-3), 1),
                mixed_speedup_vs_plain=round(sum(plain_ms) / ms32, 3), mixed_max_dFtot_vs_plain=dF32)


def _pol_summary(S):
    sI, sQ, sU, sV = (float(S[..., k].sum()) for k in range(4))
    return sI, np.hypot(sQ, sU) / sI, sV / sI, 0.5 * np.arctan2(sU, sQ)


def precision_measure(args, oracle_stride=10):
    """The mixed-precision transfer on the 384x192x192 dump (a = 0.5, seed
    1), 400x400 px, thetacam 17, 230 GHz: image Ftot, LP, CP and EVPA of the
    FP32 transfer against the FP64 path, both against the oracle on an
    evenly strided pixel sample, and the device time of each."""
    from ipole_amd import ipgpu as G
    from ipole_amd import synth
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    t0 = time.time()
    h, prims = synth.make_dump(384, 192, 192, a=0.5, seed=1)
    log("precision: dump generated in %.1fs" % (time.time() - t0))
    ctx = G.Context(0)
    g, m = G.iharm_geom(h), G.iharm_model()
    ctx.set_geometry(g)
    ctx.set_model(m)
    ctx.upload_prims(prims, G.iharm_units(3e26, trat_large=40.))
    # one image on a dedicated device: a render-memory budget of 85% of it
    # holds the 400x400 image's records and packets at once (~237 GB; under
    # the library's default of 60% it renders in two chunks, ~10% slower)
    ctx.set_render_memory(int(0.85 * G.device_total_memory(0)))
    res, ms = {}, {}
    for fp32 in (0, 1):
        im, scale, _ = G.make_image(g, m.L_unit, nx=400, ny=400, thetacam=17., freqcgs=230e9, transfer_fp32=fp32)
        for _ in range(3):   # the first renders size and allocate the record and packet arenas
            ctx.render(im, fetch=False)
        t = []
        for _ in range(3):
            ctx.render(im, fetch=False)
            t.append(ctx.kernel_times()[2])
        res[fp32] = ctx.render(im)
        ms[fp32] = float(np.mean(t))
        chunks = ctx.render_info()[0]
    ctx.close()
    a, b = _pol_summary(res[0]["imageS"]), _pol_summary(res[1]["imageS"])
    err = dict(Ftot=abs(b[0] / a[0] - 1), LP=abs(b[1] - a[1]), CP=abs(b[2] - a[2]), EVPA_rad=abs(b[3] - a[3]))
    # oracle on a pixel sample
    orc, _ = O.iharm_oracle(h, prims, M_unit=3e26, trat_large=40.)
    p = O.default_params(nx=400, ny=400, thetacam=17., freqcgs=230e9)
    cam = orc.camera(p)
    pix = np.array([(i, j) for i in range(0, 400, oracle_stride) for j in range(0, 400, oracle_stride)], np.int32)
    nthreads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    t0 = time.time()
    rc, out7, ns, fl = orc.render_pixels(p, cam, pix, nthreads=nthreads)
    log("precision: oracle %d pixels in %.1fs" % (len(pix), time.time() - t0))
    f3 = 230e9 ** 3
    ref = np.column_stack([out7[:, 1], -out7[:, 2], -out7[:, 3], out7[:, 4]]) * f3
    den = np.maximum(np.abs(ref[:, 0]), 1e-6 * np.abs(ref[:, 0]).max())
    px = {}
    for fp32 in (0, 1):
        S = res[fp32]["imageS"][pix[:, 0], pix[:, 1], :4]
        px[fp32] = float((np.abs(S - ref) / den[:, None]).max())
    return dict(workload="384x192x192 FMKS dump a=0.5 seed 1, 400x400 px, thetacam 17, 230 GHz",
                fp64_ms_device=round(ms[0], 3), mixed_ms_device=round(ms[1], 3), chunks=int(chunks),
                render_memory="85% of the device",
                mixed_speedup=round(ms[0] / ms[1], 3), mixed_vs_fp64=err,
                oracle_sample_pixels=int(len(pix)), max_pixel_err_vs_oracle_fp64=px[0],
                max_pixel_err_vs_oracle_mixed=px[1])


def adjoint_measure(args):
    """The adjoint on the 96x64x64 dump (the 160x160 workload's generator),
    128x128 px, thetacam 163, 230 GHz: loss against a target rendered at
    perturbed parameters (M_unit x1.2, R_high x0.9, n_e scaled by a seeded
    lognormal field of 5%), device time of the adjoint against one forward
    render, and its dL/dM_unit against a central finite difference."""
    from ipole_amd import ipgpu as G
    from ipole_amd import synth
    h, prims = synth.make_dump(96, 64, 64, a=0.9375, seed=1)
    ctx = G.Context(0)
    g, m = G.iharm_geom(h), G.iharm_model()
    ctx.set_geometry(g)
    ctx.set_model(m)
    im, _, _ = G.make_image(g, m.L_unit, nx=128, ny=128, thetacam=163., freqcgs=230e9)
    M, RH = 3e26, 40.
    rng = np.random.default_rng(3)
    ctx.upload_prims(prims, G.iharm_units(1.2 * M, trat_large=0.9 * RH))
    ctx.scale_density(np.exp(0.05 * rng.normal(size=(96, 64, 64))))
    target = ctx.render(im)["imageS"]
    ctx.upload_prims(prims, G.iharm_units(M, trat_large=RH))
    ctx.render(im, fetch=False)
    t_fwd = []
    for _ in range(2):
        ctx.render(im, fetch=False)
        t_fwd.append(ctx.kernel_times()[2])
    r = ctx.adjoint(im, target)
    t_adj = []
    for _ in range(2):
        r = ctx.adjoint(im, target)
        t_adj.append(ctx.kernel_times()[2])

    def loss(Mu):
        ctx.upload_prims(prims, G.iharm_units(Mu, trat_large=RH))
        S = ctx.render(im)["imageS"]
        return float(np.sum((S[..., :4] - target[..., :4]) ** 2))
    hh = 1e-5
    fd = (loss(M * (1 + hh)) - loss(M * (1 - hh))) / (2 * hh * M)
    ctx.close()
    return dict(workload="96x64x64 FMKS dump seed 1, 128x128 px, thetacam 163, 230 GHz; target at M_unit x1.2, "
                         "R_high x0.9, 5% lognormal n_e scale",
                forward_ms_device=round(float(np.mean(t_fwd)), 3), adjoint_ms_device=round(float(np.mean(t_adj)), 3),
                adjoint_over_forward=round(float(np.mean(t_adj) / np.mean(t_fwd)), 3), loss=r["loss"],
                dL_dM_unit=r["dM_unit"], dL_dM_unit_fd=fd, rel_err_vs_fd=abs(r["dM_unit"] / fd - 1),
                dL_dR_high=r["dR_high"], dL_dR_low=r["dR_low"],
                density_grad_abs_max=float(np.abs(r["density"]).max()))


def adaptive_measure(ctx, im, full, scale, cfg, reps=3):
    """The same image refined adaptively from 20x20 with the reference's
    refinement test settings (tests/refinement/refinement.par: refine_rel
    1e-2, refine_abs 1e-6): wall time per render, rays traced, and the flux
    against the full render."""
    from ipole_amd import ipgpu as G
    rf = G.make_refine(cfg["nx"] // 8, refine_abs=1e-6, refine_rel=1e-2)
    ctx.render_adaptive(im, rf)
    t0 = time.perf_counter()
    for _ in range(reps):
        a = ctx.render_adaptive(im, rf)
    ms = 1e3 * (time.perf_counter() - t0) / reps
    F_full = full["imageS"][..., 0].sum() * scale
    F_a = a["imageS"][..., 0].sum() * scale
    out = dict(nx_min=cfg["nx"] // 8, refine_rel=1e-2, refine_abs=1e-6, ms_wall=round(ms, 3),
               levels=a["levels"], traced=[int(x) for x in a["traced"]], rays_traced=int(sum(a["traced"])),
               Ftot_rel_to_full=round(F_a / F_full - 1, 8))
    log("adaptive %dx%d from %dx%d: %.3f ms wall, %d rays traced %s, Ftot %+.2e vs the full render" %
        (cfg["nx"], cfg["ny"], cfg["nx"] // 8, cfg["ny"] // 8, ms, out["rays_traced"], out["traced"],
         out["Ftot_rel_to_full"]))
    return out


def bench_batch(args):
    """--workload batch: the batched figures as the bench line's value
    (FP64 batched pixels/s)."""
    r = batch_measure(args)
    out = dict(metric=METRIC, value=r["fp64_pixels_per_s"], unit="pixels/s", n_gpus=1, steps=args.steps,
               warmup=args.warmup, ms_per_step=r["fp64_ms_device"], higher_is_better=True, scaling="weak",
               vs_baseline=None, dtype="f64", data="synthetic",
               config=dict(r, workload="batched FP64: " + r["workload"]))
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched (64-job) figures")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the mixed-precision (400x400) and adjoint (128x128) figures")
    ap.add_argument("--cpu-stride", type=int, default=2,
                    help="CPU baseline sample: every n-th pixel row (all columns)")
    ap.add_argument("--small", action="store_true", help="small dump for quick checks")
    ap.add_argument("--gather-staging", action="store_true",
                    help="the 160x160 render's step-packet kernels read the dump cells staged in LDS "
                         "(ipgpu_set_gather_staging) instead of straight from the dump")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run one plain device-to-device copy of MIB MiB (the reference "
                         "kernel for scaling the FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--workload", choices=["c1", "batch", "extras", "adjoint", "precision"], default="c1",
                    help="c1: one 160x160 image (default); batch: the 64-job batch against the "
                         "same jobs rendered one by one; adjoint, precision: those figures alone; "
                         "extras: both")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this path is data-parallel over images, and no "
              "multi-GPU figure is produced (got --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copied %d MiB" % args.copy_calibration)
        return

    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "extras":
        print(json.dumps(dict(precision=precision_measure(args), adjoint=adjoint_measure(args))), flush=True)
        return
    if args.workload == "adjoint":
        print(json.dumps(dict(adjoint=adjoint_measure(args))), flush=True)
        return
    if args.workload == "precision":
        print(json.dumps(dict(precision=precision_measure(args))), flush=True)
        return

    from ipole_amd import ipgpu as G
    from ipole_amd import synth

    cfg = dict(C1)
    if args.small:
        cfg.update(N=(96, 64, 64), nx=64, ny=64)
    t0 = time.time()
    h, prims = synth.make_dump(*cfg["N"], a=cfg["a"], seed=cfg["seed"])
    log("dump generated in %.1fs" % (time.time() - t0))
    ctx = G.Context(0)
    g, m = G.iharm_geom(h), G.iharm_model()
    ctx.set_geometry(g)
    ctx.set_model(m)
    ctx.upload_prims(prims, G.iharm_units(cfg["M_unit"], trat_large=cfg["trat_large"]))
    im, scale, _ = G.make_image(g, m.L_unit, nx=cfg["nx"], ny=cfg["ny"], thetacam=cfg["thetacam"],
                                freqcgs=cfg["freqcgs"])
    npix = cfg["nx"] * cfg["ny"]
    if args.gather_staging:
        ctx.set_gather_staging(True)

    tuned = ctx.autotune_render(im)
    log("autotuned streamed render: %s (%.2f ms)" % (tuned, ctx.kernel_times()[2]))
    for w in range(args.warmup):
        ctx.render(im, fetch=False)
        log("warmup %d done: %.2f ms" % (w, ctx.kernel_times()[2]))
    ms_geo, ms_xfer, ms_tot, ms_stage = [], [], [], []
    t0 = time.perf_counter()
    for s in range(args.steps):
        ctx.render(im, fetch=False)   # synchronizes on the library's stream before returning
        a, b, c = ctx.kernel_times()
        ms_geo.append(a)
        ms_xfer.append(b)
        ms_tot.append(c)
        st_, streamed = ctx.stage_times_mode()
        ms_stage.append(st_)
    wall = time.perf_counter() - t0
    n_geo, n_rad = ctx.step_counts()
    res = ctx.render(im)
    # the one-pass render of the same image, for the log
    ctx.set_tuning(trace_budget=0)
    one = []
    for s in range(3):
        ctx.render(im, fetch=False)
        one.append(ctx.kernel_times()[2])
    log("one-pass render of the same image: %.3f ms" % np.mean(one[1:]))
    ctx.set_tuning(trace_budget=tuned["trace_budget"])
    adaptive = adaptive_measure(ctx, im, res, scale, cfg)
    st = G.image_stats(res, scale, cfg["freqcgs"], 16.9e6)
    ns_img = res["nstep"]
    tiles = ns_img[:ns_img.shape[0] // 8 * 8, :ns_img.shape[1] // 8 * 8].reshape(
        ns_img.shape[0] // 8, 8, ns_img.shape[1] // 8, 8).max(axis=(1, 3))
    log("nstep: max %d, mean %.1f, p99 %d; sum over 8x8 tiles of the tile max %d" %
        (ns_img.max(), ns_img.mean(), np.percentile(ns_img, 99), tiles.sum()))
    for T in sorted({1024, 2048, 4096, tuned["seg_min_steps"]}):
        sel = ns_img > T
        log("  rays > %d steps: %d, their steps %d (%.1f%%)" %
            (T, sel.sum(), ns_img[sel].sum(), 100. * ns_img[sel].sum() / ns_img.sum()))
    ctx.close()
    ms_step = 1e3 * wall / args.steps
    st_ms = np.mean(np.array(ms_stage), axis=0)
    if streamed:
        log("per step: wall %.3f ms, device %.3f ms (streamed: trace launches %.3f, span of the finished "
            "rays' transfer %.3f, tail after the last trace launch %.3f, k_finalize %.3f)" %
            ((ms_step, np.mean(ms_tot)) + tuple(st_ms)))
    else:
        log("per step: wall %.3f ms, device %.3f ms (k_trace %.3f, k_prep %.3f, k_chain %.3f, k_finalize %.3f)" %
            ((ms_step, np.mean(ms_tot)) + tuple(st_ms)))
    log("image: Ftot %.6g Jy, unpol %.6g Jy, LP %.3f%%, CP %.3f%%; geodesic steps %d, radiating %d" %
        (st["Ftot"], st["Ftot_unpol"], st["LP"], st["CP"], n_geo, n_rad))

    roof = roofline(np.mean(np.array(ms_stage), axis=0), n_geo, streamed)
    batched = precision = adjoint = None
    if not args.no_batch and not args.small:
        batched = batch_measure(args, h, prims)
    if not args.no_extras and not args.small:
        precision = precision_measure(args)
        adjoint = adjoint_measure(args)
    cpu = None
    if not args.no_cpu:
        nthreads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        cpu = cpu_baseline(h, prims, cfg, args.cpu_stride, nthreads)
    value = npix / (np.mean(ms_tot) * 1e-3)
    out = dict(metric=METRIC, value=round(value, 1), unit="pixels/s", n_gpus=1, steps=args.steps,
               warmup=args.warmup, ms_per_step=round(ms_step, 4), higher_is_better=True,
               scaling="weak", vs_baseline=None, dtype="f64", data="synthetic",
               config=dict(workload="plain FP64 GPU path: synthetic FMKS dump %dx%dx%d a=%g seed %d, "
                                    "%dx%d px, thetacam %g, %g GHz, R_high %g, M_unit %g" %
                                    (cfg["N"] + (cfg["a"], cfg["seed"], cfg["nx"], cfg["ny"],
                                                 cfg["thetacam"], cfg["freqcgs"] / 1e9,
                                                 cfg["trat_large"], cfg["M_unit"])),
                           pixels=npix, Ftot_Jy=round(st["Ftot"], 9), geodesic_steps=n_geo,
                           radiating_steps=n_rad, ms_device=round(float(np.mean(ms_tot)), 4),
                           ms_geodesic=round(float(np.mean(ms_geo)), 4),
                           ms_transfer=round(float(np.mean(ms_xfer)), 4),
                           streamed=bool(streamed), autotuned=tuned, gather_staging=bool(args.gather_staging),
                           ms_one_pass=round(float(np.mean(one[1:])), 4)),
               roofline=roof, cpu_baseline=cpu, batched=batched, mixed_precision=precision,
               adjoint=adjoint, adaptive=adaptive)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
