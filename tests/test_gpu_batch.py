"""Batched render (ipgpu_render_batch) against the oracle: jobs with
different cameras, image sizes, frequencies, M_unit and polarization modes
on one dump, in one call, at the plain path's parity bar.  Jobs that share a
camera reuse one geodesic trace at another frequency (K and dl rescaled);
M_unit enters as n_e and B scale factors.  Also: several passes over the
rays when the trajectory memory is small, and the maxnstep error."""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import gpu_image_kwargs, small_dump, stokes_errors
from ipole_amd import ipgpu as G
from test_gpu_parity import check, spread

pytestmark = pytest.mark.gpu

M_UNIT = 3e26
JOBS = [  # thetacam, nx, ny, freq, M_unit scale, unpol, emission_type
    (163., 24, 24, 230e9, 1.0, 0, 4),
    (163., 24, 24, 86e9, 1.0, 0, 4),
    (163., 24, 24, 345e9, 0.3, 0, 4),
    (60., 16, 24, 230e9, 2.0, 0, 4),
    (60., 16, 24, 230e9, 1.0, 1, 4),
    (90., 20, 20, 86e9, 0.5, 0, 1),
]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup():
    h, prims = small_dump()
    c = G.Context(0)
    g, m = G.iharm_geom(h), G.iharm_model()
    c.set_geometry(g)
    c.set_model(m)
    c.upload_prims(prims, G.iharm_units(M_UNIT))
    yield h, prims, c, g, m
    c.close()


def oracle_ref(h, prims, th, nx, ny, f, s, unpol, et):
    orc, _ = O.iharm_oracle(h, prims, M_unit=M_UNIT * s)
    p = O.default_params(nx=nx, ny=ny, thetacam=th, freqcgs=f, only_unpolarized=unpol, emission_type=et)
    return orc.render(p), p, (None if unpol else spread(h, prims, M_UNIT * s, p))


@pytest.mark.parametrize("traj_mib", [48 * 1024, 2])
def test_batch_vs_oracle(setup, traj_mib):
    h, prims, c, g, m = setup
    say("case batch of %d jobs, trajectory memory %d MiB" % (len(JOBS), traj_mib))
    c.set_batch_memory(traj_mib << 20)
    jobs, scales, refs, floors = [], [], [], []
    for (th, nx, ny, f, s, unpol, et) in JOBS:
        ref, p, fl = oracle_ref(h, prims, th, nx, ny, f, s, unpol, et)
        im, scale, _ = G.make_image(g, m.L_unit, **gpu_image_kwargs(p))
        jobs.append((im, s))
        scales.append(scale)
        refs.append(ref)
        floors.append(fl)
    got = c.render_batch(jobs)
    c.set_batch_memory(48 << 30)
    for k, (ref, out, scale, fl) in enumerate(zip(refs, got, scales, floors)):
        say("  job %d %s vs oracle, floor %s" % (k, JOBS[k], fl))
        check(ref, out, scale, unpol=bool(JOBS[k][5]), floor=fl)
    # against the plain path on the same job (dump derived at the job's
    # M_unit): the same kernels' arithmetic up to the rescaling roundoff
    for k, ((im, s), out) in enumerate(zip(jobs, got)):
        c.upload_prims(prims, G.iharm_units(M_UNIT * s))
        plain = c.render(im)
        assert np.array_equal(plain["nstep"], out["nstep"])
        e = stokes_errors(plain["imageS"], plain["image"], out["imageS"], out["image"])
        F = abs(out["image"].sum() / plain["image"].sum() - 1)
        say("  job %d vs plain: dF %.2e %s" % (k, F, e))
        assert F < 1e-9
        for name, (mx, frac) in e.items():
            if JOBS[k][5] and name != "unpol":
                continue
            assert mx < 1e-6, (k, name, mx)
    c.upload_prims(prims, G.iharm_units(M_UNIT))


def test_batch_chunks_identical(setup):
    """With little trajectory memory the rays are counted, then traced and
    transferred in chunks; with much, their records stay where the first
    pass put them: the images are the same."""
    h, prims, c, g, m = setup
    say("case batch chunking: 1 MiB, 3 MiB, 48 GiB of trajectory memory")
    jobs = []
    for (th, nx, ny, f, s, unpol, et) in JOBS:
        p = O.default_params(nx=nx, ny=ny, thetacam=th, freqcgs=f, only_unpolarized=unpol, emission_type=et)
        im, _, _ = G.make_image(g, m.L_unit, **gpu_image_kwargs(p))
        jobs.append((im, s))
    outs = []
    for mib in (1, 3, 48 * 1024):
        c.set_batch_memory(mib << 20)
        outs.append(c.render_batch(jobs))
    c.set_batch_memory(48 << 30)
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            for k in ("image", "imageS", "taus", "nstep", "flag"):
                if a.get(k) is not None:
                    assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("arena_mib", [8192, 1, 0])
@pytest.mark.parametrize("slot", [64, 301])
def test_batch_slot_records(setup, slot, arena_mib, monkeypatch):
    """The first trace pass stores each ray's first `slot` records in a slot
    of its own and, while the arena lasts, the rest in chunks of an arena;
    the rays longer than the slot get full-length regions, filled from the
    slot and their chunks, or, for the rays that could not get all their
    chunks (a 1 MiB arena holds 16; none when it is off), by continuing the
    ray on lane pairs from its last slot record.  The images equal those of
    the count-then-store passes (a small trajectory memory forces them)."""
    h, prims, c, g, m = setup
    say("case batch slot records, slot %d, arena %d MiB" % (slot, arena_mib))
    jobs = []
    for (th, nx, ny, f, s, unpol, et) in JOBS:
        p = O.default_params(nx=nx, ny=ny, thetacam=th, freqcgs=f, only_unpolarized=unpol, emission_type=et)
        im, _, _ = G.make_image(g, m.L_unit, **gpu_image_kwargs(p))
        jobs.append((im, s))
    c.set_batch_memory(1 << 20)
    ref = c.render_batch(jobs)
    steps_ref = c.step_counts()[0]
    c.set_batch_memory(48 << 30)
    monkeypatch.setenv("IPGPU_BATCH_SLOT", str(slot))
    monkeypatch.setenv("IPGPU_BATCH_ARENA_MIB", str(arena_mib))
    c2 = G.Context(0)
    try:
        c2.set_geometry(g)
        c2.set_model(m)
        c2.upload_prims(prims, G.iharm_units(M_UNIT))
        got = c2.render_batch(jobs)
        assert max(int(o["nstep"].max()) for o in got) >= slot      # some rays continued
        for a, b in zip(ref, got):
            for k in ("image", "imageS", "taus", "nstep", "flag"):
                if a.get(k) is not None:
                    assert np.array_equal(a[k], b[k]), k
        assert c2.step_counts()[0] == steps_ref
    finally:
        c2.close()


def test_batch_maxnstep(setup):
    h, prims, c, g, m = setup
    say("case batch maxnstep")
    ok, _, _ = G.make_image(g, m.L_unit, nx=8, ny=8, thetacam=163., freqcgs=230e9)
    bad, _, _ = G.make_image(g, m.L_unit, nx=8, ny=8, thetacam=60., freqcgs=230e9, maxnstep=40)
    with pytest.raises(G.IpgpuError) as e:
        c.render_batch([(ok, 1.), (bad, 1.)])
    assert e.value.code == G.ERR_MAXNSTEP and "job 1" in str(e.value)


def batch_jobs_of(g, m, scales=None, fp32=0):
    jobs = []
    for k, (th, nx, ny, f, s, unpol, et) in enumerate(JOBS):
        p = O.default_params(nx=nx, ny=ny, thetacam=th, freqcgs=f, only_unpolarized=unpol, emission_type=et)
        kw = gpu_image_kwargs(p)
        kw["transfer_fp32"] = fp32
        im, _, _ = G.make_image(g, m.L_unit, **kw)
        jobs.append((im, s if scales is None else scales[k]))
    return jobs


def same_outputs(a, b, tag):
    for x, y in zip(a, b):
        for k in ("image", "imageS", "taus", "nstep", "flag"):
            assert np.array_equal(x[k], y[k]), (tag, k)


def test_batch_queue_launch(setup):
    """The transfer as a persistent work-queue launch (waves take chunks of
    64 items from an atomic ticket) gives the images of the one-wave-per-64
    launch, for any number of waves."""
    h, prims, c, g, m = setup
    jobs = batch_jobs_of(g, m)
    try:
        c.set_batch_options(queue=0)
        ref = c.render_batch(jobs)
        for waves in (1, 7, 0):
            say("case batch queue launch, %d waves" % waves)
            c.set_batch_options(queue=1, queue_waves=waves)
            same_outputs(ref, c.render_batch(jobs), waves)
    finally:
        c.set_batch_options(queue=1, queue_waves=0)


def test_rerender_batch(setup):
    """ipgpu_rerender_batch: the last batch's records reused with new M_unit
    scales equal a fresh batch at those scales, bit for bit, eagerly and from
    the captured graph; the graph is replayed several times with other work
    on the stream and host synchronisations between the replays, and every
    replay is checked.  A new dump of the same grid is allowed; a render that
    reuses the ray buffers ends the resident batch."""
    h, prims, c, g, m = setup
    sa = [1.0, 1.0, 0.3, 2.0, 1.0, 0.5]
    sb = [0.7, 1.9, 0.45, 1.1, 3.0, 0.25]
    ja, jb = batch_jobs_of(g, m, sa), batch_jobs_of(g, m, sb)
    say("case rerender: fresh batches at two scale sets")
    fresh_b = c.render_batch(jb)
    fresh_a = c.render_batch(ja)
    say("case rerender: eager")
    c.set_batch_options(rerender_graph=0)
    same_outputs(fresh_b, c.rerender_batch(jb), "eager b")
    same_outputs(fresh_a, c.rerender_batch(ja), "eager a")
    c.set_batch_options(rerender_graph=1)
    say("case rerender: graph replays")
    for k, (jobs, ref) in enumerate([(jb, fresh_b), (ja, fresh_a), (jb, fresh_b), (jb, fresh_b), (ja, fresh_a)]):
        same_outputs(ref, c.rerender_batch(jobs), "replay %d" % k)
        c.download_fields()                        # other work on the stream between the replays
        assert c.step_counts()[0] == 0
    say("case rerender: new dump of the same grid")
    c.upload_prims(prims, G.iharm_units(2 * M_UNIT))
    half = c.rerender_batch(batch_jobs_of(g, m, [s / 2 for s in sa]))
    c.upload_prims(prims, G.iharm_units(M_UNIT))
    for k, (x, y) in enumerate(zip(fresh_a, half)):
        assert np.array_equal(x["nstep"], y["nstep"])
        assert abs(y["image"].sum() / x["image"].sum() - 1) < 1e-9, k
        e = stokes_errors(x["imageS"], x["image"], y["imageS"], y["image"])
        for name, (mx, frac) in e.items():
            if JOBS[k][5] and name != "unpol":
                continue
            assert mx < 1e-6, (k, name, mx)
    say("case rerender: a different image, a plain render in between")
    other = list(ja)
    im0, _, _ = G.make_image(g, m.L_unit, nx=8, ny=8, thetacam=30., freqcgs=230e9)
    other[0] = (im0, 1.)
    with pytest.raises(G.IpgpuError) as e:
        c.rerender_batch(other)
    assert e.value.code == G.ERR_ARG
    c.render(im0)
    with pytest.raises(G.IpgpuError) as e:
        c.rerender_batch(ja)
    assert e.value.code == G.ERR_STATE
    say("case rerender: a batch in passes is not resident")
    c.set_batch_memory(1 << 20)
    c.render_batch(ja)
    with pytest.raises(G.IpgpuError) as e:
        c.rerender_batch(ja)
    assert e.value.code == G.ERR_STATE
    c.set_batch_memory(48 << 30)


def test_rerender_batch_mixed(setup):
    """The mixed-precision batch gathers an FP32 copy of the cell records,
    written by a kernel before each transfer launch, also inside the
    rerender's captured graph: replays equal fresh mixed batches bit for bit,
    and a replay after a new dump of the same grid reads the new dump (twice
    the M_unit at half the scales gives the same images to the FP32 rounding
    of the records; a stale copy would give half the density)."""
    h, prims, c, g, m = setup
    sa = [1.0, 1.0, 0.3, 2.0, 1.0, 0.5]
    sb = [0.7, 1.9, 0.45, 1.1, 3.0, 0.25]
    ja, jb = batch_jobs_of(g, m, sa, fp32=1), batch_jobs_of(g, m, sb, fp32=1)
    say("case mixed rerender: fresh batches at two scale sets")
    fresh_b = c.render_batch(jb)
    fresh_a = c.render_batch(ja)
    c.set_batch_options(rerender_graph=1)
    say("case mixed rerender: graph replays")
    for k, (jobs, ref) in enumerate([(jb, fresh_b), (ja, fresh_a), (jb, fresh_b), (ja, fresh_a)]):
        same_outputs(ref, c.rerender_batch(jobs), "mixed replay %d" % k)
        c.download_fields()
    say("case mixed rerender: new dump of the same grid")
    c.upload_prims(prims, G.iharm_units(2 * M_UNIT))
    try:
        half = c.rerender_batch(batch_jobs_of(g, m, [s / 2 for s in sa], fp32=1))
    finally:
        c.upload_prims(prims, G.iharm_units(M_UNIT))
    for k, (x, y) in enumerate(zip(fresh_a, half)):
        assert np.array_equal(x["nstep"], y["nstep"])
        assert abs(y["image"].sum() / x["image"].sum() - 1) < 1e-5, k
        if not JOBS[k][5]:
            assert abs(y["imageS"][..., 0].sum() / x["imageS"][..., 0].sum() - 1) < 1e-5, k


def test_autotune_batch(setup):
    """ipgpu_autotune_batch picks a slot and a transfer launch by timing them,
    keeps them, and leaves the batch resident: the images are those of the
    default options."""
    h, prims, c, g, m = setup
    say("case batch autotune")
    jobs = batch_jobs_of(g, m)
    ref = c.render_batch(jobs)
    try:
        ch = c.autotune_batch(jobs)
        say("  chosen %s" % ch)
        assert ch["slot"] in (1024, 2048, 4096) and ch["queue"] in (0, 1)
        assert ch["resident_waves_per_cu"] >= 1
        same_outputs(ref, c.rerender_batch(jobs), "resident after tuning")
        same_outputs(ref, c.render_batch(jobs), "tuned options")
    finally:
        c.set_batch_options(slot=2048, queue=1, queue_waves=0)
