"""Counts the algorithmic FLOPs per unit of each pipeline kernel with the
op-counting CPU build (tests/kernel_emu/opcount.cpp) on the 160x160 FMKS
workload (synthetic 288x128x128 dump, a = 0.9375, seed 1, thetacam 163,
230 GHz, R_high 40, M_unit 3e26), over an evenly strided pixel sample, and
writes ipole_amd/flops_per_step.json, which bench.py reads for the
roofline's achieved FLOP/s.

    python tests/kernel_emu/count_flops.py [--stride 8]
"""
import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

from ipole_amd import ipgpu as G  # noqa: E402
from ipole_amd import synth  # noqa: E402
import emu  # noqa: E402
import oracle as O  # noqa: E402


def count(geom, model, image, dump, pix):
    subprocess.run(["make", "-s", "-C", HERE, "libopcount.so"], check=True)
    L = C.CDLL(os.path.join(HERE, "libopcount.so"))
    L.opcount_render.argtypes = [C.POINTER(G.Geom), C.POINTER(G.Model), C.POINTER(G.Image), G.dp, G.ip,
                                 C.c_int, G.dp, G.dp, G.dp]
    out = np.zeros(12)
    pix = np.ascontiguousarray(pix, np.int32)
    per_ray = np.zeros((len(pix), 3))
    per_ray_multi = np.zeros((len(pix), 3))
    rc = L.opcount_render(C.byref(geom), C.byref(model), C.byref(image),
                          G.P(dump) if dump is not None else None, G.IP(pix), len(pix), G.P(out), G.P(per_ray),
                          G.P(per_ray_multi))
    assert rc == 0
    return out, per_ray, per_ray_multi


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--stride", type=int, default=8)
    args = ap.parse_args()
    t0 = time.time()
    h, prims = synth.make_dump(288, 128, 128, a=0.9375, seed=1)
    orc, fields = O.iharm_oracle(h, prims, M_unit=3e26, trat_large=40.)
    dump = emu.pack_dump(fields)
    g, m = G.iharm_geom(h), G.iharm_model()
    pix = np.array([(i, j) for i in range(0, 160, args.stride) for j in range(0, 160, args.stride)])
    res = {}
    for unpol in (0, 1):
        im, _, _ = G.make_image(g, m.L_unit, nx=160, ny=160, thetacam=163., freqcgs=230e9,
                                only_unpolarized=unpol)
        o, per_ray, per_ray_multi = count(g, m, im, dump, pix)
        n_geo, n_step, n_rad = o[0], o[1], o[2]
        # the fused transfer's flops per ray as a radiating and a non-radiating
        # step cost plus a per-ray cost (far-end coefficients, projection),
        # least squares over the rays
        A = np.column_stack([per_ray[:, 1], per_ray[:, 0] - per_ray[:, 1], np.ones(len(per_ray))])
        (f_rad, f_non, f_ray), *_ = np.linalg.lstsq(A, per_ray[:, 2], rcond=None)
        fit_err = float(np.abs(A @ [f_rad, f_non, f_ray] - per_ray[:, 2]).sum() / per_ray[:, 2].sum())
        # the multi-job transfer: per ray, flops of 1, 2, 3 jobs = shared + n x per job
        jm = (per_ray_multi[:, 2] - per_ray_multi[:, 0]) / 2.
        gm = per_ray_multi[:, 0] - jm
        (g_rad, g_non, g_ray), *_ = np.linalg.lstsq(A, gm, rcond=None)
        (j_rad, j_non, j_ray), *_ = np.linalg.lstsq(A, jm, rcond=None)
        lin_err = float(np.abs(per_ray_multi[:, 1] - (gm + 2. * jm)).sum() / per_ray_multi[:, 1].sum())
        res["unpol" if unpol else "pol"] = dict(
            geodesic_steps=int(n_geo), transfer_steps=int(n_step), radiating_steps=int(n_rad), rays=int(o[8]),
            F_trace_per_step=o[3] / n_geo,
            F_prep_per_radiating_step=o[4] / n_rad,
            F_prep_per_nonradiating_step=o[5] / max(n_step - n_rad, 1),
            F_prep_per_step=(o[4] + o[5]) / n_step,
            F_chain_per_step=(o[6] + o[7]) / n_step,
            F_fused_per_step=o[9] / n_step,
            F_fused_per_radiating_step=f_rad, F_fused_per_nonradiating_step=f_non, F_fused_per_ray=f_ray,
            F_fused_fit_rel_err=fit_err,
            F_seg_per_column_step=o[10] / (17 * n_step), F_seg_intensity_per_step=o[11] / n_step,
            F_multi_shared_per_radiating_step=g_rad, F_multi_shared_per_nonradiating_step=g_non,
            F_multi_shared_per_ray=g_ray, F_multi_job_per_radiating_step=j_rad,
            F_multi_job_per_nonradiating_step=j_non, F_multi_job_per_ray=j_ray,
            F_multi_linearity_rel_err=lin_err,
        )
        print(("unpol" if unpol else "pol"), json.dumps(res["unpol" if unpol else "pol"]), flush=True)
    out = dict(
        source="op-counting CPU build of the kernel code (tests/kernel_emu/opcount.cpp) over %d pixels "
               "(every %dth row and column) of the 160x160 FMKS workload; +,-,*,/,sqrt = 1, fma = 2, "
               "each transcendental (exp, log, sin, cos, pow, acos, cbrt, tanh) = 20"
               % (len(pix), args.stride),
        command="python tests/kernel_emu/count_flops.py --stride %d" % args.stride,
        units=dict(F_trace_per_step="FLOP per geodesic step of k_trace (stepsize + midpoint push)",
                   F_prep_per_step="FLOP per transfer step of k_prep (mix of radiating and not)",
                   F_chain_per_step="FLOP per transfer step of k_chain, including its per-ray "
                                    "far-end coefficients and projection",
                   F_fused_per_step="FLOP per transfer step of the FP64 fused transfer of the batch "
                                    "(k_xfer_batch / k_xfer_queue), far-end coefficients included",
                   F_fused_per_radiating_step="the fused transfer's FLOP per ray fitted (least squares over "
                                              "the rays) as radiating steps x this + non-radiating steps x "
                                              "F_fused_per_nonradiating_step + F_fused_per_ray; "
                                              "F_fused_fit_rel_err is the fit's summed |residual| over the "
                                              "summed count",
                   F_seg_per_column_step="FLOP per step of one of the 17 N-map columns of k_seg "
                                         "(segment_column), 384-step segments",
                   F_seg_intensity_per_step="FLOP per step of k_seg's unpolarized column (segment_intensity)",
                   F_multi_shared_per_radiating_step="the multi-job fused transfer (k_xfer_multi) of n jobs of a "
                                                     "ray costs the shared part (per radiating step, "
                                                     "non-radiating step and ray) plus n times the per-job part "
                                                     "(F_multi_job_*), fitted from 1 and 3 jobs; "
                                                     "F_multi_linearity_rel_err checks 2 jobs against the fit"),
        **res)
    path = os.path.join(ROOT, "ipole_amd", "flops_per_step.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote %s in %.0fs" % (path, time.time() - t0))


if __name__ == "__main__":
    main()
