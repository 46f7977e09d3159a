#!/usr/bin/env python3
"""Benchmark: two-photon luminosity grid points/s on one MI355X.

Workload (one step): the full upcgen luminosity path for one configuration,
Pb-Pb 5.02 TeV, Woods-Saxon form-factor flux (FLUX_POINT 0), XNXN breakup
(BREAKUP_MODE 2), M 3.56-50 GeV x 1001, Y -6..6 x 121, z x 100, dimuons:
  tables (rho0, G_AA, 10^6-knot form factor, breakup)  -> UpcCrossSection::init
  flux rows + QAGS form-factor flux + L(M, Y) grid      -> prepareTwoPhotonLumi
  sigma(z, M), sigma(Y, M), total                        -> computeNuclXsection
  sampler CDFs (2-D (Y, M) and 1001 z rows)              -> UpcSampler1D/2D
value = 121,121 grid points / step time.  Inputs are configuration scalars:
everything the step reads lives in HBM before the timed region starts.

Prints one JSON line.  `--gpus` above 1 is not supported (one GPU per box).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NM, NY, NZ = 1001, 121, 100
POINTS = NM * NY
# Algorithmic work per (M, Y) point of the reference quadrature
# (src/UpcCrossSection.cpp:252-267): 2*5*120*120 + 2*120*120 + 2*120 FLOP.
W_ALG = 2 * 5 * 120 * 120 + 2 * 120 * 120 + 2 * 120
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0
FP32_MFMA_PEAK_TFLOPS = 157.3
# one MFMA tile step of the MFMA forms: 16 points x 16 columns x 32 deep
KSTEP_FLOP = 2 * 16 * 16 * 32
# FP64 operations of one form-factor flux integrand evaluation q^2 F(t)/t J1(bq/hc)
# (fluxFormIntegrand, src/UpcCrossSection.cpp:181-191) plus its QK21 accumulation,
# counted on the J1 branch x > 8 that most evaluations take (FMA = 2), as the
# device evaluates it:
#   t and the Q^2 clamp 2, cubic-spline lookup 11, q^2 F / t 3, x = b q / hc 2,
#   J1 modulus/phase form: 1/sqrt(x) and 1/x 9, two 13-term polynomials 48,
#   v and the Q scale 3, sin/cos with argument reduction 39, combination 6,
#   product 1, QK21 Gauss/Kronrod/abs sums 8
FLOP_PER_QAGS_EVAL = 132
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (one
# counter per pass) on this workload, scaled by a 1 GiB copy's readings; the
# commands and raw counts are in profiles/pmc_traffic_1001x121_xnxn.txt.
TRAFFIC_BYTES = {"k_flux_qags": 3.194e10, "k_lumi_fp64": 1.270e8, "k_mfma_lumi_split": 8.85e8}
HBM_PEAK_GBS = 8000.0


def workload_config(**over):
    import upcgen_amd as U
    kw = dict(Z=82, A=208, R=6.68, a=0.447, sqrts=5020.0, flux_point=0, breakup_mode=2,
              use_polarized=0, nm=NM, mmin=3.56, mmax=50.0, ny=NY, ymin=-6.0, ymax=6.0,
              nz=NZ, zmin=-1.0, zmax=1.0, proc_id=13, lep_a=0.0)
    kw.update(over)
    return U.Config.make(**kw)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline():
    """The CPU oracle (C restatement of the reference's OpenMP path, -O2 -fopenmp,
    static M-row split) on the host, with as many OpenMP threads as this process
    may use (its CPU affinity, capped by OMP_NUM_THREADS when that is set: the
    share of the host that one GPU's job gets).

    The timed sample of the workload is a 126 x 121 grid over the same M and Y
    ranges (15,246 of the 121,121 bins); the tables (form factor, G_AA, breakup)
    are built at their full size, and each phase is timed on its own:
    form-factor table, G_AA, breakup table, the L(M, Y) sample, sigma(z, M),
    sigma(Y, M) and the sampler CDFs.  Config 1 (101 x 41, point flux) is also
    timed on one thread and on all threads."""
    import numpy as np
    import oracle as O
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = nproc
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(affinity, int(omp)) if omp.isdigit() and int(omp) > 0 else affinity
    nm_s = 126
    o = O.Oracle(flux_point=0, breakup_mode=2, nm=nm_s, ny=NY, nz=NZ, proc_id=13, num_threads=threads)
    tt = o.table_times()
    t1 = time.perf_counter()
    g, nbad = o.lumi_grid()
    t_grid = time.perf_counter() - t1
    t2 = time.perf_counter()
    zm = o.cs_zm(0)
    t_zm = time.perf_counter() - t2
    t3 = time.perf_counter()
    cs, _, tot = o.nuc_cs_ym(lumi=np.nan_to_num(g, nan=0.0))
    t_cs = time.perf_counter() - t3
    t4 = time.perf_counter()
    O.pdf_init(cs.ravel())
    for r in range(zm.shape[0]):
        O.pdf_init(zm[r])
    t_cdf = time.perf_counter() - t4
    o.close()
    n = nm_s * NY
    cfg1 = {}
    for nt in (1, threads):
        o1 = O.Oracle(flux_point=1, breakup_mode=1, nm=101, ny=41, nz=50, proc_id=13, num_threads=nt)
        t5 = time.perf_counter()
        o1.lumi_grid()
        cfg1[nt] = 101 * 41 / (time.perf_counter() - t5)
        o1.close()
    return {"value": n / t_grid, "unit": "grid points/s", "cores": threads, "kind": "port",
            "sample": f"{nm_s}x{NY} grid over the same M, Y ranges ({n} of {POINTS} bins), OpenMP static "
                      f"M-row split on {threads} threads; L-grid phase {t_grid:.2f} s (the full grid would "
                      f"take about {t_grid * POINTS / n:.0f} s); tables built at full size",
            "nproc": nproc, "affinity_cores": affinity, "cpu_model": _cpu_model(),
            "phases_s": {"form_factor_table": round(tt["form_factor"], 4), "gaa_table": round(tt["gaa"], 4),
                         "breakup_table": round(tt["breakup"], 4), "lumi_grid_sample": round(t_grid, 3),
                         "cs_zm": round(t_zm, 4), "nuc_cs_ym_sample": round(t_cs, 4),
                         "cdf_sample": round(t_cdf, 4)},
            "config1_101x41_point_flux_pts_per_s": {"1_thread": round(cfg1[1], 1),
                                                     f"{threads}_threads": round(cfg1[threads], 1)},
            "qags_bad": nbad}


def batch_bench(seed, reps):
    """64 polarised jobs (form-factor flux) in one graph-captured launch
    sequence; the launch shapes come from the autotuner."""
    import torch
    import upcgen_amd as U
    jobs = U.mixed_systems_jobs(seed=seed)
    t0 = time.perf_counter()
    b = U.Batch(jobs)
    torch.cuda.synchronize()
    t_create = time.perf_counter() - t0
    tune = b.autotune()
    b.run(graph=True)  # capture + warm
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(reps):
        b.run(graph=True)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t1) * 1e3 / reps
    st = b.stats()
    res = {"workload": "64 jobs: Pb-Pb 5.02 TeV, Xe-Xe 5.44 TeV, O-O 7 TeV, Au-Au 200 GeV x BREAKUP_MODE 1-4, "
                       "USE_POLARIZED_CS 1, FLUX_POINT 0, nm in {101..2001}, ny in {41, 81, 121}",
           "seed": seed, "points": st.points, "rows": st.rows, "ms_per_batch": round(ms, 2),
           "points_per_s": round(st.points / ms * 1e3, 1), "replays": reps,
           "qags_calls": st.qags_calls, "qags_evals": st.qags_evals, "qags_bad": st.qags_bad,
           "tuned": {"qags_blocks_per_cu": tune.qags_blocks_per_cu, "lumi_waves": tune.lumi_waves,
                     "qags_ms": dict(zip(map(str, tune.qags_candidates), [round(x, 2) for x in tune.qags_ms])),
                     "lumi_ms": dict(zip(map(str, tune.lumi_candidates), [round(x, 2) for x in tune.lumi_ms]))},
           "create_s": round(t_create, 2)}
    # the same jobs with the MFMA contraction (FP64 flux rows, split-bf16 W.D.W^T)
    flat = lambda outs: [t.clone() for o in outs for t in (o if isinstance(o, tuple) else (o,))]
    ref = flat(b.outputs)
    b.set_form(U.FORM_MFMA)
    tune = b.autotune()
    b.run(graph=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(reps):
        b.run(graph=True)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t1) * 1e3 / reps
    got = flat(b.outputs)
    # bins whose flux integrals stopped at a form-factor domain error are NaN in both forms
    nan_match = all(bool((torch.isnan(g) == torch.isnan(r)).all()) for g, r in zip(got, ref))
    err = max(float(((g - r).abs() / r.abs())[torch.isfinite(r)].max()) for g, r in zip(got, ref))
    res["mfma"] = {"ms_per_batch": round(ms, 2), "points_per_s": round(st.points / ms * 1e3, 1),
                   "fp64_guard_items": b.stats().mfma_fp64_items,
                   "max_rel_err_vs_fp64": err, "nan_bins": int(sum(int(torch.isnan(r).sum()) for r in ref)),
                   "nan_bins_match": nan_match, "mfma_tile_group": tune.mfma_tile_group,
                   "mfma_ms": dict(zip(map(str, tune.mfma_candidates), [round(x, 2) for x in tune.mfma_ms]))}
    b.close()
    return res


def other_configs():
    """The small plain-method grid (Pb-Pb 5.02 TeV, point flux, mode 1, 101x41, PROC_ID 13) and
    end-to-end ditau events (PROC_ID 15, XNXN, form-factor flux, sigma(M,Y,z) 1001x121x100
    from the MFMA luminosity, 10^8 Philox draws, seed 12345): GPU times beside the headline."""
    import torch
    import upcgen_amd as U
    out = {}
    c1 = U.Context(flux_point=1, breakup_mode=1, nm=101, ny=41, nz=50, proc_id=13)
    L1 = c1.lumi_grid()
    torch.cuda.synchronize()
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        c1.lumi_grid(out=L1)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    out["config1_101x41_point_flux"] = {"ms_per_grid": round(ms, 3), "grid_points_per_s": round(101 * 41 / ms * 1e3, 1)}
    c1.close()
    c5 = U.Context(flux_point=0, breakup_mode=2, nm=1001, ny=121, nz=100, proc_id=15)
    t0 = time.perf_counter()
    L5 = torch.nan_to_num(c5.lumi_grid(U.FORM_MFMA), nan=0.0)
    cs, _, _ = c5.nuc_cs_ym(L5)
    sum_ym = c5.pdf_init(cs.reshape(1, -1))[0].contiguous()
    sum_z = c5.pdf_init(c5.cs_zm(0)).contiguous()
    torch.cuda.synchronize()
    t_tables = time.perf_counter() - t0
    n_total, chunk = 100_000_000, 25_000_000
    c5.sample_events(sum_ym, sum_z=sum_z, n=1000, seed=12345)  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for off in range(0, n_total, chunk):
        c5.sample_events(sum_ym, sum_z=sum_z, n=chunk, seed=12345, offset=off)
    torch.cuda.synchronize()
    t_draw = time.perf_counter() - t0
    c5.generate_pairs(sum_ym, sum_z=sum_z, n=1000, seed=12345)  # warm
    torch.cuda.synchronize()
    n_pairs = 20_000_000
    t0 = time.perf_counter()
    for off in range(0, n_pairs, 10_000_000):
        c5.generate_pairs(sum_ym, sum_z=sum_z, n=10_000_000, seed=12345, offset=off)
    torch.cuda.synchronize()
    t_pairs = time.perf_counter() - t0
    out["config5_ditau_events"] = {
        "tables_s": round(t_tables, 3), "draws": n_total, "draw_s": round(t_draw, 4),
        "draws_per_s": round(n_total / t_draw, 1),
        "events_with_kinematics": n_pairs, "kinematics_s": round(t_pairs, 4),
        "events_with_kinematics_per_s": round(n_pairs / t_pairs, 1),
        "note": "(Y, M, z) draws by inverse CDF over the FP64 prefix sums; the kinematics add the "
                "lepton 4-momenta of each pair (UpcGenerator::generateEvent with NON_ZERO_GAM_PT 0 "
                "and no cuts)"}
    c5.close()
    return out


def surrogate_form(ctx, ref, steps=20000, fit_points=32768, fit_steps=15000, save_path=None, load_path=None):
    """The learned surrogate as a grid form.  The general network is trained
    on 2^20 FP64 quadrature points inside its domain (256 random systems x
    4096, `steps` Adam steps of 8192 rows, split-bf16 products) and checked
    on 2^16 held-out points of 32 new systems; then fine-tuned on
    `fit_points` quadrature points of this system and served through the
    FP64 anchor lattice with a 1e-3 bound.  Training happens before the
    timed grids; its costs are reported beside them.  load_path: a
    system-fitted surrogate saved by an earlier run (--surrogate-out) is
    served instead, without training (for profiling the grid form)."""
    import numpy as np
    import torch
    import upcgen_amd as U
    from upcgen_amd import surrogate as S
    if load_path:
        return surrogate_grid(ctx, ref, S.Surrogate.load(load_path), None)
    t0 = time.perf_counter()
    X, r, _, _ = S.make_dataset(256, 4096, seed=11)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    mu, sig = float(r.mean()), float(r.std())
    s = S.Surrogate(11)
    s.norm = (mu, sig)
    t1 = time.perf_counter()
    loss = s.train(X.contiguous(), ((r - mu) / sig).float().contiguous(), steps=steps, batch=8192, lr=2e-3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    Xc, _, lgc, L0c = S.make_dataset(0, 2048, seed=31, systems=S.sample_systems(16, 33))
    s.calibrate(Xc.contiguous(), L0c, lgc)
    Xh, _, lgh, L0h = S.make_dataset(0, 2048, seed=16, systems=S.sample_systems(32, 18))
    pred, spread = s.log_lumi(Xh.contiguous(), L0h)
    err = (pred - lgh).abs().cpu().numpy()
    served_h = (s.err_scale * spread).cpu().numpy() <= 1e-3
    f, t_label, t_fit = s.fit_system(ctx.config, n=fit_points, steps=fit_steps, lr=5e-4)
    if save_path:
        f.save(save_path)
    training = {"rows": int(X.shape[0]), "steps": steps, "batch": 8192, "data_s": round(t_data, 2),
                "train_s": round(t_train, 2),
                "heldout_rows": int(Xh.shape[0]),
                "heldout_abs_dlogL_median": float(np.median(err)),
                "heldout_abs_dlogL_p99": float(np.quantile(err, 0.99)),
                "heldout_spread_gate_served_frac": float(served_h.mean()),
                "system_fit_points": fit_points, "system_fit_steps": fit_steps,
                "system_fit_label_s": round(t_label, 2), "system_fit_train_s": round(t_fit, 2)}
    return surrogate_grid(ctx, ref, f, training)


def surrogate_grid(ctx, ref, f, training):
    """Times the UPC_FORM_SURROGATE grid served by the system-fitted
    surrogate f and checks every bin against the FP64 grid ref."""
    import numpy as np
    import torch
    import upcgen_amd as U
    from upcgen_amd import surrogate as S
    cfg = ctx.config
    f.attach(ctx, max_err=1e-3)
    L = torch.empty_like(ref)
    ctx.lumi_grid(U.FORM_SURROGATE, out=L)  # warm
    torch.cuda.synchronize()
    reps = 5
    t2 = time.perf_counter()
    for _ in range(reps):
        ctx.lumi_grid(U.FORM_SURROGATE, out=L)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t2) * 1e3 / reps
    ctx.set_timing(True)
    ctx.timings(reset=True)
    for _ in range(reps):
        ctx.lumi_grid(U.FORM_SURROGATE, out=L)
    ph = ctx.timings(reset=True)
    ctx.set_timing(False)
    inf = ctx.info
    ok = ~torch.isnan(ref)
    assert torch.equal(ok, ~torch.isnan(L)), "surrogate"
    e = ((L - ref).abs() / ref.abs())[ok].cpu().numpy()
    served = e > 1e-12
    M = np.linspace(cfg.mmin, cfg.mmax, cfg.nm, endpoint=False)
    Y = np.linspace(cfg.ymin, cfg.ymax, cfg.ny, endpoint=False)
    inside = int((S.xi_max(cfg, M[:, None], Y[None, :]) <= S.XI_MAX).sum())
    return {"ms_per_grid": round(ms, 3), "grid_points_per_s": round(POINTS / ms * 1e3, 1),
            "qags_ms": round(ph["qags"][0] / reps, 3),
            "contraction_ms": round((ph["lumi"][0] + ph["project"][0]) / reps, 3),
            "max_rel_err_vs_fp64": float(e.max()),
            "median_rel_err_vs_fp64": float(np.median(e)),
            "served_bins": inf.surrogate_points, "fallback_bins": inf.surrogate_fallback,
            "bins_inside_domain": inside,
            "served_median_rel_err": float(np.median(e[served])) if served.any() else None,
            "served_frac_within_1e-3": float(np.mean(e[served] <= 1e-3)) if served.any() else None,
            "nan_bins": int((~ok).sum()),
            "gate": "FP64 anchors every 16 M x 4 Y bins; served where the network is within 5e-4 at the 4 "
                    "anchors around the bin; outside the domain (xi > 4) and elsewhere the FP64 grid of "
                    "those bins",
            "training": training if training else "loaded (no training in this run)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--form", default="fp64", choices=["fp64", "mfma", "mfma_fp32"],
                    help="the form of the timed steps (every form's grid is timed beside them)")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip timing the other forms' grids and their error against the FP64 form")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch", action="store_true",
                    help="also time the 64-job mixed-systems batch (Pb/Xe/O/Au x breakup modes 1-4, "
                         "polarised, uneven grids) as one graph-captured launch sequence")
    ap.add_argument("--batch-seed", type=int, default=2024)
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the surrogate form (training it takes about a minute)")
    ap.add_argument("--surrogate-out", default=None, help="save the system-fitted surrogate (.npz)")
    ap.add_argument("--surrogate-in", default=None,
                    help="serve a saved system-fitted surrogate (.npz from --surrogate-out) instead of training")
    ap.add_argument("--no-other-configs", action="store_true",
                    help="skip the 101x41 point-flux grid and the 10^8 ditau event draws")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run three 1 GiB device-to-device copies (scale factor for the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the luminosity grid "
              "shards by M rows with no exchange, but multi-GPU runs are not part of this build",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import upcgen_amd as U

    if args.calibrate_copy:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read_per_copy": src.numel() * 8,
                          "bytes_written_per_copy": dst.numel() * 8}))
        return

    forms = {"fp64": U.FORM_FP64, "mfma": U.FORM_MFMA, "mfma_fp32": U.FORM_MFMA_FP32,
             "surrogate": U.FORM_SURROGATE}
    form = forms[args.form]
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    ctx = U.Context(workload_config())
    L = torch.empty((NM, NY), dtype=torch.float64, device="cuda")

    def step():
        ctx.rebuild_tables()
        ctx.lumi_grid(form, out=L)
        # the four bins whose form-factor integrals leave the form-factor table
        # (NaN; the reference process stops there) are emptied, as the
        # UpcCrossSection facade does before the cross sections
        torch.nan_to_num_(L, nan=0.0)
        zm = ctx.cs_zm(0)
        cs, _, tot = ctx.nuc_cs_ym(L)
        ctx.pdf_init(cs.reshape(1, -1))
        ctx.pdf_init(zm)
        return tot

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctx.set_timing(True)
    ctx.timings(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tot = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ph = ctx.timings(reset=True)
    ctx.set_timing(False)
    inf = ctx.info

    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = POINTS / (ms_per_step / 1e3)
    per_launch = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in ph.items()}
    dominant = max(("qags", "lumi", "tables", "flux", "project"), key=lambda k: ph[k][0])
    lumi_ms = per_launch["lumi"]
    # The dominant kernel: the QAGS form-factor flux integrals (the same in every
    # form).  Its algorithmic work is the per-point figure W_ALG times the points
    # of one launch (one grid), over its average launch time, against the FP64
    # vector peak.  W_ALG counts the reference's node contraction only, so this
    # fraction is small by construction; the QAGS integrand's own FP64 work is
    # reported beside it as an estimate (FLOP_PER_QAGS_EVAL per evaluation).
    qags_ms = per_launch["qags"]
    alg_flop = W_ALG * POINTS
    alg_tf = alg_flop / (qags_ms * 1e-3) / 1e12 if qags_ms > 0 else 0.0
    qags_flop = inf.qags_evals * FLOP_PER_QAGS_EVAL
    qags_tf = qags_flop / (qags_ms * 1e-3) / 1e12 if qags_ms > 0 else 0.0
    roof_qags = {"bound": "fp64 valu",
                 "kernel": "k_flux_qags (QAGS form-factor flux integrals, one launch per grid, FP64 VALU)",
                 "achieved": round(alg_tf, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                 "frac": round(alg_tf / FP64_PEAK_TFLOPS, 5), "traffic": TRAFFIC_BYTES["k_flux_qags"],
                 "traffic_unit": "bytes per launch (rocprofv3 --pmc, copy-calibrated)",
                 "algorithmic_flop_per_launch": alg_flop, "algorithmic_flop_per_point": W_ALG,
                 "points_per_launch": POINTS, "avg_launch_ms": round(qags_ms, 4),
                 "qags_integrand_estimate": {
                     "integrand_evals_per_launch": inf.qags_evals, "flop_per_eval": FLOP_PER_QAGS_EVAL,
                     "flop_per_launch": qags_flop, "achieved_tflops": round(qags_tf, 3),
                     "frac_of_fp64_peak": round(qags_tf / FP64_PEAK_TFLOPS, 5)}}
    if form == U.FORM_FP64:
        # algorithmic FLOP of the reference quadrature per launch
        flop = W_ALG * POINTS
        tflops = flop / (lumi_ms * 1e-3) / 1e12 if lumi_ms > 0 else 0.0
        roof = {"bound": "fp64 valu", "kernel": "k_lumi_fp64 (luminosity contraction, FP64 VALU)",
                "achieved": round(tflops, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(tflops / FP64_PEAK_TFLOPS, 5), "traffic": TRAFFIC_BYTES["k_lumi_fp64"],
                "algorithmic_flop_per_launch": flop, "avg_launch_ms": round(lumi_ms, 4)}
    else:
        # MFMA FLOP issued per launch: tile steps x 16x16x32 x products per step
        # (split-bf16: hi*hi + hi*lo + lo*hi; FP32: one) for each luminosity (1 or 2)
        # The banded GEMM does 4e10 MFMA FLOP against ~0.6 GB of operands, so
        # it is bound by HBM: its algorithmic bytes are every projected flux
        # row read once (4 bytes a column: bf16 hi + lo, or FP32) and the band
        # of D; the MFMA rate is reported beside them.
        split = form == U.FORM_MFMA
        nprod = 3 if split else 1
        flop = inf.mfma_ksteps * KSTEP_FLOP * nprod
        peak = BF16_PEAK_TFLOPS if split else FP32_MFMA_PEAK_TFLOPS
        tflops = flop / (lumi_ms * 1e-3) / 1e12 if lumi_ms > 0 else 0.0
        h = inf.mfma_spacing
        cext = -math.cos(math.pi * 0.9739065285171717)  # the unpolarised grid's extreme node cosine
        nu = int(math.floor(20.0 / math.sqrt(1.0 - cext * cext) * (1.0 + 1e-12) / h)) + 4
        nu_pad = (nu + 31) // 32 * 32
        band = int(math.ceil(20.0 / h)) + 1
        rows = POINTS + NM  # k1 rows of every bin + k2 rows of the iy = 0 bins (mirror sharing)
        alg_bytes = rows * nu_pad * 4 + nu_pad * (2 * band - 1) * 4
        gbs = alg_bytes / (lumi_ms * 1e-3) / 1e9 if lumi_ms > 0 else 0.0
        roof = {"bound": "hbm", "kernel": f"k_mfma_lumi ({'split-bf16' if split else 'fp32'} banded GEMM)",
                "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 5), "algorithmic_bytes_per_launch": alg_bytes,
                "traffic": TRAFFIC_BYTES["k_mfma_lumi_split"] if split else None,
                "mfma_tflops": round(tflops, 3), "mfma_peak_tflops": peak, "mfma_frac": round(tflops / peak, 5),
                "mfma_flop_per_launch": flop, "ksteps_per_launch": inf.mfma_ksteps,
                "avg_launch_ms": round(lumi_ms, 4), "shared_grid_spacing_fm": h}
    compare = None
    if not args.no_compare:
        compare = {}
        ref = torch.empty_like(L)
        ctx.lumi_grid(U.FORM_FP64, out=ref)
        for name, f in forms.items():
            if f == U.FORM_SURROGATE:
                continue
            ctx.lumi_grid(f, out=L)  # warm
            torch.cuda.synchronize()
            reps = 5
            t2 = time.perf_counter()
            for _ in range(reps):
                ctx.lumi_grid(f, out=L)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t2) * 1e3 / reps
            # the same grids' phases (HIP events): the forms share the QAGS flux rows
            ctx.set_timing(True)
            ctx.timings(reset=True)
            for _ in range(reps):
                ctx.lumi_grid(f, out=L)
            ph_f = ctx.timings(reset=True)
            ctx.set_timing(False)
            ok = ~torch.isnan(ref)
            assert torch.equal(ok, ~torch.isnan(L)), name
            err = ((L - ref).abs() / ref.abs())[ok]
            compare[name] = {"ms_per_grid": round(ms, 3), "grid_points_per_s": round(POINTS / ms * 1e3, 1),
                             "qags_ms": round(ph_f["qags"][0] / reps, 3),
                             "contraction_ms": round((ph_f["lumi"][0] + ph_f["project"][0]) / reps, 3),
                             "max_rel_err_vs_fp64": float(err.max()),
                             "median_rel_err_vs_fp64": float(err.median()),
                             "nan_bins": int((~ok).sum())}
            if f in (U.FORM_MFMA, U.FORM_MFMA_FP32):
                compare[name]["fp64_guard_items"] = ctx.info.mfma_fp64_items
        if not args.no_surrogate:
            compare["surrogate"] = surrogate_form(ctx, ref, save_path=args.surrogate_out,
                                                  load_path=args.surrogate_in)
        for name, v in compare.items():
            v["speedup_vs_fp64_grid"] = round(compare["fp64"]["ms_per_grid"] / v["ms_per_grid"], 3)
    out = {
        "metric": "lumi grid pts/s at 1001x121 XNXN: FP64 vs MFMA vs surrogate at stated rel err",
        "value": round(value, 1),
        "unit": "grid points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (configuration scalars only; the grid is deterministic)",
        "config": {"workload": "Pb-Pb 5.02 TeV, Woods-Saxon form-factor flux, XNXN (mode 2), "
                               "1001x121 (M 3.56-50 GeV, Y -6..6), z x 100, PROC_ID 13",
                   "form": args.form, "grid_points": POINTS, "parallelism": "single GPU"},
        "roofline": roof_qags,
        "roofline_contraction": roof,
        "phases_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in ph.items()},
        "dominant_phase": dominant,
        "forms": compare,
        "qags": {"calls": inf.qags_calls, "integrand_evals": inf.qags_evals, "bad": inf.qags_bad},
        "tot_cs_mb": tot,
    }
    if not args.no_other_configs:
        out["other_configs"] = other_configs()
    if args.batch:
        out["batch"] = batch_bench(args.batch_seed, max(1, min(args.steps, 3)))
    if not args.no_cpu_baseline:
        cb = cpu_baseline()
        out["cpu_baseline"] = {k: v for k, v in cb.items() if k != "qags_bad"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
