"""Benchmark of the k-space time-stepping path on one MI355X.

Workload (default): the large synthetic Wannier model -- 32 bands, 961 lattice
vectors (|R1|,|R2| <= 15, hexagonal, a = 2.5 A, exp(-|R|/3 A) decay, seed 4004),
512x512 k-grid, 8 filled bands, two gaussian pumps along z (0.4 eV at 0.4 V/A and
0.8 eV at 0.1 V/A, 20 fs intensity FWHM), RK4 with dt = 0.04 a.u.  A step is one
RK4 step (4 fused RHS+stage launches) over all k, with the state and every
operator resident in HBM before the timed region.

The line carries the forms of the stage kernel: the plain FP64 path (the
drop-in, `value`), the mixed form (FP32 operators and products, FP64 state;
`mixed`) and the mixed form with FP32 stage storage (`mixed32`: FP32 stage
inputs and histories, FP64 state at t_n and accumulator), each mixed form with
its J(t), final conduction population and harmonic-intensity errors against the
FP64 path over a window of steps at the pulse peak.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload ...] [--forms fp64,mixed,mixed32,bf16x3] [--no-cpu]

Prints one JSON line.
"""
import argparse
import json
import os
import platform
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_MATRIX_PEAK_TFLOPS = 78.6
FP32_MATRIX_PEAK_TFLOPS = 157.3
BF16_MATRIX_PEAK_TFLOPS = 2500.0  # dense bf16 matrix peak (no sparsity)
METRIC = "k-point·steps/s, FP64 vs mixed/batched, at stated J(t) error; HBM GB/s, MFMA %"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["wannier32", "wannier", "hbn", "hbn_coulomb", "ab5", "batch"], default="wannier32",
                    help="wannier32: 32 bands, |R| <= 15, 512x512, two pumps (default); wannier: 16 bands, |R| <= 10, "
                         "200x200; hbn: 3-band hBN tight binding, 60x60, T2 = 10 fs (the CPU-runnable case); "
                         "hbn_coulomb: 3-band hBN with the mean field, 300x300, AB4; ab5: the reference's own AB5 "
                         "example run end to end through the drop-in (all 16591 steps, observables and files), the "
                         "workload of its one published time; batch: the batched scan of independent instances")
    ap.add_argument("--nk", type=int, default=0, help="grid points per direction (0: the workload's own)")
    ap.add_argument("--nw", type=int, default=16)
    ap.add_argument("--rmax", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="fp64,mixed,mixed32,bf16x3",
                    help="stage-kernel forms to time, the first one is `value`: fp64 (the drop-in), mixed (FP32 "
                         "operators and products, FP64 state), mixed32 (mixed plus FP32 stage inputs and histories, "
                         "FP64 master state and accumulator), bf16x3 (mixed with BF16x3 split products); the mixed "
                         "forms need 5 bands or more, bf16x3 9 or more")
    ap.add_argument("--precision", choices=["fp64", "mixed", "mixed32", "bf16x3"], default=None,
                    help="time one form only (same as --forms <form>)")
    ap.add_argument("--stepper", choices=["rk4", "taylor4"], default="rk4",
                    help="wannier workloads: classic RK4 (four stage launches a step) or the Taylor-4 multistep "
                         "integrator (one launch a step)")
    ap.add_argument("--acc-steps", type=int, default=200, help="steps of the mixed-vs-FP64 J(t) comparison")
    ap.add_argument("--contraction", choices=["fp64", "fp32"], default="fp64",
                    help="wannier workloads: Wannier contraction on the FP64 or the FP32 matrix cores (the J(t) "
                         "error of the mixed form is then taken against an FP64-contracted setup)")
    ap.add_argument("--instances", type=int, default=256, help="batch workload: number of instances")
    ap.add_argument("--end-to-end", action="store_true",
                    help="time the drop-in's own loop (observables and output files included) for --steps steps")
    ap.add_argument("--cpu-steps", type=int, default=0, help="oracle steps of the CPU sample (0: auto)")
    ap.add_argument("--cpu-subgrid", type=int, default=128,
                    help="grid points per direction of the CPU sample when the workload's grid is above 200x200")
    ap.add_argument("--no-probe", action="store_true", help="skip the streaming-rate probe")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in launch schedule instead of timing the candidates (edus_autotune)")
    ap.add_argument("--graph", action="store_true",
                    help="RK4 steps from a device-resident field table through the captured step graph "
                         "(edus_set_field_table + edus_step_rk4_graph) instead of per-call stage launches")
    ap.add_argument("--workdir", default=None)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the library's streaming probe on arrays of MIB MiB (its copy kernel is the "
                         "known-size kernel of the HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: runs on one GPU only (the k-grid of one problem lives in one HBM); "
              "sharding the grid in k-slabs over several GPUs is not built.", file=sys.stderr)
        sys.exit(2)
    if args.precision:
        args.forms = args.precision

    if args.copy_calibration:
        from edus_amd import native as N
        r = N.stream_probe(args.copy_calibration << 20, 5)
        print(json.dumps({"probe_mib": args.copy_calibration, "gbs": r}))
        return
    if args.workload == "ab5":
        bench_ab5()
        return
    if args.workload == "batch":
        bench_batch(args)
        return
    bench_single(args)


def workload_input(args, wd, nk_override=None):
    """Input file and name of a single-problem workload."""
    from edus_amd import workloads

    wstep = "RK4" if args.stepper == "rk4" else "Taylor"
    nsteps_total = args.steps + args.warmup
    if args.workload == "wannier":
        nkd = nk_override or args.nk or 200
        inp = workloads.wannier_synthetic(wd, nw=args.nw, rmax=args.rmax, nk=nkd, nsteps=max(4000, nsteps_total),
                                          dt=0.05, stepper=wstep, e_v_ang=0.5, nvalence=4, seed=1001, decay_ang=2.0)
        return inp, f"wannier_nw{args.nw}_R{2 * args.rmax + 1}x{2 * args.rmax + 1}_{nkd}x{nkd}_{args.stepper}_dt0.05"
    if args.workload == "wannier32":
        nkd = nk_override or args.nk or 512
        inp = workloads.wannier_synthetic(wd, nw=32, rmax=15, nk=nkd, nsteps=max(5000, nsteps_total), dt=0.04,
                                          stepper=wstep, e_v_ang=0.4, nvalence=8, seed=4004, decay_ang=3.0,
                                          photon_ev=0.4, fwhm_fs=20.0, extra_pumps=((0.8, 0.1),))
        return inp, f"wannier_nw32_R31x31_{nkd}x{nkd}_2pumps_{args.stepper}_dt0.04"
    if args.workload == "hbn_coulomb":
        nkd = nk_override or args.nk or 300
        inp = workloads.hbn_coulomb(wd, os.path.join(ROOT, "tests", "golden", "coulomb"), "G_num502657_Ncut20_", nk=nkd,
                                    nsteps=max(6000, nsteps_total), dt=0.03, order=4)
        return inp, f"hbn_coulomb_haldane_corebn_{nkd}x{nkd}_ncut20_ab4_dt0.03"
    nkd = nk_override or args.nk or 60
    inp = workloads.hbn_tb(wd, nk=nkd, nsteps=max(8000, nsteps_total), dt=0.05, stepper="RK4")
    return inp, f"hbn_gen2d_3band_{nkd}x{nkd}_T2_10fs_rk4_dt0.05"


def stage_model(info, nd, form):
    """Bytes and matrix-core flops one stage launch must move / do per k in the
    engine's own layout: packed Hermitian state (8 nb^2 B in FP64), packed H and
    the nd dipole components the field needs (8 nb^2 B each in FP64, 4 nb^2 B in
    the mixed form), U full complex (16 / 8 nb^2 B) when dephasing is on.  State
    transfers per launch: RK4 3, 5, 5, 3 (mean 4); Adams-Bashforth / Taylor of
    order s: s + 2 (read y_n, write f_n, read s - 1 old f, write y_n+1), +1 for
    the Coulomb pass.  Flops: one complex NBT^3 product for the commutator
    (NBT the padded tile, 16 or 32 on the matrix cores), four more with
    dephasing.  Also the full-complex figure S (4 + n_op) per RK4 launch
    (S = 16 nb^2, every operator full complex, all three dipole components),
    reported as survey_bytes_*."""
    from edus_amd import native as N

    nb = info["nb"]
    rk4 = info["stepper"] == N.STEPPER_RK4
    opw = 8 if form == "fp64" else 4
    diss = info["n_diss"] > 0
    if form == "mixed32":
        # FP32 stage inputs / histories (4 nb^2 B), FP64 state at t_n and accumulator (8 nb^2 B):
        # RK4 stages move 20, 32, 32, 20 nb^2 B of state; order-s multistep 16 + 4 s
        state_k = (26 if rk4 else 16 + 4 * info["order"]) * nb * nb
    else:
        transfers = 4 if rk4 else info["order"] + 2 + (1 if info["ncut"] else 0)
        state_k = transfers * 8 * nb * nb
    per_k = state_k + (1 + nd) * opw * nb * nb + (2 * opw * nb * nb if diss else 0)
    nbt = nb if nb <= 4 else 8 if nb <= 8 else 16 if nb <= 16 else 32
    flop_k = (8 * nbt ** 3) * (5 if diss else 1) if nbt >= 16 else 0
    S = 16 * nb * nb
    n_op = 4 + (1 if diss else 0)
    survey_k = (4 + n_op) * S if rk4 else (info["order"] + 2 + n_op + (1 if info["ncut"] else 0)) * S
    kern = (f"k_stage_wave<{nbt}>" if nb > 4 else f"k_stage_small<{nb}>")
    stepper = "RK4" if rk4 else (f"Taylor{info['order']}" if info["stepper"] == N.STEPPER_TAYLOR else f"AB{info['order']}")
    return {"bytes_k": per_k, "flop_k": flop_k, "survey_bytes_k": survey_k, "kernel": kern, "stepper": stepper}


def time_form(s, form, f, g, args):
    """Engine of one precision form: warm-up, then args.steps timed steps (wall
    clock around a synchronised call; kernel time from the engine's HIP events)."""
    from edus_amd import native as N

    prec = {"fp64": N.PRECISION_FP64, "mixed": N.PRECISION_MIXED, "mixed32": N.PRECISION_MIXED32,
            "bf16x3": N.PRECISION_BF16X3}[form]
    t0 = time.perf_counter()
    eng = N.Engine(s.problem(), precision=prec)
    t_create = time.perf_counter() - t0
    schedule = None
    if not getattr(args, "no_autotune", False):  # the timed steps' own field, before any step
        schedule = eng.autotune(f[args.warmup, 0], bool(g[args.warmup]), reps=3)
    if getattr(args, "graph", False) and eng.stepper == N.STEPPER_RK4:
        eng.set_field_table(f, g)
        if args.warmup:
            eng.step_graph(args.warmup)
        t0 = time.perf_counter()
        eng.step_graph(args.steps)
        wall = time.perf_counter() - t0
    else:
        if args.warmup:
            eng.step(f[:args.warmup], g[:args.warmup])
        t0 = time.perf_counter()
        eng.step(f[args.warmup:], g[args.warmup:])
        wall = time.perf_counter() - t0
    dev_ms, nlaunch = eng.last_step_time()
    return eng, {"wall_s": wall, "dev_ms": dev_ms, "nlaunch": nlaunch, "create_s": t_create, "schedule": schedule}


def bench_single(args):
    from edus_amd import native as N

    wd = args.workdir or tempfile.mkdtemp(prefix="edus_bench_")
    forms = [x.strip() for x in args.forms.split(",") if x.strip()]
    if args.stepper != "rk4" and args.workload not in ("wannier", "wannier32"):
        sys.exit("--stepper taylor4 applies to the wannier workloads")
    if args.contraction == "fp32" and args.workload not in ("wannier", "wannier32"):
        sys.exit("--contraction fp32 applies to the wannier workloads")
    t_setup = time.perf_counter()
    inp, wname = workload_input(args, wd)
    s = N.Setup(inp, fp32_contraction=args.contraction == "fp32")
    info = s.info
    t_setup = time.perf_counter() - t_setup
    nk, nb = info["nk"], info["nb"]
    if nb < 5:
        forms = [x for x in forms if x == "fp64"] or ["fp64"]
    elif nb <= 8:  # the BF16x3 products run on the matrix-core tiles of nb >= 9
        forms = [x for x in forms if x != "bf16x3"] or ["fp64"]
    if args.end_to_end:
        s.run(os.path.join(wd, "warm"), max_steps=args.warmup)
        t0 = time.perf_counter()
        n = s.run(os.path.join(wd, "run"), max_steps=args.steps)
        el = time.perf_counter() - t0
        print(json.dumps({
            "metric": METRIC, "value": nk * n / el, "unit": "k-point·steps/s", "n_gpus": 1, "steps": n,
            "warmup": args.warmup, "ms_per_step": el * 1e3 / n, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": wname, "nk": nk, "bands": nb, "it_resolution": info["it_resolution"],
                       "includes": "the drop-in's time loop: steps, observables and output files from t0",
                       "seconds": round(el, 3)}}))
        return

    probe = None
    if not args.no_probe:
        probe = N.stream_probe(1 << 30, 5)

    # Field schedule of the steps around the pulse peak, computed before timing.
    nsteps_total = args.steps + args.warmup
    it0 = max(0, (info["itfi"] + 1) // 2 - nsteps_total // 2)
    f, g = s.fields(it0, nsteps_total)
    ft = f[args.warmup:]
    nd = sum(1 for x in range(3) if np.any(ft[..., x] != 0) or np.any(ft[..., 3 + x] != 0))

    legs, engines = {}, {}
    for form in forms:
        eng, t = time_form(s, form, f, g, args)
        engines[form] = eng
        m = stage_model(info, nd, form)
        avg_launch_s = t["dev_ms"] * 1e-3 / t["nlaunch"]
        bytes_launch = m["bytes_k"] * nk
        achieved = bytes_launch / avg_launch_s / 1e9
        if achieved > HBM_PEAK_GBS:
            raise SystemExit(f"bench.py: {form} stage kernel at {achieved:.0f} GB/s above the HBM peak: the byte model is wrong")
        peak_mx = {"fp64": FP64_MATRIX_PEAK_TFLOPS, "bf16x3": BF16_MATRIX_PEAK_TFLOPS}.get(form, FP32_MATRIX_PEAK_TFLOPS)
        # BF16x3: three bf16 matrix products per FP32 one
        tflops = m["flop_k"] * (3 if form == "bf16x3" else 1) * nk / avg_launch_s / 1e12
        legs[form] = {
            "value": nk * args.steps / t["wall_s"], "ms_per_step": t["wall_s"] * 1e3 / args.steps,
            "device_ms_per_step": t["dev_ms"] / args.steps, "avg_launch_us": round(avg_launch_s * 1e6, 3),
            "launches_per_step": t["nlaunch"] // args.steps, "engine_create_s": round(t["create_s"], 2),
            "schedule": (None if t["schedule"] is None else
                         {("strip_width" if nb > 4 else "workgroup_threads"): t["schedule"], "picked_by": "edus_autotune"}),
            "kernel": m["kernel"] + (" (RK4 stages)" if m["stepper"] == "RK4" else f" ({m['stepper']} steps)")
                      + {"fp64": "", "mixed": ", FP32 operators and products",
                         "mixed32": ", FP32 operators, products and stage storage",
                         "bf16x3": ", FP32 operators, BF16x3 split products"}[form],
            "bytes_per_launch": bytes_launch, "achieved_gbs": round(achieved, 2),
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "mfma": ({"flop_per_launch": m["flop_k"] * (3 if form == "bf16x3" else 1) * nk, "achieved_tflops": round(tflops, 2),
                      "peak_tflops": peak_mx, "frac": round(tflops / peak_mx, 4)} if m["flop_k"] else None),
            "survey_bytes_per_launch": m["survey_bytes_k"] * nk,
        }
    head = forms[0]
    lead = legs[head]
    traffic, tsrc = None, None
    suffix = "" if head == "fp64" else "_mixed"
    tfile = os.path.join("profiles", f"traffic_stage_wave_nb{nb}_nk{nk}{suffix}.json")
    if os.path.exists(os.path.join(ROOT, tfile)):
        traffic = json.load(open(os.path.join(ROOT, tfile))).get("bytes_per_launch")
        tsrc = tfile
    roof = {"bound": "hbm", "achieved": lead["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": lead["frac"],
            "traffic": traffic, "traffic_source": tsrc, "kernel": lead["kernel"],
            "bytes_per_launch": lead["bytes_per_launch"], "bytes_model": "packed Hermitian state and operators, "
            f"{nd} dipole component(s) read, mean state transfers per launch", "avg_launch_us": lead["avg_launch_us"],
            "mfma": lead["mfma"],
            "survey_bytes_per_launch": lead["survey_bytes_per_launch"],
            "survey_rate_gbs": round(lead["survey_bytes_per_launch"] / (lead["avg_launch_us"] * 1e-6) / 1e9, 2)}
    if probe:
        roof["stream_probe_gbs"] = {k: round(v, 1) for k, v in probe.items()}
        roof["frac_of_copy"] = round(lead["achieved_gbs"] / probe["copy"], 4)

    mixed_legs = {}
    ref_setup = N.Setup(inp) if args.contraction == "fp32" and any(f != "fp64" for f in legs) else None
    if info["stepper"] != N.STEPPER_RK4:
        # multistep error legs start fresh engines (zero histories): the timed
        # ones are released first, so that two engines, not six, hold HBM
        for e in engines.values():
            e.close()
        engines = {}
    for form in legs:
        if form == "fp64":
            continue
        leg = dict(legs[form])
        if "fp64" in legs:
            leg["speedup_vs_fp64"] = round(legs[form]["value"] / legs["fp64"]["value"], 4)
        prec_of = {"mixed": N.PRECISION_MIXED, "mixed32": N.PRECISION_MIXED32, "bf16x3": N.PRECISION_BF16X3}[form]
        leg["error_vs_fp64"] = (mixed_j_error(s, engines.get(form), it0, args.acc_steps, ref=engines.get("fp64"),
                                              ref_setup=ref_setup, eng_precision=prec_of) if args.acc_steps > 0 else None)
        mixed_legs[form] = leg
    for e in engines.values():
        e.close()

    cpu = None
    rk4 = info["stepper"] == N.STEPPER_RK4
    if not args.no_cpu and rk4 and not info["ncut"]:  # the CPU sample runs the plain RK4 oracle
        if nk > 200 * 200:
            sub_inp, _ = workload_input(args, os.path.join(wd, "cpu_subgrid"), nk_override=args.cpu_subgrid)
            cpu = cpu_baseline(N.Setup(sub_inp), args.cpu_steps, subgrid=True)
        else:
            cpu = cpu_baseline(s, args.cpu_steps)

    out = {
        "metric": METRIC,
        "value": lead["value"],
        "unit": "k-point·steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": lead["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f64/f32",
        "forms": {"fp64": "the drop-in: FP64 throughout",
                  "mixed": "FP32 operators and commutator/dephasing products on the matrix cores, FP64 state",
                  "mixed32": "mixed plus FP32 stage inputs and derivative histories; FP64 state at t_n, FP64 "
                             "accumulator and updates",
                  "bf16x3": "mixed with the products as BF16x3 split products (lo*hi + hi*lo + hi*hi) on the bf16 "
                            "matrix cores, FP32 accumulation"},
        "data": "synthetic",
        "config": {"workload": wname, "nk": nk, "bands": nb, "stepper": stepper_name(info),
                   "form": head, "launch": "captured step graph" if args.graph else "stream launches per call",
                   "wannier_contraction": args.contraction if args.workload in ("wannier", "wannier32") else None,
                   "state_layout": "packed-hermitian", "setup_s": round(t_setup, 2),
                   "device_ms_per_step": lead["device_ms_per_step"], "schedule": lead["schedule"]},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    for form, leg in mixed_legs.items():
        if form == head:
            out["j_error_vs_fp64"] = leg["error_vs_fp64"]
        else:
            out[form] = leg
    print(json.dumps(out))


def stepper_name(info):
    from edus_amd import native as N

    if info["stepper"] == N.STEPPER_RK4:
        return "RK4"
    return f"Taylor{info['order']}" if info["stepper"] == N.STEPPER_TAYLOR else f"AB{info['order']}"


def currents(ob):
    """J(t) per Cartesian component: Re J1 + J2 (J2 as the output file writes it)."""
    return ob["j1"][0::2] - 2.0 * ob["j2"][1::2]


def mixed_j_error(s, eng, it0, nsteps, every=2, ref=None, ref_setup=None, eng_precision=None):
    """The mixed form's error against the FP64 drop-in path over nsteps from rho0
    in the pulse-peak window, both engines given the same fields: J(t) (every
    `every` steps), the final conduction population, and the harmonic spectrum
    |FFT(J w)|^2 (Hann window w) of the J(t) series.  ref: an FP64 engine of the
    same setup (made here if None); ref_setup: the FP64 path's own tables (an
    FP64 Wannier contraction) when `s` was built otherwise; both start from
    their own rho0."""
    from edus_amd import native as N

    info = s.info
    rs = ref_setup or s
    own = ref is None or ref_setup is not None
    multistep = info["stepper"] != N.STEPPER_RK4
    fresh = None
    if multistep:
        # Adams-Bashforth / Taylor carry a derivative history: both runs start from
        # zero histories (as the reference's run does), not from the timed steps'
        fresh = N.Engine(s.problem(), precision=eng_precision)
        eng = fresh
        own = True
    if own:
        ref = N.Engine(rs.problem(), precision=N.PRECISION_FP64)
    P0 = s.array("rho0", True).reshape(-1, info["nb"], info["nb"])
    P0r = rs.array("rho0", True).reshape(-1, info["nb"], info["nb"])
    f, g = s.fields(it0, nsteps)
    jm, jr = [], []
    eng.set_state(P0)
    ref.set_state(P0r)
    obs = {}
    for a in range(0, nsteps, every):
        b = min(a + every, nsteps)
        for e, acc in ((eng, jm), (ref, jr)):
            e.step(f[a:b], g[a:b])
            o = e.observables()
            acc.append(currents(o))
            obs[id(e)] = o
    jm, jr = np.array(jm), np.array(jr)
    scale = max(float(np.max(np.abs(jr))), 1e-300)
    # conduction population: the bands above the filled ones, summed
    ncc = info["ncc"]
    cm, cr = float(np.sum(obs[id(eng)]["pop"][ncc:])), float(np.sum(obs[id(ref)]["pop"][ncc:]))
    # harmonic intensities of each Cartesian component, relative to the largest
    win = np.hanning(len(jr))[:, None]
    im = np.abs(np.fft.rfft(jm * win, axis=0)) ** 2
    ir = np.abs(np.fft.rfft(jr * win, axis=0)) ** 2
    pm, pr = eng.get_state(), ref.get_state()
    if own:
        ref.close()
    if fresh is not None:
        fresh.close()
    return {"J_rel_err": float(np.max(np.abs(jm - jr)) / scale), "J_max": scale,
            "conduction_population_rel_err": abs(cm - cr) / max(abs(cr), 1e-300), "conduction_population": cr,
            "harmonic_intensity_rel_err": float(np.max(np.abs(im - ir)) / max(float(np.max(ir)), 1e-300)),
            "rho_rel_err": float(np.max(np.abs(pm - pr)) / np.max(np.abs(pr))),
            "steps": nsteps, "observed_every": every, "reference": "FP64 path, same steps and fields"}


# The reference's published run (example AB5: 70x70 k, 3 bands, AB5, 16591 steps,
# observables every 5 steps): 787 s wall on 1 MPI rank x 7 OpenMP threads.
REF_AB5_RATE = 4900 * 16591 / 787.0


def bench_ab5():
    """The drop-in end to end on the reference's AB5 input: setup excluded, the time
    loop with its observables and output files timed, as the reference's log times it."""
    from edus_amd import native as N

    src = os.path.join(ROOT, "tests", "golden", "ab5", "input.txt")
    wd = tempfile.mkdtemp(prefix="edus_ab5_")
    s = N.Setup(src)
    info = s.info
    s.run(wd, max_steps=200)  # warm-up: first launches, file creation
    t0 = time.perf_counter()
    n = s.run(wd)
    el = time.perf_counter() - t0
    value = info["nk"] * n / el
    print(json.dumps({
        "metric": "k-point\u00b7steps/s, FP64 vs mixed/batched, at stated J(t) error; HBM GB/s, MFMA %",
        "value": value, "unit": "k-point\u00b7steps/s", "n_gpus": 1, "steps": n, "warmup": 200,
        "ms_per_step": el * 1e3 / n, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": value / REF_AB5_RATE, "dtype": "f64", "data": "the reference's AB5 example input",
        "config": {"workload": "ab5_haldane_corebn_70x70_ab5_dt0.1_16591steps", "nk": info["nk"], "bands": info["nb"],
                   "stepper": "AB5", "it_resolution": info["it_resolution"], "includes": "observables and output files",
                   "reference_seconds": 787.0, "seconds": round(el, 3)}}))


def bench_batch(args):
    """The batched scan: independent RK4 instances of uneven grids and band counts,
    each with its own pulse, advanced by one launch per stage and kernel class.
    The same instances stepped one context at a time are timed beside it."""
    from edus_amd import native as N
    from edus_amd import workloads

    wd = args.workdir or tempfile.mkdtemp(prefix="edus_batch_")
    t_setup = time.perf_counter()
    paths = workloads.batch_scan(wd, n=args.instances, nsteps=max(3000, args.steps + args.warmup + 10))
    setups = []
    for i, p in enumerate(paths):
        setups.append(N.Setup(p))
        if i % 16 == 15 or i == len(paths) - 1:
            print(f"setup {i + 1}/{len(paths)} instances, {time.perf_counter() - t_setup:.1f} s", file=sys.stderr, flush=True)
    prec = N.PRECISION_MIXED if args.precision == "mixed" else None
    problems = [s.problem() for s in setups]
    if prec is not None:  # mixed precision applies to the wave-kernel members
        for p, s in zip(problems, setups):
            if s.info["nb"] >= 5:
                p.precision = prec
    batch = N.Batch(problems)
    t_setup = time.perf_counter() - t_setup
    nk = np.array([s.info["nk"] for s in setups])
    nb = np.array([s.info["nb"] for s in setups])
    n_op = np.array([4 + (1 if s.info["n_diss"] > 0 else 0) for s in setups])
    total = int(args.warmup + args.steps)
    it0 = 0
    fg = [s.fields(it0, total) for s in setups]
    batch.step([f[:args.warmup] for f, _ in fg], [g[:args.warmup] for _, g in fg])
    t0 = time.perf_counter()
    batch.step([f[args.warmup:] for f, _ in fg], [g[args.warmup:] for _, g in fg])
    dt_wall = time.perf_counter() - t0
    dev_ms, nlaunch = batch.last_step_time()
    value = float(nk.sum()) * args.steps / dt_wall
    # the same instances, one context after another (per-instance launches)
    seq_steps = max(1, min(args.steps, 20))
    t1 = time.perf_counter()
    for m, (f, g) in zip(batch.members, fg):
        m.step(f[:seq_steps], g[:seq_steps])
    seq_wall = time.perf_counter() - t1
    seq_value = float(nk.sum()) * seq_steps / seq_wall
    S = 16 * nb * nb
    survey_step = float(np.sum((16 + 4 * n_op) * S * nk))   # survey figure: S (16 + 4 n_op) per k and RK4 step
    # own layout (as for the single workloads): 16 packed FP64 state transfers per RK4
    # step, and per stage packed H, the dipole components the instance's fields use
    # and full complex U with dephasing, at 4 bytes for members in mixed precision
    own = []
    for s_, (f, _), p in zip(setups, fg, problems):
        b = s_.info["nb"]
        ft = f[args.warmup:]
        nd = sum(1 for x in range(3) if np.any(ft[..., x] != 0) or np.any(ft[..., 3 + x] != 0))
        opw = 4 if p.precision != N.PRECISION_FP64 else 8
        per_k = 16 * 8 * b * b + 4 * (opw * b * b * (1 + nd) + (2 * opw * b * b if s_.info["n_diss"] > 0 else 0))
        own.append(per_k * s_.info["nk"])
    bytes_step = float(sum(own))
    achieved = bytes_step * args.steps / (dev_ms * 1e-3) / 1e9
    classes = sorted({(int(b) if b <= 4 else (8 if b <= 8 else 16 if b <= 16 else 32)) for b in nb})
    print(json.dumps({
        "metric": "k-point\u00b7steps/s, FP64 vs mixed/batched, at stated J(t) error; HBM GB/s, MFMA %",
        "value": value, "unit": "k-point\u00b7steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt_wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec is None else "f64/f32", "data": "synthetic",
        "config": {"workload": f"batch_scan_{len(setups)}_instances_seed3003_rk4_dt0.05", "instances": len(setups),
                   "total_nk": int(nk.sum()), "bands": {str(b): int(np.sum(nb == b)) for b in sorted(set(nb.tolist()))},
                   "kernel_classes": classes, "launches_per_step": nlaunch // args.steps,
                   "precision": "fp64" if prec is None else "mixed for nb >= 5", "setup_s": round(t_setup, 1),
                   "device_ms_per_step": dev_ms / args.steps,
                   "unbatched": {"value": seq_value, "steps": seq_steps, "speedup_of_batch": value / seq_value}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_stage_small_batch / k_stage_wave_batch (RK4 stages)",
                     "bytes_per_step": bytes_step,
                     "bytes_model": "packed Hermitian state and operators, the dipole components each member's fields use",
                     "survey_bytes_per_step": survey_step,
                     "survey_rate_gbs": round(survey_step * args.steps / (dev_ms * 1e-3) / 1e9, 2)},
        "cpu_baseline": None}))


def cpu_model():
    """CPU model name of the host (from /proc/cpuinfo), for the baseline's record."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.lower().startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(s, steps, subgrid=False):
    """The CPU oracle (FP64 restatement of the reference loops, OpenMP over k) on a
    bounded sample of the workload: a few RK4 steps of the setup's whole grid (the
    workload's own grid, or a sub-grid of the same model when the full grid is too
    large for the host) from the initial state at the pulse peak.  A GPU engine of
    the same setup repeats the same steps, and the J(t) error of the GPU path is
    stated from the two final states."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from edus_amd import native as N
    from oracle import oracle_py as O
    import helpers as Hh

    info = s.info
    nb = info["nb"]
    osys = Hh.oracle_system(s, use_oracle_tables=False, bruteforce_neighbours=False)
    cores = O.max_threads()
    P0 = s.array("rho0", True).reshape(-1, nb, nb)
    pumps = [s.pulse(i) for i in range(info["n_pumps"])]
    probe = s.pulse(info["n_pumps"])
    it0 = (info["itfi"] + 1) // 2
    t_start = s.time(it0)
    n = steps or 2
    t0 = time.perf_counter()
    _, Pc = osys["sys"].run(P0, 0, 4, info["dt"], t_start, n, 10 ** 9, pumps, probe)
    el = time.perf_counter() - t0
    n_left = info["itfi"] + 1 - it0  # the steps the run has from it0 on
    if not steps and el < 10.0:  # scale the sample to ~10-30 s of CPU work, within the run
        n = max(2, min(n_left, int(n * 12.0 / max(el, 1e-3))))
        t0 = time.perf_counter()
        _, Pc = osys["sys"].run(P0, 0, 4, info["dt"], t_start, n, 10 ** 9, pumps, probe)
        el = time.perf_counter() - t0
    # the same steps on the GPU, then J1 of both final states (GPU reductions)
    eng = N.Engine(s.problem(), precision=N.PRECISION_FP64)
    f, g = s.fields(it0, n)
    g[0] = 1  # the oracle's run starts with E_non0 = true
    eng.set_state(P0)
    eng.step(f, g)
    Pg = eng.get_state()
    jg = eng.observables()["j1"].copy()
    eng.set_state(Pc)
    jc = eng.observables()["j1"].copy()
    eng.close()
    j_err = float(np.max(np.abs(jg - jc)) / max(np.max(np.abs(jc)), 1e-300))
    p_err = float(np.max(np.abs(Pg - Pc)) / np.max(np.abs(Pc)))
    grid = f"{info['nk1']}x{info['nk2']}"
    return {"value": info["nk"] * n / el, "unit": "k-point\u00b7steps/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"{n} RK4 steps of the {'sub-grid ' if subgrid else 'full '}{grid} grid ({nb} bands"
                      f"{', the same model' if subgrid else ''}) from rho0 at t = {t_start:.3f} au, oracle, "
                      f"{cores} OpenMP threads",
            "seconds": round(el, 2), "gpu_vs_oracle": {"J1_rel_err": j_err, "rho_rel_err": p_err, "steps": n, "grid": grid}}


if __name__ == "__main__":
    main()
