"""Benchmark: mixed 2-256-atom request stream on one MI355X.

Workload: the serving stream -- ``--requests`` (2000) structures with Poisson
arrivals at ``--rate`` per second (10000/s by default, above the engine's
capacity, so that the value is the capacity; an untimed pass at
``--latency-rate``, 2000/s, reports the latencies below saturation), atom counts log-uniform in [2, 256], 1-4
species drawn from Z = 1..65 with seeded synthetic tables, random cells,
k-mesh N_a = max(1, ceil(|b*_a| / 0.2 A^-1)) with the crystallographic b*
(no 2 pi). Per request the device computes H(k)/S(k), all eigenvalues, the
Fermi level (bisection), band edges / gap and the DOS on the 2001-point
[-20, 10] eV grid. One "step" replays the whole stream; value is successful
structures per second over the timed steps (failed requests are counted and
excluded); p50/p99 latency (arrival to result) are pooled over the timed
steps.

Modes: ``serve`` (continuous batching engine) or ``plain`` (one structure at a
time, synchronous). Beside the timed steps, outside the timed region, the
serve line carries: the rooflines of the dominant kernels from a separate
sampled pass (execution windows of every 8th launch, plus HIP events on the
launch stream), the FP64 serving pass over the same stream with the gap /
band-edge errors of the benched (mixed) results against it, the plain
per-structure GPU path on a bounded prefix of the stream, and a CPU baseline
(the FP64 oracle on the host cores over a bounded sample).
"""
import argparse
import json
import os
import sys
import time

# The serving engine runs one batch per HIP stream; HIP maps streams onto this
# many hardware queues per process (its default, 4, serialises the engine's 8
# streams pairwise). Read once at HIP initialisation, so it is set before torch
# is imported.
SERVE_STREAMS = 8
os.environ["GPU_MAX_HW_QUEUES"] = str(SERVE_STREAMS)

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "structures/s + p50/p99 latency, mixed 2-256-atom stream, <=1 meV gap err"
STREAM_SEED = 2026
TABLE_SEED = 7
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=1)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--requests", type=int, default=2000)
    p.add_argument("--rate", type=float, default=10000.0,
                   help="Poisson arrival rate, requests/s (the default is above the engine's capacity, so value "
                        "is the capacity; --latency-rate gives the latencies below saturation)")
    p.add_argument("--latency-rate", type=float, default=2000.0,
                   help="arrival rate of the extra untimed pass that reports latency below saturation (0 = skip)")
    p.add_argument("--mode", choices=["serve", "plain", "fit", "large"], default="serve",
                   help="fit: the fitting workload (32 cells of 8-64 atoms, 3 species, DOS + gap loss, Adam "
                        "steps); large: one 1024-atom cubic cell (9216 orbitals) at Gamma + 4 complex k")
    p.add_argument("--fit-steps", type=int, default=200, help="Adam steps timed in --mode fit")
    p.add_argument("--fit-lr", type=float, default=1e-5,
                   help="AdamW learning rate in --mode fit (the reference's default, optim.py:1575)")
    p.add_argument("--sample-every", type=int, default=8,
                   help="sample every n-th tridiagonalisation mat-vec launch for the roofline")
    p.add_argument("--sample-capacity", type=int, default=8192,
                   help="event-sampled launches at most (0: count launches and bytes only, no events)")
    p.add_argument("--copy-mib", type=int, default=0,
                   help="run only a plain device copy of this many MiB (counter calibration) and exit")
    p.add_argument("--cpu-seconds", type=float, default=20.0, help="CPU baseline sample budget (0 = skip)")
    p.add_argument("--no-roofline", action="store_true")
    p.add_argument("--streams", type=int, default=SERVE_STREAMS, help="HIP streams of the serving engine")
    p.add_argument("--complex-only", action="store_true",
                   help="solve every batch as complex Hermitian (no real-symmetric batches at k with 2k integer)")
    p.add_argument("--graph-max-ld", type=int, default=256,
                   help="buckets up to this dimension replay their eigensolve from a HIP graph (0 = never)")
    p.add_argument("--plain-seconds", type=float, default=10.0,
                   help="budget of the plain per-structure pass over a prefix of the stream (0 = skip)")
    p.add_argument("--precision", choices=["mixed", "fp64"], default="mixed",
                   help="mixed: FP64 Cholesky/inverse + FP32 reduction/tridiagonalisation + FP64 bisection")
    return p.parse_args()


def make_stream(n):
    from slakonet_amd import synth
    return [synth.stream_structure(STREAM_SEED, i) for i in range(n)]


def stream_tables(structs, device):
    from slakonet_amd import synth
    from slakonet_amd.tables import TableSet
    species = sorted({int(z) for st in structs for z in st.z})
    skfs = synth.make_skf_set(species, seed=TABLE_SEED)
    return TableSet(skfs, synth.shell_dict_spd(species), device=device), skfs


def to_input(st):
    from slakonet_amd.engine import StructureInput
    from slakonet_amd.units import AA_BOHR
    return StructureInput(st.z, st.positions * AA_BOHR, st.cell * AA_BOHR, st.kpoints, st.k_weights)


PRECISION = "mixed"


def run_plain(tables, inputs, arrivals, grid, budget_s=None):
    """Sequential replay: each request waits for its arrival and the previous
    one. Returns (wall seconds, latency per request, results, requests done);
    with ``budget_s`` it stops after the request that crosses the budget."""
    import torch
    from slakonet_amd import engine
    lat, results = [], {}
    t0 = time.perf_counter()
    for i, inp in enumerate(inputs):
        now = time.perf_counter() - t0
        if now < arrivals[i]:
            time.sleep(arrivals[i] - now)
        try:
            plan = engine.plan_build(tables, [inp])
            H, S = engine.run_build(tables, plan)
            wsl = []
            w, _, info = engine.eigh(H, S, precision=PRECISION, check=False, workspace=wsl,
                                     real=REAL_SYMMETRIC and engine.real_kpoints(inp.kpoints))
            refine = (H, S, wsl[0]) if PRECISION == "mixed" else None
            edges, dos, stats = engine.post(plan, w, fermi_mode=0, dos_grid=grid, sigma=0.1, refine=refine)
            row = edges[0].cpu().numpy()
            res = engine.edges_dict(row)
            res["ok"] = not bool(info.cpu().numpy().any())
        except Exception as exc:
            res = {"ok": False, "error": f"{type(exc).__name__}: {exc}"}
        results[i] = res
        lat.append(time.perf_counter() - t0 - arrivals[i])
        if budget_s is not None and time.perf_counter() - t0 > budget_s:
            break
    torch.cuda.synchronize()
    return time.perf_counter() - t0, np.array(lat), results, len(lat)


N_STREAMS = SERVE_STREAMS
GRAPH_MAX_LD = 256
REAL_SYMMETRIC = True
GRAPH_STATS = {"captures": 0, "replays": 0}


def run_serve(tables, inputs, arrivals, grid, precision=None, quiet=False):
    from slakonet_amd.serving import ServingEngine
    eng = ServingEngine(tables, dos_grid=grid, precision=precision or PRECISION, n_streams=N_STREAMS,
                        graphs=GRAPH_MAX_LD > 0, graph_max_ld=GRAPH_MAX_LD, real_symmetric=REAL_SYMMETRIC)
    res = eng.replay(inputs, arrivals)
    eng.close()
    GRAPH_STATS["captures"] += eng.graph_captures
    GRAPH_STATS["replays"] += eng.graph_replays
    if not quiet:
        by_ld = {}
        for ld, nmat, ns, t0, t1 in eng.batch_log:
            e = by_ld.setdefault(ld, [0, 0, 0, 0.0])
            e[0] += 1
            e[1] += nmat
            e[2] += ns
            e[3] += t1 - t0
        summary = " ".join(f"{ld}:{e[0]}b/{e[2]}s/{e[3]:.2f}s" for ld, e in sorted(by_ld.items()))
        print(f"# batches by dimension (batches/structures/in-flight seconds): {summary}", file=sys.stderr,
              flush=True)
    return res["wall_s"], res["latency_s"], res["results"], len(inputs)


def edge_errors(mixed, ref):
    """Max |gap|, |VBM|, |CBM| differences (eV) between two result sets over
    the requests that succeeded in both and are insulators in the reference."""
    gap = vbm = cbm = 0.0
    n = flips = 0
    flip_gaps = []
    for i, r in ref.items():
        m = mixed.get(i)
        if m is None or not m.get("ok") or not r.get("ok"):
            continue
        if (r["metal"] > 0) != (m["metal"] > 0):
            # a gap at the metal threshold (1 meV) can land on either side
            flips += 1
            flip_gaps.append([r["cbm"] - r["vbm"], m["cbm"] - m["vbm"]])
            continue
        if r["metal"] > 0:
            continue
        n += 1
        vbm = max(vbm, abs(m["vbm"] - r["vbm"]))
        cbm = max(cbm, abs(m["cbm"] - r["cbm"]))
        gap = max(gap, abs((m["cbm"] - m["vbm"]) - (r["cbm"] - r["vbm"])))
    return {"max_gap_err_eV": gap, "max_vbm_err_eV": vbm, "max_cbm_err_eV": cbm, "insulators_compared": n,
            "metal_flag_flips": flips, "flip_edge_gaps_eV_fp64_mixed": flip_gaps[:8]}


HETRD_SAMPLE_EVERY = 8
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hetrd_step_traffic.json")
GEMM_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "gemm_f64_traffic.json")
FUSED_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fused_update_av_traffic.json")
FP64_MFMA_PEAK_TFS = 78.6
FP32_MFMA_PEAK_TFS = 157.3
# one compute unit's FP32 VALU rate: 4 SIMDs x 16 lanes x 2 flops (FMA) x 2.4 GHz
CU_FP32_VALU_PEAK_TFS = 4 * 16 * 2 * 2.4e9 / 1e12
KINDS = ("hetrd_step", "build", "gemm_f64", "chase", "fused_update_av", "dos", "fused_update_av_real",
         "chase_real")


def start_sampling(every, capacity):
    """Live sampling (inside the library): every launch of the sampled kinds
    is counted with its algorithmic amount; every ``every``-th launch of each
    records its execution window (earliest workgroup start, latest workgroup
    end on the GPU's constant-rate wall clock) and is bracketed by a HIP
    start/stop event pair on its launch stream. ``avg_launch_us`` is the
    execution window; ``avg_launch_us_events`` the event pair, which under
    four concurrent streams also holds the wait for a dispatch slot."""
    from slakonet_amd import _lib, engine
    _lib.check(_lib.lib().slk_sampler_reset(1, every, capacity), "slk_sampler_reset")
    engine.SAMPLING = True


def stop_sampling():
    import ctypes as C
    from slakonet_amd import _lib, engine
    lib = _lib.lib()
    res = {}
    for kind, name in enumerate(KINDS):
        buf = (C.c_double * 9)()
        _lib.check(lib.slk_sampler_read(kind, buf), "slk_sampler_read")
        res[name] = tuple(buf)
    _lib.check(lib.slk_sampler_reset(0, 0, 0), "slk_sampler_reset")
    engine.SAMPLING = False
    return res


def _ratio(path, key):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get(key)
    return None


def _window(sample):
    """(sampled launches, avg window ms, avg event ms, amount per sampled
    launch, launches counted, amount of all counted launches, aux per sampled
    launch)."""
    n, ms, amount, n_all, amount_all, ev_ms, _, aux, _ = sample
    if n <= 0 or ms <= 0:
        return None
    return n, ms / n, ev_ms / n, amount / n, n_all, amount_all, aux / n


def _mfma_roofline(sample, peak, kernel, traffic_file, traffic_key):
    g = _window(sample)
    if g is None:
        return None
    n, ms, ev, fl, n_all, fl_all, _ = g
    ach = fl / (ms * 1e-3) / 1e12
    ratio = _ratio(traffic_file, traffic_key)
    return {"bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
            "traffic": ratio * fl if ratio else None, "kernel": kernel,
            "avg_launch_us": 1000.0 * ms, "avg_launch_us_events": 1000.0 * ev,
            "algorithmic_flops_per_launch": fl, "launches_sampled": int(n), "sample_every": HETRD_SAMPLE_EVERY,
            "launches_total": int(n_all), "algorithmic_flops_total": fl_all, "est_kernel_ms_total": ms * n_all}


def roofline_fields(samples):
    """Rooflines from the live samples of the sampled pass.

    roofline: the FP32 fused stage-1 update with the next panel's A22 V
    (fused_update_av_kernel<float, true>, the real-symmetric batches, which
    carry most of its flops; the complex instantiation is
    roofline_fused_complex), the largest consumer of the device in the
    serving stream. Real: 16 flop per algorithmic byte, against HBM
    bandwidth; complex: 32 flop/B, against the FP32 dense MFMA peak.
    traffic = measured HBM bytes per algorithmic flop (counter runs) times
    the flops per launch.
    roofline_gemm_f64: the FP64 MFMA GEMM (Cholesky / L^-1 recursion, FP64
    reduction). roofline_chase: the bulge chase, one workgroup per matrix,
    against one CU's FP32 VALU rate (flops are counted per workgroup).
    build_roofline / dos_roofline: the H/S build and the DOS against HBM
    bandwidth (the DOS also as Gaussians per second of the full sum)."""
    out = {}
    rc = _mfma_roofline(samples["fused_update_av"], FP32_MFMA_PEAK_TFS,
                        "fused_update_av_kernel<float, false> (complex batches: stage-1 rank-2k update of the "
                        "lower triangle + next A22 V)",
                        FUSED_TRAFFIC_FILE, "hbm_bytes_per_flop")
    if rc is not None:
        rc["algorithmic_bytes_per_flop"] = 1.0 / 32.0   # lower triangle read and written once, 8 B per element
    rr = _mfma_roofline(samples["fused_update_av_real"], FP32_MFMA_PEAK_TFS,
                        "fused_update_av_kernel<float, true> (real-symmetric batches: stage-1 rank-2k update of "
                        "the lower triangle + next A22 V, one real MFMA per block)",
                        FUSED_TRAFFIC_FILE, "hbm_bytes_per_flop_real")
    if rr is not None:
        # 64 m^2 flops per matrix against the real lower triangle read and written once in FP32
        # (m^2/2 x 4 B each way): 16 flop/B, below the FP32 ridge (157.3 TF/s / 8 TB/s = 19.7 flop/B),
        # so this instantiation is bound by HBM. The kernel keeps the complex64 layout (zero imaginary
        # parts), so its measured traffic is at least twice these bytes.
        fl = rr["algorithmic_flops_per_launch"]
        t_s = rr["avg_launch_us"] * 1e-6
        rr.update({"bound": "hbm", "achieved": fl / 16.0 / t_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": fl / 16.0 / t_s / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": fl / 16.0,
                   "algorithmic_bytes_per_flop": 1.0 / 16.0, "achieved_tflops": fl / t_s / 1e12,
                   "frac_of_fp32_mfma_peak": fl / t_s / 1e12 / FP32_MFMA_PEAK_TFS})
    # the instantiation with more flops in the stream is the line's roofline
    if rr is not None and (rc is None or rr["algorithmic_flops_total"] >= rc["algorithmic_flops_total"]):
        out["roofline"] = rr
        if rc is not None:
            out["roofline_fused_complex"] = rc
    elif rc is not None:
        out["roofline"] = rc
        if rr is not None:
            out["roofline_fused_real"] = rr
    r = _mfma_roofline(samples["gemm_f64"], FP64_MFMA_PEAK_TFS,
                       "gemm_mfma_kernel<double> (Cholesky trailing updates, L^-1)", GEMM_TRAFFIC_FILE,
                       "hbm_bytes_per_flop")
    if r is not None:
        out["roofline_gemm_f64"] = r
    ch = _window(samples["chase"])
    if ch is not None:
        n, ms, ev, fl, n_all, fl_all, _ = ch
        ach = fl / (ms * 1e-3) / 1e12
        out["roofline_chase"] = {"bound": "valu (one CU)", "achieved": ach, "peak": CU_FP32_VALU_PEAK_TFS,
                                 "unit": "TFLOP/s", "frac": ach / CU_FP32_VALU_PEAK_TFS,
                                 "frac_of_device_fp32": ach / FP32_MFMA_PEAK_TFS,
                                 "kernel": "band_chase_kernel<float> (band -> tridiagonal, one workgroup per matrix)",
                                 "avg_launch_us": 1000.0 * ms, "avg_launch_us_events": 1000.0 * ev,
                                 "algorithmic_flops_per_workgroup": fl, "launches_sampled": int(n),
                                 "launches_total": int(n_all), "est_kernel_ms_total": ms * n_all}
    ch = _window(samples["chase_real"])
    if ch is not None:
        n, ms, ev, fl, n_all, fl_all, _ = ch
        ach = fl / (ms * 1e-3) / 1e12
        out["roofline_chase_real"] = {"bound": "valu (one CU)", "achieved": ach, "peak": CU_FP32_VALU_PEAK_TFS,
                                      "unit": "TFLOP/s", "frac": ach / CU_FP32_VALU_PEAK_TFS,
                                      "kernel": "band_chase_kernel<float, *, true> (real-symmetric batches)",
                                      "avg_launch_us": 1000.0 * ms, "avg_launch_us_events": 1000.0 * ev,
                                      "algorithmic_flops_per_workgroup": fl, "launches_sampled": int(n),
                                      "launches_total": int(n_all), "est_kernel_ms_total": ms * n_all}
    h = _window(samples["hetrd_step"])
    if h is not None:
        n, ms, ev, nbytes, n_all, _, _ = h
        ach = nbytes / (ms * 1e-3) / 1e9
        ratio = _ratio(TRAFFIC_FILE, "hbm_bytes_per_algorithmic_byte")
        out["roofline_hetrd_step"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                      "frac": ach / HBM_PEAK_GBS, "traffic": ratio * nbytes if ratio else None,
                                      "kernel": "hetrd_step_kernel<float> (one-stage trailing mat-vec)",
                                      "avg_launch_us": 1000.0 * ms, "launches_sampled": int(n),
                                      "launches_total": int(n_all)}
    b = _window(samples["build"])
    if b is not None:
        n, ms, ev, nbytes, n_all, _, _ = b
        ach = nbytes / (ms * 1e-3) / 1e9
        out["build_roofline"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": ach / HBM_PEAK_GBS,
                                 "kernel": "slk_fill_kernel + slk_build_rows_kernel / slk_build_kernel "
                                           "(complex128; the window spans the launch sequence)",
                                 "avg_launch_us": 1000.0 * ms, "avg_launch_us_events": 1000.0 * ev,
                                 "algorithmic_bytes_per_launch": nbytes,
                                 "launches_sampled": int(n), "launches_total": int(n_all)}
    d = _window(samples["dos"])
    if d is not None:
        n, ms, ev, nbytes, n_all, _, gauss = d
        ach = nbytes / (ms * 1e-3) / 1e9
        out["dos_roofline"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": ach / HBM_PEAK_GBS, "kernel": "dos_kernel (windowed Gaussian sum)",
                               "gaussians_per_s": gauss / (ms * 1e-3),
                               "gaussians_per_launch": gauss, "algorithmic_bytes_per_launch": nbytes,
                               "avg_launch_us": 1000.0 * ms, "avg_launch_us_events": 1000.0 * ev,
                               "launches_sampled": int(n), "launches_total": int(n_all)}
    return out


def run_fit(args):
    """Fitting workload: teacher targets, then timed AdamW steps (adjoint backward)."""
    import torch
    from slakonet_amd import fit, synth
    from slakonet_amd.optim import MultiElementSkfParameterOptimizer
    from slakonet_amd.tables import TableSet
    dev = torch.device("cuda:0")
    species = (6, 14, 31)
    sp = sorted(species)
    structs = synth.fitting_batch(seed=3, n_cells=32, species=species)
    grid = torch.linspace(-20.0, 10.0, 2001, dtype=torch.float64, device=dev)
    teacher = TableSet(synth.make_skf_set(sp, seed=TABLE_SEED, teacher_seed=TABLE_SEED + 1),
                       synth.shell_dict_spd(sp), device=dev)
    ds = fit.teacher_datasets(structs, teacher, grid)
    model = MultiElementSkfParameterOptimizer(skf_dicts=synth.make_skf_set(sp, seed=TABLE_SEED), device=dev).to(dev)
    shells = synth.shell_dict_spd(sp)
    fit.train_skf_parameters(model, ds, shells, num_epochs=max(1, args.warmup), learning_rate=args.fit_lr,
                             loss="dos_gap", log=None)
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats(dev)
    t0 = time.perf_counter()
    _, hist = fit.train_skf_parameters(model, ds, shells, num_epochs=args.fit_steps, learning_rate=args.fit_lr,
                                       loss="dos_gap", early_stopping_patience=10 ** 9,
                                       log=lambda m: print("# " + m, file=sys.stderr, flush=True), log_every=20)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    n_mat = sum(len(st.kpoints) for st in structs)
    out = {
        "metric": "fitting Adam steps/s (32 cells, DOS + gap loss, adjoint backward)", "value": len(hist) / total,
        "unit": "steps/s", "n_gpus": 1, "steps": len(hist), "warmup": args.warmup,
        "ms_per_step": 1000.0 * total / max(1, len(hist)), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "loss_first": hist[0]["total_loss"], "loss_last": hist[-1]["total_loss"],
        "loss_min": min(h["total_loss"] for h in hist),
        "loss_every_20": [round(h["total_loss"], 6) for h in hist[::20]],
        "peak_hbm_bytes": int(torch.cuda.max_memory_allocated(dev)),
        "config": {"workload": f"fitting: 32 random cells 8-64 atoms, species {species}, {n_mat} k-matrices, "
                               "teacher tables (A,b,B,c x (1+N(0,0.1))), loss MSE(DOS 2001 pts) + 0.1 gap err^2, "
                               f"AdamW lr {args.fit_lr:g}, clip 2.0"},
    }
    print(json.dumps(out))


def run_large(args):
    """1024-atom cubic cell (L = 22 A, 3 species, d_min 1.9 A): 9216x9216 at
    Gamma and four complex k. Time to solution of the precision form (FP32
    bulk + FP64 Cholesky / tridiagonal / edge refinement) and its errors
    against the all-FP64 device path on the same matrices."""
    import torch
    from slakonet_amd import engine, synth
    from slakonet_amd.tables import TableSet
    from slakonet_amd.units import AA_BOHR
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    st = synth.large_cubic()
    print(f"# structure generated in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    species = sorted({int(z) for z in st.z})
    tab = TableSet(synth.make_skf_set(species, seed=TABLE_SEED), synth.shell_dict_spd(species), device=dev)
    inp = engine.StructureInput(st.z, st.positions * AA_BOHR, st.cell * AA_BOHR, st.kpoints, st.k_weights)
    grid = torch.linspace(-20.0, 10.0, 2001, dtype=torch.float64, device=dev)
    res = {}
    for prec in ("mixed", "fp64"):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        plan = engine.plan_build(tab, [inp])
        H, S = engine.run_build(tab, plan)
        torch.cuda.synchronize()
        t_build = time.perf_counter() - t1
        stage = []
        wsl = []
        w, _, _ = engine.eigh(H, S, precision=prec, stage_ms=stage, workspace=wsl)
        refine = (H, S, wsl[0]) if prec == "mixed" else None
        edges, dos, _ = engine.post(plan, w, dos_grid=grid, refine=refine)
        torch.cuda.synchronize()
        t_all = time.perf_counter() - t1
        res[prec] = dict(w=w[:, :plan.n_orb[0]].cpu().numpy(), edges=engine.edges_dict(edges[0].cpu().numpy()),
                         dos=dos[0].cpu().numpy(), t=t_all, t_build=t_build, stages=stage[0])
        print(f"# {prec}: {t_all:.2f} s (build {t_build:.2f} s; stages ms {stage[0]})", file=sys.stderr,
              flush=True)
        del H, S, wsl, w
        torch.cuda.empty_cache()
    n = 9 * len(st.z)
    m, f = res["mixed"], res["fp64"]
    eig_err = float(np.abs(m["w"] - f["w"]).max()) * 27.211
    e = grid.cpu().numpy()
    l1 = float(np.trapezoid(np.abs(m["dos"] - f["dos"]), e) / np.trapezoid(f["dos"], e))
    nk = len(st.kpoints)
    # complex multiply-adds per matrix, x8 real flops each: Cholesky n^3/6 and L^-1 n^3/6 (FP64, one
    # recursion timed as "potrf"); reduce n^3/2 (W H on and below the diagonal with W lower, n^3/3,
    # then (W H) W^H on and below it, n^3/6; FP32 in the precision form); hetrd 2n^3/3
    fl = {"potrf": 8 * n ** 3 / 3, "reduce": 4 * n ** 3, "hetrd": 16 * n ** 3 / 3}
    mfma = {k: nk * fl[k] / (m["stages"][k] * 1e-3) / 1e12 for k in fl}
    out = {
        "metric": "time to solution, 1024-atom cell (9216 orbitals), Gamma + 4 k, eigenvalues + gap + DOS",
        "value": m["t"], "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": 1000.0 * m["t"],
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "fp64_path_s": f["t"],
        "errors_vs_fp64_path": {"max_eig_err_eV": eig_err,
                                "vbm_err_eV": abs(m["edges"]["vbm"] - f["edges"]["vbm"]),
                                "cbm_err_eV": abs(m["edges"]["cbm"] - f["edges"]["cbm"]),
                                "dos_l1_rel": l1, "metal": bool(f["edges"]["metal"])},
        "stage_ms": m["stages"], "stage_tflops": mfma,
        "mfma_utilisation": {"potrf": mfma["potrf"] / FP64_PEAK_TFLOPS, "reduce": mfma["reduce"] / 157.3},
        "config": {"workload": f"cubic L=22 A, {len(st.z)} atoms (Z 6/8/14), d_min 1.9 A, {nk} k-points, "
                               f"n={n}, 2001-pt DOS"},
    }
    print(json.dumps(out))


def host_info():
    """CPU model, affinity core count and the BLAS the oracle's numpy/scipy use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    blas = []
    try:
        from threadpoolctl import threadpool_info
        import scipy.linalg  # noqa: F401  (loads scipy's own BLAS too)
        blas = [{"api": d.get("internal_api"), "version": d.get("version"), "threads": d.get("num_threads")}
                for d in threadpool_info()]
    except Exception:
        pass
    return {"cpu_model": model, "affinity_cores": len(os.sched_getaffinity(0)), "blas": blas}


def cpu_baseline(structs, skfs, budget_s):
    """FP64 oracle (numpy build + LAPACK via numpy/scipy) over the stream in
    order until the budget is spent, on the host cores of this run. The Python
    loops are single-threaded; the LAPACK calls use the BLAS thread pool, whose
    size is reported as the cores used."""
    from oracle import build as obuild, post as opost, solve as osolve
    from slakonet_amd.units import AA_BOHR
    from slakonet_amd import synth
    info = host_info()
    threads = max([b["threads"] or 1 for b in info["blas"]] or [1])
    t0 = time.perf_counter()
    done = 0
    grid_args = dict(energy_range=(-20.0, 10.0), num_points=2001, sigma=0.1)
    for st in structs:
        species = sorted({int(z) for z in st.z})
        sk = {(a, b): skfs[(a, b)] for a in species for b in species}
        hs = obuild.build_hs(st.z, st.positions * AA_BOHR, st.cell * AA_BOHR, synth.shell_dict_spd(species), sk,
                             st.kpoints)
        w, _ = osolve.solve_all(hs["H"], hs["S"])
        ne = obuild.n_electrons(st.z, sk)
        mu = opost.fermi_bisect(w, ne, st.k_weights)
        opost.band_gap(w, mu)
        opost.dos(w, st.k_weights, mu, **grid_args)
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    return {"value": done / el, "unit": "structures/s", "cores": int(threads), "kind": "port",
            "sample": f"first {done} requests of the stream, in order (FP64 oracle: numpy build + LAPACK "
                      f"Cholesky/eigh per k + Fermi bisection + DOS), {el:.1f} s",
            "cpu_model": info["cpu_model"], "affinity_cores": info["affinity_cores"], "blas": info["blas"]}


def main():
    args = parse()
    global PRECISION, N_STREAMS, GRAPH_MAX_LD, REAL_SYMMETRIC
    PRECISION = args.precision
    N_STREAMS = args.streams
    GRAPH_MAX_LD = args.graph_max_ld
    REAL_SYMMETRIC = not args.complex_only
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the stream is data-parallel "
              "per request and would shard by request across GPUs)", file=sys.stderr)
        sys.exit(2)
    import torch
    from slakonet_amd import synth
    torch.cuda.set_device(0)
    global HETRD_SAMPLE_EVERY
    HETRD_SAMPLE_EVERY = max(1, args.sample_every)
    if args.copy_mib > 0:
        # plain copy of a known size: HBM counter calibration (reads and writes copy_mib MiB)
        src = torch.ones(args.copy_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 4}))
        return
    if args.mode == "fit":
        run_fit(args)
        return
    if args.mode == "large":
        run_large(args)
        return
    dev = torch.device("cuda:0")
    t_gen = time.perf_counter()
    structs = make_stream(args.requests)
    tables, skfs = stream_tables(structs, dev)
    inputs = [to_input(st) for st in structs]
    arrivals = synth.poisson_arrivals(args.requests, args.rate, STREAM_SEED)
    grid = torch.linspace(-20.0, 10.0, 2001, dtype=torch.float64, device=dev)
    mean_nk = float(np.mean([len(st.kpoints) for st in structs]))
    print(f"# stream generated in {time.perf_counter() - t_gen:.1f} s; mode={args.mode}", file=sys.stderr,
          flush=True)
    runner = run_serve if args.mode == "serve" else run_plain
    # warmup steps: the same pass over the stream as a timed step (every
    # bucket's shapes reach the caching allocator and the kernels' one-time
    # attributes are set before the timed region)
    for _ in range(args.warmup):
        runner(tables, inputs, arrivals, grid)
        print("# warmup done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    if args.steps == 0:
        # counter runs: only the sampled pass (launch counts and algorithmic
        # amounts per kind, to divide the counters by), no timed region
        start_sampling(HETRD_SAMPLE_EVERY, args.sample_capacity)
        run_serve(tables, inputs, arrivals, grid, quiet=True)
        torch.cuda.synchronize()
        samples = stop_sampling()
        print(json.dumps({"sampler_counts": {k: [int(v[3]), v[4], v[8]] for k, v in samples.items()},
                          **roofline_fields(samples)}))
        return
    walls, lats, n_ok, n_failed, results = [], [], 0, 0, None
    t_start = time.perf_counter()
    for s in range(args.steps):
        wall, lat, results, n_done = runner(tables, inputs, arrivals, grid)
        walls.append(wall)
        lats.append(lat)
        ok = sum(1 for r in results.values() if r.get("ok"))
        n_ok += ok
        n_failed += n_done - ok
        print(f"# step {s}: {wall:.2f} s ({n_done - ok} failed)", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    total = time.perf_counter() - t_start
    lat_all = np.concatenate(lats)
    value = n_ok / total
    out = {
        "metric": METRIC, "value": value, "unit": "structures/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1000.0 * total / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64" if PRECISION == "mixed" else "f64",
        "data": "synthetic",
        "p50_latency_ms": float(np.percentile(lat_all, 50) * 1000),
        "p99_latency_ms": float(np.percentile(lat_all, 99) * 1000),
        "n_failed": int(n_failed),
        "config": {"workload": f"serving stream: {args.requests} requests, atoms log-uniform 2-256, 1-4 species "
                               f"Z<=65, Poisson {args.rate:g}/s, k-mesh N_a = max(1, ceil(|b*_a| / 0.2 A^-1)) "
                               f"(b* without 2 pi; mean {mean_nk:.2f} k-points), full spectrum + Fermi + gap + "
                               "2001-pt DOS",
                   "mode": args.mode, "requests": args.requests, "rate_per_s": args.rate,
                   "latency": "arrival to result on the host, pooled over the timed steps"},
    }
    if args.mode == "serve":
        from slakonet_amd import engine as _eng
        n_real = sum(1 for inp in inputs if _eng.real_kpoints(inp.kpoints))
        out["config"]["serving"] = {"streams": N_STREAMS, "graph_max_ld": GRAPH_MAX_LD,
                                    "real_symmetric_batches": REAL_SYMMETRIC,
                                    "requests_at_real_k": n_real,
                                    "graph_captures": GRAPH_STATS["captures"],
                                    "graph_replays": GRAPH_STATS["replays"]}
    print(json.dumps({"partial": True, "value": value, "p99_latency_ms": out["p99_latency_ms"]}), file=sys.stderr,
          flush=True)
    if args.mode == "serve":
        # --- outside the timed region -------------------------------------
        if not args.no_roofline:
            start_sampling(HETRD_SAMPLE_EVERY, args.sample_capacity)
            run_serve(tables, inputs, arrivals, grid, quiet=True)
            torch.cuda.synchronize()
            samples = stop_sampling()
            out.update(roofline_fields(samples))
            out["sampler_counts"] = {k: [int(v[3]), v[4]] for k, v in samples.items()}
            print("# sampled pass done", file=sys.stderr, flush=True)
        # latency below saturation: the same stream arriving at --latency-rate
        if args.latency_rate > 0:
            arr_l = synth.poisson_arrivals(args.requests, args.latency_rate, STREAM_SEED)
            wall_l, lat_l, res_l, _ = run_serve(tables, inputs, arr_l, grid, quiet=True)
            ok_l = sum(1 for r in res_l.values() if r.get("ok"))
            out["at_latency_rate"] = {"rate_per_s": args.latency_rate, "value": ok_l / wall_l,
                                      "unit": "structures/s",
                                      "p50_latency_ms": float(np.percentile(lat_l, 50) * 1000),
                                      "p99_latency_ms": float(np.percentile(lat_l, 99) * 1000)}
            print("# latency-rate pass done", file=sys.stderr, flush=True)
        # the accuracy guard and the FP64 serving figure: the same stream in FP64
        wall64, lat64, res64, _ = run_serve(tables, inputs, arrivals, grid, precision="fp64", quiet=True)
        ok64 = sum(1 for r in res64.values() if r.get("ok"))
        out["fp64_serving"] = {"value": ok64 / wall64, "unit": "structures/s",
                               "p50_latency_ms": float(np.percentile(lat64, 50) * 1000),
                               "p99_latency_ms": float(np.percentile(lat64, 99) * 1000), "dtype": "f64"}
        if PRECISION == "mixed":
            out["errors_vs_fp64"] = edge_errors(results, res64)
        print("# fp64 pass done", file=sys.stderr, flush=True)
        # the plain per-structure GPU path (same precision) on a bounded prefix,
        # every request present at t = 0: structures/s and service latency
        if args.plain_seconds > 0:
            wallp, latp, resp, n_p = run_plain(tables, inputs, np.zeros(len(inputs)), grid,
                                               budget_s=args.plain_seconds)
            svc = np.diff(np.concatenate([[0.0], latp]))
            okp = sum(1 for r in resp.values() if r.get("ok"))
            out["plain_path"] = {"value": okp / wallp, "unit": "structures/s", "requests": n_p,
                                 "p50_service_ms": float(np.percentile(svc, 50) * 1000),
                                 "p99_service_ms": float(np.percentile(svc, 99) * 1000),
                                 "dtype": out["dtype"],
                                 "sample": f"first {n_p} requests of the stream, one structure at a time"}
            out["serving_over_plain"] = value / out["plain_path"]["value"]
            print("# plain pass done", file=sys.stderr, flush=True)
    if args.cpu_seconds > 0:
        out["cpu_baseline"] = cpu_baseline(structs, skfs, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
