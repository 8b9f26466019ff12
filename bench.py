#!/usr/bin/env python
"""Benchmark of the CCCS beamlet-dose path on one MI355X.

Metric: beamlets/s of the whole path (Terma -> CCCS -> BEV->PVCS resampling into CSR rows) at
<= 0.5 % per-beamlet max-dose error against the FP64 oracle, with p50/p99 per-beam latency.

Headline (`value`): the batched 4pi workload -- a 200 x 200 x 160 body at 2.5 mm (smoothed seeded
noise, 0.2-1.8 g/cc inside an ellipsoid, 0.0012 outside), 4 isocenters in an ellipsoidal PTV, 180
Fibonacci-sphere beams with 20 x 20 fluence grids keeping the beamlets that intersect the PTV
(flagged on the device), 8 zenith x 8 azimuth cones, FP32.  A step is one batched call over all
180 beams (one matrix per beam), with the launch setting the autotuner picks on every other beam;
density, tables and beamlet inputs are resident before the timed region.  The per-beam latency
beside it comes from the serving workload (220 x 220 x 150 body, 1160 candidate beams, Poisson
arrivals at 0.9 x capacity).  Extra legs state each other workload with its error against the FP64
oracle: plain one-beam-per-call on the same 180 beams, the 64^3 water beam, the 128^3 slab beam,
the precision setting (1e-3 threshold, BF16/FP16 values, 256 fluence vectors) and serving at
several loads and batch sizes.  The CPU baseline is the FP64 oracle on a bounded sample of the
headline's beamlets, timed on this host's cores.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FLOPS_PER_SEGMENT = 40  # algorithmic flops of one (ray, segment) CCCS step, exp counted as 1
N_CU, CLOCK_HZ = 256, 2.4e9  # MI355X compute units and peak engine clock
TOL = 0.005


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * q / 100.0
    lo, hi = int(k), min(int(k) + 1, len(xs) - 1)
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def host_threads() -> int:
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(n, 64))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            return next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        return ""


def profiled_value(name: str, key: str):
    """A `key: value` line of a committed rocprofv3 summary under profiles/, or None."""
    import re

    path = ROOT / "profiles" / name
    if not path.exists():
        return None
    m = re.search(rf"{key}:\s*([0-9.eE+]+)", path.read_text())
    return int(float(m.group(1))) if m else None


def profiled_traffic(name: str):
    """HBM bytes per launch of a kernel from a committed rocprofv3 counter summary (FETCH_SIZE +
    WRITE_SIZE, scaled by the same counters on a copy of known size), or None."""
    return profiled_value(name, "traffic_bytes_per_launch")


# ---------------------------------------------------------------------------- oracle side (checker)

def oracle_rows(cfg, items):
    """FP64 oracle rows (Terma -> CCCS -> resample) of (beam, fluence) items on the host cores."""
    from concurrent.futures import ThreadPoolExecutor

    import oracle as orc

    ovol, ospec, okern = (orc.cast(orc.Volume, cfg.volume), orc.cast(orc.Spectrum, cfg.spectrum),
                          orc.cast(orc.Kernel, cfg.kernel))

    def one(it):
        return orc.beamlet_pipeline(orc.cast(orc.Beam, it[0]), it[1], ovol, cfg.density, ospec, okern, cfg.extent)

    with ThreadPoolExecutor(max_workers=host_threads()) as ex:
        return list(ex.map(one, items))


def row_error(cols, vals, ref):
    """max |GPU - FP64| / max FP64 over the union of columns (missing = 0), and column-set equality."""
    import numpy as np

    gc = np.asarray(cols).astype(np.int64)
    u = np.union1d(gc, ref["cols"])
    dg, dr = np.zeros(len(u)), np.zeros(len(u))
    dg[np.searchsorted(u, gc)] = np.asarray(vals, dtype=np.float64)
    dr[np.searchsorted(u, ref["cols"])] = ref["vals"]
    return float(np.abs(dg - dr).max() / dr.max()), bool(np.array_equal(gc, ref["cols"]))


def accuracy(cfg, items, rows):
    """GPU rows [(cols, vals)] of items [(beam, fluence)] against the FP64 oracle."""
    refs = oracle_rows(cfg, items)
    worst, same = 0.0, True
    for (c, v), ref in zip(rows, refs):
        e, s = row_error(c, v, ref)
        worst, same = max(worst, e), same and s
    return {"beamlets": len(items), "max_err_rel_beamlet_max": worst, "within_0p5pct": worst <= TOL,
            "columns_equal": same, "longest_L": int(max(it[0].long_dim for it in items))}


def cpu_baseline(cfg, items, what: str):
    """The FP64 oracle on the host cores (OpenMP over beamlets): beamlets/s on the sample, plus a
    one-core figure on its first 4 beamlets."""
    import numpy as np

    import oracle as orc

    threads = host_threads()
    beams = [b for b, _ in items]
    flu = np.stack([f for _, f in items])
    assert np.all(flu == flu[0]), "the oracle baseline takes one sub-fluence pattern for all beamlets"
    args = (orc.cast(orc.Volume, cfg.volume), cfg.density, orc.cast(orc.Spectrum, cfg.spectrum),
            orc.cast(orc.Kernel, cfg.kernel), cfg.extent, 0)
    t0 = time.perf_counter()
    orc.dose_beamlets(beams, flu[0], *args, threads)
    dt = time.perf_counter() - t0
    n1 = min(4, len(beams))
    t1 = time.perf_counter()
    orc.dose_beamlets(beams[:n1], flu[0], *args, 1)
    dt1 = time.perf_counter() - t1
    return {
        "value": round(len(beams) / dt, 3),
        "unit": "beamlets/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{len(beams)} {what}, full FP64 path Terma->CCCS->resample, OpenMP over beamlets on {threads} "
                  f"threads, {dt:.1f} s wall",
        "one_core": {"value": round(n1 / dt1, 3), "beamlets": n1, "wall_s": round(dt1, 2)},
        "host": {"cpu_model": cpu_model(), "nproc": os.cpu_count()},
    }


# ---------------------------------------------------------------------------- GPU helpers

def nseg_of(ctx, fd, beams):
    """CCCS (ray, segment) steps of every beamlet.  The lattice walk depends only on the beamlet size,
    the longitudinal spacing and long_dim, so each distinct triple is walked once."""
    import numpy as np

    key_rep = {}
    keys = []
    for i in range(len(beams)):
        b = beams[i]
        k = (float(b.beamlet_size[0]), float(b.beamlet_size[1]), float(b.long_spacing), int(b.long_dim))
        keys.append(k)
        key_rep.setdefault(k, i)
    reps = list(key_rep.items())
    counts = ctx.cccs_segments(fd.beam_array([beams[i] for _, i in reps]))
    per = {k: int(c) for (k, _), c in zip(reps, counts)}
    return np.array([per[k] for k in keys], dtype=np.int64)


def matrix_rows(dm, store, idx):
    """Host (cols, vals) of global beamlets idx from a DoseMat with one matrix per store beam."""
    cache, out = {}, []
    for g in idx:
        k, r = store.beam_of(g)
        if k not in cache:
            cache[k] = dm.matrix(k)
        off, col, val = cache[k]
        out.append((col[int(off[r]):int(off[r + 1])], val[int(off[r]):int(off[r + 1])]))
    return out


def fourpi(synth, which):
    if which == 3:
        return synth.fourpi_config("batched-4pi-200x200x160", (200, 200, 160), 1003, 180, (22.0, 15.0, 13.0),
                                   with_flags=False)
    return synth.fourpi_config("serving-4pi-220x220x150", (220, 220, 150), 1005, 1160, (26.0, 18.0, 15.0),
                               with_flags=False)


# ---------------------------------------------------------------------------- headline: batched 4pi

def leg_headline(args, fd, synth, workload, torch, np):
    t0 = time.perf_counter()
    cfg = fourpi(synth, 3)
    ctx = fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel, options=fd.FD_OPT_TIMING)
    flags = workload.device_flags(ctx, cfg, torch)
    store = workload.BeamStore(cfg, flags)
    n = store.n
    nvox = cfg.volume.dim[0] * cfg.volume.dim[1] * cfg.volume.dim[2]
    arr, flu, cnt = store.request(range(len(cfg.bundles)))
    log(f"batched 4pi: {len(cfg.bundles)} beams, {n} beamlets ({store.counts.min()}..{store.counts.max()} per "
        f"beam), setup {time.perf_counter() - t0:.1f} s")
    # launch setting: tuned on every other beam (setup, not timed); graph replay off so that the HIP
    # events of every kernel launch are recorded inside the timed steps
    t = time.perf_counter()
    sa, sf, sc = store.request(range(0, len(cfg.bundles), 2))
    tuning = ctx.autotune(sa, sf, sc, cfg.extent)
    ctx.set_tuning(tuning["batch_beamlets"], 0, tuning["terma_segments"])
    log(f"batched 4pi: autotune {tuning} in {time.perf_counter() - t:.1f} s; timed with graph replay off")
    nnz = 0
    for w in range(max(args.warmup, 1)):
        dm = fd.DoseMat(nvox, nnz)
        dm.add_beams(ctx, arr, flu, cnt, 0, cfg.extent)
        torch.cuda.synchronize()
        nnz = dm.info()[2]
        dm.close()
        log(f"warmup {w}: nnz {nnz}")
    kt, kn, lat = {}, {}, []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        dm = fd.DoseMat(nvox, nnz)
        dm.add_beams(ctx, arr, flu, cnt, 0, cfg.extent)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - ts) * 1e3)
        if k < args.steps - 1:
            dm.close()
        for name, ms in ctx.timings().items():
            kt.setdefault(name, []).append(ms)
        for name, c in ctx.timing_launches().items():
            kn.setdefault(name, []).append(c)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    log(f"timed {args.steps} steps in {elapsed:.3f} s")
    kstep = {k: float(np.mean(v)) for k, v in kt.items()}           # ms per step, summed over launches
    launches = {k: int(round(float(np.mean(v)))) for k, v in kn.items()}
    nseg = int(nseg_of(ctx, fd, arr).sum())
    cccs_launch_ms = kstep["cccs"] / launches["cccs"]
    # batches of many beamlets per march geometry run the table-fed CCCS kernel (its tables are
    # timed in their own slot, "cccs_tables")
    cccs_name = "cccs_tab_kernel" if "cccs_tables" in kstep else "cccs_kernel"
    achieved = FLOPS_PER_SEGMENT * nseg / (kstep["cccs"] * 1e-3) / 1e12
    res = {
        "value": n * args.steps / elapsed,
        "ms_per_step": elapsed / args.steps * 1e3,
        "step_ms": {"p50": round(percentile(lat, 50), 3), "p99": round(percentile(lat, 99), 3)},
        "beams": len(cfg.bundles), "beamlets": n, "nnz": int(nnz),
        "beamlets_per_beam": [int(store.counts.min()), int(store.counts.max())],
        "long_dim_max": int(max(arr[i].long_dim for i in range(n))),
        "tuning": {k: (round(v, 1) if isinstance(v, float) else v) for k, v in tuning.items()},
        "kernels_ms_per_step": {k: round(v, 3) for k, v in kstep.items()},
        "kernel_launches_per_step": launches,
        "roofline": {
            "bound": "valu",
            "kernel": cccs_name,
            "achieved": round(achieved, 3),
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
            "traffic": profiled_traffic(f"{cccs_name}_batched4pi_traffic.md"),
            "algorithmic": f"{FLOPS_PER_SEGMENT} flop x {nseg // launches['cccs']} (ray, segment) steps per launch "
                           f"(average of {launches['cccs']} launches per step, {nseg} per step); "
                           f"average launch {cccs_launch_ms:.3f} ms",
            # the vector instructions the kernel issues (counter summary of the same workload): wave64
            # instructions per CU per clock at the bench's own launch time
            "issue": issue_block(f"{cccs_name}_batched4pi_counters.md", cccs_launch_ms),
        },
    }
    # the last step's matrices for the accuracy check, kept with the beamlet store
    return res, (cfg, ctx, store, dm, arr, flu, cnt)


def issue_block(name: str, launch_ms: float):
    valu = profiled_value(name, "valu_instr_per_launch")
    if valu is None:
        return None
    return {"valu_instr_per_launch": valu, "trans_instr_per_launch": profiled_value(name, "trans_instr_per_launch"),
            "valu_per_cu_cycle": round(valu / (N_CU * CLOCK_HZ * launch_ms * 1e-3), 3),
            "source": f"profiles/{name} (SQ_INSTS_VALU), {N_CU} CUs at {CLOCK_HZ / 1e9:.1f} GHz"}


def leg_plain_4pi(fd, torch, np, state):
    """The same 180 beams one beam per call (default launch setting)."""
    cfg, ctx, store, _, _, _, _ = state
    nvox = cfg.volume.dim[0] * cfg.volume.dim[1] * cfg.volume.dim[2]
    ctx.set_tuning(8192, 0, 16)
    dm = fd.DoseMat(nvox)
    a, f, _ = store.request([0])
    dm.add_beam(ctx, a, f, cfg.extent)
    dm.close()
    dm = fd.DoseMat(nvox)
    torch.cuda.synchronize()
    lat = []
    t = time.perf_counter()
    for k in range(len(cfg.bundles)):
        a, f, _ = store.request([k])
        tb = time.perf_counter()
        dm.add_beam(ctx, a, f, cfg.extent)
        lat.append((time.perf_counter() - tb) * 1e3)
    s = time.perf_counter() - t
    nnz = dm.info()[2]
    dm.close()
    return {"beamlets_per_s": round(store.n / s, 1), "nnz": int(nnz),
            "beam_latency_ms": {"p50": round(percentile(lat, 50), 3), "p99": round(percentile(lat, 99), 3)}}


# ---------------------------------------------------------------------------- plain beams (configs 1, 2)

def plain_stage_run(fd, torch, np, cfg, ctx, beams, flu, steps):
    """The split stages (fd_terma, fd_cccs, fd_resample_csr) on one beam, steps times."""
    import ctypes as C

    dev = torch.device("cuda:0")
    n = len(beams)
    Lmax = max(b.long_dim for b in beams)
    stride = Lmax * 256
    flu_d = torch.from_numpy(np.ascontiguousarray(flu)).to(dev)
    T = torch.zeros(n * stride, dtype=torch.float32, device=dev)
    R = torch.zeros_like(T)
    D = torch.zeros_like(T)
    offsets = torch.zeros(n + 1, dtype=torch.int64, device=dev)
    lib = fd.load()
    arr = fd.beam_array(beams)
    st = {"cap": 0, "cols": None, "vals": None}

    def step():
        fd._lib.call("fd_terma", ctx.handle, arr, n, C.c_void_p(flu_d.data_ptr()), C.c_void_p(T.data_ptr()),
                     C.c_void_p(R.data_ptr()), stride, None)
        fd._lib.call("fd_cccs", ctx.handle, arr, n, C.c_void_p(T.data_ptr()), C.c_void_p(R.data_ptr()),
                     C.c_void_p(D.data_ptr()), stride, None)
        csr = fd.CSR(0, 0, st["cap"], offsets.data_ptr(), st["cols"].data_ptr() if st["cols"] is not None else None,
                     st["vals"].data_ptr() if st["vals"] is not None else None)
        rc = lib.fd_resample_csr(ctx.handle, arr, n, C.c_void_p(D.data_ptr()), stride, cfg.extent, C.byref(csr), None)
        if rc == fd._lib.FD_ERR_CAPACITY:
            st["cap"] = int(csr.nnz * 1.25) + 1024
            st["cols"] = torch.empty(st["cap"], dtype=torch.int64, device=dev)
            st["vals"] = torch.empty(st["cap"], dtype=torch.float32, device=dev)
            return step()
        fd._lib.check(rc, "fd_resample_csr")
        return int(csr.nnz)

    for _ in range(2):
        nnz = step()
        torch.cuda.synchronize()
    lat, kt = [], {}
    t_start = time.perf_counter()
    for _ in range(steps):
        t0 = time.perf_counter()
        nnz = step()
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t0) * 1e3)
        for name, ms in ctx.timings().items():
            kt.setdefault(name, []).append(ms)
    elapsed = time.perf_counter() - t_start
    off = offsets.cpu().numpy()
    cols, vals = st["cols"][:nnz].cpu().numpy(), st["vals"][:nnz].cpu().numpy()
    rows = [(cols[int(off[i]):int(off[i + 1])], vals[int(off[i]):int(off[i + 1])]) for i in range(n)]
    return {"beamlets_per_s": n * steps / elapsed, "ms_per_step": elapsed / steps * 1e3,
            "latency": lat, "kernels_ms": {k: float(np.mean(v)) for k, v in kt.items()}, "nnz": nnz, "rows": rows}


def leg_slab(fd, synth, torch, np, steps):
    """128^3 slab phantom, one seeded non-coplanar beam, 16 x 16 beamlets: plain FP32 stages."""
    cfg = synth.config2()
    beams, flu = fd.beamlets_init(cfg.bundles[0], cfg.flags[0], cfg.volume)
    n = len(beams)
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel, options=fd.FD_OPT_TIMING) as ctx:
        r = plain_stage_run(fd, torch, np, cfg, ctx, beams, flu, steps)
        nseg = int(nseg_of(ctx, fd, beams).sum())
    achieved = FLOPS_PER_SEGMENT * nseg / (r["kernels_ms"]["cccs"] * 1e-3) / 1e12
    idx = list(range(0, n, 32)) + [n - 1]
    acc = accuracy(cfg, [(beams[i], flu[i]) for i in idx], [r["rows"][i] for i in idx])
    log(f"slab: {r['beamlets_per_s']:.0f} beamlets/s, CCCS {r['kernels_ms']['cccs']:.3f} ms, "
        f"err vs FP64 {acc['max_err_rel_beamlet_max']:.2e}")
    return {
        "workload": "plain one-beam-per-call FP32 stages; 128^3 slab phantom (water/lung 0.25/bone 1.8/water) at "
                    "2.5 mm; 1 seeded non-coplanar beam (gantry %.1f, couch %.1f deg); 16x16 beamlets of 5 mm"
                    % (cfg.notes["gantry_deg"], cfg.notes["couch_deg"]),
        "beamlets": n, "beamlets_per_s": round(r["beamlets_per_s"], 1), "ms_per_beam": round(r["ms_per_step"], 3),
        "beam_latency_ms": {"p50": round(percentile(r["latency"], 50), 3),
                            "p99": round(percentile(r["latency"], 99), 3)},
        "kernels_ms": {k: round(v, 4) for k, v in r["kernels_ms"].items()},
        "cccs_roofline": {"achieved": round(achieved, 3), "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                          "traffic": profiled_traffic("cccs_kernel_slab128_traffic.md"),
                          "algorithmic": f"{FLOPS_PER_SEGMENT} flop x {nseg} (ray, segment) steps per launch"},
        "accuracy_vs_fp64": acc,
    }


def leg_small(fd, synth, torch, np, steps):
    """64^3 water, one beam at gantry 0, 8 x 8 beamlets: plain FP32 stages vs the FP64 oracle on every
    beamlet, and the oracle's own throughput on the host cores."""
    cfg = synth.config1()
    beams, flu = fd.beamlets_init(cfg.bundles[0], cfg.flags[0], cfg.volume)
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel, options=fd.FD_OPT_TIMING) as ctx:
        r = plain_stage_run(fd, torch, np, cfg, ctx, beams, flu, steps)
    items = [(beams[i], flu[i]) for i in range(len(beams))]
    acc = accuracy(cfg, items, r["rows"])
    cpu = cpu_baseline(cfg, items, "beamlets (all of the beam)")
    log(f"small: {r['beamlets_per_s']:.0f} beamlets/s, oracle {cpu['value']:.1f} beamlets/s, "
        f"err vs FP64 {acc['max_err_rel_beamlet_max']:.2e}")
    return {"workload": "64^3 water at 2.5 mm, 1 beam at gantry 0, 8x8 beamlets of 5 mm, plain FP32 stages",
            "beamlets": len(beams), "beamlets_per_s": round(r["beamlets_per_s"], 1),
            "ms_per_beam": round(r["ms_per_step"], 3), "accuracy_vs_fp64": acc, "cpu_oracle": cpu}


# ---------------------------------------------------------------------------- precision (config 4)

def leg_precision(fd, synth, torch, np, k=256, reps=10):
    """Slab beam with the 1e-3 relative threshold; per-beamlet max-dose error of the FP32, BF16 and FP16
    columns against the FP64 oracle (all 256 beamlets), and D^T X with 256 seeded U[0,1] fluence
    vectors (FP32 accumulation) against the same product through the oracle's matrix."""
    import scipy.sparse as sp

    cfg = synth.config2()
    beams, flu = fd.beamlets_init(cfg.bundles[0], cfg.flags[0], cfg.volume)
    n = len(beams)
    nvox = 128 ** 3
    refs = oracle_rows(cfg, [(beams[i], flu[i]) for i in range(n)])

    def csr_of(thr):
        ip, cs, vs = [0], [], []
        for ref in refs:
            c, v = ref["cols"], ref["vals"]
            if thr > 0:
                keep = v >= thr * v.max()
                c, v = c[keep], v[keep]
            cs.append(c)
            vs.append(v)
            ip.append(ip[-1] + len(c))
        return sp.csr_matrix((np.concatenate(vs), np.concatenate(cs), np.array(ip)), shape=(n, nvox))

    D_full, D_thr = csr_of(0.0), csr_of(1e-3)
    out = {"threshold": 1e-3, "rows": n, "k": k, "types": {}}
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        ctx.set_threshold(1e-3)
        dm = fd.dose_mat_construction(ctx, cfg.bundles, cfg.flags, cfg.extent)
        off, col, val = dm.matrix(0)
        rows_of = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
        gen = torch.Generator(device="cuda").manual_seed(4)
        X = torch.rand((n, k), dtype=torch.float32, device="cuda", generator=gen)
        sel = np.arange(0, k, 16)
        Xh = X.cpu().numpy().astype(np.float64)[:, sel]
        Yfull, Ythr = D_full.T @ Xh, D_thr.T @ Xh
        ymax = float(np.abs(Yfull).max())
        out["threshold_cost_DtX"] = float(np.abs(Ythr - Yfull).max()) / ymax
        out["nnz"] = int(len(col))
        for name, vt, vb in (("f32", fd.FD_VAL_F32, 4), ("bf16", fd.FD_VAL_BF16, 2), ("fp16", fd.FD_VAL_FP16, 2)):
            if vt == fd.FD_VAL_F32:
                v = val
            else:
                dm.pack(vt)
                scale, q = dm.packed()
                h = (q.astype(np.uint32) << 16).view(np.float32) if vt == fd.FD_VAL_BF16 else \
                    q.view(np.float16).astype(np.float32)
                v = h * scale[rows_of]
            worst = 0.0
            for i in range(n):
                a, b = int(off[i]), int(off[i + 1])
                worst = max(worst, row_error(col[a:b], v[a:b], refs[i])[0])
            with fd.Spmm(dm, vt) as p:
                Y = p.run(X)
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(reps):
                    p.run(X, Y)
                e1.record()
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / reps
                Yg = Y[:, torch.from_numpy(sel).cuda()].cpu().numpy().astype(np.float64)
            bytes_alg = nvox * k * 4 + len(col) * (vb + 4) + (nvox + 1) * 8 + n * k * 4
            out["types"][name] = {
                "beamlet_max_err_vs_fp64": worst, "within_0p5pct": worst <= TOL,
                "DtX_ms": round(ms, 4), "DtX_GB_per_s": round(bytes_alg / ms / 1e6, 1),
                "DtX_err_vs_fp64_thresholded": float(np.abs(Yg - Ythr).max()) / ymax,
                "DtX_err_vs_fp64_unthresholded": float(np.abs(Yg - Yfull).max()) / ymax}
        dm.close()
        # the FP64 radiological path: no ray of this beam is longer than 256 voxels, so the
        # FD_OPT_RPL_FP64 rule leaves it in FP32; FD_OPT_RPL_FP64_ALL forces FP64 on every ray
        idx = [0, 85, 170, 255]
        dev = torch.device("cuda:0")
        sub = fd.beam_array([beams[i] for i in idx])
        stride = max(b.long_dim for b in sub) * 256
        terma = {}
        for opt in (fd.FD_OPT_TIMING, fd.FD_OPT_TIMING | fd.FD_OPT_RPL_FP64_ALL):
            ctx.set_options(opt)
            T = torch.zeros(len(idx) * stride, device=dev)
            R = torch.zeros_like(T)
            ctx.terma(sub, torch.from_numpy(np.ascontiguousarray(flu[idx])).to(dev), T, R, stride)
            torch.cuda.synchronize()
            Th = T.cpu().numpy().reshape(len(idx), -1, 256)
            e = max(float(np.abs(Th[j, :sub[j].long_dim] - refs[i]["terma"]).max() / refs[i]["terma"].max())
                    for j, i in enumerate(idx))
            terma["fp64_all" if opt & fd.FD_OPT_RPL_FP64_ALL else "fp32"] = {
                "terma_ms_4_beamlets": round(ctx.timings()["terma"], 4), "terma_err_vs_fp64": e}
        out["radiological_path"] = {"longest_ray_voxels": int(max(b.long_dim for b in beams)), "terma": terma}
    log(f"precision: {json.dumps(out['types'])}")
    return out


# ---------------------------------------------------------------------------- serving (config 5)

def leg_serving(fd, synth, workload, torch, np, capacity_beamlets_s, n_req=60):
    """220 x 220 x 150 body with 1160 candidate beams; Poisson arrivals of 1-20-beam requests at
    several loads and batch sizes.  Per-beam latency = submit -> beam ready (library), request
    latency = submit -> fetched; each request is fetched as soon as it completes, inside the span."""
    import threading

    t0 = time.perf_counter()
    cfg = fourpi(synth, 5)
    ctx = fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel)
    flags = workload.device_flags(ctx, cfg, torch)
    store = workload.BeamStore(cfg, flags)
    nvox = cfg.volume.dim[0] * cfg.volume.dim[1] * cfg.volume.dim[2]
    log(f"serving: {len(cfg.bundles)} candidate beams, {store.n} beamlets, setup {time.perf_counter() - t0:.1f} s")
    rng = np.random.default_rng(1005)
    reqs = [list(rng.choice(len(cfg.bundles), size=int(rng.integers(1, 21)), replace=False)) for _ in range(n_req)]
    payload = [store.request(ks) for ks in reqs]
    mean_req = sum(sum(c) for _, _, c in payload) / n_req
    out = {"candidate_beams": len(cfg.bundles), "requests_per_run": n_req, "beams_per_request": [1, 20],
           "beamlets_per_beam": [int(store.counts.min()), int(store.counts.max())], "runs": []}
    # plain one beam per call on the first 40 beams of the requests
    flat = [k for ks in reqs for k in ks][:40]
    dm = fd.DoseMat(nvox)
    dm.add_beam(ctx, *store.request([flat[0]])[:2], cfg.extent)
    lat, nb = [], 0
    t = time.perf_counter()
    for k in flat:
        a, f, c = store.request([k])
        tb = time.perf_counter()
        dm.add_beam(ctx, a, f, cfg.extent)
        lat.append((time.perf_counter() - tb) * 1e3)
        nb += c[0]
    plain_s = time.perf_counter() - t
    plain_dm = dm
    out["plain"] = {"beamlets_per_s": round(nb / plain_s, 1), "beam_latency_ms": {
        "p50": round(percentile(lat, 50), 3), "p99": round(percentile(lat, 99), 3)}}
    first_req = None
    warm = sorted(range(n_req), key=lambda i: -sum(payload[i][2]))[:6]
    for batch, slots, loads in ((4096, 3, (0.5, 0.9, 1.5)), (1024, 3, (0.9,)), (2048, 3, (0.9,)), (1024, 2, (0.9,)),
                                (2048, 2, (0.9,))):
        srv = fd.Server(ctx, n_slots=slots, max_batch_beamlets=batch, extent=cfg.extent)
        # warm-up: the largest requests submitted together, so that the slots' batches reach full
        # size (and their scratch its working size) before the timed runs
        for w in [srv.submit(*payload[i]) for i in warm]:
            srv.fetch(w).close()
        for mult in loads:
            rate = mult * capacity_beamlets_s / mean_req  # requests per second
            gaps = rng.exponential(1.0 / rate, size=n_req)
            srv.reset_stats()
            ids, t_sub, t_got, mats = [None] * n_req, [0.0] * n_req, [0.0] * n_req, [None] * n_req
            submitted = [0]

            def producer():
                t_next = time.perf_counter()
                for i in range(n_req):
                    t_next += gaps[i]
                    d = t_next - time.perf_counter()
                    if d > 0:
                        time.sleep(d)
                    t_sub[i] = time.perf_counter()
                    ids[i] = srv.submit(*payload[i])
                    submitted[0] = i + 1

            th = threading.Thread(target=producer)
            t_begin = time.perf_counter()
            th.start()
            pending = set()
            nseen = 0
            while nseen < n_req or pending:  # fetch each request as soon as it completes
                while nseen < submitted[0]:
                    pending.add(nseen)
                    nseen += 1
                done = [i for i in sorted(pending) if srv.poll(ids[i]) != 0]
                for i in done:
                    m = fd.DoseMat(nvox)
                    srv.fetch(ids[i], m)
                    t_got[i] = time.perf_counter()
                    mats[i] = m
                    pending.discard(i)
                if not done:
                    time.sleep(2e-4)
            th.join()
            span = max(t_got) - min(t_sub)
            st = srv.stats()
            rl = [(b - a) * 1e3 for a, b in zip(t_sub, t_got)]
            out["runs"].append({
                "max_batch_beamlets": batch, "slots": slots, "load": mult, "requests_per_s": round(rate, 2),
                "beams": st["beams_done"], "beamlets": st["beamlets_done"],
                "beamlets_per_s": round(st["beamlets_done"] / span, 1),
                "beam_latency_ms": {"p50": round(st["p50_ms"], 3), "p99": round(st["p99_ms"], 3)},
                "request_latency_ms": {"p50": round(percentile(rl, 50), 3), "p99": round(percentile(rl, 99), 3)},
                "batches": st["batches"], "peak_inflight_batches": st["peak_inflight_batches"],
                "wall_s": round(time.perf_counter() - t_begin, 3)})
            log(f"serving batch {batch} slots {slots} load {mult}: {st['beamlets_done'] / span:.0f} beamlets/s, beam p50 "
                f"{st['p50_ms']:.1f} ms, p99 {st['p99_ms']:.1f} ms")
            if first_req is None:
                first_req = (reqs[0], mats[0])
                mats[0] = None
            for m in mats:
                if m is not None:
                    m.close()
        srv.close()
    # served beamlets of the first request: identical to the plain path, and against the FP64 oracle
    ks, m = first_req
    same = True
    pl = fd.DoseMat(nvox)
    for j, k in enumerate(ks):
        a, f, _ = store.request([k])
        pl.add_beam(ctx, a, f, cfg.extent)
        same = same and all(np.array_equal(x, y) for x, y in zip(m.matrix(j), pl.matrix(j)))
    pl.close()
    plain_dm.close()
    gl = [int(store.first[k]) + r for k in ks for r in range(int(store.counts[k]))]
    Ls = np.array([store.beams[g].long_dim for g in gl])
    idx = sorted(set(gl[:: max(1, len(gl) // 5)][:5]) | {gl[int(Ls.argmax())]})
    rows = []
    for g in idx:
        k, r = store.beam_of(g)
        off, col, val = m.matrix(ks.index(k))
        rows.append((col[int(off[r]):int(off[r + 1])], val[int(off[r]):int(off[r + 1])]))
    out["served_equal_plain_bitwise"] = bool(same)
    out["accuracy_vs_fp64"] = accuracy(cfg, [(store.beamlet(g), store.flu[g]) for g in idx], rows)
    m.close()
    ctx.close()
    return out


# ---------------------------------------------------------------------------- main

def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-sample", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="headline only (no plain/slab/small/precision/serving)")
    ap.add_argument("--only-slab", action="store_true", help="run only the 128^3 slab leg and print its JSON")
    ap.add_argument("--only-serving", type=float, default=0.0, metavar="CAPACITY",
                    help="run only the serving leg, loads relative to CAPACITY beamlets/s, and print its JSON")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 256 MiB device copy first (reference for HBM traffic counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the beamlet-dose path runs on a single MI355X in this build "
              "(beams shard across GPUs as independent processes); run with --gpus 1", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import fastdose_amd as fd
    from fastdose_amd import synth, workload

    if not torch.cuda.is_available():
        log("no HIP device")
        return 1
    if args.calib_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.only_slab:
        print(json.dumps({"slab": leg_slab(fd, synth, torch, np, args.steps)}), flush=True)
        return 0
    if args.only_serving > 0:
        print(json.dumps({"serving": leg_serving(fd, synth, workload, torch, np, args.only_serving)}), flush=True)
        return 0

    head, state = leg_headline(args, fd, synth, workload, torch, np)
    cfg, ctx, store, dm, arr, flu, cnt = state
    out = {
        "metric": "beamlets/s CCCS dose at <=0.5% err vs FP64; p50/p99 per-beam latency",
        "value": round(head["value"], 2),
        "unit": "beamlets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "batched 4pi FP32: 200x200x160 body at 2.5 mm (sigma-4 smoothed seeded noise -> 0.2-1.8 g/cc "
                        "in an ellipsoid, 0.0012 outside), 4 isocenters in an ellipsoidal PTV, 180 Fibonacci-sphere "
                        "beams, 20x20 fluence of 5 mm beamlets keeping those that intersect the PTV; 12-bin spectrum; "
                        "8 zenith x 8 azimuth cones; one step = one batched call over all 180 beams (one CSR matrix "
                        "per beam)",
            "beams_per_step": head["beams"],
            "beamlets_per_step": head["beamlets"],
            "beamlets_per_beam": head["beamlets_per_beam"],
            "long_dim_max": head["long_dim_max"],
            "nnz_per_step": head["nnz"],
            "tuning": head["tuning"],
        },
        "latency_ms": None,
        "latency_source": "per-beam latency (submit -> beam ready) of the serving workload at 0.9 x capacity, "
                          "batches of 4096 beamlets",
        "step_ms": head["step_ms"],
        "kernels_ms_per_step": head["kernels_ms_per_step"],
        "kernel_launches_per_step": head["kernel_launches_per_step"],
        "roofline": head["roofline"],
        "cpu_baseline": None,
    }
    # accuracy of the headline's own matrices on sampled beamlets (every 1/8 of the store, plus the longest)
    Ls = np.array([store.beams[g].long_dim for g in range(store.n)])
    idx = sorted(set(range(0, store.n, store.n // 8)) | {int(Ls.argmax())})
    out["accuracy_vs_fp64"] = accuracy(cfg, [(store.beamlet(g), store.flu[g]) for g in idx],
                                       matrix_rows(dm, store, idx))
    dm.close()
    log(f"headline accuracy: {out['accuracy_vs_fp64']}")
    if not args.no_cpu_baseline:
        sample = np.linspace(0, store.n - 1, args.cpu_sample).astype(int)
        out["cpu_baseline"] = cpu_baseline(cfg, [(store.beamlet(int(g)), store.flu[int(g)]) for g in sample],
                                           f"of the {store.n} beamlets (evenly spaced)")
        log(f"cpu baseline: {out['cpu_baseline']['value']} beamlets/s")
    if not args.no_extras:
        plain = leg_plain_4pi(fd, torch, np, state)
        plain["batched_over_plain"] = round(head["value"] / plain["beamlets_per_s"], 3)
        out["plain_4pi"] = plain
        log(f"plain 4pi: {plain['beamlets_per_s']} beamlets/s (batched x{plain['batched_over_plain']})")
    ctx.close()
    if not args.no_extras:
        out["slab"] = leg_slab(fd, synth, torch, np, 20)
        out["small_water"] = leg_small(fd, synth, torch, np, 10)
        out["precision"] = leg_precision(fd, synth, torch, np)
        out["serving"] = leg_serving(fd, synth, workload, torch, np, head["value"])
        run = next(r for r in out["serving"]["runs"]
                   if r["max_batch_beamlets"] == 4096 and r["slots"] == 3 and r["load"] == 0.9)
        out["latency_ms"] = run["beam_latency_ms"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
