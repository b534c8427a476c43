"""Benchmark: fixed-rate 512^3 f32 ZFP compress + decompress on one MI355X.

A step is one compress and one decompress of the whole field through the C
ABI (zfp_compress / zfp_decompress), field and stream resident in HBM.  value
is uncompressed GB/s over the step: 2 x 512 MiB of field data handled per
step (compress reads it, decompress writes it) / wall time per step.

Also reported:
  roofline      the dominant kernel (the slower of encode and decode):
                algorithmic bytes per launch (field bytes + stream bytes) /
                its average duration from HIP events recorded on the stream
                the kernel runs on, against MI355X's 8 TB/s HBM peak, and
                against the measured ceiling of a device-to-device copy of
                the field (copy_ceiling, frac_of_copy);
  cpu_baseline  the oracle's threaded CPU codec (C-OMP chunked compress,
                per-chunk-seek fixed-rate decompress) on a 512x512x64 slab of
                the same field, timed on this host, and its stream checked
                byte for byte against the GPU's;
  batched       4096 mixed arrays (shapes, types, modes) compressed and
                decompressed by the batched entry points: GB/s, the step's
                p50/p99 latency, and the same work as one call per array;
  serving       Poisson arrivals (1000 req/s, 16 KiB-256 MiB arrays) through
                the continuous-batching loop and one call per request:
                GB/s and latency p50/p99; the same trace at 200 and 2000
                req/s, and with every request arriving at once;
  variable_rate the variable-rate configurations (384^3 f64 accuracy and
                precision, 8192^2 f32 precision, 256^3 i32 reversible):
                compress and decompress GB/s through the offset index;
  four_d        4-D fields (64^4 f32 rate 8 and precision 16, 32^4 f64
                reversible): compress and decompress GB/s.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--rate 8] [--n 512]
"""
import argparse
import ctypes
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
for p in (ROOT / "zfp-rs_amd", ROOT / "oracle"):
    sys.path.insert(0, str(p))

METRIC = "compress+decompress GB/s (uncompressed), 512^3 f32 rate 8; batched p99 ms"
HBM_PEAK_GBPS = 8000.0


def cpu_baseline(field_gpu, rate, gpu_words_fn, budget_s=12.0):
    """Time the oracle (threaded) on a slab; returns the cpu_baseline object."""
    import numpy as np

    import oracle_py as O

    slab = field_gpu[:64].contiguous().cpu().numpy()
    cfg = O.fixed_rate(rate, O.FLOAT, 3)
    threads = max(1, min(16, os.cpu_count() or 1))
    # the GPU stream of the same slab must be byte-identical
    on, oerr, ow = O.compress(slab, cfg, threads=threads)
    gw = gpu_words_fn(slab)
    same = oerr == 0 and np.array_equal(gw[: on // 8], ow[: on // 8])
    best = None
    t_end = time.perf_counter() + budget_s
    reps = 0
    while reps < 2 or (time.perf_counter() < t_end and reps < 20):
        t0 = time.perf_counter()
        n, err, words = O.compress(slab, cfg, threads=threads)
        t1 = time.perf_counter()
        out, m, derr = O.decompress(words, slab.shape, slab.dtype, cfg, threads=threads)
        t2 = time.perf_counter()
        assert err == 0 and derr == 0
        step = t2 - t0
        best = step if best is None or step < best else best
        reps += 1
    gbps = 2 * slab.nbytes / best / 1e9
    # the same work on one thread (best of 2)
    best1 = None
    for _ in range(2):
        t0 = time.perf_counter()
        n1, err1, words1 = O.compress(slab, cfg, threads=1)
        O.decompress(words1, slab.shape, slab.dtype, cfg, threads=1)
        step1 = time.perf_counter() - t0
        best1 = step1 if best1 is None or step1 < best1 else best1
    import platform
    cpu_model = platform.processor() or "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": round(gbps, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "single_thread_value": round(2 * slab.nbytes / best1 / 1e9, 3),
            "sample": f"512x512x64 f32 slab of the same field, rate {rate}, compress+decompress, "
                      f"best of {reps}, oracle threaded C code, {threads} threads on {cpu_model} "
                      f"(nproc {os.cpu_count()}); the oracle is a plain restatement, about 3x "
                      f"slower per core than zfp-rs's published serial rates (339/390 Melem/s "
                      f"compress/decompress at rate 8, 3-D f32), so this understates zfp-rs",
            "stream_matches_gpu": bool(same)}


def batched_leg(n_arrays, reps, seed=4096):
    """4096 independent arrays (dims 1-3 uniform, each axis log-uniform in
    [3, 300], type uniform over f32/f64/i32/i64, mode uniform over fixed rate
    4-24, fixed precision 8-48 and fixed accuracy 1e-9..1e-2 log-uniform,
    integers coded at full precision under fixed accuracy, each filled from
    its own seed with the k^-3 field), device-resident.  One step = compress
    then decompress every array.  Times the batched entry points against one
    zfp_compress / zfp_decompress call per array, checks that both give the
    same bytes, and reports the batched step's p50/p99 over `reps` steps; then
    the same batch planned once (zfp_hip_batch_plan: tiling tuned, device
    work captured into a HIP graph) and replayed, with its own p50/p99."""
    import math

    import numpy as np
    import torch

    import zfp_hip as Z
    from synth import batch_workload

    def note(msg):
        print(f"[batched] {msg}", file=sys.stderr, flush=True)

    arrays, outs, streams, bufs = [], [], [], []
    field_bytes = 0
    classes = set()
    for x, cfg, mode in batch_workload(n_arrays, seed, "cuda", note):
        cap = Z.maximum_size(cfg, x)
        buf = torch.zeros((cap + 7) // 8, dtype=torch.int64, device="cuda")
        arrays.append(x)
        outs.append(torch.empty_like(x))
        bufs.append(buf)
        streams.append(Z.Stream(buf, cfg))
        field_bytes += x.numel() * x.element_size()
        classes.add((str(x.dtype), x.dim(), mode == 0))
    torch.cuda.synchronize()
    note(f"{n_arrays} arrays, {field_bytes / 1e9:.2f} GB")

    def rewind():
        for st in streams:
            st.rewind()

    # one call per array
    single_ms = []
    for _ in range(2):
        rewind()
        t0 = time.perf_counter()
        sizes = [st.compress(x) for st, x in zip(streams, arrays)]
        rewind()
        for st, o in zip(streams, outs):
            st.decompress(o)
        torch.cuda.synchronize()
        single_ms.append((time.perf_counter() - t0) * 1e3)
        note(f"one call per array: {single_ms[-1]:.1f} ms")
    sample = list(range(0, n_arrays, max(1, n_arrays // 64)))
    ref_words = {k: bufs[k][: sizes[k] // 8].cpu().clone() for k in sample}
    ref_out = {k: outs[k].cpu().clone() for k in sample}
    stream_bytes = sum(sizes)

    batch = Z.Batch(streams, arrays, outs)
    step_ms, comp_ms, decomp_ms = [], [], []
    ok = True
    for r in range(reps + 2):
        rewind()
        for k in sample:
            bufs[k].zero_()
            outs[k].zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nc = batch.compress()
        t1 = time.perf_counter()
        ok = ok and nc == n_arrays and list(batch.sizes) == sizes
        rewind()
        t2 = time.perf_counter()
        nd = batch.decompress()
        t3 = time.perf_counter()
        ok = ok and nd == n_arrays
        note(f"batched step {r}: {(t1 - t0 + t3 - t2) * 1e3:.2f} ms")
        if r >= 2:
            comp_ms.append((t1 - t0) * 1e3)
            decomp_ms.append((t3 - t2) * 1e3)
            step_ms.append((t1 - t0 + t3 - t2) * 1e3)
    # the batched bytes and values equal the single calls'
    rewind()
    batch.compress()
    same = ok and all(bufs[k][: sizes[k] // 8].cpu().equal(ref_words[k]) for k in sample)
    rewind()
    batch.decompress()
    torch.cuda.synchronize()
    same = same and all(outs[k].cpu().equal(ref_out[k]) for k in sample)

    # the same batch planned once (tiling tuned by timing it) and replayed
    # from captured graphs, one for compress and one for decompress
    rewind()
    t0 = time.perf_counter()
    enc_plan = Z.BatchPlan(True, streams, arrays)
    rewind()
    dec_plan = Z.BatchPlan(False, streams, outs)
    plan_ms = (time.perf_counter() - t0) * 1e3
    note(f"planned: encode tiling {enc_plan.tiling()}, decode tiling {dec_plan.tiling()}, "
         f"{plan_ms:.0f} ms to plan")
    graph_ms, gcomp_ms, gdecomp_ms = [], [], []
    gok = True
    for r in range(reps + 2):
        for k in sample:
            bufs[k].zero_()
            outs[k].zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nc, gsizes = enc_plan.run()
        t1 = time.perf_counter()
        nd, _ = dec_plan.run()
        t2 = time.perf_counter()
        gok = gok and nc == n_arrays and nd == n_arrays and gsizes == sizes
        note(f"graph step {r}: {(t2 - t0) * 1e3:.2f} ms")
        if r >= 2:
            gcomp_ms.append((t1 - t0) * 1e3)
            gdecomp_ms.append((t2 - t1) * 1e3)
            graph_ms.append((t2 - t0) * 1e3)
    torch.cuda.synchronize()
    gsame = gok and all(bufs[k][: sizes[k] // 8].cpu().equal(ref_words[k]) for k in sample)
    gsame = gsame and all(outs[k].cpu().equal(ref_out[k]) for k in sample)
    et, es, _ = enc_plan.tiling()
    dt_, ds, _ = dec_plan.tiling()
    enc_plan.close()
    dec_plan.close()
    for st in streams:
        st.close()
    best_single = min(single_ms)
    p50 = float(np.percentile(step_ms, 50))
    p99 = float(np.percentile(step_ms, 99))
    g50 = float(np.percentile(graph_ms, 50))
    g99 = float(np.percentile(graph_ms, 99))
    res = {
        "arrays": n_arrays, "classes": len(classes), "uncompressed_bytes": field_bytes,
        "compressed_bytes": stream_bytes,
        "batched_GBps": round(2 * field_bytes / (p50 * 1e-3) / 1e9, 2),
        "per_call_GBps": round(2 * field_bytes / (best_single * 1e-3) / 1e9, 2),
        "speedup_vs_per_call": round(best_single / p50, 2),
        "step_p50_ms": round(p50, 3), "step_p99_ms": round(p99, 3),
        "compress_p50_ms": round(float(np.percentile(comp_ms, 50)), 3),
        "decompress_p50_ms": round(float(np.percentile(decomp_ms, 50)), 3),
        "per_call_step_ms": round(best_single, 3), "steps": reps,
        "bytes_match_per_call": bool(same),
        "graph": {
            "GBps": round(2 * field_bytes / (g50 * 1e-3) / 1e9, 2),
            "speedup_vs_per_call": round(best_single / g50, 2),
            "step_p50_ms": round(g50, 3), "step_p99_ms": round(g99, 3),
            "compress_p50_ms": round(float(np.percentile(gcomp_ms, 50)), 3),
            "decompress_p50_ms": round(float(np.percentile(gdecomp_ms, 50)), 3),
            "plan_ms": round(plan_ms, 1),
            "tiling": {"compress": {"blocks_per_workgroup": et, "lds_slice": es},
                       "decompress": {"blocks_per_workgroup": dt_, "lds_slice": ds}},
            "bytes_match_per_call": bool(gsame),
        },
    }
    del arrays, outs, bufs, batch
    torch.cuda.empty_cache()
    return res


def serving_leg(n_requests, rate, seed=2000):
    """Poisson arrivals of compress and decompress requests (50/50), array
    sizes log-uniform from 16 KiB to 256 MiB with the batched leg's mix of
    dimensionality, type and mode, fields and streams resident in HBM.  The
    same trace is replayed through the continuous-batching loop (admission by
    an HBM budget of half the free memory) with one batch at a time and with
    two batches in flight on separate HIP streams, and one call per request;
    also replayed with every request arriving at once (throughput when
    saturated).  Reports uncompressed GB/s and latency p50/p99."""
    import math

    import numpy as np
    import torch

    import serve
    import zfp_hip as Z
    from synth import serving_workload

    def note(msg):
        print(f"[serving] {msg}", file=sys.stderr, flush=True)

    trace = []
    for kind, x, cfg, t in serving_workload(n_requests, rate, seed):
        esz = x.element_size()
        cap = Z.maximum_size(cfg, x)
        buf = torch.zeros((cap + 7) // 8, dtype=torch.int64, device="cuda")
        st = Z.Stream(buf, cfg)
        if kind == "decompress":
            st.compress(x)  # the stream a decompress request reads
            st.rewind()
            req = serve.Request("decompress", st, torch.empty_like(x), x.numel() * esz, cap, t)
        else:
            req = serve.Request("compress", st, x, x.numel() * esz, x.numel() * esz + cap, t)
        trace.append(req)
    torch.cuda.synchronize()
    total = sum(r.nbytes for r in trace)
    note(f"{n_requests} requests, {total / 1e9:.2f} GB, {rate:g} req/s")
    budget = torch.cuda.mem_get_info()[0] // 2
    codec = serve.GpuCodec()
    arrivals = [r.arrival for r in trace]

    def replay(fn, at_once=False):
        for r, a in zip(trace, arrivals):
            r.stream.rewind()
            r.done = None
            r.arrival = 0.0 if at_once else a
        torch.cuda.synchronize()
        extra = fn()
        torch.cuda.synchronize()
        return extra

    replay(lambda: serve.serve_batched(trace, codec, budget))  # warm the workspace
    stats = replay(lambda: serve.serve_batched(trace, codec, budget))
    batched = serve.summary(trace)
    batched.update(batches=stats["batches"], max_batch=stats["max_batch"])
    note(f"batched: {batched}")
    # two batches in flight on separate HIP streams
    replay(lambda: serve.serve_batched(trace, codec, budget, inflight=2))  # warm the pool
    ostats = replay(lambda: serve.serve_batched(trace, codec, budget, inflight=2))
    overlapped = serve.summary(trace)
    overlapped.update(batches=ostats["batches"], max_batch=ostats["max_batch"],
                      submitted_while_another_ran=ostats["overlapped"])
    note(f"overlapped: {overlapped}")
    replay(lambda: serve.serve_one_at_a_time(trace, codec))
    single = serve.summary(trace)
    note(f"one call per request: {single}")
    # the same trace at the other ends of the arrival-rate range (arrival
    # times scaled), batched and one call per request
    sweep = []
    for r_rps in (200.0, 2000.0):
        scaled = [a * rate / r_rps for a in arrivals]

        def replay_at(fn):
            for r, a in zip(trace, scaled):
                r.stream.rewind()
                r.done = None
                r.arrival = a
            torch.cuda.synchronize()
            fn()
            torch.cuda.synchronize()

        replay_at(lambda: serve.serve_batched(trace, codec, budget))
        sb = serve.summary(trace)
        replay_at(lambda: serve.serve_one_at_a_time(trace, codec))
        ss = serve.summary(trace)
        sweep.append({"rate_rps": r_rps, "batched": sb, "per_request": ss})
        note(f"{r_rps:g} req/s: batched {sb}, one call per request {ss}")
    replay(lambda: serve.serve_batched(trace, codec, budget), at_once=True)
    sat_b = serve.summary(trace)
    replay(lambda: serve.serve_batched(trace, codec, budget, inflight=2), at_once=True)
    sat_o = serve.summary(trace)
    replay(lambda: serve.serve_one_at_a_time(trace, codec), at_once=True)
    sat_s = serve.summary(trace)
    note(f"saturated: batched {sat_b['GBps']} GB/s, two in flight {sat_o['GBps']} GB/s, "
         f"one call per request {sat_s['GBps']} GB/s")
    for r in trace:
        r.stream.close()
    del trace
    torch.cuda.empty_cache()
    return {"rate_rps": rate, "requests": n_requests, "uncompressed_bytes": total,
            "hbm_budget_bytes": budget, "batched": batched, "overlapped": overlapped,
            "per_request": single, "rate_sweep": sweep,
            "saturated": {"batched_GBps": sat_b["GBps"], "overlapped_GBps": sat_o["GBps"],
                          "per_request_GBps": sat_s["GBps"],
                          "speedup": round(max(sat_b["GBps"], sat_o["GBps"]) / sat_s["GBps"], 2),
                          "batched_p99_ms": sat_b["p99_ms"], "overlapped_p99_ms": sat_o["p99_ms"]}}


def variable_rate_leg(steps=5):
    """The variable-rate configurations: 384^3 f64 k^-3 field scaled into
    [-1e3, 1e3] at fixed accuracy 1e-6 and fixed precision 32; 8192^2 f32 at
    fixed precision 20; 256^3 i32 random walk in reversible mode.  Compress
    (length pass, scan, write pass; the offset index kept) and decompress
    through the index; best step of `steps`; round trip checked (reversible
    exact, fixed accuracy within the tolerance)."""
    import numpy as np
    import torch

    import zfp_hip as Z
    from synth import grf

    def note(msg):
        print(f"[variable] {msg}", file=sys.stderr, flush=True)

    cases = []
    v1 = (grf((384, 384, 384), seed=11, dtype=np.float64, device="cuda") * 300.0).clamp_(-1e3, 1e3)
    cases.append(("384^3 f64 fixed-accuracy 1e-6", v1, Z.ZfpConfig.fixed_accuracy(1e-6), 1e-6))
    cases.append(("384^3 f64 fixed-precision 32", v1, Z.ZfpConfig.fixed_precision(32), None))
    v2 = grf((8192, 8192), seed=12, dtype=np.float32, device="cuda")
    cases.append(("8192^2 f32 fixed-precision 20", v2, Z.ZfpConfig.fixed_precision(20), None))
    g = torch.Generator(device="cuda")
    g.manual_seed(13)
    v3 = torch.randint(-3, 4, (256 ** 3,), generator=g, device="cuda", dtype=torch.int32)
    v3 = torch.cumsum(v3, 0, dtype=torch.int32).reshape(256, 256, 256)
    cases.append(("256^3 i32 random walk reversible", v3, Z.ZfpConfig.reversible(), 0))
    out = _time_cases(cases, steps, note, indexless=True)
    del v1, v2, v3, cases
    torch.cuda.empty_cache()
    return out


def four_d_leg(steps=3):
    """4-D fields: 64^4 f32 (a smooth separable field plus 1e-3 noise) at fixed
    rate 8 and fixed precision 16, and 32^4 f64 reversible; compress and
    decompress GB/s as for the variable-rate configurations.  4-D blocks run
    one wave per SIMD, a 32-bit block's 256 values in registers."""
    import torch

    import zfp_hip as Z

    def note(msg):
        print(f"[four_d] {msg}", file=sys.stderr, flush=True)

    t = torch.linspace(0.0, 6.283185307179586, 64, device="cuda")
    g = torch.Generator(device="cuda")
    g.manual_seed(44)
    v = (torch.sin(t)[None, None, None, :] * torch.cos(2 * t)[None, None, :, None]
         + torch.sin(3 * t)[None, :, None, None] * torch.cos(t)[:, None, None, None])
    v = (v + 1e-3 * torch.randn(v.shape, generator=g, device="cuda")).contiguous()
    w = v[:32, :32, :32, :32].to(torch.float64).contiguous()
    cases = [("64^4 f32 fixed-rate 8", v, Z.ZfpConfig.fixed_rate(8.0, Z.TYPE_FLOAT, 4), None),
             ("64^4 f32 fixed-precision 16", v, Z.ZfpConfig.fixed_precision(16), None),
             ("32^4 f64 reversible", w, Z.ZfpConfig.reversible(), 0)]
    out = _time_cases(cases, steps, note)
    del v, w, cases
    torch.cuda.empty_cache()
    return out


def _time_cases(cases, steps, note, indexless=False):
    """Best compress and decompress time of `steps` (after one untimed step)
    for each (name, device tensor, config, tolerance) case; tolerance 0 checks
    an exact round trip, a float checks the largest error.  Each direction's
    roofline fraction is its algorithmic bytes (field plus stream) per second
    over the HBM peak.  `indexless` also times one decode of the same words
    without their offset index (block lengths read on the host, runs decoded
    on the GPU as they are found) and checks it equals the indexed decode."""
    import torch

    import zfp_hip as Z

    out = []
    for name, x, cfg, tol in cases:
        cap = Z.maximum_size(cfg, x)
        buf = torch.zeros((cap + 7) // 8, dtype=torch.int64, device="cuda")
        y = torch.empty_like(x)
        s = Z.Stream(buf, cfg)
        best_c = best_d = None
        for _ in range(steps + 1):
            s.rewind()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            n = s.compress(x)
            t1 = time.perf_counter()
            s.rewind()
            s.decompress(y)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            best_c = t1 - t0 if best_c is None else min(best_c, t1 - t0)
            best_d = t2 - t1 if best_d is None else min(best_d, t2 - t1)
        nb = x.numel() * x.element_size()
        if tol == 0:
            ok = bool(torch.equal(x, y))
        elif tol is not None:
            ok = bool((x - y).abs().max().item() <= tol)
        else:
            ok = True
        s.close()
        r = {"config": name, "uncompressed_bytes": nb, "stream_bytes": n,
             "ratio": round(nb / n, 3), "compress_GBps": round(nb / best_c / 1e9, 2),
             "decompress_GBps": round(nb / best_d / 1e9, 2), "round_trip_ok": ok,
             "roofline": {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBPS,
                          "traffic_per_call": nb + n,
                          "compress_achieved": round((nb + n) / best_c / 1e9, 1),
                          "compress_frac": round((nb + n) / best_c / 1e9 / HBM_PEAK_GBPS, 4),
                          "decompress_achieved": round((nb + n) / best_d / 1e9, 1),
                          "decompress_frac": round((nb + n) / best_d / 1e9 / HBM_PEAK_GBPS, 4)}}
        if indexless:
            y2 = torch.empty_like(x)
            s2 = Z.Stream(buf, cfg)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            s2.decompress(y2)
            torch.cuda.synchronize()
            t_il = time.perf_counter() - t0
            s2.close()
            iv = torch.int32 if x.element_size() == 4 else torch.int64
            r["indexless_decompress_ms"] = round(t_il * 1e3, 1)
            r["indexed_decompress_ms"] = round(best_d * 1e3, 3)
            r["indexless_decompress_GBps"] = round(nb / t_il / 1e9, 2)
            r["indexless_equals_indexed"] = bool(torch.equal(y2.view(iv), y.view(iv)))
            del y2
        note(str(r))
        out.append(r)
        del buf, y
    return out


KERNELS = {"encode": "zx::k_encode_fx<float, 3, false, true>",
           "decode": "zx::k_decode_fx<float, 3, false, true>"}


def measured_traffic(which, n, rate):
    """HBM bytes per launch of the dominant kernel from the committed counter
    profile (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes, each scaled by
    what the counter reads for a 512 MiB device copy), or None when no
    profile covers this workload."""
    path = ROOT / "profiles" / f"traffic_{n}cube_f32_rate{rate:g}.json"
    if not path.exists():
        return None
    prof = json.loads(path.read_text())
    k = prof["kernels"].get(KERNELS[which])
    if k is None:
        return None
    return k["hbm_bytes"], f"{path.relative_to(ROOT)} ({KERNELS[which]})"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--rate", type=float, default=8.0)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batched-arrays", type=int, default=4096,
                    help="arrays in the batched leg (0 skips it)")
    ap.add_argument("--batched-steps", type=int, default=20)
    ap.add_argument("--serving-requests", type=int, default=600,
                    help="requests in the serving leg (0 skips it)")
    ap.add_argument("--serving-rate", type=float, default=1000.0)
    ap.add_argument("--no-variable-rate", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one device-to-device copy of the field (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the codec runs on a single MI355X here and "
              "no multi-GPU figure is produced; use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import zfp_hip as Z
    from synth import grf

    torch.cuda.set_device(0)
    n = args.n
    shape = (n, n, n)
    x = grf(shape, seed=2024, zero_blocks=0.05, subnormals=0.01, device="cuda")
    assert x.dtype == torch.float32 and x.is_contiguous()
    out = torch.empty_like(x)
    cfg = Z.ZfpConfig.fixed_rate(args.rate, Z.TYPE_FLOAT, 3)
    nbytes_max = Z.maximum_size(cfg, x)
    buf = torch.zeros((nbytes_max + 7) // 8, dtype=torch.int64, device="cuda")
    stream = torch.cuda.current_stream()
    s = Z.Stream(buf, cfg)
    s.set_hip_stream(stream.cuda_stream)
    s.set_timing(True)

    enc_ms, dec_ms = [], []
    # the steps call the C ABI directly, as a C caller would: fields described
    # once, then zfp_stream_rewind + zfp_compress / zfp_decompress per step
    L = Z.lib()
    zs = s.zs
    fin, fout = Z.make_field(x), Z.make_field(out)
    pin, pout = ctypes.byref(fin), ctypes.byref(fout)
    call_ms = ctypes.c_double(0)

    def step(record):
        L.zfp_stream_rewind(zs)
        c = L.zfp_compress(zs, pin)
        if record:
            enc_ms.append(L.zfp_stream_kernel_ms(zs, ctypes.byref(call_ms)))
        L.zfp_stream_rewind(zs)
        d = L.zfp_decompress(zs, pout)
        if record:
            dec_ms.append(L.zfp_stream_kernel_ms(zs, ctypes.byref(call_ms)))
        if not (c and d):
            raise RuntimeError("zfp_compress / zfp_decompress failed")
        return c, d

    print(f"[main] field ready, {args.warmup} warmup + {args.steps} steps", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        c_bytes, d_bytes = step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        c_bytes, d_bytes = step(True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()

    if args.calib_copy:
        out.copy_(x)
        torch.cuda.synchronize()

    # the achievable ceiling: a device-to-device copy of the field (the same
    # bytes read and written as a step's worth of one direction), best of 5
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    copy_ms = None
    for _ in range(5):
        ev0.record()
        out.copy_(x)
        ev1.record()
        ev1.synchronize()
        ms = ev0.elapsed_time(ev1)
        copy_ms = ms if copy_ms is None else min(copy_ms, ms)

    field_bytes = x.numel() * 4
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = 2 * field_bytes / (ms_per_step * 1e-3) / 1e9
    stream_bytes = c_bytes
    enc = float(np.mean(enc_ms))
    dec = float(np.mean(dec_ms))
    alg = field_bytes + stream_bytes  # per launch, either direction
    dominant = "encode" if enc >= dec else "decode"
    dom_ms = max(enc, dec)
    achieved = alg / (dom_ms * 1e-3) / 1e9
    traffic = measured_traffic(dominant, n, args.rate)
    # sanity: a decode of the compressed stream round-trips the block count
    blocks = ((n + 3) // 4) ** 3
    assert stream_bytes == (blocks * cfg.maxbits + 63) // 64 * 8

    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded k^-3 Gaussian random field, unit variance + 1e-3 noise, "
                "5% zero blocks, 1% subnormals; field and stream resident in HBM",
        "config": {"workload": f"{n}^3 f32 fixed-rate {args.rate:g} compress+decompress "
                               "(zfp_compress/zfp_decompress, device-resident)",
                   "rate": args.rate, "shape": list(shape), "rounding": "Never",
                   "stream_bytes": stream_bytes},
        "compress_GBps": round(field_bytes / (enc * 1e-3) / 1e9, 2),
        "decompress_GBps": round(field_bytes / (dec * 1e-3) / 1e9, 2),
        "encode_kernel_ms": round(enc, 4),
        "decode_kernel_ms": round(dec, 4),
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": traffic[0] if traffic else None,
                     "traffic_source": traffic[1] if traffic else None,
                     "algorithmic_bytes_per_launch": alg,
                     "copy_ceiling": round(2 * field_bytes / (copy_ms * 1e-3) / 1e9, 1),
                     "frac_of_copy": round(achieved / (2 * field_bytes / (copy_ms * 1e-3) / 1e9), 4)},
        "cpu_baseline": None,
    }

    print(f"[main] encode {enc:.4f} ms, decode {dec:.4f} ms", file=sys.stderr, flush=True)
    if args.batched_arrays > 0:
        result["batched"] = batched_leg(args.batched_arrays, args.batched_steps)
    if args.serving_requests > 0:
        result["serving"] = serving_leg(args.serving_requests, args.serving_rate)
    if not args.no_variable_rate:
        result["variable_rate"] = variable_rate_leg()
        result["four_d"] = four_d_leg()

    if not args.no_cpu_baseline:
        def gpu_words(slab):
            b, nb = Z.compress(slab, cfg)
            return b
        result["cpu_baseline"] = cpu_baseline(x, args.rate, gpu_words)
    s.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
