#!/usr/bin/env python3
"""Benchmark: reads/s classified (D=10240, k=12, 100k references, top-5).

One step = encode a batch of reads on the GPU (HVCache::compute semantics)
and search it against the 100k reference prototypes with the MFMA form,
returning the top-5 (index, cosine) per read.  Inputs are resident in HBM
before the timed region; the reference database is encoded once, outside it.

Also measured in the same run: the plain (VALU) form over the same reads
(median of three runs), the main-kernel durations (HIP events on the launch
stream) for the roofline, the CPU baseline (the C restatement of the
reference's bit-packed path, timed on this host) on a bounded sample, whose
top-5 must equal the GPU's, the board's power and graphics clock on two
untimed steps (amd-smi), the vendor library's int8 GEMM (torch._int_mm on
hipBLASLt) at K = D on the same card, and the card's int8 MFMA rate on
operands spread like the search's rows with no memory traffic
(hvk_calib_mfma_i8), the two rates this card sustains under its power cap.
After the headline line is measured, the default run also records the
batching form (1M reads of 100-1,600 nt) and the hierarchical descent
(262,144 reads) at reduced size under "also" (--no-extras skips them);
--workload runs each at full size on its own.
"""
import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bacterial-classification_amd", "python"))

DIM, KMER, TOPK = 10240, 12, 5
MFMA_INT8_PEAK_TOPS = 5000.0  # MI355X dense int8 MFMA, TOPS
HBM_PEAK_GBS = 8000.0
PMC_PROFILE = os.path.join(ROOT, "profiles", "mfma_kernel_config1_pmc.md")


def pmc_traffic(kernel, reads, refs):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (copy-calibrated); None when the
    profile does not cover this kernel and size."""
    try:
        text = open(PMC_PROFILE).read()
    except OSError:
        return None
    for m in re.finditer(r"traffic_per_launch: kernel=(\S+) reads=(\d+) refs=(\d+) bytes=(\d+)", text):
        if m.group(1) == kernel and int(m.group(2)) == reads and int(m.group(3)) == refs:
            return float(m.group(4))
    return None


VALU_PROFILE = os.path.join(ROOT, "profiles", "encode_config1_pmc.md")
NOMINAL_CLOCK_GHZ = 2.4  # MI355X peak engine clock


def valu_counters(kernel, reads):
    """Per-launch SQ counters of `kernel` on `reads` reads from the committed
    rocprofv3 --pmc summary, or None when it does not cover them."""
    try:
        text = open(VALU_PROFILE).read()
    except OSError:
        return None
    for m in re.finditer(r"valu_per_launch: kernel=(\S+) reads=(\d+) (.*)", text):
        if m.group(1) == kernel and int(m.group(2)) == reads:
            return {k: float(v) for k, v in re.findall(r"(\w+)=([0-9.eE+]+)", m.group(3))}
    return None


def cpu_rows(w, n_take, db_host, dbn, n_refs, dim, kmer, topk, gpu_idx, cpu_seconds):
    """The CPU baseline: the C restatement of the reference's bit-packed path
    (ctz bundling of on-the-fly bound k-mer HVs, the int64 cosine loop, top-k)
    built here with the reference's Release flags for this host's CPU, timed
    on one thread and on every core this process may use (at most 16, the
    box's CPU share per GPU), each on the first reads of the workload (sample
    sized for about cpu_seconds / 2 each).  Top-k must equal the GPU's."""
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as orc
    tmp = tempfile.mkdtemp(prefix="hvk_cpu_")
    O, flags = orc.release_classifier(tmp)
    if O is None:
        O, flags = orc.oracle(), "prebuilt liboracle.so (-O3 -march=x86-64-v3): " + flags
    nproc, avail, model = orc.host_cpu()
    n_all = max(1, min(16, avail))

    def run(n, threads):
        seqs = w.reads[: w.read_offsets[n]]
        offs = np.ascontiguousarray(w.read_offsets[: n + 1])
        oi = np.zeros((n, topk), np.int32)
        oc = np.zeros((n, topk), np.float64)
        sec = O.orc_classify_reads(dim, kmer, seqs, offs.ctypes.data, n, db_host.ctypes.data,
                                   dbn.ctypes.data, n_refs, topk, oi.ctypes.data, oc.ctypes.data,
                                   threads)
        return sec, oi

    rows = {}
    for name, threads in (("cpu_1", 1), ("cpu_all", n_all)):
        t1, _ = run(min(threads, n_take), threads)  # one read per thread: calibrates the sample
        n = int(max(min(threads, n_take),
                    min(n_take, (cpu_seconds / 2.0) / t1 * threads // threads * threads)))
        sec, oi = run(n, threads)
        rows[name] = {"value": n / sec, "unit": "reads/s", "threads": threads, "reads": n,
                      "seconds": sec, "topk_equal_gpu": bool(np.array_equal(oi, gpu_idx[:n]))}
        print(f"[bench] {name}: {n / sec:.2f} reads/s ({n} reads, {threads} threads, {sec:.1f}s)",
              file=sys.stderr, flush=True)
    a = rows["cpu_all"]
    return {"value": a["value"], "unit": "reads/s", "cores": n_all, "kind": "port",
            "sample": f"first {a['reads']} reads x all {n_refs} refs, top-{topk} ({a['seconds']:.1f}s "
                      f"wall, {n_all} threads); 1-thread row: first {rows['cpu_1']['reads']} reads",
            "flags": flags, "nproc": nproc, "cores_available": avail, "cpu_model": model,
            "cpu_1": rows["cpu_1"], "cpu_all": a,
            "topk_equal_gpu": bool(a["topk_equal_gpu"] and rows["cpu_1"]["topk_equal_gpu"])}


POWER_SAMPLING = True  # --no-power turns it off (profiled runs: amd-smi is a script whose
                       # interpreter start would be an exec under the profiler's preload)


class PowerSampler:
    """Samples the board's socket power and graphics clocks (amd-smi metric)
    about once a second on a host thread while GPU work runs (kept out of
    timed regions: a query reaches the card's power controller).  Without
    amd-smi, or with --no-power, it records nothing."""

    def __init__(self):
        self.samples = []
        self._stop = threading.Event()
        self._t = None

    def __enter__(self):
        if POWER_SAMPLING and shutil.which("amd-smi"):
            self._t = threading.Thread(target=self._run, daemon=True)
            self._t.start()
        return self

    def _run(self):
        while not self._stop.is_set():
            try:
                out = subprocess.run(["amd-smi", "metric", "-p", "-c"], capture_output=True,
                                     text=True, timeout=10).stdout
            except (OSError, subprocess.SubprocessError):
                return
            p = re.search(r"SOCKET_POWER:\s*([\d.]+)\s*W", out)
            c = re.findall(r"GFX_\d+:\s*\n\s*CLK:\s*([\d.]+)\s*MHz", out)
            if p and c:
                self.samples.append((float(p.group(1)), float(np.mean([float(x) for x in c]))))
            self._stop.wait(1.0)

    def __exit__(self, *exc):
        self._stop.set()
        if self._t:
            self._t.join(timeout=20)

    def summary(self):
        s = self.samples[1:] if len(self.samples) > 2 else self.samples  # skip the ramp
        if not s:
            return None
        return {"socket_power_w_median": float(np.median([x[0] for x in s])),
                "gfx_clk_mhz_median": float(np.median([x[1] for x in s])),
                "samples": len(s), "source": "amd-smi metric -p -c, ~1/s, untimed repeat of the same work"}


def library_int8_gemm(m=16384, n=16384, k=DIM, seconds=3.0):
    """int8 x int8 -> int32 GEMM through torch._int_mm (hipBLASLt on this
    card), operands uniform in [-40, 40] like clamped counts, timed with CUDA
    events over ~`seconds` of back-to-back calls; one 64 x 64 block is
    checked against int64 CPU arithmetic."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(1)
    a = torch.randint(-40, 41, (m, k), dtype=torch.int8, device="cuda", generator=g)
    b = torch.randint(-40, 41, (n, k), dtype=torch.int8, device="cuda", generator=g).t()
    c = torch._int_mm(a, b)
    torch.cuda.synchronize()
    t0 = time.time()
    c = torch._int_mm(a, b)
    torch.cuda.synchronize()
    it = max(3, int(seconds / max(time.time() - t0, 1e-4)))
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(it):
        c = torch._int_mm(a, b)
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / it
    with PowerSampler() as pw:  # power and clock on as many calls again, untimed
        for _ in range(it):
            c = torch._int_mm(a, b)
        torch.cuda.synchronize()
    exact = bool(torch.equal(a[:64].cpu().long() @ b[:, :64].cpu().long(), c[:64, :64].cpu().long()))
    del a, b, c
    return {"op": "torch._int_mm (hipBLASLt), int8 x int8 -> int32", "m": m, "n": n, "k": k,
            "calls": it, "ms_per_call": ms, "tops": 2.0 * m * n * k / (ms / 1000.0) / 1e12,
            "exact": exact, "power": pw.summary()}


def mfma_power_ceiling(q_sd, r_sd, seconds=4.0):
    """The int8 MFMA rate this card sustains under its power and clock limits
    on operands spread like the search's (hvk_calib_mfma_i8: MFMAs from
    registers, operands ~ round(N(0, sd)) clamped to +-127, no memory
    traffic, two waves per SIMD), timed with CUDA events on torch's current
    stream, which the launch uses; power and clock on an untimed repeat."""
    import torch
    import hdc_amd as H
    L = H.lib()
    sink = torch.zeros(1, dtype=torch.int32, device="cuda")
    ops = H.ctypes.c_double()

    def launch(iters):
        H.check(L.hvk_calib_mfma_i8(2, iters, int(q_sd), int(r_sd), H._ptr(sink),
                                    H.ctypes.byref(ops), H._stream()), "hvk_calib_mfma_i8")

    launch(64)
    torch.cuda.synchronize()
    t0 = time.time()
    launch(2048)
    torch.cuda.synchronize()
    iters = int(min(1 << 24, max(64, 2048 * seconds / max(time.time() - t0, 1e-4))))
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    launch(iters)
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e)
    with PowerSampler() as pw:
        launch(iters)
        torch.cuda.synchronize()
    return {"op": "hvk_calib_mfma_i8: v_mfma_i32_32x32x32_i8 from registers, 2 waves per SIMD",
            "q_sd": int(q_sd), "r_sd": int(r_sd), "ms": ms,
            "tops": ops.value / (ms / 1000.0) / 1e12, "power": pw.summary()}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--reads", type=int, default=1_000_000)
    ap.add_argument("--ancestors", type=int, default=2000)
    ap.add_argument("--plain-reads", type=int, default=1_000_000,
                    help="reads of the plain-form runs (median of --plain-runs; at most --reads)")
    ap.add_argument("--plain-runs", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline budget")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-two-limb", action="store_true")
    ap.add_argument("--no-gemm-ref", action="store_true",
                    help="skip the library int8 GEMM reference point")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the reduced-size batching-form and descent runs of the default line")
    ap.add_argument("--no-ceiling", action="store_true",
                    help="skip the power-capped int8 MFMA calibration (hvk_calib_mfma_i8)")
    ap.add_argument("--no-power", action="store_true",
                    help="no amd-smi power sampling (for runs under rocprofv3)")
    ap.add_argument("--workload", choices=["classify", "precision", "descent", "batch", "config0"],
                    default="classify",
                    help="classify: the reads/s line (1M reads x 100k refs, top-5); config0: D=8192, "
                         "k=10, 1,000 refs, 500 reads of 250 nt, top-1 (the CPU-parity workload); "
                         "precision: "
                         "per-result error vs FP64 at 500k refs, 2M reads of 250 nt, top-10; "
                         "descent: top-3 beam down a 7-rank taxonomy of 60,000 leaves, 1M reads; "
                         "batch: the batching form on 4M reads of 100-1,600 nt, 100k refs, top-5")
    ap.add_argument("--batch-reads", type=int, default=4_000_000)
    ap.add_argument("--batch-sample", type=int, default=1 << 20, help="autotuner sample")
    ap.add_argument("--descent-reads", type=int, default=1_000_000)
    ap.add_argument("--descent-beam", type=int, default=3)
    ap.add_argument("--descent-chunk", type=int, default=262144,
                    help="chunk of the batching form of the descent")
    ap.add_argument("--no-batch-descend", action="store_true")
    ap.add_argument("--precision-reads", type=int, default=2_000_000)
    ap.add_argument("--precision-ancestors", type=int, default=5000)
    ap.add_argument("--copy-calib-mb", type=int, default=0,
                    help="first run a device-to-device copy of this many MiB (a kernel of known "
                         "traffic, for scaling HBM byte counters)")
    return ap.parse_args()


def main():
    global POWER_SAMPLING
    a = parse()
    POWER_SAMPLING = not a.no_power
    if a.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path "
              "(the search would shard by reference prototypes).", file=sys.stderr)
        sys.exit(2)
    import torch
    import hdc_amd as H

    torch.cuda.set_device(0)
    if a.workload == "precision":
        run_precision(a)
        return
    if a.workload == "descent":
        run_descent(a)
        return
    if a.workload == "batch":
        run_batch(a)
        return
    if a.workload == "config0":
        run_config0(a)
        return
    if a.copy_calib_mb > 0:
        nbytes = a.copy_calib_mb << 20
        src = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] copy calibration: {nbytes} bytes read and written", file=sys.stderr,
              flush=True)
        del src, dst
    t0 = time.time()
    spec = H.make_spec(seed=0, n_ancestors=a.ancestors, desc_per_ancestor=50, anc_len_min=1300,
                       anc_len_max=1600, div_min=0.02, div_max=0.08, indel_rate=0.01,
                       n_reads=a.reads, read_len_min=150, read_len_max=300, err_min=0.005,
                       err_max=0.02)
    w = H.generate(spec)
    n_refs = len(w.ref_offsets) - 1
    print(f"[bench] generated {n_refs} refs, {a.reads} reads in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)

    ctx = H.Context(DIM, KMER)
    ctx.set_timing(True)
    ref_seqs = H.DeviceSeqs(w.refs, w.ref_offsets)
    torch.cuda.synchronize()
    t1 = time.time()
    db = H.encode(ctx, ref_seqs, counts=True, limbs=True, hi=True, clamp=True)
    torch.cuda.synchronize()
    db_encode_s = time.time() - t1
    db_two_limbs = int(db.maxabs.max().item()) > 127
    read_seqs = H.DeviceSeqs(w.reads, w.read_offsets)
    torch.cuda.synchronize()
    print(f"[bench] DB encoded in {db_encode_s:.2f}s (max|count| {int(db.maxabs.max())})",
          file=sys.stderr, flush=True)

    nq, ld = a.reads, ctx.ld
    L = H.lib()
    db_bundles = H.bundles(db, db_two_limbs)
    reads = H._alloc(nq, ld, counts=False, limbs=True, hi=True, clamp=True)
    reads_out = H._out_struct(reads, ld)
    out = H._topk_bufs(nq, TOPK)

    def encode_reads():
        H.check(L.hvk_encode_ex(ctx.h, H._ptr(read_seqs.packed), H._ptr(read_seqs.invalid),
                                H._ptr(read_seqs.offsets), nq, read_seqs.max_len,
                                H.ctypes.byref(reads_out), H._stream()), "hvk_encode_ex")

    encode_reads()
    torch.cuda.synchronize()
    q_two = int(reads.maxabs.max().item()) > 127
    q_bundles = H.bundles(reads, q_two)

    def step_bounded():
        encode_reads()
        H.check(L.hvk_search_mfma_bounded(ctx.h, H.ctypes.byref(q_bundles),
                                          H.ctypes.byref(db_bundles), ld, TOPK, H._ptr(out.idx),
                                          H._ptr(out.score), H._ptr(out.cos64),
                                          H._ptr(out.n_rescored), H._stream()),
                "hvk_search_mfma_bounded")

    def step_two_limb():
        encode_reads()
        H.search_mfma(ctx, reads, db, TOPK, q_two_limbs=q_two, r_two_limbs=db_two_limbs, out=out2)

    def timed(step, label, steps, warmup):
        for i in range(warmup):
            step()
            torch.cuda.synchronize()
            print(f"[bench] {label} warmup {i} done", file=sys.stderr, flush=True)
        enc_ms, srch_ms = [], []
        torch.cuda.synchronize()
        t_start = time.time()
        for i in range(steps):
            ts = time.time()
            step()
            torch.cuda.synchronize()
            e, s_ = ctx.last_timing()
            enc_ms.append(e)
            srch_ms.append(s_)
            print(f"[bench] {label} step {i}: {time.time() - ts:.3f}s encode {e:.1f}ms "
                  f"search {s_:.1f}ms", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        return (time.time() - t_start) / steps, enc_ms, srch_ms

    sec_step, enc_ms, srch_ms = timed(step_bounded, "mfma-1limb-bounded", a.steps, a.warmup)
    # power and clock on two more (untimed) steps: the sampler stays out of the timed region
    with PowerSampler() as power:
        for _ in range(2):
            step_bounded()
            torch.cuda.synchronize()
    ms_per_step = 1000.0 * sec_step
    value = nq / sec_step
    n_res = int(out.n_rescored.item())

    # the two-limb MFMA form on the same reads (same results required)
    out2 = H._topk_bufs(nq, TOPK)
    two = None
    if not a.no_two_limb:
        sec2, _, srch2 = timed(step_two_limb, "mfma-2limb", 1, 1)
        two = {"reads_per_s": nq / sec2, "kernel_ms": float(np.mean(srch2)),
               "top5_equal": bool(torch.equal(out2.idx, out.idx))}

    # roofline of the dominant kernel (the MFMA search)
    s_kernel = float(np.mean(srch_ms)) / 1000.0
    algo_ops = 2.0 * DIM * n_refs * nq  # logical MACs x2, limb passes not counted
    achieved_tops = algo_ops / s_kernel / 1e12
    e_kernel = float(np.mean(enc_ms)) / 1000.0
    lens = np.diff(w.read_offsets)
    # packed input + lo/hi/clamped rows + per-row norm, max, residual, k-mer count
    enc_bytes = float(np.sum((lens + 3) // 4)) + nq * (3.0 * ld) + nq * 20
    enc_gbs = enc_bytes / e_kernel / 1e9

    # the vendor library's int8 GEMM on this card, for scale
    gemm = None
    if not a.no_gemm_ref:
        gemm = library_int8_gemm()
        gemm["search_tops_over_library"] = achieved_tops / gemm["tops"]
        print(f"[bench] library int8 GEMM: {gemm['tops']:.0f} TOPS (search kernel "
              f"{achieved_tops:.0f} TOPS)", file=sys.stderr, flush=True)

    # the card's int8 MFMA rate under its power cap on operands spread like
    # these (standard deviation of the clamped rows, over a sample of rows)
    ceiling = None
    if not a.no_ceiling:
        q_sd = float(reads.clamp8[: min(nq, 4096)].float().std().item())
        r_sd = float(db.clamp8[: min(n_refs, 4096)].float().std().item())
        ceiling = mfma_power_ceiling(round(q_sd), round(r_sd))
        ceiling["search_tops_over_ceiling"] = achieved_tops / ceiling["tops"]
        print(f"[bench] MFMA ceiling on sd ({round(q_sd)}, {round(r_sd)}) operands: "
              f"{ceiling['tops']:.0f} TOPS (search kernel {achieved_tops:.0f} TOPS)",
              file=sys.stderr, flush=True)

    # plain form over the same reads (all of them by default): median of plain_runs
    # timed runs, the first of which also pays the kernel's first launch
    plain = None
    if not a.no_plain:
        npl = min(a.plain_reads, nq)
        sub = H.DeviceSeqs(w.reads[: w.read_offsets[npl]], w.read_offsets[: npl + 1])
        qp = H.encode(ctx, sub, counts=True, limbs=False)
        torch.cuda.synchronize()
        resp = H._topk_bufs(npl, TOPK)
        runs, kms = [], []
        for i in range(a.plain_runs):
            tp = time.time()
            resp = H.search_plain(ctx, qp, db, TOPK, out=resp)
            torch.cuda.synchronize()
            runs.append(time.time() - tp)
            kms.append(ctx.last_timing()[1])
            print(f"[bench] plain run {i}: {runs[-1]:.2f}s ({npl} reads)", file=sys.stderr,
                  flush=True)
        plain_s = float(np.median(runs))
        plain_kernel_ms = float(np.median(kms))
        same = bool(torch.equal(resp.idx, out.idx[:npl]))
        plain = {"reads": npl, "runs": a.plain_runs, "reads_per_s": npl / plain_s,
                 "reads_per_s_runs": [npl / t for t in runs],
                 "kernel_ms": plain_kernel_ms, "top5_equal_mfma": same,
                 "valu_int_tops": 2.0 * DIM * n_refs * npl / (plain_kernel_ms / 1000.0) / 1e12}
        print(f"[bench] plain: {npl / plain_s:.0f} reads/s median of {a.plain_runs} "
              f"(top-5 equal: {same})", file=sys.stderr, flush=True)
        del qp, resp

    # CPU baseline: the oracle's restatement of the reference's bit-packed path
    cpu = None
    if not a.no_cpu:
        db_host = np.ascontiguousarray(db.counts.cpu().numpy()[:, :DIM])
        cpu = cpu_rows(w, nq, db_host, db.sqnorm.cpu().numpy(), n_refs, DIM, KMER, TOPK,
                       out.idx.cpu().numpy(), a.cpu_seconds)

    # the bounded form keeps KP-entry lists: 8 up to top-5, 16 up to top-13, else 32
    bounded_kernel = f"mfma_kernel<1,1,{8 if TOPK <= 5 else (16 if TOPK <= 13 else 32)},true>"
    line = {
        "metric": "reads/s classified (D=10240, 100k refs, top-5): MFMA form vs plain; % roofline",
        "value": value, "unit": "reads/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int8",
        "data": "synthetic (deterministic generator, seed 0)",
        "config": {"workload": f"config 1: D={DIM}, k={KMER}, {n_refs} refs "
                               f"(2,000 ancestors x 50, 1,300-1,600 nt, 2-8% div, 1% indel), "
                               f"{nq} reads 150-300 nt, 0.5-2% error, top-{TOPK}",
                   "reads": nq, "refs": n_refs, "dim": DIM, "k": KMER, "topk": TOPK,
                   "form": "MFMA, one clamped int8 limb per side + exact bound, FP64 re-score",
                   "read_limbs": 2 if q_two else 1, "ref_limbs": 2 if db_two_limbs else 1},
        "roofline": {"bound": "mfma", "achieved": achieved_tops, "peak": MFMA_INT8_PEAK_TOPS,
                     "unit": "TFLOP/s", "frac": achieved_tops / MFMA_INT8_PEAK_TOPS,
                     "traffic": pmc_traffic(bounded_kernel, nq, n_refs),
                     "traffic_source": "rocprofv3 --pmc FETCH_SIZE x2.0 + WRITE_SIZE x1.0 "
                                       "(copy-calibrated), profiles/mfma_kernel_config1_pmc.md",
                     "kernel": bounded_kernel,
                     "kernel_ms": float(np.mean(srch_ms)),
                     "power": power.summary(),
                     "library_int8_gemm": gemm,
                     "power_capped_ceiling": ceiling},
        "encode": encode_line(enc_ms, enc_gbs, nq, lens),
        "mfma_two_limb": two,
        "plain": plain,
        "mfma_vs_plain": (value / plain["reads_per_s"]) if plain else None,
        "n_rescored": n_res,
        "db_encode_s": db_encode_s,
        "cpu_baseline": cpu,
    }
    if not a.no_extras:
        line["also"] = extras(a)
    print(json.dumps(line), flush=True)


def extras(a):
    """The batching form and the hierarchical descent at reduced size, run after
    the headline measurement (outside its timed region) so that one default run
    records them too: 1M reads of the batching form's length mix (its full
    workload is 4M; tuned on a 524,288-read sample), and 262,144 reads of the
    descent (its full workload is 1M).
    Each keeps its own checks: the batching form's top-5 equal to the direct
    path, the descent's batching form equal to the direct descent."""
    import copy
    out = {}
    b = copy.copy(a)
    b.no_cpu = True
    # a half-size sample lets the tuner reach 524,288-read chunks (two per step), the
    # size at which a chunk's search runs the same 16 slices as the direct path
    b.batch_reads, b.batch_sample = 1 << 20, 1 << 19
    b.descent_reads = 1 << 18
    print("[bench] extras: batching form, 1M reads", file=sys.stderr, flush=True)
    bl = run_batch(b, emit=False)
    out["batch_form"] = {"workload": bl["config"]["workload"], "reads_per_s": bl["value"],
                         "ms_per_step": bl["ms_per_step"], "steps": bl["steps"],
                         "autotune": {k: v for k, v in bl["autotune"].items() if k != "table"},
                         "direct_path": bl["direct_path"], "n_rescored": bl["n_rescored"]}
    print("[bench] extras: descent, 262,144 reads", file=sys.stderr, flush=True)
    dl = run_descent(b, emit=False)
    out["descent"] = {"workload": dl["config"]["workload"], "reads_per_s": dl["value"],
                      "ms_per_step": dl["ms_per_step"], "steps": dl["steps"],
                      "encode_ms": dl["encode_ms"], "descent_ms": dl["descent_ms"],
                      "leaf_accuracy": dl["leaf_accuracy"],
                      "batch_form": dl["batch_form"]}
    return out


def encode_line(enc_ms, enc_gbs, nq, lens):
    """The encoder's measurements: HBM bytes (its algorithmic traffic) and,
    from the committed SQ counter profile of the same launch, its VALU issue
    against the chip's peak (one wave64 VALU instruction per CU per clock at
    the nominal clock)."""
    import torch
    kernel = "encode_wave_kernel<12,11,1>"
    t = float(np.mean(enc_ms)) / 1000.0
    line = {"kernel": kernel, "kernel_ms": 1000.0 * t, "achieved_gbs": enc_gbs,
            "frac_hbm": enc_gbs / HBM_PEAK_GBS, "roofline": None}
    c = valu_counters(kernel, nq)
    if c:
        n_cu = torch.cuda.get_device_properties(0).multi_processor_count
        peak = n_cu * NOMINAL_CLOCK_GHZ  # G wave-instructions per second
        achieved = c["SQ_INSTS_VALU"] / t / 1e9
        kmers = float(np.sum(np.maximum(lens - KMER + 1, 0)))
        line["roofline"] = {
            "bound": "valu", "achieved": achieved, "peak": peak, "unit": "G wave-instr/s",
            "frac": achieved / peak,
            "valu_insts_per_kmer_word": c["SQ_INSTS_VALU"] * 64 / (kmers * ((DIM + 31) // 32)),
            "lds_insts_per_valu": c.get("SQ_INSTS_LDS", 0.0) / c["SQ_INSTS_VALU"],
            "lds_busy": (c["SQ_LDS_IDX_ACTIVE"] / n_cu / (c["GRBM_GUI_ACTIVE"] / 8)
                         if "SQ_LDS_IDX_ACTIVE" in c and "GRBM_GUI_ACTIVE" in c else None),
            "source": "rocprofv3 --pmc SQ_INSTS_VALU et al., profiles/encode_config1_pmc.md"}
    return line


def run_precision(a):
    """Precision form: 500k references (5,000 ancestors x 100), 2M reads of 250 nt, top-10.

    Both MFMA forms run on the same encoded reads: the bounded one-limb form (clamped int8
    operands, FP32 keys carrying an exact upper bound) and the two-limb form (counts split into
    hi/lo int8 limbs, exact int32 partial products).  Each returned score is float(c) of the
    FP64 c recomputed from the exact int64 dot; the report gives, for every one of the
    2M x 10 results, its relative error against that FP64 value, and checks a sample of reads
    against the oracle's FP64 brute force (indices identical, FP64 values equal)."""
    import torch
    import hdc_amd as H
    K10 = 10
    t0 = time.time()
    spec = H.make_spec(seed=0, n_ancestors=a.precision_ancestors, desc_per_ancestor=100,
                       anc_len_min=1300, anc_len_max=1600, div_min=0.02, div_max=0.08,
                       indel_rate=0.01, n_reads=a.precision_reads, read_len_min=250,
                       read_len_max=250, err_min=0.005, err_max=0.02)
    w = H.generate(spec)
    n_refs, nq = len(w.ref_offsets) - 1, a.precision_reads
    print(f"[precision] generated {n_refs} refs, {nq} reads in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    ctx = H.Context(DIM, KMER)
    ctx.set_timing(True)
    db = H.encode(ctx, H.DeviceSeqs(w.refs, w.ref_offsets), counts=True, limbs=True, hi=True,
                  clamp=True)
    read_seqs = H.DeviceSeqs(w.reads, w.read_offsets)
    reads = H.encode(ctx, read_seqs, counts=False, limbs=True, hi=True, clamp=True)
    torch.cuda.synchronize()
    print(f"[precision] encoded (max|count| refs {int(db.maxabs.max())}, reads "
          f"{int(reads.maxabs.max())})", file=sys.stderr, flush=True)

    res = {}
    for form in ("bounded", "two_limb"):
        out = H._topk_bufs(nq, K10)
        run = (lambda: H.search_mfma_bounded(ctx, reads, db, K10, out=out)) if form == "bounded" \
            else (lambda: H.search_mfma(ctx, reads, db, K10, out=out))
        run()  # warm-up
        torch.cuda.synchronize()
        ts = time.time()
        run()
        torch.cuda.synchronize()
        sec = time.time() - ts
        _, kms = ctx.last_timing()
        res[form] = (out, sec, kms)
        print(f"[precision] {form}: {nq / sec:.0f} reads/s (search kernel {kms:.1f} ms), "
              f"re-scored {int(out.n_rescored.item())}", file=sys.stderr, flush=True)

    out, sec, kms = res["bounded"]
    sc = out.score.double()
    c64 = out.cos64
    ok = out.idx >= 0
    rel = ((sc - c64).abs() / c64.abs().clamp_min(1e-300))[ok]
    absd = (sc - c64).abs()[ok]
    same_forms = bool(torch.equal(out.idx, res["two_limb"][0].idx))
    # margin of the k-th result over the (k+1)-th: how near the ties are
    gap = (c64[:, K10 - 2] - c64[:, K10 - 1])[ok[:, K10 - 1]]

    cpu = None
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as orc
        ncores = min(16, os.cpu_count() or 1)
        db_host = np.ascontiguousarray(db.counts.cpu().numpy()[:, :DIM])
        dbn = db.sqnorm.cpu().numpy()
        O = orc.oracle()

        def cpu_run(n):
            seqs = w.reads[: w.read_offsets[n]]
            offs = np.ascontiguousarray(w.read_offsets[: n + 1])
            oi = np.zeros((n, K10), np.int32)
            oc = np.zeros((n, K10), np.float64)
            t = O.orc_classify_reads(DIM, KMER, seqs, offs.ctypes.data, n, db_host.ctypes.data,
                                     dbn.ctypes.data, n_refs, K10, oi.ctypes.data, oc.ctypes.data,
                                     ncores)
            return t, oi, oc

        t1, _, _ = cpu_run(ncores)
        n_cpu = int(max(ncores, min(nq, (a.cpu_seconds / (t1 / ncores)) // ncores * ncores)))
        tc, oi, oc = cpu_run(n_cpu)
        gi = out.idx[:n_cpu].cpu().numpy()
        gc = c64[:n_cpu].cpu().numpy()
        cpu = {"reads": n_cpu, "cores": ncores, "reads_per_s": n_cpu / tc,
               "top10_equal": bool(np.array_equal(oi, gi)),
               "cos64_max_abs_diff": float(np.abs(oc - gc).max()),
               "score_max_rel_err": float((np.abs(out.score[:n_cpu].cpu().numpy().astype(np.float64)
                                                  - oc) / np.abs(oc)).max())}
        print(f"[precision] oracle sample: {n_cpu} reads, top-10 equal {cpu['top10_equal']}",
              file=sys.stderr, flush=True)

    line = {
        "metric": "per-result error vs FP64 (precision form, D=10240, 500k refs, 2M reads, top-10)",
        "value": float(rel.max().item()), "unit": "max relative error of float(c) vs FP64 c",
        "n_gpus": 1, "higher_is_better": False, "dtype": "int8",
        "data": "synthetic (deterministic generator, seed 0)",
        "config": {"workload": f"config 2: D={DIM}, k={KMER}, {n_refs} refs "
                               f"({a.precision_ancestors} ancestors x 100), {nq} reads of 250 nt, "
                               f"top-{K10}", "reads": nq, "refs": n_refs, "topk": K10},
        "results": int(ok.sum().item()),
        "rel_err": {"max": float(rel.max().item()), "mean": float(rel.mean().item()),
                    "bound": 2.0 ** -24},
        "abs_err_max": float(absd.max().item()),
        "kth_gap_min": float(gap.min().item()) if gap.numel() else None,
        "kth_gap_below_2e-6": int((gap < 2e-6).sum().item()),
        "bounded": {"reads_per_s": nq / sec, "kernel_ms": kms,
                    "n_rescored": int(out.n_rescored.item())},
        "two_limb": {"reads_per_s": nq / res["two_limb"][1], "kernel_ms": res["two_limb"][2],
                     "n_rescored": int(res["two_limb"][0].n_rescored.item()),
                     "top10_equal_bounded": same_forms},
        "oracle_sample": cpu,
    }
    print(json.dumps(line), flush=True)


def run_descent(a, emit=True):
    """Hierarchical descent: a 7-rank taxonomy (1 root, 40 phyla, 200 classes, 800 orders,
    3,000 families, 12,000 genera, 60,000 species leaves of 1,500 nt, each 1% from its genus
    centroid), parent prototypes the integer sums of their children's bundles, 1M reads of
    150-1,600 nt (log-uniform, clamped to the leaf length) with 1% substitutions; top-3 beam
    per rank, identify's scoring and tie order, threshold -1 (the integer-sum prototypes of the
    top ranks score reads below identify's 0.0 cut).  One step = encode the reads + the six
    levels of the descent; inputs resident in HBM; prototypes built once, outside the step."""
    import torch
    import hdc_amd as H
    t0 = time.time()
    nq, beam, thr = a.descent_reads, a.descent_beam, -1.0
    tax, rd, roff, truth = H.make_taxonomy(n_reads=nq, seed=0)
    print(f"[descent] generated {tax.n_cat} categories, {nq} reads "
          f"({len(rd) / 1e6:.0f} Mnt) in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    ctx = H.Context(DIM, KMER)
    ctx.set_timing(True)
    t1 = time.time()
    leaf_enc = H.encode(ctx, H.DeviceSeqs(tax.leaf_seqs, tax.leaf_offsets), counts=True,
                        limbs=False, hi=False)
    tree = H.build_prototypes(ctx, tax, leaf_enc.counts, tax.ranks[-1])
    torch.cuda.synchronize()
    build_s = time.time() - t1
    del leaf_enc
    read_seqs = H.DeviceSeqs(rd, roff)
    L = H.lib()
    reads = H._alloc(nq, ctx.ld, counts=True, limbs=False, hi=False, clamp=False)
    reads_out = H._out_struct(reads, ctx.ld)

    def encode_reads():
        H.check(L.hvk_encode_ex(ctx.h, H._ptr(read_seqs.packed), H._ptr(read_seqs.invalid),
                                H._ptr(read_seqs.offsets), nq, read_seqs.max_len,
                                H.ctypes.byref(reads_out), H._stream()), "hvk_encode_ex")

    encode_reads()
    torch.cuda.synchronize()
    q_maxabs = int(reads.maxabs.max().item())
    out = H.Descent(torch.full((nq, 6), -1, dtype=torch.int32, device="cuda"),
                    torch.zeros((nq, 6), dtype=torch.float32, device="cuda"),
                    torch.zeros(nq, dtype=torch.int32, device="cuda"),
                    torch.zeros(nq, dtype=torch.int32, device="cuda"))

    def step():
        encode_reads()
        H.descend(ctx, tree, reads, beam=beam, threshold=thr, levels=6, q_maxabs=q_maxabs, out=out)

    for _ in range(a.warmup):
        step()
        torch.cuda.synchronize()
    enc_ms, desc_ms = [], []
    ts = time.time()
    for i in range(a.steps):
        t = time.time()
        step()
        torch.cuda.synchronize()
        e, d = ctx.last_timing()
        enc_ms.append(e)
        desc_ms.append(d)
        print(f"[descent] step {i}: {time.time() - t:.3f}s encode {e:.1f}ms descent {d:.1f}ms",
              file=sys.stderr, flush=True)
    sec = (time.time() - ts) / a.steps
    path = out.path.cpu().numpy()
    ex = out.exit_level.cpu().numpy()
    acc = float((path[:, 5] == truth).mean())

    # the batching form: hvk_batch_descend over length-ordered chunks replayed from graphs
    batch_line = None
    if not a.no_batch_descend:
        bat = H.Batch(DIM, KMER, chunk=a.descent_chunk)
        bout = H.Descent(torch.full((nq, 6), -1, dtype=torch.int32, device="cuda"),
                         torch.zeros((nq, 6), dtype=torch.float32, device="cuda"),
                         torch.zeros(nq, dtype=torch.int32, device="cuda"),
                         torch.zeros(nq, dtype=torch.int32, device="cuda"))
        bat.descend(tree, read_seqs, beam=beam, threshold=thr, levels=6, out=bout)  # capture
        torch.cuda.synchronize()
        tb = []
        for _ in range(a.steps):
            t = time.time()
            bat.descend(tree, read_seqs, beam=beam, threshold=thr, levels=6, out=bout)
            torch.cuda.synchronize()
            tb.append(time.time() - t)
        same_b = bool(torch.equal(bout.path, out.path) and torch.equal(bout.length, out.length) and
                      torch.equal(bout.score.view(torch.int32), out.score.view(torch.int32)) and
                      torch.equal(bout.exit_level, out.exit_level))
        batch_line = {"reads_per_s": nq / float(np.mean(tb)), "ms_per_step": 1000.0 * float(np.mean(tb)),
                      "chunk": a.descent_chunk, "equal_direct": same_b,
                      "note": "hvk_batch_descend: encode + descent per length-ordered chunk, "
                              "graph replays, results in input order"}
        print(f"[descent] batch form: {nq / float(np.mean(tb)):.0f} reads/s, equal {same_b}",
              file=sys.stderr, flush=True)
        bat.close()
        del bout

    cpu = None
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as orc
        nproc, avail, model = orc.host_cpu()
        ncores = max(1, min(16, avail))
        proto = tree.proto.cpu().numpy()[:, :DIM]
        qc_all = reads.counts  # device
        n_cpu = ncores
        tc = 0.0
        while True:  # grow the sample until it takes about cpu_seconds
            qc = qc_all[:n_cpu].cpu().numpy()[:, :DIM]
            t = time.time()
            ref = orc.descend(proto, tax.child_off, tax.child_idx, tax.roots, qc, beam, thr, 6,
                              threads=ncores)
            tc = time.time() - t
            if tc * 4 > a.cpu_seconds or n_cpu * 4 > nq:
                break
            n_cpu *= 4
        same = bool(np.array_equal(ref[0], path[:n_cpu]) and
                    np.array_equal(ref[1].view(np.int32), out.score[:n_cpu].cpu().numpy().view(np.int32)))
        cpu = {"value": n_cpu / tc, "unit": "reads/s", "cores": ncores, "kind": "port",
               "sample": f"first {n_cpu} reads (encoded), six levels ({tc:.1f}s wall)",
               "nproc": nproc, "cores_available": avail, "cpu_model": model,
               "flags": "prebuilt liboracle.so (-O3 -march=x86-64-v3)",
               "paths_and_scores_equal_gpu": same}
        print(f"[descent] cpu: {n_cpu / tc:.1f} reads/s, equal {same}", file=sys.stderr, flush=True)

    # the scoring work: reads x candidates x D multiply-adds (limb passes not counted)
    n_cand = (len(tax.roots) + beam * 5 + beam * 4 + beam * 3.75 + beam * 4 + beam * 5)
    line = {
        "metric": "reads/s classified by hierarchical descent (D=10240, 76,040 categories, "
                  "top-3 beam per rank)",
        "value": nq / sec, "unit": "reads/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1000.0 * sec, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int8/int16",
        "data": "synthetic (deterministic generator, seed 0)",
        "config": {"workload": f"config 4: 7-rank taxonomy 1-40-200-800-3000-12000-60000, "
                               f"leaves 1500 nt, {nq} reads 150-1600 nt, beam {beam}",
                   "reads": nq, "categories": tax.n_cat, "beam": beam, "threshold": thr},
        "encode_ms": float(np.mean(enc_ms)), "descent_ms": float(np.mean(desc_ms)),
        "prototype_build_s": build_s,
        "leaf_accuracy": acc, "exited_early": int((ex >= 0).sum()),
        "approx_int_ops_per_read": 2.0 * DIM * n_cand,
        "batch_form": batch_line,
        "cpu_baseline": cpu,
    }
    if emit:
        print(json.dumps(line), flush=True)
    return line


def run_batch(a, emit=True):
    """Batching form: 100k references (2,000 ancestors x 50), 4M reads of log-uniform length
    100-1,600 nt (clamped to the source reference) with 1% substitutions, top-5.  Reads are
    ordered by length into chunks, each chunk's encode -> MFMA search -> finalize -> scatter is
    replayed from a captured hipGraph; the autotuner picks form / slices / chunk on a sample.
    One step = one hvk_batch_classify over all reads (graph replays), inputs resident in HBM."""
    import torch
    import hdc_amd as H
    t0 = time.time()
    nq = a.batch_reads
    spec = H.make_spec(seed=0, n_ancestors=a.ancestors, desc_per_ancestor=50, anc_len_min=1300,
                       anc_len_max=1600, div_min=0.02, div_max=0.08, indel_rate=0.01,
                       n_reads=nq, read_len_min=100, read_len_max=1600, read_len_log_uniform=1,
                       err_min=0.01, err_max=0.01)
    w = H.generate(spec)
    n_refs = len(w.ref_offsets) - 1
    lens = np.diff(w.read_offsets)
    print(f"[batch] generated {n_refs} refs, {nq} reads (mean {lens.mean():.0f} nt) in "
          f"{time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    ctx = H.Context(DIM, KMER)
    db = H.encode(ctx, H.DeviceSeqs(w.refs, w.ref_offsets), counts=True, limbs=True, hi=True,
                  clamp=True)
    dbb = H.bundles(db)
    seqs = H.DeviceSeqs(w.reads, w.read_offsets)
    torch.cuda.synchronize()
    b = H.Batch(DIM, KMER, 524288)
    out = H._topk_bufs(nq, TOPK)
    t1 = time.time()
    table = []
    ms_per_m = b.autotune(seqs, dbb, TOPK, out, a.batch_sample, table=table)
    cfg = b.get_config()
    print(f"[batch] autotune {time.time() - t1:.1f}s: {cfg} ({ms_per_m:.0f} ms per M reads)",
          file=sys.stderr, flush=True)
    for i in range(max(1, a.warmup)):  # direct runs + captures
        b.classify(seqs, dbb, TOPK, out=out)
        torch.cuda.synchronize()
    ts = time.time()
    for i in range(a.steps):
        t = time.time()
        b.classify(seqs, dbb, TOPK, out=out)
        torch.cuda.synchronize()
        print(f"[batch] step {i}: {time.time() - t:.3f}s ({b.get_config()['graphs']} graphs)",
              file=sys.stderr, flush=True)
    sec = (time.time() - ts) / a.steps

    # the direct path (one encode + one bounded search, reads in input order) on a subset:
    # same results, and its rate
    nd = min(nq, 1_000_000)
    sub = H.DeviceSeqs(w.reads[: w.read_offsets[nd]], w.read_offsets[: nd + 1])
    ctx.set_timing(True)
    q = H.encode(ctx, sub, counts=False, limbs=True, hi=True, clamp=True)
    direct = H._topk_bufs(nd, TOPK)
    H.search_mfma_bounded(ctx, q, db, TOPK, out=direct)
    torch.cuda.synchronize()
    td = time.time()
    q = H.encode(ctx, sub, counts=False, limbs=True, hi=True, clamp=True)
    H.search_mfma_bounded(ctx, q, db, TOPK, out=direct)
    torch.cuda.synchronize()
    direct_s = time.time() - td
    same = bool(torch.equal(direct.idx, out.idx[:nd]) and torch.equal(direct.score, out.score[:nd]))
    del q

    cpu = None
    if not a.no_cpu:
        db_host = np.ascontiguousarray(db.counts.cpu().numpy()[:, :DIM])
        cpu = cpu_rows(w, nq, db_host, db.sqnorm.cpu().numpy(), n_refs, DIM, KMER, TOPK,
                       out.idx.cpu().numpy(), a.cpu_seconds)

    line = {
        "metric": "reads/s classified, batching form (D=10240, 100k refs, 4M reads of "
                  "100-1,600 nt, top-5)",
        "value": nq / sec, "unit": "reads/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1000.0 * sec, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int8", "data": "synthetic (deterministic generator, seed 0)",
        "config": {"workload": f"config 3: D={DIM}, k={KMER}, {n_refs} refs, {nq} reads "
                               f"log-uniform 100-1600 nt (mean {lens.mean():.0f}), 1% error, "
                               f"top-{TOPK}", "reads": nq, "refs": n_refs, "topk": TOPK},
        "autotune": {"form": ["bounded", "two_limb"][cfg["form"]], "shape": cfg["shape"],
                     "splits": cfg["splits"], "chunk": cfg["chunk"],
                     "ms_per_mread_sample": ms_per_m, "sample": a.batch_sample,
                     "table": table},
        "graphs": b.get_config()["graphs"],
        "direct_path": {"reads": nd, "reads_per_s": nd / direct_s, "top5_equal_batch": same},
        "n_rescored": int(out.n_rescored.item()),
        "cpu_baseline": cpu,
    }
    if emit:
        print(json.dumps(line), flush=True)
    return line


def run_config0(a):
    """The CPU-parity workload: D=8192, k=10; 20 random ancestors of 1,500 nt,
    50 descendants each at 3% substitution and 1% indels (1,000 references),
    500 reads of 250 nt sampled from the references with 1% substitutions;
    top-1.  One step = encode the 500 reads + the bounded MFMA search with its
    exact finalize; inputs resident in HBM.  Also timed: the plain and
    two-limb GPU forms, and the CPU baseline on all 500 reads (its own
    workload: the config names the reference's CPU method)."""
    import torch
    import hdc_amd as H
    D0, K0, T0 = 8192, 10, 1
    w = H.generate(H.make_spec())
    n_refs, nq = len(w.ref_offsets) - 1, len(w.read_offsets) - 1
    ctx = H.Context(D0, K0)
    db = H.encode(ctx, H.DeviceSeqs(w.refs, w.ref_offsets), counts=True, limbs=True, hi=True,
                  clamp=True)
    seqs = H.DeviceSeqs(w.reads, w.read_offsets)
    L = H.lib()
    reads = H._alloc(nq, ctx.ld, counts=True, limbs=True, hi=True, clamp=True)
    ro = H._out_struct(reads, ctx.ld)
    out = H._topk_bufs(nq, T0)
    torch.cuda.synchronize()
    db_b = H.bundles(db)

    def encode():
        H.check(L.hvk_encode_ex(ctx.h, H._ptr(seqs.packed), H._ptr(seqs.invalid),
                                H._ptr(seqs.offsets), nq, seqs.max_len, H.ctypes.byref(ro),
                                H._stream()), "hvk_encode_ex")

    encode()
    torch.cuda.synchronize()
    q_b = H.bundles(reads)

    def step():
        encode()
        H.check(L.hvk_search_mfma_bounded(ctx.h, H.ctypes.byref(q_b), H.ctypes.byref(db_b), ctx.ld,
                                          T0, H._ptr(out.idx), H._ptr(out.score), H._ptr(out.cos64),
                                          H._ptr(out.n_rescored), H._stream()),
                "hvk_search_mfma_bounded")

    steps = max(a.steps, 200)
    for _ in range(max(a.warmup, 5)):
        step()
    torch.cuda.synchronize()
    t0 = time.time()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    sec = (time.time() - t0) / steps
    forms = {}
    for name, fn in (("plain", lambda: H.search_plain(ctx, reads, db, T0)),
                     ("two_limb", lambda: H.search_mfma(ctx, reads, db, T0))):
        r = fn()
        torch.cuda.synchronize()
        ts = time.time()
        for _ in range(50):
            encode()
            r = fn()
        torch.cuda.synchronize()
        forms[name] = {"reads_per_s": nq * 50 / (time.time() - ts),
                       "top1_equal_bounded": bool(torch.equal(r.idx, out.idx))}
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as orc
    qc = reads.counts.cpu().numpy()[:, :D0]
    rc = db.counts.cpu().numpy()[:, :D0]
    ref_idx, ref_cos = orc.topk16(qc, rc, T0)
    sc = out.score.cpu().numpy().astype(np.float64)
    cpu = None
    if not a.no_cpu:
        cpu = cpu_rows(w, nq, np.ascontiguousarray(rc), db.sqnorm.cpu().numpy(), n_refs, D0, K0,
                       T0, out.idx.cpu().numpy(), a.cpu_seconds)
    line = {
        "metric": "reads/s classified (D=8192, k=10, 1,000 refs, 500 reads of 250 nt, top-1)",
        "value": nq / sec, "unit": "reads/s", "n_gpus": 1, "steps": steps, "warmup": a.warmup,
        "ms_per_step": 1000.0 * sec, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int8", "data": "synthetic (deterministic generator, seed 0)",
        "config": {"workload": "config 0: D=8192, k=10, 20 ancestors x 50 refs of 1,500 nt "
                               "(3% subst, 1% indel), 500 reads of 250 nt (1% error), top-1",
                   "reads": nq, "refs": n_refs, "dim": D0, "k": K0, "topk": T0},
        "top1_equal_oracle": bool(np.array_equal(out.idx.cpu().numpy(), ref_idx)),
        "score_max_rel_err": float((np.abs(sc - ref_cos) / np.abs(ref_cos)).max()),
        "truth_hit_rate": float((out.idx.cpu().numpy()[:, 0] == w.truth).mean()),
        "forms": forms,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
