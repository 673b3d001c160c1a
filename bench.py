"""Throughput of the Polya-urn LDA iteration on one MI355X.

Workload: the large synthetic corpus (V=100,000 Zipf(1.07), K=1,024, D=1,000,000 documents with
lognormal(median 200, sigma 1.0) lengths capped at 4,096, alpha=50/K, beta=0.01, seed 2),
generated in memory and made resident in HBM before timing.  One step is one full Gibbs iteration:
Poisson Polya-urn draw of phi, per-word alias tables, the z-sweep of every token and the
word-sorted n_wk reduction.

Prints one JSON line.  `value` is tokens sampled per second over the K timed iterations (the
corpus and model already in HBM).  `roofline` describes the z-sweep kernel (the dominant kernel):
algorithmic bytes per launch = N * (18 + 4 * S_bar) with S_bar the mean number of doc-topic map
entries read per token, over its mean device time measured with HIP events on the library's
stream.  `cpu_baseline` is the CPU oracle (OpenMP) on a bounded sample of the same iteration, run
from the same model state; the sampled documents are also checked bit for bit against the GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "sweep_large_traffic.json")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="large", choices=["tiny", "medium", "large", "batched"])
    ap.add_argument("--docs", type=int, default=0, help="override D (smaller runs for profiling)")
    ap.add_argument("--cpu-docs", type=int, default=20000, help="documents in the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--schedule", default="auto", choices=["auto", "buckets", "persistent", "block", "split"],
                    help="sweep schedule; auto = the autotuner's choice after warmup")
    ap.add_argument("--waves-per-cu", type=int, default=32,
                    help="wavefronts per CU of the persistent schedules when --schedule is not auto")
    ap.add_argument("--split-len", type=int, default=256,
                    help="--schedule split: documents longer than this get the large-LDS launch")
    ap.add_argument("--no-count-rows", action="store_true",
                    help="sweep reads the f32 phi rows only (not the u16 count rows where they fit)")
    ap.add_argument("--no-serve", action="store_true", help="skip the fold-in serving measurement")
    ap.add_argument("--serve-requests", type=int, default=4000)
    ap.add_argument("--serve-serial", type=int, default=100, help="requests in the serial baseline sample")
    ap.add_argument("--serve-waves", type=int, default=2, help="worker wavefronts per CU of the fold-in server")
    ap.add_argument("--no-abi", action="store_true",
                    help="skip timing the five-call ABI path (host buffers, z over PCIe each iteration)")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched-form measurement (power-law corpus: hipGraph replay against "
                         "per-bucket launches, identical Philox streams, z compared)")
    ap.add_argument("--batched-steps", type=int, default=5)
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus above 1 is not supported "
                                   "(documents shard across GPUs with one n_wk all-reduce per iteration; "
                                   "see the sharding design notes)"}))
        sys.exit(2)

    from gpupolyaurnlda_amd import corpus
    from gpupolyaurnlda_amd.resident import ResidentCorpus

    spec = corpus.SPECS[args.workload]
    if args.docs:
        spec = spec.scaled(args.docs)
    t0 = time.time()
    c = corpus.generate(spec)
    N = len(c["w"])
    log(f"generated {spec.name}: N={N} tokens in {time.time() - t0:.1f}s")

    t0 = time.time()
    rc = ResidentCorpus(c["w"], c["d_len"], spec.K, spec.V, spec.alpha, spec.beta, seed=spec.seed, z=c["z"])
    log(f"resident in HBM in {time.time() - t0:.1f}s")
    if args.no_count_rows:
        rc.count_rows(False)

    if args.warmup:
        rc.iterate(args.warmup)
        log(f"warmup {args.warmup} done")
    tune = None
    if args.schedule == "auto":
        rows, chosen = rc.autotune()
        tune = {"candidates": rows, "chosen": rows[chosen]}
        log(f"autotune: chose {rows[chosen]} from {rows}")
    else:
        kind = {"buckets": 0, "persistent": 1, "block": 2, "split": 3}[args.schedule]
        rc.schedule(kind, args.waves_per_cu, {2: 4096, 3: args.split_len}.get(kind, 0))
    n_before = rc.counts() if not args.no_cpu else None
    z_before = rc.z() if not args.no_cpu else None
    rc.reset_stats()
    it0 = rc.iteration

    t0 = time.perf_counter()
    stage_ms = rc.iterate(args.steps)  # synchronises before returning
    elapsed = time.perf_counter() - t0
    visited = rc.visited()
    log(f"timed {args.steps} iterations in {elapsed:.3f}s; stage ms {stage_ms.tolist()}")
    # count rows as the timed iterations ran them (the autotuner may switch them off), and the last
    # Phi step's u16 count rows against its f32 rows
    on, bad, fit = rc.check_count_rows()
    count_rows = {"on": on, "mismatches": bad, "words_fit": fit} if on else {"on": False}
    log(f"count rows: {count_rows}")

    tokens_per_s = N * args.steps / elapsed
    s_bar = visited / (N * args.steps)
    sweep_ms = float(stage_ms[2])
    alg_bytes = N * (18.0 + 4.0 * s_bar)
    achieved = alg_bytes / (sweep_ms * 1e-3) / 1e9
    # HBM traffic per launch comes from rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this same command
    # (same warmup and steps, so the same chain state); it is used only when that run's S_bar is
    # this run's, and is reported with its own S_bar and algorithmic bytes either way
    traffic, traffic_run = None, None
    if os.path.exists(TRAFFIC_FILE):
        try:
            tr = json.load(open(TRAFFIC_FILE))
            if tr.get("workload") == spec.name:
                traffic_run = {k: tr.get(k) for k in ("kernel", "count_rows", "s_bar", "alg_bytes_per_launch",
                                                      "bytes_per_launch", "traffic_over_alg", "command")}
                same_kernel = bool(tr.get("count_rows", False)) == count_rows["on"]
                if same_kernel and tr.get("s_bar") is not None and abs(float(tr["s_bar"]) - s_bar) <= 1e-3 * s_bar:
                    traffic = tr.get("bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "z-sweep (k_sweep_*: all launches of one sweep)", "kernel_ms": round(sweep_ms, 3),
                "alg_bytes_per_launch": int(alg_bytes), "s_bar": round(s_bar, 3), "traffic_run": traffic_run}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(rc, spec, c, n_before, z_before, it0, args.steps, args.cpu_docs)

    serving = None
    if not args.no_serve:
        serving = serve_bench(rc, spec, args.serve_requests, args.serve_serial, args.serve_waves)

    rc.close()  # the library holds one model at a time
    batched = None
    if not args.no_batched:
        batched = batched_bench(args.batched_steps)
    abi = None
    if not args.no_abi:
        abi = abi_rate(c, spec, 3)

    out = {
        "metric": "tokens sampled/sec at K=1024 (% HBM roofline); fold-in docs/s + p99 latency",
        "value": round(tokens_per_s, 1),
        "unit": "tokens/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded LDA generator, resident in HBM)",
        "config": {"workload": spec.name, "V": spec.V, "K": spec.K, "D": spec.D, "N": N,
                   "alpha": spec.alpha, "beta": spec.beta, "seed": spec.seed,
                   "stage_ms": {"phi_draw": round(float(stage_ms[0]), 3), "alias": round(float(stage_ms[1]), 3),
                                "sweep": round(float(stage_ms[2]), 3), "count": round(float(stage_ms[3]), 3)},
                   "parallelism": "1 GPU",
                   "schedule": tune if tune is not None else args.schedule,
                   "count_rows": count_rows,
                   "abi_path_pcie_inclusive": abi},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "serving": serving,
        "batched": batched,
    }
    print(json.dumps(out), flush=True)


def serve_bench(rc, spec, n_req, n_serial, waves_per_cu=2):
    """Fold-in serving against the model just trained (phi and alias tables resident): requests
    with lognormal(300, 1.2) lengths clipped to [10, 5000], words from a random mixture of 1-5
    topics, Poisson arrivals at 2,000 docs/s, 50 sweeps each, admission bounded by 16 GB of HBM.
    Baseline: the first n_serial requests run one at a time through the same sampler."""
    from gpupolyaurnlda_amd import corpus
    from gpupolyaurnlda_amd.serve import FoldInServer, percentiles, run_serial, run_stream

    phi = rc.phi()
    d_len, arrival, w = corpus.requests(phi, n_req, median=300.0, sigma=1.2, lo=10, hi=5000, rate=2000.0, seed=4)
    del phi
    with FoldInServer(hbm_budget_bytes=16 << 30, sweeps=50, waves_per_cu=waves_per_cu, max_slots=16384) as srv:
        # warm the kernels once outside the measurement
        run_serial(srv, d_len[:2], w[:int(d_len[:2].sum())], id_base=1 << 30)
        mixed = run_stream(srv, d_len, arrival, w)
        n_s = min(n_serial, n_req)
        serial = run_serial(srv, d_len[:n_s], w[:int(d_len[:n_s].sum())])
    lat = mixed["latency_s"]
    short = d_len <= 1000
    docs_s = n_req / mixed["wall_s"]
    serial_docs_s = n_s / serial["service_s"].sum()
    same = all(np.array_equal(mixed["results"][i], serial["results"][i]) for i in range(n_s))
    out = {"requests": n_req, "offered_docs_per_s": 2000.0, "sweeps": 50,
           "docs_per_s": round(docs_s, 1),
           "latency_ms": {k: round(v * 1e3, 3) for k, v in percentiles(lat).items()},
           "latency_ms_le1000_tokens": {k: round(v * 1e3, 3) for k, v in percentiles(lat[short]).items()},
           "in_server_ms": {k: round(v * 1e3, 3) for k, v in percentiles(mixed["in_server_s"]).items()},
           "serial_docs_per_s": round(float(serial_docs_s), 2),
           "serial_sample": n_s,
           "speedup_vs_serial": round(float(docs_s / serial_docs_s), 1),
           "results_identical_alone_vs_mixed": bool(same),
           "mean_tokens": round(float(d_len.mean()), 1)}
    log(f"serving: {out}")
    return out


def abi_rate(c, spec, iters):
    """The reference host's path: train.rs's loop over the five calls with the whole corpus in one
    host buffer, so every iteration copies z (u32) to the device and back over PCIe.  Returns the
    PCIe-inclusive tokens/s (one warm-up iteration, which also uploads w and d once)."""
    from gpupolyaurnlda_amd.train import Sampler, train

    N, D = len(c["w"]), len(c["d_len"])
    z = c["z"].astype(np.uint32)
    w = np.ascontiguousarray(c["w"], dtype=np.uint32)
    d = np.ascontiguousarray(c["d_len"], dtype=np.uint32)
    k_d = np.zeros(D, dtype=np.uint32)
    C_counts = np.bincount(w, minlength=spec.V).astype(np.uint32)
    from gpupolyaurnlda_amd._lib import check, lib

    modes = {}
    with Sampler(spec.alpha, spec.beta, spec.K, C_counts, N, D, seed=spec.seed) as smp:
        # slice cache: 1 = reuse the device copy of w/d after verifying the whole host content
        # (the default), 2 = trusted (host buffers unchanged between calls), 0 = upload w/d on
        # every call as the reference does
        for mode, name in [(1, "content_verified"), (2, "trusted"), (0, "reupload_every_call")]:
            check(lib().gpulda_set_slice_cache(mode), "set_slice_cache")
            train(smp, z, w, d, k_d, 1, N, D)
            t0 = time.perf_counter()
            train(smp, z, w, d, k_d, iters, N, D)
            dt = (time.perf_counter() - t0) / iters
            modes[name] = {"tokens_per_s": round(N / dt, 1), "ms_per_iteration": round(dt * 1e3, 3)}
        check(lib().gpulda_set_slice_cache(1), "set_slice_cache")
    out = dict(modes["content_verified"])
    out.update({"iterations": iters, "slice_cache_modes": modes,
                "buffer": "whole corpus in one pageable host buffer (train.rs loop; z narrowed to u16 in pinned "
                          "memory, both ways per iteration)"})
    log(f"ABI path: {out}")
    return out


def batched_bench(steps):
    """The batched form on its own corpus (V=100k, K=1,024, D=2M, power-law lengths on [4, 20000]):
    after two warm-up iterations the autotuner picks the schedule (persistent and block-per-document
    launches), then `steps` iterations run as a replayed hipGraph; the same chain is replayed from
    the start with plain per-bucket launches over the same iterations, and z must come out
    identical."""
    from gpupolyaurnlda_amd import corpus
    from gpupolyaurnlda_amd.resident import ResidentCorpus

    spec = corpus.BATCHED
    t0 = time.time()
    c = corpus.generate(spec)
    N = len(c["w"])
    log(f"batched: generated N={N} tokens, longest document {int(c['d_len'].max())}, in {time.time() - t0:.1f}s")
    rc = ResidentCorpus(c["w"], c["d_len"], spec.K, spec.V, spec.alpha, spec.beta, seed=spec.seed, z=c["z"])
    rc.iterate_untimed(2)
    rows, chosen = rc.autotune()
    it0 = rc.iteration
    rc.use_graph(True)
    rc.iterate_untimed(1)  # capture
    t0 = time.perf_counter()
    stage_ms = rc.iterate(steps)
    t_graph = (time.perf_counter() - t0) / steps
    z_graph = rc.z()
    rc.close()
    plain = ResidentCorpus(c["w"], c["d_len"], spec.K, spec.V, spec.alpha, spec.beta, seed=spec.seed, z=c["z"])
    plain.schedule(0)
    plain.iterate_untimed(it0 + 1)
    t0 = time.perf_counter()
    plain.iterate_untimed(steps)
    t_plain = (time.perf_counter() - t0) / steps
    same = bool(np.array_equal(plain.z(), z_graph))
    plain.close()
    out = {"workload": spec.name, "N": N, "schedule": rows[chosen], "steps": steps,
           "graph_ms_per_iter": round(t_graph * 1e3, 3), "graph_tokens_per_s": round(N / t_graph, 1),
           "graph_stage_ms": {"phi_draw": round(float(stage_ms[0]), 3), "alias": round(float(stage_ms[1]), 3),
                              "sweep": round(float(stage_ms[2]), 3), "count": round(float(stage_ms[3]), 3)},
           "plain_buckets_ms_per_iter": round(t_plain * 1e3, 3),
           "plain_buckets_tokens_per_s": round(N / t_plain, 1), "identical_z": same}
    log(f"batched: {out}")
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(rc, spec, c, n_before, z_before, it0, steps, cpu_docs):
    """Times the oracle on the iteration after the timed ones: the full phi step (all V x K
    cells) plus the z-sweep of the first `cpu_docs` documents, scaled to the whole corpus.  The
    GPU runs the same iteration; phi and the sampled documents' z are compared bit for bit."""
    from oracle import oracle as O

    N = len(c["w"])
    n_now = rc.counts()
    z_now = rc.z()
    threads = int(O.lib().orc_threads())
    rc.iterate_untimed(1)  # GPU: iteration it0 + steps + 1
    z_gpu = rc.z()
    phi_gpu = rc.phi()
    t0 = time.perf_counter()
    tables = O.poisson_tables(spec.beta)[2]
    pt, ts = O.phi_draw(n_now, spec.beta, it0 + steps + 1, spec.seed, tables)
    phi, sig, prob, al = O.phi_finish(pt, ts, spec.alpha)
    t_phi = time.perf_counter() - t0
    D_s = min(cpu_docs, spec.D)
    off = c["doc_off"]
    n_s = int(off[D_s])
    t0 = time.perf_counter()
    z_s, _ = O.sweep(z_now[:n_s], c["w"][:n_s], off[:D_s + 1], 0, it0 + steps + 1, spec.seed, spec.K, phi, sig,
                     prob, al, threads=threads)
    t_sw = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.count(c["w"][:n_s], z_s, spec.V, spec.K)
    t_cnt = time.perf_counter() - t0
    t_iter = t_phi + (t_sw + t_cnt) * (N / n_s)
    # one thread: the z-sweep of a smaller sample (the sweep is the iteration's dominant cost)
    D_1 = min(max(cpu_docs // 8, 1), spec.D)
    n_1 = int(off[D_1])
    t0 = time.perf_counter()
    z_1, _ = O.sweep(z_now[:n_1], c["w"][:n_1], off[:D_1 + 1], 0, it0 + steps + 1, spec.seed, spec.K, phi, sig,
                     prob, al, threads=1)
    t_1 = time.perf_counter() - t0
    same_1 = bool(np.array_equal(z_1, z_s[:n_1]))
    same_phi = bool(np.array_equal(phi.view(np.uint32), phi_gpu.view(np.uint32)))
    same_z = bool(np.array_equal(z_s, z_gpu[:n_s]))
    log(f"cpu: phi {t_phi:.2f}s, sweep {t_sw:.2f}s + count {t_cnt:.2f}s for {n_s} tokens; "
        f"bit-exact phi={same_phi} z={same_z}")
    del n_before, z_before
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return {"value": round(N / t_iter, 1), "unit": "tokens/s", "cores": threads, "kind": "port",
            "sample": f"phi step over all V*K cells + z-sweep and count of the first {D_s} documents "
                      f"({n_s} tokens), scaled to N={N}; phi and sampled z bit-identical to the GPU: "
                      f"{same_phi and same_z}",
            "bit_exact_sample": same_phi and same_z,
            "model": cpu_model(), "nproc": os.cpu_count(), "affinity_cpus": affinity,
            "sweep_tokens_per_s": round(n_s / t_sw, 1),
            "sweep_tokens_per_s_1thread": round(n_1 / t_1, 1),
            "sample_1thread": f"z-sweep of the first {D_1} documents ({n_1} tokens) on one thread; "
                              f"z equal to the {threads}-thread sweep: {same_1}"}


if __name__ == "__main__":
    main()
