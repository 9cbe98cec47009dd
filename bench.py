"""Benchmark: plain collapsed-Gibbs LDA iterations on one MI355X.

A step is one full iteration of the hot path over the whole corpus: sample
every chunk (in turn) with frozen counts + FP64 LL, rebuild n_wk/n_k on the
device, rebuild every chunk's theta -- the window the reference times
(src_culda/train/SingleChunkSingleGPU.h:72-127).  Inputs are resident in
HBM before timing starts.

Default workload (the at-scale plain-CGS configuration): K=1024,
alpha=50/1024, beta=0.01; synthetic LDA corpus V=100,000 (Zipf 1.0
frequencies), D=300,000 docs with lognormal(5.6, 0.6) lengths (~100M
tokens), K_true=500, doc-topic Dirichlet(0.1), topic-word Dirichlet(0.01),
generator seed 2; 4 data_chunks.

Prints ONE JSON line.  --gpus > 1 is not supported (one device per run).
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

CONFIGS = {
    # name: generator + sampler settings
    "c2": dict(V=100_000, D=300_000, KT=500, len_kind=1, len_mean=0.0, len_mu=5.6, len_sigma=0.6,
               doc_prior=0.1, word_prior=0.01, zipf=1.0, chunks=4, gen_seed=2,
               desc="plain CGS K=1024, V=100k Zipf(1.0), D=300k lognormal(5.6,0.6) docs, K_true=500, "
                    "4 data_chunks processed in turn, on-device phi accumulation"),
    "c1": dict(V=10_000, D=10_000, KT=100, len_kind=0, len_mean=100.0, len_mu=0.0, len_sigma=0.0,
               doc_prior=0.1, word_prior=0.05, zipf=1.07, chunks=1, gen_seed=1,
               desc="plain CGS K=1024, V=10k Zipf(1.07), D=10k Poisson(100) docs, K_true=100, 1 chunk"),
}
# Batched workload: 64 independent corpora, each K=1024, Poisson doc lengths
# with mean ~U[30, 400], V ~U[5k, 50k], Zipf(1.0); D ~U[1000, 5000] so the
# total is ~40M tokens (D ~U[1000, 50000] with those lengths would be ~350M).
C3 = dict(num=64, D=(1000, 5000), V=(5000, 50000), mean=(30.0, 400.0), KT=100, doc_prior=0.1,
          word_prior=0.01, zipf=1.0, gen_seed=3,
          desc="batched plain CGS: 64 independent LDA corpora, K=1024, D~U[1000,5000], V~U[5k,50k] Zipf(1.0), "
               "Poisson doc lengths with mean~U[30,400], K_true=100; one persistent sampler launch per "
               "iteration, graph-replayed, slot size and grid autotuned")
HBM_PEAK_GBPS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS) + ["c3", "c4", "c5"])
    ap.add_argument("--no-tune", action="store_true", help="c3: skip the autotuner")
    ap.add_argument("--tune", action="store_true", help="c1/c2: run the autotuner on the warmed-up state before timing")
    ap.add_argument("--iters", type=int, default=200, help="c4: plain iterations that fix the target")
    ap.add_argument("--mh-max-iters", type=int, default=100, help="c4: learned-MH chain length (iterations)")
    ap.add_argument("--eval-every", type=int, default=5, help="c4: held-out evaluation period")
    ap.add_argument("--fold-in", type=int, default=10, help="c4: fold-in sweeps per evaluation")
    ap.add_argument("--epochs", type=int, default=2, help="c4: MLP training epochs")
    ap.add_argument("--k", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--slot-tokens", type=int, default=0)
    ap.add_argument("--fused-phi", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-ttl", action="store_true", help="skip the time-to-LL (plain vs learned MH) leg")
    ap.add_argument("--mh", action="store_true",
                    help="c5: also the bounded learned-MH leg (branched from the plain chain's state)")
    ap.add_argument("--mh-branch-iters", type=int, default=5, help="c5 --mh: iterations of each branch")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py drives exactly one GPU (the corpus is processed chunk by chunk on one device); "
              "--gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import culda_cgs_amd as cg

    if args.config == "c3":
        run_batch(cg, args)
        return
    if args.config == "c4":
        run_mh(cg, args)
        return
    if args.config == "c5":
        run_large(cg, args)
        return
    cfg = CONFIGS[args.config]
    K = args.k
    alpha, beta = 50.0 / K, 0.01
    t0 = time.time()
    corpus = cg.Corpus.generate(cfg["V"], cfg["D"], cfg["KT"], cfg["chunks"], len_kind=cfg["len_kind"],
                                len_mean=cfg["len_mean"], len_mu=cfg["len_mu"], len_sigma=cfg["len_sigma"],
                                doc_topic_prior=cfg["doc_prior"], topic_word_prior=cfg["word_prior"],
                                zipf_s=cfg["zipf"], seed=cfg["gen_seed"], num_threads=host_threads())
    info = corpus.info()
    N = info.num_tokens
    log(f"corpus {args.config}: V={info.num_words} D={info.num_docs} N={N} chunks={info.num_chunks} "
        f"({time.time() - t0:.1f}s)")

    s = cg.Sampler(k=K, alpha=alpha, beta=beta, seed=args.seed, profile=True, slot_tokens=args.slot_tokens,
                   fused_phi=bool(args.fused_phi))
    t0 = time.time()
    s.upload(corpus)
    s.init_z()
    s.rebuild_all()
    s.sync()
    log(f"upload + init {time.time() - t0:.1f}s")
    copy_gbps = s.stream_copy(1 << 30, 20)
    log(f"stream copy {copy_gbps:.0f} GB/s")

    it = 1
    warm_lls, z_warm = [], None
    # the parity check and the CPU baseline take the warm-up iterations but
    # the last: z is downloaded after them, and the last warm-up iteration
    # then runs the GPU up again after that pause, before the timed steps
    parity_iters = args.warmup - 1 if (not args.no_cpu and args.warmup > 1) else args.warmup
    if parity_iters:
        warm_lls, _ = s.iterate_n(it, parity_iters)
        it += parity_iters
    s.sync()
    tune = None
    if args.tune:
        # the autotuner (slot size from timed sweeps; with several chunks, the
        # theta rebuild's workgroups per CU from timed iterations) on the
        # warmed-up state, which it leaves unchanged; not timed.  Opt-in: at
        # C2 its candidates lie within ~1% of each other and its sweep-only
        # timing can pick a slot size that is slower beside the theta
        # rebuilds of the timed iterations
        t0 = time.time()
        tune = s.autotune()
        log(f"autotune ({time.time() - t0:.1f}s): slot_tokens={tune.slot_tokens} theta_groups={tune.theta_groups}")
    if not args.no_cpu and parity_iters:
        # the state after those warm-up iterations, for the parity check below
        z_warm = [s.get_z(c) for c in range(info.num_chunks)]
    if args.warmup > parity_iters:
        s.iterate_n(it, args.warmup - parity_iters)
        it += args.warmup - parity_iters
        s.sync()
    # the timed steps: one call enqueues them back to back (no host round
    # trip between iterations); the LL of every step comes back at the end
    t_start = time.perf_counter()
    first_timed = it
    lls, _ = s.iterate_n(it, args.steps)
    samp_ms_total, samp_bytes_total, launches = s.last_sweeps()  # sums over the timed steps
    s.sync()
    elapsed = time.perf_counter() - t_start
    it += args.steps
    if samp_ms_total == 0:
        # a one-chunk corpus replays its iteration as a graph (no per-sweep
        # events): the sampler launch is timed in one more, eager iteration
        s.iterate(it)
        it += 1
        samp_ms_total, samp_bytes_total, launches = s.last_sweeps()
    st = s.stats()
    tokens_per_s = N * args.steps / elapsed
    avg_launch_ms = samp_ms_total / launches
    bytes_per_launch = samp_bytes_total / launches
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9
    log(f"{tokens_per_s / 1e6:.1f} M tokens/s, LL/token {lls[-1]:.6f}, sampler {avg_launch_ms:.3f} ms/launch, "
        f"{achieved:.0f} GB/s algorithmic")

    traffic = None
    prof = ROOT / "profiles" / f"sampler_{args.config}_traffic.json"
    if prof.exists() and K == 1024:
        try:
            traffic = json.loads(prof.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    time_to_ll = None
    if not args.no_ttl:
        s.close()
        time_to_ll = time_to_ll_c4(cg, args, K)

    cpu, parity = None, None
    if not args.no_cpu and z_warm is not None:
        cpu, parity = cpu_baseline(corpus, info, K, alpha, beta, args.seed, parity_iters, warm_lls, z_warm,
                                   args.cpu_threads)

    line = {
        "metric": "tokens/sec at K=1024 + time to fixed LL/token: plain CGS vs learned-MH",
        "value": tokens_per_s,
        "unit": "tokens/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded LDA generative model)",
        "config": {"workload": cfg["desc"], "k": K, "alpha": alpha, "beta": beta, "num_tokens": N,
                   "num_docs": info.num_docs, "num_words": info.num_words, "chunks": info.num_chunks,
                   "iterations_timed": [first_timed, first_timed + args.steps - 1], "ll_per_token_last": lls[-1],
                   "phi_rebuild": "fused over chunks" if args.fused_phi else "chunks in turn (+=)",
                   "stream_copy_gbps": copy_gbps,
                   "sum_token_nnz_last": st.sum_token_nnz,
                   "autotune": None if tune is None else {
                       "slot_tokens": tune.slot_tokens, "theta_groups_per_cu": tune.theta_groups,
                       "sweep_ms_by_slot_tokens": tune.tried(), "two_iterations_ms_by_theta_groups": tune.theta_tried()},
                   "time_to_ll": time_to_ll},
        "parity": parity,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                     "kernel": "sample_kernel", "avg_launch_ms": avg_launch_ms,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "cpu_baseline": cpu,
    }
    s.close()
    print(json.dumps(line), flush=True)


def time_to_ll_c4(cg, args, K):
    """The second half of the metric: time to the plain sampler's
    200-iteration held-out LL/token, plain CGS vs learned MH, on the
    1M-token corpus (the protocol of --config c4, run with its defaults)."""
    try:
        out = mh_protocol(cg, args, K)
    except Exception as e:  # reported, never hides the tokens/s line
        log(f"time-to-LL run failed: {e}")
        return {"error": str(e)}
    keys = ("target_ll", "threshold_ll", "t_plain_s", "t_mh_s", "speedup", "plain_ms_per_iteration",
            "mh_ms_per_iteration", "mh_train_s", "mh_acceptance", "time_to_ll_by_tolerance", "curve_plain",
            "curve_mh", "workload")
    return {k: out[k] for k in keys}


def run_batch(cg, args):
    """64 corpora sampled together: one graph-replayed iteration per step."""
    K = args.k
    alpha, beta = 50.0 / K, 0.01
    rng = np.random.default_rng(C3["gen_seed"])
    t0 = time.time()
    corpora = []
    for i in range(C3["num"]):
        V = int(rng.integers(C3["V"][0], C3["V"][1] + 1))
        D = int(rng.integers(C3["D"][0], C3["D"][1] + 1))
        mean = float(rng.uniform(*C3["mean"]))
        corpora.append(cg.Corpus.generate(V, D, C3["KT"], 1, len_kind=0, len_mean=mean,
                                          doc_topic_prior=C3["doc_prior"], topic_word_prior=C3["word_prior"],
                                          zipf_s=C3["zipf"], seed=1000 + i,
                                          num_threads=host_threads()))
    N = sum(c.info().num_tokens for c in corpora)
    log(f"c3: {len(corpora)} corpora, N={N} ({time.time() - t0:.1f}s)")
    b = cg.Batch(k=K, profile=True, slot_tokens=args.slot_tokens)
    for i, c in enumerate(corpora):
        b.add(c, alpha, beta, args.seed + i)
    t0 = time.time()
    b.upload()
    b.init()
    log(f"upload + init {time.time() - t0:.1f}s")
    tune = None
    if not args.no_tune:
        tune = b.autotune()
        log(f"autotune: slot_tokens={tune.slot_tokens} groups={tune.persistent_groups} "
            f"({tune.best_ms:.2f} ms/sweep)")
    it = 1
    if args.warmup:
        b.iterate(it, count=args.warmup, use_graph=True)
        it += args.warmup
    t_start = time.perf_counter()
    lls, _ = b.iterate(it, count=args.steps, use_graph=True)
    elapsed = time.perf_counter() - t_start
    it += args.steps
    tokens_per_s = N * args.steps / elapsed
    # the sampler launch alone, timed with events on its stream in one
    # eager iteration after the timed region
    b.iterate(it, count=1, use_graph=False)
    it += 1
    st = b.stats()
    launch_ms = st.sampler_ms
    bytes_per_launch = b.sampler_bytes()
    achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9
    log(f"{tokens_per_s / 1e6:.1f} M tokens/s, {elapsed / args.steps * 1e3:.2f} ms/iteration, sampler "
        f"{launch_ms:.3f} ms, {achieved:.0f} GB/s algorithmic")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_batch(b, corpora, K, alpha, beta, args.seed, it, args.cpu_threads)
    line = {
        "metric": "tokens/sec at K=1024 + time to fixed LL/token: plain CGS vs learned-MH",
        "value": tokens_per_s,
        "unit": "tokens/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded LDA generative model)",
        "config": {"workload": C3["desc"], "k": K, "alpha": alpha, "beta": beta, "num_tokens": N,
                   "num_corpora": len(corpora), "num_slots": st.num_slots,
                   "slot_tokens": tune.slot_tokens if tune else args.slot_tokens,
                   "persistent_groups": tune.persistent_groups if tune else 0,
                   "autotune_tried_ms": tune.tried() if tune else None,
                   "iterations_timed": [it - 1 - args.steps, it - 2],
                   "ll_per_token_last_mean": float(np.mean(lls[-1]))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": None,
                     "kernel": "sample_kernel", "avg_launch_ms": launch_ms,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "cpu_baseline": cpu,
    }
    b.close()
    print(json.dumps(line), flush=True)


def run_large(cg, args):
    """The large corpus (V=140k Zipf(1.05), D=8.2M Poisson(90) docs, ~738M
    tokens, K_true=1000, Dir(0.05) / Dir(0.01), generator seed 5): every
    10th doc held out, the rest in 16 data_chunks resident in HBM, sampled
    in turn with on-device count accumulation; held-out LL/token at the end."""
    K = args.k
    alpha, beta = 50.0 / K, 0.01
    threads = host_threads()
    t0 = time.time()
    full = cg.Corpus.generate(140_000, 8_200_000, 1000, 1, len_kind=0, len_mean=90.0, doc_topic_prior=0.05,
                              topic_word_prior=0.01, zipf_s=1.05, seed=5, num_threads=threads)
    log(f"c5: {full.info().num_tokens} tokens generated ({time.time() - t0:.0f}s)")
    t0 = time.time()
    train, obs, ev = full.heldout_split(10, 16)
    del full
    info = train.info()
    N = info.num_tokens
    log(f"c5: train N={N} in {info.num_chunks} chunks, held-out {obs.info().num_tokens} observed / "
        f"{ev.info().num_tokens} evaluated ({time.time() - t0:.0f}s)")
    s = cg.Sampler(k=K, alpha=alpha, beta=beta, seed=args.seed, profile=True, slot_tokens=args.slot_tokens)
    t0 = time.time()
    s.upload(train)
    s.init_z()
    s.rebuild_all()
    s.sync()
    log(f"upload + init {time.time() - t0:.1f}s")
    fold = cg.Sampler(k=K, alpha=alpha, beta=beta, seed=args.seed + 1000)
    fold.upload(obs)
    fold.set_eval(ev)
    ll0 = fold.heldout_ll(s, args.fold_in)
    it = 1
    if args.warmup:
        s.iterate_n(it, args.warmup)
        it += args.warmup
    s.sync()
    t_start = time.perf_counter()
    lls, _ = s.iterate_n(it, args.steps)
    samp_ms_total, samp_bytes_total, launches = s.last_sweeps()
    s.sync()
    elapsed = time.perf_counter() - t_start
    it += args.steps
    avg_launch_ms = samp_ms_total / launches
    bytes_per_launch = samp_bytes_total / launches
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9
    hll = fold.heldout_ll(s, args.fold_in)
    log(f"{N * args.steps / elapsed / 1e6:.1f} M tokens/s, {elapsed / args.steps * 1e3:.1f} ms/iteration, "
        f"held-out LL/token {ll0:.4f} -> {hll:.4f} after {it - 1} iterations")
    mh = c5_mh_leg(args, s, fold, it, hll) if args.mh else None
    line = {
        "metric": "tokens/sec at K=1024 + time to fixed LL/token: plain CGS vs learned-MH",
        "value": N * args.steps / elapsed,
        "unit": "tokens/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded LDA generative model)",
        "config": {"workload": "plain CGS K=1024 on the large corpus: V=140k Zipf(1.05), D=8.2M Poisson(90) docs, "
                               "K_true=1000, every 10th doc held out, 16 data_chunks resident in HBM",
                   "k": K, "alpha": alpha, "beta": beta, "num_tokens": N, "chunks": info.num_chunks,
                   "iterations_timed": [it - args.steps, it - 1], "ll_per_token_last": lls[-1],
                   "heldout_ll_per_token_initial": ll0, "heldout_ll_per_token_last": hll,
                   "fold_in_sweeps": args.fold_in, "learned_mh": mh},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": None, "kernel": "sample_kernel",
                     "avg_launch_ms": avg_launch_ms, "algorithmic_bytes_per_launch": bytes_per_launch},
        "cpu_baseline": None,
    }
    fold.close()
    s.close()
    print(json.dumps(line), flush=True)


def c5_mh_leg(args, s, fold, it, hll):
    """Learned MH on the large corpus in bounded form, branched from the
    plain chain's state after iteration it - 1 (a checkpoint): the plain
    branch runs n more plain iterations; the MH branch takes two training
    snapshots (the branch state and one plain iteration later), trains the
    proposal MLP one epoch on them, then runs n - 1 MH iterations (2
    proposals per token).  Held-out LL/token after every iteration of both
    branches, each against its own wall clock (evaluations excluded)."""
    import tempfile
    n = args.mh_branch_iters
    ck = Path(tempfile.mkdtemp(prefix="culda_c5_")) / "state.ck"
    s.save_checkpoint(ck, it)

    def timed(fn, *a):
        t0 = time.perf_counter()
        out = fn(*a)
        s.sync()
        return out, time.perf_counter() - t0

    clock, plain = 0.0, [(0, 0.0, hll)]
    for j in range(n):
        _, dt = timed(s.iterate, it + j)
        clock += dt
        plain.append((j + 1, clock, fold.heldout_ll(s, args.fold_in)))
    log(f"c5 plain branch: {plain}")
    nxt = s.load_checkpoint(ck)
    ck.unlink()
    ck.parent.rmdir()
    s.rebuild_all()
    s.sync()
    _, t_enable = timed(lambda: s.mh_enable(max_snapshots=2, init_seed=args.seed))
    clock, mhc = 0.0, [(0, 0.0, fold.heldout_ll(s, args.fold_in))]
    _, dt = timed(s.mh_collect)
    clock += dt
    _, dt = timed(s.iterate, nxt)
    clock += dt
    _, dt2 = timed(s.mh_collect)
    clock += dt2
    mhc.append((1, clock, fold.heldout_ll(s, args.fold_in)))
    (loss, _), dt_train = timed(s.mh_train, 1)
    clock += dt_train
    log(f"c5 MH: 2 snapshots, trained 1 epoch in {dt_train:.1f}s, loss {loss:.4f}")
    mh_ms, rates = [], []
    for j in range(1, n):
        (rate, _), dt = timed(s.mh_iterate, nxt + j, 2)
        clock += dt
        mh_ms.append(dt * 1e3)
        rates.append(rate)
        mhc.append((j + 1, clock, fold.heldout_ll(s, args.fold_in)))
    log(f"c5 MH branch: {mhc}; {np.median(mh_ms) if mh_ms else 0:.0f} ms per MH iteration, "
        f"acceptance {np.mean(rates) if rates else 0:.3f}")
    return {"protocol": "both branches from the plain chain's state after iteration "
                        f"{it - 1}: {n} plain iterations vs 1 plain iteration + 2 snapshots + 1 training "
                        f"epoch + {n - 1} MH iterations (2 proposals per token); held-out LL/token after "
                        f"each iteration ({args.fold_in} fold-in sweeps), clocks exclude evaluations",
            "curve_plain": plain, "curve_mh": mhc, "mh_enable_s": t_enable, "mh_train_s": dt_train,
            "mh_train_loss": loss, "mh_ms_per_iteration": float(np.median(mh_ms)) if mh_ms else None,
            "mh_acceptance": float(np.mean(rates)) if rates else None}


def mh_protocol(cg, args, K):
    """Time to a fixed held-out LL/token, plain CGS vs learned MH, on the
    1M-token corpus with every 10th doc held out (document completion).
    Both clocks count iterations (and, for MH, snapshot collection and
    training), not the held-out evaluations."""
    cfg = CONFIGS["c1"]
    alpha, beta = 50.0 / K, 0.01
    full = cg.Corpus.generate(cfg["V"], cfg["D"], cfg["KT"], 1, len_kind=cfg["len_kind"], len_mean=cfg["len_mean"],
                              doc_topic_prior=cfg["doc_prior"], topic_word_prior=cfg["word_prior"],
                              zipf_s=cfg["zipf"], seed=cfg["gen_seed"], num_threads=host_threads())
    train, obs, ev = full.heldout_split(10, 1)
    N = train.info().num_tokens
    log(f"time-to-LL: train N={N}, held-out observed {obs.info().num_tokens}, evaluated {ev.info().num_tokens}")

    def timed(fn, *a):
        t0 = time.perf_counter()
        out = fn(*a)
        return out, time.perf_counter() - t0

    def sampler():
        s = cg.Sampler(k=K, alpha=alpha, beta=beta, seed=args.seed, slot_tokens=args.slot_tokens)
        s.upload(train)
        s.init_z()
        s.rebuild_all()
        s.sync()
        return s

    fold = cg.Sampler(k=K, alpha=alpha, beta=beta, seed=args.seed + 1000)
    fold.upload(obs)
    fold.set_eval(ev)
    evaluate = lambda s: fold.heldout_ll(s, args.fold_in)

    # plain CGS: the target is its held-out LL after args.iters iterations
    s = sampler()
    clock, plain = 0.0, [(0, 0.0, evaluate(s))]
    for it in range(1, args.iters + 1):
        _, dt = timed(lambda: (s.iterate(it), s.sync()))
        clock += dt
        if it % args.eval_every == 0 or it == args.iters:
            plain.append((it, clock, evaluate(s)))
    plain_iter_ms = clock / args.iters * 1e3
    s.close()
    target = plain[-1][2]
    thr = target - 0.005 * abs(target)
    t_plain = next(t for _, t, ll in plain if ll >= thr)
    log(f"plain: target {target:.5f} (threshold {thr:.5f}) reached at {t_plain:.4f}s; "
        f"{plain_iter_ms:.3f} ms/iteration")

    # learned MH: plain burn-in 1..30 collecting 10..30, training, then MH
    # iterations (a plain sweep every 10th) up to args.mh_max_iters
    m = sampler()
    m.mh_enable(init_seed=args.seed)
    clock, mhc = 0.0, [(0, 0.0, evaluate(m))]
    mh_ms, rates = [], []
    for it in range(1, 31):
        _, dt = timed(lambda: (m.iterate(it), m.sync()))
        clock += dt
        if it >= 10:
            _, dt = timed(m.mh_collect)
            clock += dt
        if it % args.eval_every == 0:
            mhc.append((it, clock, evaluate(m)))
    burn_in_end = clock
    (loss, _), dt_train = timed(m.mh_train, args.epochs)
    clock += dt_train
    log(f"learned MH: trained {args.epochs} epochs in {dt_train:.2f}s, loss {loss:.4f}")
    for it in range(31, args.mh_max_iters + 1):
        if it % 10 == 0:
            _, dt = timed(lambda: (m.iterate(it), m.sync()))
        else:
            (rate, _), dt = timed(m.mh_iterate, it, 2)
            mh_ms.append(dt * 1e3)
            rates.append(rate)
        clock += dt
        if it % args.eval_every == 0:
            mhc.append((it, clock, evaluate(m)))
    m.close()
    fold.close()
    mh_iter_ms = float(np.median(mh_ms)) if mh_ms else None

    def first_time(curve, thr):
        return next((t for _, t, ll in curve if ll >= thr), None)

    tolerances = {}
    for tol in (0.005, 0.001, 0.0005):
        thr_t = target - tol * abs(target)
        tp, tm = first_time(plain, thr_t), first_time(mhc, thr_t)
        tolerances[str(tol)] = {"threshold": thr_t, "t_plain_s": tp, "t_mh_s": tm,
                                "mh_reached_before_any_mh_iteration": tm is not None and tm <= burn_in_end,
                                "speedup": (tp / tm) if (tp and tm) else None}
    t_mh = tolerances["0.005"]["t_mh_s"]
    log(f"learned MH: {mh_iter_ms} ms per MH iteration, acceptance {np.mean(rates) if rates else 0:.3f}; "
        f"time-to-LL {tolerances}")
    return {"workload": "time to the plain sampler's 200-iteration held-out LL/token (within 0.5%), "
                        "plain CGS vs learned MH (2 proposals per token, plain sweep every 10th iteration, "
                        "MLP trained 2 epochs on iterations 10-30); 1M-token corpus (V=10k, D=10k "
                        "Poisson(100)), every 10th doc held out, document completion with "
                        f"{args.fold_in} fold-in sweeps",
            "k": K, "alpha": alpha, "beta": beta, "num_tokens_train": N,
            "target_ll": target, "threshold_ll": thr, "t_plain_s": t_plain, "t_mh_s": t_mh,
            "speedup": (t_plain / t_mh) if t_mh else None,
            "time_to_ll_by_tolerance": tolerances,
            "plain_ms_per_iteration": plain_iter_ms, "mh_ms_per_iteration": mh_iter_ms,
            "mh_train_s": dt_train, "mh_train_loss": loss,
            "mh_acceptance": float(np.mean(rates)) if rates else None,
            "curve_plain": plain, "curve_mh": mhc, "burn_in_end_s": burn_in_end}


def run_mh(cg, args):
    """--config c4: the time-to-LL comparison as its own JSON line."""
    K = args.k
    out = mh_protocol(cg, args, K)
    N = out["num_tokens_train"]
    flops = 2.0 * (384 * 512 + 512 * K) * N
    mh_iter_ms = out["mh_ms_per_iteration"]
    line = {
        "metric": "tokens/sec at K=1024 + time to fixed LL/token: plain CGS vs learned-MH",
        "value": out["speedup"] or 0.0,
        "unit": "x (plain time-to-LL / learned-MH time-to-LL)",
        "n_gpus": 1,
        "steps": args.iters,
        "warmup": 0,
        "ms_per_step": out["plain_ms_per_iteration"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (sampler), bf16 MFMA with f32 accumulation (MLP)",
        "data": "synthetic (seeded LDA generative model)",
        "config": dict(out, mlp_tflops_per_mh_iteration=flops / 1e12),
        "roofline": {"bound": "mfma", "achieved": (flops / (mh_iter_ms * 1e-3) / 1e12) if mh_iter_ms else None,
                     "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": (flops / (mh_iter_ms * 1e-3) / 1e12 / 2500.0) if mh_iter_ms else None,
                     "traffic": None, "kernel": "learned-MH iteration (MLP inference + MH + rebuilds)"},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline_batch(b, corpora, K, alpha, beta, seed, next_it, threads):
    """The C oracle timed on every corpus of the batch, one iteration each,
    from the GPU's current state."""
    import oracle

    threads = threads or host_threads()
    oracle.set_threads(threads)
    n, dt = 0, 0.0
    for m in range(len(corpora)):
        c = corpora[m]
        info = c.info()
        _, end, wi, wt = c.chunk(0)
        om = oracle.OracleModel(info.num_words, K, c.doc_cum(), [end], [wi], [wt], z=[b.get_z(m, 0)])
        om.rebuild()
        t0 = time.perf_counter()
        om.sweep_fast(alpha, beta, seed + m, next_it)
        om.rebuild()
        dt += time.perf_counter() - t0
        n += info.num_tokens
    log(f"cpu oracle: {n} tokens in {dt:.1f}s on {threads} threads")
    return {"value": n / dt, "unit": "tokens/s", "cores": threads, "kind": "port",
            "sample": f"one full iteration of all {len(corpora)} corpora ({n} tokens) from the GPU's state"}


def host_threads():
    """The host's CPU share: OMP_NUM_THREADS where set (the GPU hosts set it),
    else the cores this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline(corpus, info, K, alpha, beta, seed, iters, gpu_lls, gpu_z, threads):
    """The C oracle (the same schedule, OpenMP) on this host: iterations
    1..iters from the same initial z (the GPU's warm-up iterations), timed
    per phase on the host's CPU share, plus one thread on a sample of the
    next sweep.  Also the full-size parity check: the oracle's z after those
    iterations against the GPU's, the LL of each, and the draws of the next
    sweep that fall within 1e-6 (relative) of an FP32 bucket boundary,
    counted by the oracle's FP64 shadow on a sample of words."""
    import oracle

    threads = threads or host_threads()
    oracle.set_threads(threads)
    C = info.num_chunks
    parts = [corpus.chunk(c) for c in range(C)]
    om = oracle.OracleModel(info.num_words, K, corpus.doc_cum(), [p[1] for p in parts], [p[2] for p in parts],
                            [p[3] for p in parts])
    om.init_z(seed)
    t0 = time.perf_counter()
    om.rebuild()
    t_init = time.perf_counter() - t0
    phases = []
    lls = []
    for it in range(1, iters + 1):
        t0 = time.perf_counter()
        ll = om.sweep_fast(alpha, beta, seed, it)
        t1 = time.perf_counter()
        om.rebuild_phi()
        t2 = time.perf_counter()
        om.rebuild_theta()
        t3 = time.perf_counter()
        phases.append({"iteration": it, "sweep_s": t1 - t0, "phi_s": t2 - t1, "theta_s": t3 - t2})
        lls.append(ll / om.num_tokens)
        log(f"cpu oracle: iteration {it} in {t3 - t0:.2f}s ({threads} threads)")
    total = sum(p["sweep_s"] + p["phi_s"] + p["theta_s"] for p in phases)
    z_mismatch = int(sum(np.count_nonzero(om.z[c] != gpu_z[c]) for c in range(C)))
    ll_rel = [abs(a - b) / abs(b) for a, b in zip(gpu_lls, lls)]
    # a word sample holding ~1/32 of the tokens, from the middle of the
    # frequency range: one thread, then the FP64 shadow for boundary draws
    cnt = np.zeros(info.num_words, np.int64)
    for p in parts:
        cnt += np.diff(p[2])
    cum = np.cumsum(cnt)
    w_lo = int(np.searchsorted(cum, cum[-1] // 2))
    w_hi = int(np.searchsorted(cum, cum[-1] // 2 + cum[-1] // 32)) + 1
    zs = [z.copy() for z in om.z]
    oracle.set_threads(1)
    t0 = time.perf_counter()
    one = om.sweep_words(alpha, beta, seed, iters + 1, w_lo, w_hi)
    t_one = time.perf_counter() - t0
    om.set_z(zs)
    oracle.set_threads(threads)
    ex = om.sweep_words(alpha, beta, seed, iters + 1, w_lo, w_hi, exact=True)
    log(f"cpu oracle: {iters} iterations {total:.1f}s on {threads} threads; 1 thread {one.tokens} tokens in "
        f"{t_one:.1f}s; parity z mismatches {z_mismatch}, boundary draws {ex.boundary} of {ex.tokens}")
    cpu = {"value": info.num_tokens * iters / total, "unit": "tokens/s", "cores": threads, "kind": "port",
           "sample": f"iterations 1-{iters} of the same corpus ({info.num_tokens} tokens, {C} chunks) from the same "
                     f"initial z: sweep + phi rebuild + theta rebuild, {threads} threads",
           "per_iteration": phases, "initial_rebuild_s": t_init,
           "one_thread": {"value": one.tokens / t_one, "unit": "tokens/s", "cores": 1,
                          "sample": f"the sweep of iteration {iters + 1} over words {w_lo}..{w_hi - 1} "
                                    f"({one.tokens} tokens, ~1/32 of the corpus, mid-frequency words)"}}
    parity = {"check": f"GPU z after iterations 1-{iters} vs the oracle's from the same initial z, every chunk; "
                       f"LL/token of each iteration; boundary draws of sweep {iters + 1} on the word sample",
              "z_mismatch": z_mismatch, "tokens": int(info.num_tokens), "ll_rel": ll_rel,
              "ll_rel_max": max(ll_rel) if ll_rel else None,
              "boundary_draws": int(ex.boundary), "boundary_sample_tokens": int(ex.tokens),
              "fp32_vs_fp64_choice_differs": int(ex.mismatch)}
    return cpu, parity


if __name__ == "__main__":
    main()
