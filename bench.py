"""Benchmark of the strobealign seeding path on one MI355X.

Default workload (the GPU drop-in configuration, C2): 100 Mbp synthetic
genome (8 contigs, GC 41%, 5% repeat content from 500 families of 1-10 kbp
copied 2-50x at 1-5% divergence), 1M single-end 150 bp reads (0.5%
substitutions, 0.05% indels), fixed 65,536-read batches, one launch per
stage. A step is one batch of reads already resident in HBM through
seeding -> lookup -> hits -> NAMs (+ rescue), its NAM arrays (offsets, NAMs,
fractions, flags) left in HBM, with several batches in flight on streams of
their own (2, 3 and 4 by default, --lanes); `value` is the fastest of those
rates. The host-to-host rate is reported beside it
(host_to_host): each batch's reads copied from pinned host memory to HBM,
the same kernels, and its NAM arrays copied back, with batches of several
lanes in flight so that one batch's copies overlap another's kernels.

--config c3 runs the batched uneven-read form instead: the same genome, 2M
read pairs with lengths drawn from 50/100/150/250/300 bp (weights
0.1/0.2/0.4/0.2/0.1, insert N(400, 50)), 131,072-read batches through the
persistent fused kernel under the configuration the autotuner picks (one
batch at a time, then ranked again with the batches in flight), beside the
per-stage launches. --config c4 runs the human-scale index: a 3.1 Gbp
synthetic genome (24 contigs, 45% repeats, 2% tandem repeats) indexed on the
GPU, 10M pairs of 150 bp in 262,144-read batches, fused against per-stage,
with a bit-exact sample against the oracle over the same table. --config c5
serves timed request traces on that index (Poisson arrivals at five rates).

Prints one JSON line: reads/s (whole job, reads resident in HBM), p50/p99
batch latency (host to host: from the start of a batch's first copy to HBM
to the end of its last copy back), the roofline of the lookup kernel and
the CPU oracle's rate on every host core this process may run on.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    """The host threads this process may run on, and what the host is."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"threads": len(os.sched_getaffinity(0)), "nproc": os.cpu_count(), "model": model}


def stream_pass(G, idx, pin, n_reads, batch, lanes, mode, out=None):
    """One host-to-host pass (gis_find_nams_stream) over the first n_reads
    reads of the pinned pool: -> (seconds, MapResult, per-batch ms)."""
    s, o, l = pin
    t0 = time.perf_counter()
    res, ms = G.find_nams_stream(idx, s, o[:n_reads], l[:n_reads], batch_reads=batch, lanes=lanes, mode=mode,
                                 out=out)
    return time.perf_counter() - t0, res, ms


def host_to_host(G, idx, pin, batch, steps, warmup, mode, lane_choices):
    """The headline form: K steps (batches) from pinned host memory to NAM
    arrays in pinned host memory. Untimed: a pass that sizes every lane's
    buffers, captures its graphs and allocates the output, then W warm-up
    passes per lane count, then two timed passes each, the faster of which
    picks the count. Timed: K batches at the chosen count, as passes over
    the pool (a pass covers up to all of it). Returns the timing and the last
    pass's results."""
    n_all = len(pin[2])
    nb_all = math.ceil(n_all / batch)
    per_pass = min(steps, nb_all)
    n_pass = min(n_all, per_pass * batch)
    _, res, _ = stream_pass(G, idx, pin, n_pass, batch, max(lane_choices), mode)
    sweep = {}
    for lanes in lane_choices:
        for _ in range(max(0, warmup)):
            stream_pass(G, idx, pin, n_pass, batch, lanes, mode, res)
        best = math.inf
        for _ in range(2):
            el, res, _ = stream_pass(G, idx, pin, n_pass, batch, lanes, mode, res)
            best = min(best, el)
        sweep[lanes] = n_pass / best
    lanes = max(sweep, key=sweep.get)
    # passes of up to per_pass batches, a shorter one first, so that the
    # last pass (whose results the parity leg checks) covers n_pass reads
    passes, left = [], steps
    while left > 0:
        passes.append(min(left, per_pass))
        left -= passes[-1]
    passes.reverse()
    reads, elapsed, batch_ms = 0, 0.0, []
    for k in passes:
        n = min(n_all, k * batch)
        el, out, ms = stream_pass(G, idx, pin, n, batch, lanes, mode, res if n == n_pass else None)
        if n == n_pass:
            res = out
        elapsed += el
        reads += n
        batch_ms += list(ms)
    log(f"host to host ({'fused' if mode == G.MODE_FUSED else 'plain'}): {reads / elapsed / 1e6:.1f} M reads/s with "
        f"{lanes} lanes (sweep: " + ", ".join(f"{k}: {v / 1e6:.1f} M" for k, v in sweep.items()) + ")")
    return {"value": reads / elapsed, "unit": "reads/s", "ms_per_step": elapsed * 1e3 / steps,
            "p50_batch_ms": float(np.percentile(batch_ms, 50)), "p99_batch_ms": float(np.percentile(batch_ms, 99)),
            "lanes": lanes, "lane_sweep": {str(k): v for k, v in sweep.items()}, "reads": reads,
            "elapsed_s": elapsed, "copies": G.stream_copy_stats(idx)}, res, n_pass


def parity_leg(O, oidx, seqs, offs, lens, res, n, threads):
    """Every read of the last pass against the oracle over the same table."""
    t0 = time.time()
    eo, en, ef, efl = O.map_batch(oidx, seqs, offs[:n], lens[:n], 2 * oidx.filter_cutoff, 1000, threads)
    exact = bool(np.array_equal(res.nam_offsets, eo) and res.nams.tobytes() == en.tobytes()
                 and res.nonrepetitive_fraction.tobytes() == ef.tobytes()
                 and np.array_equal(res.flags & 1, efl & 1))
    log(f"parity: {n} reads bit-exact vs the oracle: {exact} ({len(en)} NAMs, {time.time() - t0:.1f}s)")
    return {"reads": int(n), "nams": int(len(en)), "bit_exact_vs_oracle": exact}


def resident_results(G, mappers, mode, max_batches=None):
    """Each resident batch run once more (graph replay, as timed) and its
    results downloaded before the next batch of its lane reuses the
    workspace: the NAMs of the first reads of the pool, in read order."""
    parts = []
    for m in mappers[:max_batches]:
        m.run(mode=mode)
        parts.append(m.download())
    offs, base = [np.zeros(1, dtype=np.uint64)], 0
    for r in parts:
        offs.append(r.nam_offsets[1:] + np.uint64(base))
        base += len(r.nams)
    res = G.MapResult(np.concatenate(offs), np.concatenate([r.nams for r in parts]),
                      np.concatenate([r.nonrepetitive_fraction[:len(r.nam_offsets) - 1] for r in parts]),
                      np.concatenate([r.flags[:len(r.nam_offsets) - 1] for r in parts]))
    return res, sum(len(r.nam_offsets) - 1 for r in parts)


MODE_NAMES = {0: "plain", 1: "fused"}


# the top-level p50/p99: the latency of a batch through the host-to-host
# stream, from its reads' copy to HBM to its NAM arrays' copy back
LATENCY_DEF = ("per batch, from the issue of the copy of its reads to HBM to the completion of the "
               "copy of its NAM arrays back (host_to_host); value is the resident rate")


def c2_genome(G, genome_bp):
    return G.synth_genome(genome_bp, n_contigs=8, gc=0.41, repeat_fraction=0.05, n_families=500,
                          family_len=(1000, 10000), copies=(2, 50), divergence=(0.01, 0.05), seed=1)


def build_index(G, genome, clens, threads):
    p = G.IndexParameters.from_read_length(150)
    idx = G.StrobemerIndex(None, p)
    t0 = time.time()
    idx.populate_arrays(genome, clens, f=0.0002, n_threads=threads)
    info = idx.info()
    log(f"index: {info.n_randstrobes} randstrobes, bits {info.bits}, filter_cutoff "
        f"{info.filter_cutoff}, built+uploaded in {time.time() - t0:.1f}s")
    return idx, info


def make_batches(G, idx, seqs, offs, lens, batch, lanes=1):
    """Resident batches; batch i runs on lane i % lanes. With one lane every
    batch runs on the first batch's device workspace and the context's stream
    (the batches run one after another); with more, each lane's first batch
    has a workspace and stream of its own that the lane's other batches share,
    so that runs on different lanes overlap. Only the reads stay resident per
    batch."""
    n = len(lens)
    nb = math.ceil(n / batch)
    mappers = []
    for i in range(nb):
        lo, hi = i * batch, min(n, (i + 1) * batch)
        base = int(offs[lo:hi].min())
        end = int((offs[lo:hi] + lens[lo:hi]).max())
        first = mappers[i % lanes] if i >= lanes else None
        m = G.Mapper(idx, hi - lo, end - base, share=first if lanes > 1 else (mappers[0] if mappers else None),
                     own_stream=lanes > 1 and first is None)
        m.upload(seqs[base:end], offs[lo:hi] - base, lens[lo:hi])
        mappers.append(m)
    return mappers


def pair_blocks(offs, lens, batch):
    """Adjacent mates -> per batch the pair step's layout: the batch's mate 1
    reads, then its mate 2 reads (same pairs, same batches)."""
    n = len(lens)
    order = np.empty(n, dtype=np.int64)
    for lo in range(0, n, batch):
        hi = min(n, lo + batch)
        k = (hi - lo) // 2
        order[lo: lo + k] = np.arange(lo, hi, 2)
        order[lo + k: hi] = np.arange(lo + 1, hi, 2)
    return offs[order], lens[order]


def pair_step(G, O, mappers, steps, mode):
    """The pair step (reference order, types, MAPQ, joint NAM pairs) on each
    batch's resident NAMs, timed alone; batch 0 is checked against the
    oracle's pair step on the same NAMs, which also gives the CPU rate."""
    C = G._native.C
    pp = G._pair_params(joint=True)
    nb = len(mappers)
    ms, pairs_done = [], 0
    for i in range(steps + 1):
        m = mappers[i % nb]
        m.run(mode=mode)
        t0 = time.perf_counter()
        m.ctx.check(m.ctx.L.gis_batch_pairs(m.h, C.byref(pp)))
        if i:  # the first call sizes the pair buffers
            ms.append((time.perf_counter() - t0) * 1e3)
            pairs_done += m.n // 2
    m = mappers[0]
    m.run(mode=mode)
    got = m.pairs()
    res = m.download()
    n_pairs = m.n // 2
    t0 = time.perf_counter()
    exp = O.pairs_batch(res.nam_offsets, res.nams, n_pairs)
    cpu_s = time.perf_counter() - t0
    same = (got.nams.tobytes() == exp["ord"].tobytes() and np.array_equal(got.types, exp["types"])
            and got.joint.tobytes() == exp["joint"].tobytes() and np.array_equal(got.mapq, exp["mapq"])
            and got.top_dropoff.tobytes() == exp["top_dropoff"].tobytes())
    return {"pairs_per_s": pairs_done / (sum(ms) / 1e3), "p50_batch_ms": float(np.percentile(ms, 50)),
            "p99_batch_ms": float(np.percentile(ms, 99)), "pairs_per_batch": n_pairs,
            "type_counts": list(got.type_counts), "joint_entries_batch0": int(len(got.joint)),
            "bit_exact_batch0": bool(same),
            "cpu_oracle": {"value": n_pairs / cpu_s, "unit": "pairs/s", "cores": 1, "kind": "port",
                           "sample": f"pair step of batch 0 ({n_pairs} pairs)"}}


def timed(mappers, steps, warmup, mode, graph):
    """K timed batch runs after an untimed pass over every batch (buffers
    allocated, graphs captured) and W warm-up runs."""
    nb = len(mappers)
    for m in mappers:
        m.run(mode=mode, graph=graph)
    for i in range(warmup):
        mappers[i % nb].run(mode=mode, graph=graph)
    step_ms, reads_done = [], 0
    t_start = time.perf_counter()
    for i in range(steps):
        m = mappers[i % nb]
        ts = time.perf_counter()
        m.run(mode=mode, graph=graph)  # ends with a stream synchronisation
        step_ms.append((time.perf_counter() - ts) * 1e3)
        reads_done += m.n
    elapsed = time.perf_counter() - t_start
    return reads_done / elapsed, elapsed, step_ms


def timed_lanes(mappers, lanes, steps, warmup, mode, graph):
    """K timed batch runs with up to `lanes` in flight, after an untimed pass
    over every batch and W warm-up runs: step i runs batch i % nb on lane
    (i % nb) % lanes, whose previous run is waited for before the launch.
    A batch's latency runs from its launch to the return of its wait."""
    nb = len(mappers)
    for m in mappers:
        m.run(mode=mode, graph=graph)
    pending = [None] * lanes
    step_ms = []

    def drain(lane, record):
        if pending[lane] is not None:
            m, t0 = pending[lane]
            m.wait()
            if record:
                step_ms.append((time.perf_counter() - t0) * 1e3)
            pending[lane] = None

    def go(i, record):
        b = i % nb
        lane = b % lanes
        drain(lane, record)
        pending[lane] = (mappers[b], time.perf_counter())
        mappers[b].launch(mode=mode, graph=graph)
        return mappers[b].n

    for i in range(warmup):
        go(i, False)
    for lane in range(lanes):
        drain(lane, False)
    reads_done = 0
    t_start = time.perf_counter()
    for i in range(steps):
        reads_done += go(i, True)
    for lane in range(lanes):
        drain(lane, True)
    elapsed = time.perf_counter() - t_start
    return reads_done / elapsed, elapsed, step_ms


# the fused form's launch configurations (the library autotuner's set)
FUSED_CANDIDATES = [dict(waves_per_block=w, length_split=sp, queue=q, chunk=ch, blocks_per_cu=bpc)
                    for w in (2, 4, 8) for sp in (0, 1)
                    for (q, ch, bpc) in ((0, 0, 0), (0, 0, 4), (0, 0, 8), (1, 4, 4), (1, 4, 8), (1, 16, 8),
                                         (1, 64, 4))]


def inflight_tune(G, idx, mappers, lanes, steps, warmup):
    """Rank the fused configurations again with `lanes` batches in flight.
    The library autotuner times one batch at a time; with several batches in
    flight a grid that leaves CUs to the other batches' kernels can be the
    faster one. Keeps the winner in the context -> (config, reads/s)."""
    best_c, best_r = dict(idx.ctx.fused_config), 0.0
    for c in FUSED_CANDIDATES:
        idx.ctx.fused_config = c
        r = timed_lanes(mappers, lanes, steps, warmup, G.MODE_FUSED, True)[0]
        if r > best_r:
            best_c, best_r = c, r
    idx.ctx.fused_config = best_c
    return best_c, best_r


def retune_fused(G, idx, mappers, forms, lane_sweep, lanes_max, steps, warmup):
    """inflight_tune at the sweep's largest lane count, then the fused form
    timed again under the winner; forms and lane_sweep take the faster."""
    cfg, _ = inflight_tune(G, idx, mappers, lanes_max, max(8, steps // 4), warmup)
    many = timed_lanes(mappers, lanes_max, steps, warmup, G.MODE_FUSED, True)
    log(f"fused configuration at {lanes_max} in flight: {cfg}, {many[0] / 1e6:.1f} M reads/s")
    if many[0] > lane_sweep[lanes_max][MODE_NAMES[G.MODE_FUSED]]:
        lane_sweep[lanes_max][MODE_NAMES[G.MODE_FUSED]] = many[0]
    one = forms[G.MODE_FUSED][0]
    if many[0] > forms[G.MODE_FUSED][1][0]:
        forms[G.MODE_FUSED] = (one, many, lanes_max)
    return cfg


def resident_sweep(G, idx, seqs, offs, lens, batch, lane_choices, steps, warmup, modes, setup=None):
    """The resident forms at each lane count: a fresh set of batches per
    count (the lanes' workspaces and streams), each form timed one batch at
    a time (on the first set) and with the lanes in flight. Returns
    {mode: (one, many, lanes)} at each form's fastest count, the per-count
    rates, and the last set of batches (kept open for the legs after)."""
    best, sweep, mappers, one = {}, {}, None, {}
    for li, nl in enumerate(lane_choices):
        if mappers:
            for m in mappers:
                m.close()
        mappers = make_batches(G, idx, seqs, offs, lens, batch, lanes=nl)
        if li == 0 and setup:
            setup(mappers)
        sweep[nl] = {}
        for mode in modes:
            if li == 0:
                one[mode] = timed(mappers, steps, warmup, mode, True)
            many = timed_lanes(mappers, nl, steps, warmup, mode, True) if nl > 1 else one[mode]
            sweep[nl][MODE_NAMES[mode]] = many[0]
            if mode not in best or many[0] > best[mode][1][0]:
                best[mode] = (one[mode], many, nl)
            log(f"{MODE_NAMES[mode]} (resident): {many[0] / 1e6:.1f} M reads/s with {nl} in flight")
    return best, sweep, mappers


def timed_forms(mappers, lanes, steps, warmup, mode):
    """Graph-replayed runs one batch at a time and with `lanes` in flight."""
    one = timed(mappers, steps, warmup, mode, True)
    many = timed_lanes(mappers, lanes, steps, warmup, mode, True) if lanes > 1 else one
    return one, many


def form_json(t, steps, **extra):
    rps, elapsed, ms = t
    return {"value": rps, "unit": "reads/s", "ms_per_step": elapsed * 1e3 / steps,
            "p50_batch_ms": float(np.percentile(ms, 50)), "p99_batch_ms": float(np.percentile(ms, 99)),
            **extra}


def kernel_times(G, mappers, k, mode):
    """The same runs enqueued launch by launch, with HIP events between the
    kernels on the batch's stream (graph replays carry no events); with each
    run, the algorithmic bytes k_lookup counted on the device. Full-size
    batches only (the last batch of a workload may be smaller)."""
    out = []
    full = max(m.n for m in mappers)
    for m in [m for m in mappers if m.n == full][:k]:
        m.run(mode=mode, graph=False)
        st = m.stage_times()
        if mode == G.MODE_PLAIN:
            st["_lookup_bytes"], st["_lookups"], st["_lookup_streams"] = m.lookup_bytes()
        out.append(st)
    return out


def stage_means(stages):
    return {k: float(np.mean([st[k] for st in stages])) for k in stages[0] if not k.startswith("_")}


def lookup_roofline(stages, config):
    """Algorithmic bytes of k_lookup per launch, in the byte model of the
    reference's search (16 for the bucket bounds, 16 per record a linear scan
    reads to the first hash >= key, 8 for a found seed's filter probe, 16 per
    run record of a kept seed; counted on the device in the plain eager runs
    and equal to the oracle's count), over its HIP-event time. The bytes the
    kernel streams beside the table (queries in, results out, hits written)
    are reported apart. traffic: HBM bytes per launch from the committed
    rocprofv3 counter summary of this configuration, if there is one."""
    lk_bytes = [st["_lookup_bytes"] for st in stages]
    lk_n = [st["_lookups"] for st in stages]
    lk_streams = [st["_lookup_streams"] for st in stages]
    lookup_ms = [st["lookup"] for st in stages]
    achieved = float(np.sum(lk_bytes) / (np.sum(lookup_ms) * 1e-3) / 1e9)
    log(f"lookup: {np.mean(lk_bytes) / 1e6:.1f} MB algorithmic per launch (+{np.mean(lk_streams) / 1e6:.1f} MB "
        f"streams), {np.mean(lookup_ms):.3f} ms avg, {np.mean(lk_n):.0f} lookups per launch")
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"lookup_traffic_{config}.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get("bytes_per_launch")
    return {"kernel": "k_lookup", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "algorithmic_bytes_per_launch": float(np.mean(lk_bytes)),
            "lookups_per_launch": float(np.mean(lk_n)),
            "algorithmic_bytes_per_lookup": float(np.sum(lk_bytes) / max(1, np.sum(lk_n))),
            "stream_bytes_per_launch": float(np.mean(lk_streams)),
            "avg_launch_ms": float(np.mean(lookup_ms))}


def oracle_index(genome, clens, threads):
    from oracle import oracle_lib as O
    t0 = time.time()
    oidx = O.Index.build_arrays(genome, clens, O.params_from_read_length(150), n_threads=threads)
    log(f"oracle index built in {time.time() - t0:.1f}s")
    return oidx


def cpu_baseline(oidx, seqs, offs, lens, n_cpu, threads, what, build_flags):
    """The CPU oracle on the host cores: one warm-up pass, then the median of
    five timed passes over the first n_cpu reads, and a 1-thread figure on a
    sixteenth of them."""
    from oracle import oracle_lib as O
    n_cpu = min(n_cpu, len(lens))
    end = int((offs[:n_cpu] + lens[:n_cpu]).max())
    rc = 2 * oidx.filter_cutoff
    O.time_map_batch(oidx, seqs[:end], offs[:n_cpu], lens[:n_cpu], rc, 1000, threads)
    runs = [O.time_map_batch(oidx, seqs[:end], offs[:n_cpu], lens[:n_cpu], rc, 1000, threads)
            for _ in range(5)]
    secs = float(np.median(runs))
    n1 = max(1, n_cpu // 16)
    end1 = int((offs[:n1] + lens[:n1]).max())
    s1 = O.time_map_batch(oidx, seqs[:end1], offs[:n1], lens[:n1], rc, 1000, 1)
    host = host_cores()
    log(f"cpu oracle: {n_cpu / secs / 1e6:.2f} M reads/s on {threads} threads ({host['nproc']} cores, "
        f"{host['model']}), {n1 / s1 / 1e3:.0f} k reads/s on 1 thread")
    return {"value": n_cpu / secs, "unit": "reads/s", "cores": threads, "kind": "port",
            "sample": f"first {n_cpu} reads of the {what} workload, seeding through NAMs + rescue; "
                      f"median of 5 passes ({secs:.2f} s each) after a warm-up pass",
            "one_thread_value": n1 / s1, "nproc": host["nproc"], "affinity_threads": host["threads"],
            "cpu_model": host["model"], "build": f"gcc {build_flags}"}


def run_c2(args, G):
    from oracle import oracle_lib as O
    threads = args.cpu_threads or host_cores()["threads"]
    genome, clens = c2_genome(G, args.genome)
    idx, info = build_index(G, genome, clens, threads)
    t0 = time.time()
    seqs, offs, lens, _ = G.synth_reads(genome, clens, args.reads, lengths=(150,), weights=(1.0,),
                                        sub_rate=0.005, indel_rate=0.0005, seed=2)
    log(f"reads {args.reads} in {time.time() - t0:.1f}s")
    batch = args.batch or 65536

    # host memory to host memory, batches of several lanes in flight
    pin = (G.pinned_copy(seqs), G.pinned_copy(offs), G.pinned_copy(lens))
    h2h, res, n_pass = host_to_host(G, idx, pin, batch, args.steps, args.warmup, G.MODE_PLAIN,
                                    args.stream_lanes)
    nams_per_read = len(res.nams) / n_pass
    h2h_bytes = {"h2d_per_batch": float(batch * (lens[:n_pass].mean() + 12)),
                 "d2h_per_batch": float(batch * (nams_per_read * 44 + 13))}

    # the same kernels with the reads resident in HBM and the NAMs left there
    forms, lane_sweep, mappers = resident_sweep(G, idx, seqs, offs, lens, batch, args.lanes, args.steps,
                                                args.warmup, (G.MODE_PLAIN, G.MODE_FUSED))
    (serial_rps, serial_elapsed, serial_ms), (res_rps, res_elapsed, res_ms), lanes = forms[G.MODE_PLAIN]
    fused_one, fused_many, fused_lanes = forms[G.MODE_FUSED]
    log(f"plain form (resident): {serial_rps / 1e6:.1f} M reads/s one batch at a time, "
        f"{res_rps / 1e6:.1f} M with {lanes} in flight (graph)")
    stages = kernel_times(G, mappers, args.steps, G.MODE_PLAIN)
    roof = lookup_roofline(stages, "c2")
    log("stage ms (mean): " + ", ".join(f"{k}={v:.3f}" for k, v in stage_means(stages).items()))
    res_resident, n_resident = resident_results(G, mappers, G.MODE_PLAIN)
    for m in mappers:
        m.close()

    oidx = oracle_index(genome, clens, threads)
    parity = {"resident": parity_leg(O, oidx, seqs, offs, lens, res_resident, n_resident, threads),
              "host_to_host": parity_leg(O, oidx, seqs, offs, lens, res, n_pass, threads)}
    parity["bit_exact_vs_oracle"] = bool(parity["resident"]["bit_exact_vs_oracle"]
                                         and parity["host_to_host"]["bit_exact_vs_oracle"])
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(oidx, seqs, offs, lens, args.cpu_reads, threads, "C2", args.oracle_build)
    oidx.free()

    return {
        "metric": "reads/sec to NAMs at 150bp (1 GPU, bit-exact) vs CPU; p50/p99 batch latency",
        "value": res_rps,
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res_elapsed * 1e3 / args.steps,
        "p50_batch_ms": h2h["p50_batch_ms"],
        "p99_batch_ms": h2h["p99_batch_ms"],
        "latency_definition": LATENCY_DEF,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded genome and reads)",
        "config": {
            "workload": "plain GPU drop-in: 100 Mbp synthetic genome (8 contigs, GC 41%, 5% "
                        "repeats), 1M SE 150 bp reads, 65,536-read batches, one launch per stage; "
                        "value: reads resident in HBM, NAM arrays left in HBM, " + str(lanes) +
                        " batches in flight; host_to_host: reads copied from pinned host memory, NAM "
                        "arrays copied back",
            "batches_in_flight": lanes,
            "genome_bp": args.genome, "reads": args.reads, "batch_reads": batch,
            "batches_per_pass": math.ceil(n_pass / batch), "lanes": h2h["lanes"],
            "index_randstrobes": int(info.n_randstrobes),
            "bits": int(info.bits), "filter_cutoff": int(info.filter_cutoff),
            "parallelism": "single GPU",
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "parity": parity,
        "host_to_host_value": h2h["value"],
        "host_to_host": {**h2h, "bytes": h2h_bytes, "nams_per_read": nams_per_read},
        "device_resident": {"value": res_rps, "batches_in_flight": lanes,
                            "ms_per_step": res_elapsed * 1e3 / args.steps,
                            "p50_batch_ms": float(np.percentile(res_ms, 50)),
                            "p99_batch_ms": float(np.percentile(res_ms, 99)),
                            "one_batch_at_a_time": {"value": serial_rps,
                                                    "ms_per_step": serial_elapsed * 1e3 / args.steps,
                                                    "p50_batch_ms": float(np.percentile(serial_ms, 50)),
                                                    "p99_batch_ms": float(np.percentile(serial_ms, 99))}},
        "stage_ms": stage_means(stages),
        "resident_lane_sweep": lane_sweep,
        "fused_form_resident": form_json(fused_many, args.steps, batches_in_flight=fused_lanes,
                                         one_batch_at_a_time=form_json(fused_one, args.steps)),
    }


def run_c3(args, G):
    from oracle import oracle_lib as O
    threads = args.cpu_threads or host_cores()["threads"]
    genome, clens = c2_genome(G, args.genome)
    idx, info = build_index(G, genome, clens, threads)
    pairs = args.pairs or 2_000_000
    t0 = time.time()
    seqs, offs, lens, _ = G.synth_reads(genome, clens, pairs, lengths=(50, 100, 150, 250, 300),
                                        weights=(0.1, 0.2, 0.4, 0.2, 0.1), paired=True,
                                        insert=(400.0, 50.0), sub_rate=0.005, indel_rate=0.0005,
                                        seed=2)
    log(f"reads {len(lens)} ({pairs} pairs, mean length {lens.mean():.1f}) in {time.time() - t0:.1f}s")
    batch = args.batch or 131072
    # each batch holds whole pairs laid out for the pair step (mate 1 block,
    # mate 2 block)
    offs, lens = pair_blocks(offs, lens, batch)
    steps = args.steps
    tuned = {}

    def autotune(mappers):
        tuned["best"], tuned["ms"], tuned["n"] = mappers[0].autotune()
        log(f"autotune: {tuned['n']} configurations, best {tuned['best']} at {tuned['ms']:.3f} ms per batch")

    forms, lane_sweep, mappers = resident_sweep(G, idx, seqs, offs, lens, batch, args.lanes, steps, args.warmup,
                                                (G.MODE_FUSED, G.MODE_PLAIN), setup=autotune)
    n_tried = tuned["n"]
    best = retune_fused(G, idx, mappers, forms, lane_sweep, args.lanes[-1], steps, args.warmup)
    fused_one, fused_many, lanes = forms[G.MODE_FUSED]
    plain_one, plain_many, plain_lanes = forms[G.MODE_PLAIN]
    log(f"fused (autotuned, resident): {fused_one[0] / 1e6:.1f} M reads/s one batch at a time, "
        f"{fused_many[0] / 1e6:.1f} M with {lanes} in flight")
    log(f"plain per-stage (resident): {plain_one[0] / 1e6:.1f} M reads/s one batch at a time, "
        f"{plain_many[0] / 1e6:.1f} M with {plain_lanes} in flight")
    fstage = kernel_times(G, mappers, steps, G.MODE_FUSED)
    mappers[0].run(mode=G.MODE_FUSED)
    spill_frac = float(np.mean((mappers[0].download().flags & 0x80) != 0))
    stages = kernel_times(G, mappers, steps, G.MODE_PLAIN)
    roof = lookup_roofline(stages, "c3")
    pairs_out = pair_step(G, O, mappers, steps, G.MODE_FUSED)
    log(f"pair step: {pairs_out['pairs_per_s'] / 1e6:.1f} M pairs/s, bit-exact batch 0: "
        f"{pairs_out['bit_exact_batch0']}")
    res_resident, n_resident = resident_results(G, mappers, G.MODE_FUSED)
    for m in mappers:
        m.close()

    # the headline: host memory to host memory
    pin = (G.pinned_copy(seqs), G.pinned_copy(offs), G.pinned_copy(lens))
    h2h_fused, res_f, n_pass = host_to_host(G, idx, pin, batch, steps, args.warmup, G.MODE_FUSED,
                                            args.stream_lanes)
    h2h_plain, res_p, _ = host_to_host(G, idx, pin, batch, steps, args.warmup, G.MODE_PLAIN, args.stream_lanes)
    same = bool(np.array_equal(res_f.nam_offsets, res_p.nam_offsets) and res_f.nams.tobytes() == res_p.nams.tobytes())

    oidx = oracle_index(genome, clens, threads)
    parity = parity_leg(O, oidx, seqs, offs, lens, res_f, n_pass, threads)
    parity["resident_fused"] = parity_leg(O, oidx, seqs, offs, lens, res_resident, n_resident, threads)
    parity["fused_equals_plain"] = same
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(oidx, seqs, offs, lens, args.cpu_reads, threads, "C3", args.oracle_build)
    oidx.free()

    return {
        "metric": "reads/sec to NAMs, 50-300 bp paired mixture (1 GPU, bit-exact) vs CPU; "
                  "p50/p99 batch latency",
        "value": fused_many[0],
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": fused_many[1] * 1e3 / steps,
        "p50_batch_ms": h2h_fused["p50_batch_ms"],
        "p99_batch_ms": h2h_fused["p99_batch_ms"],
        "latency_definition": LATENCY_DEF,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded genome and reads)",
        "config": {
            "workload": "batched uneven reads: 100 Mbp synthetic genome, 2M read pairs of "
                        "50/100/150/250/300 bp (0.1/0.2/0.4/0.2/0.1), insert N(400, 50), "
                        "131,072-read batches, persistent fused kernel (autotuned); value: reads "
                        "resident in HBM, NAM arrays left in HBM, " + str(lanes) + " batches in flight; "
                        "host_to_host: reads copied from pinned host memory, NAM arrays copied back",
            "batches_in_flight": lanes,
            "genome_bp": args.genome, "pairs": pairs, "reads": int(len(lens)),
            "batch_reads": batch, "lanes": h2h_fused["lanes"], "fused_config": best,
            "autotune_single_batch_pick": tuned["best"],
            "autotune_candidates": n_tried, "spilled_fraction": spill_frac,
            "index_randstrobes": int(info.n_randstrobes), "parallelism": "single GPU",
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "parity": parity,
        "host_to_host_value": h2h_fused["value"],
        "host_to_host": h2h_fused,
        "plain_host_to_host": h2h_plain,
        "fused_resident": form_json(fused_many, steps, batches_in_flight=lanes,
                                    one_batch_at_a_time=form_json(fused_one, steps)),
        "fused_stage_ms": stage_means(fstage),
        "pair_step": pairs_out,
        "plain_resident": form_json(plain_many, steps, stage_ms=stage_means(stages), batches_in_flight=plain_lanes,
                                    one_batch_at_a_time=form_json(plain_one, steps)),
        "resident_lane_sweep": lane_sweep,
    }


def c4_genome(G, genome_bp):
    return G.synth_genome(genome_bp, n_contigs=24, gc=0.41, repeat_fraction=0.45, n_families=5000,
                          family_len=(300, 6000), copies=(2, 50), divergence=(0.02, 0.20),
                          tandem_fraction=0.02, tandem_unit=(2, 50),
                          contig_len=(50_000_000, 250_000_000), seed=1)


def run_c4(args, G):
    """Human-scale index: 3.1 Gbp synthetic genome (24 contigs of 50-250 Mbp,
    45% repeats, 2% tandem repeats), 10M pairs of 150 bp; the batched fused
    form (autotuned) against the per-stage launches, NAMs bit-exact."""
    from oracle import oracle_lib as O
    genome_bp = args.genome or 3_100_000_000
    threads = args.cpu_threads or host_cores()["threads"]
    t0 = time.time()
    genome, clens = c4_genome(G, genome_bp)
    log(f"genome {genome_bp} bp, {len(clens)} contigs in {time.time() - t0:.1f}s")
    t0 = time.time()
    idx, info = build_index(G, genome, clens, threads)
    build_s = time.time() - t0
    pairs = args.pairs or 10_000_000
    t0 = time.time()
    seqs, offs, lens, _ = G.synth_reads(genome, clens, pairs, lengths=(150,), weights=(1.0,),
                                        paired=True, insert=(400.0, 50.0), sub_rate=0.005,
                                        indel_rate=0.0005, seed=2)
    log(f"reads {len(lens)} ({pairs} pairs) in {time.time() - t0:.1f}s")
    batch = args.batch or 262144
    steps = args.steps
    tuned = {}

    def autotune(mappers):
        tuned["best"], tuned["ms"], tuned["n"] = mappers[0].autotune()
        log(f"autotune: {tuned['n']} configurations, best {tuned['best']} at {tuned['ms']:.3f} ms per batch")

    forms, lane_sweep, mappers = resident_sweep(G, idx, seqs, offs, lens, batch, args.lanes, steps, args.warmup,
                                                (G.MODE_FUSED, G.MODE_PLAIN), setup=autotune)
    n_tried = tuned["n"]
    best = retune_fused(G, idx, mappers, forms, lane_sweep, args.lanes[-1], steps, args.warmup)
    fused_one, fused_many, lanes = forms[G.MODE_FUSED]
    plain_one, plain_many, plain_lanes = forms[G.MODE_PLAIN]
    log(f"fused (autotuned, resident): {fused_one[0] / 1e6:.1f} M reads/s one batch at a time, "
        f"{fused_many[0] / 1e6:.1f} M with {lanes} in flight")
    log(f"plain per-stage (resident): {plain_one[0] / 1e6:.1f} M reads/s one batch at a time, "
        f"{plain_many[0] / 1e6:.1f} M with {plain_lanes} in flight")
    stages = kernel_times(G, mappers, steps, G.MODE_PLAIN)
    roof = lookup_roofline(stages, "c4")
    log("stage ms (mean): " + ", ".join(f"{k}={v:.3f}" for k, v in stage_means(stages).items()))
    res_resident, n_resident = resident_results(G, mappers, G.MODE_FUSED, max_batches=8)
    for m in mappers:
        m.close()

    # the headline: host memory to host memory
    pin = (G.pinned_copy(seqs), G.pinned_copy(offs), G.pinned_copy(lens))
    h2h_fused, res_f, n_pass = host_to_host(G, idx, pin, batch, steps, args.warmup, G.MODE_FUSED,
                                            args.stream_lanes)
    h2h_plain, res_p, _ = host_to_host(G, idx, pin, batch, steps, args.warmup, G.MODE_PLAIN, args.stream_lanes)
    same = bool(np.array_equal(res_f.nam_offsets, res_p.nam_offsets) and res_f.nams.tobytes() == res_p.nams.tobytes())
    del pin

    # the oracle over the same (GPU-built) table: every read of the timed
    # pass, then the CPU baseline
    t0 = time.time()
    oidx = O.Index.wrap(idx.randstrobes, idx.randstrobe_start_indices, info.bits, info.filter_cutoff,
                        O.params_from_read_length(150))
    log(f"table copied to the host for the oracle in {time.time() - t0:.1f}s")
    parity = parity_leg(O, oidx, seqs, offs, lens, res_f, n_pass, threads)
    parity["resident_fused_first_8_batches"] = parity_leg(O, oidx, seqs, offs, lens, res_resident, n_resident,
                                                          threads)
    parity["fused_equals_plain"] = same
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(oidx, seqs, offs, lens, args.cpu_reads, threads, "C4", args.oracle_build)
    oidx.free()
    return {
        "metric": "reads/sec to NAMs at 150bp, 3.1 Gbp index (1 GPU, bit-exact) vs CPU; "
                  "p50/p99 batch latency",
        "value": fused_many[0],
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": fused_many[1] * 1e3 / steps,
        "p50_batch_ms": h2h_fused["p50_batch_ms"],
        "p99_batch_ms": h2h_fused["p99_batch_ms"],
        "latency_definition": LATENCY_DEF,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded genome and reads)",
        "config": {
            "workload": "human-scale index: 3.1 Gbp synthetic genome (24 contigs of 50-250 Mbp, "
                        "45% repeats, 2% tandem repeats), 10M pairs of 150 bp, 262,144-read batches, "
                        "batched fused form (autotuned) vs per-stage launches; value: reads resident "
                        "in HBM, NAM arrays left in HBM, " + str(lanes) + " batches in flight; "
                        "host_to_host: reads copied from pinned host memory, NAM arrays copied back",
            "batches_in_flight": lanes,
            "genome_bp": genome_bp, "pairs": pairs, "reads": int(len(lens)), "batch_reads": batch,
            "lanes": h2h_fused["lanes"], "fused_config": best,
            "autotune_single_batch_pick": tuned["best"],
            "autotune_candidates": n_tried, "index_randstrobes": int(info.n_randstrobes), "bits": int(info.bits),
            "filter_cutoff": int(info.filter_cutoff), "index_gb": info.device_bytes / 1e9,
            "index_build_s": build_s, "parallelism": "single GPU",
        },
        "parity": parity,
        "roofline": roof,
        "cpu_baseline": cpu,
        "host_to_host_value": h2h_fused["value"],
        "host_to_host": h2h_fused,
        "plain_host_to_host": h2h_plain,
        "fused_resident": form_json(fused_many, steps, batches_in_flight=lanes,
                                    one_batch_at_a_time=form_json(fused_one, steps)),
        "plain_resident": form_json(plain_many, steps, stage_ms=stage_means(stages), batches_in_flight=plain_lanes,
                                    one_batch_at_a_time=form_json(plain_one, steps)),
        "resident_lane_sweep": lane_sweep,
    }


def write_fastq(path, seqs, offs, lens, idx):
    """FASTQ records of the given reads (name, read, '+', quality of 'I')."""
    with open(path, "wb") as f:
        buf = []
        for k, i in enumerate(idx):
            o, n = int(offs[i]), int(lens[i])
            buf.append(b"@read%d/1 synthetic\n%s\n+\n%s\n" % (k, seqs[o: o + n].tobytes(), b"I" * n))
            if len(buf) == 4096:
                f.write(b"".join(buf))
                buf = []
        f.write(b"".join(buf))
    return os.path.getsize(path)


def fastq_leg(G, srv, seqs, offs, lens, n_se, n_pairs, in_flight):
    """FASTQ ingestion into the serving queue: two mates files of n_pairs
    records are read in 4 MiB chunks by the reader and parser threads, each
    chunk submitted as one request, up to in_flight requests outstanding.
    Reports the end-to-end rate (files to NAMs) and the reader's rate alone."""
    import tempfile
    n_pairs = min(n_pairs, (len(lens) - n_se) // 2)
    first = n_se + 2 * np.arange(n_pairs)
    d = tempfile.mkdtemp(prefix="gis_fastq_")
    p1, p2 = os.path.join(d, "r_1.fq"), os.path.join(d, "r_2.fq")
    try:
        nbytes = write_fastq(p1, seqs, offs, lens, first) + write_fastq(p2, seqs, offs, lens, first + 1)
        t0 = time.perf_counter()
        recs = 0
        c = G._native.FastqChunk()
        with G.FastqReader(p1, p2) as r:
            while True:
                code = G._native.lib().gis_fastq_next(r.h, G._native.C.byref(c))
                if code:
                    raise G.GisError(code, r._last_error())
                if not c.n_records:
                    break
                recs += c.n_records
        parse_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        reads, chunks, pending = 0, 0, []
        with G.FastqReader(p1, p2) as r:
            while True:
                rq, ch = srv.submit_fastq(r, keep=False)
                if rq is not None:
                    pending.append(rq)
                    chunks += 1
                if pending and (len(pending) >= in_flight or rq is None):
                    res, _ = srv.wait(pending.pop(0))
                    reads += len(res.nam_offsets) - 1
                if rq is None and not pending:
                    break
        e2e_s = time.perf_counter() - t0
    finally:
        for f in (p1, p2):
            if os.path.exists(f):
                os.remove(f)
        os.rmdir(d)
    log(f"fastq: {recs} pairs ({nbytes / 1e6:.0f} MB) parsed in {parse_s:.2f}s; files -> NAMs "
        f"{reads / e2e_s / 1e6:.1f} M reads/s over {chunks} chunk requests")
    return {"pairs": int(recs), "file_bytes": int(nbytes), "chunk_bytes": 4 << 20,
            "reader_reads_per_s": 2 * recs / parse_s, "reader_gb_per_s": nbytes / parse_s / 1e9,
            "end_to_end_reads_per_s": reads / e2e_s, "requests": chunks, "in_flight": in_flight}


def run_c5(args, G):
    """Serving on the 3.1 Gbp index: Poisson request arrivals at five rates
    (fractions of the saturated throughput), each request 1-8,192 reads
    (log-uniform) with lengths from the 50-300 bp mixture, paired and single
    mixed 50/50; HBM-budgeted admission, graph-replayed steps, several steps
    in flight. Reports p50/p99 request latency and sustained reads/s per rate."""
    genome_bp = args.genome or 3_100_000_000
    t0 = time.time()
    genome, clens = c4_genome(G, genome_bp)
    threads = args.cpu_threads or host_cores()["threads"]
    idx, info = build_index(G, genome, clens, threads)
    log(f"genome + index in {time.time() - t0:.1f}s")
    mix = dict(lengths=(50, 100, 150, 250, 300), weights=(0.1, 0.2, 0.4, 0.2, 0.1),
               sub_rate=0.005, indel_rate=0.0005)
    pool_pairs = args.pairs or 1_000_000
    s_pe, o_pe, l_pe, _ = G.synth_reads(genome, clens, pool_pairs, paired=True, insert=(400.0, 50.0),
                                        seed=2, **mix)
    s_se, o_se, l_se, _ = G.synth_reads(genome, clens, 2 * pool_pairs, seed=3, **mix)
    # one pool: single-end reads then pairs (mates adjacent)
    seqs = np.concatenate([s_se[: int(o_se[-1] + l_se[-1])], s_pe[: int(o_pe[-1] + l_pe[-1])]])
    offs = np.concatenate([o_se, o_pe + np.uint64(int(o_se[-1] + l_se[-1]))]).astype(np.uint64)
    lens = np.concatenate([l_se, l_pe]).astype(np.uint32)
    n_se = len(l_se)
    del s_pe, s_se
    rng = np.random.default_rng(5)

    def requests(n):
        sizes = np.clip(np.exp(rng.uniform(0, np.log(8192), size=n)).astype(np.int64), 1, 8192)
        paired = rng.random(n) < 0.5
        first = np.empty(n, dtype=np.int64)
        for i in range(n):
            k = int(sizes[i])
            if paired[i]:
                k += k & 1  # whole pairs
                sizes[i] = k
                start = n_se + 2 * int(rng.integers(0, (len(lens) - n_se - k) // 2))
            else:
                start = int(rng.integers(0, n_se - k))
            first[i] = start
        return first, sizes

    mode = G.MODE_FUSED if args.server_mode == "fused" else G.MODE_PLAIN
    srv = G.Server(idx, slots=args.slots, max_step_reads=args.step_reads, hbm_budget=32 << 30, mode=mode)
    # saturation: a burst of requests at t = 0 (after one untimed burst that
    # sizes every slot's buffers and captures its graphs)
    fw, kw = requests(args.burst)
    srv.replay(seqs, offs, lens, np.zeros(len(fw)), fw, kw)
    f0, k0 = requests(args.burst)
    lat0, wall0 = srv.replay(seqs, offs, lens, np.zeros(len(f0)), f0, k0)
    sat = float(k0.sum()) / wall0
    log(f"saturation: {len(f0)} requests ({int(k0.sum())} reads) in {wall0:.2f}s = {sat / 1e6:.1f} M reads/s")
    mean_size = float(np.mean(np.clip(np.exp(rng.uniform(0, np.log(8192), size=200000)), 1, 8192)))
    rates = []
    for frac in (0.2, 0.4, 0.6, 0.8, 0.95):
        target = frac * sat
        lam = target / mean_size
        n = max(20, int(lam * args.rate_seconds))
        fr, kr = requests(n)
        arrival = np.cumsum(rng.exponential(1.0 / lam, size=n))
        lat, wall = srv.replay(seqs, offs, lens, arrival, fr, kr)
        done = float(kr.sum())
        row = {"target_reads_per_s": target, "offered_reads_per_s": done / float(arrival[-1]),
               "sustained_reads_per_s": done / wall, "requests": n,
               "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
               "failed": int((lat < 0).sum())}
        log(f"rate {frac:.0%}: {row['sustained_reads_per_s'] / 1e6:.1f} M reads/s sustained, "
            f"p50 {row['p50_ms']:.2f} ms, p99 {row['p99_ms']:.2f} ms over {n} requests")
        rates.append(row)
    fastq = fastq_leg(G, srv, seqs, offs, lens, n_se, args.fastq_pairs, 2 * args.slots) \
        if args.fastq_pairs else None
    st = srv.stats()
    srv.close()
    best = max(rates, key=lambda r: r["sustained_reads_per_s"])
    cpu = None
    if not args.no_cpu:
        # the oracle over the same (GPU-built) table, on the pool's first
        # reads (single-end, the 50-300 bp mixture)
        from oracle import oracle_lib as O
        oidx = O.Index.wrap(idx.randstrobes, idx.randstrobe_start_indices, info.bits, info.filter_cutoff,
                            O.params_from_read_length(150))
        cpu = cpu_baseline(oidx, seqs, offs, lens, args.cpu_reads, threads, "C5", args.oracle_build)
        oidx.free()
    return {
        "metric": "served reads/sec to NAMs, 3.1 Gbp index, Poisson request arrivals (1 GPU, "
                  "bit-exact); p50/p99 request latency",
        "value": best["sustained_reads_per_s"],
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": int(st["steps"]),
        "warmup": 0,
        "ms_per_step": st["busy_s"] * 1e3 / max(1, st["steps"]),
        "p50_batch_ms": best["p50_ms"],
        "p99_batch_ms": best["p99_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded genome, reads and arrivals)",
        "config": {
            "workload": "serving: 3.1 Gbp index resident, Poisson arrivals at 5 rates (20-95% of "
                        "saturation), requests of 1-8,192 reads (log-uniform), 50-300 bp mixture, "
                        "paired and single 50/50, HBM-budgeted admission, graph-replayed steps",
            "genome_bp": genome_bp, "pool_reads": int(len(lens)), "slots": args.slots,
            "max_step_reads": args.step_reads, "server_mode": args.server_mode,
            "saturation_reads_per_s": sat, "parallelism": "single GPU",
        },
        "cpu_baseline": cpu,
        "rates": rates,
        "fastq_ingest": fastq,
        "server_stats": st,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0,
                    help="timed batches (0: whole passes over the read pool, c2 64 = 4 x 16, c3 62 = 2 x 31, "
                         "c4 64; c5 times requests instead)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", choices=("c2", "c3", "c4", "c5"), default="c2")
    ap.add_argument("--slots", type=int, default=6, help="c5: steps in flight")
    ap.add_argument("--step-reads", type=int, default=131072, help="c5: reads per step")
    ap.add_argument("--server-mode", choices=("plain", "fused"), default="plain")
    ap.add_argument("--burst", type=int, default=24000, help="c5: requests of the saturation burst")
    ap.add_argument("--rate-seconds", type=float, default=3.0, help="c5: seconds of arrivals per rate")
    ap.add_argument("--fastq-pairs", type=int, default=400000,
                    help="c5: pairs written to FASTQ and served from the files (0: skip)")
    ap.add_argument("--reads", type=int, default=1_000_000, help="c2: single-end reads")
    ap.add_argument("--pairs", type=int, default=0, help="c3/c4: read pairs (2M / 10M)")
    ap.add_argument("--batch", type=int, default=0, help="reads per batch (c2 65536, c3 131072, c4 262144)")
    ap.add_argument("--lanes", type=lambda x: tuple(int(v) for v in x.split(",")), default=None,
                    help="c2-c4 resident forms: batches in flight (each lane on its own stream and workspace); "
                         "every count given is timed and each form reports its fastest (default 2,3,4 for c2, "
                         "whose batches are short; 2,3 for c3 and c4, where a fourth batch is slower)")
    ap.add_argument("--stream-lanes", type=lambda x: tuple(int(v) for v in x.split(",")), default=(3, 4, 5, 6),
                    help="c2-c4 headline: lane counts tried for the host-to-host pipeline (the fastest is timed)")
    ap.add_argument("--genome", type=int, default=0, help="genome bp (c2/c3 100M, c4 3.1G)")
    ap.add_argument("--cpu-reads", type=int, default=1_000_000,
                    help="bounded sample of the workload timed on the CPU oracle")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: every core this process may run on")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.lanes is None:
        args.lanes = (2, 3, 4) if args.config == "c2" else (2, 3)

    if args.steps <= 0:
        args.steps = {"c2": 64, "c3": 62, "c4": 64, "c5": 64}[args.config]
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the read-sharded "
              "multi-GPU form needs no exchange and is not launched here", file=sys.stderr)
        sys.exit(2)

    # the oracle (parity legs and the CPU baseline) built for this host's CPU
    from oracle import oracle_lib as O
    args.oracle_build = O.use_native()
    import gpu_index_test_amd as G
    if args.config == "c4":
        out = run_c4(args, G)
    elif args.config == "c5":
        out = run_c5(args, G)
    else:
        args.genome = args.genome or 100_000_000
        out = run_c3(args, G) if args.config == "c3" else run_c2(args, G)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
