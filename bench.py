"""Benchmark of the secret-shared iris-matching hot path on one MI355X.

Default workload (the headline line): the full 3-party protocol on one GPU,
both eyes. Three parties each hold a left- and a right-eye DB share of
1,048,576 irises in HBM (u16 code shares of 12,800 coefficients and trimmed
mask shares of 6,400, stored as biased i8 limbs); one step is one batch of
64 queries x 31 rotations through the whole pipeline: per 32,768-row chunk
the six share GEMMs (3 parties x code, mask) with the fused correction and
ChaCha12 zero-share masks, the reshare, the 16 -> 32-bit lift, bit injection
and MSB comparison circuits at the 0.375 threshold, the open into match
bitmaps (OR over rotations), then the AND over eyes and first-match merge.
After the timed steps a serving pass on the same resident left-eye DB
measures request latency (p50 / p99) under Poisson arrivals.

Prints one JSON line (driver contract). Comparisons = query-rotation rows x
DB rows (x 2 eyes). The CPU baseline is the oracle's restatement of the same
3-party step (the reference CPU dot loop iris-mpc-cpu/src/protocol/ops.rs:
238-347 for the three parties' dot shares plus the threshold circuits) timed
on a bounded sample on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "iris-mpc_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

INT8_PEAK_TOPS = 5000.0  # MI355X dense int8 MFMA, 5.0 POPS
HBM_PEAK_GBS = 8000.0
# 32-bit integer VALU lane-ops: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
VALU_PEAK_TOPS = 39.3
# Distinct ChaCha12 blocks per (query-rotation, DB) pair per party in the
# threshold circuits: every AND gate draws one bit per element from each of
# the three pairwise streams of the trio (F(own) ^ F(prev) per party, each
# pairwise block shared by two parties), i.e. one block per 512 elements per
# party and gate: lift 32 ANDs + MSB adder 61 ANDs -> 93 / 512; bit injection
# draws 2 u16 per pair from each pairwise stream -> 4 / 64
CHACHA_BLOCKS_PER_PAIR = 93 / 512 + 4 / 64
# HBM bytes per pair per party of the circuit stage in the scan: the fused
# compare reads the code and mask a shares (4 B; every b share is the previous
# party's a, read by that party's wave), carries and corrections stay on chip;
# the opens read ~3 bits
CIRCUIT_BYTES_PER_PAIR = 4 + 3 / 8
CHACHA12_OPS = 600  # 6 double rounds x 8 quarter rounds x 12 ops + feed-forward and set-up
THRESHOLD_ANON_A = 16384  # 0.375: the prefilter threshold of the full scan
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "share_gemm_c2_traffic.json")
C3_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "share_gemm_c3_traffic.json")
INT8_OPS_PER_PAIR = 115200  # 2 x 3 limb products x (12,800 + 6,400) per pair per party
# kernel behind each rotation-aware GEMM variant id; the automatic choice (-1)
# for raw-share queries is 102 up to two queries (62 rows), else 100
KERNEL_OF_VARIANT = {100: "share_gemm_rot_kernel<4, 2, 4, 1>", 101: "share_gemm_rot_kernel<4, 2, 2, 1>",
                     102: "share_gemm_rot_kernel<2, 1, 4, 2>"}
# the scan's six share GEMMs of a chunk (3 parties x code, mask) as one launch
# of the 4-query kernel (when the scan resolves to variant 100)
JOINT_KERNEL = "share_gemm_rot_heads_kernel<4, 2, 4>"
ROT_SMALL_ROWS = 62
# int8 ops per DB byte at which the share GEMM turns from HBM- to MFMA-bound:
# 5.0 POPS / 8 TB/s; a launch does 3 Qr ops per DB byte
RIDGE_OPS_PER_BYTE = INT8_PEAK_TOPS / HBM_PEAK_GBS * 1e3


def measured_traffic(kernel, path=TRAFFIC_FILE):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (None if absent)."""
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if not any(kernel in k for k in t.get("kernel", [])):
        return None
    return t["traffic_bytes_per_launch"]


PROBE_CHUNKS = 4           # DB chunks of the profiled probe run
PROBE_COPY_BYTES = 1 << 31  # calibration copy: 2 GiB device to device


def read_counters(files, counter):
    """{kernel name: [value per dispatch]} of `counter` in rocprofv3's
    counter-collection CSV files."""
    import csv

    vals = {}
    for fn in files:
        with open(fn, newline="") as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") == counter:
                    vals.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
    return vals


def _pmc_pass(counter, kernel_regex, child_args):
    """One rocprofv3 --pmc pass of `counter` over a child process running the
    traffic probe. Returns ({kernel name: [per-dispatch values]}, None) or
    (None, reason)."""
    import glob
    import shutil
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3")
    if not prof:
        return None, "rocprofv3 not found"
    d = tempfile.mkdtemp(prefix="irismpc_pmc_")
    try:
        cmd = ["timeout", "-s", "KILL", "240", prof, "--pmc", counter, "--kernel-include-regex",
               kernel_regex + "|copyBuffer", "-f", "csv", "-d", d, "-o", "pmc", "--",
               sys.executable, os.path.abspath(__file__), "--workload", "traffic-probe"] + child_args
        r = subprocess.run(cmd, cwd=d, env=dict(os.environ, TMPDIR=d), stdout=subprocess.DEVNULL,
                           stderr=subprocess.PIPE, text=True)
        vals = read_counters(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True),
                             counter)
        # rocprofv3 may end with a signal after the program has finished: the
        # counters are used whenever the file holds them
        if not vals:
            return None, f"rc {r.returncode}, no counters ({r.stderr.strip()[-200:]})"
        return vals, None
    finally:
        shutil.rmtree(d, ignore_errors=True)


def live_traffic(kernel, child_args):
    """HBM bytes per launch of `kernel` measured in this run: two rocprofv3
    --pmc passes (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch) over the
    traffic probe, each counter scaled by what it reads for the probe's 2 GiB
    device copy in the same pass. Returns (bytes, detail) or (None, reason)."""
    name = kernel.split("<")[0]
    out = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        progress(f"rocprofv3 --pmc {counter} pass over the traffic probe")
        vals, err = _pmc_pass(counter, name, child_args)
        if err:
            return None, f"{counter}: {err}"
        gemm = [v for k, vs in vals.items() if name in k for v in vs]
        copies = [v for k, vs in vals.items() if "copyBuffer" in k for v in vs]
        if not gemm or not copies:
            return None, f"{counter}: kernels missing from the counter file"
        scale = PROBE_COPY_BYTES / (max(copies) * 1024.0)
        out[counter] = {"dispatches": len(gemm), "KiB_avg": sum(gemm) / len(gemm),
                        "copy_KiB": max(copies), "scale": scale}
    fetch, write = out["FETCH_SIZE"], out["WRITE_SIZE"]
    traffic = fetch["KiB_avg"] * 1024 * fetch["scale"] + write["KiB_avg"] * 1024 * write["scale"]
    return traffic, out


def _gemm_algorithmic_bytes(Qr):
    """HBM bytes one share GEMM of the scan must move: its chunk's DB limbs
    read once (the average of a code and a mask GEMM) and the u16 results
    written once."""
    db = 32768 * (12800 + 6400) * 2 / 2
    return db + Qr * 32768 * 2


def run_traffic_probe(args):
    """The child that the traffic passes profile: one batch of the 3-party scan
    of one eye over a DB of PROBE_CHUNKS chunks (every share GEMM launch is
    the launch of the headline workload: one 32,768-row chunk's six GEMMs, or
    with a forced variant one party's code or mask GEMM), then one 2 GiB
    device-to-device copy."""
    import torch

    import irismpc

    args.db = PROBE_CHUNKS * irismpc.DB_CHUNK_SIZE
    N, B, Qr, chunk, U, n_plant, eyes, expect = _c3_engines(args, n_eyes=1)
    e = eyes[0]
    torch.cuda.synchronize()
    irismpc.scan_db_trio(e["parties"], e["qcs"], e["qms"], N, B, irismpc.THRESHOLD_ANON_STATS_A,
                         [b.data_ptr() for b in e["bitmaps"]], e["streams"])
    torch.cuda.synchronize()
    x = torch.ones(PROBE_COPY_BYTES // 4, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print("traffic probe done", flush=True)


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo; None where absent)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def with_ratio(line):
    """Adds the host CPU model and the GPU / CPU ratio of the line's value to its cpu_baseline."""
    cb = line.get("cpu_baseline")
    if cb:
        cb["cpu_model"] = cpu_model()
        cb["gpu_over_cpu"] = line["value"] / cb["value"] if cb.get("value") else None
    return line


def progress(msg):
    """A progress line on stderr (stdout carries only the JSON result)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(n_target_s=12.0):
    """Oracle restatement of the reference CPU dot loop on a bounded sample."""
    import numpy as np

    from oracle import orc

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rng = np.random.default_rng(7)
    n = 65536  # DB share rows of the sample
    dc = rng.integers(0, 1 << 16, (n, 12800), dtype=np.uint16)
    dm = rng.integers(0, 1 << 16, (n, 6400), dtype=np.uint16)
    queries, elapsed = 0, 0.0
    while elapsed < n_target_s and queries < 64:  # at most the 64-query batch
        qc = rng.integers(0, 1 << 16, 12800, dtype=np.uint16)
        qm = rng.integers(0, 1 << 16, 6400, dtype=np.uint16)
        t0 = time.perf_counter()
        orc.cpu_rotation_aware_dot(qc, qm, dc, dm, threads=threads)
        elapsed += time.perf_counter() - t0
        queries += 1
    return {
        "value": queries * 31 * n / elapsed,
        "unit": "comparisons/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{queries} queries x 31 rotations vs {n} DB share rows (code 12800 + mask 6400 "
                  f"u16) = {queries * 31 * n} comparisons in {elapsed:.1f} s on {threads} threads; "
                  "same per-pair work as the GPU step; comparisons/s extrapolate linearly in DB rows",
    }


def cpu_baseline_protocol(n_target_s=12.0):
    """Oracle restatement of the full 3-party step on a bounded sample: the three
    parties' rotation-aware code and mask dot shares (the reference CPU dot
    loop) and the threshold circuits of all three parties (lift, bit injection,
    MSB adder, the same schedule as the GPU), for one query x 31 rotations
    against 16,384 DB rows per party; the circuits run as independent parts on
    the threads, which ctypes calls release to."""
    import threading

    import numpy as np

    from oracle import orc

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rng = np.random.default_rng(8)
    n = 16384
    dc = rng.integers(0, 1 << 16, (n, 12800), dtype=np.uint16)
    dm = rng.integers(0, 1 << 16, (n, 6400), dtype=np.uint16)
    qc = rng.integers(0, 1 << 16, 12800, dtype=np.uint16)
    qm = rng.integers(0, 1 << 16, 6400, dtype=np.uint16)
    L = 31 * n
    words = L // 64
    parts = max(1, min(threads, words // 8))
    bounds = [(words * i // parts) * 64 for i in range(parts + 1)]
    shares = [rng.integers(0, 1 << 16, L, dtype=np.uint16) for _ in range(4)]
    so = np.array([[p] * 8 for p in range(3)], np.uint32)
    sp = np.array([[(p + 2) % 3] * 8 for p in range(3)], np.uint32)

    def circuits(lo, hi):
        v = [x[lo:hi] for x in shares]
        orc.threshold([v[0]] * 3, [v[1]] * 3, [v[2]] * 3, [v[3]] * 3, (hi - lo) // 64, 16384, so, sp)

    steps, elapsed = 0, 0.0
    while elapsed < n_target_s and steps < 64:
        t0 = time.perf_counter()
        for _ in range(3):  # each party's own dot shares
            orc.cpu_rotation_aware_dot(qc, qm, dc, dm, threads=threads)
        ts = [threading.Thread(target=circuits, args=(bounds[i], bounds[i + 1])) for i in range(parts)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        elapsed += time.perf_counter() - t0
        steps += 1
    return {
        "value": steps * L / elapsed,
        "unit": "comparisons/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} x (1 query x 31 rotations vs {n} DB rows, one eye): 3 parties' code + mask dot "
                  f"shares and the 3-party threshold circuits = {steps * L} comparisons in {elapsed:.1f} s "
                  f"on {threads} threads; comparisons/s extrapolate linearly in DB rows (the work per pair "
                  "does not depend on the DB size)",
    }


def _c3_engines(args, extra_left=0, n_eyes=2):
    """Both eyes' engines of the full protocol (the left eye's alone with
    n_eyes=1): 3 parties per eye, each party holding the eye's DB share of
    args.db rows in HBM (the left eye's slices with room for extra_left more
    rows), synthetic irises with planted queries (copies of DB irises, 10-30 %
    of code bits flipped)."""
    import numpy as np
    import torch

    import irismpc

    from irismpc import synth

    torch.cuda.set_device(0)
    N, B = args.db, args.queries
    Qr, chunk = B * 31, irismpc.DB_CHUNK_SIZE
    U = min(N, 32768)  # distinct synthetic irises, tiled over the DB (row r = iris r mod U)
    rng = np.random.default_rng(3)
    n_plant = B // 2
    targets = rng.choice(U, size=n_plant, replace=False)
    flips = rng.integers(1280, 3841, size=n_plant)
    planted = [(j, int(t), int(f)) for j, t, f in zip(range(n_plant), targets, flips)]
    base = [np.full(8, p, np.uint32) for p in range(3)]
    eyes = []
    for eye in range(n_eyes):
        ds = 1 + eye  # left / right eye datasets
        ex = irismpc.Exchange(timeout_s=300)
        parties, bitmaps, streams = [], [], []
        for p in range(3):
            seed = lambda n: (irismpc.derive_seed(base[p], n + 8 * eye),  # noqa: E731
                              irismpc.derive_seed(base[(p + 2) % 3], n + 8 * eye))
            codes = irismpc.ShareDB(p, chunk, Qr, irismpc.IRIS_CODE_LENGTH, seed(0), comm=ex, tag=0)
            masks = irismpc.ShareDB(p, chunk, Qr, irismpc.MASK_CODE_LENGTH, seed(1), comm=ex, tag=1)
            if args.variant >= 0:
                codes.set_variant(args.variant)
                masks.set_variant(args.variant)
            circ = irismpc.Circuits(p, Qr * chunk, Qr * chunk // 64, seed(5), ex, tag=2)
            cmp = irismpc.DistanceComparator(Qr, N)
            cap = N + (extra_left if eye == 0 else 0)
            parties.append(irismpc.PartyEngines(codes, masks, circ, cmp, codes.alloc_db(cap), masks.alloc_db(cap)))
            bitmaps.append(torch.zeros(N * B // 64 + 2, dtype=torch.int64, device="cuda"))
            streams.append(irismpc.stream_create())
        synth.build_db([pe.code_db for pe in parties], [pe.mask_db for pe in parties], N, U, ds, ds + 30)
        _, _, shares = synth.make_queries(B, planted, ds, ds + 20, ds + 10, ds + 40)
        qcs = [parties[p].codes.upload_query_shares(shares[p][0]) for p in range(3)]
        qms = [parties[p].masks.upload_query_shares(shares[p][1]) for p in range(3)]
        eyes.append(dict(ex=ex, parties=parties, qcs=qcs, qms=qms, bitmaps=bitmaps, streams=streams))
    expect = np.full(B, 0xFFFFFFFF, np.uint32)
    for j, t, _ in planted:
        expect[j] = t  # first DB row holding iris t
    return N, B, Qr, chunk, U, n_plant, eyes, expect


def _serving_requests(rng, U, N, ds, rand_ds, counter):
    """Request builders of the serving pass over a DB whose row r holds
    synthetic iris (ds, r mod U) shared with randomness (rand_ds, r mod U):
    a quarter of the uniqueness requests are returning users (a DB iris with
    10-30 % of its code bits flipped), reauths target a DB row with 1,500
    flips, insertions add new irises, deletions hit random rows."""
    from irismpc import serving, synth

    def shares(dataset, t, flips):
        c, m = synth.synth_iris(dataset, t)
        if flips:
            c = synth.synth_flip(c, 85, counter[0], flips)
        sh = synth.shares_of(c[None], m[None], 86, counter[0])
        counter[0] += 1
        return [(sh[p][0][0], sh[p][1][0]) for p in range(3)]

    def uniq(k):
        if rng.random() < 0.25:
            return serving.Request("uniqueness", 0, shares=shares(ds, int(rng.integers(U)),
                                                                  int(rng.integers(1280, 3841))))
        return serving.Request("uniqueness", 0, shares=shares(83, counter[0], 0))

    def reauth(k):
        t = int(rng.integers(U))
        return serving.Request("reauth", 0, shares=shares(ds, t, 1500), target=t)

    def insert(k):
        return serving.Request("insert", 0, shares=shares(84, counter[0], 0))

    def delete(k):
        return serving.Request("delete", 0, target=int(rng.integers(N)))

    return uniq, reauth, insert, delete


def _latency_pass(args, parties, N, U):
    """Serving form on the resident left-eye DB shares of the full-protocol
    engines (3 parties x N rows): Poisson arrivals at args.latency_rate per
    second for args.latency_seconds, 85 % uniqueness / 10 % reauth / 5 %
    insertion or deletion, continuous batching of the scan (step sizes from
    the autotuner). Returns the latency summary."""
    import numpy as np

    import irismpc
    from irismpc import serving
    from irismpc.batching import Autotuner

    eng = serving.ServingEngines(N, N + 2 * irismpc.DB_CHUNK_SIZE, max_active=64, seed_base=100,
                                 dbs=([pe.code_db for pe in parties], [pe.mask_db for pe in parties]))
    steps = eng.tune_steps(Autotuner())
    rng = np.random.default_rng(12)
    uniq, reauth, insert, delete = _serving_requests(rng, U, N, 1, 31, [0])
    n = max(16, int(args.latency_rate * args.latency_seconds))
    reqs = serving.poisson_mix(args.latency_rate, n, rng, uniq, reauth, insert, delete)
    srv = serving.Server(eng)
    out = srv.run(reqs)
    span = max(r.done for r in out) - min(r.arrival for r in out)
    comps = sum(31 * r.db_rows for r in out if r.kind == "uniqueness") + sum(31 for r in out if r.kind == "reauth")
    return {"rate_per_s": args.latency_rate, "requests": n, "db_rows": N,
            "mix": "85 % uniqueness (vs the whole DB) / 10 % reauth (vs one row) / 5 % insert+delete",
            "all": serving.latency_stats(out), "uniqueness": serving.latency_stats(out, "uniqueness"),
            "reauth": serving.latency_stats(out, "reauth"), "comparisons_per_s": comps / span,
            "max_concurrent": srv.max_concurrent, "chunks_per_step": {str(b): k for b, k in steps.items()}}


def run_c3(args):
    """Full 3-party protocol on one GPU, both eyes: 3 parties per eye, each
    holding a left- and a right-eye DB share of args.db rows in HBM; share
    GEMM, reshare, lift, threshold circuits, open, OR over rotations, then AND
    over eyes. The share GEMM and circuit launches of the timed steps are timed
    with HIP events on their streams; a serving pass then measures request
    latency on the left eye's DB."""
    import gc

    import numpy as np
    import torch

    import irismpc

    latency = args.latency_rate > 0
    Qr = args.queries * 31
    variant = args.variant if args.variant >= 100 else 100 if Qr > ROT_SMALL_ROWS else 102
    joint = variant == 100  # one launch per chunk for the six GEMMs
    kernel = JOINT_KERNEL if joint else KERNEL_OF_VARIANT[variant]
    gemms_per_launch = 6 if joint else 1
    traffic, traffic_detail = None, "not measured (--no-live-traffic)"
    if args.live_traffic:
        # before this process takes the GPU: the passes profile a child process
        traffic, traffic_detail = live_traffic(kernel, ["--queries", str(args.queries),
                                                        "--variant", str(args.variant)])
        progress(f"share GEMM traffic per launch: {traffic / 1e9:.3f} GB" if traffic else
                 f"live traffic not measured: {traffic_detail}")
    progress("building the 3 parties x 2 eyes DB shares")
    N, B, Qr, chunk, U, n_plant, eyes, expect = _c3_engines(args, 2 * irismpc.DB_CHUNK_SIZE if latency else 0)
    progress("DB shares resident; warm-up")
    finals = [torch.zeros(B, dtype=torch.int32, device="cuda") for _ in range(3)]
    engines = [x for e in eyes for pe in e["parties"] for x in (pe.codes, pe.masks)]
    circuits = [e["parties"][0].circuits for e in eyes]  # party 0 keeps the trio's intervals
    torch.cuda.synchronize()

    def batch():
        for e in eyes:
            for b in e["bitmaps"]:
                b.zero_()
        for f in finals:
            f.fill_(-1)
        for pe in eyes[0]["parties"]:
            pe.comparator.reset()
        torch.cuda.synchronize()
        for e in eyes:  # each eye: the three parties' chunk loop
            irismpc.scan_db_trio(e["parties"], e["qcs"], e["qms"], N, B, irismpc.THRESHOLD_ANON_STATS_A,
                                 [b.data_ptr() for b in e["bitmaps"]], e["streams"])
        for p in range(3):  # each party merges its own opened bitmaps (AND over eyes)
            eyes[0]["parties"][p].comparator.join_db_matches(
                eyes[0]["bitmaps"][p].data_ptr(), eyes[1]["bitmaps"][p].data_ptr(), finals[p].data_ptr(), N)
        torch.cuda.synchronize()

    for x in engines + circuits:
        x.set_timing(True)
    for _ in range(args.warmup):
        batch()
    for x in engines + circuits:
        x.launch_times()  # drop the warm-up intervals
    progress(f"timing {args.steps} batches")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch()
    ms_step = (time.perf_counter() - t0) * 1e3 / args.steps
    progress(f"{ms_step:.1f} ms per batch")
    gemm_n, gemm_ms = 0, 0.0
    for x in engines:
        n, ms = x.launch_times()
        gemm_n, gemm_ms = gemm_n + n, gemm_ms + ms
    circ_n, circ_ms = 0, 0.0
    for x in circuits:
        n, ms = x.launch_times()
        circ_n, circ_ms = circ_n + n, circ_ms + ms
    for x in engines + circuits:
        x.set_timing(False)
    same = all(torch.equal(eyes[e]["bitmaps"][0], eyes[e]["bitmaps"][p]) for e in range(2) for p in (1, 2))
    found = all(np.array_equal(f.cpu().numpy().view(np.uint32), expect) for f in finals)

    pairs = Qr * N  # per eye and party
    gemm_ops = INT8_OPS_PER_PAIR * pairs * 3 * 2 * args.steps
    achieved = gemm_ops / (gemm_ms / 1e3) / 1e12
    circ_pairs = pairs * 3 * 2 * args.steps  # per party, summed over parties and eyes
    chacha_ops = CHACHA_BLOCKS_PER_PAIR * CHACHA12_OPS * circ_pairs
    circ_bytes = CIRCUIT_BYTES_PER_PAIR * circ_pairs
    if traffic:
        traffic_source = ("measured in this run: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes over a child "
                          f"process running one batch of the same scan on a {PROBE_CHUNKS}-chunk DB (same "
                          "launch shape: " + ("one 32,768-row chunk's six GEMMs as one launch" if joint else
                                              "one party's code or mask GEMM over one 32,768-row chunk")
                          + "), each counter scaled by its reading for a 2 GiB device copy in the same pass")
    else:
        traffic = measured_traffic(kernel, C3_TRAFFIC_FILE)
        traffic_source = (f"live passes unavailable ({traffic_detail}); rocprofv3 --pmc runs of this workload "
                          "committed as profiles/share_gemm_c3_traffic.json")
    line = {
        "metric": "3-party iris comparisons/sec on 1 MI355X + p50/p99 request latency",
        "value": 2 * pairs / (ms_step / 1e3),
        "unit": "comparisons/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int8 MFMA (i32 acc) -> u16 shares; u32 lift; GF(2) bit-sliced circuits",
        "data": f"synthetic irises (codes Bernoulli(0.5), masks 0.85) GR-shared for 3 parties on the device; "
                f"{U:,} distinct irises per eye tiled over the DB; {n_plant} of {B} queries are copies of DB "
                "irises with 10-30% of code bits flipped, matched in both eyes",
        "config": {"workload": f"full 3-party protocol on one GPU, both eyes: 3 parties x 2 eyes x {N:,} "
                               f"DB share rows in HBM, {B} queries x 31 rotations per batch (one step); "
                               "comparisons = 2 eyes x query-rotation rows x DB rows",
                   "db_rows": N, "query_rows": Qr, "bitmaps_agree_across_parties": same,
                   "planted_matches_exact": found},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": INT8_PEAK_TOPS, "unit": "TFLOP/s",
                     "frac": achieved / INT8_PEAK_TOPS,
                     "traffic": traffic,
                     "traffic_source": traffic_source,
                     "traffic_passes": traffic_detail if isinstance(traffic_detail, dict) else None,
                     "algorithmic_bytes_per_launch": gemms_per_launch * _gemm_algorithmic_bytes(Qr),
                     "kernel": kernel,
                     "launches": gemm_n, "avg_launch_ms": gemm_ms / max(gemm_n, 1),
                     "int8_ops_per_launch_avg": gemm_ops / max(gemm_n, 1),
                     "basis": ("share GEMM launches inside the timed batches, HIP events on their stream: "
                               + ("one per chunk, the chunk's six GEMMs (3 parties x code, mask) as one grid"
                                  if joint else "6 per chunk (3 parties x code, mask), in order on one stream")
                               + "; 115,200 int8 ops per pair per party / launch duration")},
        "streaming_stage": {"bound": "valu", "kernels": "compare_fused_kernel (lift, injection, MSB adder)",
                            "compares": circ_n, "avg_compare_ms": circ_ms / max(circ_n, 1),
                            "chacha12_ops": chacha_ops,
                            "achieved_Tops": chacha_ops / (circ_ms / 1e3) / 1e12 if circ_ms else None,
                            "peak_Tops": VALU_PEAK_TOPS,
                            "frac": chacha_ops / (circ_ms / 1e3) / 1e12 / VALU_PEAK_TOPS if circ_ms else None,
                            "hbm_GBps": circ_bytes / (circ_ms / 1e3) / 1e9 if circ_ms else None,
                            "hbm_frac": circ_bytes / (circ_ms / 1e3) / 1e9 / HBM_PEAK_GBS if circ_ms else None,
                            "note": f"{CHACHA_BLOCKS_PER_PAIR:.3f} distinct ChaCha12 blocks x {CHACHA12_OPS} "
                                    f"int32 ops and {CIRCUIT_BYTES_PER_PAIR:.1f} HBM bytes per pair per party; "
                                    "one compare = the fused lift + injection + MSB adder of one chunk, three parties, "
                                    "HIP events on the protocol stream while the next chunk's GEMMs run"},
    }
    if latency:
        # the right eye is done with: its DB shares leave HBM before the serving pass
        left = eyes[0]["parties"]
        del eyes[1:]
        eyes[0]["qcs"] = eyes[0]["qms"] = None
        del engines, circuits, finals
        gc.collect()
        torch.cuda.synchronize()
        progress(f"serving latency pass at {args.latency_rate:g} requests/s")
        lat = _latency_pass(args, left, N, U)
        progress(f"p50 {lat['all']['p50_ms']:.1f} ms, p99 {lat['all']['p99_ms']:.1f} ms")
        line["latency_p50_ms"] = lat["all"]["p50_ms"]
        line["latency_p99_ms"] = lat["all"]["p99_ms"]
        line["latency"] = lat
    if not args.no_cpu_baseline:
        progress("CPU baseline")
    line["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline_protocol()
    return line


def cpu_baseline_c1():
    """The CPU-runnable configuration at full size on the oracle: the three
    parties' rotation-aware code and mask dot shares (the reference CPU dot
    loop) of 32 queries x 31 rotations against 4,096 DB rows (2.34e11 u16 MACs)
    and the 3-party threshold circuits over all 992 x 4,096 elements (split
    over the threads, which ctypes calls release to). Shares are random: the
    work does not depend on them."""
    import threading

    import numpy as np

    from oracle import orc

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rng = np.random.default_rng(9)
    N, B = 4096, 32
    dc = [rng.integers(0, 1 << 16, (N, 12800), dtype=np.uint16) for _ in range(3)]
    dm = [rng.integers(0, 1 << 16, (N, 6400), dtype=np.uint16) for _ in range(3)]
    qc = [rng.integers(0, 1 << 16, (B, 12800), dtype=np.uint16) for _ in range(3)]
    qm = [rng.integers(0, 1 << 16, (B, 6400), dtype=np.uint16) for _ in range(3)]
    L = B * 31 * N
    words = L // 64
    parts = max(1, min(threads, words // 8))
    bounds = [(words * i // parts) * 64 for i in range(parts + 1)]
    shares = [rng.integers(0, 1 << 16, L, dtype=np.uint16) for _ in range(4)]
    so = np.array([[p] * 8 for p in range(3)], np.uint32)
    sp = np.array([[(p + 2) % 3] * 8 for p in range(3)], np.uint32)

    def circuits(lo, hi):
        v = [x[lo:hi] for x in shares]
        orc.threshold([v[0]] * 3, [v[1]] * 3, [v[2]] * 3, [v[3]] * 3, (hi - lo) // 64, 16384, so, sp)

    t0 = time.perf_counter()
    for p in range(3):
        for j in range(B):
            orc.cpu_rotation_aware_dot(qc[p][j], qm[p][j], dc[p], dm[p], threads=threads)
    t_dot = time.perf_counter() - t0
    t1 = time.perf_counter()
    ts = [threading.Thread(target=circuits, args=(bounds[i], bounds[i + 1])) for i in range(parts)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    t_circ = time.perf_counter() - t1
    total = t_dot + t_circ
    return {"value": B * 31 * N / total, "unit": "comparisons/s", "cores": threads, "kind": "port",
            "sample": f"the whole configuration: 3 parties' code + mask dot shares of {B} queries x 31 "
                      f"rotations vs {N} DB rows ({3 * B * 31 * N * 19200:.3g} u16 MACs) in {t_dot:.2f} s "
                      f"and the 3-party circuits over {L} elements in {t_circ:.2f} s on {threads} threads"}


def run_c1(args):
    """The CPU-runnable configuration on the GPU: DB of 4,096 synthetic irises,
    32 queries x 31 rotations, 8 of them copies of DB irises with exactly 1,920
    code bits flipped, three parties (engine seeds derive_seed(([p; 8],
    [p - 1; 8]), nonce 0 / 1 / 5)); one batch = the three parties' scan (one
    chunk). The planted targets and the expected 0.375 bitmap are the committed
    golden vectors, which every timed batch must reproduce."""
    import numpy as np
    import torch

    import irismpc
    from irismpc import synth

    torch.cuda.set_device(0)
    golden = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "config1.npz"))
    N, B, A, FLIPS = 4096, 32, 16384, 1920
    Qr, chunk = B * 31, irismpc.DB_CHUNK_SIZE
    planted = [(j, int(t), FLIPS) for j, t in zip(range(0, B, 4), golden["targets"])]
    ex = irismpc.Exchange(timeout_s=120)
    parties, bitmaps, streams = [], [], []
    for p in range(3):
        def seed(n):
            return (irismpc.derive_seed(np.full(8, p, np.uint32), n),
                    irismpc.derive_seed(np.full(8, (p + 2) % 3, np.uint32), n))

        codes = irismpc.ShareDB(p, chunk, Qr, irismpc.IRIS_CODE_LENGTH, seed(0), comm=ex, tag=0)
        masks = irismpc.ShareDB(p, chunk, Qr, irismpc.MASK_CODE_LENGTH, seed(1), comm=ex, tag=1)
        circ = irismpc.Circuits(p, Qr * chunk, Qr * chunk // 64, seed(5), ex, tag=2)
        parties.append(irismpc.PartyEngines(codes, masks, circ, irismpc.DistanceComparator(Qr, N),
                                            codes.alloc_db(N), masks.alloc_db(N)))
        bitmaps.append(torch.zeros(N * B // 64 + 2, dtype=torch.int64, device="cuda"))
        streams.append(irismpc.stream_create())
    synth.build_db([pe.code_db for pe in parties], [pe.mask_db for pe in parties], N, N, 1, 30)
    _, _, shares = synth.make_queries(B, planted, 1, 21, 11, 31)
    qcs = [parties[p].codes.upload_query_shares(shares[p][0]) for p in range(3)]
    qms = [parties[p].masks.upload_query_shares(shares[p][1]) for p in range(3)]
    torch.cuda.synchronize()
    want = golden["bitmap"]

    def batch():
        for b, s in zip(bitmaps, streams):
            irismpc.fill_u64(b.data_ptr(), b.numel(), 0, s)
        irismpc.scan_db_trio(parties, qcs, qms, N, B, A, [b.data_ptr() for b in bitmaps], streams)
        for s in streams:
            irismpc.stream_synchronize(s)

    for _ in range(args.warmup):
        batch()
    exact = True
    t = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        batch()
        t.append(time.perf_counter() - t0)
        exact = exact and all(np.array_equal(b.cpu().numpy().view(np.uint64)[:N * B // 64], want) for b in bitmaps)
    ms = float(np.median(t)) * 1e3
    return {
        "metric": "3-party iris comparisons/sec on 1 MI355X + p50/p99 request latency",
        "value": Qr * N / (ms / 1e3),
        "unit": "comparisons/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int8 MFMA (i32 acc) -> u16 shares; u32 lift; GF(2) bit-sliced circuits",
        "data": "synthetic irises GR-shared for 3 parties (the golden-vector configuration)",
        "config": {"workload": "CPU-runnable configuration: 3 parties, DB 4,096 rows, 32 queries x 31 "
                               "rotations (8 copies with 1,920 flips), one eye; median batch of the three "
                               "parties' scan, host synchronisation included",
                   "db_rows": N, "query_rows": Qr, "bitmaps_equal_golden_vectors": bool(exact)},
        "roofline": None,
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline_c1(),
    }


def run_c3_two_stage(args):
    """The reference's batch answer on the full-protocol engines: one eye's full
    scan at the 0.375 prefilter, the candidate subset, both eyes against the
    subset at 0.345, AND over eyes (actor.rs:939-1131, 1900-2244)."""
    import numpy as np
    import torch

    from irismpc.two_stage import TwoStageMatcher

    N, B, Qr, chunk, U, n_plant, eyes, expect = _c3_engines(args)
    matcher = TwoStageMatcher({"left": eyes[0]["parties"], "right": eyes[1]["parties"]}, N, B)
    queries = {"left": (eyes[0]["qcs"], eyes[0]["qms"]), "right": (eyes[1]["qcs"], eyes[1]["qms"])}
    torch.cuda.synchronize()
    res = None
    for _ in range(args.warmup):
        res = matcher.match(queries, N, B)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = matcher.match(queries, N, B)
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / args.steps
    exact = all(np.array_equal(f, expect) for f in res.final_results)
    return {
        "metric": "3-party two-stage batch on 1 MI355X: DB-row decisions per second",
        "value": 2 * Qr * N / (ms_step / 1e3),
        "unit": "comparisons/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int8 MFMA (i32 acc) -> u16 shares; u32 lift; GF(2) bit-sliced circuits",
        "data": f"synthetic irises GR-shared for 3 parties on the device; {U:,} distinct irises per eye tiled "
                f"over the DB; {n_plant} of {B} queries are copies of DB irises with 10-30% of code bits "
                "flipped, matched in both eyes",
        "config": {"workload": f"two-stage flow of the reference's batch: left eye full scan at 0.375 over "
                               f"{N:,} rows, candidate subset, both eyes on the candidates at 0.345, AND over "
                               f"eyes; 3 parties, {B} queries x 31 rotations; value counts the 2 eyes x "
                               "query-rotation rows x DB rows pairs the batch decides, the right eye only "
                               "computing the candidates",
                   "db_rows": N, "query_rows": Qr, "candidates": int(res.candidates.size),
                   "final_results_exact": exact},
        "roofline": None,
    }


def run_c4(args):
    """Batching form: one party's share GEMM (code and trimmed mask, masked)
    over a DB share of args.db rows for 200 query batches of seeded
    log-uniform size in 1..256 queries. The plain form runs batch after batch;
    the batching form packs consecutive batches (pack size and GEMM variant
    picked by the autotuner on a DB sample) and replays the whole packed
    schedule from one HIP graph. Both forms produce the same shares."""
    import numpy as np
    import torch

    import irismpc
    from irismpc.batching import Autotuner, BatchRunner, Matrix, log_uniform_batch_sizes

    torch.cuda.set_device(0)
    N, chunk = args.db, irismpc.DB_CHUNK_SIZE
    sizes = log_uniform_batch_sizes(200, seed=4)
    rng = np.random.default_rng(5)
    ex = irismpc.Exchange()
    seeds = [irismpc.derive_seed([0] * 8, 0), irismpc.derive_seed([2] * 8, 0)]
    mats = []
    for tag, (K, mult) in enumerate(((irismpc.IRIS_CODE_LENGTH, 1), (irismpc.MASK_CODE_LENGTH, 2))):
        eng = irismpc.ShareDB(0, chunk, 256 * 31, K, seeds, comm=ex, tag=tag)
        db = eng.alloc_db(N)
        db.fill_random(0, N, np.arange(8) + 100 * tag, 0)
        shares = [rng.integers(0, 1 << 16, (int(b), K), dtype=np.uint16) for b in sizes]
        mats.append(Matrix(eng, db, N, mult, shares))
    stream = irismpc.stream_create()
    torch.cuda.synchronize()

    progress(f"{N:,} DB rows filled; autotuning GEMM variants and pack size")
    tuner = Autotuner()
    sample = min(N, 4 * chunk)
    probe = mats[0].engine.upload_query_shares_rows(mats[0].shares[int(np.argmax(sizes))], int(sizes.max()) * 31)
    variants = {}
    for m in mats:
        q = probe if m is mats[0] else m.engine.upload_query_shares_rows(
            m.shares[int(np.argmax(sizes))], int(sizes.max()) * 31)
        variants[m.engine.code_length] = tuner.gemm_variant(m, q, int(sizes.max()) * 31, chunk, stream)
        m.engine.set_variant(variants[m.engine.code_length])

    del probe
    torch.cuda.empty_cache()
    runner = BatchRunner(mats, sizes, chunk, stream, max_queries=256)
    max_q = tuner.pack_size(runner, sample)
    runner.capture()
    torch.cuda.synchronize()

    progress("plain form: batch after batch")
    t0 = time.perf_counter()
    runner.run_plain()
    t_plain = time.perf_counter() - t0
    plain = [[o.clone() for o in outs] for outs in runner.outs]
    for outs in runner.outs:
        for o in outs:
            o.zero_()
    torch.cuda.synchronize()
    progress(f"plain form {t_plain:.1f} s; packed form, one launch per pack and chunk")
    t0 = time.perf_counter()
    runner.run_packed(graph=True)
    t_packed = time.perf_counter() - t0
    same = all(torch.equal(a, b) for pa, pb in zip(plain, runner.outs) for a, b in zip(pa, pb))
    # the persistent form: one launch per chunk and matrix covering every pack
    for outs in runner.outs:
        for o in outs:
            o.zero_()
    torch.cuda.synchronize()
    runner.capture(persistent=True)
    progress(f"packed form {t_packed:.1f} s; persistent form, one launch per chunk")
    t0 = time.perf_counter()
    runner.run_packed(graph=True, persistent=True)
    t_persistent = time.perf_counter() - t0
    same_persistent = all(torch.equal(a, b) for pa, pb in zip(plain, runner.outs) for a, b in zip(pa, pb))
    comps = runner.comparisons()
    ops = 115200 * comps
    del plain
    # the reference's per-batch form pads every batch's query rows to
    # max_batch_size x 31 (actor.rs:895-901, share_db.rs:800-845): every padded
    # batch does the same work, so one (the smallest batch, padded to 256
    # queries) is timed over the whole DB and the 200 batches take 200 x that
    progress(f"persistent form {t_persistent:.1f} s; one padded batch")
    small = int(np.argmin(sizes))
    pad_rows = 256 * 31
    t_pad = float("inf")
    pad_out = torch.empty(chunk * pad_rows + 64, dtype=torch.int16, device="cuda")
    pad_q = [m.engine.upload_query_shares_rows(m.shares[small], pad_rows) for m in mats]
    for rep in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for m, q in zip(mats, pad_q):
            off = 0
            while off < N:
                tot = -(-min(chunk, N - off) // 64) * 64
                m.engine.dot_reduce_range(q, m.db, off, tot, chunk, m.mult, pad_out.data_ptr(), stream)
                off += chunk
        irismpc.stream_synchronize(stream)
        t_pad = min(t_pad, time.perf_counter() - t0)
    t_padded = t_pad * len(sizes)
    return {
        "metric": "per-party share GEMM comparisons/sec on 1 MI355X, batching form (one party's code + mask dot shares)",
        "value": comps / t_packed,
        "unit": "comparisons/s",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": t_packed * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int8 MFMA (i32 acc) -> u16 shares",
        "data": "synthetic: uniformly random u16 DB and query shares",
        "config": {"workload": f"batching form, per-party share GEMM (code + trimmed mask, masked): DB "
                               f"{N:,} rows, 200 batches of log-uniform 1-256 queries x 31 rotations "
                               f"({int(sizes.sum())} queries); packed into {len(runner.plan)} launches per "
                               "chunk and matrix, one HIP graph",
                   "db_rows": N, "batches": 200, "queries": int(sizes.sum()),
                   "packs": len(runner.plan), "max_queries_per_pack": max_q,
                   "gemm_variants": variants, "packed_equals_plain": same},
        "packed_persistent": {"value": comps / t_persistent, "ms": t_persistent * 1e3,
                              "form": "the same packs, one persistent launch per chunk and matrix (a block per CU "
                                      "looping over every pack's tiles), one HIP graph",
                              "equals_plain": same_persistent},
        "plain_form": {"value": comps / t_plain, "ms": t_plain * 1e3,
                       "form": "batch after batch, each batch's own rows (unpadded)"},
        "speedup_vs_plain": t_plain / t_packed,
        "plain_form_padded": {"value": comps / t_padded, "ms": t_padded * 1e3, "ms_per_batch": t_pad * 1e3,
                              "form": "batch after batch, every batch padded to 256 queries x 31 rows as the "
                                      "reference pads to max_batch_size; one padded batch timed over the whole DB, "
                                      "x 200 (every padded batch does the same work)"},
        "speedup_vs_padded_plain": t_padded / t_packed,
        "roofline": {"bound": "mfma", "achieved": ops / t_packed / 1e12, "peak": INT8_PEAK_TOPS,
                     "unit": "TFLOP/s", "frac": ops / t_packed / 1e12 / INT8_PEAK_TOPS, "traffic": None},
        "autotune": {k: v for k, v in tuner.cache.items()},
        # same per-pair work as one party's masked code + mask dot shares
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline(),
    }


def run_c5(args):
    """Serving form against the plain per-batch form: Poisson arrivals at
    fixed rates, 85 % uniqueness (vs the whole DB), 10 % reauth (vs one row),
    5 % insertions / deletions, over the three parties' DB shares of args.db
    rows (one eye). The serving form batches continuously at chunk-step
    granularity; the plain per-batch form runs the reference's batch after
    batch, every batch padded to max_batch x 31 query rows (64, the batch of
    the full-protocol workload, and 8, the reference's deployment setting).
    Each rate serves the same request sequence in every form; the line reports
    p50 / p99 per form and rate and, per p99 bound, the largest tested rate
    each form sustains."""
    import dataclasses

    import numpy as np
    import torch

    import irismpc
    from irismpc import serving
    from irismpc.batching import Autotuner

    torch.cuda.set_device(0)
    N, U = args.db, min(args.db, 32768)
    rates = [float(r) for r in args.rates.split(",")]
    progress(f"building 3 x {N:,} DB shares")
    eng = serving.ServingEngines(N, N + 4 * irismpc.DB_CHUNK_SIZE, max_active=64)
    eng.build(U, 81, 82)
    rng = np.random.default_rng(11)
    uniq, reauth, insert, delete = _serving_requests(rng, U, N, 81, 82, [0])
    progress("tuning step sizes")
    tuner = Autotuner()
    step_table = eng.tune_steps(tuner)

    # one uniqueness request at a time, a full scan each
    alone = [uniq(k).shares for k in range(args.plain_requests)]
    eng.run_alone(alone[0])  # warm-up
    t0 = time.perf_counter()
    for sh in alone:
        eng.run_alone(sh)
    t_one = (time.perf_counter() - t0) / len(alone)

    def comparisons(out):
        return sum(31 * r.db_rows for r in out if r.kind == "uniqueness") + sum(31 for r in out if r.kind == "reauth")

    def serve(form, reqs):
        reqs = [dataclasses.replace(r) for r in reqs]
        if form == "serving":
            srv = serving.Server(eng)
        else:
            srv = serving.PerBatchServer(eng, int(form.split("_")[-1]))
        out = srv.run(reqs)
        span = max(r.done for r in out) - min(r.arrival for r in out)
        entry = {"comparisons_per_s": comparisons(out) / span, "requests_per_s": len(out) / span,
                 "all": serving.latency_stats(out), "uniqueness": serving.latency_stats(out, "uniqueness"),
                 "reauth": serving.latency_stats(out, "reauth")}
        if form == "serving":
            entry.update(max_concurrent=srv.max_concurrent, chunk_steps=srv.steps)
        else:
            entry.update(batches=srv.batches)
        return entry, out

    forms = ["serving", "per_batch_64", "per_batch_8"]
    per_rate = []
    for rate in rates:
        n = max(16, int(rate * args.seconds))
        reqs = serving.poisson_mix(rate, n, rng, uniq, reauth, insert, delete)
        entry = {"rate_per_s": rate, "requests": n}
        for form in forms:
            progress(f"{form} at {rate:g} requests/s")
            entry[form], _ = serve(form, reqs)
        per_rate.append(entry)
    # capacity: a burst of 3 x 64 uniqueness requests at once (offered load above capacity)
    burst = [uniq(k) for k in range(3 * eng.max_active)]
    for k, r in enumerate(burst):
        r.arrival, r.rid = 0.0, k
    capacity = {}
    for form in forms:
        progress(f"{form}: capacity burst")
        e, _ = serve(form, burst)
        capacity[form] = e
    # each request's matches are the same in the mix as alone (the plain form), on the DB as it is now
    mixed = [serving.Request("uniqueness", 0.001 * k, shares=sh, rid=k) for k, sh in enumerate(alone)]
    out = sorted(serving.Server(eng).run(mixed), key=lambda r: r.rid)
    same = [r.result for r in out] == [eng.run_alone(sh) for sh in alone]

    bounds = [100.0, 150.0, 200.0, 250.0, 500.0, 1000.0, 2000.0]
    sustained = {}
    for b in bounds:
        row = {}
        for form in forms:
            ok = [e for e in per_rate if e[form]["all"]["p99_ms"] <= b]
            best = max(ok, key=lambda e: e["rate_per_s"]) if ok else None
            row[form] = {"rate_per_s": best["rate_per_s"] if best else 0.0,
                         "comparisons_per_s": best[form]["comparisons_per_s"] if best else 0.0}
        sustained[f"p99<={b:g}ms"] = row
    headline_bound = "p99<=250ms"
    h = sustained[headline_bound]
    # the latency quoted beside the value is that of the rate the value is taken at
    # (the lowest tested rate when no rate meets the bound)
    quoted = next((e for e in per_rate if e["rate_per_s"] == h["serving"]["rate_per_s"]), per_rate[0])
    top = quoted["serving"]
    return {
        "metric": "3-party iris comparisons/sec on 1 MI355X + p50/p99 request latency",
        "value": h["serving"]["comparisons_per_s"],
        "unit": "comparisons/s",
        "n_gpus": 1, "steps": len(rates), "warmup": 1, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int8 MFMA (i32 acc) -> u16 shares; u32 lift; GF(2) bit-sliced circuits",
        "data": f"synthetic irises GR-shared for 3 parties ({U:,} distinct, tiled); 25 % of uniqueness "
                "requests are planted copies with 10-30 % flips",
        "config": {"workload": f"serving form, full 3-party protocol, one eye: DB {N:,} rows per party "
                               "share; Poisson arrivals, 85 % uniqueness / 10 % reauth / 5 % insert+delete; "
                               "continuous batching over 32,768-row chunk steps, up to 64 active queries; "
                               "value = comparisons/s at the largest tested rate whose p99 <= 250 ms",
                   "db_rows": N, "rates_per_s": rates, "seconds_per_rate": args.seconds,
                   "results_same_alone_and_in_mix": same},
        "latency_p50_ms": top["all"]["p50_ms"], "latency_p99_ms": top["all"]["p99_ms"],
        "latency_rate_per_s": quoted["rate_per_s"],
        "per_rate": per_rate,
        "capacity": capacity,
        "sustained_under_p99_bound": sustained,
        "speedup_vs_per_batch_at_p99_250ms": {
            f: (h["serving"]["comparisons_per_s"] / h[f]["comparisons_per_s"]) if h[f]["comparisons_per_s"] else None
            for f in ("per_batch_64", "per_batch_8")},
        "one_request_at_a_time": {"ms_per_uniqueness_request": t_one * 1e3,
                                  "comparisons_per_s": 31 * N / t_one, "requests": len(alone)},
        "autotune": {"chunks_per_step": {str(b): k for b, k in step_table.items()},
                     "log": {k: v for k, v in tuner.log}},
        "roofline": None,
        # same work per comparison as a serving request: 3 parties' dot shares + threshold circuits
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline_protocol(),
    }


def run_matmul(args):
    """The reference's own GEMM benchmark (iris-mpc-gpu/benches/matmul.rs:17-60)
    on one GPU: DB 2,400,000 x 12,800 u16 code shares, 930 arbitrary query rows
    (no rotation structure), an engine without exchange (no masks). Each
    iteration uploads and prepares the query (host -> device copy, limb split,
    sums) and runs the fused dot + correction over the whole DB, as the
    reference's iteration does (htod_transfer_query, query_sums, dot,
    dot_reduce, await_streams). Its criterion throughput counts DB x 930 / 31
    elements per iteration."""
    import numpy as np
    import torch

    import irismpc

    torch.cuda.set_device(0)
    N, Qr, K = 2_400_000, 930, irismpc.IRIS_CODE_LENGTH
    stream = irismpc.stream_create()
    xs = torch.cuda.ExternalStream(stream)
    eng = irismpc.ShareDB(0, N, Qr, K, ([0] * 8, [0] * 8))
    if args.variant != -1:
        eng.set_variant(args.variant)
    db = eng.alloc_db(N)
    db.fill_random(0, N, np.arange(8) + 7, 0, stream)
    q = np.random.default_rng(42).integers(0, 1 << 16, (Qr, K), dtype=np.uint16)
    irismpc.stream_synchronize(stream)
    kev = []

    def iteration(ev=None):
        dq = eng.upload_query(q, stream)
        if ev is not None:
            ev[0].record(xs)
        eng.dot_reduce(dq, db, N, 0, 1, stream)
        if ev is not None:
            ev[1].record(xs)
        irismpc.stream_synchronize(stream)
        return dq

    for _ in range(args.warmup):
        iteration()
    times = []
    for _ in range(args.steps):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        t0 = time.perf_counter()
        dq = iteration(ev)
        times.append(time.perf_counter() - t0)
        kev.append(ev)
        del dq
    t = sum(times) / len(times)
    gemm_ms = sum(e[0].elapsed_time(e[1]) for e in kev) / len(kev)
    ops = 6 * K * Qr * N
    achieved = ops / (gemm_ms / 1e3) / 1e12
    variant = args.variant if args.variant != -1 else 0  # automatic: 0 for rows without rotation structure
    return {
        "metric": "per-party share GEMM comparisons/sec on 1 MI355X, the reference's matmul bench shape (code only)",
        "value": Qr * N / t,
        "unit": "comparisons/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int8 MFMA (i32 acc) -> u16 shares",
        "data": "synthetic: uniformly random u16 DB and query rows",
        "config": {"workload": "the reference's matmul bench: DB 2,400,000 x 12,800 u16 code shares, "
                               "930 query rows without rotation structure, no masks; per iteration the "
                               "query upload + preparation and the fused dot + correction",
                   "db_rows": N, "query_rows": Qr, "gemm_variant": variant},
        "reference_throughput_elements_per_s": N * Qr / 31 / t,
        "gemm_ms": gemm_ms,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": INT8_PEAK_TOPS, "unit": "TFLOP/s",
                     "frac": achieved / INT8_PEAK_TOPS, "traffic": None},
    }


def run_circuits(args):
    """Threshold circuits alone at one chunk of the full protocol: the three
    parties' lift, bit injection, MSB adder and open over 32,768 DB rows x
    args.queries x 31 rotations of uniformly random code/mask dot shares."""
    import numpy as np
    import torch

    import irismpc

    torch.cuda.set_device(0)
    Qr = args.queries * 31
    L = Qr * irismpc.DB_CHUNK_SIZE
    W = L // 64
    ex = irismpc.Exchange(timeout_s=300)
    seeds = [(np.full(8, p, np.uint32), np.full(8, (p + 2) % 3, np.uint32)) for p in range(3)]
    circ = [irismpc.Circuits(p, L, W, seeds[p], ex) for p in range(3)]
    streams = [irismpc.stream_create() for _ in range(3)]
    gen = torch.Generator(device="cuda").manual_seed(7)
    # each party's a shares; its b shares are the previous party's a buffers,
    # as the scan's reshare hands them over (zero-copy views)
    bufs = [[torch.randint(-32768, 32767, (L,), dtype=torch.int16, device="cuda", generator=gen)
             for _ in range(2)] for _ in range(3)]
    torch.cuda.synchronize()
    codes = [(bufs[p][0].data_ptr(), bufs[(p + 2) % 3][0].data_ptr()) for p in range(3)]
    masks = [(bufs[p][1].data_ptr(), bufs[(p + 2) % 3][1].data_ptr()) for p in range(3)]

    def step():
        irismpc.Circuits.compare_threshold_masked_many_trio(circ, codes, masks, L, THRESHOLD_ANON_A, streams)
        irismpc.Circuits.open_trio(circ, streams)

    for _ in range(args.warmup):
        step()
    for st in streams:
        irismpc.stream_synchronize(st)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    for st in streams:
        irismpc.stream_synchronize(st)
    ms_step = (time.perf_counter() - t0) * 1e3 / args.steps
    ops = CHACHA_BLOCKS_PER_PAIR * CHACHA12_OPS * L * 3
    achieved = ops / (ms_step / 1e3) / 1e12
    return {
        "metric": "threshold-circuit comparisons/sec (3 parties, one DB chunk)",
        "value": L / (ms_step / 1e3), "unit": "comparisons/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "GF(2) bit-sliced circuits on u16/u32 shares, ChaCha12 correlated randomness",
        "data": "uniformly random u16 code and mask dot shares per party (a shares; each b share is the "
                "previous party's a buffer, as after the reshare)",
        "config": {"workload": f"lift, bit injection, MSB adder and open of 3 parties over one chunk: "
                               f"{Qr:,} query-rotation rows x {irismpc.DB_CHUNK_SIZE:,} DB rows "
                               f"({W:,} result words)", "threshold_a": THRESHOLD_ANON_A},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_TOPS, "unit": "TOP/s",
                     "frac": achieved / VALU_PEAK_TOPS, "traffic": None,
                     "stage": f"{CHACHA_BLOCKS_PER_PAIR:.3f} ChaCha12 blocks x {CHACHA12_OPS} int32 ops "
                              "per pair per party"},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--db", type=int, default=1 << 20)
    ap.add_argument("--queries", type=int, default=64)
    ap.add_argument("--variant", type=int, default=-1,
                    help="GEMM kernel variant (-1: automatic, the rotation-aware kernel)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=("c1", "c2", "c3", "c3-two-stage", "c4", "c5", "circuits", "matmul",
                                           "traffic-probe"),
                    default="c3",
                    help="c3 (the headline line): full 3-party protocol, both eyes, plus a serving latency "
                         "pass; c1: the CPU-runnable configuration (4,096 rows, 32 queries) against its "
                         "golden vectors, with the oracle at full size as CPU baseline; c2: one party's share GEMM; c4: batching form (use --db 4194304 for the "
                         "full-size DB); "
                         "c5: serving form (use --db 2097152 for the full-size DB); circuits: the "
                         "threshold circuits alone at one chunk; c3-two-stage: the reference's batch flow "
                         "(one-eye prefilter scan, both eyes on the candidates); matmul: the reference's "
                         "own GEMM benchmark shape (2.4M x 930 rows, code only, query upload included)")
    ap.add_argument("--rates", default="10,20,40,50,60,70,80", help="c5: request rates per second")
    ap.add_argument("--seconds", type=float, default=6.0, help="c5: arrival window per rate")
    ap.add_argument("--plain-requests", type=int, default=8, help="c5: requests timed in the plain form")
    ap.add_argument("--latency-rate", type=float, default=40.0,
                    help="c3: Poisson request rate of the serving latency pass (0: no pass)")
    ap.add_argument("--latency-seconds", type=float, default=6.0, help="c3: arrival window of that pass")
    ap.add_argument("--no-live-traffic", dest="live_traffic", action="store_false",
                    help="c3: skip the two rocprofv3 --pmc passes that measure the share GEMM's HBM traffic "
                         "(the committed profile's figure is reported instead)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the three parties share one GPU; no multi-GPU path is measured)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "c3":
        print(json.dumps(with_ratio(run_c3(args))))
        return
    if args.workload == "traffic-probe":
        run_traffic_probe(args)
        return
    if args.workload == "c1":
        print(json.dumps(with_ratio(run_c1(args))))
        return
    if args.workload == "circuits":
        print(json.dumps(run_circuits(args)))
        return
    if args.workload == "c3-two-stage":
        print(json.dumps(run_c3_two_stage(args)))
        return
    if args.workload == "c4":
        print(json.dumps(with_ratio(run_c4(args))))
        return
    if args.workload == "c5":
        print(json.dumps(with_ratio(run_c5(args))))
        return
    if args.workload == "matmul":
        print(json.dumps(run_matmul(args)))
        return

    import numpy as np
    import torch

    import irismpc

    torch.cuda.set_device(0)
    N, B = args.db, args.queries
    Qr = B * 31
    chunk = irismpc.DB_CHUNK_SIZE
    variant = args.variant
    stream = irismpc.stream_create()
    xs = torch.cuda.ExternalStream(stream)

    seeds = [irismpc.derive_seed([0] * 8, 0), irismpc.derive_seed([2] * 8, 0)]
    ex = irismpc.Exchange()
    engines, dbs, queries, outs = [], [], [], []
    rng = np.random.default_rng(1)
    for tag, K in enumerate((irismpc.IRIS_CODE_LENGTH, irismpc.MASK_CODE_LENGTH)):
        eng = irismpc.ShareDB(0, chunk, Qr, K, seeds, comm=ex, tag=tag)
        eng.set_variant(variant)
        db = eng.alloc_db(N)
        db.fill_random(0, N, np.arange(8) + 100 * tag, 0, stream)
        # raw query shares: Lagrange/basis preprocessing and the 31 rotations run on the device
        q = eng.upload_query_shares(rng.integers(0, 1 << 16, (B, K), dtype=np.uint16), stream)
        out = torch.empty(N * Qr + 64, dtype=torch.int16, device="cuda")
        engines.append(eng)
        dbs.append(db)
        queries.append(q)
        outs.append(out)
    irismpc.stream_synchronize(stream)

    def step(ev=None):
        for i, mult in ((0, 1), (1, 2)):
            if ev is not None:
                ev[2 * i].record(xs)
            engines[i].dot_reduce_range(queries[i], dbs[i], 0, N, chunk, mult, outs[i].data_ptr(), stream)
            if ev is not None:
                ev[2 * i + 1].record(xs)

    for _ in range(args.warmup):
        step()
    irismpc.stream_synchronize(stream)
    torch.cuda.synchronize()

    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    kev = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    t_start.record(xs)
    for k in range(args.steps):
        step(kev[k])
    t_end.record(xs)
    t_end.synchronize()
    torch.cuda.synchronize()
    total_ms = t_start.elapsed_time(t_end)
    ms_step = total_ms / args.steps
    code_ms = sum(e[0].elapsed_time(e[1]) for e in kev) / args.steps
    mask_ms = sum(e[2].elapsed_time(e[3]) for e in kev) / args.steps

    pairs = Qr * N
    ops_code, ops_mask = 6 * 12800 * pairs, 6 * 6400 * pairs
    # the code and mask launches are the same kernel; per launch on average:
    # algorithmic int8 ops / average launch duration (HIP events on its stream)
    achieved = (ops_code + ops_mask) / ((code_ms + mask_ms) / 1e3) / 1e12
    resolved = variant if variant != -1 else (102 if Qr <= ROT_SMALL_ROWS else 100)
    kernel = KERNEL_OF_VARIANT.get(resolved)
    traffic = measured_traffic(kernel) if kernel and (N, Qr) == (1 << 20, 1984) else None
    if 3 * Qr >= RIDGE_OPS_PER_BYTE:
        roofline = {"bound": "mfma", "achieved": achieved, "peak": INT8_PEAK_TOPS, "unit": "TFLOP/s",
                    "frac": achieved / INT8_PEAK_TOPS, "traffic": traffic}
    else:
        # few query rows: the launch streams the DB limbs once (and writes Qr u16
        # shares per DB row); algorithmic bytes over the average launch time
        hbm_bytes = N * (12800 + 6400) * 2 + 2 * Qr * N * 2
        gbs = hbm_bytes / 2 / ((code_ms + mask_ms) / 2 / 1e3) / 1e9
        roofline = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": gbs / HBM_PEAK_GBS, "traffic": traffic}
    line = {
        "metric": "per-party share GEMM comparisons/sec on 1 MI355X (one party's code + mask dot shares)",
        "value": pairs / (ms_step / 1e3),
        "unit": "comparisons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8 MFMA (i32 acc) -> u16 shares",
        "data": "synthetic: uniformly random u16 DB and query shares (the distribution of real "
                "Shamir shares), ChaCha20-generated on device",
        "config": {
            "workload": f"per-party share GEMM, plain form: DB {N:,} x (12,800 code + 6,400 "
                        f"trimmed-mask) u16 shares in HBM, {B} queries x 31 rotations = {Qr:,} rows, "
                        "fused correction + mask x2 + ChaCha12 zero-share epilogue, "
                        "reference chunking 32,768 rows",
            "db_rows": N,
            "query_rows": Qr,
            "gemm_variant": resolved,
        },
        "roofline": roofline,
        "kernels": {
            "dominant": kernel,
            "share_gemm_code_ms": code_ms,
            "share_gemm_mask_ms": mask_ms,
            "int8_ops_per_pair": 115200,
            "int8_ops_per_launch_avg": (ops_code + ops_mask) / 2,
        },
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(with_ratio(line)))


if __name__ == "__main__":
    main()
