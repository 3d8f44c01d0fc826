#!/usr/bin/env python3
"""Headline benchmark: SpGEMM GFLOPS of C = A^2 for the 3D 27-point stencil
on a 128^3 periodic grid (n = 2,097,152, 56,623,104 nonzeros, FP64 values
U[-1,1) from seed 1) on one MI355X.

A step is one full device pipeline (step 1 tile symbolic, step 2 bitmask
symbolic, step 3 numeric, including the pooled allocations of C) with A
already resident in HBM; GFLOPS = 2 * nnzCub / t (src/tilespgemm-cuda.h:
3185).  Prints one JSON line (see the keys below), with the batched and
served workloads under "batched" and "served".  --gpus N > 1 exits (one GPU
only).
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rw-spgemm_amd"))

METRIC = "SpGEMM GFLOPS (A², 27-pt 128³); batched/served solves/s + p99 latency"
HBM_PEAK_GBS = 8000.0


def step3_algorithmic_bytes(ta, tc_numtile, tc_nnz, tilen):
    """Compulsory bytes of one numeric-step launch (the dominant kernel):
    read A once and B once in tile format (B aliases A for A^2 but is
    counted again), read the CSC tile index, read C's step-2 outputs
    (row/col tile index, tile offsets, u8 row offsets, masks), write C's
    values and in-tile columns."""
    T, nnz, tilem = ta.numtile, ta.nnz, ta.tilem
    a_bytes = 4 * (tilem + 1) + 4 * T + 4 * (T + 1) + 8 * nnz + 16 * T + 32 * T
    b_bytes = a_bytes + 4 * (tilen + 1) + 4 * T + 4 * T
    c_meta = 4 * tc_numtile * 2 + 4 * (tc_numtile + 1) + 16 * tc_numtile + 32 * tc_numtile
    c_write = 9 * tc_nnz
    return a_bytes + b_bytes + c_meta + c_write


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(a, nnzcub_per_row, target_s=15.0):
    """The reference CPU path (src/spgemm_serialref_spa_new.h, count + fill
    passes, timed as src/external/cusparse/main.cu:191-222) restated in
    oracle/oracle.c, on a contiguous block of rows from the middle of the
    matrix sized to roughly `target_s` seconds."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    mid = a.m // 2
    rows = 4096
    t, _ = oracle.spa_structure_rows_timed(a, a, mid, mid + rows)
    while t < target_s / 8 and rows < a.m // 2:
        rows = min(a.m // 2, rows * 4)
        t, _ = oracle.spa_structure_rows_timed(a, a, mid - rows // 2, mid + rows // 2)
    r0, r1 = mid - rows // 2, mid + rows // 2
    if rows == 4096:
        r0, r1 = mid, mid + rows
    # sample's intermediate products
    cub = int(nnzcub_per_row[r0:r1].sum())
    times = [t]
    for _ in range(2):
        tt, _ = oracle.spa_structure_rows_timed(a, a, r0, r1)
        times.append(tt)
    tmin = min(times)
    # the value-computing SPA (src/external/cusparse/spgemm_serialref_spa.h
    # order, touched-entry reset) on the same rows
    vt = []
    for _ in range(2):
        t0 = time.perf_counter()
        oracle.spgemm_rows(a, a, r0, r1, values=True)
        vt.append(time.perf_counter() - t0)
    return {"value": round(2.0 * cub / tmin / 1e9, 4), "unit": "GFLOPS", "cores": oracle.num_threads(),
            "cpu_model": cpu_model(), "kind": "port",
            "sample": f"rows [{r0},{r1}) of the same A^2 ({r1 - r0} of {a.m} rows, {cub} products), "
                      f"count+fill passes, min of {len(times)} runs ({tmin:.2f} s)",
            "values_spa": {"value": round(2.0 * cub / min(vt) / 1e9, 4), "unit": "GFLOPS",
                           "sample": f"same rows, structure + FP64 values, min of {len(vt)} runs ({min(vt):.2f} s)"}}


def dirichlet_leg(ts, ctx, N, tune, warmup, steps, rowindex=False):
    """The same pipeline on the 27-point Dirichlet stencil N^3 (rows at the
    boundary keep only in-grid neighbours); the headline is the periodic
    stencil."""
    a = ts.gen_stencil27(N, N, N, periodic=False, valmode=ts.VAL_UNIFORM_PM1, seed=1)
    da = ctx.upload(a)
    ta = ctx.csr2tile(da)
    ctx.build_csc(ta)
    if rowindex:
        ctx.build_rowindex(ta)
    cub = ctx.nnzcub(da, da)
    for _ in range(warmup):
        tc, _ = ctx.spgemm(ta, ta, tune)
        ctx.free(tc)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        tc, _ = ctx.spgemm(ta, ta, tune, timed=False)
        ctx.free(tc)
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    ev = []
    for _ in range(steps):
        tc, st = ctx.spgemm(ta, ta, tune)
        ev.append(st.t_total_ms)
        ctx.free(tc)
    ctx.sync()
    out = {"workload": f"27-point Dirichlet stencil {N}^3, C = A^2", "n": a.m, "nnz": a.nnz, "nnzCub": cub,
           "nnzC": st.nnzC, "tilesC": st.numtileC, "gflops": round(2.0 * cub / (ms * 1e6), 2),
           "ms_per_step": round(ms, 4), "gflops_event_min": round(2.0 * cub / (min(ev) * 1e6), 2)}
    ctx.free(ta)
    ctx.free(da)
    ctx.trim()
    return out


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def batch_leg(ts, ctx, count=512, seed=2, nmin=1000, nmax=65536, steps=5, autotune=True, sequential=True):
    """Batched workload: `count` SpGEMMs (A^2 / AA^T, 50/50; n log-uniform in
    [nmin, nmax]; Erdos-Renyi mean nnz/row U[4,32] or banded half-bandwidth
    U{2..40}; values U[-1,1)).  Batched = one graph launch per step over
    the packed problems (tsg_batch); sequential = one plain tsg_spgemm call
    per problem on the same resident tile operands."""
    specs = ts.gen_batch_spec(count, seed=seed, nmin=nmin, nmax=nmax)
    t0 = time.perf_counter()
    mats = [ts.gen_batch_problem(sp) for sp in specs]
    ops = [sp.op for sp in specs]
    gen_s = time.perf_counter() - t0
    progress(f"batch: generated {count} problems in {gen_s:.1f} s")
    dps = [ctx.upload(m) for m in mats]
    nnz_total = sum(m.nnz for m in mats)
    del mats
    # plain operands for the sequential calls, prepared like the batch's
    # (CSC and row index on B), so both run the same kernels
    seq = []
    for d, op in zip(dps, ops):
        ta = ctx.csr2tile(d)
        ctx.build_csc(ta)
        if op == ts.OP_A2:
            ctx.build_rowindex(ta)
            seq.append((ta, ta, None))
        else:
            ctx.build_rowindex(ta)
            dt = ctx.transpose(d)
            tb = ctx.csr2tile(dt)
            ctx.build_csc(tb)
            ctx.build_rowindex(tb)
            ctx.free(dt)
            seq.append((ta, tb, tb))
    seq_s = seq_std_s = float("nan")
    if sequential:
        # the same calls on operands with the standard preparation only (tiles
        # + CSC, no element row index): what plain tsg_spgemm calls get
        std = []
        for d, op in zip(dps, ops):
            ta = ctx.csr2tile(d)
            ctx.build_csc(ta)
            if op == ts.OP_A2:
                std.append((ta, ta, None))
            else:
                dt = ctx.transpose(d)
                tb = ctx.csr2tile(dt)
                ctx.build_csc(tb)
                ctx.free(dt)
                std.append((ta, tb, tb))
        for ta, tb, _ in std:
            tc, _ = ctx.spgemm(ta, tb, timed=False)
            ctx.free(tc)
        ctx.sync()
        t0 = time.perf_counter()
        for ta, tb, _ in std:
            tc, _ = ctx.spgemm(ta, tb, timed=False)
            ctx.free(tc)
        ctx.sync()
        seq_std_s = time.perf_counter() - t0
        for ta, tb, own in std:
            ctx.free(ta)
            if own is not None:
                ctx.free(own)
        progress(f"batch: sequential, standard operands {seq_std_s * 1e3:.1f} ms")
        for ta, tb, _ in seq:  # warm the allocator cache
            tc, _ = ctx.spgemm(ta, tb, timed=False)
            ctx.free(tc)
        ctx.sync()
        t0 = time.perf_counter()
        for ta, tb, _ in seq:
            tc, _ = ctx.spgemm(ta, tb, timed=False)
            ctx.free(tc)
        ctx.sync()
        seq_s = time.perf_counter() - t0
    progress(f"batch: sequential {seq_s * 1e3:.1f} ms")
    for ta, tb, own in seq:
        ctx.free(ta)
        if own is not None:
            ctx.free(own)
    t0 = time.perf_counter()
    b = ts.Batch(ctx, dps, ops)
    create_s = time.perf_counter() - t0
    progress(f"batch: created in {create_s:.1f} s")
    tuned = None
    if autotune:
        t0 = time.perf_counter()
        tt, best_ms = b.autotune(reps=2)
        tuned = {"force_tile_parallel": tt.force_tile_parallel, "dense_acc_min_nnz": tt.dense_acc_min_nnz,
                 "best_ms": round(best_ms, 3), "autotune_s": round(time.perf_counter() - t0, 2)}
        progress(f"batch: autotuned {tuned}")
    b.run()
    ctx.sync()
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        b.run()
        ctx.sync()
        times.append(time.perf_counter() - t0)
    info = b.info()
    b.close()
    for d in dps:
        ctx.free(d)
    step_s = statistics.mean(times)
    return {"problems": count, "seed": seed, "nnz_total": nnz_total, "nnzC": info.nnzC, "tilesC": info.numtileC,
            "groups": info.groups, "graph_nodes": info.graph_nodes,
            "solves_per_s": round(count / step_s, 1), "ms_per_step": round(step_s * 1e3, 3),
            "ms_per_step_min": round(min(times) * 1e3, 3),
            "sequential_solves_per_s": round(count / seq_s, 1), "sequential_ms": round(seq_s * 1e3, 3),
            "speedup_vs_sequential": round(seq_s / step_s, 2),
            "sequential_standard_ms": round(seq_std_s * 1e3, 3),
            "speedup_vs_sequential_standard": round(seq_std_s / step_s, 2), "create_s": round(create_s, 2),
            "generate_s": round(gen_s, 2), "autotune": tuned}


def serve_leg(ts, n_req=120, lam=None, seed=3, n_er=24, n_rmat=2, rmat_scale=18, budget_gib=64):
    """Serving workload: Poisson arrivals of A^2 / AA^T requests (50/50);
    80% on Erdos-Renyi matrices (n log-uniform in [2k, 256k], mean nnz/row
    U[4,32]), 20% on R-MAT (scale 18, edge factor 16), all registered once
    (resident in HBM) under a 64 GiB budget.  Reports latency percentiles
    (arrival -> result complete in HBM) at a fixed rate, the burst capacity
    (all requests at once) and, on the same request list, one call after
    another on the same resident operands (prepared like the server's, with
    the element row index) and plain calls on operands with the drop-in
    preparation (tiles + CSC)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    ers = []
    for i in range(n_er):
        n = int(round(np.exp(rng.uniform(np.log(2000), np.log(262144)))))
        ers.append(ts.gen_er(n, float(rng.uniform(4, 32)), seed=1000 + i))
    rms = [ts.gen_rmat(rmat_scale, 16, seed=2000 + i) for i in range(n_rmat)]
    progress(f"serve: generated pool in {time.perf_counter() - t0:.1f} s")
    mats = ers + rms
    kinds = rng.uniform(size=n_req) < 0.8
    pick = [int(rng.integers(0, n_er)) if k else n_er + int(rng.integers(0, n_rmat)) for k in kinds]
    ops = [int(x) for x in rng.integers(0, 2, n_req)]
    # sequential plain calls on resident operands (same kernels, same row index)
    ctx = ts.Context(0)
    dev = []
    for a in mats:
        da = ctx.upload(a)
        ta = ctx.csr2tile(da)
        ctx.build_csc(ta)
        ctx.build_rowindex(ta)
        dat = ctx.transpose(da)
        tat = ctx.csr2tile(dat)
        ctx.build_csc(tat)
        ctx.build_rowindex(tat)
        dev.append((da, ta, dat, tat))
    plans = {}
    for i, op in zip(pick, ops):
        if (i, op) not in plans:
            _, ta, _, tat = dev[i]
            plans[(i, op)] = ctx.chunk_plan(ta, ta if op == ts.OP_A2 else tat, 1500000000)[0]
    def one(i, op):
        _, ta, _, tat = dev[i]
        b = plans[(i, op)]
        for r0, r1 in zip(b, b[1:]):
            c, _ = ctx.spgemm_tile_rows(ta, ta if op == ts.OP_A2 else tat, r0, r1, timed=False)
            ctx.free(c)
        ctx.sync()
    for key in plans:  # warm
        one(*key)
    t0 = time.perf_counter()
    for i, op in zip(pick, ops):
        one(i, op)
    seq_s = time.perf_counter() - t0
    progress(f"serve: sequential {seq_s:.2f} s for {n_req} requests")
    for t in dev:
        for o in t:
            ctx.free(o)
    ctx.trim()
    # sequential plain calls: operands with the drop-in preparation only
    # (tiles + CSC, no element row index), one timed call per distinct
    # (matrix, op) of the request list, summed over the requests
    plain_ms = {}
    for i in sorted(set(pick)):
        d = ctx.upload(mats[i])
        ta = ctx.csr2tile(d)
        ctx.build_csc(ta)
        tb = None
        for op in sorted(set(o for j, o in zip(pick, ops) if j == i)):
            if op == ts.OP_A2:
                b = ta
            else:
                if tb is None:
                    dt = ctx.transpose(d)
                    tb = ctx.csr2tile(dt)
                    ctx.build_csc(tb)
                    ctx.free(dt)
                b = tb
            bounds = ctx.chunk_plan(ta, b, 1500000000)[0]
            ctx.sync()
            t0 = time.perf_counter()
            for r0, r1 in zip(bounds, bounds[1:]):
                c, _ = ctx.spgemm_tile_rows(ta, b, r0, r1, timed=False)
                ctx.free(c)
            ctx.sync()
            plain_ms[(i, op)] = (time.perf_counter() - t0) * 1e3
        for o in (tb, ta, d):
            if o is not None:
                ctx.free(o)
        ctx.trim()
    plain_s = sum(plain_ms[(i, op)] for i, op in zip(pick, ops)) / 1e3
    progress(f"serve: sequential plain calls {plain_s:.1f} s for {n_req} requests "
             f"({len(plain_ms)} distinct products timed)")
    ctx.close()
    srv = ts.Server(0, ts.server_config(budget_bytes=budget_gib << 30))
    t0 = time.perf_counter()
    handles = [srv.register(a) for a in mats]
    reg_s = time.perf_counter() - t0
    progress(f"serve: registered {len(mats)} matrices in {reg_s:.1f} s")
    hs = [handles[i] for i in pick]

    def replay_progress(label):
        """Prints the replay's own submissions / completions, at most every 10 s."""
        last = [time.perf_counter()]

        def cb(sub, done):
            now = time.perf_counter()
            if now - last[0] >= 10.0 or done == n_req:
                progress(f"serve: {label}: {sub}/{n_req} submitted, {done} completed")
                last[0] = now
        return cb

    # burst: every request at t = 0 (capacity)
    t0 = time.perf_counter()
    _, binfo = srv.replay(hs, ops, [0.0] * n_req, keep=False, progress=replay_progress("burst"))
    burst_s = time.perf_counter() - t0
    progress(f"serve: burst {burst_s:.2f} s")
    cap = n_req / burst_s
    rate = lam if lam else 0.7 * cap
    arr = np.cumsum(rng.exponential(1000.0 / rate, n_req)).tolist()
    st0 = srv.stats()
    t0 = time.perf_counter()
    _, infos = srv.replay(hs, ops, arr, keep=False, progress=replay_progress(f"Poisson {rate:.2f} req/s"))
    open_s = time.perf_counter() - t0
    st1 = srv.stats()
    srv.close()
    lat = np.array([i.t_done_ms - i.t_arrival_ms for i in infos])
    blat = np.array([i.t_done_ms - i.t_arrival_ms for i in binfo])
    rm = np.array([not k for k in kinds])
    return {"requests": n_req, "seed": seed, "pool": {"er": n_er, "rmat": n_rmat, "rmat_scale": rmat_scale},
            "budget_gib": budget_gib, "register_s": round(reg_s, 2),
            "rate_req_per_s": round(rate, 2), "solves_per_s": round(n_req / open_s, 2),
            "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
            "p99_ms_er": round(float(np.percentile(lat[~rm], 99)), 2) if (~rm).any() else None,
            "p99_ms_rmat": round(float(np.percentile(lat[rm], 99)), 2) if rm.any() else None,
            "burst_solves_per_s": round(cap, 2), "burst_p99_ms": round(float(np.percentile(blat, 99)), 2),
            "sequential_solves_per_s": round(n_req / seq_s, 2),
            "speedup_vs_sequential": round(seq_s / burst_s, 2),
            "sequential_plain_solves_per_s": round(n_req / plain_s, 3),
            "speedup_vs_sequential_plain": round(plain_s / burst_s, 2),
            "sequential_plain_method": "one timed plain call (tiles + CSC, chunked) per distinct (matrix, op), "
                                       "summed over the request list",
            "fused_launches": st1.fused_launches - st0.fused_launches,
            "single_launches": st1.single_launches - st0.single_launches,
            "peak_reserved_gib": round(st1.peak_reserved_bytes / 2 ** 30, 2),
            "resident_gib": round(st1.resident_bytes / 2 ** 30, 2)}


def rmat_cpu_baseline(a, b, target_s=10.0):
    """The oracle's faithful SPA (count + fill passes, per-row (nB/32+1)-word
    bitmap, 64-bit row offsets since C overflows int32) on a contiguous
    block of rows of the same A*A^T, grown until it takes about target_s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    rp = a.rowptr.astype(np.int64)
    rowlen_b = np.diff(b.rowptr.astype(np.int64))
    cs = np.concatenate([[0], np.cumsum(rowlen_b[a.colidx])])
    cub_row = cs[rp[1:]] - cs[rp[:-1]]
    mid, rows = a.m // 2, 1024
    while True:
        r0, r1 = mid - rows // 2, mid + rows // 2
        t, _ = oracle.spa_structure_rows_timed(a, b, r0, r1)
        if t >= target_s / 4 or rows >= a.m // 2:
            break
        rows = min(a.m // 2, rows * 4)
    times = [t]
    for _ in range(2):
        times.append(oracle.spa_structure_rows_timed(a, b, r0, r1)[0])
    cub = int(cub_row[r0:r1].sum())
    tmin = min(times)
    return {"value": round(2.0 * cub / tmin / 1e9, 4), "unit": "GFLOPS", "cores": oracle.num_threads(),
            "cpu_model": cpu_model(), "kind": "port",
            "sample": f"rows [{r0},{r1}) of A*A^T in the generator's order ({r1 - r0} of {a.m} rows, {cub} "
                      f"products), count+fill passes with 64-bit row offsets, min of {len(times)} runs "
                      f"({tmin:.2f} s)"}


def rmat_leg(ts, scale=20, chunk_products=1500000000, cpu=True, orders=("random", "rcm", "degree")):
    """R-MAT scale-20 C = A*A^T (edge factor 16, (0.57, 0.19, 0.19, 0.05),
    symmetrized, duplicates merged, values U(0,1], seed 1) under the random,
    RCM and degree orderings, each computed on the device.  C (~2.5e10
    nonzeros) does not fit next to its operands, so it runs in tile-row
    chunks of at most chunk_products products; each chunk is computed and
    dropped.  GFLOPS = 2 * nnzCub / (sum of chunk pipeline times)."""
    import numpy as np
    t0 = time.perf_counter()
    a = ts.gen_rmat(scale, 16, seed=1)
    progress(f"rmat{scale}: generated {a.nnz} nnz in {time.perf_counter() - t0:.1f} s")
    ctx = ts.Context(0)
    d = ctx.upload(a)
    out = {"scale": scale, "n": a.m, "nnz": a.nnz, "orders": {}}
    nnzc = set()
    for algo in orders:
        if algo not in ("random", "rcm", "degree"):
            raise ValueError(f"unknown R-MAT ordering {algo!r}")
        ctx.sync()
        t0 = time.perf_counter()
        if algo == "rcm":
            old, _ = ctx.order_rcm(d)
        elif algo == "degree":
            old = ctx.order_degree(d)
        else:
            old = ctx.order_random(a.m, 1)
        t_order = time.perf_counter() - t0
        new = np.empty_like(old)
        new[old] = np.arange(a.m, dtype=np.int32)
        t0 = time.perf_counter()
        dp = ctx.permute(d, old, new)
        ta = ctx.csr2tile(dp)
        ctx.build_csc(ta)
        ctx.build_rowindex(ta)
        dpt = ctx.transpose(dp)
        tb = ctx.csr2tile(dpt)
        ctx.build_csc(tb)
        ctx.build_rowindex(tb)
        ctx.sync()
        t_prep = time.perf_counter() - t0
        bounds, prods = ctx.chunk_plan(ta, tb, chunk_products)
        tot_ms, nnz, tiles = 0.0, 0, 0
        for r0, r1 in zip(bounds, bounds[1:]):
            c, st = ctx.spgemm_tile_rows(ta, tb, r0, r1)
            tot_ms += st.t_total_ms
            nnz += st.nnzC
            tiles += st.numtileC
            ctx.free(c)
        nnzc.add(nnz)
        out["orders"][algo] = {"order_s": round(t_order, 3), "prep_s": round(t_prep, 3), "A_tiles": ta.numtile,
                               "chunks": len(bounds) - 1, "nnzCub": prods, "nnzC": nnz, "tilesC": tiles,
                               "spgemm_ms": round(tot_ms, 2), "gflops": round(2.0 * prods / (tot_ms * 1e6), 2)}
        progress(f"rmat{scale} {algo}: {out['orders'][algo]}")
        for o in (tb, dpt, ta, dp):
            ctx.free(o)
        ctx.trim()
    out["nnzC_invariant_across_orders"] = len(nnzc) == 1
    ctx.free(d)
    ctx.close()
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        tp, tc, tv = oracle.transpose(a)
        out["cpu_baseline"] = rmat_cpu_baseline(a, ts.Csr(a.n, a.m, tp, tc, tv))
        progress(f"rmat{scale}: cpu baseline {out['cpu_baseline']['value']} GFLOPS")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-only", action="store_true")
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--batch-no-sequential", action="store_true")
    ap.add_argument("--element-mode", type=int, default=None,
                    help="build B's row index and use this element_mode (0 auto, 1 never, 2 every tile-row)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--rowindex", action="store_true",
                    help="also keep B's element row index resident (as the drop-in tilespgemm() does)")
    ap.add_argument("--no-27d", action="store_true")
    ap.add_argument("--rmat-only", action="store_true")
    ap.add_argument("--no-rmat", action="store_true")
    ap.add_argument("--rmat-scale", type=int, default=20)
    ap.add_argument("--rmat-orders", default="random,rcm,degree",
                    help="orderings of the --rmat-only leg, comma-separated")
    ap.add_argument("--serve-only", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=120)
    ap.add_argument("--serve-rmat-scale", type=int, default=18)
    ap.add_argument("--dense-threshold", type=int, default=None)
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="first run the plain copy kernel over this many bytes (byte-counter reference)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only; --gpus {args.gpus} is not supported "
              "(a multi-GPU SpGEMM would shard by C tile-rows with B replicated)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import tilespgemm as ts

    if args.rmat_only:
        print(json.dumps({"rmat": rmat_leg(ts, scale=args.rmat_scale, cpu=not args.no_cpu,
                                           orders=tuple(args.rmat_orders.split(",")))}), flush=True)
        return
    if args.serve_only:
        print(json.dumps({"served": serve_leg(ts, n_req=args.serve_requests, rmat_scale=args.serve_rmat_scale)}),
              flush=True)
        return
    if args.batch_only:
        ctx = ts.Context(0)
        print(json.dumps({"batched": batch_leg(ts, ctx, count=args.batch_count, autotune=not args.no_autotune,
                                               sequential=not args.batch_no_sequential)}), flush=True)
        ctx.close()
        return
    N = args.grid
    a = ts.gen_stencil27(N, N, N, periodic=True, valmode=ts.VAL_UNIFORM_PM1, seed=1)
    ctx = ts.Context(0)
    if args.copy_calibration:
        ctx.copy_kernel_run(args.copy_calibration)
    da = ctx.upload(a)
    # operand preparation outside the timed region: tiles and the CSC tile
    # index (every tile-row of this product fits the row-wise kernels, which
    # do not read B's element row index)
    ta = ctx.csr2tile(da)
    ctx.build_csc(ta)
    if args.rowindex:
        ctx.build_rowindex(ta)
    nnzcub = ctx.nnzcub(da, da)
    tune = ts.default_tune()
    if args.element_mode is not None:
        ctx.build_rowindex(ta)
        tune.element_mode = args.element_mode
    if args.dense_threshold:
        tune.dense_acc_min_nnz = args.dense_threshold
    for _ in range(args.warmup):
        tc, st = ctx.spgemm(ta, ta, tune)
        ctx.free(tc)
    ctx.sync()
    # throughput: plain pipeline runs (no events or statistics)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tc, _ = ctx.spgemm(ta, ta, tune, timed=False)
        ctx.free(tc)
    ctx.sync()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    # breakdown and roofline: the same runs with per-step events
    steps = []
    for _ in range(args.steps):
        tc, st = ctx.spgemm(ta, ta, tune)
        steps.append(st.as_dict())
        ctx.free(tc)
    ctx.sync()
    ev_total = [s["t_total_ms"] for s in steps]
    ev_s3 = [s["t_step3_rows_ms"] for s in steps]
    last = steps[-1]
    gflops = 2.0 * nnzcub / (ms_per_step * 1e6)
    s3_bytes = step3_algorithmic_bytes(ta, last["numtileC"], last["nnzC"], ta.tilen)
    s3_ms = statistics.mean(ev_s3)
    achieved = s3_bytes / (s3_ms * 1e-3) / 1e9
    traffic = None
    tp = os.path.join(ROOT, "profiles", "step3_traffic_27p128.json")
    if os.path.exists(tp):
        with open(tp) as f:
            traffic = json.load(f).get("bytes_per_launch")
    out = {
        "metric": METRIC,
        "value": round(gflops, 2),
        "unit": "GFLOPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"27-point periodic stencil {N}^3, C = A^2", "n": a.m, "nnz": a.nnz,
                   "nnzCub": nnzcub, "nnzC": last["nnzC"], "tilesC": last["numtileC"],
                   "tilesC_empty": last["tiles_empty"], "values": "U[-1,1) seed 1",
                   "operand": "resident tiles + CSC tile index" + (" + element row index" if args.rowindex else "")},
        "roofline": {"bound": "hbm", "kernel": "k_step3_rows (numeric)", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "algorithmic_bytes": s3_bytes, "avg_ms": round(s3_ms, 4)},
        "step_ms": {"step1": round(statistics.mean(s["t_step1_ms"] for s in steps), 4),
                    "step2": round(statistics.mean(s["t_step2_ms"] for s in steps), 4),
                    "step3": round(statistics.mean(s["t_step3_ms"] for s in steps), 4),
                    "step3_rows_kernel": round(s3_ms, 4),
                    "event_total_min": round(min(ev_total), 4),
                    "event_total_median": round(statistics.median(ev_total), 4)},
        "gflops_event_min": round(2.0 * nnzcub / (min(ev_total) * 1e6), 2),
    }
    ctx.free(ta)
    ctx.free(da)
    if not args.no_27d:
        out["stencil27_dirichlet"] = dirichlet_leg(ts, ctx, N, tune, args.warmup, args.steps,
                                                   rowindex=args.rowindex)
    if not args.no_batch:
        ctx.trim()
        out["batched"] = batch_leg(ts, ctx, count=args.batch_count, autotune=not args.no_autotune)
    ctx.close()
    if not args.no_serve:
        out["served"] = serve_leg(ts, n_req=args.serve_requests, rmat_scale=args.serve_rmat_scale)
    if not args.no_rmat:
        out["rmat"] = rmat_leg(ts, scale=args.rmat_scale, cpu=not args.no_cpu)
    if not args.no_cpu:
        rp = a.rowptr.astype(np.int64)
        rowlen = np.diff(rp)
        cs = np.concatenate([[0], np.cumsum(rowlen[a.colidx])])
        cub_row = cs[rp[1:]] - cs[rp[:-1]]
        out["cpu_baseline"] = cpu_baseline(a, cub_row)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
