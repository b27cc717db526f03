"""Autotuner for the mixed form's tile shape, LDS halo and workgroup size,
and for the batched ensemble's chunk size and persistent grid.

For a workload key (dimension, grid, particles per cell, drift) it builds the
real configuration for every compiled candidate, runs a few warm-up steps,
times the push+deposit kernel with HIP events over a fixed number of steps
(re-binning on the sort cadence, as the run would), and keeps the fastest.
Results persist as JSON so a run can pick its shape without re-tuning:

    python -m aperture3_amd.autotune --dim 2 --n 2048 --ppc 32

The default file is aperture3_amd/tuning.json (override with APT_TUNING).
"""
from __future__ import annotations

import argparse
import json
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_DB = os.path.join(HERE, "tuning.json")

# (tile extent, halo, threads); must match the shapes compiled into
# kern_tile_mixed.hip (TILE2D_SHAPES and TILE3D_SHAPES)
CANDIDATES = {
    2: [((16, 16, 0), 6, 256), ((16, 16, 0), 6, 512), ((16, 16, 0), 5, 256),
        ((16, 16, 0), 4, 256), ((32, 16, 0), 6, 512), ((16, 16, 0), 7, 512),
        ((16, 16, 0), 8, 512), ((16, 16, 0), 9, 512), ((16, 16, 0), 10, 512),
        ((16, 8, 0), 8, 512), ((32, 16, 0), 7, 512), ((32, 16, 0), 8, 512),
        ((32, 32, 0), 6, 1024), ((8, 8, 0), 4, 256)],
    3: [((8, 8, 8), 3, 256), ((8, 8, 8), 4, 256), ((4, 4, 4), 3, 256),
        ((8, 8, 8), 4, 512), ((8, 8, 8), 4, 1024), ((8, 8, 8), 5, 512),
        ((8, 8, 8), 5, 1024)],
}


def db_path() -> str:
    return os.environ.get("APT_TUNING", DEFAULT_DB)


def key(dim: int, n: int, ppc: int, theta: float, drift: float, sort_interval: int) -> str:
    return f"d{dim}_n{n}_ppc{ppc}_th{theta:g}_dr{drift:g}_sort{sort_interval}"


def lookup(k: str):
    try:
        with open(db_path()) as f:
            return json.load(f).get(k)
    except (OSError, ValueError):
        return None


def store(k: str, entry: dict) -> None:
    path = db_path()
    try:
        with open(path) as f:
            db = json.load(f)
    except (OSError, ValueError):
        db = {}
    db[k] = entry
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(db, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def measure(apt, dim, n, ppc, theta, drift, tile, halo, threads, steps=12,
            sort_interval=20, warmup=2):
    """Push-kernel ms per launch and sort ms for one candidate; None if the
    shape does not apply to this mesh."""
    N = [n if d < dim else 1 for d in range(3)]
    p = apt.SimParams(N=N, guard=[2 if d < dim else 0 for d in range(3)],
                      periodic_boundary=[True] * 3, q_e=1.0,
                      max_ptc_number=(n ** dim) * ppc * 2 + 1024,
                      precision=apt.APT_PREC_MIXED, data_interval=10 ** 9,
                      tile=list(tile), tile_halo=halo, tile_threads=threads)
    env = apt.sim_environment(p)
    data = apt.sim_data(env)
    try:
        data.init_bg_B_field(2, lambda x, y, z: 1.0 + 0.0 * x)
        data.fill_multiplicity(1.0, ppc, (theta, theta), (drift, -drift), sort=False)
        off, _ = data.tile_offsets()
        if off is None:
            return None
        dt = 0.5 / n
        pu, fs = apt.ptc_updater(env), apt.field_solver(env)
        for s in range(warmup):
            pu.update_particles(data, dt, s)
            fs.update_fields(data, dt, s)
        data.sync()
        data.timing_enable(True)
        for s in range(warmup, warmup + steps):
            pu.update_particles(data, dt, s)
            fs.update_fields(data, dt, s)
            if s % sort_interval == 0:
                data.sort_particles()
        ms, cnt = data.timing_get(apt.KERNEL_PUSH_DEPOSIT)
        sms, scnt = data.timing_get(apt.KERNEL_SORT)
        return {"push_ms": ms / max(cnt, 1), "sort_ms": sms / scnt if scnt else None}
    finally:
        data.close()


def tune(dim, n, ppc, theta=0.05, drift=0.3, sort_interval=20, steps=12, verbose=True):
    import aperture3_amd as apt
    apt.load_library()
    results = []
    for tile, halo, threads in CANDIDATES[dim]:
        t0 = time.time()
        r = measure(apt, dim, n, ppc, theta, drift, tile, halo, threads, steps=steps,
                    sort_interval=sort_interval)
        if r is None:
            continue
        r.update({"tile": list(tile), "halo": halo, "threads": threads})
        results.append(r)
        if verbose:
            print(f"tile={tile} halo={halo} threads={threads}: push {r['push_ms']:.3f} ms "
                  f"({time.time() - t0:.1f} s)", flush=True)
    best = min(results, key=lambda r: r["push_ms"])
    entry = dict(best)
    entry["candidates"] = results
    store(key(dim, n, ppc, theta, drift, sort_interval), entry)
    return entry


# batched ensemble: particles per workgroup (chunk) x persistent
# workgroups per CU
BATCH_CANDIDATES = [(c, g) for c in (256, 512, 1024) for g in (2, 4, 8, 16)]


def batch_key(n_inst: int, seed: int, sort_interval: int) -> str:
    return f"batch_n{n_inst}_seed{seed}_sort{sort_interval}"


def measure_batch(apt, insts, chunk, grid_mult, steps=8, sort_interval=20, seed=1):
    """Average ms per graph-replayed ensemble step and per persistent push
    for one (chunk, grid_mult)."""
    b = apt.batch_ensemble(insts, precision=apt.APT_PREC_MIXED, seed=seed, chunk=chunk,
                           grid_mult=grid_mult, sort_interval=sort_interval)
    try:
        b.graph(True)
        b.step(2)  # step 0 (eager, E half step) and the graph capture
        b.timing(1)
        b.step(steps)
        step_ms, push_ms, n = b.timing(0)
        return {"step_ms": step_ms / n, "push_ms": push_ms / n}
    finally:
        b.close()


def tune_batch(n_inst=256, seed=1, sort_interval=20, steps=8, verbose=True):
    import aperture3_amd as apt
    apt.load_library()
    insts = apt.sample_batch(n_inst, seed)
    results = []
    for chunk, gm in BATCH_CANDIDATES:
        t0 = time.time()
        r = measure_batch(apt, insts, chunk, gm, steps=steps, sort_interval=sort_interval,
                          seed=seed)
        r.update({"chunk": chunk, "grid_mult": gm})
        results.append(r)
        if verbose:
            print(f"chunk={chunk} grid_mult={gm}: step {r['step_ms']:.3f} ms, push "
                  f"{r['push_ms']:.3f} ms ({time.time() - t0:.1f} s)", flush=True)
    best = min(results, key=lambda r: r["step_ms"])
    entry = dict(best)
    entry["candidates"] = results
    store(batch_key(n_inst, seed, sort_interval), entry)
    return entry


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--batch", type=int, default=0,
                    help="tune the batched ensemble of this many boxes instead")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--dim", type=int, default=2)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--ppc", type=int, default=32)
    ap.add_argument("--theta", type=float, default=0.05)
    ap.add_argument("--drift", type=float, default=0.3)
    ap.add_argument("--sort-interval", type=int, default=20)
    ap.add_argument("--steps", type=int, default=12)
    a = ap.parse_args()
    if a.batch:
        best = tune_batch(a.batch, a.seed, a.sort_interval, a.steps)
        print(json.dumps({k: v for k, v in best.items() if k != "candidates"}))
        return
    best = tune(a.dim, a.n, a.ppc, a.theta, a.drift, a.sort_interval, a.steps)
    print(json.dumps({k: v for k, v in best.items() if k != "candidates"}))


if __name__ == "__main__":
    main()
