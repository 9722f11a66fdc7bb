"""Offline tuning sweep of the batched solver's launch constants, per system-size bucket.

    python -m gpupfem2_amd.tune [--out gpupfem2_amd/tuning_batched.json] [--reps 2]

The 256-mesh batched pressure-Poisson workload (cases.batched_poisson_case: meshes of 10k-500k triangles)
is solved with Jacobi PCG in mixed precision to 1e-10 under candidate settings, coordinate by coordinate:
for each size bucket (systems of at most 20,000 rows, at most 100,000 rows, larger) the SpMV tile limits
(rows per tile in {64, 128, 256}, non-zeros per tile in {1024, 2048}; every tile stays within one system)
with the other buckets at their current best, then the host poll interval (iterations launched between
two reads of the per-system states, {4, 8, 16}). Each setting is timed as the best of --reps solves after a
warm-up solve. The result (the chosen tiles, the poll interval and every measured time) is written as JSON;
bench.py's batched workload applies it when the file exists (pfem_solver_set_batch_tiles,
PFEM_SOLVER_OPT_POLL).
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

DEFAULT_OUT = os.path.join(HERE, "tuning_batched.json")
BUCKETS = (20_000, 100_000, 1 << 30)
TILE_ROWS = (64, 128, 256)
TILE_NNZ = (1024, 2048)
POLLS = (4, 8, 16)


def load(path=DEFAULT_OUT):
    """The tuned settings ({"buckets": [[max_rows, tile_rows, tile_nnz], ...], "poll": p}) or None."""
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return None


def apply(solver, tuning):
    """Set a batched solver's tiles and poll interval from a tuning record (no-op for None)."""
    from gpupfem2_amd import capi
    if not tuning:
        return
    solver.set_batch_tiles([tuple(b) for b in tuning["buckets"]])
    solver.set_option(capi.SOLVER_OPT_POLL, int(tuning["poll"]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--reps", type=int, default=2)
    args = ap.parse_args()
    from gpupfem2_amd import capi, cases
    case = cases.batched_poisson_case()
    ctx = capi.Context(0)
    mesh, a, b = cases.build_batched_poisson(ctx, case)
    rows = case["sys_rows"]
    n = int(rows[-1])
    sizes = rows[1:] - rows[:-1]
    print(f"[tune] {rows.size - 1} systems, {n} rows; per bucket: "
          f"{[int(((sizes <= hi) & (sizes > lo)).sum()) for lo, hi in zip((0,) + BUCKETS[:-1], BUCKETS)]}",
          file=sys.stderr, flush=True)

    def measure(buckets, poll):
        s = capi.Solver(a, capi.SOLVER_CG, 1e-10, 20000, 0, capi.TOL_RELATIVE, capi.Preconditioner(a, capi.PRECOND_JACOBI))
        s.set_precision(capi.PRECISION_MIXED, 1e-5, 30)
        s.set_batch_tiles(buckets)
        s.set_option(capi.SOLVER_OPT_POLL, poll)
        x = ctx.zeros(n)
        best = None
        for k in range(args.reps + 1):
            x.zero()
            ctx.sync()
            t = time.perf_counter()
            st, _ = s.solve_batched(a, rows, x, b)
            ctx.sync()
            dt = time.perf_counter() - t
            assert all(q.converged for q in st)
            if k > 0:
                best = dt if best is None else min(best, dt)
        del s
        return best

    cur = [[hi, 256, 2048] for hi in BUCKETS]
    poll = 8
    log = []
    base = measure(cur, poll)
    log.append({"buckets": [list(c) for c in cur], "poll": poll, "s": base})
    print(f"[tune] default: {base * 1e3:.1f} ms", file=sys.stderr, flush=True)
    best_t = base
    for k in range(len(BUCKETS)):
        for tr in TILE_ROWS:
            for tn in TILE_NNZ:
                if [tr, tn] == cur[k][1:]:
                    continue
                trial = [list(c) for c in cur]
                trial[k][1:] = [tr, tn]
                t = measure(trial, poll)
                log.append({"buckets": trial, "poll": poll, "s": t})
                print(f"[tune] bucket {k} rows {tr} nnz {tn}: {t * 1e3:.1f} ms", file=sys.stderr, flush=True)
                if t < best_t:
                    best_t, cur = t, trial
    for p in POLLS:
        if p == poll:
            continue
        t = measure(cur, p)
        log.append({"buckets": [list(c) for c in cur], "poll": p, "s": t})
        print(f"[tune] poll {p}: {t * 1e3:.1f} ms", file=sys.stderr, flush=True)
        if t < best_t:
            best_t, poll = t, p
    out = {"workload": "256 jittered-triangle Poisson systems (10k-500k triangles), Jacobi PCG, mixed, rel. 1e-10",
           "buckets": cur, "poll": poll, "default_s": base, "tuned_s": best_t, "trials": log}
    with open(args.out, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps({k: out[k] for k in ("buckets", "poll", "default_s", "tuned_s")}))


if __name__ == "__main__":
    main()
