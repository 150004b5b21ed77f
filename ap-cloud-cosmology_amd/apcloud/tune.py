"""Launch-shape and cycle-shape autotuner for a context's kernels (the
APC_KNOB_* knobs of include/apcloud.h).

The tuner times candidate values of every knob on steps of the caller's own
workload and keeps the fastest: coordinate descent in knob order, one pass,
every trial a step from the same saved state with the library's kernel timers
on (HIP events on the context's stream).  A knob's cost is the device time of
the kernels it shapes, each timer's average launch time weighted by its
launch count in the trial of the knob's current value, so that a candidate
that changes the iteration count by a reduction-order rounding is not
rewarded or penalised for it.  A candidate replaces the current value only
when it is at least `margin` faster.  Knobs whose kernels the workload never
launches (for instance the packed boxes' tile grid on a single box) are
skipped.  These knobs change launch shapes only; results differ at most by
the summation order of reductions.

The multigrid cycle knobs (levels without pre-smoothing, W-cycle depth) are
tuned first: they change the preconditioner and so the Krylov step count,
and their cost is the wall time of the whole step (the best of `reps` runs,
timers off).  Results then differ within the solves' tolerance.
"""
import time

# cycle-shape knobs, costed by the step's wall time (tuned first, in order)
WALL = {
    "amg_nopre": [0, 1, 2],
    "amg_wlevels": [2, 3],
}

# candidate values per knob (the built-in default is always among them)
CANDIDATES = {
    "search_chunk": [64, 256, 1024, 4096],
    "mdot_blocks": [512, 1024, 2048],
    "maxpy_blocks": [512, 1024, 2048],
    "seg_blocks": [1024, 2048, 4096],
    "amg_g8_rows": [4096, 16384, 65536, 262144],
    "amg_g4_rows": [262144, 1 << 20, 1 << 30],
    "amg_rg8_rows": [65536, 1 << 20, 1 << 30],
}

# kernel timers whose device time each knob shapes
TIMERS = {
    "search_chunk": ["neighbor_search"],
    "mdot_blocks": ["gmres_mdot", "gmres_mdot32"],
    "maxpy_blocks": ["gmres_maxpy", "gmres_maxpy32"],
    "seg_blocks": ["sgmres_resid", "sgmres_resid64", "sgmres_mdot", "sgmres_mdot32",
                   "sgmres_maxpy", "sgmres_maxpy32", "sgmres_update", "sgmres_update32",
                   "sgmres_update_z"],
    "amg_g8_rows": ["amg_cycle", "amg_cycle32"],
    "amg_g4_rows": ["amg_cycle", "amg_cycle32"],
    "amg_rg8_rows": ["amg_cycle", "amg_cycle32"],
}


def autotune(ctx, step, knobs=None, slot=63, margin=0.02, reps=2, clock=time.perf_counter):
    """Tune ctx's knobs on `step()`, a callable that runs one step of the
    workload on ctx.  The state at the call is saved in checkpoint `slot`,
    every trial starts from it, and it is restored at the end.  Returns
    {"choices": {knob: value}, "defaults": {knob: value}, "trials": {knob:
    {value: ms}}, "seconds": wall time of the tuning}."""
    t0 = time.perf_counter()
    ctx.checkpoint(slot)
    defaults = ctx.knobs()

    def measure():
        ctx.checkpoint(slot, restore=True)
        ctx.set_timing(True)
        ctx.reset_kernel_times()
        step()
        ctx.sync()
        kt = ctx.kernel_times()
        ctx.set_timing(False)
        return kt

    def measure_wall():
        best = float("inf")
        for _ in range(reps):
            ctx.checkpoint(slot, restore=True)
            ctx.sync()
            t = clock()
            step()
            ctx.sync()
            best = min(best, clock() - t)
        return 1e3 * best

    trials = {}
    for name in (knobs or list(WALL)):
        if name not in WALL or name not in defaults:
            continue
        cur = ctx.knobs()[name]
        res = {cur: measure_wall()}
        for v in WALL[name]:
            if v != cur:
                ctx.set_knob(name, v)
                res[v] = measure_wall()
        best = min(res, key=res.get)
        if res[best] > (1.0 - margin) * res[cur]:
            best = cur
        ctx.set_knob(name, best)
        trials[name] = {int(v): round(ms, 3) for v, ms in sorted(res.items())}
    for name in knobs or list(CANDIDATES):
        if name not in CANDIDATES or name not in defaults:
            continue
        cur = ctx.knobs()[name]
        kt0 = measure()
        weights = {t: kt0[t][1] for t in TIMERS[name] if t in kt0 and kt0[t][1] > 0}
        if not weights:
            continue

        def cost(kt):
            return sum(kt[t][0] * w for t, w in weights.items() if t in kt)

        res = {cur: cost(kt0)}
        for v in CANDIDATES[name]:
            if v == cur:
                continue
            ctx.set_knob(name, v)
            res[v] = cost(measure())
        best = min(res, key=res.get)
        if res[best] > (1.0 - margin) * res[cur]:
            best = cur
        ctx.set_knob(name, best)
        trials[name] = {int(v): round(ms, 4) for v, ms in sorted(res.items())}
    ctx.checkpoint(slot, restore=True)
    ctx.sync()
    return {"choices": ctx.knobs(), "defaults": defaults, "trials": trials,
            "seconds": round(time.perf_counter() - t0, 2)}
