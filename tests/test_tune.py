"""The launch-shape autotuner's search (apcloud.tune) on a stand-in context
whose kernel times are a known function of the knobs: the cheapest candidate
of every knob is kept, a gain below the margin keeps the current value, knobs
whose kernels never run are skipped, launch counts of the current value weight
the candidates, and every trial starts from the saved state."""
from apcloud import tune


class FakeCtx:
    def __init__(self, cost, counts):
        self.k = {"search_chunk": 256, "mdot_blocks": 1024, "maxpy_blocks": 1024,
                  "seg_blocks": 2048, "amg_g8_rows": 65536, "amg_g4_rows": 1 << 30,
                  "amg_rg8_rows": 1 << 30}
        self.cost, self.counts = cost, counts
        self.restores, self.saved, self.timing = 0, None, False

    def knobs(self):
        return dict(self.k)

    def set_knob(self, name, v):
        self.k[name] = v

    def checkpoint(self, slot, restore=False):
        if restore:
            assert self.saved == slot
            self.restores += 1
        else:
            self.saved = slot

    def set_timing(self, on=True):
        self.timing = on

    def reset_kernel_times(self):
        pass

    def sync(self):
        pass

    def kernel_times(self):
        assert self.timing
        return {t: (self.cost(t, self.k), self.counts(t, self.k), 0.0)
                for t in ("neighbor_search", "gmres_mdot32", "gmres_maxpy32", "amg_cycle32")}


def test_the_tuner_keeps_the_cheapest_candidate_of_every_knob():
    def cost(t, k):
        if t == "neighbor_search":
            return {64: 50.0, 256: 40.0, 1024: 30.0, 4096: 35.0}[k["search_chunk"]]
        if t == "gmres_mdot32":
            return {512: 1.0, 1024: 0.995, 2048: 1.2}[k["mdot_blocks"]]  # within the margin
        if t == "gmres_maxpy32":
            return {512: 0.5, 1024: 1.0, 2048: 1.0}[k["maxpy_blocks"]]
        return 10.0 - (k["amg_g8_rows"] == 16384) * 2.0
    # the iteration count moves with mdot_blocks; the weights of the current
    # value must be used, else 2048 would win on fewer launches
    def counts(t, k):
        return 30 if k["mdot_blocks"] != 2048 else 10
    ctx = FakeCtx(cost, counts)
    step_calls = []
    res = tune.autotune(ctx, lambda: step_calls.append(1))
    ch = res["choices"]
    assert ch["search_chunk"] == 1024
    assert ch["mdot_blocks"] == 1024   # 0.5% gain is below the 2% margin
    assert ch["maxpy_blocks"] == 512
    assert ch["amg_g8_rows"] == 16384
    assert ch["seg_blocks"] == 2048 and "seg_blocks" not in res["trials"]  # never launched
    assert res["defaults"]["search_chunk"] == 256
    assert res["trials"]["search_chunk"] == {64: 50.0 * 30, 256: 40.0 * 30, 1024: 30.0 * 30,
                                             4096: 35.0 * 30}
    # one restore per trial and one at the end
    assert ctx.restores == len(step_calls) + 1
    assert not ctx.timing


class WallCtx(FakeCtx):
    """The cycle knobs: each step advances a fake clock by a function of the
    knobs (the Krylov step count they cause)."""

    def __init__(self, step_ms):
        super().__init__(lambda t, k: 1.0, lambda t, k: 1)
        self.k.update({"amg_nopre": 1, "amg_wlevels": 3})
        self.step_ms, self.now = step_ms, 0.0

    def clock(self):
        return self.now

    def step(self):
        self.now += self.step_ms(self.k) * 1e-3


def test_the_cycle_knobs_are_tuned_first_by_the_step_time():
    def step_ms(k):
        return {0: 80.0, 1: 90.0, 2: 95.0}[k["amg_nopre"]] + {2: 1.0, 3: 0.0, 4: 5.0}[k["amg_wlevels"]]
    ctx = WallCtx(step_ms)
    res = tune.autotune(ctx, ctx.step, clock=ctx.clock)
    assert res["choices"]["amg_nopre"] == 0            # 11% faster: taken
    assert res["choices"]["amg_wlevels"] == 3          # 1.2% slower / 6% slower: kept
    assert res["trials"]["amg_nopre"] == {0: 80.0, 1: 90.0, 2: 95.0}
    assert list(res["trials"])[:2] == ["amg_nopre", "amg_wlevels"]
    assert not ctx.timing
