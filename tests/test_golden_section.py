"""Golden-section search over the block count and the per-level convergence rule.

Both the library's host code (gsap_golden_step / gsap_converged, used by every
partition call) and the oracle's (orc_golden_step / orc_converged) are compared, step by
step, with a literal transcription of the reference:

  * prepare_for_partition_next, GSAP/partition.hpp:80-322, on OldData (graph.hpp:57-70)
    with the snapshot buffers' pointer exchanges and device copies modelled as tokens;
  * the convergence test, GSAP/partition.hpp:918-934.

The transcription keeps the reference's integer and float arithmetic where it decides
something: B_to_merge = (int)P.B * rate is a float32 product truncated to int
(partition.hpp:246, 875) and std::round rounds halves away from zero (partition.hpp:283,
302).  S is an FP64 value in this build (the reference's is a float); the tables use S
values that are exact in both.

The tables are chosen so that plausible mistakes fail them: each mutant of the
transcription below (a swapped branch, an off-by-one, a window that includes the current
sweep, double instead of float truncation, banker's rounding, ...) disagrees with the
implementations on at least one case (test_tables_catch_mutants).
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import gsap_amd
import oracle_lib as O

INF = float("inf")


# ------------------------------------------------------------ literal transcription
class Partition:
    """graph.hpp:48-54 (the csr sizes only size device copies; left out)."""

    def __init__(self, B=0, B_to_merge=0, S=0.0):
        self.B, self.B_to_merge, self.S = B, B_to_merge, S

    def copy(self):
        return Partition(self.B, self.B_to_merge, self.S)


class OldData:
    """graph.hpp:57-70: B = 0, S = inf in all three slots."""

    def __init__(self):
        self.large = Partition(0, 0, INF)
        self.med = Partition(0, 0, INF)
        self.small = Partition(0, 0, INF)


class Gpu:
    """The four label buffers of Gpu (graph.hpp) as one-element lists; a pointer exchange
    exchanges the lists, a cudaMemcpyAsync copies the element."""

    def __init__(self, token):
        self.partitions = [token]
        self.large_partitions = [None]
        self.med_partitions = [None]
        self.small_partitions = [None]


def c_round(x):
    """std::round: halves away from zero, evaluated exactly on the double x."""
    f = Fraction(x)
    r = math.floor(abs(f) + Fraction(1, 2))
    return r if f >= 0 else -r


def ref_rate_trunc(B, rate, mut=()):
    if "double_rate" in mut:
        return int(B * float(np.float32(rate)))
    return int(np.float32(B) * np.float32(rate))  # (int)P.B * num_block_reduction_rate


def ref_prepare_next(gpu, old, P, rate, mut=(), log=None):
    """partition.hpp:80-322.  Returns (optimal_B_found, the caller's P afterwards); `log`
    collects the branches taken (for the coverage check)."""
    log = [] if log is None else log
    optimal_B_found = False
    B = P.B
    le = (P.S < old.med.S) if "strict_le" in mut else (P.S <= old.med.S)
    log.append(("le" if le else "gt", old.med.B > B))
    if le:
        cond = (old.med.B < B) if "swap_med_branch" in mut else (old.med.B > B)
        if cond:
            old.large = old.med.copy()
            gpu.large_partitions, gpu.med_partitions = gpu.med_partitions, gpu.large_partitions
        else:
            old.small = old.med.copy()
            gpu.small_partitions, gpu.med_partitions = gpu.med_partitions, gpu.small_partitions
        old.med = P.copy()
        gpu.med_partitions[0] = gpu.partitions[0]
    else:
        cond = (old.med.B < P.B) if "swap_worse_branch" in mut else (old.med.B > P.B)
        if cond:
            old.small = P.copy()
            gpu.small_partitions[0] = gpu.partitions[0]
        else:
            old.large = P.copy()
            gpu.large_partitions[0] = gpu.partitions[0]

    if math.isinf(old.small.S):
        P = old.med.copy()
        gpu.partitions[0] = gpu.med_partitions[0]
        P.B_to_merge = ref_rate_trunc(P.B, rate, mut)
        if P.B_to_merge == 0:
            optimal_B_found = True
        log.append("unbracketed_stop" if optimal_B_found else "unbracketed")
    else:
        stop_gap = 3 if "stop_gap3" in mut else 2
        if old.large.B - old.small.B == stop_gap:
            optimal_B_found = True
            P = old.med.copy()
            gpu.partitions[0] = gpu.med_partitions[0]
            log.append("gap2_stop")
        else:
            rnd = (lambda x: int(round(x))) if "bankers_round" in mut else c_round
            ge = ((old.large.B - old.med.B) > (old.med.B - old.small.B)) if "strict_ge" in mut \
                else ((old.large.B - old.med.B) >= (old.med.B - old.small.B))
            if ge:
                next_B = old.med.B + rnd((old.large.B - old.med.B) * 0.618)
                P = old.large.copy()
                gpu.partitions[0] = gpu.large_partitions[0]
                P.B_to_merge = old.large.B - next_B
                log.append("golden_large")
            else:
                next_B = old.small.B + rnd((old.med.B - old.small.B) * 0.618)
                P = old.med.copy()
                gpu.partitions[0] = gpu.med_partitions[0]
                P.B_to_merge = old.med.B - next_B
                log.append("golden_med")
    return optimal_B_found, P


def ref_converged(dE, itr, S, old, W=3, t1=5e-4, t2=1e-4, mut=()):
    """partition.hpp:918-934: True where the reference breaks out of the sweep loop."""
    if itr >= (W - 1):
        isf = math.isfinite(old.large.S) and math.isfinite(old.med.S) and \
            math.isfinite(old.small.S)
        if "isf_inverted" in mut:
            isf = not isf
        mean = 0.0
        hi = itr + 1 if "window_includes_current" in mut else itr
        lo = itr - W + 1
        for i in range(lo, hi):
            mean += dE[i]
        mean /= float(W if "divide_by_W" in mut else W - 1)
        thr = t2 if isf else t1
        if "le" in mut:
            return -mean <= thr * S
        return -mean < thr * S
    return False


# ------------------------------------------------------------ the two implementations
def make_impl(kind, rate):
    return gsap_amd.GoldenSection(rate) if kind == "lib" else O.GoldenSection(rate)


def run_table(kind, levels, rate=0.35, N=None, log=None):
    """Feeds level results (B after merging, S) through an implementation and through the
    transcription; returns the per-step records of both.  levels: list of (B, S) or a
    callable (restored B, B_to_merge) -> (B, S) for simulated searches."""
    impl = make_impl(kind, rate)
    snaps = [None, None, None]
    gpu = Gpu("init")
    old = OldData()
    B0 = N if N is not None else levels[0][0]
    P = Partition(B0, ref_rate_trunc(B0, rate), 0.0)
    mine = dict(B=B0, B_to_merge=P.B_to_merge, token="init")
    got, want = [], []
    for lev in range(200):
        if callable(levels):
            B, S = levels(P.B, P.B_to_merge)
        else:
            if lev >= len(levels):
                break
            B, S = levels[lev]
        token = f"L{lev}"
        # the reference: merges then sweeps leave labels `token` with P.B = B, P.S = S
        gpu.partitions[0] = token
        P.B, P.S = B, S
        opt, P = ref_prepare_next(gpu, old, P, rate, log=log)
        want.append((opt, P.B, P.B_to_merge, P.S, gpu.partitions[0],
                     [(x.B, x.S) for x in (old.large, old.med, old.small)]))
        store, restore, (nB, nBm, nS), iopt = impl.step(B, mine["B_to_merge"], S)
        snaps[store] = token
        mine = dict(B=nB, B_to_merge=nBm, token=snaps[restore])
        got.append((iopt, nB, nBm, nS, snaps[restore],
                    [(b, s) for (b, _, s, _) in impl.slots()]))
        if opt or iopt:
            break
    return got, want


# hand tables: (B after the level's merges, S); each exercises named branches
TABLES = {
    # level 1 no better and smaller -> small; first bracket; golden steps to the stop
    "bracket_then_stop": [(650, 100.0), (422, 120.0), (551, 90.0), (490, 95.0), (590, 92.0),
                          (566, 91.0), (540, 89.5), (530, 89.0), (535, 89.25), (526, 89.75),
                          (533, 89.125), (529, 89.5)],
    # a whole search from N = 1000 over S(B) = 80 + (B - 37)^2 / 64 (each level's B is the
    # restored slot's B minus its B_to_merge): eight unbracketed levels, golden steps from
    # large and from med, a tie at B = 38 / 36, and the large.B - small.B == 2 stop
    "convex_search": [(650, 5951.390625), (423, 2408.0625), (275, 965.0625), (179, 395.0625),
                      (117, 180.0), (77, 105.0), (51, 83.0625), (34, 80.140625),
                      (23, 83.0625), (45, 81.0), (41, 80.25), (30, 80.765625),
                      (38, 80.015625), (36, 80.015625), (37, 80.0)],
    # descending S while shrinking: med moves to large every level (small stays inf)
    "unbracketed_descent": [(1000, 500.0), (650, 400.0), (422, 300.0), (274, 250.0),
                            (178, 200.0), (115, 150.0), (74, 140.0), (48, 130.0),
                            (31, 129.0), (20, 128.5), (13, 128.25), (8, 128.0), (5, 127.5),
                            (3, 127.25), (2, 127.0), (1, 126.5)],
    # P.S == med.S ties (the <= of partition.hpp:91) at a smaller and at a larger B
    "ties": [(650, 100.0), (422, 100.0), (274, 100.0), (178, 150.0), (240, 100.0),
             (300, 100.0), (250, 99.0), (260, 99.0)],
    # a worse result with more blocks than med replaces large (partition.hpp:204-230)
    "worse_larger": [(650, 100.0), (422, 90.0), (274, 95.0), (330, 96.0), (380, 89.0),
                     (400, 89.5), (360, 88.0), (370, 88.5), (350, 88.25)],
    # the large/med gap 250 lands on an exact half: 250 * 0.618 = 154.5 rounds up to 155
    "round_half": [(1000, 50.0), (750, 40.0), (500, 30.0), (250, 35.0), (345, 31.0)],
    # B_to_merge = (int)B * 0.35f reaching 0 ends the unbracketed search (partition.hpp:247)
    "rate_zero": [(3, 10.0), (2, 9.0)],
    # B = 180 and 20: float32 products 63.0f and 7.0f, where double products truncate to
    # 62 and 6
    "float_trunc": [(1000, 10.0), (180, 9.0), (20, 8.0)],
}


@pytest.mark.parametrize("kind", ["lib", "oracle"])
@pytest.mark.parametrize("name", sorted(TABLES))
def test_golden_tables(kind, name):
    got, want = run_table(kind, TABLES[name])
    assert got == want, name


def test_golden_table_branches_covered():
    """The hand tables reach both placements of a better and of a worse result and every
    exit of partition.hpp:163-320, including the two stops."""
    log = []
    for levels in TABLES.values():
        run_table("oracle", levels, log=log)
    for e in [("le", True), ("le", False), ("gt", True), ("gt", False), "unbracketed",
              "unbracketed_stop", "gap2_stop", "golden_large", "golden_med"]:
        assert e in log, e


def test_rate_truncation_is_float32():
    """(int)P.B * 0.35f is a float product: at B = 180 it rounds to 63.0f, while the double
    product 180 * 0.35 = 62.99999999999999 truncates to 62 (and at B = 20 the product with
    (double)0.35f, 6.99999988, to 6 instead of 7); both implementations follow float."""
    for kind in ("lib", "oracle"):
        for B, want in ((180, 63), (20, 7), (200000, 70000)):
            g = make_impl(kind, 0.35)
            _, _, (nB, Bm, _), _ = g.step(B, 0, 10.0)
            assert (nB, Bm) == (B, want)
    assert int(180 * 0.35) == 62 and int(20 * float(np.float32(0.35))) == 6
    # every B up to 2^20 where float and double products truncate differently
    B = np.arange(1, 1 << 20, dtype=np.int64)
    f32 = (B.astype(np.float32) * np.float32(0.35)).astype(np.int64)
    f64 = np.floor(B * 0.35).astype(np.int64)
    diff = B[f32 != f64]
    assert len(diff) > 0
    for kind in ("lib", "oracle"):
        for b in diff[:: max(1, len(diff) // 64)]:
            g = make_impl(kind, 0.35)
            _, _, (_, Bm, _), _ = g.step(int(b), 0, 1.0)
            assert Bm == int(np.float32(b) * np.float32(0.35)), b


def simulated_levels(seed, N):
    """A search over a noisy convex S(B) with quantised values (so ties happen) and merge
    phases that now and then fall short of B_to_merge (the build's union-find stops after B
    candidates), which can give two slots the same B."""
    rng = np.random.default_rng(seed)
    opt_B = int(rng.integers(2, max(3, N // 3)))
    scale = float(rng.uniform(1, 50))
    q = float(2.0 ** int(rng.integers(-2, 4)))
    noise = float(rng.uniform(0, 4))
    short_p = float(rng.choice([0.0, 0.0, 0.2]))

    def f(B, Bm):
        nB = max(1, B - Bm)
        if short_p and rng.random() < short_p:
            nB = min(B, nB + int(rng.integers(0, 3)))
        x = math.log(nB / opt_B)
        S = 1000.0 + scale * x * x * 10 + noise * rng.standard_normal()
        return nB, round(S / q) * q
    return f


@pytest.mark.parametrize("kind", ["lib", "oracle"])
def test_golden_simulated_searches(kind):
    n_bracketed = n_ties = 0
    for seed in range(1500):
        N = int(np.random.default_rng(seed + 99).integers(3, 5000))
        got, want = run_table(kind, simulated_levels(seed, N), N=N)
        # the one deliberate difference: a bracketed step of B_to_merge <= 0 ends the search
        # here, where the reference repeats a level without merging
        if got != want:
            assert got[:-1] == want[:-1], seed
            g, w = got[-1], want[-1]
            assert g[0] and not w[0] and w[2] <= 0 and g[1:] == w[1:], (seed, g, w)
            continue
        n_bracketed += any(math.isfinite(r[5][2][1]) for r in want)
        n_ties += sum(1 for a, b in zip(want, want[1:]) if a[3] == b[3])
    assert n_bracketed > 500 and n_ties > 50


def test_zero_merge_step_ends_search():
    """Two slots of equal B: the reference's golden step gives B_to_merge = 0 and would rerun
    the level without merging; the library and the oracle end the search there instead."""
    levels = [(10, 50.0), (3, 49.0), (4, 40.0), (3, 32.0)]
    for kind in ("lib", "oracle"):
        got, want = run_table(kind, levels)
        assert got[:-1] == want[:-1]
        assert want[-1][2] <= 0 and not want[-1][0]
        assert got[-1][0] and got[-1][1:] == want[-1][1:]


# ------------------------------------------------------------ convergence rule
def conv_cases():
    rng = np.random.default_rng(7)
    cases = []
    for _ in range(3000):
        itr = int(rng.integers(0, 12))
        W = int(rng.choice([2, 3, 3, 3, 4, 5]))
        S = float(rng.choice([1e3, 2.5e5, 9.4e7, 1.2e9]))
        scale = S * float(rng.choice([1e-5, 1e-4, 5e-4, 1e-3]))
        dE = [-abs(float(rng.standard_normal())) * scale * float(rng.choice([0.5, 1, 2]))
              for _ in range(itr + 1)]
        bracketed = bool(rng.integers(0, 2))
        cases.append((dE, itr, S, bracketed, W))
    # hand cases: the current sweep is outside the window
    S = 1 << 20
    cases.append(([-1000.0, -1000.0, -1e9], 2, float(S), False, 3))   # would not stop w/ current
    cases.append(([-1.0, -1.0, -1e9], 2, float(S), False, 3))         # stops: current ignored
    cases.append(([-1e9, -1.0, -1.0], 2, float(S), False, 3))         # sweep 0 is in the window
    # exact boundary: -mean == threshold * S is not below it
    t = 0.5 ** 10
    cases.append(([-t * S, -t * S, 0.0], 2, float(S), False, 3, t, t))
    cases.append(([-t * S * 0.5, -t * S * 1.5, 0.0], 2, float(S), True, 3, t, t))
    # threshold switch: 5e-4 before bracketing, 1e-4 after
    cases.append(([-3e-4 * S] * 3, 2, float(S), False, 3))
    cases.append(([-3e-4 * S] * 3, 2, float(S), True, 3))
    cases.append(([-3e-4 * S] * 2, 1, float(S), False, 3))           # itr < W-1: never
    cases.append(([-3e-4 * S] * 2, 1, float(S), False, 2))
    return cases


def conv_ref(case, mut=()):
    dE, itr, S, bracketed, W = case[:5]
    t1, t2 = (case[5], case[6]) if len(case) > 5 else (5e-4, 1e-4)
    old = OldData()
    if bracketed:
        old.large.S = old.med.S = old.small.S = 1.0
    return ref_converged(dE, itr, S, old, W, t1, t2, mut)


@pytest.mark.parametrize("kind", ["lib", "oracle"])
def test_convergence_rule(kind):
    f = gsap_amd.converged if kind == "lib" else O.converged
    stops = 0
    for case in conv_cases():
        dE, itr, S, bracketed, W = case[:5]
        t1, t2 = (case[5], case[6]) if len(case) > 5 else (5e-4, 1e-4)
        want = conv_ref(case)
        assert f(dE, itr, S, bracketed, W, t1, t2) == want, case
        stops += want
    assert 300 < stops < 2700


# ------------------------------------------------------------ the tables' power
GOLDEN_MUTANTS = ["strict_le", "swap_med_branch", "swap_worse_branch", "stop_gap3",
                  "bankers_round", "strict_ge", "double_rate"]
CONV_MUTANTS = ["window_includes_current", "divide_by_W", "le", "isf_inverted"]


@pytest.mark.parametrize("mut", GOLDEN_MUTANTS)
def test_tables_catch_golden_mutants(mut, monkeypatch):
    """A transcription with one planted mistake disagrees with the oracle on the tables."""
    orig = ref_prepare_next

    def mutated(gpu, old, P, rate, m=(), log=None):
        return orig(gpu, old, P, rate, (mut,), log)
    monkeypatch.setitem(globals(), "ref_prepare_next", mutated)
    caught = False
    for name, levels in TABLES.items():
        got, want = run_table("oracle", levels)
        caught |= got != want
    assert caught, mut


@pytest.mark.parametrize("mut", CONV_MUTANTS)
def test_tables_catch_convergence_mutants(mut):
    caught = False
    for case in conv_cases():
        dE, itr, S, bracketed, W = case[:5]
        t1, t2 = (case[5], case[6]) if len(case) > 5 else (5e-4, 1e-4)
        caught |= O.converged(dE, itr, S, bracketed, W, t1, t2) != conv_ref(case, (mut,))
    assert caught, mut
