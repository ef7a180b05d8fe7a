"""Parameter sweeps of the compression-ramp channel: a reproducible random
generator (splitmix64-seeded xoshiro256**, the same generator the C++ mesh
generator uses for its jitter and permutations) and the draw of instances
(cells, Mach number, ramp angle) for batched solves."""
import math
import os

from . import Case, wedge_params, write_case_files

MASK = (1 << 64) - 1


def _splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & MASK
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return x, z ^ (z >> 31)


def _rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & MASK


class Xoshiro256ss:
    """xoshiro256** (Blackman and Vigna) seeded by splitmix64."""

    def __init__(self, seed):
        st = seed & MASK
        self.s = []
        for _ in range(4):
            st, z = _splitmix64(st)
            self.s.append(z)

    def next_u64(self):
        s = self.s
        result = (_rotl((s[1] * 5) & MASK, 7) * 9) & MASK
        t = (s[1] << 17) & MASK
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = _rotl(s[3], 45)
        return result

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))


def draw_instances(n, seed, cells=(20000, 400000), mach=(1.5, 4.0), theta=(5.0, 25.0), aspect=3):
    """n wedge instances: cell count log-uniform in `cells` (nx = aspect*ny
    quads), Mach number and ramp angle (degrees) uniform."""
    rng = Xoshiro256ss(seed)
    out = []
    for _ in range(n):
        nc = math.exp(rng.uniform(math.log(cells[0]), math.log(cells[1])))
        m = rng.uniform(*mach)
        th = rng.uniform(*theta)
        ny = max(4, int(round(math.sqrt(nc / aspect))))
        out.append({"nx": aspect * ny, "ny": ny, "mach": m, "theta": th})
    return out


def theta_max(mach, gamma=1.4, n=4000):
    """Largest deflection (degrees) an attached oblique shock turns a flow of
    Mach number `mach` through: the maximum over the shock angle beta of the
    theta-beta-Mach relation tan(theta) = 2 cot(beta) (M^2 sin^2(beta) - 1) /
    (M^2 (gamma + cos(2 beta)) + 2).  A ramp steeper than this detaches the
    shock."""
    mu = math.asin(1.0 / mach)
    best = 0.0
    for k in range(1, n):
        b = mu + (math.pi / 2 - mu) * k / n
        t = math.atan(2.0 / math.tan(b) * (mach * mach * math.sin(b) ** 2 - 1.0) /
                      (mach * mach * (gamma + math.cos(2.0 * b)) + 2.0))
        best = max(best, t)
    return math.degrees(best)


def oblique_shock(mach, theta_deg, gamma=1.4, n=4000):
    """Weak oblique shock of a flow at Mach `mach` turned through
    `theta_deg`: (shock angle in degrees, Mach number behind it), or None
    when the turn exceeds theta_max (detached shock).  The smallest shock
    angle of the theta-beta-Mach relation that reaches the turn, then the
    normal-shock relation for the normal Mach component."""
    th = math.radians(theta_deg)
    mu = math.asin(1.0 / mach)
    for k in range(1, n):
        b = mu + (math.pi / 2 - mu) * k / n
        t = math.atan(2.0 / math.tan(b) * (mach * mach * math.sin(b) ** 2 - 1.0) /
                      (mach * mach * (gamma + math.cos(2.0 * b)) + 2.0))
        if t >= th:
            mn1 = mach * math.sin(b)
            mn2 = math.sqrt((1.0 + 0.5 * (gamma - 1.0) * mn1 * mn1) / (gamma * mn1 * mn1 - 0.5 * (gamma - 1.0)))
            return math.degrees(b), mn2 / math.sin(b - th)
    return None


def reflection_regime(mach, theta_deg, gamma=1.4):
    """How the ramp's shock meets the channel's top wall: "regular" when the
    flow behind it is supersonic and can be turned back through theta by a
    reflected oblique shock (theta <= theta_max(M2)), "mach" when it cannot
    (a Mach reflection, with a subsonic region behind its stem), "detached"
    when the ramp itself is steeper than theta_max(mach)."""
    sh = oblique_shock(mach, theta_deg, gamma)
    if sh is None:
        return "detached"
    m2 = sh[1]
    return "regular" if m2 > 1.0 and theta_deg <= theta_max(m2, gamma) else "mach"


def make_case(inst, workdir, name, jitter=0.1, jitter_seed=1, numbering=1, **case_kw):
    """Writes the settings files of one instance and opens it as a Case on a
    generated wedge mesh (ramp angle inst['theta'], inflow Mach inst['mach'])."""
    prefix = os.path.join(workdir, name)
    write_case_files(prefix, mach=inst["mach"], **case_kw)
    p = wedge_params(nx=inst["nx"], ny=inst["ny"], ramp_deg=inst["theta"], jitter=jitter, jitter_seed=jitter_seed,
                     numbering=numbering)
    return Case.wedge(p, prefix)


def build_cases(instances, workdir, threads=16, prefix="s", **case_kw):
    """Opens every instance as a Case, generating meshes on `threads` host
    threads (the generator runs in C++ and releases the interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor

    def one(k):
        return make_case(instances[k], workdir, "%s%d" % (prefix, k), **case_kw)

    with ThreadPoolExecutor(max(1, threads)) as ex:
        return list(ex.map(one, range(len(instances))))


def solve_to_tolerance(batch, tol=1e-8, switch_tol=1e-3, chunk=1000, budget=None, diverge=1e3, stall=0.7, log=None,
                       fallback=None, restart_growth=100.0):
    """Runs a batch until every instance's residual drop reaches `tol`.

    If the batch is in FP32, each instance first runs in FP32 until its drop
    reaches `switch_tol` or stops improving (an FP32 state floors the drop
    near 1e-4: an instance whose drop, below 1e-2, fell by less than a
    factor `stall` over the last `chunk` iterations leaves the FP32 phase);
    then the batch switches to FP64 and runs to `tol`.  An instance is given
    up when its drop exceeds `diverge` (or turns non-finite) or when it has
    used its iteration budget (`budget`, per instance; default unlimited).

    `fallback` ({instance: primitive state}) gives instances a second start:
    one that turns non-finite, exceeds `diverge`, or whose drop rises above
    `restart_growth` times its best drop so far is reset, once, to that state
    in the same batch (same residual reference), and the iterations it spent
    count towards its total.

    Returns per-instance iteration counts of both phases, the final drop,
    the outcome (0 converged, 1 budget spent, 2 diverged), whether it was
    restarted from its fallback, and the wall time of each phase (seconds)."""
    import time

    import numpy as np

    n = batch.n
    budget = np.full(n, 1 << 40, dtype=np.int64) if budget is None else np.asarray(budget, dtype=np.int64)
    iters = {32: np.zeros(n, dtype=np.int64), 64: np.zeros(n, dtype=np.int64)}
    outcome = np.full(n, -1, dtype=np.int64)
    drop = np.full(n, np.inf)
    best = np.full(n, np.inf)
    restarted = np.zeros(n, dtype=bool)
    walls = {32: 0.0, 64: 0.0}
    fallback = fallback or {}

    def phase(prec, target):
        t0 = time.perf_counter()
        running = set(int(i) for i in np.nonzero(outcome < 0)[0])
        for i in range(n):
            batch.set_enabled(i, i in running)
        prev = {}
        total = 0
        while running:
            out = batch.run(chunk, tol=target, allow_failures=True)
            total += chunk
            iters[prec] += out["iters"]
            for i in sorted(running):
                d = float(out["drop"][i])
                drop[i] = d
                used = iters[32][i] + iters[64][i]
                failed = out["status"][i] == 3 or not d < diverge
                if i in fallback and not restarted[i] and (failed or d > restart_growth * best[i]):
                    batch.set_state(i, fallback[i])
                    restarted[i] = True
                    best[i] = np.inf
                    prev.pop(i, None)
                    if used >= budget[i]:
                        outcome[i] = 1
                        batch.set_enabled(i, False)
                        running.discard(i)
                    continue
                best[i] = min(best[i], d)
                if failed:
                    outcome[i] = 2
                elif out["status"][i] == 0:
                    if prec == 64 or target <= tol:
                        outcome[i] = 0
                elif used >= budget[i]:
                    outcome[i] = 1
                elif not (prec == 32 and i in prev and d < 1e-2 and d > stall * prev[i]):
                    prev[i] = d
                    continue
                batch.set_enabled(i, False)
                running.discard(i)
            if log:
                log("fp%d phase: %d iterations, %d instances still running" % (prec, total, len(running)))
        walls[prec] = time.perf_counter() - t0

    if batch.precision == 32:
        phase(32, switch_tol)
        batch.set_precision(64)
    phase(64, tol)
    for i in range(n):
        batch.set_enabled(i, True)
    return {"iters32": iters[32], "iters64": iters[64], "drop": drop, "outcome": outcome, "restarted": restarted,
            "wall32": walls[32], "wall64": walls[64], "wall": walls[32] + walls[64]}
