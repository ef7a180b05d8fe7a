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


def make_case(inst, workdir, name, jitter=0.1, jitter_seed=1, numbering=1, **case_kw):
    """Writes the settings files of one instance and opens it as a Case on a
    generated wedge mesh (ramp angle inst['theta'], inflow Mach inst['mach'])."""
    prefix = os.path.join(workdir, name)
    write_case_files(prefix, mach=inst["mach"], **case_kw)
    p = wedge_params(nx=inst["nx"], ny=inst["ny"], ramp_deg=inst["theta"], jitter=jitter, jitter_seed=jitter_seed,
                     numbering=numbering)
    return Case.wedge(p, prefix)
