"""Synthetic request streams for the batched and served solves: a mixed-size
batch of random NACA four-digit sections and a Poisson stream of the same
draws. Both come from a 64-bit Mersenne Twister with an explicit uniform
(u = (x >> 11) * 2^-53), so the manifests are the same on every platform;
the committed JSON files under workloads/ are what the bench replays, and
``python workloads.py`` writes them again.

Every instance: chord 1, closed trailing edge, C-grid domain 15 chords upstream
and vertical and 20 downstream, k-omega SST, Tu 0.1 %, mu_t/mu 10, freestream
U = rho = 1 (coefficients are those of any U, rho at the same Re). The mesh
level picks one of three C-grid sizes; the first layer follows the y+ ~ 1 rule
2.5e-5 * (1e6 / Re)^0.9 of the reference's turbulence tests
(tests/TurbulenceTests.cpp:630-632) at the instance's Re.
"""
from __future__ import annotations

import json
import math
import os

_MASK = (1 << 64) - 1

# (surfacePoints, wakePoints, normalPoints) -> cells (2*wake + 2*surface - 2) x (normal - 1)
LEVELS = {
    "coarse": (97, 32, 65),     # 256 x 64
    "medium": (193, 64, 129),   # 512 x 128
    "fine": (385, 128, 257),    # 1024 x 256
}
LEVEL_NAMES = ("coarse", "medium", "fine")
HERE = os.path.dirname(os.path.abspath(__file__))
MIXED_BATCH_FILE = os.path.join(HERE, "workloads", "mixed_batch_64_seed0.json")
SERVING_FILE = os.path.join(HERE, "workloads", "serving_poisson20_600s_seed1.json")


class MT19937_64:
    """The 64-bit Mersenne Twister (Matsumoto and Nishimura, 2004; the
    parameters of std::mt19937_64)."""

    n, m = 312, 156
    a = 0xB5026F5AA96619E9
    upper, lower = 0xFFFFFFFF80000000, 0x7FFFFFFF

    def __init__(self, seed: int = 5489):
        self.mt = [0] * self.n
        self.mt[0] = seed & _MASK
        for i in range(1, self.n):
            prev = self.mt[i - 1]
            self.mt[i] = (6364136223846793005 * (prev ^ (prev >> 62)) + i) & _MASK
        self.index = self.n

    def _twist(self):
        mt = self.mt
        for i in range(self.n):
            x = (mt[i] & self.upper) | (mt[(i + 1) % self.n] & self.lower)
            xa = x >> 1
            if x & 1:
                xa ^= self.a
            mt[i] = mt[(i + self.m) % self.n] ^ xa
        self.index = 0

    def __call__(self) -> int:
        if self.index >= self.n:
            self._twist()
        y = self.mt[self.index]
        self.index += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & _MASK

    def u01(self) -> float:
        return (self() >> 11) * (1.0 / 9007199254740992.0)


def first_layer(reynolds: float) -> float:
    return 2.5e-5 * (1e6 / reynolds) ** 0.9


def draw_instance(g: MT19937_64) -> dict:
    """One instance, draws in the order m, p, t, alpha, log10 Re, level."""
    m = min(6, int(g.u01() * 7))
    p = 2 + min(4, int(g.u01() * 5))
    if m == 0:
        p = 0  # a symmetric section has no camber position (Naca4.cpp:83-88)
    t = 8 + min(10, int(g.u01() * 11))
    alpha = -5.0 + 17.0 * g.u01()
    lo, hi = math.log10(2e5), math.log10(5e6)
    reynolds = 10.0 ** (lo + (hi - lo) * g.u01())
    level = LEVEL_NAMES[min(2, int(g.u01() * 3))]
    return {"m": m, "p": p, "t": t, "alpha": alpha, "reynolds": reynolds, "level": level}


def mixed_batch(seed: int = 0, count: int = 64) -> list[dict]:
    g = MT19937_64(seed)
    return [draw_instance(g) for _ in range(count)]


def serving_stream(seed: int = 1, rate: float = 20.0, duration: float = 600.0) -> list[dict]:
    """Poisson arrivals: per request an exponential gap -ln(1 - u) / rate, then
    the instance draws; the stream ends at the first arrival past `duration`."""
    g = MT19937_64(seed)
    out, clock = [], 0.0
    while True:
        clock += -math.log(1.0 - g.u01()) / rate
        if clock > duration:
            return out
        item = {"arrival": clock}
        item.update(draw_instance(g))
        out.append(item)


def mesh_options(G, level: str, reynolds: float):
    """The C-grid options of one instance (the binding's cgrid_options)."""
    s, w, n = LEVELS[level]
    return G.cgrid_options("coarse", upstream_chords=15.0, vertical_chords=15.0, downstream_chords=20.0,
                           surface_points=s, wake_points=w, normal_points=n,
                           first_layer_height=first_layer(reynolds))


_FIELDS = ["m", "p", "t", "alpha", "reynolds", "level"]


def _dump(path: str, header: dict, rows: list[dict], fields: list[str]):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write(json.dumps(header, indent=1)[:-2])
        f.write(',\n "fields": ' + json.dumps(fields) + ',\n "rows": [\n')
        for q, r in enumerate(rows):
            vals = [repr(r[k]) if isinstance(r[k], float) else json.dumps(r[k]) for k in fields]
            f.write("  [" + ", ".join(vals) + "]" + (",\n" if q + 1 < len(rows) else "\n"))
        f.write(" ]\n}\n")


def load(path: str) -> list[dict]:
    with open(path) as f:
        doc = json.load(f)
    return [dict(zip(doc["fields"], row)) for row in doc["rows"]]


def write_all():
    common = {"generator": "mt19937_64, u = (x >> 11) * 2^-53; draws m, p, t, alpha, log10 Re, level",
              "section": "NACA mpt, closed trailing edge, chord 1",
              "domain_chords": {"upstream": 15, "vertical": 15, "downstream": 20},
              "levels": {k: list(v) for k, v in LEVELS.items()},
              "first_layer": "2.5e-5 * (1e6 / Re)^0.9",
              "turbulence": {"model": "k-omega SST", "intensity": 0.001, "viscosity_ratio": 10.0}}
    _dump(MIXED_BATCH_FILE, dict(common, name="mixed-size batch", seed=0, count=64), mixed_batch(0, 64), _FIELDS)
    stream = serving_stream(1, 20.0, 600.0)
    _dump(SERVING_FILE, dict(common, name="serving stream", seed=1, rate_per_s=20.0, duration_s=600.0,
                             arrivals="exponential gaps -ln(1 - u) / rate, drawn before each instance",
                             count=len(stream)), stream, ["arrival"] + _FIELDS)


if __name__ == "__main__":
    write_all()
