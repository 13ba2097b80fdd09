"""Seeded synthetic Gaussian-source workloads.

One portable generator (splitmix64 -> doubles in [0,1), Box-Muller normals)
so the same draws can be reproduced in any language.  A source set is a sum of
normalised Gaussians f(r) = sum_i c_i (beta_i/pi)^{3/2} exp(-beta_i |r-R_i|^2),
handed to project() as (coefs, exponents, centres).
"""
import math

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK64

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def log_uniform(self, lo, hi):
        return math.exp(self.uniform(math.log(lo), math.log(hi)))

    def normal(self):
        u1 = self.uniform()
        u2 = self.uniform()
        u1 = max(u1, 1e-300)
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def randint(self, lo, hi):
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1))


class GaussianSources:
    """Sum of normalised Gaussians: coefficients already include (beta/pi)^{3/2}."""

    def __init__(self, weights, exponents, centers):
        self.weights = [float(w) for w in weights]
        self.exponents = [float(b) for b in exponents]
        self.centers = [tuple(float(x) for x in c) for c in centers]

    @property
    def coefs(self):
        return [w * (b / math.pi) ** 1.5 for w, b in zip(self.weights, self.exponents)]

    def __len__(self):
        return len(self.weights)


def single_gaussian(beta=100.0, center=(0.0, 0.0, 0.0)):
    """One unit-charge Gaussian (the k=5, [-20,20]^3 CPU case)."""
    return GaussianSources([1.0], [beta], [center])


def random_sources(seed, n, center_range=4.0, beta_range=(1.0, 1000.0), normal_weights=True):
    """n Gaussians, centres U[-c,c]^3, exponents log-uniform, weights N(0,1) (or 1)."""
    rng = SplitMix64(seed)
    w, b, c = [], [], []
    for _ in range(n):
        c.append(tuple(rng.uniform(-center_range, center_range) for _ in range(3)))
        b.append(rng.log_uniform(*beta_range))
        w.append(rng.normal() if normal_weights else 1.0)
    return GaussianSources(w, b, c)


# Named workloads.  k = Legendre order, box = [-L, L]^3.
WORKLOADS = {
    "gauss1_k5": dict(sources=lambda: single_gaussian(100.0), order=5, box=20, prec=1e-5, mu=None),
    "gauss16_k7": dict(sources=lambda: random_sources(20240607, 16, 4.0), order=7, box=32, prec=1e-6, mu=None),
    "gauss1_k9_poisson": dict(sources=lambda: single_gaussian(100.0), order=9, box=20, prec=1e-7, mu=None),
    "gauss1_k9_helmholtz": dict(sources=lambda: single_gaussian(100.0), order=9, box=20, prec=1e-7, mu=1.0),
}


def batch_sources(seed=77, n_trees=64):
    """64 uneven trees: 1..32 Gaussians, centres U[-6,6]^3, prec in {1e-5,1e-6,1e-7};
    even index Poisson, odd index Helmholtz with mu ~ U[0.5,3]."""
    rng = SplitMix64(seed)
    out = []
    for t in range(n_trees):
        n = rng.randint(1, 32)
        src = random_sources(rng.next_u64(), n, 6.0)
        prec = (1e-5, 1e-6, 1e-7)[rng.randint(0, 2)]
        mu = None if t % 2 == 0 else rng.uniform(0.5, 3.0)
        out.append(dict(sources=src, order=7, box=32, prec=prec, mu=mu))
    return out


def predictor_sources(seed=2024, n_trees=250):
    """Trees for the refinement predictor: 4..32 Gaussians, centres
    U[-6,6]^3, exponents log-uniform in [1,1000], order 7, prec 1e-6; even
    index Poisson, odd index Helmholtz with mu ~ U[0.5,3]."""
    rng = SplitMix64(seed)
    out = []
    for t in range(n_trees):
        n = rng.randint(4, 32)
        src = random_sources(rng.next_u64(), n, 6.0)
        mu = None if t % 2 == 0 else rng.uniform(0.5, 3.0)
        out.append(dict(sources=src, order=7, box=32, prec=1e-6, mu=mu))
    return out
