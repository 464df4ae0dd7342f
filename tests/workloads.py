"""Synthetic workloads shared by the tests: the 64-camera random spacetime batch, the
Morris-Thorne / Alcubierre scenes and the wormhole's two celestial-sphere catalogues.

The random batch draws, for each of 64 cameras in turn, a ~ U[0, 0.99],
Q ~ U[0, 0.5 sqrt(1 - a^2)], Lambda ~ U[0, 1e-4], inclination ~ U[5, 89] deg,
r_obs ~ U[20, 100] and a size index ~ U{0..4}, from a 64-bit Mersenne Twister
(mt19937_64) seeded with 0x51121005. A uniform double is the top 53 bits of one output
scaled by 2^-53, and the size index is floor(5 u). Full-size images are {64, 128, 256, 512,
1024}^2 (about 17.9 M rays in all); `scale` divides the side lengths for parity runs that the
CPU oracle must finish in seconds.
"""
import numpy as np

from sirius_amd import abi


class MT19937_64:
    """The 64-bit Mersenne Twister (Matsumoto & Nishimura 2000; the std::mt19937_64 engine)."""
    N, M = 312, 156
    A = 0xB5026F5AA96619E9
    UPPER, LOWER = 0xFFFFFFFF80000000, 0x7FFFFFFF
    MASK = (1 << 64) - 1

    def __init__(self, seed):
        self.mt = [0] * self.N
        self.mt[0] = seed & self.MASK
        for i in range(1, self.N):
            prev = self.mt[i - 1]
            self.mt[i] = (6364136223846793005 * (prev ^ (prev >> 62)) + i) & self.MASK
        self.idx = self.N

    def _twist(self):
        mt = self.mt
        for i in range(self.N):
            x = (mt[i] & self.UPPER) | (mt[(i + 1) % self.N] & self.LOWER)
            xa = x >> 1
            if x & 1:
                xa ^= self.A
            mt[i] = mt[(i + self.M) % self.N] ^ xa
        self.idx = 0

    def next(self):
        if self.idx >= self.N:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & self.MASK

    def uniform(self, lo, hi):
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / (1 << 53)))


SIDES = (64, 128, 256, 512, 1024)


def random_spacetime_batch(seed=0x51121005, n_cam=64, scale=1, fov_deg=30.0):
    """(cameras, metrics) of the random Kerr-Newman-de Sitter batch; camera i uses metric i."""
    g = MT19937_64(seed)
    cams, metrics = [], []
    for i in range(n_cam):
        a = g.uniform(0.0, 0.99)
        Q = g.uniform(0.0, 0.5 * np.sqrt(1.0 - a * a))
        Lam = g.uniform(0.0, 1e-4)
        incl = g.uniform(5.0, 89.0)
        r_obs = g.uniform(20.0, 100.0)
        side = max(4, SIDES[min(4, int(5 * g.uniform(0.0, 1.0)))] // scale)
        metrics.append(abi.metric_ks(1.0, a, Q, Lam))
        cams.append(abi.camera(r_obs, incl, fov_deg, side, side, disk=(0.0, 30.0), metric_index=i))
    return cams, metrics


def wormhole_scene(side):
    """Ellis wormhole b0 = 1 seen from r = 10 (looking at the throat)."""
    return abi.camera(10.0, 90.0, 60.0, side, side), abi.metric_mt(1.0)


def warp_scene(side):
    """Static Alcubierre bubble v = 0.5, R = 1, sigma = 8, observer 10 units off-axis."""
    return abi.camera(10.0, 60.0, 40.0, side, side, phi_deg=30.0), abi.metric_warp(0.5, 1.0, 8.0)


def sphere_catalogue(seed, n=4096):
    """A celestial sphere of n stars (rows of SiriusHipStar: direction x, y, z, distance (pc),
    apparent magnitude, B-V colour index, temperature (K), unused), drawn from an mt19937_64
    seeded with `seed`: per star, a direction uniform on the sphere (z ~ U[-1, 1],
    phi ~ U[0, 2 pi)), a magnitude ~ Exp(1) (-ln(1 - u)), a distance ~ U[10, 1000] pc and a B-V
    index ~ U[-0.3, 1.7] with the temperature the reference's generator gives that index
    (4600 K / (0.92 + B-V), PHSF001A.h:51-73). The two spheres of the wormhole scene use seeds
    0x51121008 (near side) and 0x51121009 (far side)."""
    g = MT19937_64(seed)
    rows = np.zeros((n, 8), np.float32)
    for i in range(n):
        z = g.uniform(-1.0, 1.0)
        ph = g.uniform(0.0, 2.0 * np.pi)
        mag = -np.log(1.0 - g.uniform(0.0, 1.0))
        dist = g.uniform(10.0, 1000.0)
        bv = g.uniform(-0.3, 1.7)
        rho = np.sqrt(max(0.0, 1.0 - z * z))
        temp = 4600.0 / (0.92 + bv) if bv > -0.1 else 30000.0 + 10000.0 * (-bv - 0.1)
        rows[i] = (rho * np.cos(ph), rho * np.sin(ph), z, dist, mag, bv, min(max(temp, 2500.0), 50000.0), 0.0)
    return rows


WORMHOLE_SPHERE_SEEDS = (0x51121008, 0x51121009)
