"""The synthetic workload generators used by the GPU tests."""
import numpy as np

from workloads import MT19937_64, random_spacetime_batch


def test_mt19937_64_known_answer():
    # C++ [rand.predef]: the 10000th output of a default-seeded mt19937_64 (seed 5489)
    g = MT19937_64(5489)
    for _ in range(9999):
        g.next()
    assert g.next() == 9981545732273789042


def test_random_batch_ranges():
    cams, metrics = random_spacetime_batch()
    assert len(cams) == len(metrics) == 64
    a = np.array([m.a for m in metrics])
    Q = np.array([m.Q for m in metrics])
    assert np.all((a >= 0) & (a <= 0.99)) and np.all(Q <= 0.5 * np.sqrt(1 - a * a))
    assert np.all(1 + np.sqrt(1 - a * a - Q * Q) > 1)  # every spacetime has a horizon
    n = sum(c.width * c.height for c in cams)
    assert 5e6 < n < 4e7  # ~17.9 M rays expected for 64 cameras
    small, _ = random_spacetime_batch(scale=16)
    assert [c.width * 16 for c in small] == [max(64, c.width) for c in cams]
