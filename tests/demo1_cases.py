"""Demo1::Run schedules shared by the oracle tests and the GPU parity tests.

The solve loop of src/demo1.cc:36-59: before iteration i,
  B_i = max(2 (k - |found after i-1|) + 1, min_bins)     (demo1.cc:48-49)
  bin threshold 3 sigma / sqrt(B_i), sv threshold sigma / sqrt(B_i)  (demo1.cc:50-51)
and stale counts the iterations whose Iterate returned false, never reset
(demo1.cc:55-57); the loop ends at the max_stale_iter-th such iteration.
"""
import numpy as np

HAND_SCHEDULES = [
    # k, sigma, min_bins, max_stale, script (res, |found| after), expected bins
    (10, 0.5, 5, 3, [(1, 4), (0, 4), (1, 7), (0, 7), (0, 7), (1, 8), (1, 9)],
     [21, 13, 13, 7, 7]),
    (3, 2.0, 5, 2, [(1, 5), (1, 6), (0, 6), (0, 6), (1, 7)], [7, 5, 5, 5]),
    (50, 1e-3, 11, 1, [(1, 20), (1, 45), (0, 45), (1, 46)], [101, 61, 11]),
    (7, 1e-6, 3, 4, [(0, 0), (1, 2), (0, 2), (1, 5), (0, 5), (1, 7), (0, 7), (1, 7)],
     [15, 15, 11, 11, 5, 5, 3]),
]


# Solves (n, k, sigma, min_bins, max_stale, signal seed, solve seed) whose
# schedule has an iteration that finds modes after a stale one.
STALE_THEN_FOUND = [
    (1 << 12, 20, 1e-3, 21, 4, 3, 103),
    (1 << 12, 30, 3e-3, 31, 5, 1, 101),
    (1 << 13, 40, 1e-2, 41, 5, 4, 104),
    (1 << 12, 25, 3e-2, 25, 6, 0, 100),
]


def check_demo1_schedule(k, sigma, min_bins, max_stale, bins, found, res=None):
    """The Demo1::Run recurrences on a solve's per-iteration log."""
    prev = 0
    for i, b in enumerate(bins):
        assert b == max(2 * (k - prev) + 1, min_bins), (i, b, prev)
        prev = found[i]
    if res is not None:
        stale = np.cumsum(1 - np.asarray(res, np.int64))
        assert stale[-1] == max_stale and res[-1] == 0
        assert np.all(stale[:-1] < max_stale)
        # Iterate returned true exactly where the map grew, or where it only
        # updated modes already present (the map size then stays)
        grew = np.diff(np.concatenate([[0], np.asarray(found)])) > 0
        assert np.all(np.asarray(res, bool)[grew])
