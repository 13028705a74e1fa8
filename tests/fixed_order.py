"""The fixed summation order of the device's FP64 reductions, restated in
numpy: blocks of 256 threads, thread t adds elements t, t + 256, ... of its
range to 0.0 in turn, then the 256 thread sums fold as a tree
(sum[t] += sum[t + h] for h = 128, 64, ..., 1); the node sums first go through
1024 such blocks over contiguous ranges of 256 * ceil(n / 262144) nodes."""
import numpy as np

RED_BLOCKS = 1024


def block_sum(x):
    x = np.asarray(x, np.float64)
    s = np.zeros(256)
    for j in range(0, len(x), 256):
        seg = x[j:j + 256]
        s[:len(seg)] = s[:len(seg)] + seg
    for h in (128, 64, 32, 16, 8, 4, 2, 1):
        s[:h] = s[:h] + s[h:2 * h]
    return s[0]


def node_sum(x):
    x = np.asarray(x, np.float64)
    chunk = 256 * max(1, -(-len(x) // (256 * RED_BLOCKS)))
    return block_sum([block_sum(x[g * chunk:(g + 1) * chunk]) for g in range(RED_BLOCKS)])
