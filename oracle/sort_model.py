"""CPU model of the device key-pair merge sort (1024-element sorted blocks, then rank merges)
used to check the rank rule on the CPU, including runs of equal keys (test infrastructure)."""
import numpy as np


def merge_pass(src, n, w):
    """One merge pass over key tuples src[0..n): left-run elements count right-run keys < x,
    right-run elements count left-run keys <= x."""
    dst = [None] * n
    for p in range(n):
        x = src[p]
        s = (p // (2 * w)) * (2 * w)
        left = p < s + w
        lo, hi = (s + w, min(s + 2 * w, n)) if left else (s, s + w)
        out = p if left else p - w
        cnt = sum(1 for m in range(lo, hi) if (src[m] < x if left else src[m] <= x))
        dst[out + cnt] = x
    return dst


def device_sort(keys, block=1024):
    n = len(keys)
    a = []
    for b0 in range(0, n, block):
        a += sorted(keys[b0:b0 + block])
    w = block
    while w < n:
        a = merge_pass(a, n, w)
        w *= 2
    return a
