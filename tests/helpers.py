"""Shared helpers for the parity tests (checker side)."""
import numpy as np

from oracle import oracle as O


def lcg_stream(seed=0xDEADBEEFCAFEF00D):
    """The LCG of src/fm_index/smem.rs:807-813."""
    state = seed
    while True:
        state = (state * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        yield state >> 33


def smem_random_cases(iters=80):
    """Inputs of smems_equal_containment_filtered_mems_randomized (smem.rs:805-870)."""
    nxt = lcg_stream()
    bases = "ACGT"
    iupac = "ACGTNRY"
    out = []
    for it in range(iters):
        full = "".join(bases[next(nxt) % 4] for _ in range(200))
        a, b = full[10:60], full[40:160]
        if it % 2 == 0:
            noise = "".join(iupac[next(nxt) % 7] for _ in range(80))
        else:
            noise = "".join(bases[next(nxt) % 4] for _ in range(80))
        out.append(([a, b, noise], ["A", "B", "NOISE"], full))
    return out


def containment_max(spans):
    s = set(spans)
    return sorted(x for x in s if not any(y != x and y[0] <= x[0] and x[1] <= y[1] for y in s))


def mem_model_min_len_skip(occurs, n, min_len):
    """CPU model of the device MEM algorithm (k_mem_wave): returns spans.
    occurs(s, e) -> True iff query[s:e] occurs in the text."""
    out = []
    if n == 0 or min_len == 0:
        return out
    s0, E = 0, 0
    while True:
        j = max(s0 + min_len, E + 1)
        if j > n:
            break
        k = j
        while k > 0 and occurs(k - 1, j):
            k -= 1
        if k > j - min_len:
            s0 = k
            continue
        e = j
        while e < n and occurs(k, e + 1):
            e += 1
        out.append((k, e))
        s0, E = k + 1, e
    return out


def brute_occurs_fn(text, q, tag=0):
    cache = {}

    def occurs(s, e):
        key = (s, e)
        if key not in cache:
            cache[key] = len(O.naive_locate(text, q[s:e], tag)) > 0
        return cache[key]
    return occurs


def random_text(rng, n, alphabet="ACGT", iupac_frac=0.0):
    codes = [O.SYMBOLS.index(c) for c in alphabet]
    t = rng.choice(codes, size=n).astype(np.uint8)
    if iupac_frac > 0:
        mask = rng.random(n) < iupac_frac
        t[mask] = rng.choice(O.IUPAC_AMBIG, size=int(mask.sum())).astype(np.uint8)
    return t
