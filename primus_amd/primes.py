"""NTT-friendly prime search used to set up workloads.

The largest primes q < 2^bits with q = 1 (mod 2N), by a deterministic
Miller-Rabin test (the fixed witness set below is exact for all n < 3.3e24).
"""

_WITNESSES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41)


def is_prime(n):
    if n < 2:
        return False
    for p in _WITNESSES:
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in _WITNESSES:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def ntt_primes_below(bits, two_n, count):
    """The `count` largest primes q < 2^bits with q = 1 (mod two_n), descending."""
    out = []
    q = ((1 << bits) - 1) // two_n * two_n + 1
    if q >= (1 << bits):
        q -= two_n
    while len(out) < count and q > two_n:
        if is_prime(q):
            out.append(q)
        q -= two_n
    if len(out) < count:
        raise ValueError("not enough NTT primes below 2^%d for 2N=%d" % (bits, two_n))
    return out
