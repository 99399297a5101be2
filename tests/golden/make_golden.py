"""Generate the golden vectors in tests/golden/ntt_golden.npz.

Independent of the C oracle and of the HIP library: everything below is
textbook modular arithmetic on Python integers.

* psi: the smallest primitive 2N-th root of unity mod q.  For small q it is
  found by enumerating every x in [1, q) with x^N == -1; for large q as the
  minimum of the odd powers of g = r^((q-1)/2N), r the first quadratic
  non-residue (the odd powers of one primitive root are exactly the roots of
  X^N + 1).  This is the definition the reference's table constructor
  implements (primus_ntt/src/root.rs:89-111, primus_ntt/src/ntt/primitive.rs:118).
* forward NTT output slot k is the evaluation of a(X) at psi^(2*rev(k)+1)
  (bit-reversed order, primus_ntt/src/ntt/primitive.rs:319-328), computed by
  Horner's rule in O(N^2).
* inverse NTT: a_j = N^-1 * sum_k A_k * psi^-(2*rev(k)+1)*j, O(N^2).
* schoolbook negacyclic product a*b mod (X^N+1, q) (primus_poly/src/poly/mul.rs:111-142).
* monomial transforms from their closed form.

Run:  python tests/golden/make_golden.py   (about a minute)
"""
import os
import random

import numpy as np

# (label, q, log_n, heavy) -- heavy cases skip the O(N^2) inverse/product.
CASES = [
    ("tiny_q17", 17, 3, False),
    ("q12289_n512", 12289, 9, False),
    ("root_rs_q132120577_n1024", 132120577, 10, False),        # primus_ntt/tests/root.rs:5-6
    ("hexl32_q536813569_n2048", 536813569, 11, True),          # primus_ntt/tests/hexl.rs:20
    ("hexl52_q562949953392641_n2048", 562949953392641, 11, True),      # hexl.rs:58
    ("hexl64_q1152921504606830593_n2048", 1152921504606830593, 11, True),  # hexl.rs:96
    ("plain_q2147473409_n1024", 2147473409, 10, False),
    ("rns50_q1125899906826241_n4096", 1125899906826241, 12, True),
    ("extprod_q4294957057_n1024", 4294957057, 10, False),
    ("rlwe59_q576460752303421441_n1024", 576460752303421441, 10, False),
]


def rev(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2) if bits else 0


def is_prime(n):
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
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


def minimal_psi(q, n):
    assert (q - 1) % (2 * n) == 0 and is_prime(q)
    if q < 200000:
        roots = [x for x in range(1, q) if pow(x, n, q) == q - 1]
        assert len(roots) == n
        return min(roots)
    r = 2
    while pow(r, (q - 1) // 2, q) != q - 1:
        r += 1
    g = pow(r, (q - 1) // (2 * n), q)
    g2 = g * g % q
    best, cur = g, g
    for _ in range(n):
        best = min(best, cur)
        cur = cur * g2 % q
    assert pow(best, n, q) == q - 1
    return best


def forward(a, q, psi, log_n):
    n = 1 << log_n
    out = []
    for k in range(n):
        x = pow(psi, 2 * rev(k, log_n) + 1, q)
        acc = 0
        for c in reversed(a):
            acc = (acc * x + c) % q
        out.append(acc)
    return out


def inverse(A, q, psi, log_n):
    n = 1 << log_n
    ninv = pow(n, -1, q)
    ys = [pow(psi, -(2 * rev(k, log_n) + 1), q) for k in range(n)]
    pw = [1] * n
    out = []
    for _j in range(n):
        s = 0
        for k in range(n):
            s += A[k] * pw[k]
        out.append(s * ninv % q)
        pw = [pw[k] * ys[k] % q for k in range(n)]
    return out


def negacyclic(a, b, q):
    n = len(a)
    c = [0] * n
    for i in range(n):
        ai = a[i]
        if ai == 0:
            continue
        for j in range(n):
            k = i + j
            if k < n:
                c[k] += ai * b[j]
            else:
                c[k - n] -= ai * b[j]
    return [x % q for x in c]


def monomial(coeff, degree, q, psi, log_n):
    n = 1 << log_n
    return [coeff * pow(psi, ((2 * rev(k, log_n) + 1) * degree) % (2 * n), q) % q for k in range(n)]


def main():
    out = {}
    for idx, (label, q, log_n, heavy) in enumerate(CASES):
        n = 1 << log_n
        rng = random.Random(0x5052494D5553 + idx)
        psi = minimal_psi(q, n)
        a = [rng.randrange(q) for _ in range(n)]
        A = forward(a, q, psi, log_n)
        out[label + "/q"] = np.array([q], dtype=np.uint64)
        out[label + "/log_n"] = np.array([log_n], dtype=np.uint64)
        out[label + "/psi"] = np.array([psi], dtype=np.uint64)
        out[label + "/a"] = np.array(a, dtype=np.uint64)
        out[label + "/fwd_a"] = np.array(A, dtype=np.uint64)
        deg = rng.randrange(n)
        coeff = rng.randrange(2, q - 1)
        out[label + "/mono_params"] = np.array([coeff, deg], dtype=np.uint64)
        out[label + "/mono"] = np.array(monomial(coeff, deg, q, psi, log_n), dtype=np.uint64)
        out[label + "/mono_one"] = np.array(monomial(1, deg, q, psi, log_n), dtype=np.uint64)
        out[label + "/mono_minus_one"] = np.array(monomial(q - 1, deg, q, psi, log_n), dtype=np.uint64)
        if not heavy:
            B = [rng.randrange(q) for _ in range(n)]
            out[label + "/inv_in"] = np.array(B, dtype=np.uint64)
            out[label + "/inv_out"] = np.array(inverse(B, q, psi, log_n), dtype=np.uint64)
            b = [rng.randrange(q) for _ in range(n)]
            out[label + "/b"] = np.array(b, dtype=np.uint64)
            out[label + "/prod"] = np.array(negacyclic(a, b, q), dtype=np.uint64)
        print(label, "psi =", psi, flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ntt_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
