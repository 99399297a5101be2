"""Sampling of RLWE encryption inputs from Philox4x32-10.

CPU: the oracle's generator against the published Random123 known answers and
against rocRAND's engine (a host-only build of tests/host/philox_kat.hip), and
the oracle's samplers against their distributions.  GPU: the device sampler
(ph_rlwe_sample_device) against the oracle, bit for bit, on an uneven batch,
and the sampled batch encrypted and checked against the oracle's encryption."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))

# Random123's kat_vectors for philox4x32_10: (counter, key) -> output
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


def test_philox_known_answers():
    for ctr, key, want in KAT:
        assert tuple(int(v) for v in O.philox4x32_10(ctr, key)) == want


def test_philox_against_rocrand_engine(tmp_path):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    if not os.path.exists(hipcc) or not os.path.exists("/opt/rocm/include/rocrand/rocrand_philox4x32_10.h"):
        pytest.skip("hipcc or the rocRAND headers not found")
    exe = str(tmp_path / "philox_kat")
    b = subprocess.run([hipcc, "--offload-host-only", "-std=c++17", "-O1", "-o", exe,
                        os.path.join(HERE, "host", "philox_kat.hip")], capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-2000:]
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60).stdout.strip().splitlines()
    assert len(out) == 5
    for line in out:
        lhs, rhs = line.split("->")
        words = [int(w, 16) for w in lhs.split()]
        want = [int(w, 16) for w in rhs.split()]
        assert [int(v) for v in O.philox4x32_10(words[:4], words[4:])] == want


def test_gauss_thresholds():
    t = O.gauss_thresholds().astype(object)
    assert all(t[i] < t[i + 1] for i in range(37))
    # P(e = 0) = 1 / sum_k exp(-k^2 / (2 sigma^2)), about 1 / (sigma sqrt(2 pi))
    p0 = (int(t[19]) - int(t[18])) / 2.0 ** 64
    z = sum(np.exp(-k * k / (2 * 3.2 ** 2)) for k in range(-19, 20))
    assert abs(p0 - 1 / z) < 1e-15
    assert abs(p0 - 1 / (3.2 * np.sqrt(2 * np.pi))) < 1e-6
    # symmetric: P(e <= -k-1) == P(e >= k+1)
    for k in range(19):
        left = int(t[18 - k])
        right = 2 ** 64 - int(t[19 + k])
        assert abs(left - right) < 2 ** 20


def test_sampler_distributions():
    q = 576460752303421441
    s, e, m, a = O.rlwe_sample(0x5052494D5553, 7, 16, [q])
    n = s.size
    for v in (-1, 0, 1):
        assert abs(np.mean(s == v) - 1 / 3) < 0.01
    assert np.abs(e).max() <= 19
    assert abs(e.mean()) < 0.05
    assert abs(e.astype(np.float64).var() - 3.2 ** 2) < 0.3
    assert abs(m.astype(np.float64).mean() - 127.5) < 1.5
    assert np.bincount(m, minlength=256).min() > n / 256 * 0.7
    assert (a < q).all()
    assert abs(a.astype(np.float64).mean() / q - 0.5) < 0.01
    # a different request, purpose or seed gives unrelated values
    s2, e2, _, a2 = O.rlwe_sample(0x5052494D5553, 8, 16, [q])
    assert np.mean(s2 == s) < 0.4 and not np.array_equal(a2, a)
    s3, _, _, _ = O.rlwe_sample(0x5052494D5554, 7, 16, [q])
    assert np.mean(s3 == s) < 0.4


def _torch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.mark.gpu
def test_device_sampling_matches_oracle():
    _torch()
    import primus_amd as pa
    from primus_amd.fhe import RlweContext, rlwe_layout

    rng = np.random.default_rng(31)
    shapes = [(int(rng.integers(10, 17)), int(rng.integers(1, 5))) for _ in range(12)] + [(16, 4), (10, 1)]
    grid = {ln: pa.ntt_primes_below(59, 2 << ln, 4) for ln in range(10, 17)}
    ctx = RlweContext(10, 16, 4, grid)
    reqs, a_words, small = rlwe_layout(shapes)
    seed = 0x0123456789ABCDEF
    da, dout = pa.DeviceBuffer(a_words), pa.DeviceBuffer(a_words)
    ds, de, dm = (pa.DeviceBuffer((small + 7) // 8) for _ in range(3))
    ctx.sample_device(reqs, seed, da.ptr, ds.ptr, de.ptr, dm.ptr)
    ctx.encrypt_device(reqs, da.ptr, ds.ptr, de.ptr, dm.ptr, dout.ptr)
    pa.synchronize()
    a = da.download((a_words,))
    out = dout.download((a_words,))
    s = ds.download(((small + 7) // 8,)).view(np.int8)[:small]
    e = de.download(((small + 7) // 8,)).view(np.int8)[:small]
    m = dm.download(((small + 7) // 8,)).view(np.uint8)[:small]
    for r, (ln, lb, ao, so) in enumerate(reqs):
        n = 1 << ln
        mods = grid[ln][:lb]
        ws, we, wm, wa = O.rlwe_sample(seed, r, ln, mods)
        np.testing.assert_array_equal(s[so:so + n], ws, err_msg="s, request %d" % r)
        np.testing.assert_array_equal(e[so:so + n], we, err_msg="e, request %d" % r)
        np.testing.assert_array_equal(m[so:so + n], wm, err_msg="m, request %d" % r)
        np.testing.assert_array_equal(a[ao:ao + lb * n].reshape(lb, n), wa, err_msg="a_hat, request %d" % r)
        if r in (0, len(reqs) - 2, len(reqs) - 1):
            for j, q in enumerate(mods):
                want = O.rlwe_encrypt(O.Table(ln, q), ws, we, wm, O.rlwe_delta(mods, q), wa[j])
                np.testing.assert_array_equal(out[ao + j * n:ao + (j + 1) * n], want,
                                              err_msg="b_hat, request %d limb %d" % (r, j))


@pytest.mark.gpu
def test_plan_sampling_matches_one_shot():
    _torch()
    import primus_amd as pa
    from primus_amd.fhe import RlweContext, rlwe_layout

    shapes = [(10, 3), (13, 2), (12, 1), (16, 2), (11, 4)]
    grid = {ln: pa.ntt_primes_below(59, 2 << ln, 4) for ln in range(10, 17)}
    ctx = RlweContext(10, 16, 4, grid)
    reqs, a_words, small = rlwe_layout(shapes)
    sw = (small + 7) // 8
    one = [pa.DeviceBuffer(a_words)] + [pa.DeviceBuffer(sw) for _ in range(3)]
    two = [pa.DeviceBuffer(a_words)] + [pa.DeviceBuffer(sw) for _ in range(3)]
    ctx.sample_device(reqs, 99, *[b.ptr for b in one])
    ctx.plan(reqs).sample_device(99, *[b.ptr for b in two])
    pa.synchronize()
    for x, y in zip(one, two):
        np.testing.assert_array_equal(x.download(), y.download())
