"""GPU checks of the roofline probes: the random gather that defines G (the
successor of random_access<512>, src/random_access.cu:53-66), the dependent
pointer chase, the streaming copy and the streaming read.  Their outputs are compared
with numpy, so the denominator of every "% of gather roofline" figure is a verified
kernel."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

M1 = np.uint64(0xFF51AFD7ED558CCD)
M2 = np.uint64(0xC4CEB9FE1A85EC53)


def murmur3_fmix64(x: np.ndarray) -> np.ndarray:
    """MurmurHash3's 64-bit finaliser over a uint64 array (wrapping arithmetic)."""
    x = x.astype(np.uint64, copy=True)
    s = np.uint64(33)
    with np.errstate(over="ignore"):
        x ^= x >> s
        x *= M1
        x ^= x >> s
        x *= M2
        x ^= x >> s
    return x


@pytest.mark.parametrize("items", [1, 2, 4, 8])
@pytest.mark.parametrize("m", [1, 1000, 262147])
def test_gather_reads_hashed_slot(gj, orc, gpu, items, m):
    print(f"[case] gather items={items} m={m}", flush=True)
    rng = np.random.default_rng(items * 31 + m)
    n = 1 << 16
    arr = rng.integers(0, 2 ** 64 - 1, size=n, dtype=np.uint64)
    probes = rng.integers(0, 2 ** 64 - 1, size=m, dtype=np.uint64)
    # the finaliser here is the oracle's, whose constants the oracle tests pin by inversion
    assert [int(v) for v in murmur3_fmix64(probes[:64])] == [orc.murmur3_64(int(v)) for v in probes[:64]]
    d_arr, d_p = gj.DeviceArray.from_numpy(arr), gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(m, np.uint32)
    out.fill(0x5A)
    gj.gather(d_arr, d_p, out, m, 256, items)
    gj.synchronize()
    want = (arr[murmur3_fmix64(probes) & np.uint64(n - 1)] & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    assert np.array_equal(out.numpy(), want)


def test_pointer_chase_follows_chain_and_continues(gj, gpu):
    n = 100_003
    rng = np.random.default_rng(5)
    perm = rng.permutation(n).astype(np.uint32)
    nxt = np.empty(n, dtype=np.uint32)
    nxt[perm] = np.roll(perm, -1)  # one cycle through every slot
    d_next = gj.DeviceArray.from_numpy(nxt)
    chains, steps = 100, 777
    out = gj.DeviceArray(chains, np.uint32)
    gj.pointer_chase(d_next, chains, steps, out)
    gj.synchronize()
    # default start of chain c: (c * 0x9E3779B97F4A7C15 mod 2^64) mod n
    start = np.array([((c * 0x9E3779B97F4A7C15) % (1 << 64)) % n for c in range(chains)], dtype=np.int64)
    cur = start.copy()
    for _ in range(steps):
        cur = nxt[cur].astype(np.int64)
    assert np.array_equal(out.numpy().astype(np.int64), cur)
    # continue every chain where it stopped (start = out)
    gj.pointer_chase(d_next, chains, steps, out, start=out)
    gj.synchronize()
    for _ in range(steps):
        cur = nxt[cur].astype(np.int64)
    assert np.array_equal(out.numpy().astype(np.int64), cur)
    # and a cycle of length n returns to its start after n steps
    one = gj.DeviceArray.from_numpy(np.array([perm[0]], dtype=np.uint32))
    gj.pointer_chase(d_next, 1, n, one, start=one)
    gj.synchronize()
    assert int(one.numpy()[0]) == int(perm[0])


@pytest.mark.parametrize("nbytes", [16, 4096, (1 << 20) + 48, (64 << 20) + 16])
def test_stream_copy_is_exact(gj, gpu, nbytes):
    src = np.random.default_rng(nbytes).integers(0, 256, size=nbytes, dtype=np.uint8)
    d_src = gj.DeviceArray.from_numpy(src)
    d_dst = gj.DeviceArray(nbytes + 64, np.uint8)
    d_dst.fill(0xA5)
    gj.stream_copy(d_src, d_dst, nbytes)
    gj.synchronize()
    got = d_dst.numpy()
    assert np.array_equal(got[:nbytes], src)
    assert np.all(got[nbytes:] == 0xA5)  # nothing written past the end


@pytest.mark.parametrize("nbytes", [16, 4096, (1 << 20) + 48, (64 << 20) + 16])
def test_stream_read_xor(gj, gpu, nbytes):
    # every 8-byte word XORed into the caller's value, whatever the grid and the unrolled
    # main loop / remainder split
    words = np.random.default_rng(nbytes + 1).integers(0, 2 ** 64 - 1, size=nbytes // 8, dtype=np.uint64)
    d_src = gj.DeviceArray.from_numpy(words)
    seed = np.uint64(0x0123456789ABCDEF)
    out = gj.DeviceArray.from_numpy(np.array([seed], dtype=np.uint64))
    gj.stream_read(d_src, nbytes, out)
    gj.synchronize()
    assert int(out.numpy()[0]) == int(np.bitwise_xor.reduce(words) ^ seed)
