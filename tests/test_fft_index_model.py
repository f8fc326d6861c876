"""Host model of the index arithmetic of the FFT kernels (poisson.hip) and of
the block orders (common.hpp), checked on the CPU for every plan the library
accepts (nx-2, nz-2 even, factors 2, 3, 5, at most 2048).

What is modelled, with the kernels' own formulas:
  * `Quo16`: the quotient by one high multiply, exhaustively against integer
    division for every divisor 1..2048 and numerator 0..2^16-1, and the
    numerators the passes feed it stay below 2^16;
  * the Stockham passes (`fft_pass`, `fft_first_pass` with either lane order,
    `fft_last_pass`): every lane's LDS and element indices, including the
    lanes past the last butterfly, stay inside the transform and the batch;
    each pass reads and writes every element exactly once; and the passes run
    in numpy with those index maps reproduce numpy.fft for both signs;
  * the inverse x transform's spectrum reads (mode = n or N-n, at most N/2)
    and the pair table's spectrum offsets stay inside the half spectrum;
  * `xcd_linear` is a permutation of the block indices.
"""
import numpy as np
import pytest

FFT_THREADS = 256
FFT_MAXNB = 2048


def plans():
    """Every even n <= 2048 that factors into 2, 3 and 5 (nx-2 and nz-2)."""
    out = []
    for n in range(2, 2049, 2):
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            out.append(n)
    return out


def factor_plan(n):
    """chan_create's factor_plan: radix 4 first, then 2, 3, 5."""
    r, m = [], n
    while m > 1:
        q = 4 if m % 4 == 0 else 2 if m % 2 == 0 else 3 if m % 3 == 0 else 5
        r.append(q)
        m //= q
    assert len(r) <= 12
    return r


def batch(n):
    """Transforms per block (pb.bx, pb.bz)."""
    return max(1, min(16, 2048 // n))


def quo16_m(dv):
    return (0xFFFFFFFF // dv + 1) if dv > 1 else 0


def quo(q, dv):
    q = np.asarray(q, dtype=np.uint64)
    if dv <= 1:
        return q.astype(np.int64)
    return ((q * np.uint64(quo16_m(dv))) >> np.uint64(32)).astype(np.int64)


def test_quo16_exhaustive():
    q = np.arange(1 << 16, dtype=np.uint64)
    for dv in range(1, 2049):
        assert np.array_equal(quo(q, dv), (q // np.uint64(dv)).astype(np.int64)), dv


def mb(R):
    return (FFT_MAXNB // R + FFT_THREADS - 1) // FFT_THREADS


def twiddles(N):
    t = np.arange(N)
    return np.exp(-2j * np.pi * t / N)


def dft(v, R, N, sign, tw):
    """The kernels' R-point butterflies (v: (R, lanes))."""
    def muli(a):
        return -1j * a if sign < 0 else 1j * a
    if R == 2:
        return np.stack([v[0] + v[1], v[0] - v[1]])
    if R == 4:
        s02, d02 = v[0] + v[2], v[0] - v[2]
        s13, d13 = v[1] + v[3], muli(v[1] - v[3])
        return np.stack([s02 + s13, d02 + d13, s02 - s13, d02 - d13])
    if R == 3:
        s12, d12 = v[1] + v[2], v[1] - v[2]
        t1 = v[0] - 0.5 * s12
        t2 = muli(0.86602540378443864676 * d12)
        return np.stack([v[0] + s12, t1 + t2, t1 - t2])
    st = N // R
    o = []
    for qq in range(R):
        acc = v[0].copy()
        for r in range(1, R):
            w = tw[((r * qq) % R) * st]
            acc = acc + v[r] * (w if sign < 0 else np.conj(w))
        o.append(acc)
    return np.stack(o)


def lanes(R):
    return np.arange(mb(R) * FFT_THREADS)


def fft_pass(a, N, B, ld, Ns, R, sign, tw):
    """fft_pass<R>: in place over a (B*ld), returns the new array."""
    NR, nbf, tws = N // R, N // R * B, N // (Ns * R)
    q = lanes(R)
    assert q.max() < (1 << 16)
    b = quo(q, NR)
    jj = q - b * NR
    ok = q < nbf
    b, jj = b[ok], jj[ok]
    src = np.stack([b * ld + jj + r * NR for r in range(R)])
    assert src.min() >= 0 and np.all((src % ld) < N) and np.all(src // ld < B)
    assert len(np.unique(src)) == B * N
    v = a[src]
    js = quo(jj, Ns)
    k = jj - js * Ns
    assert np.all(k >= 0) and np.all(k < Ns)
    if Ns > 1:
        for r in range(1, R):
            w = tw[r * k * tws]
            v[r] = v[r] * (w if sign < 0 else np.conj(w))
    v = dft(v, R, N, sign, tw)
    dst = np.stack([b * ld + js * Ns * R + k + r * Ns for r in range(R)])
    assert np.all((dst % ld) < N) and len(np.unique(dst)) == B * N
    out = a.copy()
    out[dst] = v
    return out


def first_pass(x, N, B, ld, R, sign, tw, bfast):
    """fft_first_pass<R, SIGN, BFAST>: element (b, n) of x (B, N) -> LDS."""
    NR, nbf = N // R, N // R * B
    q = lanes(R)
    assert q.max() < (1 << 16)
    qc = np.where(q >= nbf, 0, q)   # lanes past the last butterfly take butterfly 0
    if bfast:
        jj = quo(qc, B)
        b = qc - jj * B
    else:
        b = quo(qc, NR)
        jj = qc - b * NR
    n = np.stack([jj + r * NR for r in range(R)])
    # every lane's element index, the spare lanes' included, is inside the batch
    assert n.min() >= 0 and n.max() < N and b.min() >= 0 and b.max() < B
    ok = q < nbf
    v = dft(x[b[ok][None, :], n[:, ok]], R, N, sign, tw)
    dst = np.stack([b[ok] * ld + jj[ok] * R + r for r in range(R)])
    assert np.all((dst % ld) < N) and len(np.unique(dst)) == B * N
    a = np.zeros(B * ld, complex)
    a[dst] = v
    return a, n


def last_pass(a, N, B, ld, R, sign, tw):
    """fft_last_pass<R, SIGN>: LDS -> out (B, N)."""
    NR, nbf = N // R, N // R * B
    q = lanes(R)
    b = quo(q, NR)
    k = q - b * NR
    ok = q < nbf
    b, k = b[ok], k[ok]
    src = np.stack([b * ld + k + r * NR for r in range(R)])
    assert np.all((src % ld) < N) and len(np.unique(src)) == B * N
    v = a[src]
    for r in range(1, R):
        w = tw[r * k]
        v[r] = v[r] * (w if sign < 0 else np.conj(w))
    v = dft(v, R, N, sign, tw)
    out = np.zeros((B, N), complex)
    for r in range(R):
        n = k + r * NR
        assert n.max() < N
        out[b, n] = v[r]
    return out


def run_block_fft(x, sign, bfast):
    """The kernels' pass sequence (first pass from registers, middle passes
    in LDS, last pass to registers; a single pass through LDS)."""
    B, N = x.shape
    ld, tw, rs = N + 1, twiddles(N), factor_plan(N)
    if len(rs) == 1:
        a = np.zeros(B * ld, complex)
        for b in range(B):
            a[b * ld:b * ld + N] = x[b]
        a = fft_pass(a, N, B, ld, 1, rs[0], sign, tw)
        return np.stack([a[b * ld:b * ld + N] for b in range(B)])
    a, _ = first_pass(x, N, B, ld, rs[0], sign, tw, bfast)
    Ns = rs[0]
    for R in rs[1:-1]:
        a = fft_pass(a, N, B, ld, Ns, R, sign, tw)
        Ns *= R
    assert Ns * rs[-1] == N
    return last_pass(a, N, B, ld, rs[-1], sign, tw)


@pytest.mark.parametrize("N", plans())
def test_passes_reproduce_numpy_fft(N):
    B = batch(N)
    assert (N + (N + 1) * B) * 16 <= 160 * 1024   # fft_lds fits the LDS of a CU
    rng = np.random.default_rng(N)
    x = rng.standard_normal((B, N)) + 1j * rng.standard_normal((B, N))
    for bfast in (False, True):
        got = run_block_fft(x, -1, bfast)
        assert np.allclose(got, np.fft.fft(x, axis=1), rtol=0, atol=1e-9 * np.sqrt(N))
        got = run_block_fft(x, +1, bfast)
        assert np.allclose(got, np.fft.ifft(x, axis=1) * N, rtol=0, atol=1e-9 * np.sqrt(N))


@pytest.mark.parametrize("N", plans())
def test_inverse_x_spectrum_reads_in_bounds(N):
    """k_fftx_inv's elem: mode = n (n <= N/2) or N-n, read at spec[b] +
    mode*nzp (+1); with the spare lanes' n of the first pass."""
    B = batch(N)
    rs = factor_plan(N)
    if len(rs) < 2:
        return
    _, n = first_pass(np.zeros((B, N), complex), N, B, N + 1, rs[0], 1, twiddles(N), True)
    half = N // 2
    mode = np.where(n <= half, n, N - n)
    assert mode.min() >= 0 and mode.max() <= half   # nkx = N/2 + 1 modes


@pytest.mark.parametrize("faces", [(34, 33, 34), (194, 97, 194), (386, 193, 386), (66, 49, 66), (98, 65, 98),
                                   (10, 4, 10), (22, 13, 32), (12, 98, 10), (2050, 5, 6)])
def test_pair_tables_inside_the_half_spectrum(faces):
    """fill_pairs: the clamped pair of every table slot, its spectrum offset
    and the farthest element an x transform touches stay inside ph."""
    nx, ny, nz = faces
    nxp, nzp = nx - 2, nz - 2
    nkx = nxp // 2 + 1
    B = batch(nxp)
    npairs = (ny - 1) * (nzp // 2)
    nblocks = -(-npairs // B)
    size = (ny - 1) * nkx * nzp
    pair0 = np.arange(nblocks)[:, None] * B
    pr = np.minimum(pair0 + np.arange(B)[None, :], npairs - 1)
    jr = pr // (nzp // 2)
    kz = 2 * (pr % (nzp // 2))
    spec = jr * nkx * nzp + kz
    assert spec.min() >= 0
    assert (spec + (nkx - 1) * nzp + 1).max() < size
    # the valid pairs cover every (row, z-row pair) once
    valid = (pair0 + np.arange(B)[None, :]) < npairs
    assert len(np.unique(pr[valid])) == npairs


def xcd_linear(n):
    h = np.arange(n)
    q, r, x = n // 8, n % 8, h % 8
    return x * q + np.minimum(x, r) + h // 8


def test_xcd_order_is_a_permutation():
    for n in list(range(1, 600)) + [4096, 18624, 100001]:
        assert np.array_equal(np.sort(xcd_linear(n)), np.arange(n)), n
