"""The circulant lengths of the FFT convolutions: L = 2n - 2 is enough.

fft.hip embeds each kernel in a circulant of length L >= 2n - 2 per axis
(`fft_embed_min`), where the reference uses 2n (compute_FFT_mutual_coupling_
mats.jl:130-155).  Along an axis convolved in full the kernel is even, so
the offsets n-1 and -(n-1) share index n-1 with the same value; a half-cell
reduction (k_embed_poly) with parity offset delta = +-1 spans only 2n - 2
offsets, and index L/2 takes the one inside the row (poly_offset).  These
tests restate both index rules in 1-D and compare the circular convolution
with the direct linear sum over every parity pair, so a wrong choice at
index L/2 or a length one too short fails here."""
import numpy as np


def poly_offset(i, L, nb, half, delta, nf):
    """fft.hip poly_offset: the row offset stored at circulant index i."""
    if i <= nb - 1:
        f = abs(2 * i + delta) if half else i
        if f < nf:
            return f
    if L - i <= nb - 1:
        m = i - L
        f = abs(2 * m + delta) if half else -m
        if f < nf:
            return f
    return -1


def even_offset(i, L, n):
    """fft.hip k_embed: |offset| at index i, -1 outside the row."""
    return i if i < n else (L - i if L - i < n else -1)


def _circ(c, x, nout):
    L = len(c)
    X = np.zeros(L, complex)
    X[:len(x)] = x
    return np.fft.ifft(np.fft.fft(X) * np.fft.fft(c))[:nout]


def test_even_kernel_at_2n_minus_2():
    rng = np.random.default_rng(3)
    for n in range(2, 40):
        row = rng.standard_normal(n)
        for L in (2 * n - 2, 2 * n - 1, 2 * n):
            c = np.array([row[even_offset(i, L, n)] if even_offset(i, L, n) >= 0 else 0.0 for i in range(L)])
            x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
            y = _circ(c, x, n)
            yd = np.array([sum(row[abs(i - j)] * x[j] for j in range(n)) for i in range(n)])
            assert np.abs(y - yd).max() <= 1e-12 * np.abs(yd).max(), (n, L)
        # one shorter wraps the extreme offsets onto the outputs
        L = 2 * n - 3
        if n >= 3:
            c = np.array([row[even_offset(i, L, n)] if even_offset(i, L, n) >= 0 else 0.0 for i in range(L)])
            x = rng.standard_normal(n) + 0j
            yd = np.array([sum(row[abs(i - j)] * x[j] for j in range(n)) for i in range(n)])
            assert np.abs(_circ(c, x, n) - yd).max() > 1e-6 * np.abs(yd).max()


def test_half_cell_reduction_at_2nb_minus_2():
    """Fine positions 0 .. nf-1 (nf = 2 nb - 1); parity 0 has nb of them,
    parity 1 nb - 1; a block from parity pin to pout = pin + delta uses the
    fine offsets |2 (i - j) + delta|."""
    rng = np.random.default_rng(4)
    for nb in range(2, 30):
        nf = 2 * nb - 1
        row = rng.standard_normal(nf)
        L = 2 * nb - 2
        for delta, pin in ((-1, 1), (0, 0), (0, 1), (1, 0)):
            pout = pin + delta
            nin = nb if pin == 0 else nb - 1
            nout = nb if pout == 0 else nb - 1
            c = np.array([row[poly_offset(i, L, nb, 1, delta, nf)] if poly_offset(i, L, nb, 1, delta, nf) >= 0
                          else 0.0 for i in range(L)])
            x = rng.standard_normal(nin) + 1j * rng.standard_normal(nin)
            y = _circ(c, x, nout)
            yd = np.array([sum(row[abs(2 * (i - j) + delta)] * x[j] for j in range(nin)) for i in range(nout)])
            assert np.abs(y - yd).max() <= 1e-12 * np.abs(yd).max(), (nb, delta, pin)
            if delta == 1 and nb >= 3:
                # the rule before the 2n - 2 lengths (offset i at every
                # i <= nb - 1) leaves index L/2 empty here and misses the
                # offset -(nb-1) the block uses
                c_old = c.copy()
                c_old[L // 2] = 0.0
                assert np.abs(_circ(c_old, x, nout) - yd).max() > 1e-6 * np.abs(yd).max()
