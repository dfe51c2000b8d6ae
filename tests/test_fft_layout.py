"""LDS layout of the y-direction FFT exchanges (mhd2d_fft.h, fft_lds).

Each exchange between two passes writes positions (t / NS) NS R + t % NS +
r NS from thread t and reads t + b N/E + r N/R2.  The kernel forms one base
word per thread and exchange side and reaches the other words through
compile-time offsets, so the layout must be additive in those offsets; it
must also be injective within the padded array of N + N/E words and free of
LDS bank conflicts (16 lanes of 8-byte words per LDS cycle for the FP64
transform; 32 lanes of 4-byte words for FP32; the lane grouping was
measured with SQ_LDS_BANK_CONFLICT on gfx950).  The formula is
restated here from the header and checked exhaustively for every supported
line length.
"""
import math

import pytest


def fft_lds(E, NS, G, p):
    if NS == 1:
        return p + p // G
    if NS < G:
        return p + (p // (E * NS)) * NS
    return p


def exchanges(N, E):
    M, LE = int(math.log2(N)), int(math.log2(E))
    PE, REM = M // LE, M % LE
    NP = PE + (1 if REM else 0)
    radix = lambda p: E if p < PE else 1 << REM  # noqa: E731
    TPR = N // E
    NS = 1
    for P in range(NP - 1):
        R, R2 = radix(P), radix(P + 1)
        writes = [[(t // NS) * NS * R + t % NS + r * NS for t in range(TPR)] for r in range(R)]
        reads = [[t + b * TPR + r * (N // R2) for t in range(TPR)]
                 for b in range(E // R2) for r in range(R2)]
        yield NS, TPR, writes, reads
        NS *= R


def worst_conflict(words, lanes):
    worst = 1
    for w0 in range(0, len(words), 64):
        wave = words[w0:w0 + 64]
        for g in range(0, len(wave), lanes):
            slots = {}
            for a in set(wave[g:g + lanes]):
                slots[a % lanes] = slots.get(a % lanes, 0) + 1
            worst = max(worst, max(slots.values()))
    return worst


# (elements per thread, LDS lanes per cycle = 128 B / word size, N)
CASES = [(E, G, N) for E in (8, 16) for G in (16, 32)
         for N in (16, 32, 64, 128, 256, 512, 1024, 2048, 4096) if N >= E]


@pytest.mark.parametrize("E,G,N", CASES)
def test_layout_injective_additive_conflict_free(E, G, N):
    for NS, TPR, writes, reads in exchanges(N, E):
        L = [fft_lds(E, NS, G, p) for p in range(N)]
        assert len(set(L)) == N and max(L) < N + N // E
        for group in (writes, reads):
            base = [fft_lds(E, NS, G, p) for p in group[0]]
            for pat in group:
                # the offset the kernel uses is the layout of the t = 0 position
                off = fft_lds(E, NS, G, pat[0])
                assert all(fft_lds(E, NS, G, p) - b == off for p, b in zip(pat, base))
                if TPR >= G:
                    # rows of fewer than one lane group share groups; only
                    # full groups are priced
                    assert worst_conflict([fft_lds(E, NS, G, p) for p in pat], G) == 1


def test_conflict_counter_catches_a_strided_layout():
    # the identity layout leaves the stride-E writes of the first exchange
    # 8-way (FP64) in conflict: the check has teeth
    writes = next(exchanges(2048, 8))[2]
    assert worst_conflict(writes[0], 16) == 8
