"""The LDS FFT (glass_amd/csrc/fft_lds.h, lds_fft) orders its radix-4
Stockham stages with wavefront-scope fences instead of workgroup barriers
when N = 4^a, N/4 <= 64 and the workgroup is whole waves.  That is sound
only if every element a wave reads in a stage was written, in the previous
stage, by that same wave, and no other wave touches the elements a wave
reads and then overwrites within a stage.  This checks both by enumerating
the stages' butterfly -> thread assignment (butterfly b = tid + p*NTHR of
series b / (N/4), reading j + m*N/4 and writing (j/Ns)*Ns*4 + j%Ns + m*Ns)
for every size and workgroup shape the condition admits, and that the
checker does see a cross-wave dependency where the condition rejects one.
"""
import itertools

import pytest


def stage_access(N, NL, NTHR, Ns):
    """(wave, series, element) of every read and write of one radix-4 stage."""
    R = 4
    nb = NL * (N // R)
    per = (nb + NTHR - 1) // NTHR
    reads, writes = [], []
    for tid in range(NTHR):
        for p in range(per):
            b = tid + p * NTHR
            if b >= nb:
                continue
            l, j = divmod(b, N // R)
            k = j % Ns
            base = (j // Ns) * Ns * R + k
            for m in range(R):
                reads.append((tid // 64, l, j + m * (N // R)))
                writes.append((tid // 64, l, base + m * Ns))
    return reads, writes


def wave_local(N, NL, NTHR):
    """True when each wave only reads elements it wrote in the stage before
    and each element is read and written by one wave within a stage."""
    stages = []
    Ns = 1
    while Ns < N:
        stages.append(stage_access(N, NL, NTHR, Ns))
        Ns *= 4
    for reads, writes in stages:
        owner = {}
        for w, l, e in reads + writes:
            if owner.setdefault((l, e), w) != w:
                return False
    for (_, w_prev), (r_next, _) in zip(stages, stages[1:]):
        writer = {(l, e): w for w, l, e in w_prev}
        if any(writer[(l, e)] != w for w, l, e in r_next):
            return False
    return True


def condition(N, NTHR):
    """The compile-time condition in lds_fft (radix-4 stages only)."""
    log2 = N.bit_length() - 1
    return log2 % 2 == 0 and N // 4 <= 64 and NTHR % 64 == 0


@pytest.mark.parametrize("N,NL,NTHR", [c for c in itertools.product([4, 16, 64, 256], [1, 2, 3, 6, 12],
                                                                      [64, 128, 256, 512, 1024])
                                       if condition(c[0], c[2])])
def test_admitted_shapes_are_wave_local(N, NL, NTHR):
    assert wave_local(N, NL, NTHR)


def test_checker_sees_cross_wave_stages():
    # 1024 points: N/4 = 256 butterflies per series span four waves
    assert not condition(1024, 512)
    assert not wave_local(1024, 6, 512)
