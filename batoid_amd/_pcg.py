"""Host bookkeeping for GPU draws from a numpy PCG64 stream.

numpy's ``Generator(PCG64)`` (what ``np.random.default_rng`` returns) is a
128-bit LCG, state <- state * M + inc mod 2^128, whose 64-bit outputs are
the XSL-RR permutation of each new state; ``uniform(lo, hi)`` returns
lo + (hi - lo) * ((next64 >> 11) * 2^-53).  The GPU generator computes draw
k of the stream from the state before the first draw by jump-ahead, so the
host only has to hand over that state and, afterwards, move the numpy
generator past the draws the GPU consumed (as if numpy had drawn them).
"""
import numpy as np

MULT = (2549297995355413924 << 64) + 4865540595714422341
_MASK = (1 << 128) - 1
_M64 = (1 << 64) - 1


def stream_words(rng):
    """((state_hi, state_lo), (inc_hi, inc_lo)) of a PCG64 generator, or None
    for any other bit generator."""
    bg = rng.bit_generator
    if not isinstance(bg, np.random.PCG64):
        return None
    st = bg.state["state"]
    s, inc = int(st["state"]), int(st["inc"])
    return (s >> 64, s & _M64), (inc >> 64, inc & _M64)


def jump(state, inc, k):
    """State after k LCG steps (Brown 1994 arbitrary-stride jump)."""
    acc_mult, acc_plus = 1, 0
    cur_mult, cur_plus = MULT, inc
    while k > 0:
        if k & 1:
            acc_mult = (acc_mult * cur_mult) & _MASK
            acc_plus = (acc_plus * cur_mult + cur_plus) & _MASK
        cur_plus = ((cur_mult + 1) * cur_plus) & _MASK
        cur_mult = (cur_mult * cur_mult) & _MASK
        k >>= 1
    return (acc_mult * state + acc_plus) & _MASK


def consume(rng, k):
    """Advance a PCG64 generator by k 64-bit draws, keeping its buffered
    32-bit half-word (numpy's own ``advance`` clears it)."""
    bg = rng.bit_generator
    st = bg.state
    st["state"]["state"] = jump(int(st["state"]["state"]), int(st["state"]["inc"]), int(k))
    bg.state = st
