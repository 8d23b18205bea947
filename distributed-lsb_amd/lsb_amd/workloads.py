"""Workload definitions for the batched and serving benchmarks.

Sizes and traces are drawn from pcg64 (the generator of the reference's
keygen, mpi/mpi_lsbsort.cpp:686-690), restated here with Python integers so
a trace is the same on every machine.  Device data for each job comes from
lsb_generate.
"""
import math

_MUL = (2549297995355413924 << 64) | 4865540595714422341
_INC = (6364136223846793005 << 64) | 1442695040888963407
_M128 = (1 << 128) - 1


def pcg64_stream(seed, count):
    """Outputs 0..count-1 of pcg64(seed) (pcg-cpp seeding and XSL-RR output)."""
    s = ((seed + _INC) * _MUL + _INC) & _M128
    out = []
    for _ in range(count):
        s = (s * _MUL + _INC) & _M128
        x = ((s >> 64) ^ s) & ((1 << 64) - 1)
        r = s >> 122
        out.append(((x >> r) | (x << ((64 - r) & 63))) & ((1 << 64) - 1))
    return out


def uniform01(seed, count):
    """53-bit uniforms in [0, 1) from pcg64(seed)."""
    return [(x >> 11) * (1.0 / 9007199254740992.0) for x in pcg64_stream(seed, count)]


def batched_sizes(nseg=4096, seed=2024, lo_log2=10, span_log2=12):
    """Segment sizes floor(2^(lo + span*u)): log-uniform over
    [2^lo, 2^(lo+span)]."""
    return [int(math.floor(2.0 ** (lo_log2 + span_log2 * u))) for u in uniform01(seed, nseg)]


def serving_trace(jobs, seed=77, lo_log2=12, span_log2=16):
    """Job list for the serving stream: size log-uniform over
    [2^lo, 2^(lo+span)], key width 32 or 64 (50/50), uniform or Zipf(1.1)
    keys (50/50), and a unit-rate exponential inter-arrival gap to be scaled
    by the arrival rate."""
    u = uniform01(seed, 4 * jobs)
    out = []
    for j in range(jobs):
        size = int(math.floor(2.0 ** (lo_log2 + span_log2 * u[4 * j])))
        out.append({"n": size,
                    "key_bits": 32 if u[4 * j + 1] < 0.5 else 64,
                    "zipf": u[4 * j + 2] < 0.5,
                    "gap": -math.log(1.0 - u[4 * j + 3])})
    return out
