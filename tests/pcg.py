"""numpy restatement of the PCG32 streams the GPU kernels draw from (checker
side): PBRT-v3's PCG32 as in the reference's rng_pcg (libgi/random.h:59-96),
seeded per (frame seed, path, level) through a SplitMix64 finaliser."""
import numpy as np

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
MULT = np.uint64(0x5851F42D4C957F2D)
DEFAULT_STATE = np.uint64(0x853C49E6748FEA9B)


def stream_id(seed, path, level):
    with np.errstate(over="ignore"):
        path = np.asarray(path, np.uint64)
        z = np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + (path << np.uint64(8)) + np.uint64(level) + np.uint64(1)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


class Pcg32:
    def __init__(self, initseq):
        with np.errstate(over="ignore"):
            initseq = np.asarray(initseq, np.uint64)
            self.state = np.zeros_like(initseq)
            self.inc = (initseq << np.uint64(1)) | np.uint64(1)
            self.next_u32()
            self.state = self.state + DEFAULT_STATE
            self.next_u32()

    def next_u32(self):
        with np.errstate(over="ignore"):
            old = self.state
            self.state = old * MULT + self.inc
            xs = (((old >> np.uint64(18)) ^ old) >> np.uint64(27)) & np.uint64(0xFFFFFFFF)
            rot = (old >> np.uint64(59)).astype(np.uint64)
            left = (xs << ((~rot + np.uint64(1)) & np.uint64(31))) & np.uint64(0xFFFFFFFF)
            return ((xs >> rot) | left).astype(np.uint32)

    def next_float(self):
        f = self.next_u32().astype(np.float32) * np.float32(2.0 ** -32)
        return np.minimum(f, np.float32(1.0 - 2.0 ** -24))
