"""Writes tests/golden/rng_golden.npz: TEA seeds and PCG streams computed by an
independent numpy-uint32 restatement of assets/glsl/sampling.glsl:18-47
(written from the GLSL text, not from the oracle), used to pin the oracle's
sr_tea / sr_next_rand.

    python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
U = np.uint32


def tea(v0, v1):
    """sampling.glsl:18-32, vectorised over uint32 arrays"""
    v0 = np.asarray(v0, U).copy()
    v1 = np.asarray(v1, U).copy()
    s0 = np.zeros_like(v0)
    with np.errstate(over="ignore"):
        for _ in range(16):
            s0 = s0 + U(0x9E3779B9)
            v0 = v0 + (((v1 << U(4)) + U(0xA341316C)) ^ (v1 + s0) ^ ((v1 >> U(5)) + U(0xC8013EA4)))
            v1 = v1 + (((v0 << U(4)) + U(0xAD90777D)) ^ (v0 + s0) ^ ((v0 >> U(5)) + U(0x7E95761E)))
    return v0


def pcg_stream(state, n):
    """sampling.glsl:35-42: LCG step (multiplier 747796405, increment 1) and the
    RXS-M-XS 32/32 output permutation; returns (words, floats)"""
    state = np.asarray(state, U).copy()
    words, floats = [], []
    with np.errstate(over="ignore"):
        for _ in range(n):
            state = state * U(747796405) + U(1)
            w = ((state >> ((state >> U(28)) + U(4))) ^ state) * U(277803737)
            w = (w >> U(22)) ^ w
            words.append(w.copy())
            # float(word) / 4294967295.0f: the divisor rounds to 2^32 in fp32
            floats.append((w.astype(np.float32) / np.float32(4294967295.0)).astype(np.float32))
    return np.stack(words, -1), np.stack(floats, -1)


def main():
    rng = np.random.default_rng(2024)
    W = 1920
    xs = np.concatenate([np.arange(8), rng.integers(0, W, 56)]).astype(U)
    ys = np.concatenate([np.arange(8), rng.integers(0, 1080, 56)]).astype(U)
    frames = np.concatenate([np.zeros(8), rng.integers(0, 1 << 20, 56)]).astype(U)
    pixel = (xs + ys * U(W)).astype(U)
    seeds = tea(pixel, frames)
    words, floats = pcg_stream(seeds, 40)
    edge_states = np.array([0, 1, 0xFFFFFFFF, 0x80000000, 0x7FFFFFFF, 123456789], U)
    ew, ef = pcg_stream(edge_states, 8)
    np.savez(os.path.join(HERE, "rng_golden.npz"), pixel=pixel, frames=frames, seeds=seeds, words=words,
             floats=floats, edge_states=edge_states, edge_words=ew, edge_floats=ef)
    print("wrote rng_golden.npz", seeds.shape, floats.shape)


if __name__ == "__main__":
    main()
