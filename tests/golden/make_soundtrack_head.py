"""Writes ``soundtrack_head.npz``: the first second of the reference's rendered
soundtrack and the two loudness normalisers that go with it.

Source: ``renders/simulation_soundtrack.wav`` of the reference (data only),
produced by the default ``src/main.rs`` run: a 36x52x36 grid, 120 steps of
dt 0.08 at timeline times 15 * i / 119, vorticity eps 6.5 and 28 CG
iterations (main.rs:82-94), whose telemetry frames go through
``synthesize_fluid_soundtrack`` (audio.rs:397-560) and the TPDF-dithered
24-bit encoder (audio.rs:593-639).

The first second of audio reads only telemetry frames 0-8 (audio.rs:375-379
interpolates at u = t / 15 * 119) plus two normalisers: the peak kinetic
energy and the peak vorticity over all 120 frames (audio.rs:401-406). Those
peaks fall late in the run (frame ~118), where ulp-level differences between
platforms have been amplified by the flow's instability, so they cannot be
recomputed; they are fitted here instead, by Nelder-Mead on the RMS sample
difference over the first 0.4 s, with the oracle's first 11 frames as input.
Everything else in the comparison is computed.

Stored arrays:
  pcm_head    int32 (16000, 2): samples 0, 3, 6, ... 47997 (left, right).
  stride      3
  nsamples    48000
  norm        float64 (2,): fitted (max_ke, max_vort).
  steps, dims, dt, eps, iters: the run that produced the WAV.

Usage: python tests/golden/make_soundtrack_head.py [path/to/reference]
"""
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

NSAMPLES = 48000
STRIDE = 3
FIT_SAMPLES = 19200
DIMS, STEPS, DT, EPS, ITERS = (36, 52, 36), 120, 0.08, 6.5, 28
NFRAMES = 11


def read_wav24(path):
    """RIFF/WAVE reader for 24-bit PCM: returns (rate, channels, int32 samples
    shaped (frames, channels))."""
    b = open(path, "rb").read()
    if b[:4] != b"RIFF" or b[8:12] != b"WAVE":
        raise ValueError("not a RIFF/WAVE file")
    pos, fmt, data = 12, None, None
    while pos + 8 <= len(b):
        cid, size = b[pos:pos + 4], struct.unpack("<I", b[pos + 4:pos + 8])[0]
        body = b[pos + 8:pos + 8 + size]
        if cid == b"fmt ":
            fmt = struct.unpack("<HHIIHH", body[:16])
        elif cid == b"data":
            data = body
        pos += 8 + size + (size & 1)
    tag, ch, rate, _, align, bits = fmt
    if tag != 1 or bits != 24 or align != 3 * ch:
        raise ValueError(f"expected 24-bit PCM, got format {tag}, {bits} bits")
    raw = np.frombuffer(data, dtype=np.uint8).reshape(-1, 3).astype(np.int32)
    v = raw[:, 0] | (raw[:, 1] << 8) | (raw[:, 2] << 16)
    v = np.where(v >= 1 << 23, v - (1 << 24), v)
    return rate, ch, v.reshape(-1, ch)


def oracle_frames(nframes=NFRAMES):
    """The first telemetry frames of the run, padded to STEPS rows with the
    last one (rows past frame 9 do not reach the first second of audio)."""
    from oracle.oracle import OracleFluid
    s = OracleFluid(*DIMS)
    s.vorticity_eps, s.solver_iters = EPS, ITERS
    fr = np.array([s.step(DT, 15.0 * (i / (STEPS - 1))) for i in range(nframes)])
    return np.concatenate([fr, np.repeat(fr[-1:], STEPS - nframes, 0)])


def main():
    from scipy.optimize import minimize
    from oracle.oracle import soundtrack
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    rate, ch, pcm = read_wav24(os.path.join(ref_root, "renders", "simulation_soundtrack.wav"))
    assert rate == 48000 and ch == 2 and len(pcm) == 720000, (rate, ch, len(pcm))
    head = pcm[:NSAMPLES].astype(np.int64)
    frames = oracle_frames()

    def err(p):
        out = soundtrack(frames, p[0], p[1], FIT_SAMPLES).astype(np.int64)
        return float(np.sqrt(((out - head[:FIT_SAMPLES]) ** 2).mean()))

    best = min(((err((mk, mv)), mk, mv) for mk in np.linspace(0.03, 0.12, 19)
                for mv in np.linspace(100.0, 400.0, 13)))
    fit = minimize(err, [best[1], best[2]], method="Nelder-Mead",
                   options=dict(xatol=1e-10, fatol=1e-6, maxiter=400))
    norm = np.array(fit.x, dtype=np.float64)
    d = soundtrack(frames, norm[0], norm[1], NSAMPLES).astype(np.int64) - head
    print(f"fitted max_ke {norm[0]:.10e} max_vort {norm[1]:.10e}; "
          f"first second: max |diff| {np.abs(d).max()} LSB, exact {np.mean(d == 0):.5f}")
    np.savez_compressed(os.path.join(HERE, "soundtrack_head.npz"),
                        pcm_head=pcm[:NSAMPLES:STRIDE].astype(np.int32), stride=STRIDE,
                        nsamples=NSAMPLES, norm=norm, steps=STEPS, dims=np.array(DIMS),
                        dt=DT, eps=EPS, iters=ITERS)


if __name__ == "__main__":
    main()
