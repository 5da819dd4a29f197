"""Writes tests/golden/c3_oracle_512.npz (or the path given as the first argument): the FP64
oracle's field after the 512^3 precision study (run from the repo root; about 15 minutes on 16
cores, two hours on 8 slower ones).

The run: 3D 512^3, eps = 0.3, dx = pi/4, dt = 0.5, psi0 = -0.25 + U[-0.05, 0.05] from seed 11
(the oracle's own Philox initial field), noise off, cubic split, 500 semi-implicit steps of
oracle/oracle.c's orc_pfc_run.  The whole field is 1.07 GB in FP64, so the fixture holds:

  sample     psi[::8, ::8, ::8]                       every 8th cell per axis, 64^3 FP64
  block      psi[B:B+32, B:B+32, B:B+32], B = 240      a contiguous 32^3 block at full resolution
  sumsq, total, vmax, vmin                            sum psi^2, sum psi, max and min over all cells
  mean_hist  the mean density after every step (k = 0 mode / N), 501 values
  params     n, eps, dt, psi0, amp, seed, steps, chunk

The steps run in chunks of `chunk` (noise off, so the chunks compose; each chunk starts from the
real field, a projection onto the Hermitian part the fused stepper keeps anyway).
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import oracle  # noqa: E402

N, EPS, DT, PSI0, AMP, SEED, STEPS, CHUNK = 512, 0.3, 0.5, -0.25, 0.05, 11, 500, 50
STRIDE, B, BW = 8, 240, 32


def main():
    oracle.lib()
    h = [np.pi / 4] * 3
    psi = oracle.init_field((N, N, N), PSI0, AMP, SEED)
    means = [float(psi.mean())]
    t0 = time.time()
    for s0 in range(0, STEPS, CHUNK):
        psi, m = oracle.pfc_run(psi, h, EPS, DT, 0.0, 0, s0, CHUNK, 0, want_mass=True)
        means += [float(v) for v in m[1:]]
        print(f"oracle 512^3: {s0 + CHUNK}/{STEPS} steps, {time.time() - t0:.0f} s", flush=True)
    out = dict(
        sample=np.ascontiguousarray(psi[::STRIDE, ::STRIDE, ::STRIDE]),
        block=np.ascontiguousarray(psi[B:B + BW, B:B + BW, B:B + BW]),
        block_origin=np.array(B), stride=np.array(STRIDE),
        sumsq=np.array(float(np.sum(psi * psi))), total=np.array(float(np.sum(psi))),
        vmax=np.array(float(psi.max())), vmin=np.array(float(psi.min())),
        mean_hist=np.array(means),
        params=np.array([N, EPS, DT, PSI0, AMP, SEED, STEPS, CHUNK], dtype=np.float64),
    )
    path = sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "c3_oracle_512.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: L2 {np.sqrt(out['sumsq']):.6e}, mean {out['total'] / N ** 3:.6e}, "
          f"max {out['vmax']:.6f}, min {out['vmin']:.6f}", flush=True)


if __name__ == "__main__":
    main()
