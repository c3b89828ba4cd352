"""Writes full_run_h-N3000-L81.npz: the oracle's whole N=3000, L=81 run
(the benchmark workload).

Hydrogen, order-7 B-splines, R=600, l_max=80, 8-cycle sin^2 pulse
(E0=0.0755, omega=0.057), dt=0.05, absorber every step: all 17638 steps of
the driver loop (sphTDSE.cpp:784-833) from the 1s state of the l=0 block,
on the oracle's operators.  Stored: the initial vector c0, the final
coefficients, and the norm and <z> observed before every step.  The oracle
runs on the CPU (several minutes with OpenMP); the GPU test replays the
same run on the device from c0 with the oracle's operators uploaded.

    python tests/golden/make_full_run_h_n3000.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, ".."))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))

from _cfg import H_N3000_L81, atable, make_oracle  # noqa: E402

OUT = os.path.join(HERE, "full_run_h-N3000-L81.npz")


def main():
    o = make_oracle(H_N3000_L81)
    nt, a = atable(H_N3000_L81)
    _, c0 = o.eigenstate(0, 0, 0)
    c = np.ascontiguousarray(c0.copy())
    tic = time.perf_counter()
    obs = []
    for i0 in range(0, nt, 500):
        obs.append(o.run(c, a, nt, i0, min(nt, i0 + 500), use_mask=True, observe=True))
        print(f"step {min(nt, i0 + 500)} of {nt}, {time.perf_counter() - tic:.0f} s", flush=True)
    obs = np.concatenate(obs)
    np.savez_compressed(OUT, c0=c0, final=c, norm=obs[:, 0], dipole_z=obs[:, 1])
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
