"""Workload definitions shared by tests."""
import numpy as np

import pyoracle as po

# hydrogen, N=1000, order-7 B-splines (n_rank=7), R=200, L=31, 4-cycle sin^2
# pulse (E0=0.0534, omega=0.057; e0 holds A0 = E0/omega), dt=0.05, cos^(1/8)
# mask over the last 20 a.u.
H_N1000_L31 = dict(N=1000, p=7, L=31, rmax=200.0, dt=0.05, mask_rmin=180.0, mask_rmax=200.0, mask_fact=0.125,
            e0=0.0534 / 0.057, w0=0.057, nc=4.0)
# hydrogen, N=3000, R=600, L=81, 8 cycles, E0=0.0755
H_N3000_L81 = dict(N=3000, p=7, L=81, rmax=600.0, dt=0.05, mask_rmin=580.0, mask_rmax=600.0, mask_fact=0.125,
            e0=0.0755 / 0.057, w0=0.057, nc=8.0)


def make_oracle(cfg):
    return po.Oracle(N=cfg["N"], p=cfg["p"], L=cfg["L"], M=cfg.get("M", 1), m0=cfg.get("m0", 0),
                     rmax=cfg["rmax"], dt=cfg["dt"], mask_rmin=cfg["mask_rmin"], mask_rmax=cfg["mask_rmax"],
                     mask_fact=cfg["mask_fact"])


def shape(N, L, rmax, M=1, m0=0, nc=4.0, p=7, dt=0.05):
    """A small hydrogen workload of the given shape (mask over the last tenth of the box)."""
    return dict(N=N, p=p, L=L, M=M, m0=m0, rmax=rmax, dt=dt, mask_rmin=0.9 * rmax, mask_rmax=rmax,
                mask_fact=0.125, e0=0.0534 / 0.057, w0=0.057, nc=nc)


def atable(cfg):
    """A(t) on the time grid t_k = k*dt, k = 0..para_nt (field.h:243, 279)."""
    nt = po.para_nt(cfg["w0"], cfg["nc"], 0.0, cfg["dt"])
    a = np.array([po.sinen2(cfg["e0"], cfg["w0"], cfg["nc"], 0.0, 0.0, k * cfg["dt"]) for k in range(nt + 1)])
    return nt, a


def mixed_state(o, seed=1, amp=1e-2):
    """1s ground state plus a small seeded random admixture in every channel, S-normalised."""
    _, c = o.eigenstate(0, 0, 0)
    rng = np.random.default_rng(seed)
    n = c.size
    c = c + amp * (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(n)
    c /= np.sqrt(o.norm(c))
    return np.ascontiguousarray(c)


def upload_from_oracle(prop, o):
    prop.upload_radial(o.get(po.QOR_S), o.get(po.QOR_CHOL), o.get(po.QOR_RINV), o.get(po.QOR_RDIF),
                       o.get(po.QOR_RPOW1), o.get(po.QOR_MASK))
    prop.upload_hc(o.get(po.QOR_UPPER), o.get(po.QOR_LOWER))
    prop.upload_angular(o.get(po.QOR_YEIG), o.get(po.QOR_YPRJ), o.get(po.QOR_QEIG), o.get(po.QOR_QPRJ),
                        o.get(po.QOR_MERGED), o.get(po.QOR_Y10))


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
