"""GPU parity of t-SURFF (spectrum_c_tsf::accumulate, momentum_1e.hpp:609-719)
against the oracle's restatement (qor_tsurff_accumulate, pinned by the
free-packet and constant-field physics checks of test_tsurff.py): single
accumulates on seeded states (with and without a field term, several m), and
whole runs across the pulse end with t-SURFF inside qpc_run (after every
step, before the absorber, as binsrc/sphTDSE.cpp:798-804)."""
import numpy as np
import pytest

from _cfg import atable, make_oracle, mixed_state, rel_l2, shape, upload_from_oracle

pytestmark = pytest.mark.gpu


def _ctx(cfg):
    import qpc_gpu
    o = make_oracle(cfg)
    prop = qpc_gpu.Propagator(cfg["N"], cfg["p"], cfg["L"], cfg.get("M", 1), cfg.get("m0", 0), cfg["dt"])
    upload_from_oracle(prop, o)
    st = qpc_gpu.Setup(cfg["N"], cfg["p"], cfg["L"], cfg.get("M", 1), cfg.get("m0", 0), cfg["dt"], 0.0, cfg["rmax"],
                       1.0, 1.0, cfg["mask_rmin"], cfg["mask_rmax"], cfg["mask_fact"])
    return o, prop, st


@pytest.mark.parametrize("N,L,rmax,M,m0", [(400, 8, 80.0, 1, 0), (240, 5, 48.0, 3, -1)])
def test_tsurff_accumulate_matches_oracle(N, L, rmax, M, m0):
    cfg = shape(N, L, rmax, M=M, m0=m0)
    print(f"case: tsurff accumulate N={N} L={L} M={M}", flush=True)
    o, prop, st = _ctx(cfg)
    t = st.tsurff(0.6 * rmax, 0.1, 2.0, 33, 0.0, np.pi, 7, 0.0, 5.0, 2 if M > 1 else 1)
    prop.upload_tsurff(t)
    state, work = np.zeros(2), np.zeros((33, 7, 2 if M > 1 else 1), np.complex128)
    for seed, (a0, a1) in ((3, (0.0, 0.0)), (4, (0.7, -0.4)), (5, (1.3, 1.1))):
        c = mixed_state(o, seed=seed, amp=0.8)
        prop.set_state(c)
        prop.tsurff_accumulate(a0, a1)
        o.tsurff_accumulate(c, a0, a1, cfg["dt"], t, state, work)
    got, tim, disz = prop.tsurff_read()
    err = rel_l2(got.ravel(), work.ravel())
    print(f"  rel L2 {err:.3e}, time {tim} disz {disz}", flush=True)
    assert err <= 1e-12
    assert tim == state[0] and disz == state[1]


def test_tsurff_run_matches_oracle_across_pulse_end():
    cfg = shape(400, 8, 80.0, nc=1.0)
    print("case: tsurff run N=400 L=8 across the pulse end", flush=True)
    o, prop, st = _ctx(cfg)
    nt, a = atable(cfg)
    t = st.tsurff(50.0, 0.2, 2.5, 40, 0.0, np.pi, 9)
    prop.upload_tsurff(t)
    prop.tsurff(True)
    c = mixed_state(o, seed=11, amp=0.3)
    it0, it1 = nt - 150, nt + 60
    prop.set_state(c)
    prop.run(a, nt, it0, it1, use_mask=True)
    got, tim, disz = prop.tsurff_read()
    state, work = np.zeros(2), np.zeros((40, 9, 1), np.complex128)
    for it in range(it0, it1):
        if it < nt:
            o.step(c, a[it], a[it + 1])
            o.tsurff_accumulate(c, a[it], a[it + 1], cfg["dt"], t, state, work)
        else:
            o.step_free(c)
            o.tsurff_accumulate(c, 0.0, 0.0, cfg["dt"], t, state, work)
        o.mask(c)
    err = rel_l2(got.ravel(), work.ravel())
    serr = rel_l2(prop.get_state(), c)
    print(f"  work rel L2 {err:.3e}, state rel L2 {serr:.3e}", flush=True)
    assert serr <= 1e-11 and err <= 1e-10
    assert abs(tim - state[0]) < 1e-12 and abs(disz - state[1]) < 1e-12
    # the length gauge refuses t-SURFF (sphTDSE.cpp:288)
    import qpc_gpu
    with pytest.raises(qpc_gpu.QpcError):
        prop.run_lg(a, nt, it0, it0 + 1)
    prop.tsurff(False)
    prop.close()


def test_tsurff_xy_accumulate_matches_oracle():
    """accumulate (xy) (momentum_1e.hpp:716-776) on seeded states of an (m,l)
    grid, with and without A_x, A_y, against qor_tsurff_accumulate_xy."""
    cfg = shape(240, 5, 48.0, M=5, m0=-2)
    print("case: tsurff xy accumulate N=240 L=5 M=5", flush=True)
    o, prop, st = _ctx(cfg)
    t = st.tsurff(0.6 * cfg["rmax"], 0.1, 2.0, 17, 0.0, np.pi, 7, 0.0, 5.0, 5)
    prop.upload_tsurff(t)
    state, work = np.zeros(3), np.zeros((17, 7, 5), np.complex128)
    for seed, f in ((3, (0.0, 0.0, 0.0, 0.0)), (4, (0.7, -0.4, 0.2, 0.5)), (5, (1.3, 1.1, -0.8, -0.6))):
        c = mixed_state(o, seed=seed, amp=0.8)
        prop.set_state(c)
        prop.tsurff_accumulate_xy(*f)
        o.tsurff_accumulate_xy(c, *f, cfg["dt"], t, t["pmat1"], t["pmat2"], state, work)
    got, tim, _ = prop.tsurff_read()
    dis = prop.tsurff_displacement()
    err = rel_l2(got.ravel(), work.ravel())
    print(f"  rel L2 {err:.3e}, time {tim} dis {dis}", flush=True)
    assert err <= 1e-12
    assert tim == state[0] and dis[0] == state[1] and dis[1] == state[2] and dis[2] == 0.0


def test_tsurff_ep_run_matches_oracle_across_pulse_end():
    """t-SURFF inside qpc_run_ep (binsrc/sphTDSE.cpp:939-945: the xy form with
    A_x, A_y during the pulse, the field-free form after it, before the
    absorber) against the oracle's loop of vgwr steps, accumulates and masks."""
    import pyoracle as po
    cfg = shape(300, 6, 60.0, M=5, m0=-2, nc=1.0)
    print("case: tsurff xy in the EP run N=300 L=6 M=5 across the pulse end", flush=True)
    o, prop, st = _ctx(cfg)
    ep = po.ep_angular(cfg["M"], cfg["L"], cfg["m0"])
    o.set_ep(ep, isub=-1)
    prop.upload_angular_ep(ep, isub=-1)
    nt = po.para_nt(cfg["w0"], cfg["nc"], 0.0, cfg["dt"])
    a0 = cfg["e0"] / np.sqrt(1.25)
    ax = np.array([po.sinen2(a0, cfg["w0"], cfg["nc"], 0.0, 0.0, k * cfg["dt"]) for k in range(nt + 1)])
    ay = np.array([po.sinen2(0.5 * a0, cfg["w0"], cfg["nc"], 0.0, 0.5, k * cfg["dt"]) for k in range(nt + 1)])
    t = st.tsurff(40.0, 0.2, 2.5, 24, 0.0, np.pi, 7, 0.0, 5.5, 6)
    prop.upload_tsurff(t)
    prop.tsurff(True)
    c = mixed_state(o, seed=13, amp=0.3)
    it0, it1 = nt - 40, nt + 12
    prop.set_state(c)
    prop.run_ep(ax, ay, nt, it0, it1, use_mask=True)
    got, tim, _ = prop.tsurff_read()
    dis = prop.tsurff_displacement()
    state, work = np.zeros(3), np.zeros((24, 7, 6), np.complex128)
    for it in range(it0, it1):
        f = (ax[it], ax[it + 1], ay[it], ay[it + 1]) if it < nt else (0.0, 0.0, 0.0, 0.0)
        if it < nt:
            o.step_ep(c, *f)
        else:
            o.step_free(c)
        o.tsurff_accumulate_xy(c, *f, cfg["dt"], t, t["pmat1"], t["pmat2"], state, work)
        o.mask(c)
    err = rel_l2(got.ravel(), work.ravel())
    serr = rel_l2(prop.get_state(), c)
    print(f"  work rel L2 {err:.3e}, state rel L2 {serr:.3e}", flush=True)
    assert serr <= 1e-11 and err <= 1e-10
    assert abs(tim - state[0]) < 1e-12 and abs(dis[0] - state[1]) < 1e-12 and abs(dis[1] - state[2]) < 1e-12
    prop.tsurff(False)
    prop.close()


@pytest.mark.timeout(300)
def test_tsurff_ep_constant_field_keeps_the_free_spectrum():
    """Physics on the device path: a free outgoing packet (l = 0) propagated by
    the elliptically polarised step under a constant vector potential in the
    x-y plane, with t-SURFF inside qpc_run_ep: the canonical momentum
    distribution is that of the field-free packet, |<k|psi(0)>|^2, at every
    polar and azimuthal angle (this exercises the sin e^(+-i phi) couplings
    and the Volkov phase k sin(theta) (cos(phi) int A_x + sin(phi) int A_y))."""
    import pyoracle as po
    import qpc_gpu
    from test_tsurff import _packet, _projection_l0
    N, p, rmax, r0, dt = 400, 7, 200.0, 100.0, 0.05
    L, M, m0 = 16, 25, -12
    rc, sigma, k0 = 60.0, 5.0, 1.5
    A = (0.024, -0.018)
    print(f"case: tsurff xy physics N={N} L={L} M={M} A={A}", flush=True)
    st = qpc_gpu.Setup(N, p, L, M, m0, dt, 0.0, rmax, 0.0, 1.0, 180.0, 200.0, 0.125)
    prop = qpc_gpu.Propagator(N, p, L, M, m0, dt)
    prop.upload_setup(st)
    prop.upload_angular_ep(qpc_gpu.angular_ep(M, L, m0), isub=-1)
    o = po.Oracle(N=N, p=p, L=1, rmax=rmax, zc=0.0, dt=dt, mask_rmin=180.0, mask_rmax=200.0)
    c0, u, _ = _packet(o, N, p, rmax, rc, sigma, k0)
    c = np.zeros(M * L * N, np.complex128)
    ch = (-m0) * L
    c[ch * N:(ch + 1) * N] = c0
    t = st.tsurff(r0, 1.2, 1.8, 7, 0.3, 2.8, 4, 0.0, 4.7, 4)
    prop.upload_tsurff(t)
    prop.tsurff(True)
    nsteps = int(110.0 / dt)
    r = np.linspace(1e-6, rmax - 1e-6, 40001)
    scale = np.sqrt(np.trapezoid(np.abs(u(r)) ** 2, r))
    want = np.abs(_projection_l0(u, t["kr"], rc, sigma) / scale) ** 2
    g = {}
    for a in ((0.0, 0.0), A):
        prop.tsurff_clear()
        prop.set_state(c)
        prop.run_ep(np.full(nsteps + 1, a[0]), np.full(nsteps + 1, a[1]), nsteps, 0, nsteps, use_mask=True)
        got, tim, _ = prop.tsurff_read()
        dis = prop.tsurff_displacement()
        g[a] = np.abs(got) ** 2
        dev = np.abs(g[a] - want[:, None, None]).max() / want.max()
        print(f"  A={a}: time {tim:.2f} dis {dis}, max rel dev from |<k|psi(0)>|^2 {dev:.2e}", flush=True)
        assert abs(dis[0] - a[0] * nsteps * dt) < 1e-9 and abs(dis[1] - a[1] * nsteps * dt) < 1e-9
        assert dev < 2e-3
    # the field's own effect, apart from the free propagation's discretisation
    dfa = np.abs(g[A] - g[(0.0, 0.0)]).max() / want.max()
    print(f"  field vs field-free spectrum: max rel dev {dfa:.2e}", flush=True)
    assert dfa < 5e-4
    prop.tsurff(False)
    prop.close()
