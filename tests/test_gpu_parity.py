"""GPU parity: the HIP propagator (through the C ABI) against the CPU oracle on
the same operands and the same initial vector.  Bar: relative L2 of the
coefficients <= 1e-10 over a driver-loop run (the CN integrator's tolerance),
<= 1e-12 for single steps and short runs."""
import numpy as np
import pytest

from _cfg import H_N1000_L31, H_N3000_L81, atable, make_oracle, mixed_state, rel_l2, shape, upload_from_oracle

pytestmark = pytest.mark.gpu


def _gpu_prop(o, cfg):
    import qpc_gpu
    prop = qpc_gpu.Propagator(cfg["N"], cfg["p"], cfg["L"], cfg.get("M", 1), cfg.get("m0", 0), cfg["dt"])
    upload_from_oracle(prop, o)
    return prop


@pytest.fixture(scope="module")
def small_run():
    o = make_oracle(H_N1000_L31)
    return o, _gpu_prop(o, H_N1000_L31)


def _single_pieces(o, prop, label, a=(0.7, 0.8), seed=3):
    c0 = mixed_state(o, seed=seed)
    for name, fn_o, fn_g in [
        ("vgwz", lambda c: o.step(c, *a), lambda: prop.propagate(*a)),
        ("free", lambda c: o.step_free(c), lambda: prop.propagate()),
        ("mask", lambda c: o.mask(c), lambda: prop.mask()),
    ]:
        c = c0.copy()
        fn_o(c)
        prop.set_state(c0)
        fn_g()
        err = rel_l2(prop.get_state(), c)
        print(f"  {label} {name}: rel L2 = {err:.3e}", flush=True)
        assert err <= 1e-12, (label, name, err)


def test_single_step_pieces(small_run):
    o, prop = small_run
    print("case: N=1000 L=31 single with-field step / free step / mask", flush=True)
    _single_pieces(o, prop, "N=1000 L=31")


# shapes at the edges of the partitioned solver: fewer partitions than the
# carry look-back (N=96), one to four wave groups of partitions (N=1000,
# 4000, 6144 = the largest supported), several m blocks (M=3, m0=-1) and a
# partition count that leaves padding rows (N=1000, 4000), and 600 channels
# (M=15, L=40), a grid large enough that the solves read the transfer sets
# from global memory instead of staging them in LDS
@pytest.mark.parametrize("N,L,rmax,M,m0", [(96, 4, 20.0, 1, 0), (4000, 6, 800.0, 1, 0),
                                           (6144, 3, 1200.0, 1, 0), (240, 5, 48.0, 3, -1),
                                           (240, 40, 48.0, 15, -7)])
def test_single_step_shapes(N, L, rmax, M, m0):
    cfg = shape(N, L, rmax, M=M, m0=m0)
    print(f"case: single step pieces N={N} L={L} M={M} m0={m0}", flush=True)
    o = make_oracle(cfg)
    prop = _gpu_prop(o, cfg)
    _single_pieces(o, prop, f"N={N} L={L} M={M}", a=(0.9, 1.1), seed=N)


def test_observables(small_run):
    o, prop = small_run
    c = mixed_state(o, seed=5, amp=0.3)
    prop.set_state(c)
    import qpc_gpu
    n_g = prop.observe(qpc_gpu.QPC_OBS_NORM)
    z_g = prop.observe(qpc_gpu.QPC_OBS_DIPOLE_Z)
    assert abs(n_g - o.norm(c)) <= 1e-13
    assert abs(z_g - o.dipole_z(c, 0)) <= 1e-12 * max(1.0, abs(o.dipole_z(c, 0)))


def test_run_n1000_l31_first_400_steps(small_run):
    """First 400 steps of the N=1000, L=31 run (pulse on, 13 factorisation
    batches on the second stream), absorber every step."""
    o, prop = small_run
    nt, a = atable(H_N1000_L31)
    assert nt == 8819
    c = mixed_state(o, seed=7)
    prop.set_state(c)
    steps = (0, 400)
    print(f"case: N=1000 L=31 driver loop steps {steps}", flush=True)
    o.run(c, a, nt, steps[0], steps[1], use_mask=True)
    prop.run(a, nt, steps[0], steps[1], use_mask=True)
    err = rel_l2(prop.get_state(), c)
    print(f"  rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-10


def test_full_run_n1000_l31_against_golden(small_run):
    """All 8819 steps of the N=1000, L=31 run (pulse, absorber every step)
    from the 1s state, against the oracle's run stored in
    tests/golden/full_run_h-N1000-L31.npz: the final coefficients, and the
    norm and <z> observed before every step."""
    import os
    import qpc_gpu
    o, prop = small_run
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "full_run_h-N1000-L31.npz"))
    nt, a = atable(H_N1000_L31)
    print("case: N=1000 L=31 whole run, 8819 steps from the 1s state", flush=True)
    prop.set_state(g["c0"])
    prop.run(a, nt, 0, nt, use_mask=True)
    err = rel_l2(prop.get_state(), g["final"])
    print(f"  final rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-10
    prop.set_state(g["c0"])
    nrm, z = np.zeros(nt), np.zeros(nt)
    for it in range(nt):
        nrm[it] = prop.observe(qpc_gpu.QPC_OBS_NORM)
        z[it] = prop.observe(qpc_gpu.QPC_OBS_DIPOLE_Z)
        prop.run(a, nt, it, it + 1, use_mask=True)
    dn, dz = np.abs(nrm - g["norm"]).max(), np.abs(z - g["dipole_z"]).max()
    print(f"  norm history max diff {dn:.3e}, <z> history max diff {dz:.3e} (max |<z>| "
          f"{np.abs(g['dipole_z']).max():.3f})", flush=True)
    # the norm is quadratic in the state: a state within 1e-10 allows 2e-10
    assert dn <= 1e-11
    assert dz <= 1e-10 * np.abs(g["dipole_z"]).max()
    assert rel_l2(prop.get_state(), g["final"]) <= 1e-10


def test_run_across_pulse_end():
    """A run that crosses the end of the pulse: with-field steps, then
    field-free steps, absorber every step (1-cycle pulse, N=400, L=10)."""
    cfg = shape(400, 10, 80.0, nc=1.0)
    o = make_oracle(cfg)
    prop = _gpu_prop(o, cfg)
    nt, a = atable(cfg)
    steps = (nt - 45, nt + 20)
    print(f"case: N=400 L=10 driver loop steps {steps}, pulse ends at {nt}", flush=True)
    c = mixed_state(o, seed=13)
    prop.set_state(c)
    o.run(c, a, nt, steps[0], steps[1], use_mask=True)
    prop.run(a, nt, steps[0], steps[1], use_mask=True)
    err = rel_l2(prop.get_state(), c)
    print(f"  rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-12


def test_run_n3000_l81_pulse_peak():
    """N=3000, L=81 operators: 3 with-field steps at the pulse peak + absorber."""
    o = make_oracle(H_N3000_L81)
    prop = _gpu_prop(o, H_N3000_L81)
    nt, a = atable(H_N3000_L81)
    assert nt == 17638
    c = mixed_state(o, seed=11)
    prop.set_state(c)
    it0 = nt // 2
    print(f"case: N=3000 L=81 driver loop steps [{it0},{it0 + 3})", flush=True)
    o.run(c, a, nt, it0, it0 + 3, use_mask=True)
    prop.run(a, nt, it0, it0 + 3, use_mask=True)
    err = rel_l2(prop.get_state(), c)
    print(f"  rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-12


def test_extreme_field_runs_exactly_and_nonfinite_is_refused(small_run):
    """Any finite A(t): |A| = 40 a.u. on the N=1000 grid (D-part scalar
    e/h ~ 2.5, beyond any partitioned carry) plans the sequential solve and a
    long factorisation restart, and the step matches the oracle's sequential
    sweeps; a non-finite A is refused before any kernel runs."""
    import qpc_gpu
    o, prop = small_run
    c0 = mixed_state(o, seed=1)
    c = c0.copy()
    print("case: N=1000 L=31 single step at |A| = 40", flush=True)
    o.step(c, 40.0, -38.0)
    prop.set_state(c0)
    prop.propagate(40.0, -38.0)
    err = rel_l2(prop.get_state(), c)
    print(f"  rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-11
    with pytest.raises(qpc_gpu.QpcError):
        prop.propagate(float("nan"), 1.0)


@pytest.mark.parametrize("N,L,M,m0,nsel", [(1000, 31, 1, 0, 20), (240, 5, 3, -1, 6)])
def test_projection_matches_oracle(N, L, M, m0, nsel):
    """qpc_project (split-K FP64 MFMA GEMM per l) against eigenstate_sph
    projection restated in the oracle, same basis, reference layout
    (stride N), on a state with amplitude in every channel."""
    cfg = shape(N, L, 20.0 * N / 100.0, M=M, m0=m0)
    print(f"case: projection N={N} L={L} M={M} nsel={nsel}", flush=True)
    o = make_oracle(cfg)
    prop = _gpu_prop(o, cfg)
    W, _ = o.projection_basis(nsel)
    prop.upload_projection(W)
    c = mixed_state(o, seed=N + nsel, amp=0.3)
    prop.set_state(c)
    g = prop.project(stride=N)
    ref = o.projection(c, W, stride=N)
    err = rel_l2(g, ref)
    print(f"  rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-13


@pytest.mark.parametrize("N,L,M,m0,kmax", [(400, 8, 1, 0, 24), (240, 5, 3, -1, 96), (240, 5, 1, 0, 64)])
def test_krylov_step_matches_oracle(N, L, M, m0, kmax):
    """qpc_krylov_step (device Lanczos) against propagator_arnoldi::propagate1
    restated in the oracle: same subspace size, same criterion, same state.
    Bar: 1e-10 (below the 1e-9 of a Krylov run).  With m0 = -1 the m = -1
    and m = +1 blocks share one operator, and an unconverged step of the
    thresholded-reorthogonalisation Lanczos is then ill-conditioned in the
    algorithm itself (a 1e-15 relative perturbation of the input moves the
    oracle's own output by ~1e-2 at kmax = 16), so that case runs with a
    subspace that converges (eps < 1e-20, where the same perturbation moves
    the output by ~1e-11)."""
    import pyoracle as po
    cfg = shape(N, L, 0.2 * N, M=M, m0=m0)
    print(f"case: krylov step N={N} L={L} M={M} kmax={kmax}", flush=True)
    o = make_oracle(cfg)
    prop = _gpu_prop(o, cfg)
    prop.upload_krylov(o.get(po.QOR_H), o.get(po.QOR_Y10), o.get(po.QOR_Q10))
    c = mixed_state(o, seed=N + kmax, amp=0.3)
    for az, dt in [(0.8, 0.05), (-1.3, 0.2)]:
        ref = c.copy()
        crit_o, m_o, eps_o, _ = o.krylov_step(ref, az, dt, kmax=kmax, emin=1e-14, emax=1e-10)
        prop.set_state(c)
        crit_g, m_g, eps_g = prop.krylov_step(az, dt, kmax=kmax, emin=1e-14, emax=1e-10)
        err = rel_l2(prop.get_state(), ref)
        print(f"  az={az} dt={dt}: m={m_g}/{m_o} crit={crit_g}/{crit_o} eps={eps_g:.2e}/{eps_o:.2e} rel L2={err:.3e}",
              flush=True)
        assert m_g == m_o and crit_g == crit_o
        assert err <= 1e-10
        c = ref


def test_krylov_adaptive_step_matches_oracle():
    """qpc_krylov_step_adaptive stops at the first subspace whose estimate
    meets the tolerance; with that size the oracle's propagate1 (kmax = m)
    builds the same Lanczos vectors, so the states agree."""
    import pyoracle as po
    cfg = shape(400, 8, 80.0)
    print("case: adaptive krylov step N=400 L=8", flush=True)
    o = make_oracle(cfg)
    prop = _gpu_prop(o, cfg)
    prop.upload_krylov(o.get(po.QOR_H), o.get(po.QOR_Y10), o.get(po.QOR_Q10))
    c = mixed_state(o, seed=41, amp=0.3)
    for az, dt, tol in [(0.8, 0.02, 1e-20), (-1.1, 0.05, 1e-16), (0.3, 0.01, 1e-24)]:
        prop.set_state(c)
        m, eps, ab = prop.krylov_step_adaptive(az, dt, 4, 64, tol)
        ref = c.copy()
        _, m_o, eps_o, ab_o = o.krylov_step(ref, az, dt, kmax=m)
        err = rel_l2(prop.get_state(), ref)
        print(f"  az={az} dt={dt} tol={tol:.0e}: m={m} eps={eps:.2e}/{eps_o:.2e} rel L2={err:.3e}", flush=True)
        assert m_o == m and 4 <= m < 64 and eps <= tol
        assert np.abs(ab - ab_o).max() < 1e-10 * np.abs(ab_o).max()
        assert err <= 1e-10
        c = ref


def test_krylov_propagate_matches_converged_oracle_steps():
    """qpc_krylov_propagate without a controller (CF4 Magnus, fixed step,
    adaptive subspaces) against the same CF4 sequence of converged oracle
    Lanczos steps (kmax 64) on a sin^2 pulse."""
    cfg = shape(400, 8, 80.0)
    print("case: krylov propagate N=400 L=8, 12 CF4 steps", flush=True)
    o = make_oracle(cfg)
    prop = _gpu_prop(o, cfg)
    import pyoracle as po
    prop.upload_krylov(o.get(po.QOR_H), o.get(po.QOR_Y10), o.get(po.QOR_Q10))
    pulse = (cfg["e0"], cfg["w0"], cfg["nc"], 0.0, 0.0)
    c = mixed_state(o, seed=43, amp=0.1)
    prop.set_state(c)
    t0, h, n = 150.0, 0.04, 12
    stats, _ = prop.krylov_propagate(pulse, t0, t0 + n * h, h, 1e-26, kmin=4, kmax=64)
    r3 = np.sqrt(3.0)
    a1, a2 = (3 - 2 * r3) / 12, (3 + 2 * r3) / 12
    A = lambda t: po.sinen2(*pulse, t)
    for k in range(n):
        t = t0 + k * h
        A1, A2 = A(t + (0.5 - r3 / 6) * h), A(t + (0.5 + r3 / 6) * h)
        for a in (2 * (a2 * A1 + a1 * A2), 2 * (a1 * A1 + a2 * A2)):
            o.krylov_step(c, a, h / 2, kmax=64)
    err = rel_l2(prop.get_state(), c)
    print(f"  stats {stats}, rel L2 = {err:.3e}", flush=True)
    assert stats["steps"] == n and stats["rejected"] == 0 and stats["exponentials"] == 2 * n
    assert err <= 1e-10


def test_learned_controller_against_plain_krylov():
    """The committed controller (qpc-tdse-cpc_amd/krylov_controller_params.npy,
    trained by bench.py --krylov-tts --train) against the method without it:
    0.2 a.u. of the N=600, L=12 run from the 1s state at the field maximum,
    both with estimate tolerance 1e-22, against a tight reference (step
    0.0025, tolerance 1e-26).  Both stay within 1e-8 (wavefunction L2 norm);
    the work ratio is printed (no gain is asserted: on this workload the
    trained controller matches the plain method's work)."""
    import os
    import qpc_gpu
    import krylov_controller as kc
    path = os.path.join(os.path.dirname(qpc_gpu.__file__), "krylov_controller_params.npy")
    print("case: learned controller vs plain Krylov, N=600 L=12", flush=True)
    st = qpc_gpu.Setup(600, 7, 12, 1, 0, 0.05, 0.0, 120.0, 1.0, 1.0, 100.0, 120.0, 0.125)
    S = st.get(qpc_gpu.SETUP_S)
    _, c0 = st.eigenstate(0, 0, 0)
    prop = qpc_gpu.Propagator(600, 7, 12, 1, 0, 0.05)
    prop.upload_setup(st)
    net = qpc_gpu.Mlp(seed=1)
    net.params = np.load(path)
    pulse = (0.0534 / 0.057, 0.057, 4.0, 0.0, 0.0)
    t0, T = 193.759, 0.2
    out = {}
    for name, dt, tol, use in (("ref", 0.0025, 1e-26, False), ("plain", 0.01, 1e-22, False),
                               ("learned", 0.01, 1e-22, True)):
        prop.set_state(c0)
        stats, _ = prop.krylov_propagate(pulse, t0, t0 + T, dt, tol, kmin=8, kmax=96, net=net if use else None)
        out[name] = (prop.get_state(), stats)
    ref = out["ref"][0]
    for name in ("plain", "learned"):
        err = kc.snorm(S, out[name][0] - ref) / kc.snorm(S, ref)
        print(f"  {name}: {out[name][1]}, rel err (S) {err:.2e}", flush=True)
        assert err <= 1e-8 and out[name][1]["rejected"] == 0
    print(f"  vectors learned / plain = {out['learned'][1]['vectors'] / out['plain'][1]['vectors']:.3f}", flush=True)


def test_learned_krylov_replayed_by_oracle():
    """The learned-Krylov path (qpc_krylov_propagate with the committed
    controller choosing each step's size and starting subspace) against the
    oracle: the accepted steps, recorded without changing the propagation,
    are replayed as the same CF4 Magnus steps of converged oracle Lanczos
    exponentials (propagator_arnoldi::propagate1, la_1e.hpp:312-335).  Bar:
    relative L2 <= 1e-9 (the learned-Krylov tolerance).  N=600, L=12, 0.2 a.u.
    from the 1s state at the field maximum of the 4-cycle pulse."""
    import os
    import qpc_gpu
    import pyoracle as po
    path = os.path.join(os.path.dirname(qpc_gpu.__file__), "krylov_controller_params.npy")
    cfg = shape(600, 12, 120.0)
    print("case: learned Krylov N=600 L=12 replayed by the oracle", flush=True)
    o = make_oracle(cfg)
    prop = _gpu_prop(o, cfg)
    prop.upload_krylov(o.get(po.QOR_H), o.get(po.QOR_Y10), o.get(po.QOR_Q10))
    st = qpc_gpu.Setup(600, 7, 12, 1, 0, 0.05, 0.0, 120.0, 1.0, 1.0, 100.0, 120.0, 0.125)
    prop.upload_outer(st.get(qpc_gpu.SETUP_SOUTER))  # the controller's outer-norm feature, r > 100
    net = qpc_gpu.Mlp(seed=1)
    net.params = np.load(path)
    pulse = (0.0534 / 0.057, 0.057, 4.0, 0.0, 0.0)
    t0, T = 193.759, 0.2
    _, c0 = o.eigenstate(0, 0, 0)
    prop.set_state(c0)
    stats, rec = prop.krylov_propagate(pulse, t0, t0 + T, 0.01, 1e-22, kmin=8, kmax=96, net=net, record=256,
                                       record_ab=False)
    got = prop.get_state()
    prop.set_state(c0)
    stats2, _ = prop.krylov_propagate(pulse, t0, t0 + T, 0.01, 1e-22, kmin=8, kmax=96, net=net)
    same = rel_l2(prop.get_state(), got)
    hs, ms = rec["h"], rec["m"]
    print(f"  stats {stats}; {len(hs)} steps, h in [{hs.min():.4g}, {hs.max():.4g}], first subspaces "
          f"{ms.min()}-{ms.max()}; unrecorded run differs by {same:.1e}", flush=True)
    assert stats["rejected"] == 0 and len(hs) == stats["steps"] and abs(hs.sum() - T) < 1e-12
    assert stats2["steps"] == stats["steps"] and same <= 1e-12
    # the controller's schedule is not the fixed base step
    assert len(set(np.round(hs, 12))) > 1 or abs(hs[0] - 0.01) > 1e-12
    r3 = np.sqrt(3.0)
    a1, a2 = (3 - 2 * r3) / 12, (3 + 2 * r3) / 12
    A = lambda t: po.sinen2(*pulse, t)
    c, t = c0.copy(), t0
    for h in hs:
        A1, A2 = A(t + (0.5 - r3 / 6) * h), A(t + (0.5 + r3 / 6) * h)
        for a in (2 * (a2 * A1 + a1 * A2), 2 * (a1 * A1 + a2 * A2)):
            o.krylov_step(c, a, h / 2, kmax=64)
        t += h
    err = rel_l2(got, c)
    print(f"  learned Krylov vs oracle replay: rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-9


@pytest.mark.parametrize("gauge", ["velocity", "length"])
def test_step_by_step_loop_matches_oracle_run(gauge):
    """The driver's loop through the per-step entry points (qpc_step or
    qpc_step_lg, then qpc_mask, every call queued without a host wait): each
    call after the first hands over A(t_i) = the previous A(t_f), so the
    library factorises only A(t_f) and reuses the tables it kept.  A break in
    the chain (a call whose A(t_i) is not the previous A(t_f)) and a qpc_run
    in the middle make it factorise both time points again.  Against the
    oracle's own run of the same steps: <= 1e-12."""
    cfg = shape(400, 8, 80.0)
    print(f"case: step-by-step loop, {gauge} gauge, N=400 L=8", flush=True)
    o = make_oracle(cfg)
    prop = _gpu_prop(o, cfg)
    nt, a = atable(cfg)
    if gauge == "length":
        a = -np.gradient(a, cfg["dt"])  # an E(t) table of the same pulse
    step_o = (lambda c, x, y: o.step(c, x, y)) if gauge == "velocity" else (lambda c, x, y: o.step_lg(c, x, y))
    step_g = prop.propagate if gauge == "velocity" else prop.propagate_lg
    run_o = o.run if gauge == "velocity" else o.run_lg
    run_g = prop.run if gauge == "velocity" else prop.run_lg
    c = mixed_state(o, seed=61, amp=0.1)
    prop.set_state(c)
    it = nt // 2 - 40
    for k in range(30):                      # the chain: A(t_i) of a call = A(t_f) of the one before
        step_g(a[it + k], a[it + k + 1])
        prop.mask()
        step_o(c, a[it + k], a[it + k + 1])
        o.mask(c)
    it += 30
    step_g(a[it + 5], a[it + 6])             # a break in the chain
    prop.mask()
    step_o(c, a[it + 5], a[it + 6])
    o.mask(c)
    run_g(a, nt, it + 6, it + 16, use_mask=True)   # a run in between
    run_o(c, a, nt, it + 6, it + 16, use_mask=True)
    for k in range(16, 26):                  # the chain again after the run
        step_g(a[it + k], a[it + k + 1])
        prop.mask()
        step_o(c, a[it + k], a[it + k + 1])
        o.mask(c)
    err = rel_l2(prop.get_state(), c)
    print(f"  {gauge}: rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-12
