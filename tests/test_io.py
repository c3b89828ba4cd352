"""The driver's output layout (binsrc/sphTDSE.cpp:837-864 and the save_*
functions it calls) and load_initial (:119-160) over lib/libqpc_io.so: every
dataset name and element type the reference writes for a z-polarised run
with spectra, t-SURFF, coefficients and recorded observables; values round
trip; load_initial's checks and return codes; a file written from a device
run reloads into a context and propagates on."""
import os

import numpy as np
import pytest

qpc_io = pytest.importorskip("qpc_io")
if not os.path.exists(qpc_io.IO_PATH):
    pytest.skip("libqpc_io.so not built (no HDF5)", allow_module_level=True)

# (name, element type) as sphTDSE.cpp writes them: size_t -> U64, long -> I64,
# int -> I32 (the dataset "gauge" has no leading slash, :551)
LAYOUT = [("/para_dt", qpc_io.F64), ("/para_nt", qpc_io.U64), ("/nr", qpc_io.U64), ("/nl", qpc_io.U64),
          ("/nm", qpc_io.U64), ("/m0", qpc_io.I64), ("/mr", qpc_io.U64), ("/rmin", qpc_io.F64),
          ("/rmax", qpc_io.F64), ("/fieldz_0", qpc_io.F64), ("/pr", qpc_io.F64), ("/pt", qpc_io.F64),
          ("/pp", qpc_io.F64), ("/tsurff_zc", qpc_io.F64), ("/tsurff_r0", qpc_io.F64),
          ("/knot_sequence", qpc_io.F64), ("/coef1", qpc_io.F64), ("/lmd1", qpc_io.F64), ("/pmd1", qpc_io.F64),
          ("/pmd2", qpc_io.F64), ("/proj_val", qpc_io.F64), ("/proj_nnz", qpc_io.U64), ("/t_list", qpc_io.F64),
          ("/z_list", qpc_io.F64), ("gauge", qpc_io.I32), ("/fz_list", qpc_io.F64)]


def _run_data(N=120, p=7, L=4, M=1, m0=0, rng=None):
    rng = rng or np.random.default_rng(3)
    c = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    return dict(N=N, n_rank=p, L=L, M=M, m0=m0, rmin=0.0, rmax=50.0, para_dt=0.05, para_nt=1234,
                knots=np.linspace(0, 50, N + p + 1), fieldz=[np.array([0.0534, 0.057, 4.0, 0.0, 0.0])],
                coef=c(M * L * N), lmd1=c(7, M, L), pmd1=c(7, 3, 2), pmd2=c(7, 3, 2),
                axes=(np.linspace(0.1, 2, 7), np.linspace(0, np.pi, 3), np.linspace(0, 2 * np.pi, 2)),
                tsurff=(0.0, 40.0), proj=(c(9), np.array([3, 2, 4, 0])), t_list=np.arange(10) * 0.05,
                z_list=rng.standard_normal(10), fz_list=rng.standard_normal(10), gauge=0)


def test_layout_names_types_and_values(tmp_path):
    path = str(tmp_path / "run.h5")
    d = _run_data()
    qpc_io.save_realtime(path, **d)
    with qpc_io.H5File(path, "r") as f:
        names = f.names()
        assert sorted(names) == sorted(n.lstrip("/") for n, _ in LAYOUT)
        for name, typ in LAYOUT:
            n, t = f.info(name)
            assert t == typ, name
        assert f.load("/nr", qpc_io.U64)[0] == 120 and f.load("/m0", qpc_io.I64)[0] == 0
        assert f.load("/para_nt", qpc_io.U64)[0] == 1234 and f.load("gauge", qpc_io.I32)[0] == 0
        assert np.array_equal(f.load("/coef1", qpc_io.F64).view(np.complex128), d["coef"])
        assert np.array_equal(f.load("/pmd2", qpc_io.F64).view(np.complex128), d["pmd2"].ravel())
        assert np.array_equal(f.load("/lmd1", qpc_io.F64).view(np.complex128), d["lmd1"].ravel())
        assert np.array_equal(f.load("/proj_nnz", qpc_io.U64), [3, 2, 4, 0])
        assert np.array_equal(f.load("/fieldz_0", qpc_io.F64), d["fieldz"][0])
        assert np.array_equal(f.load("/knot_sequence", qpc_io.F64), d["knots"])
        # a size mismatch and a missing dataset are errors, not silent reads
        with pytest.raises(qpc_io.H5Error):
            f.load("/nr", qpc_io.U64, 2)
        with pytest.raises(qpc_io.H5Error):
            f.load("/wave1", qpc_io.F64)


def test_load_initial_codes_and_channel_placement(tmp_path):
    path = str(tmp_path / "gs.h5")
    d = _run_data(N=100, L=3, M=1, m0=0)
    qpc_io.save_realtime(path, **d)
    # same layout: the coefficients as written
    rc, c = qpc_io.load_initial(path, 100, 7, 3, 1, 0, 0.0, 50.0)
    assert rc == 0 and np.array_equal(c, d["coef"])
    # more l channels and m channels in the run: the file's block in place, zeros elsewhere
    rc, c = qpc_io.load_initial(path, 100, 7, 5, 3, -1, 0.0, 50.0)
    assert rc == 0
    c = c.reshape(3, 5, 100)
    assert np.array_equal(c[1, :3], d["coef"].reshape(3, 100)) and not c[0].any() and not c[2].any()
    assert not c[1, 3:].any()
    # the reference's return codes
    assert qpc_io.load_initial(path, 101, 7, 3, 1, 0, 0.0, 50.0)[0] == -1
    assert qpc_io.load_initial(path, 100, 7, 3, 1, 0, 0.0, 60.0)[0] == -1
    assert qpc_io.load_initial(path, 100, 8, 3, 1, 0, 0.0, 50.0)[0] == -2
    assert qpc_io.load_initial(path, 100, 7, 3, 1, 1, 0.0, 50.0)[0] == -3
    assert qpc_io.load_initial(path, 100, 7, 2, 1, 0, 0.0, 50.0)[0] == -4


def test_setup_ground_state_round_trip(tmp_path):
    """A ground state from the host set-up written in the driver's layout
    (with the set-up's own knot sequence) and read back by load_initial."""
    import qpc_gpu
    st = qpc_gpu.Setup(300, 7, 4, 1, 0, 0.05, 0.0, 100.0, 1.0, 1.0, 90.0, 100.0, 0.125)
    e, c0 = st.eigenstate(0, 0, 0)
    knots = st.get(qpc_gpu.SETUP_KNOTS)
    assert len(knots) == 300 + 7 + 1 and knots[0] == 0.0 and knots[-1] == 100.0
    assert np.all(np.diff(knots) >= 0)
    path = str(tmp_path / "gs.h5")
    qpc_io.save_realtime(path, N=300, n_rank=7, L=4, rmax=100.0, para_dt=0.05, para_nt=0, knots=knots, coef=c0)
    rc, c = qpc_io.load_initial(path, 300, 7, 4, 1, 0, 0.0, 100.0)
    assert rc == 0 and np.array_equal(c, c0)
