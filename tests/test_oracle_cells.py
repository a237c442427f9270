"""Pins the oracle's simulation cell (NQCBase's PeriodicCell behind DynamicsUtils/callbacks.jl:4-22) and its
classical VelocityVerlet path (Simulation{Classical}, classical.jl:80-86):
  - the wrap maps every periodic fractional coordinate into [0, 1) and moves each atom by an integer
    combination of the lattice vectors; inside configurations are left alone; non-periodic dimensions are
    kept; a hexagonal cell (docs/src/initialconditions/ebk.md:47) checks the use of the inverse, not the
    transpose;
  - VelocityVerlet on a harmonic well follows its exact discrete solution x_n = x0 cos(n theta) +
    dt v0 sin(n theta) / sin(theta), cos(theta) = 1 - (omega dt)^2 / 2;
  - a free particle in a periodic cell is the unwrapped straight line taken modulo the cell;
  - the reference's own cell test (test/Dynamics/cell_boundary_callback.jl): a carbon atom in a unit cell
    leaves it without the callback and stays in it with CellBoundaryCallback;
  - TerminateCellCallback ends each trajectory at the first step that leaves the cell."""
import numpy as np
import pytest

HEX = np.array([[5.883, -2.942, 0.0], [0.0, 5.095, 0.0], [0.0, 0.0, 20.0]])  # columns: lattice vectors
C_MASS = 12.011 * 1822.888486


def test_wrap_one_dimension(oracle):
    c = oracle.make_cell(1, [[2.5]])
    for x in (0.0, 1.3, 2.5):
        out, y = oracle.cell_apply(c, [x])
        assert not out and y[0] == x                       # inside (0 <= f <= 1) is left alone
    for x in (-0.1, 2.6, 7.9, -11.2):
        out, y = oracle.cell_apply(c, [x])
        assert out and 0.0 <= y[0] < 2.5
        k = (x - y[0]) / 2.5
        assert abs(k - round(k)) < 1e-12 and round(k) != 0
    t = oracle.make_cell(2, [[2.5]])
    out, y = oracle.cell_apply(t, [2.6])
    assert out and y[0] == 2.6                             # TerminateCellCallback does not move atoms


def test_wrap_hexagonal_cell_and_periodicity(oracle):
    rng = np.random.default_rng(5)
    inv = np.linalg.inv(HEX)
    for per in ([1, 1, 1], [1, 1, 0]):
        c = oracle.make_cell(1, HEX, per)
        for _ in range(200):
            x = rng.uniform(-15, 25, size=6)               # two atoms, coordinate fastest
            out, y = oracle.cell_apply(c, x)
            f_in = (inv @ x.reshape(2, 3).T).T
            assert out == bool(np.any((f_in < 0) | (f_in > 1)))
            if not out:
                assert np.array_equal(y, x)
                continue
            f = (inv @ y.reshape(2, 3).T).T
            for j in range(3):
                if per[j]:
                    assert np.all((f[:, j] >= -1e-12) & (f[:, j] < 1 + 1e-12))
                else:
                    assert np.allclose(f[:, j], f_in[:, j], atol=1e-12)   # not wrapped
            n = f_in - f                                                  # lattice translation per atom
            assert np.allclose(n, np.round(n), atol=1e-9)
    # a cell built from the transpose (vectors as rows) wraps differently: the check above would fail
    bad = oracle.make_cell(1, HEX.T)
    fails = 0
    for _ in range(50):
        _, yb = oracle.cell_apply(bad, rng.uniform(-15, 25, size=6))
        fb = (inv @ yb.reshape(2, 3).T).T
        fails += bool(np.any((fb < -1e-12) | (fb >= 1 + 1e-12)))
    assert fails > 10


def test_velocity_verlet_harmonic_closed_form(oracle):
    m = oracle.make_model(oracle.MODEL_HARMONIC, 1, 1, [1.0, 0.7, 0.3])    # V = 0.49 (x - 0.3)^2 / 2
    mass, dt, n = 1.0, 0.25, 400
    x0, v0 = np.array([1.1, -0.4]), np.array([0.2, 0.9])
    out = oracle.classical_run(m, mass, x0, v0, dt=dt, nsteps=n, save_every=1)
    w2 = 0.49 / mass
    th = np.arccos(1 - w2 * dt * dt / 2)
    k = np.arange(n + 1)
    for i in range(2):
        ref = 0.3 + (x0[i] - 0.3) * np.cos(k * th) + dt * v0[i] * np.sin(k * th) / np.sin(th)
        assert np.allclose(out["save_x"][i], ref, atol=1e-12)
    assert np.all(out["steps"] == n)
    e = out["ekin"] + out["epot"]
    assert np.ptp(e, axis=1).max() < 0.01 * e[:, 0].max()               # bounded energy error at dt w = 0.175


def test_free_particle_in_a_periodic_cell(oracle):
    m = oracle.make_model(oracle.MODEL_HARMONIC, 1, 1, [0.0, 1.0, 0.0])    # V = 0: a free particle
    c = oracle.make_cell(1, [[1.7]])
    x0, v0 = np.array([0.2, 1.5, 0.9]), np.array([0.13, -0.07, 0.031])
    n, dt = 500, 0.5
    out = oracle.classical_run(m, 1.0, x0, v0, dt=dt, nsteps=n, save_every=1, cell=c)
    free = oracle.classical_run(m, 1.0, x0, v0, dt=dt, nsteps=n, save_every=1)
    assert np.allclose(np.mod(free["save_x"], 1.7)[:, -1], out["x"], atol=1e-12)
    # saves precede the callbacks: a saved position is the free line modulo the cell, or that plus one cell
    # at the step that left it
    d = (free["save_x"] - out["save_x"]) / 1.7
    assert np.allclose(d, np.round(d), atol=1e-9)
    assert np.abs(np.diff(np.round(d), axis=1)).max() <= 1


def test_reference_cell_boundary_callback_test(oracle):
    """test/Dynamics/cell_boundary_callback.jl: Atoms([:C]), Harmonic(), PeriodicCell(hcat(1)), v = 1, r = 0,
    tspan (0, 300), dt = 0.1, OutputPosition saved every step."""
    m = oracle.make_model(oracle.MODEL_HARMONIC, 1, 1, [1.0, 1.0, 0.0])
    free = oracle.classical_run(m, C_MASS, [0.0], [1.0], dt=0.1, nsteps=3000, save_every=1)
    assert free["save_x"].max() > 1.0                                     # the atom leaves the cell
    c = oracle.make_cell(1, [[1.0]])
    out = oracle.classical_run(m, C_MASS, [0.0], [1.0], dt=0.1, nsteps=3000, save_every=1, cell=c)
    assert np.isclose(out["save_x"].min(), 0, atol=1e-4) or out["save_x"].min() >= 0
    assert 0.0 <= out["x"][0] <= 1.0


def test_terminate_cell_callback(oracle):
    m = oracle.make_model(oracle.MODEL_HARMONIC, 1, 1, [0.0, 1.0, 0.0])
    c = oracle.make_cell(2, [[1.0]])
    x0, v0 = np.array([0.5, 0.5, 0.95]), np.array([0.0123, -0.0031, 0.0])
    out = oracle.classical_run(m, 1.0, x0, v0, dt=1.0, nsteps=400, cell=c)
    # the first step past the boundary: 0.5 + 0.0123 k > 1 at k = 41, 0.5 - 0.0031 k < 0 at k = 162
    assert out["steps"].tolist() == [41, 162, 400]
    assert out["x"][0] > 1.0 and out["x"][1] < 0.0


@pytest.mark.parametrize("mode", [1, 2])
def test_fssh_cell_callbacks_follow_the_hop(oracle, key, mode):
    """FSSH on Tully's first model in the cell [0, 8): runs with CellBoundaryCallback end inside it and equal
    the plain runs up to the first step that leaves it; runs with TerminateCellCallback stop at that step."""
    m = oracle.make_model(oracle.MODEL_TULLY1, 2, 1, [0.01, 1.6, 0.005, 1.0])
    ntraj = 64
    rng = np.random.default_rng(2)
    r0 = rng.uniform(1.0, 7.0, size=(ntraj, 1))
    v0 = rng.uniform(-30, 30, size=(ntraj, 1)) / 2000
    sig0 = np.zeros((ntraj, 2, 2), complex); sig0[:, 0, 0] = 1
    run = lambda cell: oracle.run_ensemble(0, m, 2000.0, r0, v0, sig0, state0=np.zeros(ntraj, np.int32), dt=1.0,
                                           nsteps=600, key=key, save_every=1, nsave=601, cell=cell)
    plain, out = run(None), run(oracle.make_cell(mode, [[8.0]]))
    x = plain["pop_ad"]  # any per-save array marks the slots; the positions come from a shorter plain run
    first = []
    for i in range(ntraj):   # first step whose end position leaves [0, 8]
        p = oracle.run_ensemble(0, m, 2000.0, r0[i:i + 1], v0[i:i + 1], sig0[i:i + 1], state0=np.zeros(1, np.int32),
                                dt=1.0, nsteps=600, key=key, traj_ids=[i], term_x=0.0, term_kind=3, term_dof=0)
        q = oracle.run_ensemble(0, m, 2000.0, r0[i:i + 1], v0[i:i + 1], sig0[i:i + 1], state0=np.zeros(1, np.int32),
                                dt=1.0, nsteps=600, key=key, traj_ids=[i], term_x=8.0, term_kind=2, term_dof=0)
        k = min(p["steps"][0], q["steps"][0])
        first.append(k)
    first = np.array(first)
    assert x.shape[0] == ntraj and np.any(first < 600)
    if mode == 2:
        assert np.array_equal(out["steps"], first)
    else:
        assert np.all((out["r"][:, 0] >= 0) & (out["r"][:, 0] <= 8))
        stay = first == 600
        assert np.array_equal(out["r"][stay], plain["r"][stay]) and np.array_equal(out["state"][stay], plain["state"][stay])
