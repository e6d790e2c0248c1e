"""Real-space multigrids <-> plane-wave grids on the device
(include/grid_rs_pw.h) against the oracle (oracle/rs_pw.py, pinned by closed
forms and the adjoint identity).  Floating point through FFTs of different
libraries (hipFFT here, numpy's pocketfft in the oracle): results agree to
1e-12 of the largest value; the copies (transfer) are exact.  At the full
size of the 64-H2O and 1024-atom grids the adjoint identity
<rho(rs), v> = sum_l <rs_l, v_l> holds to 1e-10 relative."""
import numpy as np
import pytest

from cp2k_amd import grid, synth
from oracle import oracle as O
from oracle import rs_pw as R

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module", autouse=True)
def lib():
    from cp2k_amd import build
    if not build.LIB.exists():
        build.build()
    grid.set_precision(grid.FP64)
    return grid.lib()


def _close(a, b, tol=TOL):
    return np.max(np.abs(a - b)) <= tol * max(1.0, np.max(np.abs(b)))


@pytest.mark.parametrize("shapes", [[(27, 25, 21), (15, 15, 9), (9, 7, 5)],
                                    [(24, 20, 18), (12, 10, 8), (6, 6, 4)],
                                    [(16, 16, 16)]])
def test_density_and_potential_vs_oracle(shapes):
    print("rs_pw", shapes, flush=True)
    rng = np.random.default_rng(sum(map(sum, shapes)))
    rs = [rng.normal(size=s) for s in shapes]
    rho_r, rho_g = grid.density_rs2pw([grid.DeviceArray.from_numpy(g) for g in rs])
    wr, wg = R.density_rs2pw(rs)
    assert _close(rho_r.numpy(), wr)
    g = rho_g.numpy()
    assert _close(g[..., 0] + 1j * g[..., 1], wg)
    v = rng.normal(size=shapes[0])
    got = grid.potential_pw2rs(grid.DeviceArray.from_numpy(v), shapes)
    for a, b in zip(got, R.potential_pw2rs(v, shapes)):
        assert _close(a.numpy(), b)


def test_transfer_halo():
    rng = np.random.default_rng(5)
    pw = rng.normal(size=(10, 9, 12))
    rs = grid.transfer_pw2rs(grid.DeviceArray.from_numpy(pw), (3, 2, 1))
    assert np.array_equal(rs.numpy(), R.transfer_pw2rs(pw, (3, 2, 1)))
    back = grid.transfer_rs2pw(rs, (3, 2, 1))
    assert np.array_equal(back.numpy(), pw)


def test_collocate_then_rs2pw():
    """The device chain collocate -> density_rs2pw against the oracle's
    collocation followed by the oracle's rs2pw."""
    args, pab, V = synth.config("h2o2", seed=2, n=24)
    print("collocate -> rs2pw", args.ntasks, flush=True)
    gt = grid.GridTaskList(args)
    pab_d = grid.DeviceArray.from_numpy(pab)
    rho_lv = [grid.DeviceArray(s) for s in args.grid_shapes()]
    gt.collocate_dev(100, pab_d, rho_lv)
    grid.synchronize()
    rho_r, rho_g = grid.density_rs2pw(rho_lv)
    want_r, want_g = R.density_rs2pw(O.TaskList(args).collocate(100, pab))
    assert _close(rho_r.numpy(), want_r, 1e-11)
    g = rho_g.numpy()
    assert _close(g[..., 0] + 1j * g[..., 1], want_g, 1e-11)


@pytest.mark.parametrize("levels", [[128, 72, 42, 24], [200, 112, 64, 36]])
def test_adjoint_identity_full_size(levels):
    shapes = [(n, n, n) for n in levels]
    print("rs_pw adjoint", levels, flush=True)
    rng = np.random.default_rng(levels[0])
    rs = [rng.normal(size=s) for s in shapes]
    v = rng.normal(size=shapes[0])
    rho_r, _ = grid.density_rs2pw([grid.DeviceArray.from_numpy(g) for g in rs], want_g=False)
    vl = grid.potential_pw2rs(grid.DeviceArray.from_numpy(v), shapes)
    lhs = float(np.sum(rho_r.numpy() * v))
    rhs = sum(float(np.sum(g * w.numpy())) for g, w in zip(rs, vl))
    assert abs(lhs - rhs) <= 1e-10 * (abs(lhs) + 1.0)
    # charge: the mean of rho_r is the sum of the levels' means
    assert abs(rho_r.numpy().mean() - sum(g.mean() for g in rs)) < 1e-12
