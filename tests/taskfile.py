"""Reader for CP2K "#Grid task v10" known-answer files.

Field order and meaning follow src/grid/grid_replay.c:242-344.  The reader is
independent of the reference's C parser and returns numpy arrays.
"""
import gzip
import pathlib

import numpy as np

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden" / "sample_tasks"


def ncoset(l):
    return 0 if l < 0 else (l + 1) * (l + 2) * (l + 3) // 6


class Task:
    """One single-task fixture: inputs plus the reference's outputs."""


def _open(path):
    path = pathlib.Path(path)
    if path.suffix == ".gz":
        return gzip.open(path, "rt")
    return open(path, "rt")


def load(path):
    with _open(path) as fh:
        lines = fh.read().split("\n")
    pos = 0

    def nxt():
        nonlocal pos
        line = lines[pos]
        pos += 1
        return line.split()

    assert lines[0] == "#Grid task v10", lines[0]
    pos = 1
    t = Task()
    t.name = pathlib.Path(path).name.replace(".task", "").replace(".gz", "")

    def scalar(key, conv):
        f = nxt()
        assert f[0] == key, (key, f)
        return conv(f[1])

    def vec(key, conv):
        f = nxt()
        assert f[0] == key, (key, f)
        return np.array([conv(x) for x in f[1:4]])

    def mat(key):
        m = np.zeros((3, 3))
        for i in range(3):
            f = nxt()
            assert f[0] == key and int(f[1]) == i, (key, f)
            m[i] = [float(x) for x in f[2:5]]
        return m

    t.orthorhombic = bool(scalar("orthorhombic", int))
    t.border_mask = scalar("border_mask", int)
    t.func = scalar("func", int)
    t.la_max = scalar("la_max", int)
    t.la_min = scalar("la_min", int)
    t.lb_max = scalar("lb_max", int)
    t.lb_min = scalar("lb_min", int)
    t.zeta = scalar("zeta", float)
    t.zetb = scalar("zetb", float)
    t.rscale = scalar("rscale", float)
    t.dh = mat("dh")
    t.dh_inv = mat("dh_inv")
    t.ra = vec("ra", float)
    t.rab = vec("rab", float)
    t.npts_global = vec("npts_global", int).astype(np.int32)
    t.npts_local = vec("npts_local", int).astype(np.int32)
    t.shift_local = vec("shift_local", int).astype(np.int32)
    t.border_width = vec("border_width", int).astype(np.int32)
    t.radius = scalar("radius", float)
    t.o1 = scalar("o1", int)
    t.o2 = scalar("o2", int)
    t.n1 = scalar("n1", int)
    t.n2 = scalar("n2", int)
    t.pab = np.zeros((t.n2, t.n1))
    for i in range(t.n2):
        for j in range(t.n1):
            f = nxt()
            assert f[0] == "pab" and int(f[1]) == i and int(f[2]) == j
            t.pab[i, j] = float(f[3])
    nx, ny, nz = t.npts_local
    t.grid = np.zeros((nz, ny, nx))
    nnz = scalar("ngrid_nonzero", int)
    for _ in range(nnz):
        f = nxt()
        assert f[0] == "grid"
        i, j, k = int(f[1]), int(f[2]), int(f[3])
        t.grid[k, j, i] = float(f[4])
    t.hab = np.zeros((t.n2, t.n1))
    for i in range(t.o2, ncoset(t.lb_max) + t.o2):
        for j in range(t.o1, ncoset(t.la_max) + t.o1):
            f = nxt()
            assert f[0] == "hab" and int(f[1]) == i and int(f[2]) == j
            t.hab[i, j] = float(f[3])
    t.forces = np.stack([vec("force_a", float), vec("force_b", float)])
    t.virial = mat("virial")
    assert lines[pos] == "#THE_END", lines[pos]
    # integrate replays use compute_tau iff func == DADB (grid_replay.c:252)
    t.compute_tau = t.func == 200
    return t


def all_fixtures():
    return sorted(GOLDEN.glob("*.task.gz"))
