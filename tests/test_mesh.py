"""Interface/knot/stencil bookkeeping: bit-exact against the reference's recorded run
(Output/Debug/G.csv pattern, solution.csv positions, cond.txt sides) and against the
oracle's independent restatement."""
import numpy as np
import pytest

from conftest import FIX_FAC, FIX_NE, FIX_NNODE, FIX_NSUB, FIX_SW
from oracle import oracle as O


def _mesh(sw=FIX_SW, ne=FIX_NE, nsub=FIX_NSUB, nnode=FIX_NNODE, fac=FIX_FAC):
    from pdds_amd.engine import Mesh
    return Mesh(sw, ne, nsub, nnode, fac)


def knot_count(n, k):
    # (n-1) vertical lines of n(k-1)+1 knots, (n-1) rows of n(k-1)+1-(n-1) new knots
    return (n - 1) * (n * (k - 1) + 1) + (n - 1) * (n * (k - 2) + 2)


def test_fixture_knots_and_csr_pattern(golden):
    m = _mesh()
    assert m.n_knots == 165 == knot_count(4, 8)
    assert m.on_boundary.sum() == 12
    assert m.nnz == 6960
    for k in range(m.n_knots):
        assert list(m.row_cols(k)) == golden["rows"].get(k, []), k
    # boundary knots carry only the diagonal in G.csv
    for k in np.flatnonzero(m.on_boundary):
        assert k not in golden["rows"]


def test_fixture_positions(golden):
    m = _mesh()
    for r in golden["solution"]:
        k = int(r["Knot_index"])
        assert abs(m.px[k] - float(r["x"])) <= 5e-7 * max(1.0, abs(float(r["x"])))
        assert abs(m.py[k] - float(r["y"])) <= 5e-7 * max(1.0, abs(float(r["y"])))


def test_fixture_side_sizes_kinds_and_cond(golden):
    """Every (direction, kind, size) recorded in cond.txt appears in our stencils, with the
    same Frobenius condition number to the precision Eigen's FullPivLU leaves."""
    m = _mesh()
    names = ["north", "south", "east", "west"]
    ours = {}
    for k in range(m.n_knots):
        st = m.stencil(k)
        for s in range(4):
            if st["msz"][s]:
                ours.setdefault((names[s], int(st["kind"][s]), int(st["msz"][s])), []).append(st["cond"][s])
    # cond.txt was appended to by several MPI workers at once; a few lines are spliced from
    # two writers (a "dir,kind," prefix of one line and "cond,size" of another).  Such a
    # line still carries a condition number of some side of ours with the same size.
    unmatched = 0
    for d, kind, cond, size in golden["cond"]:
        if (d, kind, size) not in ours:
            unmatched += 1
            alt = [c for (dd, kk, ss), cs in ours.items() if ss == size for c in cs]
            assert min(abs(c - cond) / cond for c in alt) < 1e-4, (d, kind, size, cond)
            continue
        best = min(abs(c - cond) / cond for c in ours[(d, kind, size)])
        assert best < 1e-4, (d, kind, size, cond)
    assert unmatched <= 3
    # the knot-1 stencil from the survey: North side of 18 (kind 2), East of 11 (kind 0)
    st = m.stencil(1)
    assert list(st["msz"]) == [18, 0, 11, 0] and st["kind"][0] == 2 and st["kind"][2] == 0


@pytest.mark.parametrize("cfg", [(FIX_SW, FIX_NE, 4, 8), ((-1.0, -1.0), (1.0, 1.0), 2, 16),
                                 ((-1.0, -1.0), (1.0, 1.0), 8, 32), ((-2.0, -1.0), (3.0, 1.5), 5, 10)])
def test_product_mesh_equals_oracle_mesh(cfg):
    sw, ne, n, k = cfg
    fac = [1.0, 1.0, 1.0, 1.0]
    m = _mesh(sw, ne, n, k, fac)
    o = O.Mesh(sw, ne, n, k)
    assert m.n_knots == o.n_knots == knot_count(n, k)
    x0, y0, nx, ny, onb = o.knots()
    assert np.array_equal(m.x0, x0) and np.array_equal(m.y0, y0)
    assert np.array_equal(m.px, nx) and np.array_equal(m.py, ny)
    assert np.array_equal(m.on_boundary, onb)
    step = max(1, m.n_knots // 300)
    for q in range(0, m.n_knots, step):
        a = m.stencil(q)
        b = o.stencil(q)
        assert np.array_equal(a["rect"], b["rect"]), q
        assert np.array_equal(a["msz"], b["msz"]) and np.array_equal(a["kind"], b["kind"]), q
        assert np.array_equal(a["idx"], b["idx"]) and np.array_equal(a["px"], b["px"]) and np.array_equal(a["py"], b["py"])


def test_config2_sizes():
    """8x8 subdomains x 32 knots: 3437 knots, 28 on the boundary (3409 Monte Carlo knots)."""
    m = _mesh((-1.0, -1.0), (1.0, 1.0), 8, 32, [4.2 / 31 * 0.25 * 4] * 4)
    assert m.n_knots == 3437
    assert m.on_boundary.sum() == 28
    # every interior row references only knots of its stencil, sorted, no duplicates
    for k in range(0, m.n_knots, 37):
        c = m.row_cols(k)
        assert np.all(np.diff(c) > 0)


def test_psi_inverse_against_oracle():
    m = _mesh()
    for k in (1, 40, 100):
        st = m.stencil(k)
        off = 0
        for s in range(4):
            n = int(st["msz"][s])
            if n == 0:
                continue
            c = FIX_FAC[min(max(int(st["kind"][s]), 0), 3)]
            px, py = st["px"][off:off + n], st["py"][off:off + n]
            stat, ip, cond = O.stencil_ipsi(px, py, c)
            assert stat == 0
            assert abs(cond - st["cond"][s]) / cond < 1e-6
            # both inverses reproduce the identity to the precision cond allows
            psi = np.exp(-((px[:, None] - px[None, :]) ** 2 + (py[:, None] - py[None, :]) ** 2) / c ** 2)
            for inv in (ip, st["ipsi"][s]):
                assert np.abs(psi @ inv - np.eye(n)).max() < cond * 1e-15
            off += n


def test_psi_inverse_every_side_of_config2():
    """The 8x8 x 32-knot mesh computes Psi^-1 once per distinct side (the same interface segments
    bound several stopping rectangles) and copies it to the stencils sharing that side: every side
    of every distinct stencil must hold the inverse of its own Psi, with its own condition number."""
    from pdds_amd.workloads import relative_fac
    fac = relative_fac((-1.0, -1.0), (1.0, 1.0), 8, 32, 0.8)
    m = _mesh((-1.0, -1.0), (1.0, 1.0), 8, 32, fac)
    seen, sides, shared = set(), 0, {}
    for k in np.where(~m.on_boundary)[0]:
        st = m.stencil(int(k))
        key = st["rect"].tobytes() + st["msz"].tobytes()
        if key in seen:
            continue
        seen.add(key)
        off = 0
        for s in range(4):
            n = int(st["msz"][s])
            if n == 0:
                continue
            px, py, c = st["px"][off:off + n], st["py"][off:off + n], st["cshape"][s]
            off += n
            stat, ip, cond = O.stencil_ipsi(px, py, c)
            assert stat == 0
            assert abs(cond - st["cond"][s]) / cond < 1e-6
            psi = np.exp(-((px[:, None] - px[None, :]) ** 2 + (py[:, None] - py[None, :]) ** 2) / c ** 2)
            assert np.abs(psi @ st["ipsi"][s] - np.eye(n)).max() < cond * 1e-15
            sk = px.tobytes() + py.tobytes()
            if sk in shared:
                assert np.array_equal(shared[sk], st["ipsi"][s])
            shared[sk] = st["ipsi"][s]
            sides += 1
    assert len(seen) == 161 and sides == 556 and len(shared) == 210


def test_center_knot_of_2x2_has_no_stencil():
    m = _mesh((-1.0, -1.0), (1.0, 1.0), 2, 16, [0.25] * 4)
    assert m.n_knots == 61
    centre = [k for k in range(61) if abs(m.x0[k]) < 1e-12 and abs(m.y0[k]) < 1e-12]
    assert len(centre) == 1
    st = m.stencil(centre[0])
    assert list(st["msz"]) == [0, 0, 0, 0]
    assert np.array_equal(st["rect"], [-1.0, -1.0, 1.0, 1.0])
    assert m.row_nnz(centre[0]) == 0
