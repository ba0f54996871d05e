"""Several blocks (src/sub_init.f90:102-321,636-823; src/GKUA_boundary.f90:400-633;
src/Backup_write_read.f90:56-209) and the symmetry BC (src/sub_boundary.f90:410-471).

The node copy and the inner-face cell exchange are pinned by geometry: one
skewed block is split in two along i or j (the second part optionally turned
by 180 degrees, which makes the faces meet in reversed orientation), and every
ghost node or ghost cell that an inner face receives must be the node or cell
of the whole block at the same place."""
import math

import numpy as np
import pytest

from helpers import load_oracle, load_pkg


def _skewed(ni=9, nj=8):
    s = np.linspace(0.0, 1.0, ni) ** 1.2
    t = np.linspace(0.0, 1.0, nj) ** 1.1
    S, T = np.meshgrid(s, t, indexing="ij")
    return np.asfortranarray(S + 0.25 * T ** 2 + 0.1 * S * T), np.asfortranarray(T + 0.15 * S ** 2)


def _whole_bcs(ni, nj):
    cs = load_pkg().cases
    return [cs.bc_from_range(1, ni, 1, 1, 2, ni, nj), cs.bc_from_range(ni, ni, 1, nj, 4, ni, nj),
            cs.bc_from_range(1, ni, nj, nj, 2, ni, nj), cs.bc_from_range(1, 1, 1, nj, 4, ni, nj)]


def _node_map(ni, nj, m, axis, rot):
    """Block k's node (i, j) -> the whole block's node, written from the split's
    geometry (the second part starts at node m; turned, its indices run back)."""
    n = (ni, nj)[axis]
    n2 = n - m + 1

    def f(k, i, j):
        if k == 0:
            return i, j
        if axis == 0:
            return (m + n2 - i, nj + 1 - j) if rot else (i + m - 1, j)
        return (ni + 1 - i, m + n2 - j) if rot else (i, j + m - 1)
    return f


def _cell_map(nmap):
    """Block cell (i, j) spans nodes i..i+1, j..j+1; the whole cell is the one
    whose lowest node indices those nodes map to."""
    def f(k, i, j):
        a = [nmap(k, i + di, j + dj) for di in (0, 1) for dj in (0, 1)]
        return min(p[0] for p in a), min(p[1] for p in a)
    return f


CASES = [(0, False), (0, True), (1, False), (1, True)]


@pytest.mark.parametrize("axis,rot", CASES)
def test_split_blocks_and_inner_ghost_nodes_continue_the_whole_mesh(axis, rot):
    """Interior nodes of both parts are the whole block's; the LAP = 2 ghost
    nodes of each inner face (Update_coordinate_buffer_onemesh) are the whole
    block's nodes beyond the split line, and so are the ghost cell centres."""
    mb = load_pkg().multiblock
    cs = load_pkg().cases
    xn, yn = _skewed()
    ni, nj = xn.shape
    m = 4 if axis == 0 else 5
    nodes, bcs = mb.split_nodes(xn, yn, _whole_bcs(ni, nj), m, axis, rot)
    links = mb.get_neighb_info(bcs)
    assert [(ln.block, ln.nblock, ln.orient) for ln in links] == [(1, 2, 0 if rot else 2), (2, 1, 0 if rot else 2)]
    meshes, _ = mb.build_block_meshes(nodes, bcs)
    whole = cs.build_mesh(xn, yn, _whole_bcs(ni, nj))
    wx1, wy1 = whole.x1y1()
    nmap = _node_map(ni, nj, m, axis, rot)
    cmap = _cell_map(nmap)
    for k in (0, 1):
        bx, by = nodes[k]
        for i in range(1, bx.shape[0] + 1):
            for j in range(1, bx.shape[1] + 1):
                wi, wj = nmap(k, i, j)
                assert bx[i - 1, j - 1] == xn[wi - 1, wj - 1] and by[i - 1, j - 1] == yn[wi - 1, wj - 1]
        me = meshes[k]
        bni, bnj = bx.shape
        inner = [b for b in bcs[k] if b.code >= 1]
        assert len(inner) == 1
        b = inner[0]
        x1, y1 = me.x1y1()
        for l in (1, 2):
            # ghost nodes of the face
            for p in range((b.jst if b.face in (1, 3) else b.ist), (b.jend if b.face in (1, 3) else b.iend) + 1):
                gi, gj = {1: (1 - l, p), 3: (bni + l, p), 2: (p, 1 - l), 4: (p, bnj + l)}[b.face]
                wi, wj = nmap(k, gi, gj)
                assert me.x[gi + 1, gj + 1] == xn[wi - 1, wj - 1]
                assert me.y[gi + 1, gj + 1] == yn[wi - 1, wj - 1]
            # ghost cell centres of the face
            for p in range((b.jst if b.face in (1, 3) else b.ist), (b.jend if b.face in (1, 3) else b.iend)):
                ci, cj = {1: (1 - l, p), 3: (bni + l - 1, p), 2: (p, 1 - l), 4: (p, bnj + l - 1)}[b.face]
                wi, wj = cmap(k, ci, cj)
                assert math.isclose(x1[ci + 1, cj + 1], wx1[wi + 1, wj + 1], rel_tol=1e-14, abs_tol=1e-14)
                assert math.isclose(y1[ci + 1, cj + 1], wy1[wi + 1, wj + 1], rel_tol=1e-14, abs_tol=1e-14)


def _block_arrays(k, nodes, cmap, W, WS, nv=(3, 2), sentinel=-7.0):
    """FRD and S_GKUA of block k in the reference layout: interior cells from
    the whole block's arrays W, WS, everything else ``sentinel``."""
    bni, bnj = nodes[k][0].shape
    frd = np.full((2, bni + 7, bnj + 7) + nv, sentinel, order="F")
    s = np.full((11, bni + 7, bnj + 7), sentinel, order="F")
    for i in range(1, bni):
        for j in range(1, bnj):
            wi, wj = cmap(k, i, j)
            frd[:, i + 3, j + 3] = W[:, wi + 3, wj + 3]
            s[:, i + 3, j + 3] = WS[:, wi + 3, wj + 3]
    return frd, s


@pytest.mark.parametrize("axis,rot", CASES)
def test_inner_exchange_fills_ghost_cells_with_the_whole_blocks_cells(axis, rot):
    """GKUA_inner_mesco_boundary + GKUA_inner_macro_boundary (oracle restatement):
    after both exchanges the two ghost layers of each inner face hold FRD(1:2)
    for every velocity and S_GKUA(1:10) of the whole block's cells at the
    same place; S_GKUA(11) and every other ghost cell keep their values."""
    mb = load_pkg().multiblock
    orc = load_oracle()
    xn, yn = _skewed()
    ni, nj = xn.shape
    m = 4 if axis == 0 else 5
    nodes, bcs = mb.split_nodes(xn, yn, _whole_bcs(ni, nj), m, axis, rot)
    links = mb.get_neighb_info(bcs)
    rng = np.random.default_rng(11 + axis + 2 * rot)
    W = np.asfortranarray(rng.standard_normal((2, ni + 7, nj + 7, 3, 2)))
    WS = np.asfortranarray(rng.standard_normal((11, ni + 7, nj + 7)))
    cmap = _cell_map(_node_map(ni, nj, m, axis, rot))
    arrs = [_block_arrays(k, nodes, cmap, W, WS) for k in (0, 1)]
    before = [(f.copy(), s.copy()) for f, s in arrs]
    for ln in links:
        fd, sd = arrs[ln.block - 1]
        fs, ss = arrs[ln.nblock - 1]
        orc.inner_exchange_arrays(fd, sd, bcs[ln.block - 1][ln.ksub - 1], fs, ss,
                                  bcs[ln.nblock - 1][ln.msub - 1], ln.orient)
    for k in (0, 1):
        f, s = arrs[k]
        f0, s0 = before[k]
        bni, bnj = nodes[k][0].shape
        b = [c for c in bcs[k] if c.code >= 1][0]
        touched = np.zeros((bni + 7, bnj + 7), bool)
        for l in (1, 2):
            for p in range((b.jst if b.face in (1, 3) else b.ist), (b.jend if b.face in (1, 3) else b.iend)):
                ci, cj = {1: (1 - l, p), 3: (bni + l - 1, p), 2: (p, 1 - l), 4: (p, bnj + l - 1)}[b.face]
                wi, wj = cmap(k, ci, cj)
                np.testing.assert_array_equal(f[:, ci + 3, cj + 3], W[:, wi + 3, wj + 3])
                np.testing.assert_array_equal(s[:10, ci + 3, cj + 3], WS[:10, wi + 3, wj + 3])
                assert s[10, ci + 3, cj + 3] == -7.0
                touched[ci + 3, cj + 3] = True
        np.testing.assert_array_equal(f[:, ~touched], f0[:, ~touched])
        np.testing.assert_array_equal(s[:, ~touched], s0[:, ~touched])


def test_get_neighb_info_orientation_rules():
    """Get_neighb_info: equal range -> orient 2; i or j reversed -> orient 0;
    the last matching subface wins; no match is an error."""
    mb = load_pkg().multiblock
    cs = load_pkg().cases
    a = cs.bc_from_range(5, 5, 1, 4, -1, 5, 4)
    a.code, a.nrange = 2, (1, 1, 1, 4)
    b_wall = cs.bc_from_range(1, 5, 1, 1, 2, 5, 4)
    b_face = cs.bc_from_range(1, 1, 1, 4, 2, 5, 4)
    assert mb.get_neighb_info([[a], [b_wall, b_face]]) == [mb.Link(1, 1, 2, 2, 2)]
    a.nrange = (1, 1, 4, 1)
    assert mb.get_neighb_info([[a], [b_wall, b_face]])[0].orient == 0
    c = cs.bc_from_range(1, 5, 4, 4, -1, 5, 4)
    c.code, c.nrange = 2, (5, 1, 1, 1)
    assert mb.get_neighb_info([[c], [b_wall, b_face]]) == [mb.Link(1, 1, 2, 1, 0)]
    dup = cs.bc_from_range(1, 1, 1, 4, 4, 5, 4)
    a.nrange = (1, 1, 1, 4)
    assert mb.get_neighb_info([[a], [b_face, b_wall, dup]])[0].msub == 3
    a.nrange = (1, 1, 2, 4)
    with pytest.raises(ValueError):
        mb.get_neighb_info([[a], [b_wall, b_face]])


GENERIC_INP_TWO_BLOCKS = """ two blocks side by side
2
5 4
2
4
 1 5 1 1 2
 5 5 1 4 -1
 1 1 1 4 2
 1 5 4 4 2
 1 1 1 4 4
6 4
2
4
 1 6 1 1 2
 6 6 1 4 4
 1 6 4 4 2
 1 1 1 4 -1
 5 5 1 4 1
"""


def test_generic_inp_with_inner_faces(tmp_path):
    """read_bcin: a code of -1 is an inner face (neighb = 1 after negation) and
    the next record gives the neighbour's range and block number."""
    mb = load_pkg().multiblock
    p = tmp_path / "generic.inp"
    p.write_text(GENERIC_INP_TWO_BLOCKS)
    bcs = mb.read_generic_inp_blocks(p, [(5, 4), (6, 4)])
    assert [b.code for b in bcs[0]] == [-2, 2, -2, -4]
    assert bcs[0][1].face == 3 and bcs[0][1].nrange == (1, 1, 1, 4)
    assert [b.code for b in bcs[1]] == [-2, -4, -2, 1]
    assert bcs[1][3].face == 1 and bcs[1][3].nrange == (5, 5, 1, 4)
    links = mb.get_neighb_info(bcs)
    assert links == [mb.Link(1, 2, 2, 4, 2), mb.Link(2, 4, 1, 2, 2)]
    with pytest.raises(ValueError):
        mb.read_generic_inp_blocks(p, [(5, 4), (7, 4)])


@pytest.mark.parametrize("rot", [False, True])
def test_generic_files_of_several_blocks_round_trip(tmp_path, rot):
    mb = load_pkg().multiblock
    cs = load_pkg().cases
    xn, yn = _skewed()
    nodes, bcs = mb.split_nodes(xn, yn, _whole_bcs(*xn.shape), 4, 0, rot)
    dims = [n[0].shape for n in nodes]
    mb.write_generic_dat_blocks(tmp_path / "generic.dat", nodes)
    mb.write_generic_inp_blocks(tmp_path / "generic.inp", dims, bcs)
    back = cs.read_generic_dat_blocks(tmp_path / "generic.dat")
    for (a, b), (c, d) in zip(nodes, back):
        np.testing.assert_array_equal(a, c)
        np.testing.assert_array_equal(b, d)
    bb = mb.read_generic_inp_blocks(tmp_path / "generic.inp", dims)
    for l1, l2 in zip(bcs, bb):
        assert [(b.face, b.ist, b.iend, b.jst, b.jend, b.code, b.nrange) for b in l1] == \
               [(b.face, b.ist, b.iend, b.jst, b.jend, b.code, b.nrange) for b in l2]
    with pytest.raises(ValueError):
        cs.read_generic_dat(tmp_path / "generic.dat")


def test_list_directed_reader_records_and_repeats():
    """Each READ starts on a new record and drops what is left of its last one;
    r*c repeats a value."""
    cs = load_pkg().cases
    f = cs.ListDirected("1 99\n2 3 7\n4*0.5 1d0\n2.0\n")
    assert f.read(1) == ["1"]
    assert f.read(2) == ["2", "3"]
    assert f.read(5) == ["0.5"] * 4 + ["1d0"]
    assert f.read(1) == ["2.0"]
    with pytest.raises(EOFError):
        f.read(1)


def test_restart_file_holds_one_record_per_block(tmp_path):
    """backup_Mesco_write / backup_Mesco_read: the blocks' FRD records follow
    each other in block order in the same file."""
    rs = load_pkg().restart
    rng = np.random.default_rng(5)
    a = np.asfortranarray(rng.standard_normal((2, 6, 5, 2, 3)))
    b = np.asfortranarray(rng.standard_normal((2, 7, 5, 2, 3)))
    rs.write_restart(tmp_path, 1, [a, b], 0.25, 40)
    f = tmp_path / "restart1" / "star00" / "restart00000.dat"
    assert f.stat().st_size == 8 * (a.size + b.size) + 16
    (fa, fb), tt, k, slot = rs.read_restart(tmp_path, [a.shape, b.shape])
    np.testing.assert_array_equal(fa, a)
    np.testing.assert_array_equal(fb, b)
    assert (tt, k, slot) == (0.25, 40, 1)
    first, _, _, _ = rs.read_restart(tmp_path, a.shape)
    np.testing.assert_array_equal(first, a)
    with pytest.raises(ValueError):
        rs.read_restart(tmp_path, [a.shape, a.shape])


@pytest.mark.parametrize("face", [1, 2, 3, 4])
def test_ns_symmetry_mirror_reflects_momentum_about_the_face(face):
    """Symmetry_or_slidewall: on a parallelogram mesh the i-faces lie along
    (a, 1), so their unit normal is (1, -a)/sqrt(1+a^2), and the j-faces along
    (1, 0).  Every ghost cell written from an interior one keeps its density
    and energy, its momentum is the interior one reflected in the face:
    the sum has no normal part, the difference no tangential part."""
    cs = load_pkg().cases
    capi = load_pkg().capi
    a = 0.4
    ni, nj = 7, 6
    s = np.linspace(0.0, 1.0, ni)
    t = np.linspace(0.0, 0.8, nj)
    S, T = np.meshgrid(s, t, indexing="ij")
    xn, yn = np.asfortranarray(S + a * T), np.asfortranarray(T)
    ranges = [None, (1, 1, 1, nj), (1, ni, 1, 1), (ni, ni, 1, nj), (1, ni, nj, nj)]
    # the symmetry face, and walls elsewhere so that every ghost node is set
    bcs = [cs.bc_from_range(*ranges[face], 3, ni, nj)]
    bcs += [cs.bc_from_range(*ranges[f], 2, ni, nj) for f in (1, 2, 3, 4) if f != face]
    assert bcs[0].code == capi.BC_SYMMETRY
    mesh = cs.build_mesh(xn, yn, bcs)
    v, w = cs.nc_velocity_grid(4, 3.0)
    case = cs.Case("sym", mesh, bcs, v, w, v.copy(), w.copy(), cs.constants_from_kn(0.1), "rest")
    g = np.random.default_rng(face)
    u = np.zeros((4, ni + 3, nj + 3), order="F")
    u[0] = 1.0 + g.random((ni + 3, nj + 3))
    u[1:3] = g.standard_normal((2, ni + 3, nj + 3))
    u[3] = 2.0 + g.random((ni + 3, nj + 3))
    out = cs.ns_boundary_init(case, u, np.zeros(1))
    n = np.array([1.0, -a]) / math.sqrt(1 + a * a) if face in (1, 3) else np.array([0.0, 1.0])
    tv = np.array([-n[1], n[0]])
    pairs = []
    if face in (1, 3):
        for j in range(1, nj + 1):
            for k in (1, 2):
                pairs.append(((k, j), (1 - k, j)) if face == 1 else ((ni - k, j), (ni + k - 1, j)))
    else:
        for i in range(1, ni + 1):
            for k in (1, 2):
                pairs.append(((i, k), (i, 1 - k)) if face == 2 else ((i, nj - k), (i, nj + k - 1)))
    changed = np.zeros((ni + 3, nj + 3), bool)
    for (i1, j1), (i2, j2) in pairs:
        qi, qg = out[:, i1 + 1, j1 + 1], out[:, i2 + 1, j2 + 1]
        assert qg[0] == qi[0] and qg[3] == qi[3]
        assert abs((qg[1:3] + qi[1:3]) @ n) < 1e-13
        assert abs((qg[1:3] - qi[1:3]) @ tv) < 1e-13
        changed[i2 + 1, j2 + 1] = True
    np.testing.assert_array_equal(out[:, ~changed], u[:, ~changed])


def test_passive_faces_are_accepted_by_the_parameter_checks():
    """gsis_create takes symmetry (-3) and inner (>= 1) codes; the outlet (-6)
    and the reference's undefined -1 stay refused.  A last, malformed BC makes
    every call stop inside the parameter checks, before any device work: a
    code that passed is reported at that last BC, a refused one at its own."""
    import ctypes

    capi = load_pkg().capi
    cs = load_pkg().cases
    lib = capi.load_library()
    base = cs.cavity_case(6, 4, 0.1)
    for code, ok in ((capi.BC_SYMMETRY, True), (2, True), (-6, False), (-1, False)):
        bcs = list(base.bcs)
        bcs[1] = cs.BC(bcs[1].face, bcs[1].ist, bcs[1].iend, bcs[1].jst, bcs[1].jend, code)
        bcs.append(cs.BC(1, 2, 2, 1, base.ny, capi.BC_WALL))  # face 1 must sit at i = 1
        case = cs.Case("v", base.mesh, bcs, base.vi, base.wi, base.vj, base.wj, base.const, "rest")
        pack = cs.ParamPack(case)
        h = ctypes.c_void_p()
        rc = lib.gsis_create(ctypes.byref(pack.p), ctypes.byref(h))
        msg = lib.gsis_last_error().decode()
        if ok:
            assert rc == capi.GSIS_EARG and msg.startswith("BC 4:"), msg
        else:
            assert rc == capi.GSIS_EUNSUPPORTED and msg.startswith("BC 1:"), msg


def _split_cavity_cases(n=8, nv=8, rot=True):
    g = load_pkg()
    whole = g.cases.cavity_case(n, nv, kn=0.1, lid=0.15, dt=0.5)
    nx, ny = whole.nx, whole.ny
    xn = np.asfortranarray(whole.mesh.x[2:nx + 2, 2:ny + 2])
    yn = np.asfortranarray(whole.mesh.y[2:nx + 2, 2:ny + 2])
    nodes, bcs = g.multiblock.split_nodes(xn, yn, whole.bcs, nx // 2 + 1, 0, rot)
    return g.multiblock.block_cases(nodes, bcs, whole)


def test_multiblock_driver_save_and_resume_on_oracle_blocks(tmp_path):
    """MultiBlockSolver's restart path (backup_Mesco_write / backup_Mesco_read,
    src/Backup_write_read.f90:56-209) on oracle blocks: one record per block in
    block order, the step counter from flag.dat, and the inner-face exchange
    after every block's distribution is back, so the inner ghost cells hold the
    neighbour's saved cells (the exchange itself is pinned above)."""
    g = load_pkg()
    orc = load_oracle()
    cases, links = _split_cavity_cases()

    def make(c):
        return orc.OracleSolver(c)

    a = g.multiblock.MultiBlockSolver(cases, links, make=make)
    a.initialize()
    a.step(3)
    a.save_restart(tmp_path, 2)
    saved = a.frd()
    res_a = a.residual()
    a.close()
    b = g.multiblock.MultiBlockSolver(cases, links, make=make)
    b.initialize()
    assert b.resume(tmp_path) == 2
    assert b.counters()[0] == 3
    exp = [f.copy() for f in saved]
    for ln in links:
        c1, c2 = cases[ln.block - 1], cases[ln.nblock - 1]
        orc.inner_exchange_arrays(exp[ln.block - 1], np.zeros((11, c1.nx + 7, c1.ny + 7)), c1.bcs[ln.ksub - 1],
                                  saved[ln.nblock - 1], np.zeros((11, c2.nx + 7, c2.ny + 7)), c2.bcs[ln.msub - 1],
                                  ln.orient)
    for f, e, f0 in zip(b.frd(), exp, saved):
        np.testing.assert_array_equal(f[:, 1:-1, 1:-1], e[:, 1:-1, 1:-1])
        assert not np.array_equal(e, f0)  # the exchange changed the inner ghost cells
    assert np.all(res_a > 0)
    b.close()
